"""CPU restatement of the learned split-cost MLP.

TEST INFRASTRUCTURE ONLY: the parity tests compare the GPU library's
inference and gradient kernels against these functions. Nothing in the
product imports this module.

The reference has no learned component: its split score is the binned-SAH
`BinInfoT::best` (kernels/builders/heuristic_binning.h:340-393). This module
restates the network the GPU library substitutes for the half-area term:
8 inputs -> 32 ReLU -> 32 ReLU -> 1, weights and activations rounded to
FP16, products accumulated in higher precision, FP32 biases; loss
sum_i w_i (y_i - t_i)^2 / n. With quantize=False every value stays float64,
which is the form the finite-difference test pins.

Parameter vector: W1[32][8] | b1[32] | W2[32][32] | b2[32] | w3[32] | b3.
"""
import numpy as np

IN, H = 8, 32
OFF_W1, OFF_B1 = 0, H * IN
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1


def unpack(p):
    p = np.asarray(p, np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, IN), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H), p[OFF_B2:OFF_W3],
            p[OFF_W3:OFF_B3], p[OFF_B3])


def _q(quantize):
    if quantize:
        return lambda a: np.asarray(a, np.float64).astype(np.float16).astype(np.float64)
    return lambda a: np.asarray(a, np.float64)


def forward(p, X, quantize=True):
    """Returns y and the cached activations (x, z1, h1, z2, h2)."""
    q = _q(quantize)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    if quantize:  # biases are FP32 on the device
        b1 = b1.astype(np.float32).astype(np.float64)
        b2 = b2.astype(np.float32).astype(np.float64)
        b3 = float(np.float32(b3))
    x = q(X)
    z1 = x @ q(W1).T + b1
    h1 = q(np.maximum(z1, 0.0))
    z2 = h1 @ q(W2).T + b2
    h2 = q(np.maximum(z2, 0.0))
    y = h2 @ q(w3) + b3
    return y, (x, z1, h1, z2, h2)


def loss_and_grad(p, X, T, Wt=None, quantize=True):
    """Loss sum w (y - t)^2 / n and its gradient; back-propagation through
    the (rounded) weights, the rounding itself treated as identity."""
    q = _q(quantize)
    X = np.asarray(X, np.float64).reshape(-1, IN)
    T = np.asarray(T, np.float64)
    n = len(X)
    Wt = np.ones(n) if Wt is None else np.asarray(Wt, np.float64)
    W1, b1, W2, b2, w3, b3 = unpack(p)
    y, (x, z1, h1, z2, h2) = forward(p, X, quantize)
    e = y - T
    loss = float((Wt * e * e).sum() / n)
    dy = 2.0 * Wt * e / n
    g = np.zeros(NPARAM)
    g[OFF_W3:OFF_B3] = h2.T @ dy
    g[OFF_B3] = dy.sum()
    dz2 = (dy[:, None] * q(w3)[None, :]) * (z2 > 0)
    g[OFF_W2:OFF_B2] = (dz2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz1 = (dz2 @ q(W2)) * (z1 > 0)
    g[OFF_W1:OFF_B1] = (dz1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    return loss, g


def box_features(lo, hi, count, scene_lo, scene_hi):
    """Features of boxes (n x 3 lo/hi) holding `count` primitives, relative to
    the scene box, as the builder and the trainer form them."""
    lo = np.asarray(lo, np.float32).reshape(-1, 3)
    hi = np.asarray(hi, np.float32).reshape(-1, 3)
    slo = np.asarray(scene_lo, np.float32)
    ext = np.asarray(scene_hi, np.float32) - slo
    inv = np.where(ext > 0, np.float32(1) / np.where(ext > 0, ext, 1), 0).astype(np.float32)
    d = hi - lo
    ha = lambda v: v[:, 0] * (v[:, 1] + v[:, 2]) + v[:, 1] * v[:, 2]
    sa = ext[0] * (ext[1] + ext[2]) + ext[1] * ext[2]
    f = np.zeros((len(lo), IN), np.float32)
    f[:, 0:3] = ((lo + hi) * np.float32(0.5) - slo) * inv
    f[:, 3:6] = d * inv
    f[:, 6] = np.log2(ha(d) / sa + 2.0 ** -30) / 16
    f[:, 7] = np.log2(np.maximum(np.asarray(count, np.float64), 1)) / 16
    return f
