"""CPU restatement of the learned split-cost MLP.

TEST INFRASTRUCTURE ONLY: the parity tests compare the GPU library's
inference and gradient kernels against these functions. Nothing in the
product imports this module.

The reference has no learned component: its split score is the binned-SAH
`BinInfoT::best` (kernels/builders/heuristic_binning.h:340-393). This module
restates the network with which the GPU library corrects the half-area
term, and the features it is fed (the per-bin statistics of a split
candidate): 12 inputs -> 32 ReLU -> 32 ReLU -> 1, weights and activations
rounded to
FP16, products accumulated in higher precision, FP32 biases; loss
sum_i w_i (y_i - t_i)^2 / n. With quantize=False every value stays float64,
which is the form the finite-difference test pins.

Parameter vector: W1[32][12] | b1[32] | W2[32][32] | b2[32] | w3[32] | b3.
"""
import numpy as np

IN, H = 12, 32
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


def split_features(clo, chi, ccount, plo, phi, pcount, scene_lo, scene_hi):
    """Features of child boxes C (n x 3 lo/hi, ccount primitives) of parent
    boxes S (plo/phi, pcount), as the builder forms them for each side of a
    split candidate from its bins and the trainer for each (parent, child)
    pair of a binary tree: log2(A_C/A_S + 2^-30)/16, n_C/n_S, centre offset
    over the parent's extent (3), extent ratio (3, 1 on a flat axis), the
    parent's centre in the scene box (3), log2(A_S/A_scene + 2^-30)/16;
    float32 arithmetic in the builder's order."""
    f32 = np.float32
    clo, chi = (np.asarray(a, f32).reshape(-1, 3) for a in (clo, chi))
    plo, phi = (np.asarray(a, f32).reshape(-1, 3) for a in (plo, phi))
    slo = np.asarray(scene_lo, f32)
    sext = np.asarray(scene_hi, f32) - slo
    sinv = np.where(sext > 0, f32(1) / np.where(sext > 0, sext, 1), 0).astype(f32)
    ha = lambda d: d[:, 0] * (d[:, 1] + d[:, 2]) + d[:, 1] * d[:, 2]
    sa = f32(sext[0] * (sext[1] + sext[2]) + sext[1] * sext[2])
    s_inv_area = f32(1) / sa if sa > 0 else f32(0)
    ac, ap = ha(chi - clo), ha(phi - plo)
    n = len(clo)
    f = np.zeros((n, IN), f32)
    with np.errstate(divide="ignore", invalid="ignore"):
        f[:, 0] = np.where(ap > 0, np.log2(ac / np.where(ap > 0, ap, 1) + f32(2.0 ** -30)) * f32(0.0625), 0)
        pc = np.asarray(pcount, np.float64) * np.ones(n)
        f[:, 1] = np.where(pc > 0, np.asarray(ccount, np.float64) / np.where(pc > 0, pc, 1), 1)
        ep = phi - plo
        inv = np.where(ep > 0, f32(1) / np.where(ep > 0, ep, 1), 0).astype(f32)
        f[:, 2:5] = ((chi + clo) - (phi + plo)) * f32(0.5) * inv
        f[:, 5:8] = np.where(ep > 0, (chi - clo) * inv, 1)
        f[:, 8:11] = ((phi + plo) * f32(0.5) - slo) * sinv
        f[:, 11] = np.log2(ap * s_inv_area + f32(2.0 ** -30)) * f32(0.0625)
    return f
