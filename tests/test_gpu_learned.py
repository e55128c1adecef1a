"""Learned split cost on the GPU: inference and gradient kernels against the
CPU restatement (oracle/mlp.py), training, visit collection, and trees built
with a trained model giving the same hits as the oracle.

Tolerances: the GPU forward pass rounds weights and activations to FP16 and
accumulates in FP32 (v_dot2_f32_f16); the oracle rounds the same values to
FP16 but accumulates in float64, so an activation that lands on an FP16
rounding boundary may differ by one FP16 ulp. Predictions are compared
within 2e-3 (1 + |y|) and gradients within 5e-3 of their norm."""
import numpy as np
import pytest

from conftest import mismatches
from oracle import mlp

pytestmark = pytest.mark.gpu
INV = 0xFFFFFFFF


def rand_params(seed):
    rng = np.random.default_rng(seed)
    p = np.zeros(mlp.NPARAM, np.float32)
    p[mlp.OFF_W1:mlp.OFF_B1] = rng.uniform(-1, 1, mlp.H * mlp.IN) * np.sqrt(6 / mlp.IN)
    p[mlp.OFF_B1:mlp.OFF_W2] = rng.uniform(-0.2, 0.2, mlp.H)
    p[mlp.OFF_W2:mlp.OFF_B2] = rng.uniform(-1, 1, mlp.H * mlp.H) * np.sqrt(6 / mlp.H)
    p[mlp.OFF_B2:mlp.OFF_W3] = rng.uniform(-0.2, 0.2, mlp.H)
    p[mlp.OFF_W3:mlp.OFF_B3] = rng.uniform(-1, 1, mlp.H) * np.sqrt(3 / mlp.H)
    p[mlp.OFF_B3] = -3.0
    return p


@pytest.fixture(scope="module")
def model(rtc, device):
    m = rtc.SplitModel(device, seed=7)
    yield m
    m.release()


def test_param_layout(rtc, model):
    assert rtc.SplitModel.nparam() == mlp.NPARAM
    p = rand_params(1)
    model.set_params(p)
    assert np.array_equal(model.get_params(), p)


def test_predict_matches_oracle(model):
    p = rand_params(2)
    model.set_params(p)
    X = np.random.default_rng(3).uniform(-1, 1, (1000, mlp.IN)).astype(np.float32)
    y = model.predict(X)
    ref, _ = mlp.forward(p, X)
    assert np.abs(y - ref).max() <= 2e-3 * (1 + np.abs(ref).max())


@pytest.mark.parametrize("n", [1, 63, 64, 1000])
def test_gradient_matches_oracle(model, n):
    p = rand_params(4 + n)
    model.set_params(p)
    rng = np.random.default_rng(n)
    X = rng.uniform(-1, 1, (n, mlp.IN)).astype(np.float32)
    T = rng.normal(-4, 2, n).astype(np.float32)
    W = rng.uniform(0.5, 2, n).astype(np.float32)
    loss, g = model.gradient(X, T, W)
    rloss, rg = mlp.loss_and_grad(p, X, T, W)
    assert abs(loss - rloss) <= 2e-3 * rloss + 1e-6
    assert np.linalg.norm(g - rg) <= 5e-3 * np.linalg.norm(rg)
    # every parameter block carries gradient (a dropped term shows here)
    for a, b in [(mlp.OFF_W1, mlp.OFF_B1), (mlp.OFF_B1, mlp.OFF_W2), (mlp.OFF_W2, mlp.OFF_B2),
                 (mlp.OFF_B2, mlp.OFF_W3), (mlp.OFF_W3, mlp.OFF_B3), (mlp.OFF_B3, mlp.NPARAM)]:
        assert np.linalg.norm(g[a:b] - rg[a:b]) <= 1e-2 * np.linalg.norm(rg[a:b]) + 1e-7


def heightfield_case(rtc, n=256):
    v, ix = rtc.gen_heightfield(n, seed=5, amplitude=0.2, octaves=6)
    cam = rtc.gen_camera(160, 120, (0.5, -0.4, 0.45), (0.5, 0.45, 0.0), up=(0, 0, 1), fov=50.0)
    return v, ix, cam


def test_collect_and_train(rtc, device, oracle):
    v, ix, cam = heightfield_case(rtc)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    m = rtc.SplitModel(device, seed=11)
    d = device.malloc(cam.nbytes)
    try:
        device.memcpy(d, cam.ctypes.data, cam.nbytes)
        m.collect(sc, d, len(cam))
        X, T = m.samples()
        assert X.shape[1] == mlp.IN and len(X) == len(T) > 100
        # one sample per (parent, child) pair of a binary tree: the target is
        # log2 of the measured P(C | S) minus log2 of the area ratio, and a
        # child is entered only after its parent, so their sum is at most 0
        logp = T + 16 * X[:, 0]
        assert logp.max() <= 1e-4 and X[:, 0].max() <= 1e-6
        assert (X[:, 1] > 0).all() and (X[:, 1] < 1).all()              # a child holds part of the parent
        assert (X[:, 5:8] >= -1e-6).all() and (X[:, 5:8] <= 1 + 1e-6).all()  # and lies inside it
        assert (np.abs(X[:, 2:5]) <= 0.5 + 1e-6).all()
        assert (X[:, 8:11] >= -1e-6).all() and (X[:, 8:11] <= 1 + 1e-6).all()
        l0, _ = m.gradient(X, T)
        last = m.train(300, batch=4096, lr=3e-3, seed=1)
        l1, _ = m.gradient(X, T)
        assert l1 < 0.5 * l0 and last < 0.5 * l0
        pred = m.predict(X)
        # conditional entry rates are noisy (a few rays per small node) and
        # the sample order follows the node layout, which the collapse
        # allocates in no fixed order, so the fit varies between runs (the
        # correlation has measured 0.16 to 0.4): it must still track them
        assert np.corrcoef(pred, T)[0, 1] > 0.1
        # a scene built with the trained model: same records as the oracle
        sc2 = rtc.Scene(device, [(v, ix)])
        rtc.lib().rtcgSetSceneSplitModel(sc2.h, m.h)
        sc2.commit()
        out = cam.copy()
        sc2.intersect1M(out)
        ref = cam.copy()
        oracle.BVH([(v, ix)]).intersect(ref)
        bad = mismatches(out, ref)
        assert len(bad) == 0, bad[:5]
        assert (out["geomID"] != INV).mean() > 0.5
        sc2.release()
    finally:
        device.free(d)
        m.release()
        sc.release()


def test_sah_slack_bounds_the_tree(rtc, device):
    """The model chooses only among split candidates whose plain binned-SAH
    cost is within (1 + slack) of the segment's least: slack 0 gives the plain
    tree back, a small slack keeps the tree's SAH cost near the plain tree's,
    and a negative or NaN slack is an invalid argument."""
    import ctypes
    v, ix, cam = heightfield_case(rtc)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    m = rtc.SplitModel(device, seed=11)
    d = device.malloc(cam.nbytes)
    try:
        device.memcpy(d, cam.ctypes.data, cam.nbytes)
        m.collect(sc, d, len(cam))
        m.train(300, batch=4096, lr=3e-3, seed=1)
        plain = sc.stats()
        got = {}
        for slack in (0.0, 0.05, float("inf")):
            rtc.lib().rtcgSetSplitModelSahSlack(m.h, ctypes.c_float(slack))
            device.check("slack")
            s2 = rtc.Scene(device, [(v, ix)])
            rtc.lib().rtcgSetSceneSplitModel(s2.h, m.h)
            s2.commit()
            got[slack] = s2.stats()
            s2.release()
        assert got[0.0]["num_nodes"] == plain["num_nodes"]
        assert abs(got[0.0]["sah"] - plain["sah"]) <= 1e-4 * plain["sah"], (got[0.0]["sah"], plain["sah"])
        assert got[0.05]["sah"] <= 1.05 * plain["sah"], (got[0.05]["sah"], plain["sah"])
        for bad in (-0.5, float("nan")):
            with pytest.raises(rtc.RTCError) as e:
                rtc.lib().rtcgSetSplitModelSahSlack(m.h, ctypes.c_float(bad))
                device.check("slack")
            assert e.value.code == rtc.RTC_ERROR_INVALID_ARGUMENT
    finally:
        device.free(d)
        m.release()
        sc.release()


def _counted_nodes(rtc, device, sc, rays):
    """Mean node visits per ray (the library's counted traversal) and the
    records it leaves."""
    import ctypes  # noqa: F401
    n = len(rays)
    d_rays, d_cn, d_ct = device.malloc(rays.nbytes), device.malloc(4 * n), device.malloc(4 * n)
    try:
        device.memcpy(d_rays, rays.ctypes.data, rays.nbytes)
        rtc.lib().rtcgTraceCounted(sc.h, d_rays, n, 80, 0, d_cn, d_ct)
        device.check("rtcgTraceCounted")
        cn = np.zeros(n, np.uint32)
        device.memcpy(cn.ctypes.data, d_cn, 4 * n)
        out = np.zeros(n, rays.dtype)
        device.memcpy(out.ctypes.data, d_rays, rays.nbytes)
        return float(cn.mean()), out
    finally:
        for p in (d_rays, d_cn, d_ct):
            device.free(p)


@pytest.mark.parametrize("width,bound", [(0.3, 0.9), (0.05, 0.75)])
def test_held_out_gain_for_a_point_source(rtc, device, oracle, width, bound):
    """Where the rays are far from the uniform distribution the SAH assumes,
    a model trained on half of them builds a tree that the other half (held
    out) crosses with fewer node visits than the plain binned-SAH tree (at
    most `bound` times as many), with the same records, equal to the
    oracle's on a sample. Workload: the 2^20-triangle soup, 2^22 rays from
    the point (0.5, 0.5, -2) into the central cube `width` wide; even rays
    train, odd rays are counted. Training is deterministic (samples in
    content order, gradients summed in a fixed order), so the ratio is
    reproducible: 0.80 (width 0.3) and 0.69 (width 0.05) measured with 200
    Adam steps at SAH slack 0.5."""
    import ctypes
    v, ix = rtc.gen_soup(1 << 20, seed=1, h=0.005)
    rng = np.random.default_rng(7)
    n = 1 << 22
    tgt = (0.5 - width / 2) + width * rng.random((n, 3), dtype=np.float32)
    org = np.array([0.5, 0.5, -2.0], np.float32)
    d = tgt - org
    Q = rtc.new_rayhits(n)
    Q["org_x"], Q["org_y"], Q["org_z"] = org
    Q["dir_x"], Q["dir_y"], Q["dir_z"] = d.T
    train, held = np.ascontiguousarray(Q[::2]), np.ascontiguousarray(Q[1::2])
    plain = rtc.Scene(device, [(v, ix)]).commit()
    m = rtc.SplitModel(device, seed=1)
    d_train = device.malloc(train.nbytes)
    try:
        device.memcpy(d_train, train.ctypes.data, train.nbytes)
        m.collect(plain, d_train, len(train))
        m.train(200, batch=65536, lr=3e-3, seed=1)
        rtc.lib().rtcgSetSplitModelSahSlack(m.h, ctypes.c_float(0.5))
        learned = rtc.Scene(device, [(v, ix)])
        rtc.lib().rtcgSetSceneSplitModel(learned.h, m.h)
        learned.commit()
        pn, pout = _counted_nodes(rtc, device, plain, held)
        ln, lout = _counted_nodes(rtc, device, learned, held)
        print(f"point source into a {width}-wide cube, held out: plain {pn:.3f} nodes/ray, learned {ln:.3f} "
              f"({ln / pn:.4f})")
        assert ln <= bound * pn, (ln, pn)
        assert np.array_equal(lout.view(np.uint8), pout.view(np.uint8))
        sample = np.ascontiguousarray(held[::512])
        ref = sample.copy()
        oracle.BVH([(v, ix)]).intersect(ref)
        bad = mismatches(np.ascontiguousarray(lout[::512]), ref)
        assert len(bad) == 0, bad[:5]
        learned.release()
    finally:
        device.free(d_train)
        m.release()
        plain.release()
