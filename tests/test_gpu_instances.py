"""GPU parity of two-level scenes (instances) against the instance
restatement of the oracle, through the C-ABI.

Bar: every RTCRayHit / RTCRay record byte-identical to the oracle's. Both
move the ray into object space with the same world-to-object transform
(the reference's rcp(local2world), formed on the host) and the same FMA
sequence (kernels/geometry/instance_intersector.cpp:52-98), test the
object's triangles with the original [tnear, tfar], and pick the
lexicographic minimum of (t, instID, geomID, primID); Ng stays in object
space and instID[0] is the instance's geomID in the parent."""
import numpy as np
import pytest

from conftest import mismatches

pytestmark = pytest.mark.gpu
INV = 0xFFFFFFFF


def assert_same(a, b, what):
    bad = mismatches(a, b)
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(a)} records differ, first {bad[:5]}: {a[bad[:2]]} vs {b[bad[:2]]}"


def occ_of(rtc, rays, tfar=None):
    o = np.zeros(len(rays), rtc.RTCRay)
    for k in rtc.RTCRay.names:
        o[k] = rays[k]
    if tfar is not None:
        o["tfar"] = tfar
    return o


def aimed_rays(rtc, transforms, n, seed, spread=1.5):
    """Rays from random points toward random instance centres (plenty of hits)."""
    rng = np.random.default_rng(seed)
    r = rtc.new_rayhits(n)
    org = rng.uniform(-6.0, 6.0, (n, 3)).astype(np.float32)
    tgt = transforms[rng.integers(0, len(transforms), n), 9:12] + rng.uniform(-spread, spread, (n, 3))
    d = (tgt - org).astype(np.float32)
    r["org_x"], r["org_y"], r["org_z"] = org.T
    r["dir_x"], r["dir_y"], r["dir_z"] = d.T
    return r


class TwoLevel:
    """The same two-level scene on the GPU and in the oracle."""

    def __init__(self, rtc, oracle, device, n_inst=40, seed=3, own=True, robust_obj=False, extent=4.0):
        self.rtc = rtc
        sph = rtc.gen_sphere(12)
        hf = rtc.gen_heightfield(14, seed=seed)
        soup = rtc.gen_soup(300, seed=seed + 1, h=0.08)
        own_soup = rtc.gen_soup(200, seed=seed + 2, h=0.05)
        own_soup = (own_soup[0] * np.float32(6.0) - np.float32(3.0), own_soup[1])
        fl = rtc.RTC_SCENE_FLAG_ROBUST if robust_obj else 0
        self.objs = [rtc.Scene(device, [sph], flags=fl).commit(),
                     rtc.Scene(device, [hf, soup]).commit()]
        self.top = rtc.Scene(device, [own_soup] if own else [])
        self.T = rtc.random_transforms(n_inst, seed=seed + 3, extent=extent)
        self.ids = []
        self.O = oracle.Instances()
        oa = self.O.add_object([sph], robust=robust_obj)
        ob = self.O.add_object([hf, soup])
        if own:
            self.O.add_identity(self.O.add_object([own_soup]))
        for k, t in enumerate(self.T):
            gid = self.top.add_instance(self.objs[k % 2], t)
            self.ids.append(gid)
            self.O.add_instance(oa if k % 2 == 0 else ob, t, gid)
        self.top.commit()

    def release(self):
        self.top.release()
        for o in self.objs:
            o.release()


def test_instances_closest_hit_parity(rtc, oracle, device):
    s = TwoLevel(rtc, oracle, device)
    rays = np.concatenate([aimed_rays(rtc, s.T, 3000, seed=9), rtc.gen_random_rays(1000, seed=10, lo=-6, hi=6)])
    got = rays.copy()
    s.top.intersect1M(got)
    exp = rays.copy()
    s.O.intersect(exp)
    assert_same(got, exp, "two-level closest hit")
    hit = exp["geomID"] != INV
    assert hit.mean() > 0.3
    inst = exp["instID"][hit]
    assert (inst == INV).any() and (inst != INV).sum() > 100 and len(np.unique(inst)) > 10
    s.release()


def test_instances_occlusion_parity(rtc, oracle, device):
    s = TwoLevel(rtc, oracle, device, seed=5)
    rays = occ_of(rtc, np.concatenate([aimed_rays(rtc, s.T, 3000, seed=11),
                                        rtc.gen_random_rays(1000, seed=12, lo=-6, hi=6)]))
    for tfar in (np.inf, 2.0):
        r = rays.copy()
        r["tfar"] = tfar
        got = r.copy()
        s.top.occluded1M(got)
        exp = r.copy()
        s.O.occluded(exp)
        assert_same(got, exp, f"two-level occlusion tfar={tfar}")
        assert 0 < (exp["tfar"] < 0).sum() < len(r)
    s.release()


def test_instances_without_own_triangles_and_robust_object(rtc, oracle, device):
    s = TwoLevel(rtc, oracle, device, seed=7, own=False, robust_obj=True)
    rays = aimed_rays(rtc, s.T, 3000, seed=13)
    got = rays.copy()
    s.top.intersect1M(got)
    exp = rays.copy()
    s.O.intersect(exp)
    assert_same(got, exp, "instances only, robust object")
    assert (exp["geomID"] != INV).mean() > 0.3
    s.release()


def test_instance_bvh_oracle_agrees_with_brute_force(rtc, oracle, device):
    # the per-object-BVH restatement that the large-scene checks use
    s = TwoLevel(rtc, oracle, device, seed=17)
    rays = aimed_rays(rtc, s.T, 2000, seed=18)
    a, b = rays.copy(), rays.copy()
    s.O.intersect(a)
    s.O.intersect_bvh(b)
    assert_same(a, b, "oracle instance BVH vs brute force")
    got = rays.copy()
    s.top.intersect1M(got)
    assert_same(got, b, "two-level vs oracle instance BVH")
    s.release()


def test_full_instance_stack_skips_instances(rtc, oracle, device):
    # a context whose instID[0] is already set: every instance push fails
    # (kernels/common/instance_stack.h:21-38) and only own triangles are hit
    s = TwoLevel(rtc, oracle, device, seed=19)
    rays = np.concatenate([aimed_rays(rtc, s.T, 1000, seed=20), rtc.gen_random_rays(2000, seed=21, lo=-3, hi=3)])
    ctx = rtc.init_context()
    ctx.instID[0] = 77
    got = rays.copy()
    s.top.intersect1M(got, ctx=ctx)
    exp = rays.copy()
    s.O.intersect(exp, ctx_inst=77)
    assert_same(got, exp, "full instance stack")
    hit = exp["geomID"] != INV
    assert hit.any() and (exp["instID"][hit] == 77).all()
    s.release()


def test_instance_transform_formats_and_bounds(rtc, device):
    L = rtc.lib()
    obj = rtc.Scene(device, [rtc.gen_sphere(8)]).commit()
    top = rtc.Scene(device)
    g = L.rtcNewGeometry(device.h, rtc.RTC_GEOMETRY_TYPE_INSTANCE)
    L.rtcSetGeometryInstancedScene(g, obj.h)
    m = np.array([[0, -2, 0, 5], [2, 0, 0, -1], [0, 0, 2, 3]], np.float32)  # 3x4 row major: x -> 2y etc.
    L.rtcSetGeometryTransform(g, 0, rtc.RTC_FORMAT_FLOAT3X4_ROW_MAJOR, m.ctypes.data)
    col = np.zeros(12, np.float32)
    L.rtcGetGeometryTransform(g, 0.0, rtc.RTC_FORMAT_FLOAT3X4_COLUMN_MAJOR, col.ctypes.data)
    assert col.tolist() == m.T.reshape(-1).tolist()
    c44 = np.zeros(16, np.float32)
    L.rtcGetGeometryTransform(g, 0.0, rtc.RTC_FORMAT_FLOAT4X4_COLUMN_MAJOR, c44.ctypes.data)
    assert c44.reshape(4, 4)[:, :3].reshape(-1).tolist() == col.tolist() and c44[3::4].tolist() == [0, 0, 0, 1]
    row = np.zeros(12, np.float32)
    L.rtcGetGeometryTransform(g, 0.0, rtc.RTC_FORMAT_FLOAT3X4_ROW_MAJOR, row.ctypes.data)
    assert row.tolist() == m.reshape(-1).tolist()
    L.rtcSetGeometryTransform(g, 0, rtc.RTC_FORMAT_FLOAT4X4_ROW_MAJOR, c44.ctypes.data)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_OPERATION
    L.rtcSetGeometryTransform(g, 1, rtc.RTC_FORMAT_FLOAT3X4_ROW_MAJOR, m.ctypes.data)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_OPERATION
    L.rtcCommitGeometry(g)
    L.rtcAttachGeometry(top.h, g)
    L.rtcReleaseGeometry(g)
    top.commit()
    b = top.bounds()
    # unit sphere scaled by 2 around (5, -1, 3)
    lo = np.array([b.lower_x, b.lower_y, b.lower_z]); hi = np.array([b.upper_x, b.upper_y, b.upper_z])
    assert np.allclose(lo, [3, -3, 1], atol=1e-5) and np.allclose(hi, [7, 1, 5], atol=1e-5)
    # a ray down the instance's z axis (just off the pole vertex) meets the
    # scaled sphere's top at z = 3 + 2, i.e. t = 10 - 5
    r = rtc.new_rayhits(1)
    r["org_x"], r["org_y"], r["org_z"] = 5.01, -1.02, 10.0
    r["dir_x"], r["dir_y"], r["dir_z"] = 0.0, 0.0, -1.0
    top.intersect1M(r)
    assert r["geomID"][0] == 0 and r["instID"][0] == 0 and abs(r["tfar"][0] - 5.0) < 0.05
    top.release()
    obj.release()


def test_instance_error_paths(rtc, device):
    L = rtc.lib()
    obj = rtc.Scene(device, [rtc.gen_sphere(8)])  # not committed
    top = rtc.Scene(device)
    top.add_instance(obj, np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32))
    L.rtcCommitScene(top.h)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_OPERATION
    obj.commit()
    top.commit()
    r = rtc.new_rayhits(4)
    r["org_x"], r["org_y"], r["org_z"] = 0.01, 0.02, 5.0
    r["dir_z"] = -1.0
    top.intersect1M(r)
    assert (r["geomID"] == 0).all() and (r["instID"] == 0).all()
    # the instanced scene changes and is committed again: the parent keeps
    # the instanced scene as it was at the parent's commit until the parent
    # is committed again
    g0 = L.rtcGetGeometry(obj.h, 0)
    L.rtcDisableGeometry(g0)
    obj.commit()
    r2 = rtc.new_rayhits(4)
    r2["org_x"], r2["org_y"], r2["org_z"] = 0.01, 0.02, 5.0
    r2["dir_z"] = -1.0
    top.intersect1M(r2)
    assert (r2["geomID"] == 0).all()
    top.commit()
    r2 = rtc.new_rayhits(4)
    r2["org_x"], r2["org_y"], r2["org_z"] = 0.01, 0.02, 5.0
    r2["dir_z"] = -1.0
    top.intersect1M(r2)
    assert (r2["geomID"] == INV).all()
    L.rtcEnableGeometry(g0)
    obj.commit()
    # a triangle geometry takes no transform
    g = L.rtcNewGeometry(device.h, rtc.RTC_GEOMETRY_TYPE_TRIANGLE)
    L.rtcSetGeometryTransform(g, 0, rtc.RTC_FORMAT_FLOAT3X4_COLUMN_MAJOR, np.zeros(12, np.float32).ctypes.data)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_OPERATION
    L.rtcReleaseGeometry(g)
    # an instance of an empty scene and a disabled instance contribute nothing
    empty = rtc.Scene(device).commit()
    top2 = rtc.Scene(device)
    top2.add_instance(empty, np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32))
    top2.add_instance(obj, np.array([1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0], np.float32), enabled=False)
    top2.commit()
    r3 = rtc.new_rayhits(4)
    r3["org_x"], r3["org_y"], r3["org_z"] = 0.01, 0.02, 5.0
    r3["dir_z"] = -1.0
    top2.intersect1M(r3)
    assert (r3["geomID"] == INV).all()
    for x in (top2, empty, top, obj):
        x.release()


def test_instances_host_stride_and_device_records(rtc, oracle, device):
    s = TwoLevel(rtc, oracle, device, seed=23, n_inst=16)
    rays = aimed_rays(rtc, s.T, 1500, seed=24)
    exp = rays.copy()
    s.O.intersect(exp)
    # strided host records
    wide = np.zeros(len(rays), dtype=[("r", rtc.RTCRayHit), ("pad", "u1", 16)])
    wide["r"] = rays
    s.top.intersect1M(wide, stride=wide.itemsize)
    assert_same(wide["r"].copy(), exp, "strided host records")
    # device records through the asynchronous entry point
    L = rtc.lib()
    d = device.malloc(rays.nbytes)
    device.memcpy(d, rays.ctypes.data, rays.nbytes)
    L.rtcgIntersect1MAsync(s.top.h, None, d, len(rays), 80, None)
    L.rtcgSynchronize(device.h)
    out = np.zeros_like(rays)
    device.memcpy(out.ctypes.data, d, out.nbytes)
    device.free(d)
    assert_same(out, exp, "device records")
    s.release()


def test_learned_split_cost_at_both_levels(rtc, oracle, device):
    """The split model of a two-level scene: samples collected for the top
    tree (object None) and for each instanced scene's tree from the visits of
    the two-level traversal (rtcgSplitModelCollectInstanced), one model per
    tree trained, the objects and a new top scene built with them; the
    records equal the oracle's, and samples for a scene the top tree does not
    instance, or from a scene without instances, are refused."""
    import ctypes
    s = TwoLevel(rtc, oracle, device, n_inst=60)
    rays = np.concatenate([aimed_rays(rtc, s.T, 6000, seed=21), rtc.gen_random_rays(2000, seed=22, lo=-6, hi=6)])
    d = device.malloc(rays.nbytes)
    models = []
    try:
        device.memcpy(d, rays.ctypes.data, rays.nbytes)
        for o in s.objs:
            m = rtc.SplitModel(device, seed=1)
            m.collect_instanced(s.top, o, d, len(rays))
            X, T = m.samples()
            assert len(X) > 20 and np.isfinite(X).all() and np.isfinite(T).all()
            # a child is entered only after its parent (the target's bound)
            assert (T + 16 * X[:, 0]).max() <= 1e-4
            m.train(100, batch=4096, lr=3e-3, seed=1)
            models.append(m)
        mt = rtc.SplitModel(device, seed=1)
        mt.collect_instanced(s.top, None, d, len(rays))
        assert mt.samples()[0].shape[0] > 10
        mt.train(100, batch=4096, lr=3e-3, seed=1)
        models.append(mt)
        L = rtc.lib()
        for o, m in zip(s.objs, models):
            L.rtcgSetSceneSplitModel(o.h, m.h)
            o.commit()
        top2 = rtc.Scene(device, [])
        for k, t in enumerate(s.T):
            top2.add_instance(s.objs[k % 2], t)
        L.rtcgSetSceneSplitModel(top2.h, mt.h)
        top2.commit()
        # top2 has no own triangles: compare with an oracle scene without them
        O = oracle.Instances()
        oa = O.add_object([rtc.gen_sphere(12)])
        ob = O.add_object([rtc.gen_heightfield(14, seed=3), rtc.gen_soup(300, seed=4, h=0.08)])
        for k, t in enumerate(s.T):
            O.add_instance(oa if k % 2 == 0 else ob, t, k)
        got = rays.copy()
        top2.intersect1M(got)
        exp = rays.copy()
        O.intersect(exp)
        assert_same(got, exp, "learned two-level closest hit")
        other = rtc.Scene(device, [rtc.gen_sphere(8)]).commit()
        with pytest.raises(RuntimeError):
            models[0].collect_instanced(s.top, other, d, len(rays))
        with pytest.raises(RuntimeError):
            models[0].collect_instanced(other, None, d, len(rays))
        other.release()
        top2.release()
        for o in s.objs:
            L.rtcgSetSceneSplitModel(o.h, None)
    finally:
        device.free(d)
        for m in models:
            m.release()
        s.release()
