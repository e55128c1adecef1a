"""GPU parity: the HIP library against the CPU restatement on the same seeded
inputs, through the C-ABI (rtcore.py is a ctypes mirror).

Bar: every RTCRayHit / RTCRay record byte-identical to the oracle's
(geomID/primID exact, t/u/v/Ng bit-exact; both compute t = T/absDen with
IEEE division and pick the lexicographic minimum of (t, geomID, primID)).
Against the reference's own recorded answers the bar is the north-star
tolerance: t within 1e-6 relative, u/v within 1e-5."""
import ctypes

import numpy as np
import pytest

from conftest import mismatches, records_equal

pytestmark = pytest.mark.gpu
INV = 0xFFFFFFFF


def trace(rtc, scene, rays, occluded=False, stride=None):
    r = rays.copy()
    if occluded:
        scene.occluded1M(r, stride=stride)
    else:
        scene.intersect1M(r, stride=stride)
    return r


def oracle_trace(oracle, meshes, rays, occluded=False, brute=False, robust=False):
    r = rays.copy()
    if brute:
        (oracle.occluded_brute if occluded else oracle.intersect_brute)(meshes, r, robust=robust)
    else:
        bvh = oracle.BVH(meshes)
        assert bvh.validate() == 0
        (bvh.occluded if occluded else bvh.intersect)(r, robust=robust)
    return r


def assert_same(a, b, what):
    bad = mismatches(a, b)
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(a)} records differ, first {bad[:5]}: {a[bad[:2]]} vs {b[bad[:2]]}"


# ------------------------------------------------------------------ basics --

def test_device_properties(rtc, device):
    L = rtc.lib()
    assert L.rtcGetDeviceProperty(device.h, rtc.RTC_DEVICE_PROPERTY_VERSION) == 31201
    assert L.rtcGetDeviceProperty(device.h, rtc.RTC_DEVICE_PROPERTY_RAY_STREAM_SUPPORTED) == 1
    assert L.rtcGetDeviceProperty(device.h, rtc.RTC_DEVICE_PROPERTY_TRIANGLE_GEOMETRY_SUPPORTED) == 1
    L.rtcGetDeviceProperty(device.h, 12345)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_ARGUMENT
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_NONE


def test_single_triangle_reference_answer(rtc, device, oracle, golden_probe):
    g = golden_probe["single_triangle"]
    v = np.array(g["vertices"], np.float32)
    ix = np.array(g["indices"], np.uint32)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    r = rtc.new_rayhits(2)
    r["org_x"], r["org_y"], r["org_z"] = g["ray"]["org"]
    r["dir_x"], r["dir_y"], r["dir_z"] = g["ray"]["dir"]
    out = trace(rtc, sc, r)
    e = g["expected"]
    assert (out["geomID"] == e["geomID"]).all() and (out["primID"] == e["primID"]).all()
    assert np.all(np.abs(out["tfar"] - e["t"]) <= 1e-6 * e["t"])
    assert np.all(np.abs(out["u"] - e["u"]) <= 1e-5) and np.all(np.abs(out["v"] - e["v"]) <= 1e-5)
    assert np.array([out["Ng_x"][0], out["Ng_y"][0], out["Ng_z"][0]], np.float32).tobytes() == \
        np.array(e["Ng"], np.float32).tobytes()
    assert records_equal(out, oracle_trace(oracle, [(v, ix)], r, brute=True))
    sc.release()


def test_rtcIntersect1_and_rtcOccluded1(rtc, device):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    ix = np.array([[0, 1, 2]], np.uint32)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    r = rtc.new_rayhits(1)
    r["org_x"], r["org_y"], r["org_z"] = 0.2, 0.3, -2.0
    r["dir_z"] = 1.0
    ctx = rtc.init_context()
    rtc.lib().rtcIntersect1(sc.h, ctypes.byref(ctx), r.ctypes.data)
    device.check()
    assert r["primID"][0] == 0 and r["tfar"][0] == np.float32(2.0)
    o = np.zeros(1, rtc.RTCRay)
    o["org_x"], o["org_y"], o["org_z"], o["dir_z"], o["tfar"] = 0.2, 0.3, -2.0, 1.0, 1.5
    rtc.lib().rtcOccluded1(sc.h, ctypes.byref(ctx), o.ctypes.data)
    assert o["tfar"][0] == np.float32(1.5)  # hit at 2 is beyond tfar
    o["tfar"] = 2.5
    rtc.lib().rtcOccluded1(sc.h, ctypes.byref(ctx), o.ctypes.data)
    assert np.isneginf(o["tfar"][0])
    sc.release()


# --------------------------------------------------------- C1: sphere 65k --

@pytest.fixture(scope="module")
def c1(rtc, device):
    v, ix = rtc.gen_sphere(128)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    yield v, ix, sc
    sc.release()


def test_c1_sphere_parity_and_reference_hit_count(rtc, oracle, c1, golden_probe):
    v, ix, sc = c1
    rays = rtc.gen_pinhole(1024, 1024)
    out = trace(rtc, sc, rays)
    ref = oracle_trace(oracle, [(v, ix)], rays)
    assert int((out["geomID"] != INV).sum()) == golden_probe["uv_sphere_pinhole"]["expected_hits"]
    assert_same(out, ref, "C1 closest hit")


def test_c1_device_pointer_path(rtc, device, c1):
    """Records already in HBM are traced in place (same results as host records)."""
    v, ix, sc = c1
    rays = rtc.gen_pinhole(256, 256)
    host = trace(rtc, sc, rays)
    d = device.malloc(rays.nbytes)
    try:
        device.memcpy(d, rays.ctypes.data, rays.nbytes)
        ctx = rtc.init_context()
        rtc.lib().rtcIntersect1M(sc.h, ctypes.byref(ctx), d, len(rays), 80)
        device.check()
        back = np.zeros_like(rays)
        device.memcpy(back.ctypes.data, d, rays.nbytes)
        assert records_equal(back, host)
        # async extension on the device's stream
        device.memcpy(d, rays.ctypes.data, rays.nbytes)
        rtc.lib().rtcgIntersect1MAsync(sc.h, None, d, len(rays), 80, None)
        rtc.lib().rtcgSynchronize(device.h)
        device.check()
        device.memcpy(back.ctypes.data, d, rays.nbytes)
        assert records_equal(back, host)
    finally:
        device.free(d)


def test_c1_bvh_structure(rtc, c1):
    """Every triangle appears in exactly one leaf, every decoded child box
    contains its subtree, triangle records equal v0, v0-v1, v2-v0."""
    v, ix, sc = c1
    nodes, tris = sc.read_bvh()
    check_bvh(nodes, tris, [(v, ix)])


def check_bvh(nodes, tris, meshes):
    N = nodes.reshape(-1, 80)
    T = tris.reshape(-1, 48)
    nn, nt = len(N), len(T)
    f = N[:, :12].view(np.float32)
    e = N[:, 12:15].astype(np.uint32)
    imask = N[:, 15]
    base = N[:, 16:24].view(np.uint32)
    meta = N[:, 24:32]
    q = N[:, 32:80].reshape(nn, 3, 2, 8).astype(np.float32)  # axis, lo/hi, slot
    scale = (e << 23).view(np.float32)
    planes = q * scale[:, :, None, None] + f[:, :, None, None]  # fma vs mul+add: conservative check below
    tf = T.view(np.float32).reshape(-1, 12)
    tu = T.view(np.uint32).reshape(-1, 12)
    seen = np.zeros(nt, np.int32)
    visited = np.zeros(nn, np.int32)
    stack = [0]
    while stack:
        n = stack.pop()
        visited[n] += 1
        for c in range(8):
            m = int(meta[n, c])
            if m == 0:
                continue
            lo = planes[n, :, 0, c] - 1e-6 * np.abs(planes[n, :, 0, c]) - 1e-30
            hi = planes[n, :, 1, c] + 1e-6 * np.abs(planes[n, :, 1, c]) + 1e-30
            assert bool((imask[n] >> c) & 1) == bool(m & 0x80), (n, c)
            if (imask[n] >> c) & 1:
                stack.append(int(base[n, 0]) + (m & 0x7F))
            else:
                t = int(base[n, 1]) + m - 1
                while True:
                    seen[t] += 1
                    v0 = tf[t, 0:3]
                    v1 = v0 - tf[t, 4:7]
                    v2 = v0 + tf[t, 8:11]
                    for p in (v0, v1, v2):
                        assert (p >= lo - 1e-6).all() and (p <= hi + 1e-6).all(), (n, c, t)
                    if tu[t, 11] & 1:
                        break
                    t += 1
    assert (visited == 1).all(), "every node reached exactly once"
    assert (seen == 1).all(), "every triangle record in exactly one leaf"
    # records hold the right triangles: (geomID, primID) exactly once each, edges as the reference forms them
    ids = set()
    for t in range(nt):
        g, p = int(tu[t, 7]), int(tu[t, 3])
        ids.add((g, p))
        vv, ii = meshes[g]
        a, b, cc = vv[ii[p]]
        assert np.array_equal(tf[t, 0:3], a)
        assert np.array_equal(tf[t, 4:7], (a - b).astype(np.float32))
        assert np.array_equal(tf[t, 8:11], (cc - a).astype(np.float32))
    assert len(ids) == nt


# ------------------------------------------------------ soup: small cases --

def soup_case(rtc, n, seed, h=0.005):
    return rtc.gen_soup(n, seed=seed, h=h)


@pytest.mark.parametrize("quality", [0, 1, 2])
def test_soup_parity_with_edge_rays(rtc, device, oracle, quality):
    v, ix = soup_case(rtc, 20000, 5, h=0.02)
    sc = rtc.Scene(device, [(v, ix)], quality=quality).commit()
    rays = rtc.gen_random_rays(100000, seed=6, lo=-0.05, hi=1.05)
    rays["tfar"][::5] = 0.05
    rays["tnear"][1::9] = 0.02
    rays["tnear"][2::13] = 3.0
    rays["tfar"][2::13] = 1.0          # tnear > tfar: untouched
    rays["dir_x"][3::17] = np.nan      # NaN direction: never hits
    rays["org_z"][4::19] = np.inf
    rays["dir_y"][5::23] = 0.0         # axis-parallel components
    rays["dir_z"][5::23] = 0.0
    out = trace(rtc, sc, rays)
    ref = oracle_trace(oracle, [(v, ix)], rays)
    assert_same(out, ref, f"soup quality {quality}")
    untouched = np.zeros(len(rays), bool)
    untouched[2::13] = True
    assert records_equal(out[untouched], rays[untouched])
    assert (out["geomID"] != INV).mean() > 0.2
    sc.release()


@pytest.mark.parametrize("robust", [False, True])
def test_negative_tnear_parity(rtc, device, oracle, robust):
    """Intervals that start behind the origin (tnear < 0), origins inside
    the soup: hits at negative t count and the closest is the least t. The
    boxes behind the origin have negative entry distances, whose child-order
    keys sort as farthest and whose stack entries keep -inf as their bound
    (never culled by the pop test). Closest and any hit equal the oracle's."""
    v, ix = soup_case(rtc, 20000, 9, h=0.02)
    sc = rtc.Scene(device, [(v, ix)], flags=ROBUST if robust else 0).commit()
    rays = rtc.gen_random_rays(100000, seed=10, lo=0.2, hi=0.8)
    rays["tnear"][0::3] = -0.3
    rays["tnear"][1::3] = -3.0
    rays["tfar"][1::6] = 0.05
    rays["tfar"][2::7] = -0.1          # a wholly negative interval
    out = trace(rtc, sc, rays)
    ref = oracle_trace(oracle, [(v, ix)], rays, robust=robust)
    assert_same(out, ref, f"negative tnear, robust={robust}")
    hit = out["geomID"] != INV
    assert (out["tfar"][hit] < 0).sum() > 1000, "hits behind the origin"
    occ = np.zeros(len(rays), rtc.RTCRay)
    for k in rtc.RTCRay.names:
        occ[k] = rays[k]
    assert_same(trace(rtc, sc, occ, occluded=True),
                oracle_trace(oracle, [(v, ix)], occ, occluded=True, robust=robust),
                f"negative tnear occlusion, robust={robust}")
    sc.release()


@pytest.fixture(scope="module")
def sah_device(rtc):
    d = rtc.Device("builder=sah")
    yield d
    d.release()


@pytest.mark.parametrize("n", [2, 3, 64, 65, 4096, 4097, 70000])
def test_sah_builder_sizes(rtc, sah_device, oracle, n):
    """Binned-SAH tree at the size-class boundaries (one wave per segment up
    to 64 primitives, one workgroup up to 4096, chunked binning above):
    structure and closest/any hits identical to the oracle."""
    v, ix = soup_case(rtc, n, 100 + n, h=0.03 if n > 1000 else 0.2)
    sc = rtc.Scene(sah_device, [(v, ix)]).commit()
    st = sc.stats()
    assert st["num_prims"] == n and st["max_stack"] <= 512
    nodes, tris = sc.read_bvh()
    check_bvh(nodes, tris, [(v, ix)])
    rays = rtc.gen_random_rays(20000, seed=n, lo=-0.05, hi=1.05)
    assert_same(trace(rtc, sc, rays), oracle_trace(oracle, [(v, ix)], rays), f"SAH n={n}")
    occ = rays.copy()
    occ["tfar"] = 0.05
    assert_same(trace(rtc, sc, occ), oracle_trace(oracle, [(v, ix)], occ), f"SAH n={n} short rays")
    sc.release()


def test_sah_builder_degenerate_and_multi_geometry(rtc, sah_device, oracle):
    """Identical centroids force the object-median fallback; duplicated
    triangles across geometries exercise the tie rule on the SAH tree."""
    v = np.array([[0.1, 0.1, 0.1], [0.9, 0.1, 0.1], [0.1, 0.9, 0.5]], np.float32)
    same = np.tile(np.array([[0, 1, 2]], np.uint32), (5000, 1))
    rays = rtc.gen_random_rays(20000, seed=21)
    sc = rtc.Scene(sah_device, [(v, same)]).commit()
    assert_same(trace(rtc, sc, rays), oracle_trace(oracle, [(v, same)], rays, brute=True), "SAH identical triangles")
    sc.release()
    v1, i1 = soup_case(rtc, 30000, 22, h=0.03)
    v2, i2 = rtc.gen_sphere(64, (0.5, 0.5, 0.5), 0.3)
    meshes = [(v1, i1), (v2, i2), (v1, i1[:3000])]
    sc = rtc.Scene(sah_device, meshes).commit()
    nodes, tris = sc.read_bvh()
    check_bvh(nodes, tris, meshes)
    assert_same(trace(rtc, sc, rays), oracle_trace(oracle, meshes, rays), "SAH multi geometry")
    sc.release()


def test_soup_occlusion_parity(rtc, device, oracle):
    v, ix = soup_case(rtc, 20000, 7, h=0.02)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    rh = rtc.gen_random_rays(100000, seed=8, tfar=0.1)
    rays = np.zeros(len(rh), rtc.RTCRay)
    for k in rtc.RTCRay.names:
        rays[k] = rh[k]
    out = trace(rtc, sc, rays, occluded=True)
    ref = oracle_trace(oracle, [(v, ix)], rays, occluded=True)
    assert_same(out, ref, "soup occlusion")
    assert 0.05 < np.isneginf(out["tfar"]).mean() < 0.95
    sc.release()


@pytest.mark.parametrize("config", ["dual=0", "dual=1", "refill_min=64,refill_min_occ=1,blocks_per_cu=2"])
def test_traversal_variants_parity(rtc, oracle, config):
    # every persistent-kernel form (with and without the triangle cursor,
    # extreme refill thresholds, a small grid) returns the oracle's records
    dev = rtc.Device(config)
    v, ix = soup_case(rtc, 20000, 17, h=0.02)
    sc = rtc.Scene(dev, [(v, ix)]).commit()
    rh = rtc.gen_random_rays(60000, seed=18)
    assert_same(trace(rtc, sc, rh), oracle_trace(oracle, [(v, ix)], rh), f"closest hit {config}")
    rays = np.zeros(len(rh), rtc.RTCRay)
    for k in rtc.RTCRay.names:
        rays[k] = rh[k]
    rays["tfar"] = 0.1
    assert_same(trace(rtc, sc, rays, occluded=True), oracle_trace(oracle, [(v, ix)], rays, occluded=True),
                f"occlusion {config}")
    sc.release()
    dev.release()


def test_builds_are_deterministic(rtc):
    # the same scene built on fresh devices gives the same tree (node count
    # and SAH statistic), for both builders
    v, ix = soup_case(rtc, 200000, 19, h=0.01)
    for cfg in ("builder=sah", "builder=morton"):
        stats = set()
        for _ in range(4):
            dev = rtc.Device(cfg)
            sc = rtc.Scene(dev, [(v, ix)]).commit()
            st = sc.stats()
            stats.add((st["num_nodes"], st["sah"], st["max_depth"]))
            sc.release()
            dev.release()
        assert len(stats) == 1, (cfg, stats)


def test_brute_force_small(rtc, device, oracle):
    v, ix = soup_case(rtc, 3000, 9, h=0.05)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    rays = rtc.gen_random_rays(5000, seed=10)
    assert_same(trace(rtc, sc, rays), oracle_trace(oracle, [(v, ix)], rays, brute=True), "brute force")
    sc.release()


def test_multi_geometry_ties_and_invalid(rtc, device, oracle):
    rng = np.random.default_rng(3)
    v1, i1 = soup_case(rtc, 4000, 11, h=0.03)
    v2, i2 = rtc.gen_sphere(32, (0.5, 0.5, 0.5), 0.3)
    # duplicates of some triangles of mesh 1 (exact ties across primIDs and geomIDs)
    i3 = np.concatenate([i1[:500], i1[:500]])
    # invalid triangles: index out of range, NaN vertex, huge vertex
    v4 = np.concatenate([v1[:30], np.array([[np.nan, 0, 0], [3e18, 0, 0]], np.float32)])
    i4 = np.array([[0, 1, 2], [0, 1, 99], [30, 1, 2], [31, 1, 2], [3, 4, 5]], np.uint32)
    meshes = [(v1, i1), (v2, i2), (v1, i3), (v4, i4)]
    sc = rtc.Scene(device, meshes).commit()
    assert sc.stats()["num_prims"] == len(i1) + len(i2) + len(i3) + 2
    rays = rtc.gen_random_rays(50000, seed=12)
    out = trace(rtc, sc, rays)
    ref = oracle_trace(oracle, meshes, rays)
    assert_same(out, ref, "multi geometry")
    assert set(np.unique(out["geomID"])) >= {0, 1, INV}
    # a disabled / detached geometry disappears after the next commit
    g1 = rtc.lib().rtcGetGeometry(sc.h, 1)
    rtc.lib().rtcDisableGeometry(g1)
    rtc.lib().rtcDetachGeometry(sc.h, 3)
    sc.commit()
    out2 = trace(rtc, sc, rays)
    ref2 = oracle_trace(oracle, [(v1, i1), (v2, i2[:0]), (v1, i3)], rays)
    assert_same(out2, ref2, "after disable/detach")
    sc.release()


def test_empty_and_degenerate_scenes(rtc, device, oracle):
    rays = rtc.gen_random_rays(1000, seed=13)
    empty = rtc.Scene(device, []).commit()
    assert records_equal(trace(rtc, empty, rays), rays)
    b = empty.bounds()
    assert b.lower_x == np.inf and b.upper_x == -np.inf
    empty.release()
    # one triangle, all triangles identical, zero-area triangles
    v = np.array([[0.1, 0.1, 0.1], [0.9, 0.1, 0.1], [0.1, 0.9, 0.5]], np.float32)
    for ix in (np.array([[0, 1, 2]], np.uint32), np.tile(np.array([[0, 1, 2]], np.uint32), (100, 1)),
               np.array([[0, 0, 0], [1, 1, 2], [0, 1, 2]], np.uint32)):
        sc = rtc.Scene(device, [(v, ix)]).commit()
        assert_same(trace(rtc, sc, rays), oracle_trace(oracle, [(v, ix)], rays, brute=True), f"degenerate {len(ix)}")
        sc.release()


def test_strided_and_shared_buffers(rtc, device, oracle):
    v, ix = soup_case(rtc, 5000, 14, h=0.03)
    sc = rtc.Scene(device, [(v, ix)], shared=True).commit()
    rays = rtc.gen_random_rays(20000, seed=15)
    ref = oracle_trace(oracle, [(v, ix)], rays)
    padded = np.dtype({"names": list(rays.dtype.names) + ["pad"],
                       "formats": [rays.dtype.fields[k][0] for k in rays.dtype.names] + ["<u4"],
                       "offsets": [rays.dtype.fields[k][1] for k in rays.dtype.names] + [80], "itemsize": 96})
    pr = np.zeros(len(rays), padded)
    for k in rays.dtype.names:
        pr[k] = rays[k]
    pr["pad"] = 0xDEADBEEF
    sc.intersect1M(pr, stride=96)
    got = np.zeros(len(rays), rays.dtype)
    for k in rays.dtype.names:
        got[k] = pr[k]
    assert_same(got, ref, "stride 96")
    assert (pr["pad"] == 0xDEADBEEF).all()
    sc.release()


def test_error_paths(rtc, device):
    L = rtc.lib()
    assert not L.rtcNewGeometry(device.h, rtc.RTC_GEOMETRY_TYPE_QUAD)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_OPERATION
    g = L.rtcNewGeometry(device.h, rtc.RTC_GEOMETRY_TYPE_TRIANGLE)
    buf = np.zeros(30, np.float32)
    L.rtcSetSharedGeometryBuffer(g, rtc.RTC_BUFFER_TYPE_VERTEX, 0, rtc.RTC_FORMAT_FLOAT4, buf.ctypes.data, 0, 16, 2)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_OPERATION
    L.rtcSetSharedGeometryBuffer(g, rtc.RTC_BUFFER_TYPE_VERTEX, 0, rtc.RTC_FORMAT_FLOAT3, buf.ctypes.data + 2, 0, 12, 2)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_ARGUMENT
    L.rtcReleaseGeometry(g)
    sc = L.rtcNewScene(device.h)
    r = rtc.new_rayhits(4)
    L.rtcIntersect1M(sc, None, r.ctypes.data, 4, 80)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_OPERATION  # not committed
    # error callback sees the code
    seen = []
    cb = rtc.ERROR_FUNC(lambda p, code, msg: seen.append(code))
    L.rtcSetDeviceErrorFunction(device.h, cb, None)
    L.rtcSetSceneBuildQuality(sc, 7)
    assert seen == [rtc.RTC_ERROR_INVALID_ARGUMENT]
    L.rtcSetDeviceErrorFunction(device.h, rtc.ERROR_FUNC(0), None)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_ARGUMENT
    # progress monitor cancellation and memory monitor refusal
    v, ix = soup_case(rtc, 100, 1)
    s2 = rtc.Scene(device, [(v, ix)])
    pm = rtc.PROGRESS_FUNC(lambda p, n: False)
    L.rtcSetSceneProgressMonitorFunction(s2.h, pm, None)
    L.rtcCommitScene(s2.h)
    assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_CANCELLED
    s2.release()
    d2 = rtc.Device("")
    mm = rtc.MEMORY_MONITOR_FUNC(lambda p, b, post: False)
    L.rtcSetDeviceMemoryMonitorFunction(d2.h, mm, None)
    s3 = rtc.Scene(d2, [(v, ix)])
    L.rtcCommitScene(s3.h)
    assert L.rtcGetDeviceError(d2.h) == rtc.RTC_ERROR_OUT_OF_MEMORY
    s3.release()
    d2.release()
    L.rtcReleaseScene(sc)


# -------------------------------------------------- C2: full-size soup --

@pytest.mark.slow
def test_c2_full_size_parity(rtc, device, oracle):
    """2^20-triangle soup, 2^24 incoherent rays: closest hit and the tfar=0.1
    occlusion pass, every record identical to the oracle's."""
    v, ix = rtc.gen_soup(1 << 20, seed=1, h=0.005)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    rays = rtc.gen_random_rays(1 << 24, seed=2)
    bvh = oracle.BVH([(v, ix)])
    out = trace(rtc, sc, rays)
    ref = rays.copy()
    bvh.intersect(ref)
    assert_same(out, ref, "C2 closest hit")
    hitrate = (out["geomID"] != INV).mean()
    del out
    # the binned-SAH tree gives the same records
    d2 = rtc.Device("builder=sah")
    sc2 = rtc.Scene(d2, [(v, ix)]).commit()
    assert_same(trace(rtc, sc2, rays), ref, "C2 closest hit, SAH tree")
    sc2.release()
    d2.release()
    del ref
    occ = np.zeros(len(rays), rtc.RTCRay)
    for k in rtc.RTCRay.names:
        occ[k] = rays[k]
    occ["tfar"] = 0.1
    del rays
    o_gpu = trace(rtc, sc, occ, occluded=True)
    bvh.occluded(occ)
    assert_same(o_gpu, occ, "C2 occlusion")
    assert 0.5 < hitrate < 1.0
    sc.release()


# ------------------------------------------ robust scenes (Pluecker test) --

ROBUST = 1 << 2  # RTC_SCENE_FLAG_ROBUST


def test_robust_single_triangle(rtc, device, oracle):
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    ix = np.array([[0, 1, 2]], np.uint32)
    sc = rtc.Scene(device, [(v, ix)], flags=ROBUST).commit()
    r = rtc.new_rayhits(2)
    r["org_x"], r["org_y"], r["org_z"], r["dir_z"] = 0.25, 0.25, -1.0, 1.0
    r["tnear"][1] = 1.0  # t == tnear is inside for the Pluecker test
    out = trace(rtc, sc, r)
    assert (out["geomID"] == 0).all() and (out["tfar"] == 1.0).all()
    assert (out["u"] == 0.25).all() and (out["v"] == 0.25).all() and (out["Ng_z"] == 1.0).all()
    assert_same(out, oracle_trace(oracle, [(v, ix)], r, brute=True, robust=True), "robust single triangle")
    sc.release()


@pytest.mark.parametrize("builder", ["morton", "sah"])
def test_robust_soup_parity(rtc, oracle, builder):
    d = rtc.Device(f"builder={builder}")
    v, ix = soup_case(rtc, 20000, 41, h=0.02)
    sc = rtc.Scene(d, [(v, ix)], flags=ROBUST).commit()
    rays = rtc.gen_random_rays(100000, seed=42, lo=-0.05, hi=1.05)
    rays["tfar"][::5] = 0.05
    rays["tnear"][1::9] = 0.02
    rays["dir_x"][3::17] = np.nan
    rays["dir_y"][5::23] = 0.0
    rays["dir_z"][5::23] = 0.0
    out = trace(rtc, sc, rays)
    assert_same(out, oracle_trace(oracle, [(v, ix)], rays, robust=True), f"robust soup {builder}")
    occ = np.zeros(len(rays), rtc.RTCRay)
    for k in rtc.RTCRay.names:
        occ[k] = rays[k]
    occ["tfar"] = np.minimum(occ["tfar"], 0.1)
    assert_same(trace(rtc, sc, occ, occluded=True), oracle_trace(oracle, [(v, ix)], occ, occluded=True, robust=True),
                f"robust occlusion {builder}")
    sc.release()
    d.release()


def test_robust_watertight_c1_sphere(rtc, device, oracle):
    """Rays from the centre of the closed C1 sphere through every edge
    midpoint: none misses, and every record equals the oracle's."""
    v, ix = rtc.gen_sphere(128)
    e = np.concatenate([ix[:, [0, 1]], ix[:, [1, 2]], ix[:, [2, 0]]])
    mids = (0.5 * (v[e[:, 0]] + v[e[:, 1]])).astype(np.float32)
    r = rtc.new_rayhits(len(mids))
    r["org_x"] = r["org_y"] = r["org_z"] = 0.0
    r["dir_x"], r["dir_y"], r["dir_z"] = mids[:, 0], mids[:, 1], mids[:, 2]
    sc = rtc.Scene(device, [(v, ix)], flags=ROBUST).commit()
    out = trace(rtc, sc, r)
    assert (out["geomID"] == 0).all()
    assert_same(out, oracle_trace(oracle, [(v, ix)], r, robust=True), "robust watertight")
    sc.release()


# ------------------------------------------------------------ batched scenes --

def _batch_case(rtc):
    """Displaced spheres of uneven sizes (one empty scene, one of one
    triangle, one above the 4096-reference chunk size) with rays from a
    radius-2 sphere towards the unit ball."""
    phis = [2, 3, 5, 8, 13, 21, 40, 70, 4, 6]
    scenes = []
    for i, p in enumerate(phis):
        scenes.append(rtc.gen_displaced_sphere(p, seed=50 + i, amplitude=0.1))
    v0, ix0 = scenes[0]
    scenes.insert(3, (v0, ix0[:0]))       # empty
    scenes.insert(5, (v0, ix0[:1]))       # one triangle
    rays = [rtc.gen_sphere_rays(500 + 700 * (i % 4), seed=80 + i) for i in range(len(scenes))]
    return scenes, rays


def test_batched_commit_trace_and_graph(rtc, device, oracle):
    meshes, rays = _batch_case(rtc)
    scs = [rtc.Scene(device, [m] if len(m[1]) else []) for m in meshes]
    rtc.commit_scenes(scs)
    refs = [oracle_trace(oracle, [m], r, brute=len(m[1]) < 200) if len(m[1]) else r.copy()
            for m, r in zip(meshes, rays)]
    # each scene on its own through the standard entry point
    for i, (sc, r) in enumerate(zip(scs, rays)):
        assert_same(trace(rtc, sc, r), refs[i], f"batched scene {i}")
    # one persistent launch over all streams, and the same captured in a graph
    ptrs = [device.malloc(r.nbytes) for r in rays]
    try:
        def upload():
            for p_, r in zip(ptrs, rays):
                device.memcpy(p_, r.ctypes.data, r.nbytes)

        def download():
            outs = []
            for p_, r in zip(ptrs, rays):
                o = np.zeros_like(r)
                device.memcpy(o.ctypes.data, p_, r.nbytes)
                outs.append(o)
            return outs

        items = rtc.batch_items([(sc, p_, len(r), 80) for sc, p_, r in zip(scs, ptrs, rays)])
        upload()
        rtc.lib().rtcgTraceBatch(items.ctypes.data, len(items), 0)
        device.check("rtcgTraceBatch")
        for i, o in enumerate(download()):
            assert_same(o, refs[i], f"batched launch stream {i}")
        g = rtc.lib().rtcgNewBatchGraph(items.ctypes.data, len(items), 0)
        device.check("rtcgNewBatchGraph")
        try:
            for rep in range(3):
                upload()
                rtc.lib().rtcgLaunchBatchGraph(g, None)
                device.check("rtcgLaunchBatchGraph")
                for i, o in enumerate(download()):
                    assert_same(o, refs[i], f"graph replay {rep} stream {i}")
        finally:
            rtc.lib().rtcgReleaseBatchGraph(g)
        # occlusion streams (RTCRay records)
        occ = []
        for r in rays:
            o = np.zeros(len(r), rtc.RTCRay)
            for k in rtc.RTCRay.names:
                o[k] = r[k]
            o["tfar"] = 2.0
            occ.append(o)
        optrs = [device.malloc(o.nbytes) for o in occ]
        try:
            for p_, o in zip(optrs, occ):
                device.memcpy(p_, o.ctypes.data, o.nbytes)
            oitems = rtc.batch_items([(sc, p_, len(o), 48) for sc, p_, o in zip(scs, optrs, occ)])
            rtc.lib().rtcgTraceBatch(oitems.ctypes.data, len(oitems), 1)
            device.check("occluded batch")
            for i, (p_, o, m) in enumerate(zip(optrs, occ, meshes)):
                got = np.zeros_like(o)
                device.memcpy(got.ctypes.data, p_, o.nbytes)
                ref = oracle_trace(oracle, [m], o, occluded=True, brute=True) if len(m[1]) else o.copy()
                assert_same(got, ref, f"occluded stream {i}")
        finally:
            for p_ in optrs:
                device.free(p_)
    finally:
        for p_ in ptrs:
            device.free(p_)
    # a scene committed again on its own leaves the batch
    scs[7].commit()
    assert_same(trace(rtc, scs[7], rays[7]), refs[7], "recommitted scene")
    assert_same(trace(rtc, scs[8], rays[8]), refs[8], "batch after a member left")
    for sc in scs:
        sc.release()


# ------------------------------------------------------------- autotuner --

def test_bvh_width4_structure_and_parity(rtc, oracle):
    d = rtc.Device("bvh_width=4,builder=sah")
    v, ix = soup_case(rtc, 20000, 61, h=0.02)
    sc = rtc.Scene(d, [(v, ix)]).commit()
    nodes, tris = sc.read_bvh()
    check_bvh(nodes, tris, [(v, ix)])
    meta = nodes.reshape(-1, 80)[:, 24:32]
    assert ((meta != 0).sum(1) <= 4).all()
    rays = rtc.gen_random_rays(50000, seed=62)
    assert_same(trace(rtc, sc, rays), oracle_trace(oracle, [(v, ix)], rays), "width 4")
    sc.release()
    d.release()


def test_autotune_keeps_results(rtc, oracle):
    d = rtc.Device("")
    v, ix = soup_case(rtc, 20000, 63, h=0.02)
    sc = rtc.Scene(d, [(v, ix)]).commit()
    rays = rtc.gen_random_rays(100000, seed=64)
    dp = d.malloc(rays.nbytes)
    try:
        d.memcpy(dp, rays.ctypes.data, rays.nbytes)
        buf = ctypes.create_string_buffer(256)
        mr = rtc.lib().rtcgAutotune(sc.h, dp, len(rays), 80, 0, 1, buf, 256)
        d.check("rtcgAutotune")
        cfg = dict(kv.split("=") for kv in buf.value.decode().split(","))
        assert mr > 0 and set(cfg) >= {"bvh_width", "max_leaf", "cost_tri", "trace_stack", "blocks_per_cu"}
        back = np.zeros_like(rays)
        d.memcpy(back.ctypes.data, dp, rays.nbytes)
        assert np.array_equal(back.view(np.uint8), rays.view(np.uint8))  # sample rays untouched
        # the tuned device (and a fresh one made from the returned string) give the oracle's records
        ref = oracle_trace(oracle, [(v, ix)], rays)
        assert_same(trace(rtc, sc, rays), ref, "after autotune")
        d2 = rtc.Device(buf.value.decode())
        sc2 = rtc.Scene(d2, [(v, ix)]).commit()
        assert_same(trace(rtc, sc2, rays), ref, "tuned config string")
        sc2.release()
        d2.release()
    finally:
        d.free(dp)
    sc.release()
    d.release()


# ------------------------------------------------ packet and SoA entry points --

def _soa(rays, K, fields):
    """AoS records -> SoA packets of K lanes (float32/uint32 bit patterns)."""
    P = len(rays) // K
    out = np.zeros((P, len(fields), K), np.uint32)
    for f, name in enumerate(fields):
        out[:, f, :] = rays[name][:P * K].view(np.uint32).reshape(P, K)
    return out


def _aos_from_soa(soa, K, template, fields):
    r = template.copy()
    P = soa.shape[0]
    for f, name in enumerate(fields):
        r[name][:P * K] = soa[:, f, :].reshape(-1).view(r.dtype[name])
    return r


def test_packet_and_soa_entry_points(rtc, device, oracle):
    """Packet (4/8/16), NM, Np and 1Mp forms against the oracle. A packet
    lane is active only where its valid word is -1, as the reference reads
    the mask (kernels/bvh/bvh_intersector_hybrid.cpp:131): lanes holding 0 or
    1 are left untouched."""
    v, ix = soup_case(rtc, 5000, 71, h=0.03)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    L = rtc.lib()
    ctx = rtc.init_context()
    base = rtc.gen_random_rays(1600, seed=72)
    base["tnear"][7::29] = 5.0  # tnear > tfar: skipped
    base["tfar"][7::29] = 1.0
    ref = oracle_trace(oracle, [(v, ix)], base)
    fields = list(rtc.RTCRayHit.names)
    rng = np.random.default_rng(73)
    for K, fn in ((4, L.rtcIntersect4), (8, L.rtcIntersect8), (16, L.rtcIntersect16)):
        soa = _soa(base, K, fields)
        u = rng.random(soa.shape[0] * K)
        valid = np.where(u < 0.7, -1, np.where(u < 0.85, 0, 1)).astype(np.int32).reshape(-1, K)
        for p in range(soa.shape[0]):
            pk = np.ascontiguousarray(soa[p])
            fn(valid[p].ctypes.data, sc.h, ctypes.byref(ctx), pk.ctypes.data)
            soa[p] = pk
        device.check(f"rtcIntersect{K}")
        got = _aos_from_soa(soa, K, base, fields)
        on = valid.reshape(-1) == -1
        assert (valid == 1).any()
        n = soa.shape[0] * K
        assert records_equal(got[:n][on], ref[:n][on]), K
        assert records_equal(got[:n][~on], base[:n][~on]), K
    # occlusion packets
    occ = base.copy()
    occ["tfar"] = np.minimum(occ["tfar"], 0.2)
    oref = np.zeros(len(occ), rtc.RTCRay)
    for k in rtc.RTCRay.names:
        oref[k] = occ[k]
    oref = oracle_trace(oracle, [(v, ix)], oref, occluded=True)
    soa = _soa(occ, 8, list(rtc.RTCRay.names))
    valid = np.full(8, -1, np.int32)
    for p in range(soa.shape[0]):
        pk = np.ascontiguousarray(soa[p])
        L.rtcOccluded8(valid.ctypes.data, sc.h, ctypes.byref(ctx), pk.ctypes.data)
        soa[p] = pk
    got = _aos_from_soa(soa, 8, occ, list(rtc.RTCRay.names))
    assert np.array_equal(got["tfar"][:len(oref)].view(np.uint32), oref["tfar"].view(np.uint32))
    # NM: runtime N = 5, packets 100 bytes apart beyond their 20*5*4 = 400 B
    N, M = 5, 300
    soa = _soa(base, N, fields)[:M]
    stride = 20 * N * 4 + 100
    buf = np.zeros(M * stride, np.uint8)
    for p in range(M):
        buf[p * stride:p * stride + 400] = soa[p].reshape(-1).view(np.uint8)
    L.rtcIntersectNM(sc.h, ctypes.byref(ctx), buf.ctypes.data, N, M, stride)
    device.check("rtcIntersectNM")
    for p in range(M):
        soa[p] = buf[p * stride:p * stride + 400].view(np.uint32).reshape(20, N)
    assert records_equal(_aos_from_soa(soa, N, base, fields)[:N * M], ref[:N * M])
    # Np: one array per field
    arrs = [np.ascontiguousarray(base[name]) for name in fields]
    ptrs = (ctypes.c_void_p * 20)(*[a.ctypes.data for a in arrs])
    L.rtcIntersectNp(sc.h, ctypes.byref(ctx), ptrs, len(base))
    device.check("rtcIntersectNp")
    got = base.copy()
    for name, a in zip(fields, arrs):
        got[name] = a
    assert records_equal(got, ref)
    # 1Mp: pointers to records
    recs = [base[i:i + 1].copy() for i in range(0, 400)]
    pp = (ctypes.c_void_p * len(recs))(*[r.ctypes.data for r in recs])
    L.rtcIntersect1Mp(sc.h, ctypes.byref(ctx), pp, len(recs))
    device.check("rtcIntersect1Mp")
    assert records_equal(np.concatenate(recs), ref[:400])
    sc.release()
