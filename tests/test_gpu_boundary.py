"""The drop-in boundary under the ways a reference caller uses it: the
complete C program of INTEGRATION.md compiled against the headers and run,
rtcIntersect1M called concurrently from 8 host threads of one process (the
reference's rtcIntersect1M is reentrant and its benchmark calls it from every
worker, kernels/common/rtcore.cpp:595-623, tutorials/verify/verify.cpp:4656-4683),
the tri_builder / builder device keys (kernels/common/state.cpp:309-310,
kernels/bvh/bvh8_factory.cpp:865-878), and a batch graph whose scenes were
committed again. Every ray record is compared byte for byte with the oracle."""
import ctypes
import os
import re
import subprocess
import threading

import numpy as np
import pytest

from conftest import ROOT, mismatches

pytestmark = pytest.mark.gpu
INV = 0xFFFFFFFF


def _same(a, b, what):
    bad = mismatches(a, b)
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(a)} records differ, first {bad[:5]}"


def _oracle(oracle, meshes, rays, occluded=False):
    r = rays.copy()
    bvh = oracle.BVH(meshes)
    (bvh.occluded if occluded else bvh.intersect)(r)
    return r


def test_integration_c_program(rtc, oracle, tmp_path):
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```c example-app\n(.*?)```", text, re.S)
    assert m, "INTEGRATION.md has no example-app block"
    src = tmp_path / "app.c"
    src.write_text(m.group(1))
    libdir = os.path.dirname(rtc.LIB_PATH)
    exe = tmp_path / "app"
    subprocess.run(["gcc", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir,
                    "-l:libembree3.so", f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    v, ix = rtc.gen_soup(100000, seed=301, h=0.01)
    rays = rtc.gen_random_rays(300000, seed=302)
    rays["tfar"][::13] = 0.02
    rays["tnear"][5::17] = 3.0  # tnear > tfar: left untouched
    paths = [tmp_path / n for n in ("v.f32", "i.u32", "rays.bin", "out.bin")]
    v.astype(np.float32).tofile(paths[0])
    ix.astype(np.uint32).tofile(paths[1])
    rays.tofile(paths[2])
    subprocess.run([str(exe)] + [str(p) for p in paths], check=True, timeout=120)
    got = np.fromfile(paths[3], dtype=rtc.RTCRayHit)
    assert len(got) == len(rays)
    _same(got, _oracle(oracle, [(v, ix)], rays), "INTEGRATION.md C program")
    assert (got["geomID"] != INV).mean() > 0.3


@pytest.mark.parametrize("where", ["host", "device"])
def test_concurrent_rtcIntersect1M_from_8_threads(rtc, device, oracle, where):
    v, ix = rtc.gen_soup(200000, seed=311, h=0.01)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    n, T = 1 << 20, 8
    rays = rtc.gen_random_rays(n, seed=312)
    rays["tfar"][1::9] = 0.05
    ref = _oracle(oracle, [(v, ix)], rays)
    L = rtc.lib()
    ctx = rtc.init_context()
    bounds = np.linspace(0, n, T + 1).astype(np.int64)
    work = rays.copy()
    d = device.malloc(rays.nbytes) if where == "device" else None
    if d:
        device.memcpy(d, rays.ctypes.data, rays.nbytes)
    errors = []
    start = threading.Barrier(T)

    def worker(k):
        try:
            lo, hi = int(bounds[k]), int(bounds[k + 1])
            base = (d if d else work.ctypes.data) + 80 * lo
            start.wait()
            for _ in range(3):  # the same slice again: each call reads what the previous wrote
                L.rtcIntersect1M(sc.h, ctypes.byref(ctx), base, hi - lo, 80)
                code = L.rtcGetDeviceError(device.h)  # per-thread error slot
                if code:
                    errors.append((k, code))
        except Exception as e:  # noqa: BLE001
            errors.append((k, repr(e)))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(T)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    if d:
        device.memcpy(work.ctypes.data, d, rays.nbytes)
        device.free(d)
    _same(work, ref, f"8 concurrent threads, {where} records")
    sc.release()


def test_host_records_in_chunks(rtc, device, oracle):
    """Host records pass through pinned chunks of 2^20 records: a query of
    2^21 + 12,345 records (two full chunks and a partial one), contiguous and
    with a 96 B stride, gives the same bytes as the same query on device
    records (closest hit and occlusion); a sample equals the oracle."""
    v, ix = rtc.gen_soup(200000, seed=331, h=0.01)
    sc = rtc.Scene(device, [(v, ix)]).commit()
    L = rtc.lib()
    ctx = rtc.init_context()
    n = (1 << 21) + 12345
    rays = rtc.gen_random_rays(n, seed=332)
    rays["tfar"][1::9] = 0.05
    rays["tnear"][2::13] = 3.0
    rays["tfar"][2::13] = 1.0  # tnear > tfar: untouched
    host = rays.copy()
    L.rtcIntersect1M(sc.h, ctypes.byref(ctx), host.ctypes.data, n, 80)
    assert L.rtcGetDeviceError(device.h) == 0
    d = device.malloc(rays.nbytes)
    try:
        device.memcpy(d, rays.ctypes.data, rays.nbytes)
        L.rtcIntersect1M(sc.h, ctypes.byref(ctx), d, n, 80)
        assert L.rtcGetDeviceError(device.h) == 0
        dev_out = np.empty_like(rays)
        device.memcpy(dev_out.ctypes.data, d, rays.nbytes)
    finally:
        device.free(d)
    assert host.tobytes() == dev_out.tobytes(), "host and device records differ"
    # strided host records: the padding stays as it was
    padded = np.dtype({"names": list(rays.dtype.names) + ["pad"],
                       "formats": [rays.dtype.fields[k][0] for k in rays.dtype.names] + ["<u4"],
                       "offsets": [rays.dtype.fields[k][1] for k in rays.dtype.names] + [80], "itemsize": 96})
    pr = np.zeros(n, padded)
    for k in rays.dtype.names:
        pr[k] = rays[k]
    pr["pad"] = 0xDEADBEEF
    L.rtcIntersect1M(sc.h, ctypes.byref(ctx), pr.ctypes.data, n, 96)
    assert L.rtcGetDeviceError(device.h) == 0
    got = np.zeros(n, rays.dtype)
    for k in rays.dtype.names:
        got[k] = pr[k]
    assert got.tobytes() == host.tobytes(), "strided host records differ"
    assert (pr["pad"] == 0xDEADBEEF).all()
    # occlusion records
    occ = np.zeros(n, rtc.RTCRay)
    for k in rtc.RTCRay.names:
        occ[k] = rays[k]
    ho = occ.copy()
    L.rtcOccluded1M(sc.h, ctypes.byref(ctx), ho.ctypes.data, n, 48)
    assert L.rtcGetDeviceError(device.h) == 0
    d = device.malloc(occ.nbytes)
    try:
        device.memcpy(d, occ.ctypes.data, occ.nbytes)
        L.rtcOccluded1M(sc.h, ctypes.byref(ctx), d, n, 48)
        do = np.empty_like(occ)
        device.memcpy(do.ctypes.data, d, occ.nbytes)
    finally:
        device.free(d)
    assert ho.tobytes() == do.tobytes(), "host and device occlusion records differ"
    pick = np.sort(np.random.default_rng(333).choice(n, 20000, replace=False))
    _same(host[pick], _oracle(oracle, [(v, ix)], rays[pick]), "host records in chunks")
    sc.release()


def _build_kind(rtc, cfg, quality=None):
    d = rtc.Device(cfg)
    v, ix = rtc.gen_soup(30000, seed=321, h=0.01)
    kw = {} if quality is None else {"quality": quality}
    sc = rtc.Scene(d, [(v, ix)], **kw).commit()
    st = sc.stats()
    sc.release()
    d.release()
    return st["num_nodes"], round(st["sah"], 6)


def test_tri_builder_keys(rtc, device, oracle):
    morton = _build_kind(rtc, "builder=morton")
    sah = _build_kind(rtc, "builder=sah")
    assert morton != sah
    # tri_builder is the reference's name for the same key
    assert _build_kind(rtc, "tri_builder=morton") == morton
    assert _build_kind(rtc, "tri_builder=sah") == sah
    # default: the quality chooses (LOW: Morton); a named builder applies at every quality
    assert _build_kind(rtc, "", quality=rtc.RTC_BUILD_QUALITY_LOW) == morton
    assert _build_kind(rtc, "", quality=rtc.RTC_BUILD_QUALITY_MEDIUM) == sah
    assert _build_kind(rtc, "tri_builder=sah", quality=rtc.RTC_BUILD_QUALITY_LOW) == sah
    # an unknown builder name fails the commit, as the reference's factory does
    L = rtc.lib()
    d = rtc.Device("tri_builder=bogus")
    v, ix = rtc.gen_soup(1000, seed=322)
    sc = rtc.Scene(d, [(v, ix)])
    L.rtcCommitScene(sc.h)
    assert L.rtcGetDeviceError(d.h) == rtc.RTC_ERROR_INVALID_ARGUMENT
    assert b"unknown builder bogus" in L.rtcgGetErrorString()
    r = rtc.new_rayhits(4)
    L.rtcIntersect1M(sc.h, None, r.ctypes.data, 4, 80)
    assert L.rtcGetDeviceError(d.h) == rtc.RTC_ERROR_INVALID_OPERATION  # never committed
    sc.release()
    d.release()


def test_batch_graph_refused_after_recommit(rtc, device, oracle):
    meshes = [rtc.gen_displaced_sphere(p, seed=330 + p, amplitude=0.1) for p in (6, 10, 14)]
    scs = [rtc.Scene(device, [m]) for m in meshes]
    rtc.commit_scenes(scs)
    rays = [rtc.gen_sphere_rays(3000, seed=340 + i) for i in range(3)]
    L = rtc.lib()
    ptrs = [device.malloc(r.nbytes) for r in rays]
    try:
        for p_, r in zip(ptrs, rays):
            device.memcpy(p_, r.ctypes.data, r.nbytes)
        items = rtc.batch_items([(sc, p_, len(r), 80) for sc, p_, r in zip(scs, ptrs, rays)])
        g = L.rtcgNewBatchGraph(items.ctypes.data, len(items), 0)
        device.check("rtcgNewBatchGraph")
        try:
            L.rtcgLaunchBatchGraph(g, None)
            device.check("first launch")
            for i, (p_, r, m) in enumerate(zip(ptrs, rays, meshes)):
                got = np.zeros_like(r)
                device.memcpy(got.ctypes.data, p_, r.nbytes)
                _same(got, _oracle(oracle, [m], r), f"graph stream {i}")
            scs[1].commit()  # leaves the batch: new arrays of its own
            L.rtcgLaunchBatchGraph(g, None)
            assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_OPERATION
            # a new batch commit of all three moves every scene to new arrays
            rtc.commit_scenes(scs)
            L.rtcgLaunchBatchGraph(g, None)
            assert L.rtcGetDeviceError(device.h) == rtc.RTC_ERROR_INVALID_OPERATION
        finally:
            L.rtcgReleaseBatchGraph(g)
        # a graph made now works
        g = L.rtcgNewBatchGraph(items.ctypes.data, len(items), 0)
        for p_, r in zip(ptrs, rays):
            device.memcpy(p_, r.ctypes.data, r.nbytes)
        L.rtcgLaunchBatchGraph(g, None)
        device.check("new graph")
        L.rtcgReleaseBatchGraph(g)
        for i, (p_, r, m) in enumerate(zip(ptrs, rays, meshes)):
            got = np.zeros_like(r)
            device.memcpy(got.ctypes.data, p_, r.nbytes)
            _same(got, _oracle(oracle, [m], r), f"new graph stream {i}")
    finally:
        for p_ in ptrs:
            device.free(p_)
    for sc in scs:
        sc.release()
