"""Full-size parity for the batched (256 scenes), learned-cost heightfield
(4096^2 vertices) and two-level instanced (10,000 instances) workloads: the
GPU traces every ray of the workload at its real size, and a seeded
subsample of the records (at least 64k queries per pass, or every scene's
stream in 16 scenes) is compared byte for byte with the oracle's records for
the same input rays. The reference's stream entry points are the parity
boundary (kernels/common/rtcore.cpp:595-623, :848-874)."""
import ctypes
import sys

import numpy as np
import pytest

from conftest import mismatches

pytestmark = [pytest.mark.gpu, pytest.mark.slow]
INV = 0xFFFFFFFF


def say(msg):
    print(msg, file=sys.stderr, flush=True)


def _same(a, b, what):
    bad = mismatches(a, b)
    assert len(bad) == 0, f"{what}: {len(bad)} of {len(a)} records differ, first {bad[:5]}"


def _trace_device(rtc, device, scene, rays, occluded=False):
    """rtcIntersect1M / rtcOccluded1M over records in device memory."""
    L = rtc.lib()
    d = device.malloc(rays.nbytes)
    try:
        device.memcpy(d, rays.ctypes.data, rays.nbytes)
        ctx = rtc.init_context()
        f = L.rtcOccluded1M if occluded else L.rtcIntersect1M
        f(scene.h, ctypes.byref(ctx), d, len(rays), rays.itemsize)
        device.check("trace")
        out = np.empty_like(rays)
        device.memcpy(out.ctypes.data, d, rays.nbytes)
        return out
    finally:
        device.free(d)


def test_c5_instanced_full_size(rtc, device, oracle):
    """10,000 instances of five base meshes (2.5 M triangles, 5 G instanced);
    closest hit: 4096 x 4096 camera rays and 2^24 cosine bounce rays from
    their hit points; one shadow ray per closest-hit ray (2^25) towards the
    point light (0, 100, 0): 2^26 rays; 65,536 of each kind checked."""
    meshes = rtc.instance_workload(seed=11)
    objs = [rtc.Scene(device, [m]).commit() for m in meshes]
    T = rtc.random_transforms(10000, seed=5, scale=(0.5, 2.0), extent=50.0)
    which = np.random.default_rng(6).integers(0, len(objs), 10000)
    top = rtc.Scene(device)
    for k in range(10000):
        assert top.add_instance(objs[which[k]], T[k]) == k
    top.commit()
    cam = rtc.gen_camera(4096, 4096, org=(0.0, 20.0, -160.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0)
    assert len(cam) == 1 << 24
    cam_hits = _trace_device(rtc, device, top, cam)
    bounce = rtc.gen_bounce_n(cam_hits, 1 << 24, seed=31)
    bounce_hits = _trace_device(rtc, device, top, bounce)
    hits = np.concatenate([cam_hits, bounce_hits])
    sh = rtc.gen_shadow_rays(hits, light=(0.0, 100.0, 0.0), centre=(0.0, 0.0, 0.0))
    assert len(sh) == 1 << 25
    shadow = _trace_device(rtc, device, top, sh, occluded=True)
    hit_rate, shadowed = (hits["geomID"] != INV).mean(), (shadow["tfar"] < 0).mean()
    say(f"C5: hit rate {hit_rate:.4f} (camera {(cam_hits['geomID'] != INV).mean():.4f}), shadowed {shadowed:.4f}")
    assert 0.2 < hit_rate < 0.9 and 0.05 < shadowed < 0.95
    I = oracle.Instances()
    oid = [I.add_object([m]) for m in meshes]
    for k in range(10000):
        I.add_instance(oid[which[k]], T[k], k)
    rng = np.random.default_rng(501)
    for name, src, out in (("camera", cam, cam_hits), ("bounce", bounce, bounce_hits)):
        pick = np.sort(rng.choice(len(src), 65536, replace=False))
        ref = src[pick].copy()
        I.intersect_bvh(ref)
        _same(out[pick], ref, f"C5 {name} rays")
    spick = np.sort(rng.choice(len(sh), 65536, replace=False))
    sref = sh[spick].copy()
    I.occluded_bvh(sref)
    _same(shadow[spick], sref, "C5 shadow rays")
    top.release()
    for o in objs:
        o.release()


def test_c4_heightfield_full_size_plain_and_learned(rtc, oracle):
    """4096 x 4096-vertex heightfield (33,538,050 triangles, Perlin fBm of 6
    octaves, amplitude 0.2), 1920 x 1080 camera rays and three cosine
    bounces. The plain binned-SAH tree and a tree built with a split model
    trained on the plain tree's visits of the even-indexed queries give
    byte-identical records for every query; 65,536 queries (16,384 of each
    depth) equal the oracle's."""
    v, ix = rtc.gen_heightfield(4096, seed=11, amplitude=0.2, octaves=6)
    assert len(ix) == 33538050
    dev = rtc.Device("builder=sah")
    plain = rtc.Scene(dev, [(v, ix)]).commit()
    cam = rtc.gen_camera(1920, 1080, (0.5, -0.35, 0.5), (0.5, 0.55, 0.0), up=(0.0, 0.0, 1.0), fov=50.0)
    depth, cur = [cam], cam
    outs = []
    for b in range(4):
        o = _trace_device(rtc, dev, plain, cur)
        outs.append(o)
        if b < 3:
            cur = rtc.gen_bounce(o, seed=100 + b)
            depth.append(cur)
    Q = np.concatenate(depth)
    plain_out = np.concatenate(outs)
    say(f"C4: {len(Q)} queries by depth {[len(q) for q in depth]}")
    train = np.ascontiguousarray(Q[::2])
    d_train = dev.malloc(train.nbytes)
    dev.memcpy(d_train, train.ctypes.data, train.nbytes)
    model = rtc.SplitModel(dev, seed=1)
    model.collect(plain, d_train, len(train))
    dev.free(d_train)
    model.train(400, batch=65536, lr=3e-3, seed=1)
    learned = rtc.Scene(dev, [(v, ix)])
    rtc.lib().rtcgSetSceneSplitModel(learned.h, model.h)
    learned.commit()
    assert learned.stats()["num_nodes"] != plain.stats()["num_nodes"] or \
        learned.stats()["sah"] != plain.stats()["sah"], "the model did not change the tree"
    _same(_trace_device(rtc, dev, learned, Q), plain_out, "C4 learned tree vs plain tree")
    bvh = oracle.BVH([(v, ix)])
    rng = np.random.default_rng(502)
    offs = np.concatenate([[0], np.cumsum([len(q) for q in depth])])
    pick = np.concatenate([offs[k] + np.sort(rng.choice(len(depth[k]), min(16384, len(depth[k])), replace=False))
                           for k in range(4)])
    ref = Q[pick].copy()
    bvh.intersect(ref)
    _same(plain_out[pick], ref, "C4 queries")
    model.release()
    learned.release()
    plain.release()
    dev.release()


def test_c3_batched_full_size(rtc, oracle):
    """256 displaced spheres (triangles log-uniform in [1e3, 2e6], 66.7 M in
    all) and 256 ray streams (log-uniform in [4e3, 1e6], 46.7 M rays): the
    tree shape autotuned on the median scene, one rtcgCommitScenes, the
    launch settings autotuned on the batch (which leaves the records as they
    were), one persistent launch captured in a hipGraph and replayed twice
    (identical records); every stream of 16 scenes, the
    smallest and the largest among them, checked against the oracle (at most
    32,768 rays per stream)."""
    work = rtc.batch_workload(256, seed=7)
    dev = rtc.Device("")
    L = rtc.lib()
    meshes, scs, rays = [], [], []
    for i, (phi, _, nray) in enumerate(work):
        m = rtc.gen_displaced_sphere(phi, seed=1000 + i, amplitude=0.1)
        meshes.append(m)
        scs.append(rtc.Scene(dev, [m]))
        rays.append(rtc.gen_sphere_rays(nray, seed=2000 + i, radius=2.0))
    tris = np.array([len(m[1]) for m in meshes])
    med = int(np.argsort(tris)[len(tris) // 2])
    scs[med].commit()
    d_s = dev.malloc(rays[med].nbytes)
    dev.memcpy(d_s, rays[med].ctypes.data, rays[med].nbytes)
    buf = ctypes.create_string_buffer(256)
    assert L.rtcgAutotune(scs[med].h, d_s, len(rays[med]), 80, 0, 1, buf, 256) > 0
    dev.check("rtcgAutotune")
    dev.free(d_s)
    say(f"C3: {tris.sum()} triangles, {sum(len(r) for r in rays)} rays, tuned {buf.value.decode()}")
    rtc.commit_scenes(scs)
    allr = np.concatenate(rays)
    offs = np.concatenate([[0], np.cumsum([len(r) for r in rays])])
    d = dev.malloc(allr.nbytes)
    try:
        items = rtc.batch_items([(sc, d + 80 * int(offs[i]), len(r), 80) for i, (sc, r) in enumerate(zip(scs, rays))])
        dev.memcpy(d, allr.ctypes.data, allr.nbytes)
        assert L.rtcgAutotuneBatch(items.ctypes.data, len(items), 0, buf, 256) > 0
        dev.check("rtcgAutotuneBatch")
        say(f"C3 launch tuned {buf.value.decode()}")
        o = np.empty_like(allr)
        dev.memcpy(o.ctypes.data, d, allr.nbytes)
        assert o.tobytes() == allr.tobytes(), "rtcgAutotuneBatch changed the records"
        g = L.rtcgNewBatchGraph(items.ctypes.data, len(items), 0)
        dev.check("rtcgNewBatchGraph")
        got = []
        for rep in range(2):
            dev.memcpy(d, allr.ctypes.data, allr.nbytes)
            L.rtcgLaunchBatchGraph(g, None)
            dev.check("rtcgLaunchBatchGraph")
            o = np.empty_like(allr)
            dev.memcpy(o.ctypes.data, d, allr.nbytes)
            got.append(o)
        L.rtcgReleaseBatchGraph(g)
    finally:
        dev.free(d)
    _same(got[1], got[0], "C3 second graph replay")
    order = np.argsort(tris)
    rng = np.random.default_rng(503)
    check = [int(order[0]), int(order[-1])] + [int(x) for x in rng.choice(order[1:-1], 14, replace=False)]
    for s in check:
        out = got[0][offs[s]:offs[s + 1]]
        sel = np.arange(len(out)) if len(out) <= 32768 else np.sort(rng.choice(len(out), 32768, replace=False))
        ref = rays[s][sel].copy()
        oracle.BVH([meshes[s]]).intersect(ref)
        _same(out[sel], ref, f"C3 scene {s} ({tris[s]} triangles)")
    for sc in scs:
        sc.release()
    dev.release()
