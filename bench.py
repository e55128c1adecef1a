"""Benchmark: incoherent closest-hit ray queries on one MI355X.

Workload (random triangle soup): 2^20 FP32 triangles with centroids uniform
in [0,1]^3 and vertex offsets uniform in [-0.005,0.005]^3; 2^24 rays with
origins uniform in the cube and directions uniform on the sphere; closest hit
with tnear=0, tfar=inf, plus an occlusion pass with tfar=0.1. Rays are
RTCRayHit records resident in HBM; each step restores them from a pristine
device copy (untimed) and traces them with rtcIntersect1M semantics; the
trace kernel is timed with HIP events on the stream it runs on.

Prints one JSON line (see the keys below). `--gpus N` with N > 1 exits: the
build and the benchmark are single-GPU.
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "embree-aarch64_amd"))
sys.path.insert(0, ROOT)

import rtcore as R  # noqa: E402

HBM_PEAK_GBS = 8000.0
NUM_TRIS = 1 << 20
NUM_RAYS = 1 << 24


def log(msg):
    print(msg, file=sys.stderr, flush=True)


CALIBRATION = os.path.join(ROOT, "tests", "golden", "cpu_calibration.json")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_threads():
    """Host threads this process may use: its CPU affinity, capped by
    OMP_NUM_THREADS where the host sets one (a GPU box's per-GPU CPU share)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = os.environ.get("OMP_NUM_THREADS", "")
    return max(1, min(n, int(omp))) if omp.isdigit() and int(omp) > 0 else max(1, n)


def cpu_baseline(v, ix, rays, sample):
    """The CPU restatement (oracle/) on the host cores: BVH build over the full
    soup (parallel subtree builds) and closest-hit traversal of a bounded ray
    sample on every thread this process may use. The restatement is scalar C,
    not the reference's SIMD code: `r` (from tests/golden/cpu_calibration.json)
    is its single-thread rate over the compiled reference's on the same
    workload and host model, so value / r estimates the reference here."""
    from oracle import oracle as O
    threads = cpu_threads()
    t0 = time.perf_counter()
    bvh = O.BVH([(v, ix)], nthreads=threads)
    build_s = time.perf_counter() - t0
    r = rays[:sample].copy()
    t0 = time.perf_counter()
    bvh.intersect(r, nthreads=threads)
    trace_s = time.perf_counter() - t0
    value = sample / trace_s / 1e6
    out = {"value": round(value, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
           "sample": f"{sample} of the {len(rays)} rays, closest hit, oracle BVH8 built over all {len(ix)} triangles",
           "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
           "build_mtris_per_s": round(len(ix) / build_s / 1e6, 3), "build_threads": threads}
    try:
        cal = json.load(open(CALIBRATION))
        out["calibration"] = {"r": cal["r"], "measured_on": cal["host_cpu"],
                              "oracle_single_thread_mrays_per_s": cal["oracle_mrays_per_s_single_thread"],
                              "reference_single_thread_mrays_per_s": cal["reference_mrays_per_s_single_thread"],
                              "workload": cal["workload"]}
        out["reference_estimate_mrays_per_s"] = round(value / cal["r"], 3)
    except (OSError, ValueError, KeyError):
        out["calibration"] = None
    return out


PROFILE = os.path.join(ROOT, "profiles", "trace_build_c2_soup1M_16Mrays.txt")
TRACE_KERNEL = "k_trace_p<false, true, false, 12, false, false>"


def profiled_traffic():
    """HBM-side bytes per launch of the closest-hit kernel (FETCH_SIZE and
    WRITE_SIZE scaled by a plain copy of known size), its L2 hit rate and
    what limits it (SQ counters: instructions per ray, active lanes per
    VALU instruction), from the committed rocprofv3 summary of this
    workload."""
    traffic = hit = limiter = None
    try:
        for line in open(PROFILE):
            if TRACE_KERNEL in line and "GB read" in line:
                tail = line.split("->")[1]
                rd = float(tail.split("GB read")[0])
                wr = float(tail.split("GB read")[1].split("GB written")[0])
                hit = float(tail.split("L2 hit")[1])
                traffic = (rd + wr) * 1e9
            if line.strip().startswith("VALU instructions per ray"):
                f = line.replace(",", " ").replace(";", " ").split()
                limiter = {"kind": "VALU issue (SQ counters: VALU busy, divergent lanes)",
                           "valu_instr_per_ray": float(f[4]), "salu_instr_per_ray": float(f[6]),
                           "lanes_active_per_valu_instr": float(f[f.index("instruction") + 1]),
                           "source": os.path.relpath(PROFILE, ROOT)}
    except (OSError, ValueError, IndexError):
        pass
    return traffic, hit, limiter


def build_bytes_per_tri(n):
    """Algorithmic HBM bytes per triangle of the binned-SAH build: every
    stage reads its inputs and writes its outputs once.
      primitive references: 3 indices + 3 vertices in, box + key out   84
      Morton codes: box in, key + id out                                40
      radix sort: 4 passes of 8 B key + id in and out                   64
      boxes gathered in sorted order: id + box in, box out              68
      copies into the level buffers: id + box in and out                72
      each big level (segments above 4096): id + box read to bin and
        again to flag, flags written and scanned, id + box + flag +
        scan read and id + box + value + segment id written            176
      each mid or small level (segments of 17..4096): box read to bin
        and to flag, id + box read and id + box + value written        140
      tiny segments (2..16): id + box read, value + links written        56
      boxes in final order                                              68
      bottom-up boxes: box read and written per inner node              64
      collapse: a 128 B node slot per ~5.9 triangles                    22
      leaf records: position written and read, value, primitive
        reference, 3 indices and 3 vertices read, 48 B record written  140
    Big levels: ceil(log2(n / 4096)); mid and small levels: log2(4096 / 16) = 8."""
    big = max(0, int(np.ceil(np.log2(max(1.0, n / 4096.0)))))
    return 84 + 40 + 64 + 68 + 72 + 176 * big + 140 * 8 + 56 + 68 + 64 + 22 + 140, big


def time_trace(L, dev, sc, d_rays, d_src, n, steps):
    ms = np.zeros(max(1, steps), np.float32)
    L.rtcgTimeTrace(sc.h, d_rays, d_src, n, 80, 0, len(ms), ms.ctypes.data)
    dev.check("rtcgTimeTrace")
    return float(ms.mean())


def counted(L, dev, sc, d_rays, d_src, n, nbytes):
    d_cn, d_ct = dev.malloc(4 * n), dev.malloc(4 * n)
    dev.memcpy(d_rays, d_src, nbytes)
    L.rtcgTraceCounted(sc.h, d_rays, n, 80, 0, d_cn, d_ct)
    dev.check("rtcgTraceCounted")
    cn, ct = np.zeros(n, np.uint32), np.zeros(n, np.uint32)
    dev.memcpy(cn.ctypes.data, d_cn, 4 * n)
    dev.memcpy(ct.ctypes.data, d_ct, 4 * n)
    out = np.zeros(n, R.RTCRayHit)
    dev.memcpy(out.ctypes.data, d_rays, nbytes)
    dev.free(d_cn)
    dev.free(d_ct)
    return float(cn.mean()), float(ct.mean()), out


def learned_vs_sah(grid, width, height, bounces, train_steps, steps):
    """Heightfield terrain (grid x grid vertices, Perlin-noise fBm of 6
    octaves, amplitude 0.2) seen by a width x height pinhole camera, followed
    by cosine-hemisphere bounces (depth `bounces` in all). Queries of every
    depth are traced on the plain binned-SAH tree; the split model is trained
    on the node visits of the even-indexed queries, and the learned tree is
    timed against the plain tree on the odd-indexed queries only (held out:
    never seen in training)."""
    L = R.lib()
    t0 = time.perf_counter()
    v, ix = R.gen_heightfield(grid, seed=11, amplitude=0.2, octaves=6)
    cam = R.gen_camera(width, height, (0.5, -0.35, 0.5), (0.5, 0.55, 0.0), up=(0.0, 0.0, 1.0), fov=50.0)
    dev = R.Device("builder=sah")
    plain = R.Scene(dev, [(v, ix)]).commit()
    build_plain = plain.stats()["build_ms"]
    queries, cur = [cam], cam
    for b in range(bounces - 1):
        out = cur.copy()
        plain.intersect1M(out)
        cur = R.gen_bounce(out, seed=100 + b)
        queries.append(cur)
    Q = np.concatenate(queries)
    train = np.ascontiguousarray(Q[::2])
    held = np.ascontiguousarray(Q[1::2])
    n, nbytes = len(held), held.nbytes
    log(f"learned-vs-SAH: {len(ix)} tris, {len(Q)} queries ({[len(q) for q in queries]}), {len(train)} for "
        f"training, {n} held out and timed, setup {time.perf_counter() - t0:.1f}s")
    d_src, d_rays = dev.malloc(nbytes), dev.malloc(nbytes)
    dev.memcpy(d_src, held.ctypes.data, nbytes)
    d_train = dev.malloc(train.nbytes)
    dev.memcpy(d_train, train.ctypes.data, train.nbytes)
    model = R.SplitModel(dev, seed=1)
    t1 = time.perf_counter()
    model.collect(plain, d_train, len(train))
    loss = model.train(train_steps, batch=65536, lr=3e-3, seed=1)
    train_s = time.perf_counter() - t1
    dev.free(d_train)
    learned = R.Scene(dev, [(v, ix)])
    L.rtcgSetSceneSplitModel(learned.h, model.h)
    learned.commit()
    build_learned = learned.stats()["build_ms"]
    res = {}
    outs = {}
    # the two trees are timed alternately, three times each, so that clock
    # drift over the run reaches both alike; each keeps its median
    times = {"plain": [], "learned": []}
    for name, sc in (("plain", plain), ("learned", learned)):
        time_trace(L, dev, sc, d_rays, d_src, n, 2)
    for _ in range(3):
        for name, sc in (("plain", plain), ("learned", learned)):
            times[name].append(time_trace(L, dev, sc, d_rays, d_src, n, steps))
    for name, sc in (("plain", plain), ("learned", learned)):
        nb, tb, outs[name] = counted(L, dev, sc, d_rays, d_src, n, nbytes)
        ms = float(np.median(times[name]))
        st = sc.stats()
        res[name] = {"mrays_per_s": round(n / (ms / 1e3) / 1e6, 2), "ms": round(ms, 3), "nodes_per_ray": round(nb, 2),
                     "tris_per_ray": round(tb, 2), "bvh_nodes": st["num_nodes"], "sah": round(st["sah"], 3)}
    same = bool(np.array_equal(outs["plain"].view(np.uint8), outs["learned"].view(np.uint8)))
    for p_ in (d_src, d_rays):
        dev.free(p_)
    model.release()
    plain.release()
    learned.release()
    dev.release()
    return {"workload": f"heightfield {grid}x{grid} vertices ({len(ix)} tris), {width}x{height} camera + "
                        f"{bounces - 1} cosine bounces = {len(Q)} closest-hit queries; model trained on the "
                        f"{len(train)} even-indexed, both trees timed on the {n} odd-indexed (held out)",
            "plain_sah": res["plain"], "learned": res["learned"],
            "speedup": round(res["learned"]["mrays_per_s"] / res["plain"]["mrays_per_s"], 4),
            "hits_identical": same,
            "train_s": round(train_s, 2), "train_loss": round(loss, 4),
            "build_ms_plain": round(build_plain, 2), "build_ms_learned": round(build_learned, 2)}


def learned_point_source(steps, width=0.3):
    """A ray distribution far from the uniform one the SAH assumes: 2^22 rays
    from the point (0.5, 0.5, -2) into the central cube `width` wide of the
    2^20-triangle soup. The split model (200 Adam steps, SAH slack 0.5) is
    trained on the even-indexed rays; both trees are timed and counted on
    the odd-indexed ones (held out)."""
    L = R.lib()
    v, ix = R.gen_soup(NUM_TRIS, seed=1, h=0.005)
    rng = np.random.default_rng(7)
    nq = 1 << 22
    tgt = (0.5 - width / 2) + width * rng.random((nq, 3), dtype=np.float32)
    org = np.array([0.5, 0.5, -2.0], np.float32)
    d = tgt - org
    Q = R.new_rayhits(nq)
    Q["org_x"], Q["org_y"], Q["org_z"] = org
    Q["dir_x"], Q["dir_y"], Q["dir_z"] = d.T
    train, held = np.ascontiguousarray(Q[::2]), np.ascontiguousarray(Q[1::2])
    dev = R.Device("builder=sah")
    plain = R.Scene(dev, [(v, ix)]).commit()
    n, nbytes = len(held), held.nbytes
    d_src, d_rays, d_train = dev.malloc(nbytes), dev.malloc(nbytes), dev.malloc(train.nbytes)
    dev.memcpy(d_src, held.ctypes.data, nbytes)
    dev.memcpy(d_train, train.ctypes.data, train.nbytes)
    model = R.SplitModel(dev, seed=1)
    model.collect(plain, d_train, len(train))
    loss = model.train(200, batch=65536, lr=3e-3, seed=1)
    dev.free(d_train)
    L.rtcgSetSplitModelSahSlack(model.h, ctypes.c_float(0.5))
    learned = R.Scene(dev, [(v, ix)])
    L.rtcgSetSceneSplitModel(learned.h, model.h)
    learned.commit()
    times = {"plain": [], "learned": []}
    for name, sc in (("plain", plain), ("learned", learned)):
        time_trace(L, dev, sc, d_rays, d_src, n, 2)
    for _ in range(3):
        for name, sc in (("plain", plain), ("learned", learned)):
            times[name].append(time_trace(L, dev, sc, d_rays, d_src, n, steps))
    res, outs = {}, {}
    for name, sc in (("plain", plain), ("learned", learned)):
        nb, tb, outs[name] = counted(L, dev, sc, d_rays, d_src, n, nbytes)
        ms = float(np.median(times[name]))
        res[name] = {"mrays_per_s": round(n / (ms / 1e3) / 1e6, 2), "nodes_per_ray": round(nb, 3),
                     "tris_per_ray": round(tb, 3), "sah": round(sc.stats()["sah"], 3)}
    same = bool(np.array_equal(outs["plain"].view(np.uint8), outs["learned"].view(np.uint8)))
    for p_ in (d_src, d_rays):
        dev.free(p_)
    model.release()
    plain.release()
    learned.release()
    dev.release()
    return {"workload": f"2^20-triangle soup, 2^22 rays from (0.5, 0.5, -2) into the central {width}-wide cube; "
                        "trained on the even-indexed, timed and counted on the odd-indexed (held out); 200 steps, "
                        "SAH slack 0.5",
            "plain_sah": res["plain"], "learned": res["learned"],
            "node_ratio": round(res["learned"]["nodes_per_ray"] / res["plain"]["nodes_per_ray"], 4),
            "speedup": round(res["learned"]["mrays_per_s"] / res["plain"]["mrays_per_s"], 4),
            "hits_identical": same, "train_loss": round(loss, 4)}


def batched_scenes(num_scenes, steps):
    """Batched scenes: num_scenes displaced UV spheres (3-octave value noise,
    amplitude 0.1) with triangle counts log-uniform in [1e3, 2e6] and ray
    streams log-uniform in [4e3, 1e6] from a radius-2 sphere towards the unit
    ball. The tree shape (node width, leaf size, collapse triangle cost) and
    the traversal launch (LDS stack entries, refill threshold, resident
    blocks per CU) are autotuned: the tree shape on the scene of median size
    with 2^20 rays drawn like the streams (rtcgAutotune), then one
    rtcgCommitScenes builds them all with it, and the launch settings are
    tuned on the batched launch itself (rtcgAutotuneBatch, on a scratch copy
    of the streams); then the tuned
    settings and one persistent launch, captured in a hipGraph, traces every
    stream."""
    L = R.lib()
    work = R.batch_workload(num_scenes, seed=7)
    dev = R.Device("")
    t0 = time.perf_counter()
    scs, rays = [], []
    for i, (phi, _, nray) in enumerate(work):
        v, ix = R.gen_displaced_sphere(phi, seed=1000 + i, amplitude=0.1)
        scs.append(R.Scene(dev, [(v, ix)]))
        rays.append(R.gen_sphere_rays(nray, seed=2000 + i, radius=2.0))
    gen_s = time.perf_counter() - t0
    # autotune on the median-size scene
    ta = time.perf_counter()
    med = int(np.argsort([t for _, t, _ in work])[len(work) // 2])
    scs[med].commit()
    tune_rays = R.gen_sphere_rays(1 << 20, seed=2999, radius=2.0)
    d_sample = dev.malloc(tune_rays.nbytes)
    dev.memcpy(d_sample, tune_rays.ctypes.data, tune_rays.nbytes)
    tuned = ctypes.create_string_buffer(256)
    tuned_mrays = L.rtcgAutotune(scs[med].h, d_sample, len(tune_rays), 80, 0, 1, tuned, 256)
    dev.check("rtcgAutotune")
    dev.free(d_sample)
    tune_s = time.perf_counter() - ta
    t1 = time.perf_counter()
    R.commit_scenes(scs)
    commit_s = time.perf_counter() - t1
    allr = np.concatenate(rays)
    nbytes, n = allr.nbytes, len(allr)
    d_src, d_rays = dev.malloc(nbytes), dev.malloc(nbytes)
    dev.memcpy(d_src, allr.ctypes.data, nbytes)
    offs = np.concatenate([[0], np.cumsum([len(r) for r in rays])])
    items = R.batch_items([(sc, d_src + 80 * int(offs[i]), len(r), 80) for i, (sc, r) in enumerate(zip(scs, rays))])
    tb = time.perf_counter()
    launch_tuned = ctypes.create_string_buffer(256)
    launch_mrays = L.rtcgAutotuneBatch(items.ctypes.data, len(items), 0, launch_tuned, 256)
    dev.check("rtcgAutotuneBatch")
    tune_batch_s = time.perf_counter() - tb
    items = R.batch_items([(sc, d_rays + 80 * int(offs[i]), len(r), 80) for i, (sc, r) in enumerate(zip(scs, rays))])
    g = L.rtcgNewBatchGraph(items.ctypes.data, len(items), 0)
    dev.check("rtcgNewBatchGraph")
    times = []
    for k in range(steps + 2):
        dev.memcpy(d_rays, d_src, nbytes)
        L.rtcgSynchronize(dev.h)
        ts = time.perf_counter()
        L.rtcgLaunchBatchGraph(g, None)
        te = time.perf_counter()
        dev.check("rtcgLaunchBatchGraph")
        if k >= 2:
            times.append(te - ts)
    out = np.zeros(n, R.RTCRayHit)
    dev.memcpy(out.ctypes.data, d_rays, nbytes)
    hit_rate = float((out["geomID"] != R.RTC_INVALID_GEOMETRY_ID).mean())
    L.rtcgReleaseBatchGraph(g)
    tri_total = sum(t for _, t, _ in work)
    for p_ in (d_src, d_rays):
        dev.free(p_)
    for sc in scs:
        sc.release()
    dev.release()
    ms = 1e3 * float(np.mean(times))
    return {"workload": f"{num_scenes} displaced spheres, {tri_total} triangles, {n} rays in {num_scenes} streams",
            "autotuned": {"config": tuned.value.decode(), "on_scene": med, "scene_tris": int(work[med][1]),
                          "tuning_rays": len(tune_rays), "mrays_per_s_on_scene": round(float(tuned_mrays), 2),
                          "seconds": round(tune_s, 2),
                          "launch_config": launch_tuned.value.decode(),
                          "launch_mrays_per_s": round(float(launch_mrays), 2),
                          "launch_seconds": round(tune_batch_s, 2)},
            "commit_ms_incl_upload": round(1e3 * commit_s, 2),
            "commit_mtris_per_s_incl_upload": round(tri_total / commit_s / 1e6, 2),
            "graph_ms": round(ms, 3), "mrays_per_s": round(n / (ms / 1e3) / 1e6, 2), "hit_rate": round(hit_rate, 4),
            "generation_s": round(gen_s, 1)}


def instanced_scene(num_instances, width, height, steps, check_rays=2048, learn=True):
    """Two-level scene: five base meshes (two displaced spheres of about 1M and
    200k triangles, three heightfield patches of about 1M, 200k and 100k)
    committed as scenes, and num_instances instances of them with a uniform
    random rotation, a uniform scale in [0.5, 2] and a translation uniform in
    [-50, 50]^3. Closest hit, half of the rays: width x height pinhole-camera
    rays from (0, 20, -160) towards the origin, and as many cosine-hemisphere
    bounce rays from the camera rays' hit points (each hit point in turn,
    with a new seed per pass). Occlusion, the other half: one shadow ray per
    closest-hit ray towards the point light (0, 100, 0), from the hit point,
    or for a miss from the ray's point nearest the origin (2^26 rays in all
    at 4096 x 4096). A subsample of each kind is checked against the
    oracle's instance restatement over per-object BVHs (the full-size parity
    test checks 64k of each)."""
    from oracle import oracle as O
    L = R.lib()
    dev = R.Device("")
    t0 = time.perf_counter()
    meshes = R.instance_workload(seed=11)
    objs = [R.Scene(dev, [m]).commit() for m in meshes]
    T = R.random_transforms(num_instances, seed=5, scale=(0.5, 2.0), extent=50.0)
    which = np.random.default_rng(6).integers(0, len(objs), num_instances)
    top = R.Scene(dev)
    for k in range(num_instances):
        top.add_instance(objs[which[k]], T[k])
    setup_s = time.perf_counter() - t0
    t1 = time.perf_counter()
    top.commit()
    commit_ms = 1e3 * (time.perf_counter() - t1)

    def timed(recs, stride, occluded):
        """Mean kernel ms over `steps` launches (after two untimed) and the
        traced records."""
        d_src, d_r = dev.malloc(max(1, recs.nbytes)), dev.malloc(max(1, recs.nbytes))
        dev.memcpy(d_src, recs.ctypes.data, recs.nbytes)
        warm = np.zeros(2, np.float32)
        L.rtcgTimeTrace(top.h, d_r, d_src, len(recs), stride, occluded, 2, warm.ctypes.data)
        ms = np.zeros(steps, np.float32)
        L.rtcgTimeTrace(top.h, d_r, d_src, len(recs), stride, occluded, steps, ms.ctypes.data)
        dev.check("rtcgTimeTrace instanced")
        out = np.zeros(len(recs), recs.dtype)
        dev.memcpy(out.ctypes.data, d_r, recs.nbytes)
        dev.free(d_src)
        dev.free(d_r)
        return float(ms.mean()), out

    cam = R.gen_camera(width, height, org=(0.0, 20.0, -160.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0), fov=45.0)
    c_ms, cam_hits = timed(cam, 80, 0)
    bounce = R.gen_bounce_n(cam_hits, len(cam), seed=31)
    b_ms, bounce_hits = timed(bounce, 80, 0)
    rays = np.concatenate([cam, bounce])
    hits = np.concatenate([cam_hits, bounce_hits])
    n = len(rays)
    hit = hits["geomID"] != R.RTC_INVALID_GEOMETRY_ID
    sh = R.gen_shadow_rays(hits, light=(0.0, 100.0, 0.0), centre=(0.0, 0.0, 0.0))
    ns = len(sh)
    o_ms, shadow = timed(sh, 48, 1)
    # oracle check on a subsample of each kind
    t2 = time.perf_counter()
    I = O.Instances()
    oid = [I.add_object([m]) for m in meshes]
    for k in range(num_instances):
        I.add_instance(oid[which[k]], T[k], k)
    rng = np.random.default_rng(12)
    same = {}
    for name, src, out in (("camera", cam, cam_hits), ("bounce", bounce, bounce_hits)):
        pick = rng.choice(len(src), min(check_rays, len(src)), replace=False)
        ref = src[pick].copy()
        I.intersect_bvh(ref)
        same[name] = bool(np.array_equal(ref.view(np.uint8).reshape(len(ref), -1),
                                         out[pick].view(np.uint8).reshape(len(ref), -1)))
    spick = rng.choice(ns, min(check_rays, ns), replace=False)
    sref = sh[spick].copy()
    I.occluded_bvh(sref)
    same["shadow"] = bool(np.array_equal(sref["tfar"], shadow[spick]["tfar"]))
    check_s = time.perf_counter() - t2
    learned = instanced_learned(dev, top, objs, which, T, rays, sh, steps) if learn else None
    top.release()
    for o in objs:
        o.release()
    dev.release()
    base = [len(m[1]) for m in meshes]
    return {"workload": f"{num_instances} instances of 5 base meshes ({sum(base)} triangles: {base}), "
                        f"{len(cam)} camera rays ({width}x{height}) + {len(bounce)} cosine bounce rays closest hit, "
                        f"{ns} shadow rays",
            "instanced_triangles": int(sum(base[w] for w in which)),
            "commit_top_ms": round(commit_ms, 2), "setup_s": round(setup_s, 1),
            "closest_mrays_per_s": round(n / ((c_ms + b_ms) / 1e3) / 1e6, 2), "closest_ms": round(c_ms + b_ms, 3),
            "camera_mrays_per_s": round(len(cam) / (c_ms / 1e3) / 1e6, 2),
            "bounce_mrays_per_s": round(len(bounce) / (b_ms / 1e3) / 1e6, 2),
            "occlusion_mrays_per_s": round(ns / (o_ms / 1e3) / 1e6, 2),
            "hit_rate": round(float(hit.mean()), 4), "shadowed": round(float((shadow["tfar"] < 0).mean()), 4),
            "oracle_check": {"rays_per_kind": int(check_rays), "camera_identical": same["camera"],
                             "bounce_identical": same["bounce"], "shadow_identical": same["shadow"],
                             "seconds": round(check_s, 1)},
            "learned_both_levels": learned}


def instanced_learned(dev, top, objs, which, T, rays, shadow, steps, train_steps=200):
    """The learned split cost at both levels of the two-level scene: one
    model per base mesh trained on the visits of all its instances, one for
    the top tree over the instances, all from the even-indexed closest-hit
    rays (camera and bounce); the plain and the learned scenes are timed on
    the odd-indexed closest-hit rays and their shadow rays (held out)."""
    L = R.lib()
    train = np.ascontiguousarray(rays[::2])
    held = np.ascontiguousarray(rays[1::2])
    hsh = np.ascontiguousarray(shadow[1::2])
    d_train = dev.malloc(train.nbytes)
    dev.memcpy(d_train, train.ctypes.data, train.nbytes)
    t0 = time.perf_counter()
    models = []
    for o in objs:
        m = R.SplitModel(dev, seed=1)
        m.collect_instanced(top, o, d_train, len(train))
        m.train(train_steps, batch=65536, lr=3e-3, seed=1)
        models.append(m)
    mt = R.SplitModel(dev, seed=1)
    mt.collect_instanced(top, None, d_train, len(train))
    mt.train(train_steps, batch=65536, lr=3e-3, seed=1)
    dev.free(d_train)
    train_s = time.perf_counter() - t0

    def timed(sc, recs, stride, occluded):
        d_src, d_r = dev.malloc(recs.nbytes), dev.malloc(recs.nbytes)
        dev.memcpy(d_src, recs.ctypes.data, recs.nbytes)
        warm = np.zeros(2, np.float32)
        L.rtcgTimeTrace(sc.h, d_r, d_src, len(recs), stride, occluded, 2, warm.ctypes.data)
        ms = np.zeros(steps, np.float32)
        L.rtcgTimeTrace(sc.h, d_r, d_src, len(recs), stride, occluded, steps, ms.ctypes.data)
        dev.check("rtcgTimeTrace learned instanced")
        out = np.zeros(len(recs), recs.dtype)
        dev.memcpy(out.ctypes.data, d_r, recs.nbytes)
        dev.free(d_src)
        dev.free(d_r)
        return len(recs) / (float(np.median(ms)) / 1e3) / 1e6, out

    res = {}
    res["plain"] = [timed(top, held, 80, 0), timed(top, hsh, 48, 1)]
    # the object scenes rebuilt with their models (the plain top tree keeps
    # the trees it was committed with), then a top scene over them with its own
    t1 = time.perf_counter()
    for o, m in zip(objs, models):
        L.rtcgSetSceneSplitModel(o.h, m.h)
        o.commit()
    top2 = R.Scene(dev)
    for k in range(len(which)):
        top2.add_instance(objs[which[k]], T[k])
    L.rtcgSetSceneSplitModel(top2.h, mt.h)
    top2.commit()
    rebuild_s = time.perf_counter() - t1
    res["learned"] = [timed(top2, held, 80, 0), timed(top2, hsh, 48, 1)]
    same = bool(np.array_equal(res["plain"][0][1].view(np.uint8), res["learned"][0][1].view(np.uint8)) and
                np.array_equal(res["plain"][1][1]["tfar"], res["learned"][1][1]["tfar"]))
    top2.release()
    for o in objs:
        L.rtcgSetSceneSplitModel(o.h, None)
    for m in models + [mt]:
        m.release()
    out = {"workload": f"models trained on the {len(train)} even-indexed closest-hit rays (camera and bounce); "
                       f"timed on the {len(held)} odd-indexed ones and their {len(hsh)} shadow rays (held out); "
                       f"{train_steps} steps",
           "hits_identical": same, "train_s": round(train_s, 2), "rebuild_s": round(rebuild_s, 2)}
    for name in ("plain", "learned"):
        out[name] = {"closest_mrays_per_s": round(res[name][0][0], 2), "occlusion_mrays_per_s": round(res[name][1][0], 2)}
    out["closest_speedup"] = round(res["learned"][0][0] / res["plain"][0][0], 4)
    out["occlusion_speedup"] = round(res["learned"][1][0] / res["plain"][1][0], 4)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rays", type=int, default=NUM_RAYS)
    ap.add_argument("--tris", type=int, default=NUM_TRIS)
    ap.add_argument("--cpu-sample", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--config", default="")
    ap.add_argument("--no-learned", action="store_true")
    ap.add_argument("--hf-grid", type=int, default=4096)
    ap.add_argument("--hf-width", type=int, default=1920)
    ap.add_argument("--hf-height", type=int, default=1080)
    ap.add_argument("--hf-depth", type=int, default=4)
    ap.add_argument("--train-steps", type=int, default=400)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--batch-scenes", type=int, default=256)
    ap.add_argument("--no-instances", action="store_true")
    ap.add_argument("--instances", type=int, default=10000)
    ap.add_argument("--inst-width", type=int, default=4096)
    ap.add_argument("--inst-height", type=int, default=4096)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "single-GPU build: run with --gpus 1 (multi-GPU sharding is a design note only)"}))
        sys.exit(2)

    L = R.lib()
    log(f"generating {args.tris} triangles, {args.rays} rays")
    v, ix = R.gen_soup(args.tris, seed=1, h=0.005)
    rays = R.gen_random_rays(args.rays, seed=2)

    dev = R.Device(args.config)
    sc = R.Scene(dev, [(v, ix)]).commit()
    st = sc.stats()
    log(f"built: {st}")

    # build throughput (rebuild of the committed scene, device-resident inputs)
    bms = np.zeros(max(3, args.steps // 4), np.float32)
    L.rtcgTimeBuild(sc.h, len(bms), bms.ctypes.data)
    dev.check("rtcgTimeBuild")
    build_ms = float(np.median(bms))
    # the Morton (LOW quality) builder on the same triangles
    msc = R.Scene(dev, [(v, ix)], quality=R.RTC_BUILD_QUALITY_LOW).commit()
    L.rtcgTimeBuild(msc.h, len(bms), bms.ctypes.data)
    dev.check("rtcgTimeBuild morton")
    build_ms_morton = float(np.median(bms))
    msc.release()

    n = len(rays)
    nbytes = rays.nbytes
    d_src = dev.malloc(nbytes)
    d_rays = dev.malloc(nbytes)
    dev.memcpy(d_src, rays.ctypes.data, nbytes)

    # per-ray traversal counters for the algorithmic byte count (untimed)
    d_cn = dev.malloc(4 * n)
    d_ct = dev.malloc(4 * n)
    dev.memcpy(d_rays, d_src, nbytes)
    L.rtcgTraceCounted(sc.h, d_rays, n, 80, 0, d_cn, d_ct)
    dev.check("rtcgTraceCounted")
    cn = np.zeros(n, np.uint32)
    ct = np.zeros(n, np.uint32)
    dev.memcpy(cn.ctypes.data, d_cn, 4 * n)
    dev.memcpy(ct.ctypes.data, d_ct, 4 * n)
    hits_out = np.zeros(n, R.RTCRayHit)
    dev.memcpy(hits_out.ctypes.data, d_rays, nbytes)
    hit_rate = float((hits_out["geomID"] != R.RTC_INVALID_GEOMETRY_ID).mean())
    del hits_out

    # closest hit: warmup + timed steps
    warm = np.zeros(max(1, args.warmup), np.float32)
    L.rtcgTimeTrace(sc.h, d_rays, d_src, n, 80, 0, len(warm), warm.ctypes.data)
    L.rtcgSynchronize(dev.h)
    ms = np.zeros(args.steps, np.float32)
    t0 = time.perf_counter()
    L.rtcgTimeTrace(sc.h, d_rays, d_src, n, 80, 0, args.steps, ms.ctypes.data)
    L.rtcgSynchronize(dev.h)
    wall = time.perf_counter() - t0
    dev.check("rtcgTimeTrace")
    ms_step = float(ms.mean())
    mrays = n / (ms_step / 1e3) / 1e6

    # HIGH quality (spatial splits) on the same triangles and rays
    hsc = R.Scene(dev, [(v, ix)], quality=R.RTC_BUILD_QUALITY_HIGH).commit()
    hst = hsc.stats()
    hbms = np.zeros(max(3, args.steps // 4), np.float32)
    L.rtcgTimeBuild(hsc.h, len(hbms), hbms.ctypes.data)
    dev.check("rtcgTimeBuild high")
    hms = time_trace(L, dev, hsc, d_rays, d_src, n, max(3, args.steps // 2))
    high = {"triangle_records": int(hst["num_prims"]), "sah": round(hst["sah"], 3),
            "build_ms": round(float(np.median(hbms)), 3),
            "build_mtris_per_s": round(st["num_prims"] / (float(np.median(hbms)) / 1e3) / 1e6, 2),
            "mrays_per_s": round(n / (hms / 1e3) / 1e6, 2)}
    hsc.release()

    # the drop-in caller's path: host records through rtcIntersect1M
    # (upload, trace, download; PCIe-inclusive, never the headline value)
    host_rays = rays.copy()
    ctx = R.init_context()
    L.rtcIntersect1M(sc.h, ctypes.byref(ctx), host_rays.ctypes.data, n, 80)
    dev.check("rtcIntersect1M host records")
    host_rays[:] = rays
    t0 = time.perf_counter()
    L.rtcIntersect1M(sc.h, ctypes.byref(ctx), host_rays.ctypes.data, n, 80)
    host_s = time.perf_counter() - t0
    dev.check("rtcIntersect1M host records")
    del host_rays

    # occlusion pass, tfar = 0.1 (RTCRay records)
    occ = np.zeros(n, R.RTCRay)
    for k in R.RTCRay.names:
        occ[k] = rays[k]
    occ["tfar"] = 0.1
    d_osrc = dev.malloc(occ.nbytes)
    dev.memcpy(d_osrc, occ.ctypes.data, occ.nbytes)
    ow = np.zeros(2, np.float32)
    L.rtcgTimeTrace(sc.h, d_rays, d_osrc, n, 48, 1, 2, ow.ctypes.data)
    oms = np.zeros(max(3, args.steps // 2), np.float32)
    L.rtcgTimeTrace(sc.h, d_rays, d_osrc, n, 48, 1, len(oms), oms.ctypes.data)
    dev.check("occlusion")
    occ_mrays = n / (float(oms.mean()) / 1e3) / 1e6
    del occ

    # roofline of the closest-hit kernel: algorithmic bytes per launch
    node_b, tri_b = st["node_bytes"], st["tri_bytes"]
    nbar, tbar = float(cn.mean()), float(ct.mean())
    bytes_per_ray = 48 + 36 + nbar * node_b + tbar * tri_b
    achieved = bytes_per_ray * n / (ms_step / 1e3) / 1e9
    traffic, l2hit, limiter = profiled_traffic()
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": round(traffic) if traffic and n == NUM_RAYS and args.tris == NUM_TRIS else None,
                "traffic_source": os.path.relpath(PROFILE, ROOT) if traffic else None,
                "l2_hit_rate": l2hit,
                "bytes_per_ray": round(bytes_per_ray, 1), "nodes_per_ray": round(nbar, 2),
                "tris_per_ray": round(tbar, 2),
                "node_bytes_per_ray": round(nbar * node_b, 1),
                "limiter": limiter}

    b_tri, big_levels = build_bytes_per_tri(st["num_prims"])
    b_ach = b_tri * st["num_prims"] / (build_ms / 1e3) / 1e9
    build_roofline = {"bound": "hbm", "achieved": round(b_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(b_ach / HBM_PEAK_GBS, 4), "bytes_per_tri": b_tri, "big_levels": big_levels,
                      "build_ms": round(build_ms, 3)}

    for p in (d_src, d_rays, d_cn, d_ct, d_osrc):
        dev.free(p)

    sc.release()
    dev.release()

    learned = learned_ps = learned_beam = None
    if not args.no_learned:
        learned = learned_vs_sah(args.hf_grid, args.hf_width, args.hf_height, args.hf_depth, args.train_steps,
                                 max(3, args.steps // 2))
        log("learned split cost, point source")
        learned_ps = learned_point_source(max(3, args.steps // 2))
        log("learned split cost, narrow beam")
        learned_beam = learned_point_source(max(3, args.steps // 2), width=0.05)

    batch = None
    if not args.no_batch:
        batch = batched_scenes(args.batch_scenes, max(3, args.steps // 4))

    inst = None
    if not args.no_instances:
        log("instanced scene")
        inst = instanced_scene(args.instances, args.inst_width, args.inst_height, max(3, args.steps // 4),
                               learn=not args.no_learned)

    cpu = None
    if not args.no_cpu:
        log("timing the CPU baseline")
        cpu = cpu_baseline(v, ix, rays, min(args.cpu_sample, n))

    out = {
        "metric": "Mrays/s closest-hit (incoherent) + build Mtris/s; learned-cost vs plain SAH",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "random triangle soup, 2^20 tris in [0,1]^3 (offsets +-0.005), 2^24 incoherent "
                               "rays, closest hit tnear=0 tfar=inf; occlusion pass tfar=0.1",
                   "tris": int(st["num_prims"]), "rays": n, "builder": "binned SAH + SAH-optimal 8-wide collapse",
                   "bvh": "8-wide quantised nodes 80 B in 128 B slots, triangles 48 B"},
        "build_mtris_per_s": round(st["num_prims"] / (build_ms / 1e3) / 1e6, 2),
        "build_ms": round(build_ms, 3),
        "build_mtris_per_s_morton": round(st["num_prims"] / (build_ms_morton / 1e3) / 1e6, 2),
        "build_roofline": build_roofline,
        "high_quality_spatial_splits": high,
        "host_records_mrays_per_s_incl_pcie": round(n / host_s / 1e6, 2),
        "host_records_ms": round(1e3 * host_s, 2),
        "occlusion_mrays_per_s": round(occ_mrays, 2),
        "hit_rate": round(hit_rate, 4),
        "sah": round(st["sah"], 3),
        "timed_wall_s_incl_restore": round(wall, 4),
        "roofline": roofline,
        "cpu_baseline": cpu,
        "learned_vs_sah": learned,
        "learned_point_source": learned_ps,
        "learned_narrow_beam": learned_beam,
        "batched_scenes": batch,
        "instanced": inst,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
