"""Writes cpu_calibration.json: the oracle's single-thread closest-hit rate
on the UV-sphere pinhole workload (256x128-segment sphere, 1024^2 rays from
(0,0,3), 45 degree FOV), timed on this host, beside the compiled
reference's single-thread rate on the same workload recorded on this host
model (9.3 Mrays/s: one rtcIntersect1M call over all rays, 0.1125 s for
1,048,576 rays, Intel Xeon @ 2.70 GHz, AVX2 build). The ratio r = oracle /
reference lets a benchmark on another host state what the reference would
reach there: reference estimate = oracle rate there / r.

Run from the repository root: python tests/golden/cpu_calibration.py"""
import json
import os
import platform
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "embree-aarch64_amd"))

REFERENCE_MRAYS_SINGLE_THREAD = 9.3  # 1048576 rays / 0.1125 s, recorded with the compiled reference


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def c1_workload():
    import rtcore as R
    v, ix = R.gen_sphere(128)
    rays = R.gen_pinhole(1024, 1024)
    return v, ix, rays


def time_oracle_c1(v, ix, rays, reps=5):
    from oracle import oracle as O
    bvh = O.BVH([(v, ix)])
    times = []
    for _ in range(reps + 1):
        r = rays.copy()
        t0 = time.perf_counter()
        bvh.intersect(r, nthreads=1)
        times.append(time.perf_counter() - t0)
    hits = int((r["geomID"] != 0xFFFFFFFF).sum())
    return len(rays) / float(np.median(times[1:])) / 1e6, hits


def main():
    v, ix, rays = c1_workload()
    rate, hits = time_oracle_c1(v, ix, rays)
    out = {"workload": "UV sphere numPhi=128 (65024 triangles), 1024x1024 pinhole rays from (0,0,3), 45 deg FOV, "
                       "closest hit, one thread",
           "host_cpu": cpu_model(), "host_cores": os.cpu_count(),
           "oracle_mrays_per_s_single_thread": round(rate, 3), "oracle_hits": hits,
           "reference_mrays_per_s_single_thread": REFERENCE_MRAYS_SINGLE_THREAD,
           "reference_host_cpu": "Intel(R) Xeon(R) Processor @ 2.70GHz (8 cores), AVX2 build",
           "r": round(rate / REFERENCE_MRAYS_SINGLE_THREAD, 4)}
    with open(os.path.join(HERE, "cpu_calibration.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
