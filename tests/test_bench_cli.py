"""bench.py's command-line contract on the CPU: more than one GPU is refused
with one JSON line and exit status 2 before anything touches a device, and
the roofline reads its traffic figures from the committed profile summary."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2, p.stderr
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert "single-GPU" in line["error"]


def test_profiled_traffic_parses_the_committed_summary():
    sys.path.insert(0, ROOT)
    import bench
    traffic, l2hit, limiter = bench.profiled_traffic()
    assert traffic is not None and 1e10 < traffic < 2e11
    assert 0.0 < l2hit < 1.0
    assert 100 < limiter["valu_instr_per_ray"] < 1000 and 1 < limiter["lanes_active_per_valu_instr"] <= 64
