"""bench.py's command line on the CPU: more than one GPU is refused with a message (one
GPU's throughput is what it measures), and the CPU-only baseline of the small
two-cell workload prints one JSON object with its timing parts."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args, timeout=300):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=timeout, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_cpu_only_baseline_small_workload():
    r = _run("--cpu-only", "--config", "C1", "--sample-every", "4")
    assert r.returncode == 0, r.stderr[-2000:]
    out = json.loads(r.stdout.strip().splitlines()[-1])
    assert out["cpu_s_per_matvec"] > 0 and out["threads"] >= 1
    assert set(out["parts"]) == {"sh_synth", "splines", "real_space_sample", "pme", "sh_anal"}
