"""bench.py's command-line contract, checked on the CPU: more than one GPU is
refused with a message (the benchmark measures one GPU), and every leg can be
switched off by its own flag."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_every_leg_has_a_switch():
    r = _bench("--help")
    assert r.returncode == 0
    for flag in ["--no-cpu", "--no-profile", "--batch-graphs", "--serve-requests", "--delaunay-points", "--no-tune",
                 "--no-cut-leg", "--no-topology-leg", "--no-udf-leg", "--no-surface-leg", "--no-flow-leg", "--no-shard-leg",
                 "--copy-probe", "--atomic-probe", "--steps", "--warmup", "--no-standin-leg",
                 "--workload", "--points", "--cpu-sample-points", "--no-small-legs", "--tet-order",
                 "--serve-pool", "--serve-p99-ms", "--gather-probe"]:
        assert flag in r.stdout, flag
