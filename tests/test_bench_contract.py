"""bench.py's command-line contract, checked without a GPU: one GPU only
(--gpus above 1 exits with status 2 and says why), and the flags the driver
passes are accepted."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def run_bench(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True, timeout=120,
                          cwd=ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    for n in ("2", "8"):
        r = run_bench("--gpus", n, "--steps", "5", "--warmup", "1")
        assert r.returncode == 2
        assert "one GPU" in r.stderr and f"--gpus {n}" in r.stderr
        assert r.stdout == ""  # no JSON line for a configuration it does not run


def test_help_lists_the_driver_flags():
    r = run_bench("--help")
    assert r.returncode == 0
    for flag in ("--gpus", "--steps", "--warmup", "--mode"):
        assert flag in r.stdout
