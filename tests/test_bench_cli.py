"""bench.py's command line, without a GPU: more than one GPU is refused with a message and
exit status 2 before anything touches a device; every workload switch is listed."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def _run(*args):
    return subprocess.run([sys.executable, BENCH, *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    for n in ("2", "8"):
        r = _run("--gpus", n)
        assert r.returncode == 2
        assert "one GPU only" in r.stderr and f"--gpus {n}" in r.stderr
        assert r.stdout == ""


def test_help_lists_every_workload():
    r = _run("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--grid128", "--spectrum-fp32", "--scft", "--plain-c4",
                 "--batch-c4", "--no-graph", "--warm", "--warm-train", "--unet", "--tol",
                 "--max-iter"):
        assert flag in r.stdout, flag
