"""bench.py's command line without a GPU: more than one GPU is refused with a message (the
library runs on one GPU), and the options it documents parse."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=60, cwd=ROOT)


def test_more_than_one_gpu_is_refused():
    for n in ("2", "8"):
        r = run("--gpus", n, "--steps", "20", "--warmup", "5")
        assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def test_options_parse():
    r = run("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--form", "--workload", "--no-cpu", "--no-extra", "--no-accuracy", "--launch",
                "omp2d", "domains", "sod65536", "circle4096", "f64f", "ref32"):
        assert opt in r.stdout, opt
    r = run("--form", "f65")
    assert r.returncode == 2 and "invalid choice" in r.stderr
