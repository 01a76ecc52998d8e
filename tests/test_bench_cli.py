"""CPU checks of bench.py's command line: it measures one GPU and says so for --gpus N > 1,
before it loads the library or touches a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=60)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = run_bench("--gpus", "2", "--steps", "1", "--warmup", "0")
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""


def test_help_lists_the_measurement_legs():
    r = run_bench("--help")
    assert r.returncode == 0
    for flag in ("--steps", "--warmup", "--gpus", "--cpu-steps", "--coupled-steps", "--ct-steps", "--refbench"):
        assert flag in r.stdout
