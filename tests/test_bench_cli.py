"""bench.py's contract that holds without a GPU: it refuses more than one GPU
with a message (the simulations are independent, each GPU would run a share
of the batch), rejects step counts that time nothing, and its K timed steps
tile the 144 h horizon exactly, whatever K is."""
import importlib.util
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT,
                          capture_output=True, text=True, timeout=120)


def test_bench_refuses_several_gpus():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_bench_rejects_empty_timing():
    assert _run("--steps", "0").returncode != 0
    assert _run("--warmup", "-1").returncode != 0


def test_bench_steps_tile_the_horizon():
    b = _bench()
    total = b.HOURS * b.HOUR
    assert total == 86400
    for k in (1, 3, 7, 20, 144, 1000):
        bounds = b.step_bounds(k)
        assert len(bounds) == k + 1
        assert bounds[0] == 0 and bounds[-1] == total
        widths = [hi - lo for lo, hi in zip(bounds, bounds[1:])]
        assert min(widths) > 0
        assert max(widths) - min(widths) <= 1
