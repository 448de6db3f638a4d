"""The measurement entry points (emt6ro_probe_lds, emt6ro_probe_grid_barrier).

CPU: invalid arguments are rejected before any device call.  GPU: the
figures they return are in the range the hardware allows, and the grid
barrier probe completes without a timed-out wait or a stale read (it fails
otherwise).
"""
import ctypes as C

import pytest

from conftest import gpu_available


def test_probes_reject_invalid_arguments():
    from emt6ro._native import lib
    out = (C.c_double * 4)()
    assert lib.emt6ro_probe_lds(0, 8, out) == -1
    assert lib.emt6ro_probe_lds(0, 4, None) == -1
    out2 = (C.c_double * 2)()
    assert lib.emt6ro_probe_grid_barrier(0, 2, 0, out2) == -1
    assert lib.emt6ro_probe_grid_barrier(0, 0, -1, out2) == -1
    assert lib.emt6ro_probe_grid_barrier(0, 0, 16385, out2) == -1
    assert lib.emt6ro_probe_grid_barrier(0, 1, 0, None) == -1


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X GPU")
def test_lds_probe_figures_are_physical():
    from emt6ro._native import probe_lds
    b16, b4 = probe_lds(16), probe_lds(4)
    print(f"[probe] LDS b128 {b16['bytes_per_clk_cu']:.1f} B/clk/CU, b32 "
          f"{b4['bytes_per_clk_cu']:.1f}, clock {b16['clock_hz'] / 1e9:.2f} GHz", flush=True)
    for r in (b16, b4):
        assert 0.5e9 < r["clock_hz"] < 3.0e9
        assert r["cus"] > 0
        assert 16.0 < r["bytes_per_clk_cu"] <= 512.0
    assert b16["bytes_per_clk_cu"] > b4["bytes_per_clk_cu"]


@pytest.mark.gpu
@pytest.mark.skipif(not gpu_available(), reason="needs an MI355X GPU")
@pytest.mark.parametrize("mode", [0, 1])
def test_grid_barrier_probe_completes(mode):
    from emt6ro._native import probe_grid_barrier
    r = probe_grid_barrier(mode=mode, words=1024)
    print(f"[probe] grid barrier mode {mode}: {r['barrier_us']:.2f} us, launch "
          f"{r['launch_us']:.2f} us", flush=True)
    assert 0.1 < r["barrier_us"] < 1000.0
    assert 0.1 < r["launch_us"] < 1000.0
