"""Uneven batches on the GPU: lattices of 33, 53 and 81 with different
horizons in one batch (one persistent launch per size class), checked bit for
bit against the oracle, and the autotuner's choice checked not to change
results."""
import sys

import numpy as np
import pytest

import oracle as O
from conftest import golden
from test_gpu_parity import compare_states, oracle_state_from_arrays

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_uneven_sizes_and_horizons_match_oracle():
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    from emt6ro.synthetic import disk_state
    params = sim.load_parameters(golden("default-parameters.json"))
    v = params.values()
    sizes = [33, 53, 81, 33, 81, 53]
    radii = [6, 10, 14, 9, 5, 12]
    tumors, orc_states = [], []
    for k, (sz, r) in enumerate(zip(sizes, radii)):
        st, arr = disk_state(params, sz, sz, radius=r, seed=50 + k)
        tumors.append(st)
        orc_states.append(oracle_state_from_arrays(arr, v))
    n = len(tumors) * 2
    tumor_index = np.repeat(np.arange(len(tumors)), 2)
    horizon = np.array([900, 1500, 600, 1500, 1200, 900, 1500, 1500, 300, 1200, 1500, 600],
                       np.uint32)
    doses = np.zeros((2, 480), np.float32)
    doses[0, 0] = 2.5
    doses[1, 2] = 1.0
    prot_index = np.arange(n) % 2
    b = Batch(params, tumors, tumor_index, prot_index, horizon, doses, 300, 144000,
              sim_id=np.arange(n), seed=99)
    log("[uneven] run 1500")
    b.run(1500)
    got = b.results().astype(np.int64)
    cnt = b.counters()
    assert np.array_equal(cnt["step"].astype(np.int64), horizon.astype(np.int64))
    for i in range(n):
        t = int(tumor_index[i])
        orc = O.OracleBatch(v, [orc_states[t]], doses, [prot_index[i]], [i], 300, 144000)
        orc.run(int(horizon[i]), seed=99)
        assert got[i] == orc.counts()[0], (i, got[i], orc.counts()[0])
        compare_states(b.state(i), orc, 0, f"uneven sim {i} ({sizes[t]}x{sizes[t]})")


def test_large_rois_match_oracle():
    """ROIs large enough that the lane-run diffuser's kept list of masked
    ROI sites overflows on some threads (its per-step loop takes over),
    next to ones that fit, on 53x53 and 81x81 lattices."""
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    from emt6ro.synthetic import disk_state
    params = sim.load_parameters(golden("default-parameters.json"))
    v = params.values()
    sizes = [53, 53, 81, 81]
    radii = [20, 24, 28, 36]
    tumors, orc_states = [], []
    for k, (sz, r) in enumerate(zip(sizes, radii)):
        st, arr = disk_state(params, sz, sz, radius=r, seed=70 + k)
        tumors.append(st)
        orc_states.append(oracle_state_from_arrays(arr, v))
    n = len(tumors)
    doses = np.zeros((1, 480), np.float32)
    doses[0, 0] = 1.5
    steps = 300
    b = Batch(params, tumors, np.arange(n), np.zeros(n, np.int64),
              np.full(n, steps, np.uint32), doses, 300, 144000, sim_id=np.arange(n), seed=5)
    log(f"[uneven] large ROIs: radii {radii}, {steps} steps")
    b.run(steps)
    got = b.results().astype(np.int64)
    for i in range(n):
        orc = O.OracleBatch(v, [orc_states[i]], doses, [0], [i], 300, 144000)
        orc.run(steps, seed=5, nthreads=4)
        assert got[i] == orc.counts()[0], (i, got[i], orc.counts()[0])
        compare_states(b.state(i), orc, 0, f"large ROI sim {i} ({sizes[i]}x{sizes[i]}, r {radii[i]})")
        log(f"  sim {i}: ROI {orc.roi[0].tolist()} living {got[i]}")


def test_autotuner_keeps_results():
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    from emt6ro.synthetic import disk_state
    from emt6ro.tune import autotune
    params = sim.load_parameters(golden("default-parameters.json"))
    tumors = [disk_state(params, 53, 53, radius=10, seed=1)[0],
              disk_state(params, 33, 33, radius=8, seed=2)[0]]
    n = 64
    b = Batch(params, tumors, np.arange(n) % 2, np.zeros(n, np.int32), np.full(n, 2 ** 32 - 1),
              np.zeros((1, 480), np.float32), 300, 144000, sim_id=np.arange(n), seed=3)
    b.run(400)
    want = b.results()
    b.reset(0)
    res = autotune(b, {56: (53, 53), 36: (33, 33)}, steps=60, use_cache=False, log=log)
    assert set(res) == {56, 36}
    b.run(400)
    assert np.array_equal(b.results(), want)
    # Re-timing the closest variants later in the run keeps results too.
    b.reset(0)
    res = autotune(b, {56: (53, 53)}, steps=60, use_cache=False, log=log, refine_warm=120,
                   refine_top=2, refine_within=1.0)
    later = res[56]["ms_later"]
    assert 1 <= len(later) <= 2 and res[56]["variant"] in later
    b.run(400)
    assert np.array_equal(b.results(), want)


def test_81x81_windows_match_oracle():
    """81x81 lattices whose ROIs fit the 36-column window, the 56-column
    window, or neither, over several one-hour launches."""
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    from emt6ro.synthetic import disk_state
    params = sim.load_parameters(golden("default-parameters.json"))
    v = params.values()
    radii = [8, 17, 20, 23]
    tumors, orc_states = [], []
    for k, r in enumerate(radii):
        st, arr = disk_state(params, 81, 81, radius=r, seed=90 + k)
        tumors.append(st)
        orc_states.append(oracle_state_from_arrays(arr, v))
    n = len(tumors)
    doses = np.zeros((1, 480), np.float32)
    doses[0, 1] = 2.0
    steps = 1800
    b = Batch(params, tumors, np.arange(n), np.zeros(n, np.int64),
              np.full(n, 2 ** 32 - 1, np.uint32), doses, 300, 144000, sim_id=np.arange(n),
              seed=6)
    log(f"[uneven] 81x81 windows: radii {radii}, {steps} steps")
    b.run(steps, steps_per_launch=600)
    got = b.results().astype(np.int64)
    for i in range(n):
        orc = O.OracleBatch(v, [orc_states[i]], doses, [0], [i], 300, 144000)
        orc.run(steps, seed=6, nthreads=4)
        assert got[i] == orc.counts()[0], (i, got[i], orc.counts()[0])
        compare_states(b.state(i), orc, 0, f"81x81 window sim {i} (r {radii[i]})")
        log(f"  sim {i}: ROI {orc.roi[0].tolist()} living {got[i]}")
