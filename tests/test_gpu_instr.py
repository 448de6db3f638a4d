"""Instrumentation counters (common/instr.h:5-58, read_instr_* of
backend.cc:221-231) on the GPU against the oracle's counters for the same
simulations: repair decisions and kills, irradiation events, divisions, the
death-probability sum and both histograms, all exact."""
import sys

import numpy as np
import pytest

import oracle as O
from conftest import golden
from test_gpu_parity import oracle_state_from_arrays

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_instr_counters_match_oracle():
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    from emt6ro.synthetic import disk_state
    params = sim.load_parameters(golden("default-parameters.json"))
    v = params.values()
    st, arr = disk_state(params, 53, 53, radius=10, seed=3)
    n, steps = 6, 9000   # repair starts at the next cycle change, ends 3.34 exp(0.149 R) h later
    doses = np.zeros((2, 480), np.float32)
    doses[0, [0, 4]] = [2.5, 2.0]
    doses[1, [1, 6]] = [4.0, 1.0]
    prot_index = np.arange(n) % 2
    sim.enable_instr(False)
    sim.reset_instr_counters()
    b0 = Batch(params, [st], np.zeros(n, np.int32), prot_index, np.full(n, steps, np.uint32),
               doses, 300, 144000, seed=5)
    b0.run(steps)
    base = b0.results()
    assert sim.read_instr_counters().irradiation_events == 0
    sim.enable_instr(True)
    try:
        assert sim.instr_enabled()
        log("[instr] instrumented run")
        b = Batch(params, [st], np.zeros(n, np.int32), prot_index,
                  np.full(n, steps, np.uint32), doses, 300, 144000, seed=5)
        b.run(steps)
        got = b.results()
        c = sim.read_instr_counters()
        hist = np.array(sim.read_instr_histograms(), np.uint64)
    finally:
        sim.enable_instr(False)
    np.testing.assert_array_equal(got, base)  # counting does not change results
    orc = O.OracleBatch(v, [oracle_state_from_arrays(arr, v)] * n, doses, prot_index,
                        np.arange(n), 300, 144000)
    orc.run(steps, seed=5, nthreads=6, instr=True)
    np.testing.assert_array_equal(got, orc.counts())
    rec = orc.instr.sum(axis=0)
    log(f"  decisions {c.repair_decisions} kills {c.repair_kills} irradiation "
        f"{c.irradiation_events} divisions {c.divisions} sum p {c.sum_death_prob:.4f}")
    assert [c.repair_decisions, c.repair_kills, c.irradiation_events, c.divisions] == \
        [int(x) for x in rec[:4]]
    assert c.sum_death_prob == float(rec[4]) / 2 ** 32
    np.testing.assert_array_equal(hist, rec[5:])
    assert c.irradiation_events > 0 and c.repair_decisions > 0 and c.divisions > 0
    assert int(hist[:300].sum()) == c.repair_decisions == int(hist[300:].sum())
    assert int(b.counters()["divisions"].sum()) == c.divisions
    sim.reset_instr_counters()
    assert sim.read_instr_counters().repair_decisions == 0
