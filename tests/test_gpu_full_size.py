"""The benchmark's full workload on the GPU: 10 synthetic 53x53 tumours x 100
random protocols x 20 replicates = 20,000 simulations, one simulated hour
(600 steps, doses at 0 and 0.5 h included).

At this size the checks are the ones that do not need the oracle on every
simulation: the persistent kernel (lane-run diffusion) and the per-step
kernels (banded diffusion, replayed from a HIP graph) are two independent
implementations of the step, so their living counts, list lengths, division
counts and sampled states must agree bit for bit; and simulations spread over
the whole batch (the first and the last, and others at different protocol,
tumour and replicate positions) match the CPU oracle bit for bit."""
import sys

import numpy as np
import pytest

import oracle as O
from conftest import golden
from test_gpu_parity import compare_states, oracle_state_from_arrays

pytestmark = pytest.mark.gpu

SAMPLE = (0, 1, 19, 20, 199, 7777, 12345, 19980, 19999)
FLOAT_FIELDS = ("cho", "ox", "gi", "time_in_repair", "irradiation", "repair_delay",
                "proliferation_time", "cycle_g1", "cycle_s", "cycle_g2", "cycle_m", "cycle_d")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def arrays_of(state):
    out = {name: np.array(state.field(name)) for name in FLOAT_FIELDS}
    out["occupancy"] = np.array(state.occupancy())
    out["mode"] = np.array(state.mode())
    out["phase"] = np.array(state.phase())
    return out


def test_full_workload_first_hour():
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    from emt6ro.workloads import batched_uniform
    params = sim.load_parameters(golden("default-parameters.json"))
    v = params.values()
    wl = batched_uniform(params, seed=0)
    n = len(wl["tumor_index"])
    assert n == 20000
    b = Batch(params, wl["tumors"], wl["tumor_index"], wl["prot_index"], wl["horizon"],
              wl["doses"], 300, wl["prot_length"], sim_id=np.arange(n), seed=0)
    log("[full] persistent, 600 steps, 20000 simulations")
    b.run(600, path="persistent")
    res_p = b.results().astype(np.int64)
    cnt_p = b.counters()
    assert (cnt_p["step"] == 600).all() and (cnt_p["status"] == 0).all()
    states_p = {i: arrays_of(b.state(i)) for i in SAMPLE}

    # Add the first simulation of a protocol dosing at 0 h and of one at 0.5 h.
    prot = np.asarray(wl["prot_index"])
    extra = [int(np.flatnonzero(prot == int(np.flatnonzero(wl["doses"][:, s] > 0)[0]))[0])
             for s in (0, 1)]
    log(f"[full] oracle on sampled simulations {list(SAMPLE) + extra}")
    for i in list(SAMPLE) + extra:
        t = int(wl["tumor_index"][i])
        orc = O.OracleBatch(v, [oracle_state_from_arrays(arrays_of(wl["tumors"][t]), v)],
                            wl["doses"], [int(wl["prot_index"][i])], [i], 300,
                            wl["prot_length"])
        orc.run(600, seed=0, nthreads=1)
        assert res_p[i] == orc.counts()[0], (i, res_p[i], orc.counts()[0])
        compare_states(b.state(i), orc, 0, f"full-size sim {i}")

    log("[full] per-step graph path, 600 steps")
    b.reset(0)
    b.run(600, path="graph")
    assert np.array_equal(b.results().astype(np.int64), res_p)
    cnt_g = b.counters()
    for k in ("step", "n_list", "divisions", "status"):
        assert np.array_equal(cnt_g[k], cnt_p[k]), k
    for i in SAMPLE:
        got = arrays_of(b.state(i))
        for name, a in states_p[i].items():
            assert np.array_equal(got[name].view(np.uint8), a.view(np.uint8)), (i, name)
    log(f"[full] mean living {res_p.mean():.2f}, divisions {int(cnt_p['divisions'].sum())}")
    assert cnt_p["divisions"].sum() > 0


def test_uneven_workload_first_hour():
    """50,000 simulations mixing 33x33, 53x53 and 81x81 lattices, horizons of
    48, 144 and 240 h and 1-20 replicates per (tumour, protocol): one
    persistent launch per size class.  Over the first hour the persistent and
    per-step paths agree bit for bit, and sampled simulations of every size
    match the oracle."""
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    from emt6ro.workloads import batched_uneven
    params = sim.load_parameters(golden("default-parameters.json"))
    v = params.values()
    wl = batched_uneven(params, seed=0)
    n = len(wl["tumor_index"])
    assert n == 50000 and set(np.unique(wl["sizes"]).tolist()) == {33, 53, 81}
    b = Batch(params, wl["tumors"], wl["tumor_index"], wl["prot_index"], wl["horizon"],
              wl["doses"], 300, wl["prot_length"], sim_id=np.arange(n), seed=3)
    log("[full] uneven: persistent, 600 steps, 50000 simulations")
    b.run(600, path="persistent")
    res_p = b.results().astype(np.int64)
    cnt_p = b.counters()
    assert (cnt_p["step"] == 600).all() and (cnt_p["status"] == 0).all()
    sample = [0, n - 1]
    for size in (33, 53, 81):
        idx = np.flatnonzero(wl["sizes"] == size)
        sample += [int(idx[0]), int(idx[len(idx) // 2])]
    sample = sorted(set(sample))
    states_p = {i: arrays_of(b.state(i)) for i in sample}
    log(f"[full] uneven: oracle on {sample}")
    for i in sample:
        t = int(wl["tumor_index"][i])
        orc = O.OracleBatch(v, [oracle_state_from_arrays(arrays_of(wl["tumors"][t]), v)],
                            wl["doses"], [int(wl["prot_index"][i])], [i], 300,
                            wl["prot_length"])
        orc.run(600, seed=3, nthreads=1)
        assert res_p[i] == orc.counts()[0], (i, res_p[i], orc.counts()[0])
        compare_states(b.state(i), orc, 0, f"uneven full-size sim {i} ({wl['sizes'][i]})")
    log("[full] uneven: per-step graph path, 600 steps")
    b.reset(0)
    b.run(600, path="graph")
    assert np.array_equal(b.results().astype(np.int64), res_p)
    cnt_g = b.counters()
    for k in ("step", "n_list", "divisions", "status"):
        assert np.array_equal(cnt_g[k], cnt_p[k]), k
    for i in sample:
        got = arrays_of(b.state(i))
        for name, a in states_p[i].items():
            assert np.array_equal(got[name].view(np.uint8), a.view(np.uint8)), (i, name)


def _to_horizon(wl, idx, seed, label):
    """Simulations `idx` of a workload (their own streams, tumours,
    protocols and horizons) run to their horizons on the persistent path,
    each checked bit for bit against the oracle over the whole run (the
    oracle runs the simulations of each lattice size in parallel)."""
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    params = sim.load_parameters(golden("default-parameters.json"))
    v = params.values()
    idx = np.asarray(idx)
    b = Batch(params, wl["tumors"], np.asarray(wl["tumor_index"])[idx],
              np.asarray(wl["prot_index"])[idx], np.asarray(wl["horizon"])[idx], wl["doses"],
              300, wl["prot_length"], sim_id=idx, seed=seed)
    hmax = int(np.asarray(wl["horizon"])[idx].max())
    log(f"[{label}] persistent path to the horizon ({hmax} steps) for {len(idx)} simulations")
    b.run(hmax, path="persistent")
    res = b.results().astype(np.int64)
    cnt = b.counters()
    assert np.array_equal(cnt["step"], np.asarray(wl["horizon"])[idx]) and (cnt["status"] == 0).all()
    dims = [wl["tumors"][int(wl["tumor_index"][i])].shape for i in idx]
    for shape in sorted(set(dims)):
        # One oracle batch per lattice size, run in stages to each horizon;
        # a simulation is checked when the batch reaches its horizon.
        grp = [k for k in range(len(idx)) if dims[k] == shape]
        states = [oracle_state_from_arrays(
            arrays_of(wl["tumors"][int(wl["tumor_index"][idx[k]])]), v) for k in grp]
        orc = O.OracleBatch(v, states, wl["doses"], [int(wl["prot_index"][idx[k]]) for k in grp],
                            [int(idx[k]) for k in grp], 300, wl["prot_length"])
        done = 0
        for h in sorted(set(int(wl["horizon"][idx[k]]) for k in grp)):
            log(f"[{label}] oracle: {len(grp)} simulations of {shape[0]}x{shape[1]} to step {h}")
            orc.run(h - done, seed=seed, nthreads=16)
            done = h
            for j, k in enumerate(grp):
                if int(wl["horizon"][idx[k]]) != h:
                    continue
                assert res[k] == orc.counts()[j], (int(idx[k]), res[k], orc.counts()[j])
                compare_states(b.state(k), orc, j, f"{label} sim {int(idx[k])} at step {h}")
    log(f"[{label}] living at the horizons: {res.tolist()}")


def test_uniform_workload_to_144h_matches_oracle():
    """Simulations spread over the 20,000 of the uniform workload, run to
    their 144 h horizon; the batch keeps the window pass, whose ROIs grow out
    of the window during the run."""
    import emt6ro.simulation as sim
    from emt6ro.workloads import batched_uniform
    params = sim.load_parameters(golden("default-parameters.json"))
    wl = batched_uniform(params, seed=0)
    _to_horizon(wl, [0, 19, 4321, 9999, 15000, 19999], 0, "uniform 144 h")


def test_uneven_workload_to_horizons_matches_oracle():
    """Two simulations of every (lattice size, horizon) pair of the 50,000
    mixed workload (33/53/81, 48/144/240 h) run to their horizons."""
    import emt6ro.simulation as sim
    from emt6ro.workloads import batched_uneven
    params = sim.load_parameters(golden("default-parameters.json"))
    wl = batched_uneven(params, seed=0)
    idx = []
    for size in (33, 53, 81):
        for h in (48, 144, 240):
            cand = np.flatnonzero((wl["sizes"] == size) & (wl["hours"] == h))
            idx += [int(cand[0]), int(cand[len(cand) // 2])]
    _to_horizon(wl, idx, 3, "uneven to horizons")
