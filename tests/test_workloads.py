"""The seeded synthetic workloads the benchmark and the full-size GPU tests
run (emt6ro.workloads): protocols inside the GA's search space
(python_ga/utils.py: 0.25 Gy steps, 0.5 h grid), the Experiment layout of
the uniform batch (backend.cc:59-73), and the mixed batch's sizes, horizons
and replicate groups."""
import numpy as np
import pytest

from conftest import golden


@pytest.fixture(scope="module")
def params():
    import emt6ro.simulation as sim
    return sim.load_parameters(golden("default-parameters.json"))


def test_random_protocols_stay_in_the_search_space():
    from emt6ro.workloads import SLOT_STEPS, random_protocols
    prots = random_protocols(500, seed=3)
    assert prots == random_protocols(500, seed=3)
    assert prots != random_protocols(500, seed=4)
    lens = [len(p) for p in prots]
    assert min(lens) >= 1 and max(lens) == 10
    for p in prots:
        steps = [s for s, _ in p]
        doses = np.array([d for _, d in p])
        assert steps == sorted(set(steps))
        assert all(s % SLOT_STEPS == 0 and 0 <= s < 72 * 600 for s in steps)
        assert np.all((doses >= 0.25) & (doses <= 1.25))
        assert np.allclose(doses / 0.25, np.round(doses / 0.25))
        assert doses.sum() <= 10.0 + 1e-9


def test_dose_table_accumulates_on_a_slot():
    from emt6ro.workloads import dose_table
    t = dose_table([[(0, 1.0), (150, 0.5), (600, 2.0)], [(299, 0.25)]], n_slots=4)
    assert t.dtype == np.float32 and t.shape == (2, 4)
    np.testing.assert_array_equal(t, [[1.5, 0, 2.0, 0], [0.25, 0, 0, 0]])


def test_batched_uniform_layout(params):
    from emt6ro.workloads import batched_uniform
    wl = batched_uniform(params, n_tumors=3, n_protocols=4, reps=5, hours=12, seed=0)
    P, T, R = 4, 3, 5
    b = np.arange(P * T * R)
    np.testing.assert_array_equal(wl["prot_index"], b // (T * R))
    np.testing.assert_array_equal(wl["tumor_index"], (b // R) % T)
    assert np.all(wl["horizon"] == 12 * 600) and wl["prot_length"] == 12 * 600
    # the table holds every dose, also those past a short horizon (they never
    # fire: a dose fires only before prot_length)
    assert wl["doses"].shape[0] == P and wl["doses"].shape[1] >= 12 * 600 // 300
    np.testing.assert_allclose(wl["doses"].sum(1), [sum(d for _, d in p) for p in wl["protocols"]],
                               rtol=1e-6)
    assert len(wl["tumors"]) == T and all(t.shape == (53, 53) for t in wl["tumors"])
    living = [int((t.occupancy() == 1).sum()) for t in wl["tumors"]]
    assert all(150 < n < 700 for n in living)   # disks of radius 8-14


def test_batched_uniform_default_is_the_benchmark_workload(params):
    from emt6ro.workloads import batched_uniform
    wl = batched_uniform(params)
    assert len(wl["tumor_index"]) == 20000 and len(wl["tumors"]) == 10
    assert wl["doses"].shape == (100, 288) and wl["prot_length"] == 86400
    assert np.all(wl["horizon"] == 86400)


def test_batched_uneven_groups(params):
    from emt6ro.workloads import batched_uneven
    wl = batched_uneven(params, n_sims=3000, seed=2)
    n = len(wl["tumor_index"])
    assert n == 3000 and len(wl["prot_index"]) == len(wl["horizon"]) == n
    assert set(np.unique(wl["sizes"])) == {33, 53, 81}
    assert set(np.unique(wl["hours"])) == {48.0, 144.0, 240.0}
    np.testing.assert_array_equal(wl["horizon"], (wl["hours"] * 600).astype(np.uint32))
    shapes = np.array([wl["tumors"][t].shape[0] for t in wl["tumor_index"]])
    np.testing.assert_array_equal(shapes, wl["sizes"])
    # groups of identical (tumour, protocol, horizon) runs of 1-20 replicates
    key = np.stack([wl["tumor_index"], wl["prot_index"], wl["horizon"]], 1)
    cuts = np.flatnonzero(np.any(key[1:] != key[:-1], axis=1)) + 1
    runs = np.diff(np.concatenate([[0], cuts, [n]]))
    assert runs.max() <= 20 * 2   # two draws of one group can fall side by side
    assert wl["doses"].shape[1] == 240 * 600 // 300 and wl["prot_length"] == 240 * 600
