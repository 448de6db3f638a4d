"""emt6ro_batch_create checks every argument before its first device call,
so a bad batch is rejected with a message on any machine, GPU or not."""
import numpy as np
import pytest

from conftest import golden


@pytest.fixture(scope="module")
def setup():
    import emt6ro.simulation as sim
    params = sim.load_parameters(golden("default-parameters.json"))
    return params, sim.TumorState.empty(params, 53, 53)


def _batch(params, tumors, ti, pi, res=300, n_prot=1, hz=None, sid=None):
    from emt6ro.batch import Batch
    n = len(ti)
    return Batch(params, tumors, np.asarray(ti), np.asarray(pi),
                 np.full(n, 10, np.uint32) if hz is None else hz,
                 np.zeros((n_prot, 4), np.float32), res, 1200,
                 sim_id=np.arange(n) if sid is None else sid, seed=0)


def test_batch_rejects_oversized_lattices(setup):
    import emt6ro.simulation as sim
    params, _ = setup
    big = sim.TumorState.empty(params, 87, 87)
    with pytest.raises(RuntimeError, match="Lattices are limited to 86 x 86"):
        _batch(params, [big], [0], [0])


def test_batch_rejects_indices_out_of_range(setup):
    params, ok = setup
    with pytest.raises(RuntimeError, match="Tumour index out of range: 1"):
        _batch(params, [ok], [0, 1], [0, 0])
    with pytest.raises(RuntimeError, match="Tumour index out of range: -1"):
        _batch(params, [ok], [-1], [0])
    with pytest.raises(RuntimeError, match="Protocol index out of range: 2"):
        _batch(params, [ok], [0, 0], [0, 2], n_prot=2)


def test_batch_rejects_bad_protocols_and_counts(setup):
    params, ok = setup
    with pytest.raises(RuntimeError, match="Protocol resolution must be positive"):
        _batch(params, [ok], [0], [0], res=0)
    with pytest.raises(RuntimeError, match="counts must be positive"):
        _batch(params, [ok], [], [])


def test_batch_rejects_arrays_of_different_lengths(setup):
    params, ok = setup
    with pytest.raises(ValueError, match="one length"):
        _batch(params, [ok], [0, 0], [0])
    with pytest.raises(ValueError, match="one length"):
        _batch(params, [ok], [0, 0], [0, 0], hz=np.full(1, 10, np.uint32))
    with pytest.raises(ValueError, match="one length"):
        _batch(params, [ok], [0, 0], [0, 0], sid=np.arange(3))


def test_experiment_rejects_bad_arguments(setup):
    """Experiment (the reference's drop-in, built on a batch) rejects the
    same way: no tumours or zero counts, lattices over 86 x 86, tumours of
    different sizes in one experiment (the reference's Experiment sizes every
    simulation by the parameters' lattice dimensions, backend.cc:37-52)."""
    import emt6ro.simulation as sim
    params, ok = setup
    for tumors, reps, prots in (([], 1, 2), ([ok], 0, 2), ([ok], 1, 0)):
        with pytest.raises(RuntimeError, match="counts must be positive"):
            sim.Experiment(params, tumors, reps, prots, simulation_steps=600, seed=0)
    with pytest.raises(RuntimeError, match="Lattices are limited to 86 x 86"):
        sim.Experiment(params, [sim.TumorState.empty(params, 87, 87)], 1, 1,
                       simulation_steps=600, seed=0)
    with pytest.raises(RuntimeError, match="same dimensions"):
        sim.Experiment(params, [ok, sim.TumorState.empty(params, 33, 33)], 1, 1,
                       simulation_steps=600, seed=0)


def test_surrogate_network_rejects_bad_shapes():
    from emt6ro.surrogate import MLP
    for shape in ((0, 512), (71, 0), (71, 100)):
        with pytest.raises(RuntimeError, match="multiple of 64"):
            MLP(*shape)
