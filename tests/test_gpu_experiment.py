"""The Experiment drop-in API on the GPU (python/emt6ro/simulation/__init__.py,
backend.cc:16-136) and GPU diffusion against the FP64 oracle."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle as O
from conftest import golden

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sim():
    import emt6ro.simulation as sim
    return sim


@pytest.fixture(scope="module")
def params(sim):
    return sim.load_parameters(golden("default-parameters.json"))


@pytest.fixture(scope="module")
def tumors(sim, params):
    return [sim.load_state(golden("tumor-lib", f"tumor-{t}.txt"), params) for t in (1, 2)]


def test_experiment_shapes_and_layout(sim, params, tumors):
    log("[experiment] shapes")
    e = sim.Experiment(params, tumors, 3, 2, simulation_steps=1200, protocol_resolution=300,
                       seed=4)
    e.add_irradiations([[(0, 1.0)], [(300, 2.0), (300, 0.5)]])
    e.run(600)
    res = e.get_results()
    assert res.shape == (2, 2, 3) and res.dtype.kind == "i"
    states = e.state()
    assert len(states) == 12
    # (protocol, tumour, run) order: sample p*T*R + t*R + r.
    for p in range(2):
        for t in range(2):
            for r in range(3):
                s = states[p * 6 + t * 3 + r]
                assert int((s.occupancy() == 1).sum()) == res[p, t, r]
    assert states[0].occupancy().flags.writeable is False


def test_experiment_seed_reproducible_and_reset_epoch(sim, params, tumors):
    log("[experiment] seeds")
    def one(seed):
        e = sim.Experiment(params, tumors, 4, 1, simulation_steps=3000, seed=seed)
        e.add_irradiations([[(0, 2.0)]])
        e.run(3000)
        a = e.get_results()
        e.reset()
        e.add_irradiations([[(0, 2.0)]])
        e.run(3000)
        return a, e.get_results()
    a1, b1 = one(11)
    a2, b2 = one(11)
    assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
    # After reset the random stream moves on, as the reference's generator
    # states do (simulation.cu:212-216), so the rerun is a new sample.
    assert not np.array_equal(a1, b1)


def test_experiment_errors(sim, params, tumors):
    e = sim.Experiment(params, tumors, 1, 2, simulation_steps=600, seed=0)
    with pytest.raises(RuntimeError, match="First you have to run the experiment."):
        e.get_results()
    with pytest.raises(RuntimeError, match=r"Check failed: .*Wrong number of protocols. Expected: 2"):
        e.add_irradiations([[(0, 1.0)]])
    e.run(10)
    with pytest.raises(RuntimeError, match="Experiment already running."):
        e.run(10)
    e.get_results()


def test_paths_agree_through_experiment(sim, params, tumors):
    log("[experiment] paths")
    out = []
    for path in (0, 1, 2):
        e = sim.Experiment(params, tumors, 2, 1, simulation_steps=2000, seed=21)
        e._experiment.set_path(path)
        e.add_irradiations([[(0, 1.5), (600, 1.5)]])
        e.run(300)
        out.append(e.get_results())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def test_ga_model_predict_is_1500_minus_mean_oracle_counts(sim, params, tumors):
    """EMT6RoModel.predict (model.py:36-54): fitness = 1500 - mean final count
    over tumours x tests of each protocol.  Its Experiment runs simulation
    b = p*T*R + t*R + r on stream b of (seed, epoch); the oracle on the same
    streams gives the counts, and the second call (after the model's reset)
    draws epoch 1."""
    import oracle as O
    from emt6ro.ga_model import EMT6RoModel
    log("[experiment] EMT6RoModel.predict")
    P, T, R, steps, seed = 2, 2, 3, 600, 5
    protocols = [[(0, 1.0), (300, 0.5), (300, 0.25)], []]
    m = EMT6RoModel(params, tumors, num_protocols=P, num_tests=R, seed=seed, num_steps=steps)
    v = O.read_params_json(golden("default-parameters.json"))
    states = [O.read_state_text(golden("tumor-lib", f"tumor-{t}.txt"), v) for t in (1, 2)]
    doses = np.zeros((P, 2), np.float32)
    doses[0] = [1.0, np.float32(0.5) + np.float32(0.25)]
    b = np.arange(P * T * R)
    for epoch in (0, 1):
        f = m.predict(protocols)
        orc = O.OracleBatch(v, [states[(i // R) % T] for i in b], doses, b // (T * R), b, 300,
                            steps)
        orc.run(steps, seed=seed, epoch=epoch, nthreads=4)
        counts = orc.counts().reshape(P, T * R)
        assert f.shape == (P,)
        assert np.array_equal(f, 1500 - counts.astype(np.float64).mean(axis=1)), (f, counts)


def test_gpu_diffusion_within_1e5_of_fp64(sim, params):
    """One step with zero metabolism changes substrates by diffusion only;
    the GPU's FP32 result must be within 1e-5 (relative to the external
    level; the error relative to each site's own value is reported too and
    must stay below 1e-4) of the FP64 oracle on fixed states: 10 library tumours and
    90 synthetic lattices of 33, 53 and 81 with random substrate fields
    (100 fixed states)."""
    from emt6ro.batch import Batch
    from emt6ro.simulation import Parameters
    from emt6ro.synthetic import disk_state
    v = params.values().copy()
    v[7:19] = 0.0
    p0 = Parameters.from_values(v)
    cases = []
    rng = np.random.default_rng(0)
    for t in range(1, 11):
        cases.append(sim.load_state(golden("tumor-lib", f"tumor-{t}.txt"), p0))
    for k in range(90):
        size = (33, 53, 81)[k % 3]
        st, _ = disk_state(p0, size, size, radius=5 + k % 9, seed=k)
        for j, name in enumerate(("cho", "ox", "gi")):
            st.set_field(name, (v[19 + j] * (0.5 + rng.random((size, size)))).astype(np.float32))
        cases.append(st)
    worst = worst_pt = 0.0
    for i, st in enumerate(cases):
        H, W = st.shape
        b = Batch(p0, [st], [0], [0], [2 ** 32 - 1], np.zeros((1, 1), np.float32), 300, 0)
        b.run(1)
        g = b.state(0)
        sub = np.stack([st.field(n) for n in ("cho", "ox", "gi")]).astype(np.float64)
        roi = (C.c_int32 * 5)()
        occ = np.ascontiguousarray(st.occupancy())
        O.lib.eo_find_roi(H, W, O.ptr(occ), roi)
        p = O.make_params(v, H, W)
        out = np.zeros_like(sub)
        O.lib.eo_diffuse_f64(C.byref(p), H, W, O.ptr(np.ascontiguousarray(sub)), roi, O.ptr(out))
        for k, n in enumerate(("cho", "ox", "gi")):
            d = np.abs(g.field(n).astype(np.float64) - out[k])
            worst = max(worst, float(np.max(d)) / abs(float(v[19 + k])))
            nz = np.abs(out[k]) > 0
            worst_pt = max(worst_pt, float(np.max(d[nz] / np.abs(out[k][nz]))))
    log(f"[diffusion] max error vs FP64: {worst:.3e} relative to the external level, "
        f"{worst_pt:.3e} relative to each site's value")
    assert worst < 1e-5
    assert worst_pt < 1e-4
