"""The surrogate on simulator data and inside the GA (GPU).

Training data: tests/golden/surrogate_fixture.npz, written by
tests/golden/make_surrogate_fixture.py from the GPU simulator (three seeded
synthetic 53x53 tumours x 2,400 random protocols, mean final count over 10
replicates at 144 h).

  * Surrogate training: held-out mean absolute error at most 5% of the
    simulated mean for the feature network the GA uses, and a clear gain over
    predicting the mean; the network on the configured raw input (144 dose
    slots + 8 descriptors) is trained the same way and must beat predicting
    the mean.
  * GA: at an equal simulator budget (initial population plus three
    generations of 100 simulated genomes, 10 replicates per tumour), the
    surrogate-screened GA's best genome, re-simulated on fresh replicates, is
    at least as fit as the plain GA's (python_ga/genetic_algorithm.py:161-252
    with EMT6RoModel fitness, model.py:36-54).
"""
import sys

import numpy as np
import pytest

from conftest import golden, gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X GPU")]


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def data():
    import emt6ro.simulation as sim
    from emt6ro.ga import experiment as X
    z = np.load(golden("surrogate_fixture.npz"))
    params = sim.load_parameters(golden("default-parameters.json"))
    ts, desc = X.tumours(params, 10, 0)
    assert np.allclose(desc[:3], z["descriptors"]), "fixture tumours differ from the GA's"
    genomes = z["dose_units"].astype(np.float32) * np.float32(0.25)
    return {"params": params, "tumours": ts[:3], "desc": z["descriptors"], "genomes": genomes,
            "mean": z["mean"].astype(np.float64)}


def _split(P, seed=0):
    return np.random.default_rng(seed).random(P) < 0.1


def _fit_eval(S, X, y, pi, held, members, in_features, params):
    te = held[pi]
    model = S.Surrogate(in_features=in_features, seed=1, members=members, params=params)
    model.fit(X[~te], y[~te], steps=500, batch=2048, lr=1e-3, seed=1)
    pred = model.predict(X[te])
    mae = float(np.abs(pred - y[te]).mean())
    base = float(np.abs(y[te] - y[~te].mean()).mean())
    return model, mae, base, float(y[te].mean())


def test_surrogate_held_out_error_on_simulator_data(data):
    from emt6ro import surrogate as S
    g, mean, desc = data["genomes"], data["mean"], data["desc"]
    P, T = mean.shape
    pi, ti = np.repeat(np.arange(P), T), np.tile(np.arange(T), P)
    y = mean.reshape(-1)
    held = _split(P)
    X = S.features(g, desc, pi, ti, data["params"])
    _, mae, base, ym = _fit_eval(S, X, y, pi, held, 5, S.N_FEATURES, data["params"])
    log(f"[surrogate] features: held-out MAE {mae:.2f} = {100 * mae / ym:.2f}% of {ym:.1f}; "
        f"predict-the-mean {base:.2f}")
    assert mae <= 0.05 * ym
    assert mae <= 0.5 * base
    # The configured input: 144 dose slots (Gy) + the 8 descriptors.
    Xr = np.concatenate([np.asarray(desc, np.float32)[ti], g[pi]], axis=1)
    _, mae_r, _, _ = _fit_eval(S, Xr, y, pi, held, 5, Xr.shape[1], data["params"])
    log(f"[surrogate] raw 152 inputs: held-out MAE {mae_r:.2f} = {100 * mae_r / ym:.2f}%")
    assert mae_r < 0.8 * base


def test_surrogate_screened_ga_beats_plain_ga_at_equal_budget(data):
    from emt6ro import surrogate as S
    from emt6ro.ga import driver as D
    from emt6ro.ga import operators as ops
    g, mean, desc = data["genomes"], data["mean"], data["desc"]
    P, T = mean.shape
    pi, ti = np.repeat(np.arange(P), T), np.tile(np.arange(T), P)
    model = S.Surrogate(seed=2, members=5, params=data["params"])
    model.fit(S.features(g, desc, pi, ti, data["params"]), mean.reshape(-1), steps=500,
              batch=2048, lr=1e-3, seed=2)
    cfg = ops.default_config()
    gens, pop, reps = 3, 100, 10
    fit_p = D.SimulatorFitness(data["params"], data["tumours"], reps, seed=11)
    hist_p = D.run_plain_ga(np.random.default_rng(1), fit_p, cfg, pop_size=pop,
                            generations=gens, log=log)
    fit_s = D.SimulatorFitness(data["params"], data["tumours"], reps, seed=12)
    hist_s = D.run_surrogate_ga(np.random.default_rng(2), fit_s, model, desc, cfg,
                                pop_size=pop, generations=gens, pool=10000, verify=pop,
                                log=log)
    assert fit_p.evaluations == fit_s.evaluations == pop * (gens + 1)
    judge = D.SimulatorFitness(data["params"], data["tumours"], 100, seed=99)

    def best(hist):
        h = max(hist, key=lambda x: x["best"])
        return np.asarray(h["best_genome"], np.float32)

    f_plain, f_sur = judge([best(hist_p), best(hist_s)])
    log(f"[ga] re-simulated best: plain {f_plain:.2f}, surrogate-screened {f_sur:.2f} "
        f"({fit_p.evaluations} genomes simulated by each)")
    assert f_sur >= f_plain
