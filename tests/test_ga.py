"""GA operators (emt6ro.ga.operators, restating python_ga/utils.py,
selection.py, crossover.py, mutation.py) and the GA drivers
(emt6ro.ga.driver) on the CPU.

The reference ships no tests for its GA; these check the invariants its code
maintains (dose grid, per-dose cap, total budget, minimum spacing between
doses, annealing schedule, selection sizes) and run both drivers against a
cheap deterministic fitness, as the reference's example_run.py does with its
MockPredictionModel (python_ga/example_run.py:11-13)."""
import numpy as np
import pytest

from emt6ro.ga import operators as ops
from emt6ro.ga import driver as D

N = 144


def cfg(**kw):
    return ops.default_config(**kw)


def gaps(g):
    nz = np.nonzero(np.asarray(g))[0]
    return np.diff(nz)


def on_grid(g, step=0.25):
    g = np.asarray(g, float)
    return np.allclose(g / step, np.round(g / step))


def test_annealing_schedule_closed_form():
    assert ops.annealing_probability(1, 0.5, 100) == 0.5
    assert ops.annealing_probability(51, 0.5, 100) == round(0.5 ** 4 * 0.5, 6)
    assert ops.annealing_probability(101, 0.5, 100) == 0.001
    assert ops.annealing_probability(100, 0.2, 100) == 0.001  # below eps


def test_selection_counts_are_even():
    for pop in range(2, 60):
        for frac in (0.1, 0.33, 0.5, 0.9):
            n = ops.even_count(frac, pop)
            assert n % 2 == 0 and round(pop * frac) - 1 <= n <= round(pop * frac)
    assert ops.even_count(0.5, 5) == 2      # round(2.5) = 2 (half to even)
    assert ops.even_count(0.7, 5) == 4      # round(3.5) = 4


def test_random_protocols_respect_grid_cap_budget_and_spacing():
    rng = np.random.default_rng(0)
    c = cfg()
    for g in ops.random_population(rng, 300, N, c):
        assert g.sum() <= c["max_value"] + 1e-9
        assert g.max() <= c["max_dose_value"] + 1e-9 and g[g > 0].min() >= 0.25
        assert on_grid(g)
        # doses more than time_interval_hours apart: > 6 slots of 0.5 h
        assert (gaps(g) > 2 * c["time_interval_hours"]).all()


def test_pairs_genome_round_trip():
    pairs = [(0, 1.25), (3000, 0.5), (42900, 2.0)]
    g = ops.pairs_to_genome(pairs, N)
    assert g[0] == 1.25 and g[10] == 0.5 and g[143] == 2.0 and g.sum() == 3.75
    assert ops.genome_to_pairs(g) == pairs


def test_refine_enforces_spacing_without_adding_dose():
    rng = np.random.default_rng(1)
    c = cfg()
    for _ in range(300):
        g = np.zeros(N)
        idx = rng.choice(N, size=int(rng.integers(1, 12)), replace=False)
        g[idx] = rng.integers(1, 11, size=len(idx)) * 0.25
        r = ops.refine_to_time_constraint(rng, g, 6, c)
        assert r.sum() <= g.sum() + 1e-9
        assert (gaps(r) >= 6).all()
        # every remaining dose is one of the original doses
        assert set(np.round(r[r > 0], 6)) <= set(np.round(g[g > 0], 6))


def test_normalize_budget():
    c = cfg()
    g = np.zeros(N)
    g[[0, 10, 20, 30, 40, 50]] = [2.5, 2.5, 2.5, 2.0, 1.5, 1.0]
    r = ops.normalize_budget(g, c)
    assert r.sum() <= 10 + 1e-9 and on_grid(r) and (r >= 0).all()
    assert r.sum() == 10.0
    ok = np.zeros(N)
    ok[[3, 30]] = [1.0, 2.0]
    np.testing.assert_array_equal(ops.normalize_budget(ok, c), ok)


@pytest.mark.parametrize("kind", ["cross_one_point", "cross_two_points", "cross_uniform",
                                  "normalized_crossover"])
def test_crossovers_keep_budget_grid_and_spacing(kind):
    rng = np.random.default_rng(2)
    c = cfg(crossover={"type": kind})
    pop = ops.random_population(rng, 40, N, c)
    cross = ops.CROSSOVERS[kind]
    for _ in range(200):
        x1 = pop[int(rng.integers(40))]
        x2 = pop[int(rng.integers(40))]
        for ch in cross(rng, x1, x2, c):
            ch = np.asarray(ch, float)
            assert len(ch) == N
            assert ch.sum() <= c["max_value"] + 1e-9 and on_grid(ch)
            assert (gaps(ch) >= 6).all()
            if kind in ("cross_uniform", "normalized_crossover"):
                # gene i comes from a parent's gene i (or is dropped, or
                # trimmed by the budget normalisation)
                assert all(v == 0 or v <= max(x1[i], x2[i]) + 1e-9 for i, v in enumerate(ch))


def test_selections():
    rng = np.random.default_rng(3)
    pop = [np.full(3, i, float) for i in range(20)]
    fit = np.arange(20, dtype=float)
    top = ops.simple_selection(rng, pop, fit, 4, {})
    assert [int(g[0]) for g in top] == [19, 18, 17, 16]
    only = np.zeros(20)
    only[7] = 5.0
    picks = ops.roulette_selection(rng, pop, only, 50, {})
    assert len(picks) == 50 and all(int(g[0]) == 7 for g in picks)
    for kind in ("tournament_selection_classic", "tournament_selection_tuned"):
        out = ops.SELECTIONS[kind](rng, pop, fit + 1, 8, {"probability": 0.9})
        assert len(out) == 8
    # tournaments favour the fit: mean pick well above the population mean
    # (4 picks from 20: tournaments of 5)
    out = [g for _ in range(100)
           for g in ops.tournament_selection_classic(rng, pop, fit + 1, 4, {"probability": 0.9})]
    assert np.mean([g[0] for g in out]) > 13.0


def test_mutations_keep_budget_cap_and_spacing():
    rng = np.random.default_rng(4)
    c = cfg(mutations={
        "mutate_dose_value": {"mut_prob": 0.1, "min_value": 0.25, "max_value": 10.0,
                              "step_value": 0.25},
        "mutate_time_value": {"mut_prob": 0.2},
        "mutate_swap": {"mut_prob": 0.05},
        "mutate_split": {"mut_prob": 0.5},
    })
    pop = ops.random_population(rng, 100, N, c)
    for name, fn in ops.MUTATIONS.items():
        out = fn(rng, pop, c)
        assert len(out) == len(pop)
        for g in out:
            g = np.asarray(g)
            assert g.sum() <= c["max_value"] + 1e-9, name
            assert g.max() <= c["max_dose_value"] + 1e-9, name
            assert (gaps(g) >= 6).all(), name
    m = ops.mutate(rng, pop, dict(c, no_of_retaining_parents=10), 1)
    assert all(np.array_equal(a, b) for a, b in zip(m[:10], pop[:10]))


def test_next_generation_sizes_and_retained_parents():
    rng = np.random.default_rng(5)
    c = cfg()
    pop = ops.random_population(rng, 100, N, c)
    fit = rng.random(100)
    nxt = ops.next_generation(rng, pop, fit, c)
    assert len(nxt) == 100
    best = np.argsort(fit)[::-1][:10]
    for k, i in enumerate(best):
        np.testing.assert_array_equal(nxt[k], pop[i])
    assert len(ops.next_generation(rng, pop, fit, c, pop_size=1000)) == 1000


# ---------------------------------------------------------------------------
# Drivers against a deterministic fitness
# ---------------------------------------------------------------------------
W = 1.0 + np.arange(N) / N          # later doses count more


def fake_counts(g, t):
    g = np.asarray(g, float)
    return 800.0 - 12.0 * (1 + 0.1 * t) * (g * W).sum() + 4.0 * np.count_nonzero(g)


class FakeSim:
    """The SimulatorFitness interface over fake_counts (T tumours)."""

    def __init__(self, T=3):
        self.T, self.calls, self.evaluations = T, 0, 0

    def mean_counts(self, genomes):
        self.calls += 1
        self.evaluations += len(genomes)
        return np.array([[fake_counts(g, t) for t in range(self.T)] for g in genomes])

    def __call__(self, genomes):
        return D.FITNESS_OFFSET - self.mean_counts(genomes).mean(axis=1)


class FakeModel:
    """A perfect surrogate: its rows are the doses and the tumour id."""

    def features(self, g, desc, pi, ti):
        g = np.asarray(g, np.float32)
        return np.concatenate([g[pi], np.asarray(desc, np.float32)[ti]], axis=1)

    def predict(self, X):
        X = np.asarray(X)
        t = np.rint(X[:, -1]).astype(int)  # descriptor column holds the tumour id
        return np.array([fake_counts(x[:N], ti) for x, ti in zip(X, t)])


def test_plain_ga_improves_and_counts_evaluations():
    rng = np.random.default_rng(6)
    sim = FakeSim()
    hist = D.run_plain_ga(rng, sim, cfg(), pop_size=40, generations=15)
    assert len(hist) == 16 and hist[-1]["evaluations"] == 40 * 16 == sim.evaluations
    so_far = [h["best_so_far"] for h in hist]
    assert all(b >= a for a, b in zip(so_far, so_far[1:]))
    assert so_far[-1] > so_far[0]


def test_surrogate_ga_beats_plain_at_equal_evaluations():
    desc = np.zeros((3, 8), np.float32)
    desc[:, -1] = np.arange(3)
    c = cfg()
    plain = D.run_plain_ga(np.random.default_rng(7), FakeSim(), c, pop_size=40, generations=20)
    sim = FakeSim()
    hist = D.run_surrogate_ga(np.random.default_rng(7), sim, FakeModel(), desc, c, pop_size=40,
                              generations=20, pool=400, verify=40)
    assert sim.evaluations == 40 * 21
    assert all(h["surrogate_mae"] < 1e-6 for h in hist[1:])
    assert hist[-1]["best_so_far"] >= plain[-1]["best_so_far"]
    # a target stops the run early
    sim2 = FakeSim()
    h2 = D.run_surrogate_ga(np.random.default_rng(7), sim2, FakeModel(), desc, c, pop_size=40,
                            generations=20, pool=400, verify=40,
                            target=plain[-1]["best_so_far"])
    assert h2[-1]["best_so_far"] >= plain[-1]["best_so_far"] and len(h2) <= len(hist)
