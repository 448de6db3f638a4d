"""emt6ro.ga.operators against vectors written by the reference GA's own
operators (python_ga/utils.py, crossover.py, mutation.py, selection.py;
tests/golden/make_ga_vectors.py).  Random operators are driven by a scripted
uniform stream that the golden script also fed to the reference through
numpy's global random functions: equal outputs and an equal number of draws
mean both made the same decisions in the same order.  CPU only."""
import gzip
import json

import numpy as np
import pytest

from emt6ro.ga import operators as ops

from conftest import golden

with gzip.open(golden("ga_vectors.json.gz"), "rt") as _f:
    V = json.load(_f)
CFG = V["config"]


class ScriptedRng:
    """numpy Generator stand-in: draw k is default_rng(seed).random()[k],
    mapped as in the golden script (integers(lo, hi) = lo + floor(u (hi -
    lo)), random() = uniform() = u, choice(a) = a[floor(u len(a))])."""

    def __init__(self, seed):
        self.u = np.random.default_rng(seed).random(200000)
        self.k = 0

    def _next(self):
        v = float(self.u[self.k])
        self.k += 1
        return v

    def integers(self, lo, hi=None, size=None):
        if hi is None:
            lo, hi = 0, lo
        lo, hi = int(lo), int(hi)
        if size is None:
            return lo + int(self._next() * (hi - lo))
        return np.array([lo + int(self._next() * (hi - lo)) for _ in range(size)])

    def random(self, size=None):
        if size is None:
            return self._next()
        return np.array([self._next() for _ in range(size)])

    def uniform(self, low=0.0, high=1.0, size=None):
        assert size is None and low == 0.0 and high == 1.0
        return self._next()

    def choice(self, a):
        a = np.asarray(a)
        return a[int(self._next() * len(a))]


def cfg():
    import copy
    return copy.deepcopy(CFG)


def arr(x):
    return np.asarray(x, dtype=float)


# ---------------------------------------------------------------------------
# Deterministic functions
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("k", range(len(V["deterministic"]["normalize_crossover"])))
def test_normalize_crossover(k):
    c = V["deterministic"]["normalize_crossover"][k]
    got = ops.normalize_budget(arr(c["genome"]), cfg())
    assert np.array_equal(got, arr(c["out"]))


def test_annealing_schedule():
    for c in V["deterministic"]["annealing"]:
        got = ops.annealing_probability(c["iteration"], c["max_value"], c["max_iter"])
        assert got == c["out"], c


def test_selection_sizes():
    for c in V["deterministic"]["round_select_n"]:
        assert ops.even_count(c["select_n"], c["pop_size"]) == c["out"], c


def test_simple_selection():
    c = V["deterministic"]["simple_selection"]
    pop = [arr(g) for g in c["population"]]
    got = ops.simple_selection(None, pop, np.asarray(c["fitness"]), c["select_n"], cfg())
    assert [list(g) for g in got] == c["out"]


# ---------------------------------------------------------------------------
# Random operators on the scripted stream
# ---------------------------------------------------------------------------
def _run(case):
    rng = ScriptedRng(case["seed"])
    a = case["args"]
    op = case["op"]
    c = cfg()
    if op == "get_random_protocol":
        n_slots = len(a[0])
        res = ops.random_protocol(rng, n_slots)
        return [[int(s), float(d)] for s, d in res], rng.k
    if op == "refine":
        return ops.refine_to_time_constraint(rng, arr(a[0]), 6, c), rng.k
    if op in ops.CROSSOVERS:
        x1, x2 = ops.CROSSOVERS[op](rng, arr(a[0]), arr(a[1]), c)
        return [x1, x2], rng.k
    if op in ops.MUTATIONS:
        return ops.MUTATIONS[op](rng, [arr(g) for g in a[0]], c), rng.k
    if op in ops.SELECTIONS:
        sel = {"probability": 0.9, "candidates_dispersion": bool(case["seed"] % 2)}
        pop = [arr(g) for g in a[0]]
        return ops.SELECTIONS[op](rng, pop, np.asarray(a[1]), 6, sel), rng.k
    raise KeyError(op)


def _norm(x):
    if isinstance(x, np.ndarray):
        return [_norm(v) for v in x.tolist()]
    if isinstance(x, (list, tuple)):
        return [_norm(v) for v in x]
    return float(x)


@pytest.mark.parametrize("i", range(len(V["random"])),
                         ids=[f"{c['op']}-{c['seed']}" for c in V["random"]])
def test_random_operator_matches_reference(i):
    case = V["random"][i]
    got, draws = _run(case)
    assert _norm(got) == _norm(case["out"])
    assert draws == case["draws"]


def test_vectors_cover_every_operator():
    ops_seen = {c["op"] for c in V["random"]}
    assert set(ops.CROSSOVERS) <= ops_seen
    assert set(ops.MUTATIONS) <= ops_seen
    assert set(ops.SELECTIONS) - {"simple_selection"} <= ops_seen
    assert {"get_random_protocol", "refine"} <= ops_seen
