#!/usr/bin/env python3
"""Writes ga_vectors.json.gz: input/output vectors of the reference GA's
operators (python_ga/utils.py, crossover.py, mutation.py, selection.py),
run from /root/reference in the build container (those four modules import
only numpy, yaml and os).

The reference draws from numpy's global random functions.  For the random
operators this script replaces np.random.randint / random / uniform / choice
with a scripted source: draw k is u_k = default_rng(seed).random()[k], mapped
as randint(lo, hi) = lo + floor(u (hi - lo)), random() = uniform() = u,
choice(a) = a[floor(u len(a))].  tests/test_ga_vectors.py feeds the same
stream to emt6ro.ga.operators through a Generator stand-in with the same
mapping, so equal outputs (and an equal number of draws) mean the two make
the same decisions in the same order.

Usage: python tests/golden/make_ga_vectors.py  (needs /root/reference)
"""
import copy
import gzip
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_GA = "/root/reference/python_ga"


class Script:
    def __init__(self, seed):
        self.u = np.random.default_rng(seed).random(200000)
        self.k = 0

    def next(self):
        v = float(self.u[self.k])
        self.k += 1
        return v

    def randint(self, lo, hi=None, size=None):
        if hi is None:
            lo, hi = 0, lo
        lo, hi = int(lo), int(hi)
        if size is None:
            return lo + int(self.next() * (hi - lo))
        return np.array([lo + int(self.next() * (hi - lo)) for _ in range(size)])

    def random(self, size=None):
        if size is None:
            return self.next()
        return np.array([self.next() for _ in range(size)])

    def uniform(self, low=0.0, high=1.0, size=None):
        assert size is None and low == 0.0 and high == 1.0
        return self.next()

    def choice(self, a):
        a = np.asarray(a)
        return a[int(self.next() * len(a))]


def patch(script):
    np.random.randint = script.randint
    np.random.random = script.random
    np.random.uniform = script.uniform
    np.random.choice = script.choice


def jl(x):
    """JSON-ready copy (lists of floats)."""
    if isinstance(x, np.ndarray):
        return [jl(v) for v in x.tolist()]
    if isinstance(x, (list, tuple)):
        return [jl(v) for v in x]
    if isinstance(x, (np.floating, float)):
        return float(x)
    if isinstance(x, (np.integer, int)):
        return int(x)
    return x


def genome(rng, n=240, doses=8, spread=True):
    g = np.zeros(n)
    idx = rng.choice(n, size=doses, replace=False) if spread else \
        np.sort(rng.choice(40, size=doses, replace=False)) + 10
    g[idx] = rng.integers(1, 11, size=doses) * 0.25
    return g


def main():
    sys.path.insert(0, REF_GA)
    import crossover as rc  # noqa: E402
    import mutation as rm  # noqa: E402
    import selection as rs  # noqa: E402
    import utils as ru  # noqa: E402
    saved = (np.random.randint, np.random.random, np.random.uniform, np.random.choice)
    cfg = {"step_value": 0.25, "max_value": 10.0, "max_dose_value": 2.5, "select_n": 0.5,
           "time_interval_hours": 3.0, "protocol_resolution": 300,
           "crossover": {"cross_point_range_percentage": [25, 75],
                         "cross_points_range_percentage": [25, 50, 75]},
           "mutations": {"mutate_dose_value": {"mut_prob": 0.05, "min_value": 0.25,
                                               "max_value": 10.0, "step_value": 0.25},
                         "mutate_time_value": {"mut_prob": 0.05},
                         "mutate_swap": {"mut_prob": 0.03},
                         "mutate_split": {"mut_prob": 0.5}}}
    inp = np.random.default_rng(2024)
    out = {"config": cfg, "mapping": "randint(lo,hi)=lo+floor(u*(hi-lo)); random()=uniform()="
           "u; choice(a)=a[floor(u*len(a))]; u_k = default_rng(seed).random(200000)[k]",
           "deterministic": {}, "random": []}

    # --- deterministic functions ----------------------------------------
    det = out["deterministic"]
    det["normalize_crossover"] = []
    for k in range(12):
        g = genome(inp, doses=int(inp.integers(3, 16)))
        g[g > 0] = inp.integers(1, 11, size=int((g > 0).sum())) * 0.25
        det["normalize_crossover"].append(
            {"genome": jl(g), "out": jl(rc.normalize_crossover(list(g), cfg))})
    det["annealing"] = [
        {"iteration": it, "max_value": mv, "max_iter": mi,
         "out": float(ru.calculate_probability_annealing(it, max_value=mv, max_iter=mi))}
        for it in (1, 2, 10, 25, 50, 75, 99, 100, 101) for mv in (0.5, 0.1, 0.05)
        for mi in (50, 100)]
    det["round_select_n"] = [
        {"select_n": sn, "pop_size": ps, "out": rs.round_select_n(sn, ps)}
        for sn in (0.1, 0.25, 0.4, 0.5, 0.55, 0.9) for ps in (7, 10, 20, 33, 40, 100)]
    pop = [genome(inp) for _ in range(10)]
    fit = np.round(inp.random(10) * 500 + 1000, 3)
    det["simple_selection"] = {"population": jl(pop), "fitness": jl(fit), "select_n": 4,
                               "out": jl(rs.simple_selection(pop, fit, 4, cfg))}

    # --- random operators with the scripted source ------------------------
    def case(name, seed, fn, args):
        script = Script(seed)
        patch(script)
        try:
            res = fn(*copy.deepcopy(args))
        finally:
            np.random.randint, np.random.random, np.random.uniform, np.random.choice = saved
        out["random"].append({"op": name, "seed": seed, "args": jl(args), "out": jl(res),
                              "draws": script.k})

    time_steps = np.arange(0, 240 * 300, 300)
    for s in range(4):
        case("get_random_protocol", 100 + s,
             lambda ts: ru.get_random_protocol(time_steps=np.asarray(ts),
                                               time_interval_steps=1800), [time_steps])
    for s in range(6):
        g = genome(inp, doses=int(inp.integers(4, 10)), spread=False)
        case("refine", 200 + s,
             lambda g_: ru.refine_genome_around_cross_point_to_time_constraint(
                 genome=list(g_), interval_in_indices=6, config=cfg), [g])
    for s in range(4):
        x1, x2 = genome(inp), genome(inp)
        case("cross_one_point", 300 + s, lambda a, b: rc.cross_one_point(list(a), list(b), cfg),
             [x1, x2])
        case("cross_two_points", 310 + s,
             lambda a, b: rc.cross_two_points(list(a), list(b), cfg), [x1, x2])
        case("cross_uniform", 320 + s, lambda a, b: rc.cross_uniform(list(a), list(b), cfg),
             [x1, x2])
        case("normalized_crossover", 330 + s,
             lambda a, b: rc.normalized_crossover(list(a), list(b), cfg), [x1, x2])
    for s in range(3):
        p = [genome(inp, doses=int(inp.integers(2, 8))) for _ in range(5)]
        case("mutate_swap", 400 + s, lambda q: rm.mutate_swap([list(g) for g in q], cfg), [p])
        case("mutate_time_value", 410 + s,
             lambda q: rm.mutate_time_value([np.array(g) for g in q], cfg), [p])
        case("mutate_dose_value", 420 + s,
             lambda q: rm.mutate_dose_value([np.array(g) for g in q], cfg), [p])
        case("mutate_split", 430 + s, lambda q: rm.mutate_split([list(g) for g in q], cfg), [p])
    for s in range(3):
        pop = [genome(inp) for _ in range(20)]
        fit = np.round(inp.random(20) * 500 + 1000, 3)
        sel = {"probability": 0.9, "candidates_dispersion": bool(s % 2)}
        case("roulette_selection", 500 + s,
             lambda q, f: rs.roulette_selection(q, np.asarray(f), 6, sel), [pop, fit])
        case("tournament_selection_classic", 510 + s,
             lambda q, f: rs.tournament_selection_classic(q, np.asarray(f), 6, sel), [pop, fit])
        case("tournament_selection_tuned", 520 + s,
             lambda q, f: rs.tournament_selection_tuned(q, np.asarray(f), 6, sel), [pop, fit])

    with gzip.open(os.path.join(HERE, "ga_vectors.json.gz"), "wt") as f:
        json.dump(out, f)
    print(f"{len(out['random'])} random cases, "
          f"{sum(len(v) if isinstance(v, list) else 1 for v in det.values())} deterministic")


if __name__ == "__main__":
    main()
