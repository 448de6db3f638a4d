"""Dose-protocol genetic algorithm operators, restated from the reference GA
(python_ga/utils.py, selection.py, crossover.py, mutation.py,
genetic_algorithm.py) with an explicit numpy Generator instead of the global
random state.

A genome is a float array of doses (Gy) per protocol slot (300 simulation
steps = 0.5 h); `pairs` are (simulation step, dose) lists as the simulator's
add_irradiations takes them.

Config keys (the reference's YAML names, python_ga/experiment_config):
step_value, max_value (total budget, Gy), max_dose_value (per dose),
select_n, retained_parent_protocols, time_interval_hours, protocol_resolution,
crossover {type, cross_point_range_percentage, cross_points_range_percentage},
selection {type, probability, candidates_dispersion},
mutations {name: {mut_prob | mut_prob_max + mut_type 'annealing', ...}},
max_iter.
"""
import numpy as np

HOUR_STEPS = 600


def default_config(**over):
    """The base experiment (experiment_config/base__tournament__dose_value_
    time_value.yaml) with retained parents 0.1 and a 3 h dose interval."""
    cfg = {
        "step_value": 0.25, "max_value": 10.0, "max_dose_value": 2.5, "select_n": 0.5,
        "retained_parent_protocols": 0.1, "time_interval_hours": 3.0,
        "protocol_resolution": 300, "max_iter": 100, "stop_fitness": 1e6,
        "crossover": {"type": "normalized_crossover"},
        "selection": {"type": "tournament_selection_classic", "probability": 0.9},
        "mutations": {
            "mutate_dose_value": {"mut_prob": 0.05, "min_value": 0.25, "max_value": 10.0,
                                  "step_value": 0.25},
            "mutate_time_value": {"mut_prob": 0.05},
        },
    }
    cfg.update(over)
    return cfg


# ---------------------------------------------------------------------------
# Representation (utils.py:131-160)
# ---------------------------------------------------------------------------
def pairs_to_genome(pairs, n_slots, res=300):
    g = np.zeros(n_slots)
    for step, dose in pairs:
        g[int(step // res)] = dose
    return g


def genome_to_pairs(genome, res=300):
    return [(int(i) * res, float(genome[i])) for i in np.nonzero(np.asarray(genome))[0]]


# ---------------------------------------------------------------------------
# Time-interval constraint (utils.py:7-62)
# ---------------------------------------------------------------------------
def assign_dose_within_time_constraint(rng, dose_steps, time_steps, interval_steps):
    """A random allowed step at least interval_steps away from every dose."""
    allowed = np.asarray(time_steps)
    top = int(np.max(time_steps))
    for s in dose_steps:
        lo, hi = max(int(s) - interval_steps, 0), min(int(s) + interval_steps, top)
        allowed = allowed[(allowed < lo) | (allowed > hi)]
    if len(allowed):
        return int(rng.choice(allowed))
    return None


def refine_to_time_constraint(rng, genome, interval_in_indices, config):
    """Moves doses closer than the interval to random allowed slots
    (refine_genome_around_cross_point_to_time_constraint, utils.py:27-62).
    The new slot is drawn from the genome's own slots; the reference draws
    from 240 slots (the 120 h default of utils.py:9) whatever the genome's
    length, which is the same for its 240-slot genomes."""
    genome = np.array(genome, dtype=float)
    nz = np.argwhere(genome > 0)[:, 0]
    gaps = np.diff(nz)
    if not (gaps < interval_in_indices).any():
        return genome
    wrong = np.argwhere(gaps < interval_in_indices + 1)[:, 0]
    keep = np.delete(nz, wrong)
    res = config["protocol_resolution"]
    time_steps = np.arange(0, len(genome) * res, res)
    for pos in wrong:
        step = assign_dose_within_time_constraint(
            rng, keep * res, time_steps, int(config["time_interval_hours"] * HOUR_STEPS))
        cur = nz[pos]
        if step is not None:
            idx = int(step // res)
            genome[idx] = genome[cur]
            keep = np.append(keep, idx)
        genome[cur] = 0
    return genome


def random_protocol(rng, n_slots, min_dose=0.25, max_dose=10.0, step_value=0.25,
                    max_dose_value=2.5, time_interval_hours=3.0, res=300):
    """get_random_protocol (utils.py:66-98) as (step, dose) pairs."""
    time_steps = np.arange(0, n_slots * res, res)
    interval = int(HOUR_STEPS * time_interval_hours)
    total, limit, pairs, steps = 0.0, max_dose_value, [], []
    while total < max_dose:
        if max_dose - total < max_dose_value:
            limit = max_dose - total
        upper = int(round((limit - min_dose) / step_value)) + 1
        dose = int(rng.integers(0, upper)) * step_value + min_dose
        step = assign_dose_within_time_constraint(rng, steps, time_steps, interval)
        if step is None:
            break
        steps.append(step)
        pairs.append((step, dose))
        total += dose
    return pairs


def random_population(rng, n, n_slots, config):
    return [pairs_to_genome(random_protocol(rng, n_slots, step_value=config["step_value"],
                                            max_dose=config["max_value"],
                                            max_dose_value=config["max_dose_value"],
                                            time_interval_hours=config["time_interval_hours"],
                                            res=config["protocol_resolution"]),
                            n_slots, config["protocol_resolution"])
            for _ in range(n)]


# ---------------------------------------------------------------------------
# Selection (selection.py)
# ---------------------------------------------------------------------------
def even_count(fraction, pop_size):
    """round(pop_size * fraction) made even by rounding down (selection.py's
    selection-size rule)."""
    n = round(pop_size * fraction)
    return n - (n % 2)


def _tournament_size(pop_size, select_n):
    """Candidates per tournament: the population over the selections,
    rounded half to even like Python's round()."""
    return round(pop_size / select_n)


def roulette_selection(rng, population, fitness, select_n, config):
    fitness = np.asarray(fitness, float)
    if config.get("candidates_dispersion"):
        disp = fitness.max() - fitness.min()
        ratio = disp / fitness.mean()
        fitness = fitness - fitness.min() * (1 - ratio)
    wheel = np.cumsum(fitness) / fitness.sum()
    out = []
    for _ in range(select_n):
        r = rng.random()
        j = int(np.searchsorted(wheel, r, side="right"))
        if j < len(population):
            out.append(population[j])
    return out


def simple_selection(rng, population, fitness, select_n, config):
    order = np.argsort(fitness)[::-1]
    return [population[i] for i in order][:select_n]


def tournament_selection_classic(rng, population, fitness, select_n, config):
    p = config["probability"]
    fitness = np.asarray(fitness)
    k = _tournament_size(len(population), select_n)
    out = []
    for _ in range(select_n):
        cand = [int(rng.integers(0, len(population))) for _ in range(k)]
        probs = [p * (1 - p) ** j for j in range(k)]
        order = [cand[i] for i in np.argsort(fitness[cand])[::-1]]
        ranked = [population[i] for i in order]
        pick = roulette_selection(rng, ranked, probs, 1, {"candidates_dispersion": False})
        out.append(pick[0] if pick else ranked[-1])
    return out


def tournament_selection_tuned(rng, population, fitness, select_n, config):
    p = config["probability"]
    k = _tournament_size(len(population), select_n)
    out = []
    for i in range(select_n):
        sp = p * (1 - p) ** i
        best = None
        for _ in range(k):
            c = int(rng.integers(0, len(population)))
            if best is None or fitness[c] > fitness[best]:
                best = c
        if rng.random() < sp:
            out.append(population[best])
        else:
            pick = roulette_selection(rng, population, fitness, 1, config)
            out.append(pick[0] if pick else population[best])
    return out


SELECTIONS = {"simple_selection": simple_selection, "roulette_selection": roulette_selection,
              "tournament_selection_classic": tournament_selection_classic,
              "tournament_selection_tuned": tournament_selection_tuned}


# ---------------------------------------------------------------------------
# Crossover (crossover.py)
# ---------------------------------------------------------------------------
def normalize_budget(genome, config):
    """normalize_crossover: take step_value off the largest doses until the
    total fits max_value."""
    genome = np.array(genome, dtype=float)
    step, cap = config["step_value"], config["max_value"]
    while genome.sum() > cap + 1e-9:
        for idx in np.argsort(genome)[::-1]:
            if genome[idx] < step:
                break
            genome[idx] -= step
            if genome.sum() <= cap + 1e-9:
                break
    return genome


def _assign_with_constraint(src, dst, idx, interval, last):
    if src[idx] == 0:
        dst.append(0.0)
    elif last is None or last + interval < idx:
        dst.append(float(src[idx]))
        last = idx
    else:
        dst.append(0.0)
    return last


def cross_one_point(rng, x1, x2, config):
    interval = int(2 * config["time_interval_hours"])
    n = len(x1)
    lo, hi = 1, n
    rng_pct = config["crossover"].get("cross_point_range_percentage")
    if rng_pct:
        lo, hi = int(n * rng_pct[0] / 100), int(n * rng_pct[1] / 100)
    cp = int(rng.integers(lo, hi))
    a = np.concatenate([x1[:cp], x2[cp:]])
    b = np.concatenate([x2[:cp], x1[cp:]])
    a = refine_to_time_constraint(rng, a, interval, config)
    b = refine_to_time_constraint(rng, b, interval, config)
    return normalize_budget(a, config), normalize_budget(b, config)


def cross_two_points(rng, x1, x2, config):
    interval = int(2 * config["time_interval_hours"])
    n = len(x1)
    lo, hi = 1, n - 1
    mid = (hi - lo) // 2
    pct = config["crossover"].get("cross_points_range_percentage")
    if pct:
        lo, mid, hi = (int(n * pct[0] / 100), int(n * pct[1] / 100), int(n * pct[2] / 100))
    c1 = int(rng.integers(lo, mid))
    c2 = int(rng.integers(c1, hi))
    a, b = np.array(x1, float), np.array(x2, float)
    a[c1:c2], b[c1:c2] = x2[c1:c2], x1[c1:c2]
    a = refine_to_time_constraint(rng, a, interval, config)
    b = refine_to_time_constraint(rng, b, interval, config)
    return normalize_budget(a, config), normalize_budget(b, config)


def cross_uniform(rng, x1, x2, config):
    interval = int(2 * config["time_interval_hours"])
    a, b, la, lb = [], [], None, None
    for i in range(len(x1)):
        p = rng.uniform()
        la = _assign_with_constraint(x1 if p < 0.5 else x2, a, i, interval, la)
        lb = _assign_with_constraint(x2 if p < 0.5 else x1, b, i, interval, lb)
    return normalize_budget(a, config), normalize_budget(b, config)


def _group_replacement(x1, x2, group, config):
    interval = int(2 * config["time_interval_hours"])
    out, last = [], None
    for i in range(len(x1)):
        last = _assign_with_constraint(x2 if group[i] else x1, out, i, interval, last)
    return out


def normalized_crossover(rng, x1, x2, config):
    group = rng.integers(0, 2, size=len(x1)) == 0
    a = _group_replacement(x1, x2, group, config)
    b = _group_replacement(x2, x1, group, config)
    return normalize_budget(a, config), normalize_budget(b, config)


CROSSOVERS = {"cross_one_point": cross_one_point, "cross_two_points": cross_two_points,
              "cross_uniform": cross_uniform, "normalized_crossover": normalized_crossover}


# ---------------------------------------------------------------------------
# Mutation (mutation.py)
# ---------------------------------------------------------------------------
def mutate_dose_value(rng, population, config):
    """Raises random genes to the largest dose the budget allows until the
    budget is over 80 % used (mutation.py:6-37; the reference applies it to
    every genome and does not read mut_prob).  As there, the raises build up
    on the genome as drawn, and the result is the time-constraint refinement
    of its last state (each raise re-refines it).  Bounded at 10,000 draws
    per genome."""
    interval = int(2 * config["time_interval_hours"])
    mc = config["mutations"]["mutate_dose_value"]
    cap_one, cap, min_dose = config["max_dose_value"], mc["max_value"], mc["min_value"]
    out = []
    for g in population:
        g = np.array(g, float)
        result = g.copy()
        for _ in range(10000):
            idx = int(rng.integers(0, len(g)))
            limit = min(cap_one, cap - g.sum()) + g[idx]
            limit = min(limit, cap_one)
            if limit > min_dose:
                g[idx] = limit
                result = refine_to_time_constraint(rng, g, interval, config)
            elif limit == 0 or g.sum() > cap * 0.8:
                break
        out.append(np.asarray(result, float))
    return out


def mutate_time_value(rng, population, config):
    """mutation.py:40-68: per gene, one uniform; below mut_prob a non-zero
    dose moves to a random slot (capped at max_dose_value, the rest stays)."""
    interval = int(2 * config["time_interval_hours"])
    prob = config["mutations"]["mutate_time_value"]["mut_prob"]
    cap_one = config["max_dose_value"]
    out = []
    for g in population:
        g = np.array(g, float)
        for idx in range(len(g)):
            if rng.uniform() < prob and g[idx] > 0:
                t = int(rng.integers(len(g)))
                v = g[t] + g[idx]
                rem = 0.0
                if v > cap_one:
                    rem, v = v - cap_one, cap_one
                g[t] = v
                g[idx] = rem
        out.append(refine_to_time_constraint(rng, g, interval, config))
    return out


def mutate_swap(rng, population, config):
    """mutation.py:71-91: per gene, one uniform; below mut_prob the gene
    swaps with another one at a random offset."""
    interval = int(2 * config["time_interval_hours"])
    prob = config["mutations"]["mutate_swap"]["mut_prob"]
    out = []
    for g in population:
        g = np.array(g, float)
        for idx in range(len(g)):
            if rng.uniform() >= prob:
                continue
            k = (idx + int(rng.integers(len(g) - 1))) % len(g)
            g[idx], g[k] = g[k], g[idx]
        out.append(refine_to_time_constraint(rng, g, interval, config))
    return out


def mutate_split(rng, population, config):
    interval = int(2 * config["time_interval_hours"])
    prob = config["mutations"]["mutate_split"]["mut_prob"]
    step, cap_one = config["step_value"], config["max_dose_value"]
    out = []
    for g in population:
        g = np.array(g, float)
        if rng.uniform() < prob:
            nz = np.nonzero(g)[0]
            if nz.size:
                idx = int(rng.choice(nz))
                k = g[idx] / step
                split = int(rng.integers(0, max(int(k), 1)))
                d1 = split * step
                d2 = g[idx] - d1
                for d in (d1, d2):
                    for _ in range(len(population)):
                        j = int(rng.integers(len(g)))
                        if g[j] + d <= cap_one:
                            g[j] += d
                            break
                g[idx] = 0
        out.append(refine_to_time_constraint(rng, g, interval, config))
    return out


MUTATIONS = {"mutate_swap": mutate_swap, "mutate_dose_value": mutate_dose_value,
             "mutate_time_value": mutate_time_value, "mutate_split": mutate_split}


def annealing_probability(iteration, max_value=0.5, max_iter=100, eps=0.001, decimals=6):
    """calculate_probability_annealing (utils.py:208-216)."""
    x = 1 - (iteration - 1) / max_iter
    r = np.round((x ** 4) * max_value, decimals)
    return r if r >= eps else eps


# ---------------------------------------------------------------------------
# One generation (genetic_algorithm.py:30-155)
# ---------------------------------------------------------------------------
def next_generation(rng, population, fitness, config, pop_size=None):
    """Selection and crossover; returns pop_size genomes (default: the
    population's size), the retained parents first."""
    n = len(population)
    pop_size = n if pop_size is None else pop_size
    select_n = even_count(config["select_n"] - config["retained_parent_protocols"], n)
    config["no_of_retaining_parents"] = even_count(config["retained_parent_protocols"], n)
    keep = simple_selection(rng, population, fitness, config["no_of_retaining_parents"], config)
    chosen = SELECTIONS[config["selection"]["type"]](rng, population, fitness, select_n,
                                                      config["selection"])
    if not chosen:
        chosen = keep or list(population)
    new = [np.array(g, float) for g in keep]
    cross = CROSSOVERS[config["crossover"]["type"]]
    for _ in range(int((pop_size - len(keep)) / 2)):
        x1 = chosen[int(rng.integers(len(chosen)))]
        x2 = chosen[int(rng.integers(len(chosen)))]
        a, b = cross(rng, np.asarray(x1, float), np.asarray(x2, float), config)
        new.append(np.asarray(a, float))
        new.append(np.asarray(b, float))
    return new


def mutate(rng, population, config, iteration):
    keep = config.get("no_of_retaining_parents", 0)
    part = population[keep:]
    for name, mc in config["mutations"].items():
        if name not in MUTATIONS:
            continue
        if mc.get("mut_prob") == "annealing" or mc.get("mut_type") == "annealing":
            mc["mut_type"] = "annealing"
            mc["mut_prob"] = annealing_probability(iteration, mc["mut_prob_max"],
                                                   config["max_iter"])
        part = MUTATIONS[name](rng, part, config)
    return population[:keep] + list(part)
