"""Dose-protocol GA drivers: the reference's loop with simulator fitness
(python_ga/genetic_algorithm.py:161-252) and a surrogate-screened variant.

Fitness is EMT6RoModel's (python/emt6ro/ga_model/model.py:36-54): 1500 minus
the mean final living-cell count over tumours and replicates.  The simulator
runs on the GPU through emt6ro.batch.Batch: one batch per population size,
re-dosed (`set_doses`) and re-seeded (`reset(epoch)`) for every evaluation, so
each call draws fresh replicate streams like the reference's Experiment.
"""
import time

import numpy as np

from .. import surrogate as S
from ..batch import Batch
from . import operators as ops

HOUR_STEPS = ops.HOUR_STEPS
FITNESS_OFFSET = 1500.0


class SimulatorFitness:
    """Fitness of genomes (doses per 0.5 h slot) on `tumors`, `reps`
    replicates each, over `hours` of simulated time.

    Simulation b of a call with P genomes is (protocol b // (T R), tumour
    (b // R) % T, replicate b % R), the Experiment layout (backend.cc:59-73).
    """

    def __init__(self, params, tumors, reps, hours=144, gpu_id=0, seed=0, res=300):
        self.params, self.tumors = params, list(tumors)
        self.T, self.R = len(self.tumors), int(reps)
        self.horizon = int(hours * HOUR_STEPS)
        self.res = int(res)
        self.n_slots = (self.horizon + self.res - 1) // self.res
        self.gpu_id, self.seed = gpu_id, seed
        self._batches = {}
        self._epoch = 0
        self.evaluations = 0          # genomes evaluated
        self.simulations = 0          # simulations run
        self.seconds = 0.0            # wall time inside the simulator

    def _batch(self, P):
        b = self._batches.get(P)
        if b is None:
            n = P * self.T * self.R
            i = np.arange(n)
            b = Batch(self.params, self.tumors, (i // self.R) % self.T, i // (self.T * self.R),
                      np.full(n, self.horizon, np.uint32),
                      np.zeros((P, self.n_slots), np.float32), self.res, self.horizon,
                      gpu_id=self.gpu_id, seed=self.seed)
            self._batches[P] = b
        return b

    def counts(self, genomes):
        """(P, T, R) final living-cell counts."""
        t0 = time.perf_counter()
        P = len(genomes)
        doses = np.zeros((P, self.n_slots), np.float32)
        for p, g in enumerate(genomes):
            g = np.asarray(g, np.float32)
            doses[p, :min(len(g), self.n_slots)] = g[:self.n_slots]
        b = self._batch(P)
        self._epoch += 1
        b.reset(self._epoch)
        b.set_doses(doses)
        b.run(self.horizon)
        out = b.results().astype(np.float64).reshape(P, self.T, self.R)
        self.seconds += time.perf_counter() - t0
        self.evaluations += P
        self.simulations += P * self.T * self.R
        return out

    def mean_counts(self, genomes):
        """(P, T) mean final living count per genome and tumour."""
        return self.counts(genomes).mean(axis=2)

    def __call__(self, genomes):
        return FITNESS_OFFSET - self.counts(genomes).mean(axis=(1, 2))


def _record(history, t0, gen, fitness, population, evals):
    k = int(np.argmax(fitness))
    best = float(fitness[k])
    prev = history[-1]["best_so_far"] if history else -np.inf
    history.append({"generation": gen, "seconds": time.perf_counter() - t0,
                    "best": best, "mean": float(np.mean(fitness)),
                    "best_so_far": max(prev, best), "evaluations": evals,
                    "best_genome": np.asarray(population[k]).tolist()})


def run_plain_ga(rng, fitness, config, pop_size=100, generations=50, n_slots=144,
                 time_limit=None, log=None):
    """The reference loop: evaluate, then per generation selection +
    crossover (next_generation), mutation, evaluation, until max_iter
    generations or stop_fitness (genetic_algorithm.py:210)."""
    config = dict(config, max_iter=generations)
    t0 = time.perf_counter()
    population = ops.random_population(rng, pop_size, n_slots, config)
    fit = fitness(population)
    history = []
    evals = len(population)
    _record(history, t0, 0, fit, population, evals)
    gen = 0
    while gen < generations and max(fit) < config["stop_fitness"]:
        if time_limit is not None and time.perf_counter() - t0 > time_limit:
            break
        gen += 1
        population = ops.next_generation(rng, population, fit, config)
        population = ops.mutate(rng, population, config, gen)
        fit = fitness(population)
        evals += len(population)
        _record(history, t0, gen, fit, population, evals)
        if log:
            log(f"plain gen {gen}: best {history[-1]['best']:.2f} "
                f"so far {history[-1]['best_so_far']:.2f} ({history[-1]['seconds']:.1f} s)")
    return history


def _spearman(a, b):
    """Rank correlation of two samples (ties broken by order)."""
    ra = np.argsort(np.argsort(a)).astype(np.float64)
    rb = np.argsort(np.argsort(b)).astype(np.float64)
    if ra.std() == 0 or rb.std() == 0:
        return 0.0
    return float(np.corrcoef(ra, rb)[0, 1])


def _rows(model, g, descriptors):
    """Input rows of `model` for every (genome, tumour) pair, genome-major:
    the model's own `features` method, else the module-level one."""
    T = len(descriptors)
    pi, ti = np.repeat(np.arange(len(g)), T), np.tile(np.arange(T), len(g))
    f = getattr(model, "features", None)
    return f(g, descriptors, pi, ti) if f is not None else S.features(g, descriptors, pi, ti)


def surrogate_fitness(model, genomes, descriptors):
    """Predicted fitness: 1500 minus the predicted mean count over tumours."""
    g = np.asarray(genomes, np.float32)
    X = _rows(model, g, descriptors)
    return FITNESS_OFFSET - model.predict(X).reshape(len(g), len(descriptors)).mean(axis=1)


def run_surrogate_ga(rng, fitness, model, descriptors, config, pop_size=100, generations=50,
                     n_slots=144, pool=10000, verify=200, refit_steps=0, time_limit=None,
                     target=None, log=None):
    """GA with surrogate screening.  Each generation breeds a pool of
    candidates from the population with the reference's selection, crossover
    and mutation operators, ranks the pool by the surrogate, simulates the
    top `verify`, and keeps the pop_size best simulated genomes of the
    retained parents and the verified children as the next population.  With
    refit_steps > 0 the surrogate is fine-tuned on every simulated genome so
    far after each generation.  Stops at `generations`, `time_limit` seconds
    or once the best simulated fitness reaches `target`."""
    config = dict(config, max_iter=generations)
    t0 = time.perf_counter()
    population = ops.random_population(rng, pop_size, n_slots, config)
    fit = fitness(population)
    history = []
    evals = len(population)
    _record(history, t0, 0, fit, population, evals)
    seen_x, seen_y = [], []
    err = []
    gen = 0
    while gen < generations:
        if time_limit is not None and time.perf_counter() - t0 > time_limit:
            break
        if target is not None and history[-1]["best_so_far"] >= target:
            break
        gen += 1
        cands = ops.next_generation(rng, population, fit, config, pop_size=pool)
        cands = ops.mutate(rng, cands, config, gen)
        pred = surrogate_fitness(model, cands, descriptors)
        top = np.argsort(-pred)[:verify]
        chosen = [cands[i] for i in top]
        counts = fitness.mean_counts(chosen) if hasattr(fitness, "mean_counts") else None
        if counts is not None:
            vfit = FITNESS_OFFSET - counts.mean(axis=1)
        else:
            vfit = fitness(chosen)
        evals += len(chosen)
        err.append(float(np.mean(np.abs(pred[top] - vfit))))
        rank_corr = _spearman(pred[top], vfit)
        keep_n = config.get("no_of_retaining_parents", 0)
        order = np.argsort(-np.asarray(fit))[:max(keep_n, 1)]
        merged = [population[i] for i in order] + chosen
        mfit = np.concatenate([np.asarray(fit)[order], vfit])
        best = np.argsort(-mfit)[:pop_size]
        population = [merged[i] for i in best]
        fit = mfit[best]
        _record(history, t0, gen, vfit, chosen, evals)
        history[-1]["surrogate_mae"] = err[-1]
        history[-1]["screen_rank_corr"] = rank_corr
        if log:
            log(f"surrogate gen {gen}: best {history[-1]['best']:.2f} so far "
                f"{history[-1]['best_so_far']:.2f} screen MAE {err[-1]:.2f} "
                f"({history[-1]['seconds']:.1f} s)")
        if refit_steps and counts is not None and hasattr(model, "refit"):
            seen_x.append(_rows(model, np.asarray(chosen, np.float32), descriptors))
            seen_y.append(counts.reshape(-1))
            model.refit(np.concatenate(seen_x), np.concatenate(seen_y), refit_steps)
    return history
