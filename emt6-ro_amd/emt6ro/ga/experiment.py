"""Surrogate data generation, training and the plain-vs-surrogate GA
comparison, as a command line tool.

  python -m emt6ro.ga.experiment data --out data.npz [--protocols 1200]
  python -m emt6ro.ga.experiment compare --data data.npz [more.npz ...] --out result.json
  python -m emt6ro.ga.experiment config4 --data data.npz [more.npz ...] --out result.json

`data` simulates random dose protocols on synthetic 53x53 tumours (disk radii
8-14) and records, per (protocol, tumour), the surrogate features and the mean
and sd of the final living count over the replicates.  Protocols are drawn
half from the GA's own initial-population generator (python_ga/utils.py:
66-98) and half from the batched workload's generator (up to 10 doses of
0.25-1.25 Gy, total <= 10 Gy, first 72 h).

`compare` trains the surrogate on 90% of the samples, reports its MAE on the
other 10%, then runs the simulator-only GA and the surrogate-screened GA on
three of the tumours and reports best fitness against wall time.

`config4` trains the configured 152-input network (and, beside it, the GA's
feature network) at batch 4096 on 90% of the protocols and reports the
held-out error against the simulator.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from .. import simulation as sim
from .. import surrogate as S
from .. import workloads as wl
from ..batch import Batch
from . import driver as D
from . import operators as ops

HOURS = 144
N_SLOTS = S.N_DOSE_SLOTS
_ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.dirname(
    os.path.abspath(__file__)))))
DEFAULT_PARAMS = os.path.join(_ROOT, "tests", "golden", "default-parameters.json")


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def make_protocols(n, seed):
    """n genomes of N_SLOTS doses (see module docstring)."""
    rng = np.random.default_rng(seed)
    cfg = ops.default_config()
    half = n // 2
    ga = ops.random_population(rng, half, N_SLOTS, cfg)
    other = wl.dose_table(wl.random_protocols(n - half, seed + 1), N_SLOTS)
    return np.concatenate([np.asarray(ga, np.float32).reshape(-1, N_SLOTS), other])


def tumours(params, n, seed):
    ts = wl.synthetic_tumors(params, n, seed)
    return ts, np.stack([S.tumour_descriptor(t) for t in ts])


def simulate_means(params, tumors, genomes, reps, seed, epoch=0):
    """(P, T) mean and sd of final living counts, one batch of P*T*reps."""
    P, T = len(genomes), len(tumors)
    horizon = HOURS * wl.STEPS_PER_HOUR
    n_slots = horizon // wl.SLOT_STEPS
    doses = np.zeros((P, n_slots), np.float32)
    doses[:, :N_SLOTS] = genomes
    i = np.arange(P * T * reps)
    b = Batch(params, tumors, (i // reps) % T, i // (T * reps), np.full(len(i), horizon, np.uint32),
              doses, wl.SLOT_STEPS, horizon, gpu_id=0, seed=seed)
    b.reset(epoch)
    ms = b.run(horizon)
    c = b.results().astype(np.float64).reshape(P, T, reps)
    return c.mean(axis=2), c.std(axis=2, ddof=1), ms


def cmd_data(a):
    params = sim.load_parameters(a.params)
    ts, desc = tumours(params, a.tumours, a.seed)
    pseed = a.seed + 100 if a.protocol_seed is None else a.protocol_seed
    genomes = make_protocols(a.protocols, pseed)
    P, T = len(genomes), len(ts)
    mean = np.zeros((P, T))
    sd = np.zeros((P, T))
    t0 = time.perf_counter()
    for p0 in range(0, P, a.chunk):
        p1 = min(P, p0 + a.chunk)
        m, s, ms = simulate_means(params, ts, genomes[p0:p1], a.reps, a.seed,
                                  epoch=a.epoch_base + p0)
        mean[p0:p1], sd[p0:p1] = m, s
        th = (p1 - p0) * T * a.reps * HOURS
        log(f"[data] protocols {p0}-{p1}: {ms / 1e3:.1f} s GPU, {th / (ms / 1e3):.0f} "
            f"tumour-h/s, mean count {m.mean():.1f}")
        np.savez_compressed(a.out, genomes=genomes[:p1], descriptors=desc, mean=mean[:p1],
                            sd=sd[:p1], reps=a.reps, hours=HOURS,
                            seconds=time.perf_counter() - t0)
    log(f"[data] {P * T} samples, {P * T * a.reps} simulations in "
        f"{time.perf_counter() - t0:.1f} s -> {a.out}")


def load_samples(paths, params=None):
    """Features, targets (mean final count) and protocol index of the
    samples in one or more `data` files of the same tumours."""
    if isinstance(paths, (str, os.PathLike)):
        paths = [paths]
    zs = [np.load(p) for p in paths]
    desc = zs[0]["descriptors"]
    for z in zs[1:]:
        if not np.array_equal(z["descriptors"], desc):
            raise ValueError("data files hold different tumours")
    g = np.concatenate([z["genomes"] for z in zs])
    mean = np.concatenate([z["mean"] for z in zs])
    P, T = mean.shape
    pi = np.repeat(np.arange(P), T)
    ti = np.tile(np.arange(T), P)
    return S.features(g, desc, pi, ti, params), mean.reshape(-1), pi, zs[0]


def train_surrogate(X, y, pi, a, params=None):
    """Split by protocol (10% of protocols held out), fit, report MAE."""
    rng = np.random.default_rng(a.seed)
    P = pi.max() + 1
    held = rng.random(P) < 0.1
    te = held[pi]
    model = S.Surrogate(seed=a.seed, members=a.members, params=params)
    t0 = time.perf_counter()
    losses = model.fit(X[~te], y[~te], steps=a.train_steps, batch=a.batch, lr=a.lr,
                       seed=a.seed)
    fit_s = time.perf_counter() - t0
    pred = model.predict(X[te])
    mae = float(np.abs(pred - y[te]).mean())
    out = {"train_samples": int((~te).sum()), "test_samples": int(te.sum()),
           "train_steps": a.train_steps, "batch": a.batch, "members": a.members,
           "train_seconds": fit_s,
           "final_loss": float(losses[-50:].mean()), "mae": mae,
           "test_mean": float(y[te].mean()), "mae_fraction_of_mean": mae / float(y[te].mean()),
           "predict_the_mean_mae": float(np.abs(y[te] - y[~te].mean()).mean())}
    log(f"[surrogate] {out}")
    return model, out


def cmd_config4(a):
    """The surrogate-training workload as configured: the 152-512-512-512-1
    network on the raw input (144 dose slots in Gy + the 8 tumour
    descriptors), batch 4096, trained on 90% of the protocols of the `data`
    files and scored by its mean absolute error against the simulated mean
    on the other 10%; beside it the GA's feature network and the
    predict-the-mean baseline.  Each network is fitted from scratch for
    every step count in --steps."""
    params = sim.load_parameters(a.params)
    X, y, pi, z = load_samples(a.data, params)
    mean = np.concatenate([np.load(p)["mean"] for p in a.data])
    genomes = np.concatenate([np.load(p)["genomes"] for p in a.data]).astype(np.float32)
    desc = np.asarray(z["descriptors"], np.float32)
    P, T = mean.shape
    ti = np.tile(np.arange(T), P)
    Xr = np.concatenate([desc[ti], genomes[pi]], axis=1)
    rng = np.random.default_rng(a.seed)
    te = (rng.random(P) < 0.1)[pi]
    base = float(np.abs(y[te] - y[~te].mean()).mean())
    ym = float(y[te].mean())
    out = {"samples": int(len(y)), "train_samples": int((~te).sum()),
           "test_samples": int(te.sum()), "protocols": int(P), "tumours": int(T),
           "replicates": int(z["reps"]), "hours": int(z["hours"]), "batch": a.batch,
           "test_mean": ym, "predict_the_mean_mae": base, "runs": []}
    log(f"[config4] {len(y)} samples ({int((~te).sum())} train, {int(te.sum())} held out), "
        f"test mean {ym:.1f}, predict-the-mean MAE {base:.2f}")
    nets = [("raw 152 inputs", Xr, 1), ("raw 152 inputs", Xr, 5), ("71 features", X, 5)]
    for name, Xi, members in nets[:a.nets]:
        for lr in a.lr:
            for steps in a.steps:
                model = S.Surrogate(in_features=Xi.shape[1], seed=a.seed + 1, members=members,
                                    params=params)
                t0 = time.perf_counter()
                losses = model.fit(Xi[~te], y[~te], steps=steps, batch=a.batch, lr=lr,
                                   seed=a.seed + 1)
                sec = time.perf_counter() - t0
                mae = float(np.abs(model.predict(Xi[te]) - y[te]).mean())
                r = {"input": name, "in_features": int(Xi.shape[1]), "members": members,
                     "steps": steps, "lr": lr, "fit_seconds": sec,
                     "final_loss": float(losses[-20:].mean()),
                     "mae": mae, "mae_fraction_of_mean": mae / ym}
                out["runs"].append(r)
                log(f"[config4] {name}, {members} network(s), lr {lr:g}, {steps} steps: "
                    f"held-out MAE {mae:.2f} = {100 * mae / ym:.2f}% ({sec:.1f} s)")
                del model
    with open(a.out, "w") as f:
        json.dump(out, f, indent=1)
    log(f"[config4] -> {a.out}")


def cmd_compare(a):
    params = sim.load_parameters(a.params)
    X, y, pi, z = load_samples(a.data, params)
    model, sur = train_surrogate(X, y, pi, a, params)
    ts, desc = tumours(params, 10, a.seed)
    ga_t, ga_d = ts[:3], desc[:3]
    cfg = ops.default_config()
    result = {"surrogate": sur, "config": {
        "pop_size": a.pop, "generations": a.generations, "reps": a.reps, "pool": a.pool,
        "verify": a.verify, "tumours": 3, "budget_gy": cfg["max_value"], "hours": HOURS}}
    reeval = D.SimulatorFitness(params, ga_t, a.reeval_reps, hours=HOURS, seed=a.seed + 99)

    def reevaluate(hist):
        """The best genome found, re-simulated with fresh replicates."""
        best = max(hist, key=lambda h: h["best"])
        return float(reeval([np.asarray(best["best_genome"], np.float32)])[0])

    def save():
        with open(a.out, "w") as f:
            json.dump(result, f, indent=1)

    if not a.skip_plain:
        fit = D.SimulatorFitness(params, ga_t, a.reps, hours=HOURS, seed=a.seed + 11)
        t0 = time.perf_counter()
        hist = D.run_plain_ga(np.random.default_rng(a.seed + 1), fit, cfg, pop_size=a.pop,
                              generations=a.generations, n_slots=N_SLOTS,
                              time_limit=a.time_limit, log=log)
        result["plain"] = {"seconds": time.perf_counter() - t0, "history": hist,
                           "evaluations": fit.evaluations, "simulations": fit.simulations,
                           "sim_seconds": fit.seconds, "best_reevaluated": reevaluate(hist)}
        target = hist[-1]["best_so_far"]
        save()
        log(f"[plain] {result['plain']['seconds']:.1f} s, best {target:.2f}, re-evaluated "
            f"{result['plain']['best_reevaluated']:.2f}")
    else:
        target = a.target
    if not a.skip_surrogate:
        fit = D.SimulatorFitness(params, ga_t, a.reps, hours=HOURS, seed=a.seed + 12)
        t0 = time.perf_counter()
        hist = D.run_surrogate_ga(np.random.default_rng(a.seed + 2), fit, model, ga_d, cfg,
                                  pop_size=a.pop, generations=a.generations, n_slots=N_SLOTS,
                                  pool=a.pool, verify=a.verify, refit_steps=a.refit_steps,
                                  time_limit=a.time_limit,
                                  target=target if a.stop_at_target else None, log=log)
        result["surrogate_ga"] = {"seconds": time.perf_counter() - t0, "history": hist,
                                  "evaluations": fit.evaluations,
                                  "simulations": fit.simulations, "sim_seconds": fit.seconds}
        result["surrogate_ga"]["best_reevaluated"] = reevaluate(hist)
        if target is not None:
            hit = [h for h in hist if h["best_so_far"] >= target]
            result["surrogate_ga"]["seconds_to_target"] = hit[0]["seconds"] if hit else None
            result["target"] = target
        log(f"[surrogate GA] {result['surrogate_ga']['seconds']:.1f} s, best "
            f"{hist[-1]['best_so_far']:.2f}, re-evaluated "
            f"{result['surrogate_ga']['best_reevaluated']:.2f}, to target "
            f"{result['surrogate_ga'].get('seconds_to_target')}")
    save()
    log(f"[compare] -> {a.out}")


def main(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    sub = p.add_subparsers(dest="cmd", required=True)
    d = sub.add_parser("data")
    d.add_argument("--out", required=True)
    d.add_argument("--protocols", type=int, default=1200)
    d.add_argument("--tumours", type=int, default=10)
    d.add_argument("--reps", type=int, default=10)
    d.add_argument("--chunk", type=int, default=200)
    d.add_argument("--seed", type=int, default=0, help="tumours (and replicate streams)")
    d.add_argument("--protocol-seed", type=int, default=None,
                   help="protocols (default: seed + 100)")
    d.add_argument("--epoch-base", type=int, default=0,
                   help="replicate-stream epoch of the first chunk (use a different one per file)")
    d.add_argument("--params", default=DEFAULT_PARAMS)
    c = sub.add_parser("compare")
    c.add_argument("--data", required=True, nargs="+", help="one or more data files")
    c.add_argument("--out", required=True)
    c.add_argument("--pop", type=int, default=100)
    c.add_argument("--generations", type=int, default=50)
    c.add_argument("--reps", type=int, default=20)
    c.add_argument("--pool", type=int, default=10000)
    c.add_argument("--verify", type=int, default=200)
    c.add_argument("--train-steps", type=int, default=500)
    c.add_argument("--batch", type=int, default=2048)
    c.add_argument("--members", type=int, default=5, help="networks in the ensemble")
    c.add_argument("--lr", type=float, default=1e-3)
    c.add_argument("--refit-steps", type=int, default=0)
    c.add_argument("--skip-plain", action="store_true")
    c.add_argument("--skip-surrogate", action="store_true")
    c.add_argument("--target", type=float, default=None)
    c.add_argument("--stop-at-target", action="store_true")
    c.add_argument("--time-limit", type=float, default=None,
                   help="stop a GA after this many seconds (the generation running finishes)")
    c.add_argument("--reeval-reps", type=int, default=100,
                   help="replicates per tumour when re-simulating each run's best genome")
    c.add_argument("--seed", type=int, default=0)
    c.add_argument("--params", default=DEFAULT_PARAMS)
    f = sub.add_parser("config4")
    f.add_argument("--data", required=True, nargs="+", help="data files of the same tumours")
    f.add_argument("--out", required=True)
    f.add_argument("--steps", type=int, nargs="+", default=[1000, 4000])
    f.add_argument("--batch", type=int, default=4096)
    f.add_argument("--lr", type=float, nargs="+", default=[1e-3])
    f.add_argument("--nets", type=int, default=3,
                   help="networks to fit: the single 152-input one, then its ensemble of 5, "
                        "then the feature ensemble")
    f.add_argument("--seed", type=int, default=0)
    f.add_argument("--params", default=DEFAULT_PARAMS)
    a = p.parse_args(argv)
    {"data": cmd_data, "compare": cmd_compare, "config4": cmd_config4}[a.cmd](a)


if __name__ == "__main__":
    main()
