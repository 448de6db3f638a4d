"""Seeded synthetic workloads: tumour sets and random dose protocols.

Protocols follow the search space of the reference's GA (python_ga/utils.py:
doses in 0.25 Gy steps on a grid of 300 simulation steps = 0.5 h): up to
`max_doses` doses of 0.25-1.25 Gy, total at most `budget` Gy, on distinct
0.5 h slots of the first 72 h.
"""
import numpy as np

from .synthetic import disk_state

STEPS_PER_HOUR = 600          # 6 s steps (default-parameters.json: time_step)
SLOT_STEPS = 300              # protocol resolution: 0.5 h


def random_protocols(n, seed, max_doses=10, budget=10.0, n_grid=144, dose_min=0.25,
                     dose_max=1.25, dose_step=0.25):
    """n protocols as lists of (step, dose) pairs."""
    rng = np.random.Generator(np.random.PCG64(int(seed)))
    levels = np.arange(dose_min, dose_max + 1e-9, dose_step)
    out = []
    for _ in range(n):
        k = int(rng.integers(1, max_doses + 1))
        slots = np.sort(rng.choice(n_grid, size=k, replace=False))
        doses = rng.choice(levels, size=k)
        while doses.sum() > budget + 1e-9:
            doses = doses[:-1]
            slots = slots[:-1]
        out.append([(int(s) * SLOT_STEPS, float(d)) for s, d in zip(slots, doses)])
    return out


def dose_table(protocols, n_slots, res=SLOT_STEPS):
    """(P, n_slots) float32 doses; doses on one slot accumulate
    (Protocol::closestDose, protocol.h:16-18)."""
    t = np.zeros((len(protocols), n_slots), np.float32)
    for i, prot in enumerate(protocols):
        for step, dose in prot:
            t[i, step // res] += np.float32(dose)
    return t


def synthetic_tumors(params, n, seed, size=53, radius=(8, 14)):
    """n seeded disk tumours of the given bordered size, radii uniform in the
    inclusive range."""
    rng = np.random.Generator(np.random.PCG64(int(seed)))
    out = []
    for i in range(n):
        r = int(rng.integers(radius[0], radius[1] + 1))
        st, _ = disk_state(params, size, size, radius=r, seed=int(seed) * 1000 + i)
        out.append(st)
    return out


def batched_uniform(params, n_tumors=10, n_protocols=100, reps=20, hours=144, seed=0):
    """10 synthetic 53x53 tumours x 100 random protocols x 20 replicates,
    144 h: simulation b = p*T*R + t*R + r (the Experiment layout,
    backend.cc:59-73)."""
    tumors = synthetic_tumors(params, n_tumors, seed)
    prots = random_protocols(n_protocols, seed + 1)
    horizon = hours * STEPS_PER_HOUR
    # every protocol's slots (doses past the horizon never fire)
    last = max(s for p in prots for s, _ in p)
    n_slots = max((horizon + SLOT_STEPS - 1) // SLOT_STEPS, last // SLOT_STEPS + 1)
    P, T, R = n_protocols, n_tumors, reps
    b = np.arange(P * T * R)
    return {
        "tumors": tumors,
        "tumor_index": (b // R) % T,
        "prot_index": b // (T * R),
        "horizon": np.full(P * T * R, horizon, np.uint32),
        "doses": dose_table(prots, n_slots),
        "protocols": prots,
        "prot_length": horizon,
        "hours": np.full(P * T * R, hours, np.float64),
    }


def batched_uneven(params, n_sims=50000, seed=0, sizes=(33, 53, 81), hours=(48, 144, 240),
                   reps=(1, 20), n_tumors_per_size=4, n_protocols=200):
    """Mixed lattice sizes, horizons and replicate counts: groups of
    (tumour, protocol, horizon) with 1-20 replicates each until n_sims."""
    rng = np.random.Generator(np.random.PCG64(int(seed)))
    tumors, tumor_size = [], []
    for si, size in enumerate(sizes):
        rmax = max(4, min(14, size // 2 - 6))
        ts = synthetic_tumors(params, n_tumors_per_size, seed * 10 + si, size=size,
                              radius=(max(3, rmax - 6), rmax))
        tumors += ts
        tumor_size += [size] * len(ts)
    prots = random_protocols(n_protocols, seed + 7)
    hmax = max(hours) * STEPS_PER_HOUR
    n_slots = (hmax + SLOT_STEPS - 1) // SLOT_STEPS
    ti, pi, hz, hh = [], [], [], []
    while len(ti) < n_sims:
        t = int(rng.integers(len(tumors)))
        p = int(rng.integers(n_protocols))
        h = int(rng.choice(hours))
        k = int(rng.integers(reps[0], reps[1] + 1))
        k = min(k, n_sims - len(ti))
        ti += [t] * k
        pi += [p] * k
        hz += [h * STEPS_PER_HOUR] * k
        hh += [h] * k
    return {
        "tumors": tumors,
        "tumor_index": np.array(ti, np.int32),
        "prot_index": np.array(pi, np.int32),
        "horizon": np.array(hz, np.uint32),
        "doses": dose_table(prots, n_slots),
        "protocols": prots,
        "prot_length": hmax,
        "hours": np.array(hh, np.float64),
        "sizes": np.array([tumor_size[t] for t in ti], np.int32),
    }
