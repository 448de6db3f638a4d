"""Replay a protocol file on every tumour of a library: the batch replay
harness of the reference (tools/run_protocols_gpu.py:40-139), on one
batched GPU run per chunk of protocols.

Output: one row per protocol of n_tumours * reps final living-cell counts,
tumour-major then replicate-major, space-separated, as the reference's
results files (tools/run_protocols_gpu.py:119-137), plus a JSON file of
metadata.  Simulation b = p*T*R + t*R + r of the whole file keeps its stream
id whatever the chunking, so a replay reproduces an Experiment over all the
protocols with the same seed.

  python -m emt6ro.replay --protocols P.txt --tumors data/tumor-lib/tumor-*.txt \\
      --params data/default-parameters.json --reps 100 --steps 144000 --out results.txt
"""
import argparse
import json
import os
import sys
import time

import numpy as np

from . import protocols as prot_io
from . import simulation as sim
from .batch import Batch


def dose_rows(protocols, steps, resolution):
    """(P, ceil(steps / resolution)) doses; doses on one slot add up
    (Protocol::closestDose, common/protocol.h:16-18)."""
    n_slots = (steps + resolution - 1) // resolution
    out = np.zeros((len(protocols), n_slots), np.float32)
    for i, prot in enumerate(protocols):
        for step, dose in prot:
            if not 0 <= step < n_slots * resolution:
                raise ValueError(f"protocol {i}: step {step} is outside the {steps}-step protocol")
            out[i, step // resolution] += np.float32(dose)
    return out


def replay(params, tumors, protocols, reps, steps=144000, resolution=300, seed=None,
           max_sims=50000, gpu_id=0, log=None):
    """(P, T, reps) final living counts of every protocol on every tumour."""
    if seed is None:
        seed = int.from_bytes(os.urandom(8), "little")
    T, P = len(tumors), len(protocols)
    doses = dose_rows(protocols, steps, resolution)
    chunk = max(1, max_sims // max(T * reps, 1))
    out = np.zeros((P, T, reps), np.int64)
    for p0 in range(0, P, chunk):
        p1 = min(P, p0 + chunk)
        n = (p1 - p0) * T * reps
        i = np.arange(n)
        t0 = time.perf_counter()
        b = Batch(params, tumors, (i // reps) % T, i // (T * reps),
                  np.full(n, steps, np.uint32), doses[p0:p1], resolution, steps,
                  sim_id=p0 * T * reps + i, gpu_id=gpu_id, seed=seed)
        b.run(steps)
        out[p0:p1] = b.results().reshape(p1 - p0, T, reps)
        if log:
            dt = time.perf_counter() - t0
            log(f"  protocols {p0}..{p1 - 1} of {P}: {p1 - p0} x {T} tumours x {reps} reps "
                f"in {dt:.1f} s ({n / dt:.0f} sims/s)")
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description="Replay a protocol file on a tumour library.")
    ap.add_argument("--protocols", required=True)
    ap.add_argument("--tumors", nargs="+", required=True)
    ap.add_argument("--params", required=True)
    ap.add_argument("--reps", type=int, default=100)
    ap.add_argument("--steps", type=int, default=144000)
    ap.add_argument("--resolution", type=int, default=300)
    ap.add_argument("--out", required=True)
    ap.add_argument("--meta-out", default=None)
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--max-sims", type=int, default=50000,
                    help="simulations per batched run")
    ap.add_argument("--limit-protocols", type=int, default=None)
    a = ap.parse_args(argv)
    if a.seed is None:
        # drawn here so that the metadata records the seed actually used
        a.seed = int.from_bytes(os.urandom(8), "little")
    meta_out = a.meta_out or a.out + ".meta.json"
    params = sim.load_parameters(a.params)
    tumors = [sim.load_state(p, params) for p in a.tumors]
    protocols = prot_io.canonical(prot_io.parse_two_line(a.protocols), a.resolution)
    if a.limit_protocols:
        protocols = protocols[:a.limit_protocols]
    print(f"{len(protocols)} protocols x {len(tumors)} tumours x {a.reps} reps", flush=True)
    t0 = time.perf_counter()
    res = replay(params, tumors, protocols, a.reps, a.steps, a.resolution, a.seed,
                 a.max_sims, log=lambda m: print(m, flush=True))
    wall = time.perf_counter() - t0
    os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
    with open(a.out, "w") as f:
        for row in res.reshape(len(protocols), -1):
            f.write(" ".join(str(int(x)) for x in row) + "\n")
    with open(meta_out, "w") as f:
        json.dump({"n_protocols": len(protocols), "n_tumors": len(tumors),
                   "reps_per_tumor": a.reps, "steps": a.steps, "resolution": a.resolution,
                   "tumors": a.tumors, "protocols_file": a.protocols, "params_file": a.params,
                   "seed": a.seed, "wall_time_sec": wall,
                   "row_layout": "values per row are tumor-major then rep-major: "
                                 "[t0r0, t0r1, ..., t0r{M-1}, t1r0, ...]"}, f, indent=2)
    print(f"wrote {a.out} ({len(protocols)} rows) and {meta_out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
