"""Autotuner for the persistent kernel.

For each lattice size class of a batch (LDS row stride 36 / 56 / 88), times
every compiled kernel variant that fits (workgroup size, diffusion band rows
= substep unroll depth, items per thread) on a short calibration run of the
batch itself, keeps the fastest, and restores the batch to its initial
state.  Choices are cached per (class, lattice size, batch size) in a JSON
file so later batches of the same shape skip the calibration.
"""
import json
import os

from .batch import fits, variants

CACHE = os.environ.get("EMT6RO_TUNE_CACHE",
                       os.path.join(os.path.expanduser("~"), ".cache", "emt6ro_amd", "tune.json"))


def _load():
    try:
        with open(CACHE) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(d):
    try:
        os.makedirs(os.path.dirname(CACHE), exist_ok=True)
        with open(CACHE, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
    except OSError:
        pass


def autotune(batch, classes, steps=150, use_cache=True, log=None, refine_warm=0,
             refine_top=3, refine_within=0.05):
    """classes: {rs: (hmax, wmax)} of the batch's size classes.
    Returns {rs: {"variant": id, "ms": {id: ms}}}; the batch is left reset
    (epoch 0) with the chosen variants set.

    With refine_warm > 0, the (at most refine_top) variants within
    refine_within of the fastest are timed again on `steps` steps that start
    after refine_warm steps run with the fastest one, since tumours grow and
    the ranking of close variants can change with the ROI size; results do
    not depend on the variant, so every candidate starts from the same
    state.  The refinement runs with the caller's window setting, so that
    the whole-lattice kernel sees the simulations it takes in production
    (those whose ROI no longer fits a window)."""
    cache = _load() if use_cache else {}
    # The first timing is of the whole-lattice kernel alone: the window pass
    # (which takes the simulations with small ROIs first, so that the
    # whole-lattice kernel would see none of them early in a run) is off
    # except during a refinement, and the caller's setting is restored at
    # the end.
    window = getattr(batch, "window", -1)
    if hasattr(batch, "set_window"):
        batch.set_window(0)
    out = {}
    vs = variants()
    for rs, (hmax, wmax) in sorted(classes.items()):
        key = f"rs{rs}_h{hmax}_w{wmax}_n{batch.n}"
        if use_cache and key in cache:
            batch.set_variant(rs, cache[key]["variant"])
            out[rs] = cache[key]
            continue
        timings = {}
        for v in vs:
            if not fits(v, hmax, wmax) or v["rs"] < hmax + 2:
                continue
            batch.reset(0)
            batch.set_variant(rs, v["id"])
            timings[v["id"]] = batch.run(steps, path="persistent")
            if log:
                log(f"[tune] rs={rs} variant={v} {timings[v['id']]:.2f} ms")
        best = min(timings, key=timings.get)
        out[rs] = {"variant": best, "ms": timings}
        if refine_warm > 0:
            close = sorted((t, v) for v, t in timings.items()
                           if t <= timings[best] * (1.0 + refine_within))[:refine_top]
            later = {}
            for _, v in close:
                batch.reset(0)
                if hasattr(batch, "set_window"):
                    batch.set_window(window)
                batch.set_variant(rs, best)
                batch.run(refine_warm, path="persistent")
                batch.set_variant(rs, v)
                later[v] = batch.run(steps, path="persistent")
                if hasattr(batch, "set_window"):
                    batch.set_window(0)
                if log:
                    log(f"[tune] rs={rs} variant {v} after {refine_warm} steps: {later[v]:.2f} ms")
            best = min(later, key=later.get)
            out[rs] = {"variant": best, "ms": timings, "ms_later": later,
                       "refine_warm": refine_warm}
        batch.set_variant(rs, best)
        cache[key] = out[rs]
    batch.reset(0)
    if hasattr(batch, "set_window"):
        batch.set_window(window)
    if use_cache:
        _save(cache)
    return out
