#!/usr/bin/env python3
"""Benchmark: simulated tumour-hours per second on one MI355X.

Workload (batched uniform): 10 seeded synthetic 53x53 tumours (occupied disks
of radius 8-14) x 100 random dose protocols x 20 replicates = 20,000
simulations, 144 simulated hours (86,400 steps of 6 s), FP32 substrates.

One bench step advances every simulation by one simulated hour (600 model
steps): one launch of the window kernel (the simulations whose ROI fits a
window of the lattice in LDS) and one of the whole-lattice persistent kernel
(the rest); --steps 144 (the default) is the whole 144 h workload.  Warm-up steps run on the same batch, which is then
reset to its initial states (untimed) before the timed steps.

Also reported, on bounded samples:
  * the per-step-launch path (the reference's structure: one launch per
    phase per model step) on the same 20,000 simulations for 1 h;
  * the CPU oracle on this host's cores for 16 of the simulations over the
    full 144 h (cpu_baseline).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "emt6-ro_amd"))

METRIC = "simulated tumour-hours/sec on 1 MI355X; GA evals/sec surrogate vs plain"
FP32_PEAK_TFLOPS = 157.3      # MI355X published FP32 peak
FLOP_PER_SITE_SUBSTEP = 36    # 3 substrates x (6 adds + 3 fma x 2)
FP32_NONPACKED_TFLOPS = 78.6  # the same without packed FP32 (the kernels are compiled without it)
# HBM bytes per simulated hour of this workload (20,000 simulations): the
# window kernel's and the persistent kernel's (variant 32) rocprofv3
# FETCH_SIZE + WRITE_SIZE over a 2 + 144 h run, divided by its 146 hours, each
# counter scaled by what it reads for a 1 GiB device copy (x2.0, x1.0).
TRAFFIC_BYTES_PER_LAUNCH = 14.22e9 + 20.82e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=144, help="timed simulated hours")
    ap.add_argument("--warmup", type=int, default=2, help="untimed simulated hours")
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-step path sample")
    ap.add_argument("--cpu-sims", type=int, default=16)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner")
    ap.add_argument("--prefer", type=int, default=-1,
                    help="whole-lattice variant to use without tuning (the window pass stays on)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="only run a 1 GiB device copy (profiler counter calibration)")
    ap.add_argument("--sims", type=int, default=0,
                    help="use only the first N simulations (profiling runs)")
    ap.add_argument("--tune-steps", type=int, default=600)
    ap.add_argument("--tune-refine-warm", type=int, default=1800,
                    help="re-time the closest variants after this many steps (0: off)")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate training / screening sample")
    ap.add_argument("--surrogate-only", action="store_true",
                    help="only run the surrogate sample (profiling runs)")
    ap.add_argument("--json-out", default=None)
    return ap.parse_args()


def cpu_baseline(params, wl, n_sims, hours):
    """The CPU oracle (test infrastructure) on this host's cores, on the first
    n_sims simulations of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    v = params.values()
    states = []
    cache = {}
    for i in range(n_sims):
        t = int(wl["tumor_index"][i])
        if t not in cache:
            st = wl["tumors"][t]
            H, W = st.shape
            s = O.empty_state(v, H, W)
            s["sub"][0], s["sub"][1], s["sub"][2] = st.cho(), st.ox(), st.gi()
            s["st"][:] = st.occupancy()
            cells = s["cells"].reshape(H, W)
            occ = st.occupancy() == 1
            cells["t_rep"][occ] = st.field("time_in_repair")[occ]
            cells["R"][occ] = st.field("irradiation")[occ]
            cells["delay"][occ] = st.field("repair_delay")[occ]
            cells["prolif"][occ] = st.field("proliferation_time")[occ]
            for k, n in enumerate(("cycle_g1", "cycle_s", "cycle_g2", "cycle_m", "cycle_d")):
                cells["times"][occ, k] = st.field(n)[occ]
            cells["mode"][occ] = st.mode()[occ]
            cells["phase"][occ] = st.phase()[occ]
            cache[t] = s
        states.append(cache[t])
    cores = max(1, min(16, os.cpu_count() or 1, n_sims))
    orc = O.OracleBatch(v, states, wl["doses"], wl["prot_index"][:n_sims],
                        np.arange(n_sims), 300, wl["prot_length"])
    t0 = time.perf_counter()
    orc.run(hours * 600, seed=0, nthreads=cores)
    dt = time.perf_counter() - t0
    return {"value": n_sims * hours / dt, "unit": "tumour-hours/s", "cores": cores,
            "kind": "port", "seconds": dt,
            "sample": f"{n_sims} of the 20000 simulations (first in batch order) over the "
                      f"full {hours} h, CPU oracle (C, FP32, OpenMP over simulations)"}


def surrogate_sample(train_steps=200, batch=4096, n_data=200000, n_screen=100000):
    """The fitness surrogate's network (N_FEATURES-512-512-512-1, BF16 on the
    matrix cores): Adam steps on random data and one screening pass, timed
    on the host."""
    import numpy as np
    from emt6ro.surrogate import MLP, N_FEATURES
    hidden = 512
    rng = np.random.default_rng(0)
    X = rng.standard_normal((n_data, N_FEATURES)).astype(np.float32)
    y = rng.standard_normal(n_data).astype(np.float32)
    net = MLP(N_FEATURES, hidden, seed=0)
    net.set_data(X, y)
    net.train(10, batch=batch, lr=1e-4, seed=1)            # warm-up
    t0 = time.perf_counter()
    net.train(train_steps, batch=batch, lr=1e-4, seed=2)
    train_s = time.perf_counter() - t0
    Xs = rng.standard_normal((n_screen, N_FEATURES)).astype(np.float32)
    net.predict(Xs[:4096])
    t0 = time.perf_counter()
    net.predict(Xs)
    screen_s = time.perf_counter() - t0
    # GEMM FLOP per sample, with the input padded to a multiple of 32 columns.
    kin = (N_FEATURES + 31) // 32 * 32
    fwd = 2 * (kin * hidden + 2 * hidden * hidden) * batch
    # forward, weight gradients (same shapes), input gradients of layers 2-3
    step_flop = 2 * fwd + 2 * (2 * hidden * hidden) * batch
    return {"train_steps_per_s": train_steps / train_s,
            "train_tflops_mfma_gemm": step_flop * train_steps / train_s / 1e12,
            "batch": batch, "screen_rows_per_s": n_screen / screen_s,
            "note": "host-timed; includes the host<->device copies of predict"}


def copy_calibration(gib=1):
    """A plain device copy of known size, to scale rocprofv3's FETCH_SIZE and
    WRITE_SIZE counters (run under the profiler with --copy-calibrate)."""
    import torch
    n = gib * (1 << 30) // 4
    src = torch.ones(n, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4, "copies": 3}), flush=True)


def main():
    args = parse()
    if args.copy_calibrate:
        copy_calibration()
        return
    if args.surrogate_only:
        print(json.dumps({"surrogate": surrogate_sample()}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the simulations are "
              "independent, so N GPUs would each run 1/N of the batch.", file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch, variants
    from emt6ro.workloads import batched_uniform

    params = sim.load_parameters(os.path.join(ROOT, "tests", "golden", "default-parameters.json"))
    wl = batched_uniform(params, seed=0)
    if args.sims:
        for k in ("tumor_index", "prot_index", "horizon"):
            wl[k] = wl[k][:args.sims]
    n = len(wl["tumor_index"])
    b = Batch(params, wl["tumors"], wl["tumor_index"], wl["prot_index"], wl["horizon"],
              wl["doses"], 300, wl["prot_length"], sim_id=np.arange(n), seed=0)
    hour = 600
    tuned = None
    if args.prefer >= 0:
        b.set_variant(56, args.prefer)
    if args.variant < 0 and not args.no_tune:
        from emt6ro.tune import autotune
        tuned = autotune(b, {56: (53, 53)}, steps=args.tune_steps, use_cache=False,
                         log=lambda m: print(m, file=sys.stderr, flush=True),
                         refine_warm=args.tune_refine_warm)
    for _ in range(args.warmup):
        b.run(hour, path="persistent", variant=args.variant)
    b.reset(0)
    # Timed region: every simulation advances args.steps hours.
    event_ms = 0.0
    before = b.counters()["site_substeps"].astype(np.float64).sum()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        event_ms += b.run(hour, path="persistent", variant=args.variant)
    wall = time.perf_counter() - t0
    after = b.counters()["site_substeps"].astype(np.float64).sum()
    living = b.results()
    value = n * args.steps / wall
    flops = (after - before) * FLOP_PER_SITE_SUBSTEP
    launch_s = event_ms / 1e3 / max(args.steps, 1)
    achieved = flops / max(args.steps, 1) / launch_s / 1e12
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "tumour-hours/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched uniform: 10 synthetic 53x53 tumours (disk radius 8-14) x 100 "
                        "random protocols (<=10 doses of 0.25-1.25 Gy, <=10 Gy, 0.5 h grid, "
                        "first 72 h) x 20 replicates = 20000 simulations, 144 h horizon",
            "sims": n, "lattice": "53x53", "hours_timed": args.steps, "path": "persistent",
            "variant": args.variant, "mean_living": float(living.mean()),
            "tuned": tuned,
        },
        "roofline": {
            "bound": "mfma", "pipe": "valu_fp32",
            "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS, "traffic": TRAFFIC_BYTES_PER_LAUNCH,
            "frac_of_nonpacked_fp32": achieved / FP32_NONPACKED_TFLOPS,
            "hbm_gbps": TRAFFIC_BYTES_PER_LAUNCH / launch_s / 1e9,
            "kernel": "window_kernel + persistent_kernel (one hour of every simulation)",
            "launch_ms": launch_s * 1e3,
            "site_substeps_per_launch": (after - before) / max(args.steps, 1),
            "flop_per_site_substep": FLOP_PER_SITE_SUBSTEP,
        },
    }
    if not args.no_plain:
        # Capture and instantiate the per-step graph outside the timed region.
        b.reset(0)
        b.run(128, path="graph", variant=args.variant)
        b.reset(0)
        t0 = time.perf_counter()
        b.run(hour, path="graph", variant=args.variant)
        plain = n / (time.perf_counter() - t0)
        out["plain_per_step_launch"] = {
            "value": plain, "unit": "tumour-hours/s",
            "sample": "the same 20000 simulations, first hour (600 steps x 3 launches, "
                      "replayed from a captured HIP graph)"}
        b.reset(0)
        t0 = time.perf_counter()
        b.run(hour, path="persistent", variant=args.variant)
        pers1 = n / (time.perf_counter() - t0)
        out["persistent_first_hour"] = {"value": pers1, "unit": "tumour-hours/s"}
        out["speedup_vs_plain_first_hour"] = pers1 / plain
    if not args.no_surrogate:
        sur = surrogate_sample()
        out["surrogate"] = sur
        # One GA fitness evaluation = 3 tumours x 20 replicates x 144 h.
        plain_evals = value / (3 * 20 * 144)
        out["ga_evals_per_s"] = {"plain_simulator": plain_evals,
                                 "surrogate_screen": sur["screen_rows_per_s"] / 3,
                                 "ratio": sur["screen_rows_per_s"] / 3 / plain_evals}
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(params, wl, args.cpu_sims, 144)
    line = json.dumps(out)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
