#!/usr/bin/env python3
"""Benchmark: simulated tumour-hours per second on one MI355X.

Workload (batched uniform): 10 seeded synthetic 53x53 tumours (occupied disks
of radius 8-14) x 100 random dose protocols x 20 replicates = 20,000
simulations, 144 simulated hours (86,400 steps of 6 s), FP32 substrates.

The timed region always covers the whole 144 h horizon: one bench step is
144 / --steps simulated hours of every simulation, run as one-hour launches
(the window kernel for the simulations whose ROI fits a window of the
lattice in LDS, then the whole-lattice persistent kernel for the rest).
Warm-up steps (the same size) run on the same batch, which is then reset to
its initial states (untimed) before the timed steps.

Also reported:
  * the LDS read bandwidth of the GPU (a probe kernel) and the simulator's
    LDS traffic against it;
  * the per-step-launch path (the reference's structure: per model step one
    ROI, one diffusion and one cell launch, replayed from a HIP graph) on
    every 12th simulated hour of the same run, against the persistent
    path's time for the same hours;
  * the CPU oracle on this host's cores and on one core (cpu_baseline);
  * the fitness surrogate's training and screening rates.
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
# 12 B read of the three centre values + 12 B write of the new ones per
# site-substep, neighbour reuse assumed perfect.
LDS_BYTES_PER_SITE_SUBSTEP = 24
FP32_NONPACKED_TFLOPS = 78.6  # the same without packed FP32 (the kernels are compiled without it)
HOURS = 144
HOUR = 600                    # model steps per simulated hour
# HBM bytes per simulated hour of this workload (both kernels of an hour),
# from rocprofv3 FETCH_SIZE and WRITE_SIZE over a whole run, each scaled by
# what the counter reads for a 1 GiB device copy; the profile below gives the
# commands and figures.
TRAFFIC_PROFILE = "profiles/persistent_kernel_batched_uniform_53x53.md"
TRAFFIC_BYTES_PER_HOUR = 16.4e9 + 26.4e9


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed steps; each is 144/steps simulated hours, so the timed "
                         "region is always the whole 144 h")
    ap.add_argument("--warmup", type=int, default=2, help="untimed steps (same size)")
    ap.add_argument("--variant", type=int, default=-1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-plain", action="store_true", help="skip the per-step path sample")
    ap.add_argument("--plain-every", type=int, default=12,
                    help="time the per-step path on every N-th simulated hour")
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner")
    ap.add_argument("--prefer", type=int, default=-1,
                    help="whole-lattice variant to use without tuning (the window pass stays on)")
    ap.add_argument("--copy-calibrate", action="store_true",
                    help="only run a 1 GiB device copy (profiler counter calibration)")
    ap.add_argument("--sims", type=int, default=0,
                    help="use only the first N simulations (profiling runs)")
    ap.add_argument("--tune-steps", type=int, default=600)
    ap.add_argument("--tune-refine-warm", type=int, default=HOURS * HOUR // 2,
                    help="re-time the closest variants after this many steps (0: off); "
                         "default mid-run (72 h), where the cost per hour is near the "
                         "run's mean")
    ap.add_argument("--no-surrogate", action="store_true",
                    help="skip the surrogate training / screening sample")
    ap.add_argument("--surrogate-only", action="store_true",
                    help="only run the surrogate sample (profiling runs)")
    ap.add_argument("--surrogate-graph", type=int, default=None,
                    help="training steps per replayed HIP graph (0: direct launches; "
                         "default: the library's)")
    ap.add_argument("--surrogate-bounds", action="store_true",
                    help="only measure what bounds a training step at batch 4096: a grid-wide "
                         "barrier inside a kernel against a kernel launch, and the library "
                         "GEMM (torch.matmul, hipBLASLt) on the step's product shapes")
    ap.add_argument("--json-out", default=None)
    return ap.parse_args()


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(params, wl, hours):
    """The CPU oracle (test infrastructure) on this host: the first `cores`
    simulations of the workload on `cores` OpenMP threads, one simulation per
    thread, and simulation `cores` alone on one thread, each over the full
    horizon.  cores = the CPUs this process may use, at most 16 (the GPU
    box's share of its host)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O
    v = params.values()
    cache = {}

    def state(i):
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
        return cache[t]

    def timed(idx, threads):
        orc = O.OracleBatch(v, [state(i) for i in idx], wl["doses"], wl["prot_index"][idx],
                            np.asarray(idx), 300, wl["prot_length"])
        t0 = time.perf_counter()
        orc.run(hours * HOUR, seed=0, nthreads=threads)
        return time.perf_counter() - t0

    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    cores = max(1, min(16, avail))
    dt = timed(list(range(cores)), cores)
    dt1 = timed([cores], 1)
    return {"value": cores * hours / dt, "unit": "tumour-hours/s", "cores": cores,
            "kind": "port", "seconds": dt,
            "sample": f"simulations 0-{cores - 1} of the 20000 (batch order), one per OpenMP "
                      f"thread, over the full {hours} h: CPU oracle (C, FP32)",
            "single_core": {"value": hours / dt1, "unit": "tumour-hours/s", "cores": 1,
                            "seconds": dt1,
                            "sample": f"simulation {cores} over the full {hours} h, 1 thread"},
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "cpus_available": avail}


def surrogate_sample(train_steps=200, batch=4096, n_data=200000, n_screen=100000,
                     graph_steps=None):
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
    if graph_steps is not None:
        net.set_train_graph(graph_steps)
    net.set_data(X, y)
    net.train(10, batch=batch, lr=1e-4, seed=1)            # warm-up
    t0 = time.perf_counter()
    net.train(train_steps, batch=batch, lr=1e-4, seed=2)
    train_s = time.perf_counter() - t0
    # The GA's surrogate is an ensemble of 5 such networks, trained together
    # (train_many: their steps interleaved on their own streams).
    from emt6ro.surrogate import train_many
    ens = [net] + [MLP(N_FEATURES, hidden, seed=1 + m) for m in range(4)]
    for m in ens[1:]:
        m.set_data(X, y)
    train_many(ens, 10, batch=batch, lr=1e-4, seeds=range(5))      # warm-up
    t0 = time.perf_counter()
    train_many(ens, train_steps, batch=batch, lr=1e-4, seeds=range(10, 15))
    ens_s = time.perf_counter() - t0
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
            "ensemble5_member_steps_per_s": 5 * train_steps / ens_s,
            "ensemble5_tflops_mfma_gemm": step_flop * 5 * train_steps / ens_s / 1e12,
            "batch": batch, "screen_rows_per_s": n_screen / screen_s,
            "note": "host-timed; includes the host<->device copies of predict"}


def surrogate_bounds(batch=4096, hidden=512, reps=500):
    """What bounds a surrogate training step at batch 4096.  (1) A barrier
    across the whole GPU inside one kernel (one workgroup per CU) against a
    kernel boundary, both with the same writes between them: whether the
    step's 12 launches could become one kernel.  (2) The vendor library's
    GEMM (torch.matmul on bf16, which calls hipBLASLt) on the step's product
    shapes, timed per product with events; rocprofv3 --kernel-trace gives the
    kernels' own durations."""
    import torch
    from emt6ro._native import probe_grid_barrier
    from emt6ro.surrogate import N_FEATURES
    out = {"grid_barrier": []}
    for mode, name in ((0, "one counter"), (1, "8 group counters")):
        for words in (0, 1024, 16384):
            r = probe_grid_barrier(mode=mode, words=words)
            out["grid_barrier"].append({"mode": name, "bytes_per_workgroup": 4 * words,
                                        "barrier_us": r["barrier_us"],
                                        "launch_us": r["launch_us"]})
    kin = (N_FEATURES + 31) // 32 * 32
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev).manual_seed(0)
    shapes = [("forward layer 1", batch, kin, hidden), ("forward layers 2-3, input gradients",
                                                         batch, hidden, hidden),
              ("weight gradients of layers 2-3", hidden, batch, hidden),
              ("weight gradient of layer 1", hidden, batch, kin)]
    out["library_gemm"] = []
    for name, M, K, N in shapes:
        a = torch.randn(M, K, device=dev, dtype=torch.bfloat16, generator=g)
        bt = torch.randn(N, K, device=dev, dtype=torch.bfloat16, generator=g)
        c = torch.empty(M, N, device=dev, dtype=torch.bfloat16)
        for _ in range(20):
            torch.matmul(a, bt.t(), out=c)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            torch.matmul(a, bt.t(), out=c)
        e1.record()
        torch.cuda.synchronize()
        us = e0.elapsed_time(e1) * 1e3 / reps
        out["library_gemm"].append({"product": name, "M": M, "K": K, "N": N,
                                    "us_per_call": us, "tflops": 2 * M * N * K / us / 1e6})
    return out


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


def step_bounds(steps):
    """Model-step boundaries of the timed steps: step i covers
    [b[i], b[i+1]) of the 86,400-step horizon."""
    total = HOURS * HOUR
    return [total * i // steps for i in range(steps + 1)]


def run_span(b, lo, hi, variant, hour_ms=None, path="persistent"):
    """Advance every simulation from model step lo to hi in launches that end
    on hour boundaries; per-hour device time (HIP events) goes to hour_ms."""
    t = lo
    while t < hi:
        e = min(hi, (t // HOUR + 1) * HOUR)
        ms = b.run(e - t, path=path, variant=variant)
        if hour_ms is not None:
            hour_ms[t // HOUR] = hour_ms.get(t // HOUR, 0.0) + ms
        t = e


def main():
    args = parse()
    if args.copy_calibrate:
        copy_calibration()
        return
    if args.surrogate_bounds:
        print(json.dumps({"surrogate_bounds": surrogate_bounds()}), flush=True)
        return
    if args.surrogate_only:
        print(json.dumps({"surrogate": surrogate_sample(graph_steps=args.surrogate_graph)}),
              flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the simulations are "
              "independent, so N GPUs would each run 1/N of the batch.", file=sys.stderr)
        sys.exit(2)
    if args.steps < 1 or args.warmup < 0:
        sys.exit("--steps must be >= 1 and --warmup >= 0")
    import numpy as np
    import emt6ro.simulation as sim
    from emt6ro._native import probe_lds
    from emt6ro.batch import Batch
    from emt6ro.workloads import batched_uniform

    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    params = sim.load_parameters(os.path.join(ROOT, "tests", "golden", "default-parameters.json"))
    wl = batched_uniform(params, seed=0)
    if args.sims:
        for k in ("tumor_index", "prot_index", "horizon"):
            wl[k] = wl[k][:args.sims]
    n = len(wl["tumor_index"])
    b = Batch(params, wl["tumors"], wl["tumor_index"], wl["prot_index"], wl["horizon"],
              wl["doses"], 300, wl["prot_length"], sim_id=np.arange(n), seed=0)
    tuned = None
    if args.prefer >= 0:
        b.set_variant(56, args.prefer)
    if args.variant < 0 and not args.no_tune:
        from emt6ro.tune import autotune
        tuned = autotune(b, {56: (53, 53)}, steps=args.tune_steps, use_cache=False, log=log,
                         refine_warm=args.tune_refine_warm)
    bounds = step_bounds(args.steps)
    for i in range(args.warmup):
        k = i % args.steps
        run_span(b, bounds[k], bounds[k + 1], args.variant)
        if k == args.steps - 1:
            b.reset(0)
    b.reset(0)
    log(f"[bench] warm-up done; timing {args.steps} steps of "
        f"{HOURS / args.steps:g} h ({bounds[1] - bounds[0]} model steps)")
    # Timed region: the whole 144 h of every simulation.
    hour_ms = {}
    before = b.counters()["site_substeps"].astype(np.float64).sum()
    t0 = time.perf_counter()
    for i in range(args.steps):
        run_span(b, bounds[i], bounds[i + 1], args.variant, hour_ms)
    wall = time.perf_counter() - t0
    after = b.counters()["site_substeps"].astype(np.float64).sum()
    living = b.results()
    log(f"[bench] timed {wall:.2f} s")
    event_s = sum(hour_ms.values()) / 1e3
    value = n * HOURS / wall
    sub = after - before
    achieved = sub * FLOP_PER_SITE_SUBSTEP / event_s / 1e12
    probe = {w: probe_lds(w) for w in (16, 4)}
    lds_gbps = sub * LDS_BYTES_PER_SITE_SUBSTEP / event_s / 1e9
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "tumour-hours/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "batched uniform: 10 synthetic 53x53 tumours (disk radius 8-14) x 100 "
                        "random protocols (<=10 doses of 0.25-1.25 Gy, <=10 Gy, 0.5 h grid, "
                        "first 72 h) x 20 replicates = 20000 simulations, 144 h horizon",
            "sims": n, "lattice": "53x53", "hours_timed": HOURS,
            "hours_per_step": HOURS / args.steps, "path": "persistent",
            "variant": args.variant, "mean_living": float(living.mean()),
            "tuned": tuned,
        },
        "roofline": {
            "bound": "valu", "pipe": "FP32 VALU (the diffusion stencil on LDS-resident "
                                     "substrates)",
            "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": achieved / FP32_PEAK_TFLOPS,
            "traffic": TRAFFIC_BYTES_PER_HOUR * HOURS / args.steps,
            "traffic_source": f"{TRAFFIC_PROFILE}: rocprofv3 FETCH_SIZE + WRITE_SIZE "
                              "(copy-calibrated) per simulated hour x hours per step",
            "frac_of_nonpacked_fp32": achieved / FP32_NONPACKED_TFLOPS,
            "kernel": "window_kernel + persistent_kernel, one-hour launches",
            "launch_ms_per_hour": sum(hour_ms.values()) / len(hour_ms),
            "site_substeps": sub, "flop_per_site_substep": FLOP_PER_SITE_SUBSTEP,
            "lds": {"achieved_gbps": lds_gbps,
                    "bytes_per_site_substep": LDS_BYTES_PER_SITE_SUBSTEP,
                    "peak_gbps_b128": probe[16]["bytes_per_s"] / 1e9,
                    "peak_gbps_b32": probe[4]["bytes_per_s"] / 1e9,
                    "peak_bytes_per_clk_cu_b128": probe[16]["bytes_per_clk_cu"],
                    "peak_bytes_per_clk_cu_b32": probe[4]["bytes_per_clk_cu"],
                    "probe_clock_mhz": probe[16]["clock_hz"] / 1e6,
                    "frac": lds_gbps * 1e9 / probe[16]["bytes_per_s"]},
        },
    }
    hours_sorted = sorted(hour_ms)
    out["persistent_hour_ms"] = {"first": hour_ms[hours_sorted[0]],
                                 "last": hour_ms[hours_sorted[-1]],
                                 "per_12h": [round(sum(hour_ms.get(h, 0.0)
                                                       for h in range(h0, h0 + 12)), 1)
                                             for h0 in range(0, HOURS, 12)]}
    if not args.no_plain:
        # The per-step path on every N-th hour of the same run (the other hours
        # advance on the persistent path, untimed); both paths give identical
        # states, so the persistent path's times for the same hours compare
        # like with like.  The 128-step graph is captured beforehand.
        b.reset(0)
        b.run(128, path="graph", variant=args.variant)
        b.reset(0)
        plain_ms, sampled = {}, []
        for h in range(HOURS):
            if h % args.plain_every == 0:
                run_span(b, h * HOUR, (h + 1) * HOUR, args.variant, plain_ms, path="graph")
                sampled.append(h)
            else:
                run_span(b, h * HOUR, (h + 1) * HOUR, args.variant)
        same = bool(np.array_equal(b.results(), living))
        p_ms = sum(plain_ms[h] for h in sampled)
        q_ms = sum(hour_ms[h] for h in sampled)
        out["plain_per_step_launch"] = {
            "value": n * len(sampled) / (p_ms / 1e3), "unit": "tumour-hours/s",
            "sample": f"hours {sampled[0]}, {sampled[1]}, ..., {sampled[-1]} of the same run "
                      "(600 steps x 3 launches each, replayed from a captured HIP graph; "
                      "device time)",
            "persistent_same_hours": n * len(sampled) / (q_ms / 1e3),
            "speedup": p_ms / q_ms, "final_counts_identical": same}
        log(f"[bench] per-step path sample: {p_ms / len(sampled):.0f} ms/h vs "
            f"{q_ms / len(sampled):.0f} ms/h persistent")
    if not args.no_surrogate:
        sur = surrogate_sample(graph_steps=args.surrogate_graph)
        out["surrogate"] = sur
        # One GA fitness evaluation = 3 tumours x 20 replicates x 144 h.
        plain_evals = value / (3 * 20 * HOURS)
        out["ga_evals_per_s"] = {
            "plain_simulator": plain_evals, "surrogate_screen": sur["screen_rows_per_s"] / 3,
            "note": "an evaluation is 3 tumours x 20 replicates x 144 h on the simulator, or 3 "
                    "surrogate rows; the GA's time to the plain GA's best fitness is measured "
                    "by emt6ro.ga.experiment compare"}
    if not args.no_cpu:
        log("[bench] CPU baseline")
        out["cpu_baseline"] = cpu_baseline(params, wl, HOURS)
    line = json.dumps(out)
    print(line, flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            f.write(line + "\n")


if __name__ == "__main__":
    main()
