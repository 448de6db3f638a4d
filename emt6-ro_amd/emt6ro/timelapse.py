"""Per-step readout of simulations: the reference's timelapse example
(example/timelapse.cc:23-90).

Each record is 13 numbers of one simulation's state: the three substrate
levels (cho, ox, gi) at a probe site (the lattice centre by default,
(26, 26) on 53 x 53), the number of occupied interior sites, the occupied
sites per metabolic mode (0-3) and per cycle phase (G1, S, G2, M, D).

  python -m emt6ro.timelapse --tumor data/tumor-lib/tumor-4.txt \\
      --params data/default-parameters.json --sims 4 --steps 2400 --out tl.csv
"""
import argparse
import sys

import numpy as np

COLUMNS = ("cho", "ox", "gi", "occupied", "mode0", "mode1", "mode2", "mode3",
           "g1", "s", "g2", "m", "d")


def record(state, probe=None):
    """The 13 numbers of a TumorState (float64), counted over the interior
    rows and columns 1..H-2 and 1..W-2 as timelapse.cc:29-41."""
    occ = np.asarray(state.occupancy())
    H, W = occ.shape
    r, c = probe if probe is not None else (H // 2, W // 2)
    out = np.zeros(len(COLUMNS))
    out[0] = float(state.cho()[r, c])
    out[1] = float(state.ox()[r, c])
    out[2] = float(state.gi()[r, c])
    inner = np.zeros_like(occ, dtype=bool)
    inner[1:H - 1, 1:W - 1] = True
    cell = (occ == 1) & inner
    out[3] = int(cell.sum())
    mode = np.asarray(state.mode())[cell]
    phase = np.asarray(state.phase())[cell]
    out[4:8] = np.bincount(mode, minlength=4)[:4]
    out[8:13] = np.bincount(phase, minlength=5)[:5]
    return out


def run(batch, steps, every=1, probe=None, sims=None):
    """Advance `batch` (emt6ro.batch.Batch) `steps` steps, recording every
    `every` steps (and before the first); returns (n_sims, records, 13)."""
    idx = list(range(batch.n)) if sims is None else list(sims)
    rows = [[record(batch.state(i), probe) for i in idx]]
    done = 0
    while done < steps:
        k = min(every, steps - done)
        batch.run(k)
        done += k
        rows.append([record(batch.state(i), probe) for i in idx])
    return np.transpose(np.asarray(rows), (1, 0, 2))


def main(argv=None):
    from . import simulation as sim
    from .batch import Batch
    ap = argparse.ArgumentParser(description="Per-step readout of a batch of one tumour.")
    ap.add_argument("--tumor", required=True)
    ap.add_argument("--params", required=True)
    ap.add_argument("--sims", type=int, default=4)
    ap.add_argument("--steps", type=int, default=2400)
    ap.add_argument("--every", type=int, default=1)
    ap.add_argument("--dose", action="append", default=[],
                    help="STEP:GY, on the 300-step (30 min) protocol grid; repeatable")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--out", default="-")
    a = ap.parse_args(argv)
    params = sim.load_parameters(a.params)
    tumor = sim.load_state(a.tumor, params)
    horizon = 144000
    doses = np.zeros((1, horizon // 300), np.float32)
    for d in a.dose:
        step, gy = d.split(":")
        doses[0, int(step) // 300] += float(gy)
    n = a.sims
    b = Batch(params, [tumor], [0] * n, [0] * n, [horizon] * n, doses, 300, horizon,
              sim_id=np.arange(n), seed=a.seed)
    tl = run(b, a.steps, a.every)
    out = sys.stdout if a.out == "-" else open(a.out, "w")
    out.write("sim,record," + ",".join(COLUMNS) + "\n")
    for s in range(tl.shape[0]):
        for t in range(tl.shape[1]):
            out.write(f"{s},{t}," + ",".join(f"{v:.9g}" for v in tl[s, t]) + "\n")
    if out is not sys.stdout:
        out.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
