#!/usr/bin/env python3
"""Writes config1_oracle.json: the CPU oracle's final living-cell counts for
the single-tumour workload (one seeded synthetic 53x53 tumour, occupied disk
of radius 10; 1.25 Gy at 12 h and at 36 h; 144 h = 86,400 steps;
replicates with random streams 0..99, key seed 0).

The tumour comes from emt6ro.synthetic.disk_arrays(values, 53, 53, 10,
seed=0); this script rebuilds it from the same function and runs only the
oracle.  Usage: python tests/golden/make_config1.py [threads]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "emt6-ro_amd"))

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from emt6ro.synthetic import disk_arrays  # noqa: E402

CONFIG1 = {"size": 53, "radius": 10, "tumor_seed": 0, "doses": [[7200, 1.25], [21600, 1.25]],
           "steps": 86400, "replicates": 100, "key_seed": 0, "resolution": 300}


def oracle_state(values, arrays):
    H, W = arrays["occupancy"].shape
    s = O.empty_state(values, H, W)
    s["sub"][0], s["sub"][1], s["sub"][2] = arrays["cho"], arrays["ox"], arrays["gi"]
    s["st"][:] = arrays["occupancy"]
    cells = s["cells"].reshape(H, W)
    occ = arrays["occupancy"] == 1
    cells["t_rep"][occ] = arrays["time_in_repair"][occ]
    cells["R"][occ] = arrays["irradiation"][occ]
    cells["delay"][occ] = arrays["repair_delay"][occ]
    cells["prolif"][occ] = arrays["proliferation_time"][occ]
    for k, n in enumerate(("cycle_g1", "cycle_s", "cycle_g2", "cycle_m", "cycle_d")):
        cells["times"][occ, k] = arrays[n][occ]
    cells["mode"][occ] = arrays["mode"][occ]
    cells["phase"][occ] = arrays["phase"][occ]
    return s


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else os.cpu_count()
    c = CONFIG1
    values = O.read_params_json(os.path.join(HERE, "default-parameters.json"))
    arrays = disk_arrays(values, c["size"], c["size"], c["radius"], seed=c["tumor_seed"])
    state = oracle_state(values, arrays)
    n_slots = c["steps"] // c["resolution"]
    doses = np.zeros((1, n_slots), np.float32)
    for step, d in c["doses"]:
        doses[0, step // c["resolution"]] += np.float32(d)
    R = c["replicates"]
    b = O.OracleBatch(values, [state] * R, doses, [0] * R, np.arange(R), c["resolution"],
                      c["steps"])
    t0 = time.time()
    b.run(c["steps"], seed=c["key_seed"], nthreads=threads)
    counts = b.counts().tolist()
    out = dict(c, counts=counts, mean=float(np.mean(counts)), sd=float(np.std(counts, ddof=1)),
               seconds=time.time() - t0, threads=threads)
    with open(os.path.join(HERE, "config1_oracle.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps({k: out[k] for k in ("mean", "sd", "seconds", "threads")}))


if __name__ == "__main__":
    main()
