#!/usr/bin/env python3
"""Writes rules_ab.json: the single-tumour workload of config1_oracle.json
(one seeded synthetic 53x53 tumour, disk radius 10; 1.25 Gy at 12 h and at
36 h; 144 h) run by the CPU oracle under two sets of step rules, with
replicate streams 0..N-1 of key seed 0:

  rules 0  the deterministic rules the GPU path implements (each occupied
           site stepped once per step; a division target must be vacant after
           every death of the step);
  rules 3  the reference's literal behaviour for those two points: every list
           entry of a re-occupied site steps it again (simulation.cu:86-88,
           :105), and a ready cell's target is tested right after its own
           step, in list order (simulation.cu:88-93).

Usage: python tests/golden/make_rules_ab.py [replicates] [threads] [rules ...]
(1000 replicates per rule set take about 25 min on 8 threads).
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "emt6-ro_amd"))
sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import oracle as O  # noqa: E402
from emt6ro.synthetic import disk_arrays  # noqa: E402
from make_config1 import CONFIG1, oracle_state  # noqa: E402


def main():
    n = int(sys.argv[1]) if len(sys.argv) > 1 else 1000
    threads = int(sys.argv[2]) if len(sys.argv) > 2 else os.cpu_count()
    rule_sets = [int(r) for r in sys.argv[3:]] or [0, 3]
    c = CONFIG1
    values = O.read_params_json(os.path.join(HERE, "default-parameters.json"))
    arrays = disk_arrays(values, c["size"], c["size"], c["radius"], seed=c["tumor_seed"])
    state = oracle_state(values, arrays)
    n_slots = c["steps"] // c["resolution"]
    doses = np.zeros((1, n_slots), np.float32)
    for step, d in c["doses"]:
        doses[0, step // c["resolution"]] += np.float32(d)
    path = os.path.join(HERE, "rules_ab.json")
    out = json.load(open(path)) if os.path.exists(path) else {}
    out["workload"] = {k: c[k] for k in ("size", "radius", "tumor_seed", "doses", "steps",
                                         "key_seed", "resolution")}
    for rules in rule_sets:
        key = f"rules_{rules}"
        rec = out.get(key, {"counts": [], "seconds": 0.0, "threads": threads})
        # chunks of 100 replicates, saved as they finish (a rerun resumes)
        while len(rec["counts"]) < n:
            lo = len(rec["counts"])
            hi = min(n, lo + 100)
            b = O.OracleBatch(values, [state] * (hi - lo), doses, [0] * (hi - lo),
                              np.arange(lo, hi), c["resolution"], c["steps"])
            t0 = time.time()
            b.run(c["steps"], seed=c["key_seed"], nthreads=threads, rules=rules)
            rec["counts"] += b.counts().tolist()
            rec["seconds"] += time.time() - t0
            rec["mean"] = float(np.mean(rec["counts"]))
            rec["sd"] = float(np.std(rec["counts"], ddof=1))
            out[key] = rec
            with open(path, "w") as f:
                json.dump(out, f)
            print(json.dumps({"rules": rules, "n": len(rec["counts"]), "mean": rec["mean"],
                              "sd": rec["sd"], "seconds": rec["seconds"]}), flush=True)


if __name__ == "__main__":
    main()
