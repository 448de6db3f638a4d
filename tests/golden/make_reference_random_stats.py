#!/usr/bin/env python3
"""Writes reference_random_stats.json: summary statistics of the reference's
shipped simulation results (simulation_analysis/results-gpu-random.txt and
results-cpu-random.txt, read as data): 60 random dose protocols, each on the
10 library tumours (data/tumor-lib) for 240 h, 100 (GPU) or 48 (CPU)
replicates per tumour.  Per protocol: the mean and sd of the final living
counts pooled over tumours and replicates (results-comparison.ipynb), and
the mean over tumours of the sd within a tumour.

Usage: python tests/golden/make_reference_random_stats.py  (needs /root/reference)
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/simulation_analysis"


def read(name, n, sep):
    rows = []
    with open(os.path.join(SRC, name)) as f:
        for line in f:
            r = [float(x) for x in line.split(sep)[:-1]]
            if len(r) == n:
                rows.append(np.array(r).reshape(10, -1))
    return rows


def stats(rows):
    return {"mean": [float(r.mean()) for r in rows], "sd": [float(r.std()) for r in rows],
            "within_tumour_sd": [float(r.std(axis=1).mean()) for r in rows],
            "replicates_per_tumour": int(rows[0].shape[1])}


def main():
    out = {"source": "simulation_analysis/results-{gpu,cpu}-random.txt",
           "gpu": stats(read("results-gpu-random.txt", 1000, " ")),
           "cpu": stats(read("results-cpu-random.txt", 480, ", "))}
    with open(os.path.join(HERE, "reference_random_stats.json"), "w") as f:
        json.dump(out, f)
    print({k: (min(v["mean"]), max(v["mean"])) for k, v in out.items() if k != "source"})


if __name__ == "__main__":
    main()
