"""Writes tests/golden/surrogate_fixture.npz: simulator-generated training
data for the surrogate tests (needs a GPU; run from the repository root).

Three seeded synthetic 53x53 tumours (the first three of
emt6ro.ga.experiment.tumours(params, 10, seed=0), the tumours the GA
comparison uses) x P random dose protocols (half from the GA's
initial-population generator, half from the batched workload's generator),
10 replicates each over 144 h on the persistent path.  Stored per
(protocol, tumour): the mean and sd of the final living count; per
protocol: the 144 dose slots as multiples of 0.25 Gy (uint8); per tumour:
the 8-value descriptor.

    python tests/golden/make_surrogate_fixture.py [--protocols 2400]
"""
import argparse
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, os.path.join(ROOT, "emt6-ro_amd"))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--protocols", type=int, default=2400)
    ap.add_argument("--reps", type=int, default=10)
    ap.add_argument("--chunk", type=int, default=400)
    ap.add_argument("--protocol-seed", type=int, default=700)
    ap.add_argument("--out", default=os.path.join(ROOT, "tests", "golden",
                                                  "surrogate_fixture.npz"))
    a = ap.parse_args()
    import emt6ro.simulation as sim
    from emt6ro.ga import experiment as X
    params = sim.load_parameters(X.DEFAULT_PARAMS)
    ts, desc = X.tumours(params, 10, 0)
    ts, desc = ts[:3], desc[:3]
    genomes = X.make_protocols(a.protocols, a.protocol_seed)
    units = np.rint(genomes / 0.25)
    if not np.allclose(units * 0.25, genomes) or units.max() > 255:
        raise SystemExit("doses are not multiples of 0.25 Gy below 64 Gy")
    P, T = len(genomes), len(ts)
    mean = np.zeros((P, T), np.float32)
    sd = np.zeros((P, T), np.float32)
    t0 = time.perf_counter()
    for p0 in range(0, P, a.chunk):
        p1 = min(P, p0 + a.chunk)
        m, s, ms = X.simulate_means(params, ts, genomes[p0:p1], a.reps, seed=0,
                                    epoch=1000 + p0)
        mean[p0:p1], sd[p0:p1] = m, s
        print(f"protocols {p0}-{p1}: {ms / 1e3:.1f} s GPU, mean count {m.mean():.1f}",
              flush=True)
    np.savez_compressed(a.out, dose_units=units.astype(np.uint8), mean=mean, sd=sd,
                        descriptors=desc.astype(np.float32), reps=a.reps, hours=X.HOURS,
                        protocol_seed=a.protocol_seed)
    print(f"{P * T} samples ({P * T * a.reps} simulations) in "
          f"{time.perf_counter() - t0:.1f} s -> {a.out} ({os.path.getsize(a.out)} bytes)",
          flush=True)


if __name__ == "__main__":
    main()
