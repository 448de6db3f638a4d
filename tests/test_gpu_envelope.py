"""Distribution-level check against the reference's own shipped results
(GPU).

The reference validated its GPU simulator on 60 random dose protocols, each
run on the 10 library tumours for 240 h (simulation_analysis/
results-{gpu,cpu}-random.txt, summarised in tests/golden/
reference_random_stats.json).  Its protocols are not shipped, so this test
draws 30 protocols from the GA's random-protocol generator
(python_ga/utils.py:66-126, whose restatement emt6ro.ga.operators is pinned
against the reference's own by tests/test_ga_vectors.py), runs the same 10
tumours for 240 h with 20 replicates each, and compares what does not
depend on which protocols were drawn:

  * the spread within one tumour and protocol (replicate noise): the mean
    within-tumour sd must be within 8% of the reference GPU's;
  * every per-protocol mean and sd (pooled over tumours and replicates, as
    results-comparison.ipynb computes them) must lie in the reference's range
    widened by 5%;
  * the per-protocol means must not be distinguishable from the reference's
    by a two-sample KS test (p > 0.01).
"""
import json
import sys

import numpy as np
import pytest

from conftest import golden, gpu_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not gpu_available(), reason="needs an MI355X GPU")]

HOURS, P, REPS = 240, 30, 20


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_library_tumours_under_random_protocols_match_reference_envelope():
    from scipy import stats
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    from emt6ro.ga import operators as ops
    ref = json.load(open(golden("reference_random_stats.json")))["gpu"]
    params = sim.load_parameters(golden("default-parameters.json"))
    tumors = [sim.load_state(golden("tumor-lib", f"tumor-{i}.txt"), params)
              for i in range(1, 11)]
    cfg = ops.default_config()
    # 5 days of 0.5 h slots, 10 Gy in doses of 0.25-2.5 Gy at least 3 h apart
    g = np.asarray(ops.random_population(np.random.default_rng(123), P, 240, cfg), np.float32)
    horizon = HOURS * 600
    doses = np.zeros((P, horizon // 300), np.float32)
    doses[:, :240] = g
    T = len(tumors)
    i = np.arange(P * T * REPS)
    b = Batch(params, tumors, (i // REPS) % T, i // (T * REPS),
              np.full(len(i), horizon, np.uint32), doses, 300, horizon, seed=7)
    log(f"[envelope] {len(i)} simulations x {HOURS} h")
    ms = b.run(horizon)
    c = b.results().astype(np.float64).reshape(P, T, REPS)
    log(f"[envelope] {ms / 1e3:.1f} s on the GPU")
    mean = c.reshape(P, -1).mean(1)
    sd = c.reshape(P, -1).std(1)
    within = float(c.std(axis=2).mean())
    rmean, rsd = np.array(ref["mean"]), np.array(ref["sd"])
    rwithin = float(np.mean(ref["within_tumour_sd"]))
    ks = stats.ks_2samp(mean, rmean).pvalue
    log(f"[envelope] per-protocol means {mean.min():.1f}-{mean.max():.1f} (reference "
        f"{rmean.min():.1f}-{rmean.max():.1f}), sds {sd.min():.1f}-{sd.max():.1f} (reference "
        f"{rsd.min():.1f}-{rsd.max():.1f}), within-tumour sd {within:.2f} (reference "
        f"{rwithin:.2f}), KS p {ks:.3f}")
    assert abs(within - rwithin) <= 0.08 * rwithin
    assert mean.min() >= 0.95 * rmean.min() and mean.max() <= 1.05 * rmean.max()
    assert sd.min() >= 0.95 * rsd.min() and sd.max() <= 1.05 * rsd.max()
    assert ks > 0.01
