"""The single-tumour workload at full size on the GPU (one seeded synthetic
53x53 tumour, occupied disk of radius 10; 1.25 Gy at 12 h and 36 h; 144 h =
86,400 steps) against the CPU oracle's 100 replicates in
tests/golden/config1_oracle.json.

- Replicates on the oracle's streams (0..99) give the same counts, bit for
  bit, after 86,400 steps.
- 1,000 replicates on independent streams (1000..1999) match the oracle's
  distribution: two-sample KS test at the 1% level and the mean within 2%.
"""
import json
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def cfg():
    with open(golden("config1_oracle.json")) as f:
        return json.load(f)


def _batch(cfg, sim_ids):
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    from emt6ro.synthetic import disk_state
    params = sim.load_parameters(golden("default-parameters.json"))
    st, _ = disk_state(params, cfg["size"], cfg["size"], radius=cfg["radius"],
                       seed=cfg["tumor_seed"])
    res = cfg["resolution"]
    doses = np.zeros((1, cfg["steps"] // res), np.float32)
    for step, d in cfg["doses"]:
        doses[0, step // res] += np.float32(d)
    n = len(sim_ids)
    return Batch(params, [st], np.zeros(n, np.int32), np.zeros(n, np.int32),
                 np.full(n, cfg["steps"], np.uint32), doses, res, cfg["steps"],
                 sim_id=np.asarray(sim_ids, np.uint32), seed=cfg["key_seed"])


def test_same_streams_are_bit_identical(cfg):
    R = cfg["replicates"]
    b = _batch(cfg, np.arange(R))
    log(f"[config1] {R} replicates x {cfg['steps']} steps, oracle streams")
    ms = b.run(cfg["steps"])
    got = b.results().astype(np.int64)
    log(f"  {ms:.0f} ms; mean {got.mean():.2f} (oracle {cfg['mean']:.2f})")
    assert np.array_equal(b.counters()["step"], np.full(R, cfg["steps"]))
    np.testing.assert_array_equal(got, np.asarray(cfg["counts"], np.int64))


def test_independent_streams_match_distribution(cfg):
    from scipy.stats import ks_2samp
    ref = np.asarray(cfg["counts"], np.float64)
    b = _batch(cfg, np.arange(1000, 2000))
    log("[config1] 1000 replicates, independent streams")
    b.run(cfg["steps"])
    got = b.results().astype(np.float64)
    ks = ks_2samp(got, ref)
    rel = abs(got.mean() - ref.mean()) / ref.mean()
    log(f"  GPU mean {got.mean():.2f} sd {got.std(ddof=1):.2f}; oracle mean {ref.mean():.2f} "
        f"sd {ref.std(ddof=1):.2f}; KS D={ks.statistic:.4f} p={ks.pvalue:.3f}; "
        f"mean diff {100 * rel:.2f}%")
    assert ks.pvalue > 0.01
    assert rel < 0.02
