"""The replay harness (emt6ro.replay, after tools/run_protocols_gpu.py) on
the GPU: chunked batched runs reproduce one Experiment over all protocols
with the same seed, and the output file has the reference's row layout."""
import sys

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_replay_matches_experiment_and_writes_rows(tmp_path):
    import emt6ro.simulation as sim
    from emt6ro import protocols as P
    from emt6ro import replay as R
    params = sim.load_parameters(golden("default-parameters.json"))
    paths = [golden("tumor-lib", f"tumor-{t}.txt") for t in (1, 2, 3)]
    tumors = [sim.load_state(p, params) for p in paths]
    prots = [[(1, 1.25), (3001, 1.25)], [(0, 2.0)], [(600, 0.5), (600, 0.5)], [(1200, 3.0)]]
    pf = tmp_path / "prot.txt"
    P.write_two_line(prots, pf)
    canon = P.canonical(P.parse_two_line(pf))
    steps, reps, seed = 3600, 4, 77
    log("[replay] chunked replay vs one Experiment")
    got = R.replay(params, tumors, canon, reps, steps, 300, seed, max_sims=2 * 3 * reps)
    e = sim.Experiment(params, tumors, reps, len(canon), simulation_steps=steps,
                       protocol_resolution=300, seed=seed)
    e.add_irradiations(canon)
    e.run(steps)
    np.testing.assert_array_equal(got, e.get_results())
    out = tmp_path / "res.txt"
    assert R.main(["--protocols", str(pf), "--tumors", *paths, "--params",
                   golden("default-parameters.json"), "--reps", str(reps), "--steps", str(steps),
                   "--seed", str(seed), "--out", str(out), "--max-sims", "12"]) == 0
    rows = [list(map(int, line.split())) for line in out.read_text().splitlines()]
    assert len(rows) == len(canon) == 4 and all(len(r) == 3 * reps for r in rows)
    np.testing.assert_array_equal(np.array(rows), got.reshape(len(canon), -1))
