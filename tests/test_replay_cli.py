"""emt6ro.replay's command line without a GPU: the simulations are replaced
by a stand-in that records its arguments, so what is checked is the
plumbing around them (after tools/run_protocols_gpu.py): the seed a run
used is the one its metadata records, drawn when --seed is not given; one
output row per protocol, tumour-major then replicate-major."""
import json

import numpy as np

from conftest import golden


def _run(tmp_path, monkeypatch, extra):
    from emt6ro import protocols as P
    from emt6ro import replay as R
    seen = {}

    def fake(params, tumors, protocols, reps, steps, resolution, seed, max_sims, log=None):
        seen.update(n_prot=len(protocols), n_tum=len(tumors), reps=reps, steps=steps,
                    resolution=resolution, seed=seed, max_sims=max_sims)
        # count = 1000 p + 10 t + r, in the (P, T, R) order replay() returns
        p, t, r = np.meshgrid(np.arange(len(protocols)), np.arange(len(tumors)),
                              np.arange(reps), indexing="ij")
        return (1000 * p + 10 * t + r).astype(np.uint32)

    monkeypatch.setattr(R, "replay", fake)
    pf = tmp_path / "prot.txt"
    P.write_two_line([[(0, 1.0)], [(600, 2.0), (1200, 0.5)]], pf)
    out = tmp_path / "res.txt"
    paths = [golden("tumor-lib", f"tumor-{t}.txt") for t in (1, 2)]
    assert R.main(["--protocols", str(pf), "--tumors", *paths, "--params",
                   golden("default-parameters.json"), "--reps", "3", "--steps", "1200",
                   "--out", str(out), *extra]) == 0
    rows = [list(map(int, line.split())) for line in out.read_text().splitlines()]
    meta = json.loads((tmp_path / "res.txt.meta.json").read_text())
    return seen, rows, meta


def test_replay_records_the_seed_it_drew(tmp_path, monkeypatch):
    seen, rows, meta = _run(tmp_path, monkeypatch, [])
    assert isinstance(meta["seed"], int) and meta["seed"] == seen["seed"]
    assert (meta["n_protocols"], meta["n_tumors"], meta["reps_per_tumor"]) == (2, 2, 3)
    assert rows == [[1000 * p + 10 * t + r for t in range(2) for r in range(3)]
                    for p in range(2)]


def test_replay_passes_the_given_seed(tmp_path, monkeypatch):
    seen, _, meta = _run(tmp_path, monkeypatch, ["--seed", "7", "--max-sims", "6"])
    assert seen["seed"] == 7 and meta["seed"] == 7
    assert seen["max_sims"] == 6 and seen["steps"] == 1200 and seen["resolution"] == 300
