"""tumor-ca JSON <-> text state conversion (the reference's
tools/convert_state.py:69-289 and its test tools/test_convert_state_txt.py:
43-101).  No tumor-ca document ships with the reference, so the JSON side
is pinned by a hand-worked 3x3 document (the MATLAB index arithmetic done
by hand) and by round trips through the tumour library."""
import json
import os

import numpy as np
import pytest

from conftest import golden
from emt6ro import convert_state as cs

TUMOURS = [golden("tumor-lib", f"tumor-{i}.txt") for i in range(1, 11)]


def _doc3():
    # L = 3.  MATLAB index = c * 3 + r + 1: 2 -> (1, 0), 7 -> (0, 2), 9 -> (2, 2).
    g = lambda v: [[v] * 3 for _ in range(3)]  # noqa: E731
    W = g(0)
    W[1][0] = W[0][2] = W[2][2] = 1
    st = {"W": W, "CHO": [[1, 2, 3], [4, 5, 6], [7, 8, 9]], "OX": g(0.5), "GI": g(0.25),
          "RepT": g(0.0), "R": g(0.0), "HRS": g(0.0), "gMET": g(0),
          "lMET": {"prolif": [2], "prolif_an": [], "quiesc": [], "quiesc_an": [7], "dead": [9]}}
    st["RepT"][1][0] = 1.5
    st["R"][1][0] = 2.0
    st["HRS"][0][2] = 7.25
    st["gMET"][1][0] = 3
    st["gMET"][0][2] = 1
    cyc = {k: g(0.0) for k in cs.CYCLE_FIELDS}
    for n, k in enumerate(cs.CYCLE_FIELDS):
        cyc[k][1][0] = 10.0 * (n + 1)
        cyc[k][0][2] = 100.0 + n
    return {"PARAMS": {"L": 3, "other": "kept"}, "STATE": st, "CYCLES": cyc}


def test_matlab_index_is_column_major_one_based():
    n = 7
    seen = set()
    for r in range(n):
        for c in range(n):
            i = cs.matlab_index(r, c, n)
            assert cs.from_matlab_index(i, n) == (r, c)
            seen.add(i)
    assert seen == set(range(1, n * n + 1))
    assert cs.matlab_index(1, 0, 3) == 2 and cs.matlab_index(0, 1, 3) == 4


def test_hand_worked_document(tmp_path):
    p = tmp_path / "t.json"
    p.write_text(json.dumps(_doc3()))
    out = tmp_path / "t.txt"
    info = cs.json_to_txt(p, out)
    assert info == {"L": 3, "n_occupied": 2, "n_vacant": 7, "n_dead_skipped": 1,
                    "n_w0_with_mode": 0, "scale": 1.0, "have_cycles": True}
    lines = out.read_text().split("\n")
    # (0,0) vacant; (0,1) vacant; (0,2) quiescent anaerobic (mode 3), phase 1;
    # (1,0) aerobic proliferating (mode 0), phase 3; (2,2) dead -> vacant.
    expect = (["3", "3"] + ["0", "1", "0.5", "0.25"] + ["0", "2", "0.5", "0.25"]
              + ["1", "3", "0.5", "0.25", "0", "0", "7.25", "100", "101", "102", "103", "104",
                 "3", "1"]
              + ["1", "4", "0.5", "0.25", "1.5", "2", "0", "10", "20", "30", "40", "50", "0", "3"]
              + ["0", "5", "0.5", "0.25"] + ["0", "6", "0.5", "0.25"]
              + ["0", "7", "0.5", "0.25"] + ["0", "8", "0.5", "0.25"] + ["0", "9", "0.5", "0.25"]
              + [""])
    assert lines == expect


def test_scale_applies_to_substrates_only():
    s, _ = cs.from_tumorca(_doc3(), scale=1e15)
    assert s.f["CHO"][2, 1] == 8e15 and s.f["OX"][0, 0] == 0.5e15
    assert s.f["HRS"][0, 2] == 7.25 and s.f["t_D"][1, 0] == 50.0


def test_missing_cycles_and_inconsistent_documents():
    d = _doc3()
    del d["CYCLES"]
    s, info = cs.from_tumorca(d)
    assert not info["have_cycles"] and s.f["t_G1"].max() == 0.0
    d = _doc3()
    d["STATE"]["lMET"]["prolif"] = []
    with pytest.raises(ValueError, match="no lMET entry"):
        cs.from_tumorca(d)
    d = _doc3()
    d["STATE"]["lMET"]["quiesc"] = [10]
    with pytest.raises(ValueError, match="outside"):
        cs.from_tumorca(d)
    d = _doc3()
    d["STATE"]["W"][1][0] = 0
    _, info = cs.from_tumorca(d)
    assert info["n_w0_with_mode"] == 1 and info["n_occupied"] == 1
    s = cs.TextState(2, 3)
    with pytest.raises(ValueError, match="non-square"):
        cs.to_tumorca(s, {"PARAMS": {"L": 2}})


@pytest.mark.parametrize("path", TUMOURS, ids=[os.path.basename(p) for p in TUMOURS])
def test_text_json_text_round_trip(path, tmp_path):
    """tools/test_convert_state_txt.py:43-101, with exact equality."""
    tpl = tmp_path / "tpl.json"
    tpl.write_text(json.dumps({"PARAMS": {"L": 0}}))
    mid, back = tmp_path / "mid.json", tmp_path / "back.txt"
    info1 = cs.txt_to_json(path, mid, tpl)
    info2 = cs.json_to_txt(mid, back)
    a, b = cs.TextState.read(path), cs.TextState.read(back)
    assert info1["n_occupied"] == info2["n_occupied"] == int(a.occupied.sum()) > 0
    assert (a.H, a.W) == (b.H, b.W) == (info1["L"], info1["L"])
    assert np.array_equal(a.occupied, b.occupied)
    assert np.array_equal(a.mode, b.mode) and np.array_equal(a.phase, b.phase)
    for k in a.f:
        assert np.array_equal(a.f[k], b.f[k]), k
    assert json.loads(mid.read_text())["PARAMS"]["L"] == a.H


def test_json_round_trip(tmp_path):
    """tools/convert_state.py:292-379 (roundtrip-test) on a document made from
    a library tumour, with dead cells added."""
    tpl = tmp_path / "tpl.json"
    tpl.write_text(json.dumps({"PARAMS": {"L": 0}, "extra": [1, 2]}))
    doc_path = tmp_path / "doc.json"
    cs.txt_to_json(TUMOURS[2], doc_path, tpl)
    doc = json.loads(doc_path.read_text())
    L = doc["PARAMS"]["L"]
    doc["STATE"]["lMET"]["dead"] = [1, L * L]       # corners, vacant in the library tumour
    assert doc["STATE"]["W"][0][0] == 0 and doc["STATE"]["W"][L - 1][L - 1] == 0
    doc_path.write_text(json.dumps(doc))
    assert cs.roundtrip_failures(doc_path, tmpdir=tmp_path) == []
    assert cs.main(["roundtrip-test", str(doc_path), "--tmpdir", str(tmp_path)]) == 0
    # A listed cell on a site with W = 0 does not survive the text format,
    # and the round trip reports it.
    name = next(n for n in cs.LMET_MODES if doc["STATE"]["lMET"][n])
    r, c = cs.from_matlab_index(doc["STATE"]["lMET"][name][0], L)
    doc["STATE"]["W"][r][c] = 0
    doc_path.write_text(json.dumps(doc))
    fails = cs.roundtrip_failures(doc_path, tmpdir=tmp_path)
    assert len(fails) == 1 and fails[0].startswith(f"lMET['{name}']: 1 only before")


def test_converted_text_loads_like_the_original(tmp_path, default_params_path):
    """The library's own reader sees the same tumour after JSON -> text."""
    sim = pytest.importorskip("emt6ro.simulation")
    tpl = tmp_path / "tpl.json"
    tpl.write_text(json.dumps({"PARAMS": {"L": 0}}))
    mid, back = tmp_path / "mid.json", tmp_path / "back.txt"
    cs.txt_to_json(TUMOURS[0], mid, tpl)
    cs.json_to_txt(mid, back)
    params = sim.load_parameters(default_params_path)
    a, b = sim.load_state(TUMOURS[0], params), sim.load_state(str(back), params)
    for f in ("irradiation", "occupancy", "cho", "ox", "gi", "mode", "phase"):
        assert np.array_equal(getattr(a, f)(), getattr(b, f)()), f
