"""Tumour states between tumor-ca's JSON layout and the EMT6-Ro text format
(the reference's tools/convert_state.py:69-289).

The text format is the one `load_state` reads and `TumorState.save` writes
(state/state.cc:9-82): height, width, then per site in row-major order the
occupancy flag, the three substrate levels and, for an occupied site, the
repair time, the accumulated dose, the proliferation time, the five cycle
phase boundaries (G1, S, G2, M, D), the metabolic mode and the cycle phase.

tumor-ca (an external CPU simulator whose source the reference does not
ship) keeps an L x L grid per field under "STATE" ("W" occupancy, "CHO",
"OX", "GI", "RepT", "R", "HRS", "gMET" phase), the cycle boundaries under
"CYCLES" ("t_G1" ... "t_D"), and the metabolic mode of each cell as
membership of one of the "lMET" lists, whose entries are MATLAB indices:
column-major and 1-based, index = c * L + r + 1.  Cells in the "dead" list
are vacant sites in EMT6-Ro.

Mapping (tools/convert_state.py:31-53): lMET "prolif" -> mode 0 (aerobic
proliferation), "prolif_an" -> 1, "quiesc" -> 2, "quiesc_an" -> 3; cycle
phases use the same integers 0..4 (G1, S, G2, M, D) on both sides.
Substrates can be scaled on the way (tumor-ca stores mol, EMT6-Ro about
fmol: scale 1e15 one way, 1e-15 the other); with scale 1 a round trip
returns every value, floats being written with 9 significant digits.

Command line (the reference tool's):

  python -m emt6ro.convert_state json-to-txt in.json out.txt [--scale S]
  python -m emt6ro.convert_state txt-to-json in.txt out.json --template t.json [--scale S]
  python -m emt6ro.convert_state roundtrip-test in.json
"""
import argparse
import json
import os
import sys
import tempfile

import numpy as np

LMET_MODES = ("prolif", "prolif_an", "quiesc", "quiesc_an")   # mode 0..3
CELL_FIELDS = ("RepT", "R", "HRS")                              # text order
CYCLE_FIELDS = ("t_G1", "t_S", "t_G2", "t_M", "t_D")
SUBSTRATES = ("CHO", "OX", "GI")


def matlab_index(r, c, n):
    """1-based column-major index of site (r, c) of an n x n grid."""
    return c * n + r + 1


def from_matlab_index(idx, n):
    """(r, c) of a 1-based column-major index."""
    q = int(idx) - 1
    return q % n, q // n


def _fmt(v):
    return f"{float(v):.9g}"


class TextState:
    """A tumour in the text format, as NumPy arrays of shape (H, W):
    `occupied` (bool), the substrates `CHO`, `OX`, `GI`, the cell fields
    `RepT`, `R`, `HRS`, the cycle boundaries `t_G1`..`t_D` (float64) and
    `mode`, `phase` (int).  Cell fields of vacant sites are 0."""

    def __init__(self, height, width):
        self.H, self.W = int(height), int(width)
        shape = (self.H, self.W)
        self.occupied = np.zeros(shape, bool)
        self.f = {k: np.zeros(shape) for k in SUBSTRATES + CELL_FIELDS + CYCLE_FIELDS}
        self.mode = np.zeros(shape, np.int64)
        self.phase = np.zeros(shape, np.int64)

    @classmethod
    def read(cls, path):
        with open(path) as fh:
            tok = fh.read().split()
        pos = 0

        def take(n):
            nonlocal pos
            if pos + n > len(tok):
                raise ValueError(f"{path}: the file ends early (token {len(tok)})")
            out = tok[pos:pos + n]
            pos += n
            return out

        h, w = (int(t) for t in take(2))
        s = cls(h, w)
        for r in range(h):
            for c in range(w):
                occ, cho, ox, gi = take(4)
                s.f["CHO"][r, c], s.f["OX"][r, c], s.f["GI"][r, c] = float(cho), float(ox), float(gi)
                if int(occ) == 1:
                    s.occupied[r, c] = True
                    vals = take(10)
                    for k, v in zip(CELL_FIELDS + CYCLE_FIELDS, vals[:8]):
                        s.f[k][r, c] = float(v)
                    s.mode[r, c], s.phase[r, c] = int(vals[8]), int(vals[9])
        return s

    def text(self):
        out = [f"{self.H}\n{self.W}\n"]
        for r in range(self.H):
            for c in range(self.W):
                sub = "".join(_fmt(self.f[k][r, c]) + "\n" for k in SUBSTRATES)
                if self.occupied[r, c]:
                    cell = "".join(_fmt(self.f[k][r, c]) + "\n" for k in CELL_FIELDS + CYCLE_FIELDS)
                    out.append(f"1\n{sub}{cell}{int(self.mode[r, c])}\n{int(self.phase[r, c])}\n")
                else:
                    out.append(f"0\n{sub}")
        return "".join(out)

    def write(self, path):
        with open(path, "w") as fh:
            fh.write(self.text())


def from_tumorca(doc, scale=1.0):
    """TextState of a tumor-ca document, and a summary.

    A site is occupied when W is 1, it is in one of the four mode lists and
    not in the "dead" list.  W = 1 with no list entry is an inconsistent
    document (ValueError); a listed site with W = 0 stays vacant and is
    counted in `n_w0_with_mode`.  Without a complete CYCLES block the cycle
    boundaries are written as 0 (`have_cycles` False)."""
    L = int(doc["PARAMS"]["L"])
    st = doc["STATE"]
    occ_w = np.asarray(st["W"], np.int64).reshape(L, L)
    lmet = st.get("lMET", {})
    mode = np.full((L, L), -1, np.int64)
    for m, name in enumerate(LMET_MODES):
        for idx in lmet.get(name, []):
            r, c = from_matlab_index(idx, L)
            if not (0 <= r < L and 0 <= c < L) or int(idx) < 1:
                raise ValueError(f"lMET['{name}'] index {idx} is outside a {L}x{L} grid")
            mode[r, c] = m
    dead = np.zeros((L, L), bool)
    for idx in lmet.get("dead", []):
        r, c = from_matlab_index(idx, L)
        if 0 <= r < L and 0 <= c < L and int(idx) >= 1:
            dead[r, c] = True
    bad = (occ_w == 1) & (mode < 0) & ~dead
    if bad.any():
        r, c = (int(v) for v in np.argwhere(bad)[0])
        raise ValueError(f"site ({r},{c}) has W=1 but no lMET entry: the document is inconsistent")
    occupied = (occ_w == 1) & (mode >= 0) & ~dead
    cyc = doc.get("CYCLES", {}) or {}
    have_cycles = all(cyc.get(k) is not None for k in CYCLE_FIELDS)

    s = TextState(L, L)
    s.occupied = occupied
    for k in SUBSTRATES:
        s.f[k] = np.asarray(st[k], np.float64).reshape(L, L) * scale
    for k in CELL_FIELDS:
        s.f[k] = np.where(occupied, np.asarray(st[k], np.float64).reshape(L, L), 0.0)
    for k in CYCLE_FIELDS:
        s.f[k] = (np.where(occupied, np.asarray(cyc[k], np.float64).reshape(L, L), 0.0)
                  if have_cycles else np.zeros((L, L)))
    s.mode = np.where(occupied, mode, 0)
    s.phase = np.where(occupied, np.asarray(st["gMET"], np.int64).reshape(L, L), 0)
    info = {"L": L, "n_occupied": int(occupied.sum()), "n_vacant": int((~occupied).sum()),
            "n_dead_skipped": int(dead.sum()),
            "n_w0_with_mode": int(((occ_w == 0) & (mode >= 0) & ~dead).sum()),
            "scale": scale, "have_cycles": bool(have_cycles)}
    return s, info


def to_tumorca(state, template, scale=1.0):
    """tumor-ca document of a square TextState.  The template's top-level
    fields (PARAMS in particular) are kept, with PARAMS.L set to the grid
    size; STATE and CYCLES are rebuilt.  Mode lists hold the occupied sites
    in row-major order; "dead" is empty and the "ch" (cycle changed) grid is
    0, as at step 0."""
    if "PARAMS" not in template:
        raise ValueError("the template has no PARAMS block")
    if state.H != state.W:
        raise ValueError(f"non-square grid {state.H}x{state.W}: tumor-ca grids are square")
    L = state.H
    doc = json.loads(json.dumps(template))
    doc["PARAMS"]["L"] = L
    occ = state.occupied

    def grid(a, cast=float):
        return [[cast(v) for v in row] for row in a]

    lmet = {name: [] for name in LMET_MODES}
    lmet["dead"] = []
    for r, c in zip(*np.nonzero(occ)):
        m = int(state.mode[r, c])
        if not 0 <= m < len(LMET_MODES):
            raise ValueError(f"unknown metabolic mode {m} at ({r},{c})")
        lmet[LMET_MODES[m]].append(matlab_index(int(r), int(c), L))
    doc["STATE"] = {"W": grid(occ.astype(np.int64), int)}
    for k in SUBSTRATES:
        doc["STATE"][k] = grid(state.f[k] * scale)
    for k in CELL_FIELDS:
        doc["STATE"][k] = grid(np.where(occ, state.f[k], 0.0))
    doc["STATE"]["gMET"] = grid(np.where(occ, state.phase, 0), int)
    doc["STATE"]["lMET"] = lmet
    doc["STATE"]["ch"] = grid(np.zeros((L, L), np.int64), int)
    doc["CYCLES"] = {k: grid(np.where(occ, state.f[k], 0.0)) for k in CYCLE_FIELDS}
    return doc, {"L": L, "n_occupied": int(occ.sum()), "scale": scale}


def json_to_txt(json_path, txt_path, scale=1.0):
    """tools/convert_state.py:69-184: tumor-ca JSON -> text; the summary."""
    with open(json_path) as fh:
        doc = json.load(fh)
    s, info = from_tumorca(doc, scale)
    s.write(txt_path)
    return info


def txt_to_json(txt_path, json_path, template_path, scale=1.0):
    """tools/convert_state.py:187-289: text -> tumor-ca JSON, PARAMS from
    the template; the summary."""
    with open(template_path) as fh:
        template = json.load(fh)
    doc, info = to_tumorca(TextState.read(txt_path), template, scale)
    with open(json_path, "w") as fh:
        json.dump(doc, fh, separators=(",", ":"))
    return info


def roundtrip_failures(json_path, tmpdir=None, tol=1e-6):
    """JSON -> text -> JSON with scale 1, then field-by-field comparison:
    W and the substrates on every site, the cell fields, gMET and the
    cycle boundaries on sites occupied in both, the four mode lists as
    sets.  Returns the list of differences (empty: the round trip holds)."""
    with open(json_path) as fh:
        a = json.load(fh)
    with tempfile.TemporaryDirectory(dir=tmpdir) as d:
        txt, back = os.path.join(d, "rt.txt"), os.path.join(d, "rt.json")
        json_to_txt(json_path, txt)
        txt_to_json(txt, back, json_path)
        with open(back) as fh:
            b = json.load(fh)
    L = int(a["PARAMS"]["L"])
    both = (np.asarray(a["STATE"]["W"]).reshape(L, L) == 1) & \
        (np.asarray(b["STATE"]["W"]).reshape(L, L) == 1)
    out = []

    def cmp(name, ga, gb, mask=None):
        x = np.asarray(ga, np.float64).reshape(L, L)
        y = np.asarray(gb, np.float64).reshape(L, L)
        d = np.abs(x - y) if mask is None else np.abs(x - y)[mask]
        if d.size and d.max() > tol:
            out.append(f"{name}: max|diff| {d.max():.3g} > {tol:.3g}")

    for k in ("W",) + SUBSTRATES:
        cmp(k, a["STATE"][k], b["STATE"][k])
    for k in CELL_FIELDS + ("gMET",):
        cmp(k, a["STATE"][k], b["STATE"][k], both)
    for name in LMET_MODES:
        sa = set(map(int, a["STATE"].get("lMET", {}).get(name, [])))
        sb = set(map(int, b["STATE"]["lMET"].get(name, [])))
        if sa != sb:
            out.append(f"lMET['{name}']: {len(sa - sb)} only before, {len(sb - sa)} only after")
    if "CYCLES" in a:
        for k in CYCLE_FIELDS:
            if a["CYCLES"].get(k) is not None:
                cmp("CYCLES." + k, a["CYCLES"][k], b["CYCLES"][k], both)
    return out


def main(argv=None):
    p = argparse.ArgumentParser(description="tumour states: tumor-ca JSON <-> EMT6-Ro text")
    sub = p.add_subparsers(dest="cmd", required=True)
    a = sub.add_parser("json-to-txt")
    a.add_argument("input_json")
    a.add_argument("output_txt")
    a.add_argument("--scale", type=float, default=1.0)
    b = sub.add_parser("txt-to-json")
    b.add_argument("input_txt")
    b.add_argument("output_json")
    b.add_argument("--template", required=True)
    b.add_argument("--scale", type=float, default=1.0)
    c = sub.add_parser("roundtrip-test")
    c.add_argument("input_json")
    c.add_argument("--tmpdir", default=None)
    args = p.parse_args(argv)
    if args.cmd == "json-to-txt":
        info = json_to_txt(args.input_json, args.output_txt, args.scale)
        print(f"wrote {args.output_txt}: L={info['L']} occupied={info['n_occupied']} "
              f"vacant={info['n_vacant']} dead={info['n_dead_skipped']} scale={info['scale']}")
        return 0
    if args.cmd == "txt-to-json":
        info = txt_to_json(args.input_txt, args.output_json, args.template, args.scale)
        print(f"wrote {args.output_json}: L={info['L']} occupied={info['n_occupied']} "
              f"scale={info['scale']}")
        return 0
    fails = roundtrip_failures(args.input_json, args.tmpdir)
    for f in fails:
        print("FAIL", f)
    print("PASS" if not fails else f"{len(fails)} fields differ")
    return 1 if fails else 0


if __name__ == "__main__":
    sys.exit(main())
