"""Dose-protocol files: the two-line text format shared with tumor-ca, and the
step-indexing rules of the reference's converter (tools/convert_protocols.py:
38-197).

A file holds two lines per protocol: the doses (Gy), then the simulation
steps, space-separated.  EMT6-Ro fires a dose at step s when s < length and
s % resolution == 0 (common/protocol.h:9-14); tumor-ca files number steps
from 1 (k * 300 + 1).  `canonical` turns a protocol whose steps are all
1-indexed into the 0-indexed form (tools/convert_protocols.py:148-175).

  python -m emt6ro.protocols validate FILE [--steps N] [--resolution R]
  python -m emt6ro.protocols to-emt6 FILE [--out OUT.json]
  python -m emt6ro.protocols to-two-line IN.json OUT
  python -m emt6ro.protocols diff A B
"""
import argparse
import json
import sys


def parse_two_line(path):
    """Protocols of a two-line file as lists of (step, dose) pairs; a pair of
    empty lines is skipped, a trailing odd line is ignored."""
    with open(path) as f:
        lines = f.read().splitlines()
    out = []
    for i in range(0, len(lines) - 1, 2):
        d, t = lines[i].split(), lines[i + 1].split()
        if not d and not t:
            continue
        if len(d) != len(t):
            raise ValueError(f"{path}: protocol at line {i + 1}: {len(d)} doses but "
                             f"{len(t)} steps")
        out.append([(int(s), float(x)) for s, x in zip(t, d)])
    return out


def format_two_line(protocols):
    rows = []
    for prot in protocols:
        rows.append(" ".join(str(float(d)) for _, d in prot))
        rows.append(" ".join(str(int(s)) for s, _ in prot))
    return "".join(r + "\n" for r in rows)


def write_two_line(protocols, path):
    with open(path, "w") as f:
        f.write(format_two_line(protocols))


def validate(protocols, n_steps=144000, resolution=300):
    """Counts of events in range, aligned on the 0- or 1-indexed grid,
    unaligned, and repeated steps within a protocol, with the indexing the
    file follows ("0-indexed", "1-indexed", "mixed-0-and-1", "unaligned" or
    "empty")."""
    st = {"n_protocols": len(protocols), "n_events": 0, "n_in_range": 0,
          "n_out_of_range": 0, "n_aligned_0": 0, "n_aligned_1": 0, "n_unaligned": 0,
          "duplicates_per_protocol": 0, "examples_unaligned": []}
    for prot in protocols:
        seen = set()
        for step, _ in prot:
            st["n_events"] += 1
            st["n_in_range" if 0 <= step < n_steps else "n_out_of_range"] += 1
            r = step % resolution
            if r == 0:
                st["n_aligned_0"] += 1
            elif r == 1:
                st["n_aligned_1"] += 1
            else:
                st["n_unaligned"] += 1
                if len(st["examples_unaligned"]) < 5:
                    st["examples_unaligned"].append(step)
            st["duplicates_per_protocol"] += step in seen
            seen.add(step)
    a0, a1, other = st["n_aligned_0"], st["n_aligned_1"], st["n_unaligned"]
    if st["n_events"] == 0:
        st["indexing"] = "empty"
    elif other:
        st["indexing"] = "unaligned"
    elif a0 and not a1:
        st["indexing"] = "0-indexed"
    elif a1 and not a0:
        st["indexing"] = "1-indexed"
    else:
        st["indexing"] = "mixed-0-and-1"
    return st


def canonical(protocols, resolution=300):
    """0-indexed steps: a protocol whose every step is 1 past the grid moves
    back by one step; others are kept.  Each protocol is sorted by step."""
    out = []
    for prot in protocols:
        shift = 1 if prot and all(s % resolution == 1 for s, _ in prot) else 0
        out.append(sorted(((int(s) - shift, float(d)) for s, d in prot), key=lambda e: e[0]))
    return out


def diff(a, b):
    """Number of differing protocols / events, with a line for each."""
    n, lines = 0, []
    if len(a) != len(b):
        lines.append(f"protocol counts differ: {len(a)} vs {len(b)}")
        n += 1
    for i, (pa, pb) in enumerate(zip(a, b)):
        if len(pa) != len(pb):
            lines.append(f"protocol {i}: {len(pa)} vs {len(pb)} events")
            n += 1
            continue
        for j, (ea, eb) in enumerate(zip(pa, pb)):
            if tuple(ea) != tuple(eb):
                lines.append(f"protocol {i} event {j}: {tuple(ea)} vs {tuple(eb)}")
                n += 1
    return n, lines


def main(argv=None):
    ap = argparse.ArgumentParser(description="Dose-protocol files (two-line format).")
    sub = ap.add_subparsers(dest="cmd", required=True)
    v = sub.add_parser("validate")
    v.add_argument("file")
    v.add_argument("--steps", type=int, default=144000)
    v.add_argument("--resolution", type=int, default=300)
    e = sub.add_parser("to-emt6")
    e.add_argument("file")
    e.add_argument("--out")
    t = sub.add_parser("to-two-line")
    t.add_argument("json_in")
    t.add_argument("out")
    d = sub.add_parser("diff")
    d.add_argument("a")
    d.add_argument("b")
    a = ap.parse_args(argv)
    if a.cmd == "validate":
        print(json.dumps(validate(parse_two_line(a.file), a.steps, a.resolution)))
    elif a.cmd == "to-emt6":
        txt = json.dumps(canonical(parse_two_line(a.file)))
        if a.out:
            with open(a.out, "w") as f:
                f.write(txt)
        else:
            print(txt)
    elif a.cmd == "to-two-line":
        with open(a.json_in) as f:
            write_two_line([[tuple(e) for e in p] for p in json.load(f)], a.out)
    else:
        n, lines = diff(canonical(parse_two_line(a.a)), canonical(parse_two_line(a.b)))
        for line in lines:
            print(line)
        print(f"total diffs: {n}")
        return 1 if n else 0
    return 0


if __name__ == "__main__":
    sys.exit(main())
