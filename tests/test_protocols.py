"""Protocol files and the replay harness's host side (emt6ro.protocols,
emt6ro.replay): the two-line format and the indexing rules of the
reference's converter (tools/convert_protocols.py:38-197)."""
import json

import numpy as np
import pytest

from emt6ro import protocols as P
from emt6ro.replay import dose_rows


def test_two_line_round_trip(tmp_path):
    prots = [[(0, 1.25), (7200, 2.5)], [(300, 0.5)], []]
    f = tmp_path / "p.txt"
    P.write_two_line(prots, f)
    assert f.read_text() == "1.25 2.5\n0 7200\n0.5\n300\n\n\n"
    back = P.parse_two_line(f)
    # an empty protocol is a pair of empty lines, which the reader skips
    assert back == prots[:2]


def test_parse_rejects_mismatched_lines(tmp_path):
    f = tmp_path / "bad.txt"
    f.write_text("1.0 2.0\n300\n")
    with pytest.raises(ValueError, match="2 doses but 1 steps"):
        P.parse_two_line(f)


def test_validate_and_canonical_indexing():
    one = [[(1, 1.0), (8701, 2.0)], [(301, 0.5)]]
    st = P.validate(one)
    assert st["indexing"] == "1-indexed" and st["n_aligned_1"] == 3 and st["n_events"] == 3
    canon = P.canonical(one)
    assert canon == [[(0, 1.0), (8700, 2.0)], [(300, 0.5)]]
    assert P.validate(canon)["indexing"] == "0-indexed"
    mixed = [[(0, 1.0)], [(301, 1.0)], [(150, 1.0), (150, 2.0)], [(144000, 1.0)]]
    st = P.validate(mixed)
    assert st["indexing"] == "unaligned" and st["examples_unaligned"] == [150, 150]
    assert st["duplicates_per_protocol"] == 1 and st["n_out_of_range"] == 1
    # a protocol with a mix of grid offsets is kept as it is (sorted)
    assert P.canonical([[(601, 1.0), (300, 2.0)]]) == [[(300, 2.0), (601, 1.0)]]
    assert P.validate([])["indexing"] == "empty"


def test_diff_counts_events():
    a = [[(0, 1.0), (300, 2.0)], [(600, 1.0)]]
    b = [[(0, 1.0), (300, 2.5)], [(600, 1.0), (900, 1.0)]]
    n, lines = P.diff(a, b)
    assert n == 2 and len(lines) == 2


def test_cli_to_emt6(tmp_path, capsys):
    f = tmp_path / "p.txt"
    f.write_text("1.25 1.25\n7201 21601\n")
    assert P.main(["to-emt6", str(f)]) == 0
    assert json.loads(capsys.readouterr().out) == [[[7200, 1.25], [21600, 1.25]]]


def test_dose_rows_accumulate_and_check_range():
    d = dose_rows([[(0, 1.0), (299, 0.5), (600, 2.0)], []], steps=1200, resolution=300)
    np.testing.assert_array_equal(d, [[1.5, 0, 2.0, 0], [0, 0, 0, 0]])
    with pytest.raises(ValueError, match="outside"):
        dose_rows([[(1200, 1.0)]], steps=1200, resolution=300)
