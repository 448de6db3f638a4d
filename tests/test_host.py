"""Host-side library checks that need no GPU: the C ABI loads and exports every
declared symbol, parameters and tumour states load like the reference's
loaders, and errors carry the reference's messages."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
from conftest import ROOT, golden, gpu_available


def test_library_exports_every_declared_symbol():
    from emt6ro import _native
    header = open(os.path.join(ROOT, "include", "emt6ro_amd.h")).read()
    names = re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(emt6ro_[a-z_0-9]+)\(", header, re.M)
    assert len(names) >= 30
    lib = C.CDLL(_native.LIB_PATH)
    for n in names:
        assert hasattr(lib, n), n
    assert _native.lib.emt6ro_version().startswith(b"emt6ro_amd")


def test_dummy_parameters_all_ones():
    """parameters-test.cc:16-40."""
    import emt6ro.simulation as sim
    p = sim.load_parameters(golden("dummy-parameters.json"))
    assert np.all(p.values() == 1.0)


def test_default_parameters_match_independent_reader():
    import emt6ro.simulation as sim
    p = sim.load_parameters(golden("default-parameters.json"))
    assert np.array_equal(p.values(), O.read_params_json(golden("default-parameters.json")))
    assert p.lattice_dims == (51, 51)


def test_tumor_library_matches_independent_reader():
    import emt6ro.simulation as sim
    p = sim.load_parameters(golden("default-parameters.json"))
    v = p.values()
    for t in range(1, 11):
        path = golden("tumor-lib", f"tumor-{t}.txt")
        s = sim.load_state(path, p)
        o = O.read_state_text(path, v)
        assert s.shape == (53, 53)
        assert np.array_equal(s.occupancy(), o["st"])
        for k, n in enumerate(("cho", "ox", "gi")):
            assert np.array_equal(s.field(n), o["sub"][k])
        occ = o["st"] == 1
        cells = o["cells"].reshape(53, 53)
        assert np.array_equal(s.mode()[occ], cells["mode"][occ])
        assert np.array_equal(s.phase()[occ], cells["phase"][occ])
        assert np.array_equal(s.field("proliferation_time")[occ], cells["prolif"][occ])


def test_tumor_1_known_counts():
    """tumor-1: 321 occupied; modes [87,60,100,74]; phases [241,45,24,11,0]."""
    import emt6ro.simulation as sim
    p = sim.load_parameters(golden("default-parameters.json"))
    s = sim.load_state(golden("tumor-lib", "tumor-1.txt"), p)
    occ = s.occupancy() == 1
    assert occ.sum() == 321
    assert list(np.bincount(s.mode()[occ], minlength=4)) == [87, 60, 100, 74]
    assert list(np.bincount(s.phase()[occ], minlength=5)) == [241, 45, 24, 11, 0]
    assert (s.occupancy()[0] == 2).all() and (s.occupancy()[:, -1] == 2).all()


def test_state_save_load_round_trip(tmp_path):
    import emt6ro.simulation as sim
    p = sim.load_parameters(golden("default-parameters.json"))
    s = sim.load_state(golden("tumor-lib", "tumor-4.txt"), p)
    out = tmp_path / "t.txt"
    s.save(out)
    s2 = sim.load_state(out, p)
    assert np.array_equal(s.occupancy(), s2.occupancy())
    # saveToFile writes 6 significant digits (default ostream precision).
    for n in ("cho", "ox", "gi", "proliferation_time", "cycle_g1"):
        assert np.allclose(s.field(n), s2.field(n), rtol=1e-5, atol=0)


def test_missing_file_message():
    import emt6ro.simulation as sim
    p = sim.load_parameters(golden("default-parameters.json"))
    with pytest.raises(RuntimeError, match=r"Check failed: file.good\(\); Could not open the file: /nope"):
        sim.load_state("/nope", p)


def test_synthetic_disk_has_gauss_count():
    from emt6ro.synthetic import disk_arrays
    import emt6ro.simulation as sim
    p = sim.load_parameters(golden("default-parameters.json"))
    a = disk_arrays(p.values(), 53, 53, 10, seed=0)
    assert int((a["occupancy"] == 1).sum()) == 317
    ph = a["phase"][a["occupancy"] == 1]
    assert set(np.unique(ph)) <= {0, 1, 2, 3}
    t = a["proliferation_time"][a["occupancy"] == 1]
    assert (t >= 0).all()


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU failure path")
def test_experiment_without_gpu_fails_loudly():
    import emt6ro.simulation as sim
    p = sim.load_parameters(golden("default-parameters.json"))
    s = sim.load_state(golden("tumor-lib", "tumor-1.txt"), p)
    with pytest.raises(RuntimeError, match="HIP error"):
        sim.Experiment(p, [s], 1, 1, seed=0)
