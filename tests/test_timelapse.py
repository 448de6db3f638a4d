"""Per-step readout (emt6ro.timelapse; the reference's example/timelapse.cc:
23-90): the 13 numbers of a hand-built state, and on the GPU the readout of
a batch, step by step, against the oracle's states."""
import numpy as np
import pytest

from conftest import golden

tl = pytest.importorskip("emt6ro.timelapse")


def test_record_of_a_hand_built_state(default_params_path):
    import emt6ro.simulation as sim
    params = sim.load_parameters(default_params_path)
    st = sim.TumorState.empty(params, 7, 7)
    occ = np.zeros((7, 7), np.uint8)
    mode = np.zeros((7, 7), np.uint8)
    phase = np.zeros((7, 7), np.uint8)
    # Interior cells: (1,1) mode 0 G1, (2,3) mode 1 S, (3,3) mode 3 D, (5,5) mode 2 M;
    # (0,3) and (6,6) lie on the border and are not counted.
    for (r, c, m, p) in [(1, 1, 0, 0), (2, 3, 1, 1), (3, 3, 3, 4), (5, 5, 2, 3),
                         (0, 3, 1, 2), (6, 6, 0, 2)]:
        occ[r, c], mode[r, c], phase[r, c] = 1, m, p
    st.set_field("occupancy", occ)
    st.set_field("mode", mode)
    st.set_field("phase", phase)
    cho = np.arange(49, dtype=np.float32).reshape(7, 7)
    st.set_field("cho", cho)
    st.set_field("ox", cho * 2)
    st.set_field("gi", cho / 4)
    got = tl.record(st)
    assert list(got[:3]) == [24.0, 48.0, 6.0]         # centre (3, 3)
    assert got[3] == 4
    assert list(got[4:8]) == [1, 1, 1, 1]
    assert list(got[8:13]) == [1, 1, 0, 1, 1]
    assert list(tl.record(st, probe=(1, 2))[:3]) == [9.0, 18.0, 2.25]
    assert len(tl.COLUMNS) == 13


class _OracleView:
    """An oracle simulation's state with the TumorState accessors."""

    def __init__(self, orc, b):
        H, W = orc.H, orc.W
        self._occ = orc.st[b].reshape(H, W)
        self._sub = orc.sub[b].reshape(3, H, W)
        cells = orc.cells[b].reshape(H, W)
        self._mode, self._phase = cells["mode"], cells["phase"]

    def occupancy(self):
        return self._occ

    def cho(self):
        return self._sub[0]

    def ox(self):
        return self._sub[1]

    def gi(self):
        return self._sub[2]

    def mode(self):
        return self._mode

    def phase(self):
        return self._phase


@pytest.mark.gpu
def test_batch_timelapse_matches_oracle():
    import oracle as O
    import emt6ro.simulation as sim
    from emt6ro.batch import Batch
    params = sim.load_parameters(golden("default-parameters.json"))
    v = params.values()
    path = golden("tumor-lib", "tumor-4.txt")
    tumor = sim.load_state(path, params)
    doses = np.zeros((1, 480), np.float32)
    doses[0, 0] = 2.5
    n, steps, every, seed = 2, 60, 6, 31
    b = Batch(params, [tumor], [0] * n, [0] * n, [144000] * n, doses, 300, 144000,
              sim_id=np.arange(n), seed=seed)
    print(f"[timelapse] {n} sims, {steps} steps, every {every}", flush=True)
    got = tl.run(b, steps, every)
    assert got.shape == (n, steps // every + 1, 13)
    orc = O.OracleBatch(v, [O.read_state_text(path, v)] * n, doses, [0] * n, np.arange(n), 300,
                        144000)
    want = [[tl.record(_OracleView(orc, i)) for i in range(n)]]
    for _ in range(steps // every):
        orc.run(every, seed=seed)
        want.append([tl.record(_OracleView(orc, i)) for i in range(n)])
    want = np.transpose(np.asarray(want), (1, 0, 2))
    assert np.array_equal(got, want)
    assert got[0, -1, 3] > 0 and not np.array_equal(got[0, 0], got[0, -1])


@pytest.mark.gpu
def test_timelapse_command_line(tmp_path):
    out = tmp_path / "tl.csv"
    rc = tl.main(["--tumor", golden("tumor-lib", "tumor-4.txt"),
                  "--params", golden("default-parameters.json"), "--sims", "2", "--steps", "12",
                  "--every", "6", "--dose", "0:2.5", "--out", str(out)])
    assert rc == 0
    lines = out.read_text().splitlines()
    assert lines[0] == "sim,record," + ",".join(tl.COLUMNS)
    assert len(lines) == 1 + 2 * 3
