"""GPU parity: the HIP kernels (through the C ABI) against the CPU oracle.

Integer bookkeeping (site states, list lengths, phases, modes) and every
float (substrates, cell clocks) must match the oracle bit for bit: both follow
the same operation order with the same counter-based random numbers.
"""
import sys

import numpy as np
import pytest

import oracle as O
from conftest import golden

pytestmark = pytest.mark.gpu

FLOAT_FIELDS = ("time_in_repair", "irradiation", "repair_delay", "proliferation_time",
                "cycle_g1", "cycle_s", "cycle_g2", "cycle_m", "cycle_d")
ORACLE_CELL = {"time_in_repair": ("t_rep", None), "irradiation": ("R", None),
               "repair_delay": ("delay", None), "proliferation_time": ("prolif", None),
               "cycle_g1": ("times", 0), "cycle_s": ("times", 1), "cycle_g2": ("times", 2),
               "cycle_m": ("times", 3), "cycle_d": ("times", 4)}


def log(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def sim():
    import emt6ro.simulation as sim
    return sim


@pytest.fixture(scope="module")
def params(sim):
    return sim.load_parameters(golden("default-parameters.json"))


def oracle_state_from_arrays(arrays, values):
    H, W = arrays["occupancy"].shape
    s = O.empty_state(values, H, W)
    s["sub"][0], s["sub"][1], s["sub"][2] = arrays["cho"], arrays["ox"], arrays["gi"]
    s["st"][:] = arrays["occupancy"]
    cells = s["cells"].reshape(H, W)
    occ = arrays["occupancy"] == 1
    for name, (f, k) in ORACLE_CELL.items():
        if k is None:
            cells[f][occ] = arrays[name][occ]
        else:
            cells[f][occ, k] = arrays[name][occ]
    cells["mode"][occ] = arrays["mode"][occ]
    cells["phase"][occ] = arrays["phase"][occ]
    return s


def inputs(sim, params, n_synth=2):
    """Library tumours 1-3 and synthetic disks, each as (product state, oracle state)."""
    from emt6ro.synthetic import disk_state
    v = params.values()
    out = []
    for t in (1, 2, 3):
        path = golden("tumor-lib", f"tumor-{t}.txt")
        out.append((sim.load_state(path, params), O.read_state_text(path, v)))
    for k in range(n_synth):
        st, arrays = disk_state(params, 53, 53, radius=8 + 3 * k, seed=100 + k)
        out.append((st, oracle_state_from_arrays(arrays, v)))
    return out


def compare_states(gpu_state, orc, b, where):
    H, W = gpu_state.shape
    occ_g = gpu_state.occupancy()
    occ_o = orc.st[b]
    assert np.array_equal(occ_g, occ_o), f"{where}: site states differ"
    for k, name in enumerate(("cho", "ox", "gi")):
        g = gpu_state.field(name)
        o = orc.sub[b, k]
        if not np.array_equal(g.view(np.uint32), o.view(np.uint32)):
            bad = np.argwhere(g.view(np.uint32) != o.view(np.uint32))
            r, c = bad[0]
            raise AssertionError(f"{where}: {name} differs at {len(bad)} sites, first "
                                 f"({r},{c}) gpu={g[r, c]!r} oracle={o[r, c]!r}")
    occ = occ_o == 1
    cells = orc.cells[b].reshape(H, W)
    for name, (f, k) in ORACLE_CELL.items():
        g = gpu_state.field(name)[occ]
        o = cells[f][occ] if k is None else cells[f][occ, k]
        assert np.array_equal(g.view(np.uint32), o.astype(np.float32).view(np.uint32)), \
            f"{where}: {name} differs"
    assert np.array_equal(gpu_state.mode()[occ], cells["mode"][occ]), f"{where}: mode differs"
    assert np.array_equal(gpu_state.phase()[occ], cells["phase"][occ]), f"{where}: phase differs"


def build_pair(sim, params, doses, reps, seed):
    from emt6ro.batch import Batch
    ins = inputs(sim, params)
    T = len(ins)
    n = T * reps
    tumor_index = np.repeat(np.arange(T), reps)
    prot_index = np.arange(n) % doses.shape[0]
    horizon = np.full(n, 2 ** 32 - 1, np.uint32)
    b = Batch(params, [s for s, _ in ins], tumor_index, prot_index, horizon, doses, 300, 144000,
              sim_id=np.arange(n), seed=seed)
    orc = O.OracleBatch(params.values(), [ins[t][1] for t in tumor_index], doses, prot_index,
                        np.arange(n), 300, 144000)
    return b, orc


def dose_table():
    d = np.zeros((3, 480), np.float32)
    d[0, 0] = 2.0          # dose at step 0
    d[1, 1] = 1.25         # step 300
    d[1, 3] = 3.0          # step 900
    return d


def test_graph_follows_epoch_change(sim, params):
    """The captured per-step graph holds the epoch among its launch arguments:
    after a reset to another epoch the graph path must match the persistent
    path at that epoch, not replay the old one."""
    b, _ = build_pair(sim, params, dose_table(), reps=2, seed=7)
    log("[parity] graph epoch change: capture at epoch 0, run at epoch 1")
    b.run(512, path="graph")
    ep0 = [np.array(b.state(i).cho()) for i in range(b.n)]
    b.reset(1)
    b.run(512, path="graph")
    got = b.results().copy()
    got_cho = [np.array(b.state(i).cho()) for i in range(b.n)]
    b.reset(1)
    b.run(512, path="persistent")
    assert np.array_equal(b.results(), got)
    for i in range(b.n):
        assert np.array_equal(np.array(b.state(i).cho()), got_cho[i]), i
    assert any(not np.array_equal(a, c) for a, c in zip(ep0, got_cho))


@pytest.mark.parametrize("path", ["persistent", "step", "graph"])
def test_paths_match_oracle_bitexact(sim, params, path):
    nsteps = {"persistent": 3000, "step": 400, "graph": 300}[path]
    log(f"[parity] path={path} nsteps={nsteps}")
    b, orc = build_pair(sim, params, dose_table(), reps=2, seed=12345)
    b.run(nsteps, path=path)
    orc.run(nsteps, seed=12345, nthreads=8)
    got = b.results().astype(np.int64)
    assert np.array_equal(got, orc.counts()), (got, orc.counts())
    cnt = b.counters()
    assert (cnt["step"] == nsteps).all()
    assert (cnt["status"] == 0).all()
    assert np.array_equal(cnt["n_list"].astype(np.int64), orc.nlist.astype(np.int64))
    for i in range(b.n):
        compare_states(b.state(i), orc, i, f"{path} sim {i}")


def test_long_run_with_repair_and_division_matches_oracle(sim, params):
    """Past the repair delay (> 3.3 h) so repair deaths and divisions occur."""
    log("[parity] long run 6000 steps")
    b, orc = build_pair(sim, params, dose_table(), reps=2, seed=777)
    b.run(6000)
    orc.run(6000, seed=777, nthreads=8)
    assert np.array_equal(b.results().astype(np.int64), orc.counts())
    cnt = b.counters()
    assert cnt["divisions"].sum() > 0
    for i in range(b.n):
        compare_states(b.state(i), orc, i, f"long sim {i}")


def test_split_runs_and_launch_chunks_equal_one_run(sim, params):
    log("[parity] split runs")
    d = dose_table()
    b1, _ = build_pair(sim, params, d, reps=1, seed=5)
    b2, _ = build_pair(sim, params, d, reps=1, seed=5)
    b3, _ = build_pair(sim, params, d, reps=1, seed=5)
    b1.run(1000)
    b2.run(313)
    b2.run(687, path="step")
    b3.run(1000, steps_per_launch=97)
    r1 = b1.results()
    assert np.array_equal(r1, b2.results())
    assert np.array_equal(r1, b3.results())
    for i in range(b1.n):
        s1, s2, s3 = b1.state(i), b2.state(i), b3.state(i)
        for name in ("cho", "ox", "gi", "occupancy", "proliferation_time"):
            assert np.array_equal(s1.field(name), s2.field(name))
            assert np.array_equal(s1.field(name), s3.field(name))


def test_every_variant_matches(sim, params):
    from emt6ro.batch import variants, fits
    log("[parity] variants")
    d = dose_table()
    ref = None
    for v in variants():
        if not fits(v, 53, 53):
            continue
        log(f"[parity] variant {v}")
        b, _ = build_pair(sim, params, d, reps=1, seed=3)
        b.run(500, variant=v["id"])
        st = [b.state(i) for i in range(b.n)]
        sig = [np.concatenate([s.field(n).view(np.uint32).ravel() if s.field(n).dtype != np.uint8
                               else s.field(n).astype(np.uint32).ravel()
                               for n in ("cho", "ox", "gi", "occupancy")]) for s in st]
        if ref is None:
            ref = sig
        else:
            for a, bb in zip(ref, sig):
                assert np.array_equal(a, bb), f"variant {v} differs"


def test_every_variant_matches_oracle_on_33x33(sim, params):
    """Every variant that fits 33x33 lattices (the stride-36 ones, among them
    the 128- and 64-thread lane-run kernels, and the wider ones) against the
    oracle bit for bit."""
    from emt6ro.batch import Batch, variants, fits
    from emt6ro.synthetic import disk_state
    v = params.values()
    tumors, orc_states = [], []
    for k, r in enumerate((5, 7, 9, 11)):
        st, arr = disk_state(params, 33, 33, radius=r, seed=300 + k)
        tumors.append(st)
        orc_states.append(oracle_state_from_arrays(arr, v))
    d = dose_table()
    n = len(tumors)
    steps = 500
    orc = O.OracleBatch(v, orc_states, d, np.arange(n) % d.shape[0], np.arange(n), 300, 144000)
    orc.run(steps, seed=11, nthreads=4)
    tested = 0
    for var in variants():
        if not fits(var, 33, 33):
            continue
        log(f"[parity] 33x33 variant {var}")
        b = Batch(params, tumors, np.arange(n), np.arange(n) % d.shape[0],
                  np.full(n, 2 ** 32 - 1, np.uint32), d, 300, 144000, sim_id=np.arange(n),
                  seed=11)
        b.run(steps, variant=var["id"])
        assert np.array_equal(b.results().astype(np.int64), orc.counts()), var
        for i in range(n):
            compare_states(b.state(i), orc, i, f"33x33 variant {var['id']} sim {i}")
        tested += 1
    assert tested >= 10


@pytest.mark.parametrize("rows", [-1, 24, 21])
def test_window_pass_matches_oracle(sim, params, rows):
    """The window pass (only the ROI's temporary grid of a 53x53 lattice in
    LDS) with the automatic window and with windows so small that the ROIs
    of the tumours leave them or outgrow them (the window moves, or the
    simulation finishes its launch in the whole-lattice kernel), over
    several launches."""
    log(f"[parity] window rows={rows}")
    b, orc = build_pair(sim, params, dose_table(), reps=2, seed=4242)
    b.set_window(rows)
    b.run(3000, steps_per_launch=700)
    orc.run(3000, seed=4242, nthreads=8)
    assert np.array_equal(b.results().astype(np.int64), orc.counts())
    cnt = b.counters()
    assert (cnt["step"] == 3000).all() and (cnt["status"] == 0).all()
    assert np.array_equal(cnt["n_list"].astype(np.int64), orc.nlist.astype(np.int64))
    for i in range(b.n):
        compare_states(b.state(i), orc, i, f"window rows {rows} sim {i}")


def test_wide_rois_beside_window_matches_oracle(sim, params):
    """Tumours whose ROI is too wide for the window (or outgrows it during a
    launch) next to small ones, on 53x53 lattices, over several launches."""
    from emt6ro.batch import Batch
    from emt6ro.synthetic import disk_state
    v = params.values()
    tumors, orc_states = [], []
    for k, r in enumerate((6, 15, 16, 17)):
        st, arr = disk_state(params, 53, 53, radius=r, seed=500 + k)
        tumors.append(st)
        orc_states.append(oracle_state_from_arrays(arr, v))
    d = dose_table()
    n = 2 * len(tumors)
    ti = np.repeat(np.arange(len(tumors)), 2)
    pi = np.arange(n) % d.shape[0]
    b = Batch(params, tumors, ti, pi, np.full(n, 2 ** 32 - 1, np.uint32), d, 300, 144000,
              sim_id=np.arange(n), seed=9)
    orc = O.OracleBatch(v, [orc_states[t] for t in ti], d, pi, np.arange(n), 300, 144000)
    log("[parity] wide ROIs beside the window: disks of radius 6, 15, 16, 17")
    b.run(2400, steps_per_launch=600)
    orc.run(2400, seed=9, nthreads=8)
    assert np.array_equal(b.results().astype(np.int64), orc.counts())
    for i in range(n):
        compare_states(b.state(i), orc, i, f"wider window sim {i}")
        log(f"  sim {i}: ROI {orc.roi[i].tolist()}")


def test_reset_restores_initial_state(sim, params):
    b, _ = build_pair(sim, params, dose_table(), reps=1, seed=1)
    r0 = b.results()
    b.run(700)
    b.reset(epoch=0)
    assert np.array_equal(b.results(), r0)
    b.run(700)
    first = b.results()
    b.reset(epoch=0)
    b.run(700)
    assert np.array_equal(b.results(), first)
    with pytest.raises(ValueError, match="doses must be"):
        b.set_doses(dose_table()[:, :-1])


@pytest.mark.parametrize("path", ["persistent", "step"])
def test_empty_and_tiny_lattices_match_oracle(sim, params, path):
    """A lattice without cells (no ROI: diffusion skipped, Q8) and a
    two-site-radius colony, next to each other in one batch."""
    from emt6ro.batch import Batch
    from emt6ro.synthetic import disk_state
    v = params.values()
    empty = sim.TumorState.empty(params, 53, 53)
    tiny, arr = disk_state(params, 53, 53, radius=2, seed=9)
    doses = np.zeros((1, 480), np.float32)
    doses[0, 1] = 1.0
    n, steps, seed = 4, 600, 77
    tumor_index = np.array([0, 1, 0, 1])
    b = Batch(params, [empty, tiny], tumor_index, np.zeros(n, np.int64),
              np.full(n, 2 ** 32 - 1, np.uint32), doses, 300, 144000, sim_id=np.arange(n),
              seed=seed)
    log(f"[parity] empty and tiny lattices, path={path}")
    b.run(steps, path=path)
    states = [O.empty_state(v, 53, 53), oracle_state_from_arrays(arr, v)]
    orc = O.OracleBatch(v, [states[t] for t in tumor_index], doses, np.zeros(n, np.int64),
                        np.arange(n), 300, 144000)
    orc.run(steps, seed=seed, nthreads=4)
    got = b.results().astype(np.int64)
    assert np.array_equal(got, orc.counts()), (got, orc.counts())
    assert got[0] == 0 and got[2] == 0 and got[1] > 0
    for i in range(n):
        compare_states(b.state(i), orc, i, f"{path} sim {i}")


@pytest.mark.parametrize("path", ["persistent", "graph"])
def test_colonies_dying_out_mid_run_match_oracle(sim, params, path):
    """10 Gy at step 0 on a radius-3 colony (31 cells): the repair delay
    (3.3414 exp(0.1492 R), about 15 h) ends in death with probability
    1 - exp(-0.4993 R) = 0.993, so some replicates die out between hours 32
    and 34 and the others keep a few cells.  The ROI of a dying colony shrinks
    to nothing mid-run (diffusion then skipped); every state matches the
    oracle bit for bit before (24 h) and after (48 h)."""
    from emt6ro.batch import Batch
    from emt6ro.synthetic import disk_state
    v = params.values()
    colony, arr = disk_state(params, 53, 53, radius=3, seed=9)
    doses = np.zeros((1, 480), np.float32)
    doses[0, 0] = 10.0
    n, seed = 8, 5
    b = Batch(params, [colony], np.zeros(n, np.int64), np.zeros(n, np.int64),
              np.full(n, 2 ** 32 - 1, np.uint32), doses, 300, 144000, sim_id=np.arange(n),
              seed=seed)
    orc = O.OracleBatch(v, [oracle_state_from_arrays(arr, v)] * n, doses, np.zeros(n, np.int64),
                        np.arange(n), 300, 144000)
    for hours in (24, 24):
        log(f"[parity] colonies dying out, path={path}, {hours} more hours")
        b.run(hours * 600, path=path)
        orc.run(hours * 600, seed=seed, nthreads=4)
        got = b.results().astype(np.int64)
        assert np.array_equal(got, orc.counts()), (got, orc.counts())
        for i in range(n):
            compare_states(b.state(i), orc, i, f"{path} sim {i}")
    assert (got == 0).any() and (got > 0).any(), got


@pytest.mark.parametrize("path", ["persistent", "graph"])
def test_colonies_at_the_lattice_edge_match_oracle(sim, params, path):
    """Colonies placed in a corner, against an edge, at the bottom and in the
    opposite corner of a 53x53 lattice, beside a centred one: their ROIs are
    clipped at the first or last interior row or column, so the temporary
    grid's rings (and, on the persistent path, the LDS window placed around
    the ROI) reach past the lattice.  1.5 Gy at step 300; every state matches
    the oracle bit for bit at 1 h and 12 h."""
    from emt6ro.batch import Batch
    from emt6ro.synthetic import disk_state
    v = params.values()
    spec = [((4, 4), 5), ((26, 2), 4), ((49, 30), 6), ((2, 50), 3), ((26, 26), 6)]
    tumors, states = [], []
    for k, (center, radius) in enumerate(spec):
        st, arr = disk_state(params, 53, 53, radius=radius, seed=30 + k, center=center)
        tumors.append(st)
        states.append(oracle_state_from_arrays(arr, v))
    doses = np.zeros((1, 480), np.float32)
    doses[0, 1] = 1.5
    reps, seed = 2, 11
    tumor_index = np.repeat(np.arange(len(spec)), reps)
    n = len(tumor_index)
    b = Batch(params, tumors, tumor_index, np.zeros(n, np.int64),
              np.full(n, 2 ** 32 - 1, np.uint32), doses, 300, 144000, sim_id=np.arange(n),
              seed=seed)
    orc = O.OracleBatch(v, [states[t] for t in tumor_index], doses, np.zeros(n, np.int64),
                        np.arange(n), 300, 144000)
    clipped = 0
    for hours in (1, 11):
        log(f"[parity] colonies at the lattice edge, path={path}, {hours} more hours")
        b.run(hours * 600, path=path)
        orc.run(hours * 600, seed=seed, nthreads=4)
        got = b.results().astype(np.int64)
        assert np.array_equal(got, orc.counts()), (got, orc.counts())
        for i in range(n):
            compare_states(b.state(i), orc, i, f"{path} sim {i}")
        r0, c0, h, w = (orc.roi[:, k] for k in range(4))
        clipped = max(clipped, int(((r0 == 1) | (c0 == 1) | (r0 + h == 52) | (c0 + w == 52)).sum()))
    assert clipped >= 6, orc.roi.tolist()


@pytest.mark.parametrize("path", ["persistent", "graph"])
def test_protocol_length_and_resolution_match_oracle(sim, params, path):
    """A dose fires at step s only when s is a multiple of the protocol
    resolution and s < the protocol length (protocol.h:9-14).  Resolution
    150, length 600: slots 0-3 fire and slot 4 (step 600) does not, nor does
    slot 7 beyond it.  Both paths match the oracle bit for bit after 1200
    steps; on the oracle, the doses in slots 4-7 change nothing, while a
    length of 601 lets slot 4 fire and changes the cells' irradiation."""
    from emt6ro.batch import Batch
    ins = inputs(sim, params)
    v = params.values()
    doses = np.zeros((2, 8), np.float32)
    doses[0, [0, 2, 4, 7]] = [1.0, 2.0, 3.0, 4.0]
    doses[1, :] = 0.75
    T = len(ins)
    tumor_index = np.repeat(np.arange(T), 2)
    prot_index = np.tile(np.arange(2), T)
    n, steps, seed = len(tumor_index), 1200, 21
    b = Batch(params, [s for s, _ in ins], tumor_index, prot_index,
              np.full(n, 2 ** 32 - 1, np.uint32), doses, 150, 600, sim_id=np.arange(n), seed=seed)
    log(f"[parity] protocol resolution 150, length 600, path={path}")
    b.run(steps, path=path)

    def oracle(d, length):
        orc = O.OracleBatch(v, [ins[t][1] for t in tumor_index], d, prot_index, np.arange(n),
                            150, length)
        orc.run(steps, seed=seed, nthreads=4)
        return orc

    orc = oracle(doses, 600)
    got = b.results().astype(np.int64)
    assert np.array_equal(got, orc.counts()), (got, orc.counts())
    for i in range(n):
        compare_states(b.state(i), orc, i, f"{path} sim {i}")
    cut = doses.copy()
    cut[:, 4:] = 0.0
    same = oracle(cut, 600)
    assert np.array_equal(same.sub, orc.sub) and np.array_equal(same.cells["R"], orc.cells["R"])
    assert not np.array_equal(oracle(doses, 601).cells["R"], orc.cells["R"])
