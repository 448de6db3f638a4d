"""Pins how the CPU oracle composes a simulation step, on hand-built tiny
lattices whose outcomes are worked out by hand from the reference's source
(CPU only).  The component rules (updateState, progressClock, repair, the
neighbour draw, ROI, diffusion) are pinned in test_oracle.py; these cases
check what a mistake shared by the oracle and the GPU kernels could get
wrong in putting them together:

  - list order and child append (simulation.cu:41-64, :105; grid.h GRID_FOR);
  - Site::step order: updateState, irradiate, metabolise, progressClock,
    tryRepair (site.h:62-82);
  - the vacant-neighbour snapshot taken before the step (simulation.cu:79-83);
  - division timing from the cumulative cycle ends (cell.cu:21-32), the parent
    reset and the child created with cumulative normal phase lengths and the
    parent's irradiation (cell-division.h:10-30);
  - dose firing at s % res == 0 and s < length (protocol.h:9-14);
  - one division per lattice per step, won by the reduction-tree rank of the
    list entry (simulation.cu:86-105, cuda-utils.h:23-49);
  - ROI refresh every 32 steps and list rebuild every 128 (simulation.cu:
    153-166); no diffusion on a lattice without cells (diffusion.cu:142);
  - the two step-rule switches (duplicate entries, list-order vacancy) that
    the A/B statistics use.
"""
import numpy as np
import pytest

import oracle as O

from conftest import golden

VALUES = O.read_params_json(golden("default-parameters.json"))
DT_H = np.float32(np.float32(6.0) / np.float32(3600.0))   # one step in hours
TIMES = (6.0, 16.0, 18.0, 20.0, 20.1)                    # cumulative phase ends
G1, S, G2, M, D = range(5)
N = 21                                                    # bordered lattice side


def lattice(values=VALUES, n=N):
    return O.empty_state(values, n, n)


def put(s, r, c, phase=G1, prolif=0.0, times=TIMES, t_rep=0.0, R=0.0, delay=0.0, mode=0):
    W = s["st"].shape[1]
    s["st"][r, c] = O.OCCUPIED
    cell = s["cells"][r * W + c]
    cell["phase"], cell["prolif"], cell["mode"] = phase, np.float32(prolif), mode
    cell["times"] = np.asarray(times, np.float32)
    cell["t_rep"], cell["R"], cell["delay"] = np.float32(t_rep), np.float32(R), np.float32(delay)


def batch(s, values=VALUES, doses=None, res=300, length=None, sim_id=0):
    if doses is None:
        doses = np.zeros((1, 8), np.float32)
    doses = np.asarray(doses, np.float32).reshape(1, -1)
    if length is None:
        length = doses.shape[1] * res
    return O.OracleBatch(values, [s], doses, [0], [sim_id], res, length)


def cell_at(b, r, c):
    return b.cells[0][r * b.W + c]


def site(b, r, c):
    return r * b.W + c


def neighbours(b, r, c):
    return {site(b, r + dr, c + dc) for dr in (-1, 0, 1) for dc in (-1, 0, 1) if dr or dc}


def living(b):
    return int((b.st[0] == O.OCCUPIED).sum())


# ---------------------------------------------------------------------------
# List order and the child append
# ---------------------------------------------------------------------------
def test_list_order_is_column_group_scanner_order():
    """findOccupied runs 32 scanners; scanner g visits columns g, g+32, ... row
    by row, and the lists are concatenated in scanner order.  On a 53-wide
    lattice, scanner 1 holds columns 1 and 33, scanner 2 columns 2 and 34,
    scanner 17 column 17 (and 49)."""
    s = lattice(n=53)
    cells = [(5, 33), (3, 1), (2, 17), (5, 1), (2, 2), (4, 33), (3, 34)]
    for r, c in cells:
        put(s, r, c)
    b = batch(s)
    b.run(1, seed=0)
    want = [(3, 1), (4, 33), (5, 1), (5, 33), (2, 2), (3, 34), (2, 17)]
    assert b.nlist[0] == 7
    assert [divmod(int(x), 53) for x in b.list[0, :7]] == want


def _ready_cell(s, r, c, R=0.0, delay=0.0):
    """A cell in M whose clock passes the M end in its next step."""
    put(s, r, c, phase=M, prolif=np.float32(20.0) - DT_H / 2, R=R, delay=delay)


def test_division_appends_child_and_resets_parent():
    """M -> D in this step makes the cell ready (site.h:73); with all eight
    neighbours vacant the division happens in the same step: the child site
    becomes occupied and is pushed at the end of the list; the parent goes back
    to G1 with prolif 0 and keeps its cycle ends, mode and irradiation; the
    child starts in G1 with prolif, t_rep 0, the parent's R and the repair
    delay of that R (cell-division.h:21-30)."""
    s = lattice()
    put(s, 3, 3)
    put(s, 15, 15)
    R = np.float32(0.5)
    p = O.make_params(VALUES, N, N)
    delay = np.float32(O.lib.eo_delay_time(float(R), p))
    _ready_cell(s, 10, 10, R=R, delay=delay)
    b = batch(s)
    b.run(1, seed=3)
    assert living(b) == 4
    assert b.nlist[0] == 4
    child = int(b.list[0, 3])
    assert child in neighbours(b, 10, 10)
    assert b.st[0].reshape(-1)[child] == O.OCCUPIED
    par = cell_at(b, 10, 10)
    assert par["phase"] == G1 and par["prolif"] == 0.0
    assert np.array_equal(par["times"], np.asarray(TIMES, np.float32))
    assert par["R"] == R and par["mode"] == 0
    ch = b.cells[0][child]
    assert ch["phase"] == G1 and ch["prolif"] == 0.0 and ch["t_rep"] == 0.0
    assert ch["R"] == R and ch["mode"] == 0
    # delay = 3.3414 exp(0.1492 R) (cell.cu:93-100), in closed form
    assert ch["delay"] == pytest.approx(3.3414 * np.exp(0.1492 * 0.5), rel=1e-6)
    # cumulative phase ends; G2, M and D lengths have sd 0, so they add exactly
    t = ch["times"]
    assert t[2] == np.float32(t[1] + np.float32(2.0))
    assert t[3] == np.float32(t[2] + np.float32(2.0))
    assert t[4] == np.float32(t[3] + np.float32(0.1))
    assert 0 < t[0] < t[1]
    # the child is stepped from the next step on: its clock moves once
    b.run(1, seed=3)
    assert b.cells[0][child]["prolif"] == DT_H
    assert cell_at(b, 10, 10)["prolif"] == DT_H


def test_unirradiated_child_has_no_delay():
    s = lattice()
    _ready_cell(s, 10, 10)
    b = batch(s)
    b.run(1, seed=5)
    child = int(b.list[0, 1])
    assert b.cells[0][child]["R"] == 0.0 and b.cells[0][child]["delay"] == 0.0


def test_division_step_follows_cumulative_cycle_ends():
    """A lone G1 cell at prolif 0: progressClock adds dt/3600 per step
    (float) and advances one phase when prolif passes the current phase end;
    M -> D makes it ready, and it divides in that step.  The expected step is
    the float accumulation's first crossing of the M end, G1 -> S -> G2 -> M
    crossings in between."""
    times = np.asarray([1.5, 3.0, 3.75, 4.5, 4.6], np.float32)
    p = np.float32(0.0)
    phase, k, crossings = 0, 0, []
    while phase < 4:
        p = np.float32(p + DT_H)
        if p > times[phase]:
            phase += 1
            crossings.append(k)
        k += 1
    div_step = crossings[-1]
    s = lattice()
    put(s, 10, 10, times=times)
    b = batch(s)
    done = 0
    for stop, ph in zip(crossings[:3], (S, G2, M)):
        b.run(stop + 1 - done, seed=1)
        done = stop + 1
        assert cell_at(b, 10, 10)["phase"] == ph
        assert living(b) == 1
    b.run(div_step - done, seed=1)
    assert living(b) == 1 and cell_at(b, 10, 10)["phase"] == M
    b.run(1, seed=1)
    assert living(b) == 2
    assert cell_at(b, 10, 10)["phase"] == G1 and cell_at(b, 10, 10)["prolif"] == 0.0


def test_child_cycle_lengths_are_the_parameter_normals():
    """Over many divisions, the child's G1 end is N(6, 1) and its S length
    N(10, 2) (default-parameters.json cycle_times; cell-division.h:13-17)."""
    from scipy import stats
    n = 400
    states = []
    for _ in range(n):
        s = lattice(n=9)
        _ready_cell(s, 4, 4)
        states.append(s)
    b = O.OracleBatch(VALUES, states, np.zeros((1, 8), np.float32), [0] * n, np.arange(n),
                      300, 2400)
    b.run(1, seed=11)
    assert (b.nlist == 2).all()
    kids = b.cells[np.arange(n), b.list[:, 1]]
    g1 = kids["times"][:, 0].astype(np.float64)
    s_len = (kids["times"][:, 1] - kids["times"][:, 0]).astype(np.float64)
    assert stats.kstest(g1, "norm", args=(6.0, 1.0)).pvalue > 1e-3
    assert stats.kstest(s_len, "norm", args=(10.0, 2.0)).pvalue > 1e-3


# ---------------------------------------------------------------------------
# Site::step order
# ---------------------------------------------------------------------------
def test_irradiation_comes_before_the_clock_and_repair():
    """A cell in repair (t_rep 0.5 h, R 1 Gy, long delay) gets 1 Gy at step 0.
    updateState -> irradiate: R = 1 / (1 + 0.5/0.8) + 1, t_rep = 0 ->
    metabolise -> progressClock runs because t_rep is now 0 -> tryRepair
    does nothing (t_rep 0 and no cycle change).  Had the clock or the repair
    run before the irradiation, prolif would stay 1.0 or R would differ."""
    s = lattice()
    put(s, 10, 10, prolif=1.0, t_rep=0.5, R=1.0, delay=100.0)
    doses = np.zeros(8, np.float32)
    doses[0] = 1.0
    b = batch(s, doses=doses)
    b.run(1, seed=0)
    c = cell_at(b, 10, 10)
    R = np.float32(np.float32(1.0) / np.float32(1.0 + np.float32(np.float32(0.5) / np.float32(0.8))))
    assert c["R"] == np.float32(R + np.float32(1.0))
    assert c["t_rep"] == 0.0
    assert c["prolif"] == np.float32(np.float32(1.0) + DT_H)
    assert c["delay"] == pytest.approx(3.3414 * np.exp(0.1492 * float(c["R"])), rel=1e-6)


def test_repair_stops_the_clock_without_a_dose():
    """No dose: t_rep > 0 blocks progressClock, and tryRepair advances
    t_rep by one step (cell.cu:9-12, cell.h:102-110)."""
    s = lattice()
    put(s, 10, 10, prolif=1.0, t_rep=0.5, R=1.0, delay=100.0)
    b = batch(s)
    b.run(1, seed=0)
    c = cell_at(b, 10, 10)
    assert c["prolif"] == np.float32(1.0)
    assert c["t_rep"] == np.float32(np.float32(0.5) + DT_H)
    assert c["R"] == np.float32(1.0)


@pytest.mark.parametrize("ox,mode", [(14.0, 0), (10.0, 0), (5.0, 1)])
def test_state_update_reads_substrates_before_metabolism(ox, mode):
    """A lone cell's ROI is its own site with dist 0, so the whole temporary
    grid is masked and diffusion resets the site to the external levels each
    step (diffusion.cu:150-184).  updateState then picks the mode from those
    levels and metabolise takes that mode's amounts: with ox 10 the cell is
    aerobic (10 >= 9.96), leaving 10 - 9.96; reading after the metabolism
    would have made it anaerobic."""
    v = VALUES.copy()
    v[20] = ox
    s = lattice(values=v)
    put(s, 10, 10)
    b = batch(s, values=v)
    b.run(1, seed=0)
    c = cell_at(b, 10, 10)
    assert c["mode"] == mode
    ext = v[19:22]
    met = v[7 + 3 * mode: 10 + 3 * mode]
    got = b.sub[0][:, 10, 10]
    # diffusion of the constant masked grid may move the level by an ulp
    assert np.allclose(got, (ext - met).astype(np.float32), rtol=1e-6, atol=1e-6)


def test_death_frees_the_site_in_the_same_step():
    """updateState kills a cell whose clock reached t_D - dt (cell.cu:33-36);
    the site is vacant after the step (site.h:77-78)."""
    s = lattice()
    put(s, 10, 10, phase=D, prolif=np.float32(20.1) - DT_H)
    put(s, 5, 5)
    b = batch(s)
    b.run(1, seed=0)
    assert b.st[0, 10, 10] == O.VACANT and living(b) == 1


# ---------------------------------------------------------------------------
# Vacant-neighbour snapshot
# ---------------------------------------------------------------------------
def _g1_stop_case(extra_vacant, neighbour_dies):
    """Cell C at (10, 10) in G1 just past t_G1 - 2 dt, seven stable neighbours
    and one (or two) vacant; the neighbour at (10, 9) (earlier in list order)
    optionally dies in this step."""
    s = lattice()
    ring = sorted(neighbours_rc(10, 10))
    vacant = {(9, 11)} | ({(11, 11)} if extra_vacant else set())
    for r, c in ring:
        if (r, c) in vacant:
            continue
        if (r, c) == (10, 9) and neighbour_dies:
            put(s, r, c, phase=D, prolif=np.float32(20.1) - DT_H)
        else:
            put(s, r, c)
    put(s, 10, 10, prolif=np.float32(6.0) - DT_H)
    return s


def neighbours_rc(r, c):
    return {(r + dr, c + dc) for dr in (-1, 0, 1) for dc in (-1, 0, 1) if dr or dc}


@pytest.mark.parametrize("neighbour_dies", [False, True])
def test_g1_stopping_uses_the_pre_step_neighbour_count(neighbour_dies):
    """One vacant neighbour before the step: enterG1SStopping holds
    (cell.cu:38-45), so C cannot proliferate and turns aerobic quiescent
    (mode 2) with its clock stopped, even when an earlier list entry's death
    frees a second neighbour during the step (simulation.cu:79-83: counts are
    taken before any cell is stepped)."""
    b = batch(_g1_stop_case(False, neighbour_dies))
    b.run(1, seed=0)
    if neighbour_dies:
        assert b.st[0, 10, 9] == O.VACANT
    c = cell_at(b, 10, 10)
    assert c["mode"] == 2
    assert c["prolif"] == np.float32(np.float32(6.0) - DT_H)


def test_two_vacant_neighbours_let_the_cell_proliferate():
    b = batch(_g1_stop_case(True, False))
    b.run(1, seed=0)
    c = cell_at(b, 10, 10)
    assert c["mode"] == 0
    assert c["prolif"] == np.float32(np.float32(6.0) - DT_H + DT_H)


# ---------------------------------------------------------------------------
# Dose firing
# ---------------------------------------------------------------------------
def test_dose_fires_on_multiples_of_the_resolution_only():
    """getDose(step) = data[step / res] when step < length and step % res
    == 0 (protocol.h:9-14).  Doses in slots 2 and 3 (res 300) fire at steps
    600 and 900 and at no other step."""
    s = lattice()
    put(s, 10, 10, times=(60.0, 70.0, 72.0, 74.0, 74.1))
    doses = np.zeros(8, np.float32)
    doses[2], doses[3] = 0.75, 0.5
    b = batch(s, doses=doses)
    b.run(600, seed=0)
    assert cell_at(b, 10, 10)["R"] == 0.0
    b.run(1, seed=0)
    assert cell_at(b, 10, 10)["R"] == np.float32(0.75)
    b.run(299, seed=0)
    assert cell_at(b, 10, 10)["R"] == np.float32(0.75)
    b.run(1, seed=0)
    # R / (1 + t_rep / 0.8) + d with t_rep 0 (no cycle change since)
    assert cell_at(b, 10, 10)["R"] == np.float32(np.float32(0.75) + np.float32(0.5))


def test_dose_at_or_after_length_does_not_fire():
    s = lattice()
    put(s, 10, 10, times=(60.0, 70.0, 72.0, 74.0, 74.1))
    doses = np.zeros(8, np.float32)
    doses[2] = 1.0
    b = batch(s, doses=doses, length=600)
    b.run(700, seed=0)
    assert cell_at(b, 10, 10)["R"] == 0.0
    s = lattice()
    put(s, 10, 10, times=(60.0, 70.0, 72.0, 74.0, 74.1))
    b = batch(s, doses=doses, length=601)
    b.run(601, seed=0)
    assert cell_at(b, 10, 10)["R"] == np.float32(1.0)


def test_dose_at_step_zero_fires():
    s = lattice()
    put(s, 10, 10)
    doses = np.zeros(8, np.float32)
    doses[0] = 2.0
    b = batch(s, doses=doses)
    b.run(1, seed=0)
    assert cell_at(b, 10, 10)["R"] == np.float32(2.0)


# ---------------------------------------------------------------------------
# One division per step and the arbitration rank
# ---------------------------------------------------------------------------
def test_one_division_per_lattice_per_step_by_reduction_rank():
    """List entries 1 and 2 are both ready with free neighbours.  The
    reference steps entry i on thread i and reduces with first-non-zero over
    __shfl_down offsets 16, 8, 4, 2, 1, where lane 0 takes lane 0, then 16, 8,
    24, 4, ... (lane l ranks by bitrev5(l)): lane 2 (rank 8) beats lane 1
    (rank 16).  So entry 2's cell divides at step 0 and entry 1's at step 1."""
    s = lattice(n=53)
    put(s, 5, 1)             # scanner 1, row 5: entry 0 (not ready)
    _ready_cell(s, 20, 33)   # scanner 1, row 20: entry 1
    _ready_cell(s, 20, 2)    # scanner 2: entry 2
    b = batch(s)
    b.run(1, seed=0)
    assert [divmod(int(x), 53) for x in b.list[0, :3]] == [(5, 1), (20, 33), (20, 2)]
    assert b.nlist[0] == 4 and (b.st[0] == O.OCCUPIED).sum() == 4
    child = divmod(int(b.list[0, 3]), 53)
    assert max(abs(child[0] - 20), abs(child[1] - 2)) == 1
    assert b.cells[0][20 * 53 + 2]["phase"] == G1
    assert b.cells[0][20 * 53 + 33]["phase"] == D
    b.run(1, seed=0)
    assert b.nlist[0] == 5
    child2 = divmod(int(b.list[0, 4]), 53)
    assert max(abs(child2[0] - 20), abs(child2[1] - 33)) == 1


def test_entry_zero_beats_every_other_entry():
    s = lattice(n=53)
    for r in (5, 9, 13):     # scanner 2: entries 0, 1, 2, all ready
        _ready_cell(s, r, 2)
    b = batch(s)
    b.run(1, seed=0)
    assert b.nlist[0] == 4
    child = divmod(int(b.list[0, 3]), 53)
    assert max(abs(child[0] - 5), abs(child[1] - 2)) == 1
    assert b.cells[0][5 * 53 + 2]["phase"] == G1


# ---------------------------------------------------------------------------
# Cadences
# ---------------------------------------------------------------------------
def _growing_cluster():
    """A 3x3 block whose cells reach M's end at staggered steps, so the
    cluster grows at its edges over the first hundred steps."""
    s = lattice()
    k = 0
    for r in (9, 10, 11):
        for c in (9, 10, 11):
            put(s, r, c, phase=M, prolif=np.float32(20.0) - np.float32(0.0016) * (3 + 11 * k))
            k += 1
    return s


def _roi(st):
    st = np.ascontiguousarray(st, np.uint8)
    roi = np.zeros(5, np.int32)
    O.lib.eo_find_roi(st.shape[0], st.shape[1], O.ptr(st),
                      roi.ctypes.data_as(O.C.POINTER(O.C.c_int32)))
    return roi


def test_roi_refreshes_every_32_steps_only():
    """updateROIs runs at step % 32 == 0 (simulation.cu:160-162): after step
    t the ROI is the one of the state before step 32 * (t // 32), even when
    cells born since lie outside it."""
    b = batch(_growing_cluster())
    snaps, stale = {}, 0
    for t in range(100):
        if t % 32 == 0:
            snaps[t] = b.st[0].copy()
        before = b.st[0].copy()
        b.run(1, seed=2)
        want = _roi(snaps[32 * (t // 32)])
        assert np.array_equal(b.roi[0], want), t
        stale += not np.array_equal(_roi(before), want)
    assert stale > 0
    assert not np.array_equal(_roi(snaps[0]), _roi(snaps[96]))


def test_list_rebuilds_every_128_steps_only():
    """findOccupied runs at step % 128 == 0 (simulation.cu:154-159); between
    rebuilds the list only grows by the pushed children, and the dead stay
    in it."""
    s = _growing_cluster()
    put(s, 3, 3, phase=D, prolif=np.float32(20.1) - DT_H)     # dies at step 0
    b = batch(s)
    at_rebuild = None
    births, divisions, deaths = 0, 0, 0
    for t in range(260):
        before = b.st[0].copy()
        if t % 128 == 0:
            at_rebuild = int((before == O.OCCUPIED).sum())
            births = 0
        b.run(1, seed=4, instr=True)
        born = int(b.instr[0, 3]) - divisions          # division counter
        divisions += born
        deaths += int(((before == O.OCCUPIED) & (b.st[0] == O.VACANT)).sum())
        assert born <= 1
        births += born
        assert b.nlist[0] == at_rebuild + births, t
    assert b.st[0, 3, 3] == O.VACANT
    assert divisions > 5 and deaths > 1


def test_no_cells_no_diffusion():
    """An empty lattice has ROI 0 and skips diffusion (diffusion.cu:73,
    :142): arbitrary substrates stay as they are."""
    s = lattice()
    rng = np.random.default_rng(0)
    s["sub"][:] = rng.random(s["sub"].shape).astype(np.float32)
    b = batch(s)
    before = b.sub.copy()
    b.run(5, seed=0)
    assert np.array_equal(b.sub, before)


# ---------------------------------------------------------------------------
# Step-rule switches (A/B statistics only)
# ---------------------------------------------------------------------------
def _reoccupied_site(rules):
    """Cell X at (10, 11) dies at step 0; its neighbour P at (10, 10), whose
    other seven neighbours are occupied, is ready from step 1 on and divides
    into X's site as soon as its draw picks it.  Returns the batch right after
    that division and the step it happened."""
    s = lattice()
    for r, c in neighbours_rc(10, 10):
        if (r, c) == (10, 11):
            put(s, r, c, phase=D, prolif=np.float32(20.1) - DT_H)
        else:
            put(s, r, c, times=(60.0, 70.0, 72.0, 74.0, 74.1))
    put(s, 10, 10, phase=M, prolif=np.float32(20.0) - np.float32(1.5) * DT_H,
        times=(6.0, 16.0, 18.0, 20.0, 30.0))
    b = batch(s)
    b.run(1, seed=9, rules=rules)
    assert b.st[0, 10, 11] == O.VACANT
    for t in range(1, 120):
        b.run(1, seed=9, rules=rules)
        if b.st[0, 10, 11] == O.OCCUPIED:
            return b, t
    raise AssertionError("no division into the freed site")


@pytest.mark.parametrize("rules", [0, O.RULE_Q3_EVERY_ENTRY])
def test_reoccupied_site_is_stepped_once_or_per_entry(rules):
    """The freed site keeps its old list entry until the next rebuild, and the
    child is pushed as a second entry.  Rule 0 steps the site once per step;
    the reference's literal behaviour steps it once per entry (site.h:65
    checks only that the site is occupied)."""
    b, t = _reoccupied_site(rules)
    assert t < 127
    x = site(b, 10, 11)
    entries = [i for i in range(b.nlist[0]) if b.list[0, i] == x]
    assert len(entries) == 2
    b.run(1, seed=9, rules=rules)
    per_step = 2 if rules else 1
    want = np.float32(0.0)
    for _ in range(per_step):
        want = np.float32(want + DT_H)
    assert b.cells[0][x]["prolif"] == want


PURPOSE_NEIGHBOUR = 2   # EO_PURPOSE_NEIGHBOUR of emt6ro_oracle.h


def _dying_target_case(target_later):
    """P at (10, 10) becomes ready at step 0 with all eight neighbours
    occupied; the neighbour its step-0 draw picks dies in step 0.  List order
    around P: column 9 (scanner 9), then (9, 10), P, (11, 10) (scanner 10),
    then column 11; a seed is chosen so that the dying neighbour comes after P
    (`target_later`) or before it."""
    x = 10 * N + 10
    for seed in range(500):
        u = O.lib.eo_uniform(seed, 0, 0, 0, x, PURPOSE_NEIGHBOUR)
        cr, cc = O.C.c_int(), O.C.c_int()
        O.lib.eo_choose_neighbour(10, 10, u, O.C.byref(cr), O.C.byref(cc))
        target = (cr.value, cc.value)
        later = target[1] == 11 or target == (11, 10)
        if later == target_later:
            break
    s = lattice()
    for r, c in neighbours_rc(10, 10):
        if (r, c) == target:
            put(s, r, c, phase=D, prolif=np.float32(20.1) - DT_H)
        else:
            put(s, r, c, times=(60.0, 70.0, 72.0, 74.0, 74.1))
    _ready_cell(s, 10, 10)
    return s, target, seed


@pytest.mark.parametrize("target_later", [True, False])
@pytest.mark.parametrize("rules", [0, O.RULE_Q4_LIST_ORDER])
def test_division_into_a_site_freed_in_the_same_step(rules, target_later):
    """Rule 0: a target is vacant if its cell died anywhere in this step.
    List order (the reference thread's read right after its own step,
    simulation.cu:88-93): only deaths of earlier list entries are seen."""
    s, target, seed = _dying_target_case(target_later)
    b = batch(s)
    b.run(1, seed=seed, rules=rules)
    order = [divmod(int(v), N) for v in b.list[0, :9]]
    assert (order.index(target) > order.index((10, 10))) == target_later
    divided = b.nlist[0] == 10
    assert divided == (rules == 0 or not target_later)
    assert b.st[0][target] == (O.OCCUPIED if divided else O.VACANT)
    assert cell_at(b, 10, 10)["phase"] == (G1 if divided else D)


def test_literal_q3_q4_rules_ab_on_config1():
    """tests/golden/rules_ab.json: the config-1 workload (144 h) under the
    deterministic rules (0) and the reference's literal Q3/Q4 behaviour (3),
    1,000 paired replicates each (same streams).  The living-count
    distributions are not distinguishable: KS p > 0.05, paired mean
    difference within 3 standard errors and under 1% of the mean."""
    import json
    from scipy import stats
    d = json.load(open(golden("rules_ab.json")))
    a = np.asarray(d["rules_0"]["counts"], np.float64)
    b = np.asarray(d["rules_3"]["counts"], np.float64)
    assert len(a) == len(b) == 1000
    assert d["workload"]["steps"] == 86400 and d["workload"]["size"] == 53
    assert abs(a.mean() - d["rules_0"]["mean"]) < 1e-9
    diff = b - a
    se = diff.std(ddof=1) / np.sqrt(len(diff))
    assert stats.ks_2samp(a, b).pvalue > 0.05
    assert abs(diff.mean()) < 3 * se and abs(diff.mean()) < 0.01 * a.mean()
    # Paired: the same streams give strongly related outcomes.
    assert np.corrcoef(a, b)[0, 1] > 0.5
