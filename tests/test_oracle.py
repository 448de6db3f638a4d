"""Pins the CPU oracle (oracle/emt6ro_oracle.c) against known answers:
published Philox vectors, the reference's gtest tables, closed forms of the
stencil, and brute-force restatements of the reduction and ROI mechanics.
CPU only."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle as O

from conftest import golden

lib = O.lib


def cell(**kw):
    c = np.zeros(1, O.CELL_DTYPE)
    for k, v in kw.items():
        c[k] = v
    return c


# ---------------------------------------------------------------------------
# RNG
# ---------------------------------------------------------------------------
# Known-answer vectors of Philox-4x32-10 published with Random123
# (Salmon et al., SC'11): (counter, key) -> output.
PHILOX_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(ctr, key, want):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.eo_philox4x32_10(c, k, o)
    assert list(o) == want


def test_u01_range_and_endpoints():
    assert lib.eo_u01(0) == 2.0 ** -24
    assert lib.eo_u01(0xffffffff) == 1.0
    assert lib.eo_u01(0x00000100) == 2.0 ** -23


def test_uniform_is_uniform():
    from scipy import stats
    u = np.array([lib.eo_uniform(7, 0, 3, s, x, 1) for s in range(40) for x in range(250)])
    assert u.min() > 0 and u.max() <= 1
    assert stats.kstest(u, "uniform").pvalue > 1e-3


def test_normals_are_standard_normal():
    from scipy import stats
    z = []
    buf = (C.c_float * 5)()
    for s in range(2000):
        lib.eo_normals5(11, 0, 5, s, 1234, buf)
        z.extend(buf)
    z = np.array(z)
    assert abs(z.mean()) < 0.03 and abs(z.std() - 1) < 0.03
    assert stats.kstest(z, "norm").pvalue > 1e-3


# ---------------------------------------------------------------------------
# Elementary functions against double precision
# ---------------------------------------------------------------------------
def test_expf_accuracy():
    xs = np.concatenate([np.linspace(-10, 3, 4001), [0.0, -0.4993 * 10, 0.1492 * 10]])
    got = np.array([lib.eo_expf(float(x)) for x in xs.astype(np.float32)], np.float64)
    ref = np.exp(xs.astype(np.float32).astype(np.float64))
    rel = np.abs(got - ref) / ref
    assert rel.max() < 3e-7, rel.max()


def test_logf_accuracy():
    u = np.array([lib.eo_u01(int(x)) for x in np.linspace(0, 2 ** 32 - 1, 5001).astype(np.uint64)])
    got = np.array([lib.eo_logf(float(x)) for x in u])
    ref = np.log(u.astype(np.float64))
    err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3)
    assert err.max() < 1e-6, err.max()
    assert lib.eo_logf(1.0) == 0.0


def test_sincos_accuracy():
    s, c = C.c_float(), C.c_float()
    worst = 0.0
    for u in np.linspace(2 ** -24, 1.0, 4097, dtype=np.float32):
        lib.eo_sincos2pi(float(u), C.byref(s), C.byref(c))
        a = 2 * math.pi * float(u)
        worst = max(worst, abs(s.value - math.sin(a)), abs(c.value - math.cos(a)))
    assert worst < 3e-7, worst


# ---------------------------------------------------------------------------
# Cell rules: the reference's gtest tables as known answers
# ---------------------------------------------------------------------------
def table_params():
    """cell-update-state-test.cc:20-37 parameters."""
    v = np.zeros(40, np.float32)
    v[2] = 3600.0                                   # time_step
    v[7:10] = [21, 9, 0]                            # aerobic proliferation
    v[10:13] = [54, 0, -3e-3]                       # anaerobic proliferation
    v[13:16] = [18, 8, 0]                           # aerobic quiescence
    v[16:19] = [45, 0, -3e-5]                       # anaerobic quiescence
    v[22], v[23] = 5e-2, 2e-2                       # death_gi, quiescence_gi
    return O.make_params(v, 3, 3)


G1, S, G2, M, D = range(5)
AP, ANP, AQ, ANQ = range(4)
MOCK = AQ
# (levels, vacant, g1_time, phase, prolif_time, alive, mode): the tables of
# cell-update-state-test.cc:52-89.
UPDATE_STATE = [
    ((17, 10, 0), 0, 100, G1, 50, False, MOCK),
    ((18, 10, 6e-2), 0, 100, G1, 50, False, MOCK),
    ((18, 10, 0), 0, 20, M, 19, False, MOCK),
    ((44, 7, 0), 0, 20, G1, 19, False, MOCK),
    ((45, 7, 0), 0, 20, G2, 21, True, ANQ),
    ((18, 8, 0), 0, 20, G2, 21, True, AQ),
    ((18, 10, 3e-2), 0, 20, M, 15, False, MOCK),
    ((44, 7, 3e-2), 0, 20, G1, 15, False, MOCK),
    ((45, 7, 3e-2), 0, 20, G2, 15, True, ANQ),
    ((18, 8, 3e-2), 0, 20, G2, 15, True, AQ),
    ((18, 8, 0), 0, 20, M, 15, False, MOCK),
    ((44, 7, 0), 0, 20, G1, 15, False, MOCK),
    ((45, 7, 0), 0, 20, G2, 15, True, ANQ),
    ((18, 8, 0), 0, 20, G2, 15, True, AQ),
    ((54, 7, 0), 0, 20, G1, 15, True, ANP),
    ((21, 9, 0), 0, 20, G1, 15, True, AP),
    ((21, 8, 0), 0, 20, G1, 15, True, AQ),
    ((21, 8, 0), 0, 20, M, 15, False, MOCK),
]


@pytest.mark.parametrize("levels,vac,g1,phase,prolif,alive,mode", UPDATE_STATE)
def test_update_state_table(levels, vac, g1, phase, prolif, alive, mode):
    p = table_params()
    c = cell(phase=phase, prolif=prolif)
    c["times"][0][0] = g1
    c["times"][0][4] = 100
    sub = (C.c_float * 3)(*levels)
    got = lib.eo_update_state(c.ctypes.data_as(C.c_void_p), sub, C.byref(p), vac)
    assert bool(got) == alive
    if alive:
        assert c["mode"][0] == mode


# cell-progress-clock-test.cc:45-58: time_step 360, cycle ends {1,2,3,4,5};
# (phase, expected phase, mode, prolif, expected change flag).
PROGRESS = [
    (G1, G1, AQ, 0.95, False),
    (G1, G1, ANQ, 0.95, False),
    (G1, S, AP, 0.95, True),
    (S, S, AP, 1.85, False),
    (M, D, ANP, 3.95, False),
]


@pytest.mark.parametrize("ph,want_ph,mode,t,want", PROGRESS)
def test_progress_clock_table(ph, want_ph, mode, t, want):
    c = cell(phase=ph, mode=mode, prolif=t)
    c["times"][0] = [1, 2, 3, 4, 5]
    got = lib.eo_progress_clock(c.ctypes.data_as(C.c_void_p), 360.0)
    assert bool(got) == want
    assert c["phase"][0] == want_ph
    if mode in (AP, ANP):
        assert c["prolif"][0] == np.float32(t) + np.float32(360.0) / np.float32(3600)
    else:
        assert c["prolif"][0] == np.float32(t)


@pytest.mark.parametrize("mode", [AP, ANP, AQ, ANQ])
def test_metabolise_table(mode):
    """cell-metabolise-test.cc:12-63: {1,1,1} minus the mode's amounts, exactly."""
    v = np.zeros(40, np.float32)
    v[7:19] = [0.11, 0.12, 0.13, 0.21, 0.22, 0.23, 0.31, 0.32, 0.33, 0.41, 0.42, 0.43]
    p = O.make_params(v, 3, 3)
    c = cell(mode=mode)
    sub = (C.c_float * 3)(1, 1, 1)
    lib.eo_metabolise(c.ctypes.data_as(C.c_void_p), sub, C.byref(p))
    want = np.float32(1) - v[7 + 3 * mode:10 + 3 * mode]
    assert list(np.array(sub[:], np.float32)) == list(want)


def test_irradiate_and_repair_closed_forms():
    v = O.read_params_json(golden("default-parameters.json"))
    p = O.make_params(v, 53, 53)
    c = cell(R=1.0, t_rep=0.4)
    lib.eo_irradiate(c.ctypes.data_as(C.c_void_p), 1.25, C.byref(p))
    R = 1.0 / (1 + 0.4 / 0.8) + 1.25
    assert c["R"][0] == pytest.approx(R, rel=1e-6)
    assert c["t_rep"][0] == 0
    assert c["delay"][0] == pytest.approx(3.3414 * math.exp(0.1492 * R), rel=1e-6)
    # A repair decision with u above the death probability clears the damage.
    c = cell(R=2.0, t_rep=10.0, delay=1.0)
    assert lib.eo_try_repair(c.ctypes.data_as(C.c_void_p), 0, C.byref(p), 0.999) == 1
    assert c["R"][0] == 0 and c["t_rep"][0] == 0 and c["delay"][0] == 0
    pdeath = 1 - math.exp(-0.4993 * 2.0)
    c = cell(R=2.0, t_rep=10.0, delay=1.0)
    assert lib.eo_try_repair(c.ctypes.data_as(C.c_void_p), 0, C.byref(p), pdeath - 1e-4) == 0
    # No repair clock without damage and a cycle change.
    c = cell(R=2.0, t_rep=0.0, delay=5.0)
    assert lib.eo_try_repair(c.ctypes.data_as(C.c_void_p), 0, C.byref(p), 0.0) == 1
    assert c["t_rep"][0] == 0


# ---------------------------------------------------------------------------
# Neighbour choice (cell-division.h:30-54)
# ---------------------------------------------------------------------------
DIAG = {0: (-1, -1), 1: (1, -1), 2: (-1, 1), 3: (1, 1)}
ORTHO = {0: (0, -1), 1: (-1, 0), 2: (0, 1), 3: (1, 0)}


def test_choose_neighbour_weights_and_edges():
    cr, cc = C.c_int(), C.c_int()
    counts = {}
    n = 1 << 16
    for k in range(1, n + 1):
        u = k / n
        d = lib.eo_choose_neighbour(10, 10, u, C.byref(cr), C.byref(cc))
        off = (cr.value - 10, cc.value - 10)
        want = DIAG[d] if d < 4 else ORTHO[min(d - 4, 3)] if d - 4 < 4 else ORTHO[0]
        if d < 8:
            assert off == want
        counts[off] = counts.get(off, 0) + 1
    diag = np.mean([counts[o] for o in DIAG.values()])
    ortho = np.mean([counts[o] for o in ORTHO.values()])
    assert ortho / diag == pytest.approx(math.sqrt(2), rel=2e-3)
    # u = 1 maps past the last orthogonal bucket to (r, c-1) (num 4).
    lib.eo_choose_neighbour(10, 10, 1.0, C.byref(cr), C.byref(cc))
    assert (cr.value, cc.value) == (10, 9)


# ---------------------------------------------------------------------------
# Division arbitration: brute-force model of the reference's reduction
# ---------------------------------------------------------------------------
def _warp_reduce(vals, op):
    """cuda-utils.h:23-32 on one 32-lane warp: lane i combines with
    __shfl_down(i + offset), which returns the lane's own value when the
    source is out of range."""
    v = list(vals)
    for off in (16, 8, 4, 2, 1):
        v = [op(v[i], v[i + off] if i + off < 32 else v[i]) for i in range(32)]
    return v


def _block_reduce(vals, op):
    """cuda-utils.h:34-49 for a full block of len(vals) threads."""
    n = len(vals)
    warps = [_warp_reduce(vals[w * 32:(w + 1) * 32], op) for w in range(n // 32)]
    shared = [w[0] for w in warps]
    lane_vals = [x for w in warps for x in w]
    size = n // 32
    while size > 1:
        new = lane_vals[:]
        for t in range(size):
            new[t] = shared[t]
        w0 = _warp_reduce(new[:32], op)
        shared = [w0[0]]
        lane_vals = w0 + lane_vals[32:]
        size = (size + 31) // 32
    return shared[0]


def test_block_reduce_model_matches_reference_test():
    """cuda-utils-test.cu:15-22: a sum over 1024 threads of 1..1024 is 524800."""
    assert _block_reduce(list(range(1, 1025)), lambda a, b: a + b) == 524800


def test_arbitration_key_matches_reduction_tree():
    rng = np.random.default_rng(3)
    first_nonzero = lambda a, b: a if a else b  # noqa: E731  simulation.cu:97-98
    for trial in range(300):
        n = int(rng.integers(1, 1024))
        k = int(rng.integers(1, 12))
        cand = sorted(rng.choice(n, size=min(k, n), replace=False).tolist())
        # Thread t holds the last candidate among its entries t, t+512, ...
        vals = [0] * 512
        for i in cand:
            vals[i % 512] = i + 1
        winner = _block_reduce(vals, first_nonzero) - 1
        keys = {i: lib.eo_arbitration_key(i) for i in cand}
        assert min(cand, key=lambda i: keys[i]) == winner, (cand, winner)


# ---------------------------------------------------------------------------
# Region of interest and mask (diffusion.cu:16-100)
# ---------------------------------------------------------------------------
def _roi_bruteforce(st):
    H, W = st.shape
    rs, cs = np.nonzero(st == O.OCCUPIED)
    if len(rs) == 0:
        return [0, 0, 0, 0, 0]
    minr, maxr, minc, maxc = rs.min(), rs.max(), cs.min(), cs.max()
    if maxr == 0 and maxc == 0:
        return [0, 0, 0, 0, 0]
    midr = minr + (maxr - minr + 1) // 2
    midc = minc + (maxc - minc + 1) // 2
    d2 = ((rs - midr) ** 2 + (cs - midc) ** 2).max()
    dist = math.isqrt(int(d2))
    if dist * dist < d2:
        dist += 1
    sr = 1 if midr <= dist else midr - dist
    sc = 1 if midc <= dist else midc - dist
    sw = W - 1 - sc if midc + dist >= W - 1 else midc + dist - sc + 1
    sh = H - 1 - sr if midr + dist >= H - 1 else midr + dist - sr + 1
    return [sr, sc, sh, sw, dist]


@pytest.mark.parametrize("seed", range(40))
def test_roi_matches_bruteforce(seed):
    rng = np.random.default_rng(seed)
    H, W = int(rng.integers(5, 60)), int(rng.integers(5, 60))
    st = np.zeros((H, W), np.uint8)
    st[0, :] = st[-1, :] = st[:, 0] = st[:, -1] = O.MOCKED
    k = int(rng.integers(0, 30))
    r0, c0 = int(rng.integers(1, H - 1)), int(rng.integers(1, W - 1))
    for _ in range(k):
        r = int(np.clip(r0 + rng.integers(-8, 9), 1, H - 2))
        c = int(np.clip(c0 + rng.integers(-8, 9), 1, W - 2))
        st[r, c] = O.OCCUPIED
    roi = (C.c_int32 * 5)()
    lib.eo_find_roi(H, W, O.ptr(st), roi)
    assert list(roi) == _roi_bruteforce(st)


def test_mask_float_equals_integer_test():
    """fillBorderMask's float test equals a^2 + b^2 >= 4 d^2 (the kernels')."""
    for hm in range(3, 60, 3):
        for wm in (hm, hm + 1, hm + 2):
            for d in range(0, 31, 3):
                for r in range(hm):
                    for c in range(0, wm, 2):
                        a, b = 2 * (r + 1) - hm, 2 * (c + 1) - wm
                        assert lib.eo_mask(hm, wm, d, r, c) == int(a * a + b * b >= 4 * d * d)


# ---------------------------------------------------------------------------
# Diffusion
# ---------------------------------------------------------------------------
def _diffuse64(p, sub, roi):
    H, W = sub.shape[1:]
    out = np.zeros_like(sub, dtype=np.float64)
    r = (C.c_int32 * 5)(*roi)
    src = np.ascontiguousarray(sub, np.float64)
    lib.eo_diffuse_f64(C.byref(p), H, W, O.ptr(src), r, O.ptr(out))
    return out


def _diffuse32(p, sub, roi):
    H, W = sub.shape[1:]
    out = np.ascontiguousarray(sub, np.float32).copy()
    r = (C.c_int32 * 5)(*roi)
    lib.eo_diffuse_f32(C.byref(p), H, W, O.ptr(out), r)
    return out


def test_stencil_closed_form_quadratic():
    """The 9-point scheme adds exactly 2 D dt per substep for f = a r^2 + b c^2
    at sites whose neighbourhood stays inside the unmasked region: two
    substeps (time_step = 2 dt) add 4 (a + b) D dt far from the boundary."""
    v = O.read_params_json(golden("default-parameters.json"))
    v[2] = 0.5  # time_step: 2 substeps of 0.25 s
    H = W = 61
    p = O.make_params(v, H, W)
    rr, cc = np.mgrid[0:H, 0:W].astype(np.float64)
    a, b = 0.3, 0.7
    sub = np.stack([a * (rr - 30) ** 2 + b * (cc - 30) ** 2 + 1000.0] * 3)
    roi = [1, 1, 59, 59, 29]
    out = _diffuse64(p, sub, roi)
    D = v[3:6].astype(np.float64)
    dt = float(v[6])
    inner = (slice(15, 46), slice(15, 46))
    for s in range(3):
        want = sub[s][inner] + 4 * (a + b) * D[s] * dt
        assert np.allclose(out[s][inner], want, rtol=0, atol=1e-9)
    # Linear fields are fixed points of the scheme.
    lin = np.stack([0.5 * rr - 0.25 * cc + 500.0] * 3)
    out = _diffuse64(p, lin, roi)
    assert np.allclose(out[:, inner[0], inner[1]], lin[:, inner[0], inner[1]], rtol=0, atol=1e-9)


def test_constant_field_is_fixed_point():
    v = O.read_params_json(golden("default-parameters.json"))
    p = O.make_params(v, 53, 53)
    sub = np.stack([np.full((53, 53), v[19 + s], np.float64) for s in range(3)])
    for roi in ([12, 12, 29, 29, 14], [5, 6, 41, 41, 20], [20, 20, 9, 9, 4]):
        out = _diffuse64(p, sub, roi)
        assert np.allclose(out, sub, rtol=1e-13, atol=0)
    # A ROI clipped at the lattice edge leaves unmasked sites in the ring of
    # the temporary grid next to its zero outer ring (diffusion.cu:150-162),
    # so the constant field is not preserved there.
    out = _diffuse64(p, sub, [1, 1, 51, 51, 30])
    assert not np.allclose(out, sub, rtol=1e-3, atol=0)


def _diffuse_restated(p, sub, roi):
    """diffusion.cu:135-191 with numpy slices, in float64, for small grids."""
    r0, c0, h, w, dist = roi
    ext = np.array([p.ext[s] for s in range(3)], np.float64)
    th, tw = h + 4, w + 4
    tmp = np.zeros((3, th, tw))
    act = np.zeros((th, tw), bool)
    for dr in range(th):
        for dc in range(tw):
            if dr in (0, th - 1) or dc in (0, tw - 1):
                continue
            mr, mc = dr - 1, dc - 1
            fr, fc = mr + 1.0, mc + 1.0
            masked = math.sqrt((fr - (h + 2) / 2) ** 2 + (fc - (w + 2) / 2) ** 2) >= dist
            if masked:
                tmp[:, dr, dc] = ext
            else:
                tmp[:, dr, dc] = sub[:, r0 - 2 + dr, c0 - 2 + dc]
                act[dr, dc] = True
    hs = 2 * math.sqrt(2)
    f = 4 + hs
    k = np.array([p.diff_coeffs[s] * p.diff_time_step * hs / f for s in range(3)])

    def lap(t):
        o = t[:, :-2, 1:-1] + t[:, 1:-1, :-2] + t[:, 2:, 1:-1] + t[:, 1:-1, 2:]
        d = t[:, :-2, :-2] + t[:, :-2, 2:] + t[:, 2:, :-2] + t[:, 2:, 2:]
        return (o + d / math.sqrt(2) - f * t[:, 1:-1, 1:-1]) * k[:, None, None]

    steps = int(p.time_step / p.diff_time_step)
    for _ in range(max(steps - 1, 0)):
        new = tmp.copy()
        new[:, 1:-1, 1:-1] += lap(tmp)
        tmp = np.where(act[None], new, tmp)
    new = tmp.copy()
    new[:, 1:-1, 1:-1] += lap(tmp)
    tmp[:, 2:h + 2, 2:w + 2] = new[:, 2:h + 2, 2:w + 2]
    out = sub.copy()
    out[:, r0:r0 + h, c0:c0 + w] = tmp[:, 2:h + 2, 2:w + 2]
    return out


@pytest.mark.parametrize("seed", range(12))
def test_diffusion_f64_matches_restatement(seed):
    rng = np.random.default_rng(seed)
    v = O.read_params_json(golden("default-parameters.json"))
    H, W = int(rng.integers(8, 24)), int(rng.integers(8, 24))
    p = O.make_params(v, H, W)
    st = np.zeros((H, W), np.uint8)
    st[0, :] = st[-1, :] = st[:, 0] = st[:, -1] = O.MOCKED
    for _ in range(int(rng.integers(1, 12))):
        st[int(rng.integers(1, H - 1)), int(rng.integers(1, W - 1))] = O.OCCUPIED
    roi = (C.c_int32 * 5)()
    lib.eo_find_roi(H, W, O.ptr(st), roi)
    sub = np.stack([v[19 + s] * (1 + 0.2 * rng.random((H, W))) for s in range(3)])
    got = _diffuse64(p, sub, list(roi))
    want = _diffuse_restated(p, sub, list(roi))
    assert np.allclose(got, want, rtol=1e-12, atol=1e-12)


def test_diffusion_f32_within_1e5_of_f64():
    """FP32 diffusion of fixed states against the FP64 oracle: max relative
    error (normalised by the substrate's external level) below 1e-5."""
    v = O.read_params_json(golden("default-parameters.json"))
    worst = 0.0
    for t in range(1, 11):
        s = O.read_state_text(golden("tumor-lib", f"tumor-{t}.txt"), v)
        H, W = s["st"].shape
        p = O.make_params(v, H, W)
        rng = np.random.default_rng(t)
        sub = s["sub"] * (1 + 0.3 * rng.random(s["sub"].shape)).astype(np.float32)
        roi = (C.c_int32 * 5)()
        lib.eo_find_roi(H, W, O.ptr(s["st"]), roi)
        a = _diffuse32(p, sub, list(roi)).astype(np.float64)
        b = _diffuse64(p, sub.astype(np.float64), list(roi))
        for k in range(3):
            worst = max(worst, np.max(np.abs(a[k] - b[k])) / abs(float(v[19 + k])))
    assert worst < 1e-5, worst


def test_count_living():
    st = np.random.default_rng(0).integers(0, 3, size=(20, 30)).astype(np.uint8)
    assert lib.eo_count_living(20, 30, O.ptr(st)) == int((st == 1).sum())


def test_oracle_run_is_deterministic_and_threads_agree():
    v = O.read_params_json(golden("default-parameters.json"))
    s = O.read_state_text(golden("tumor-lib", "tumor-3.txt"), v)
    prot = np.zeros((1, 480), np.float32)
    prot[0, 0] = 2.0
    runs = []
    for nt in (1, 3):
        b = O.OracleBatch(v, [s] * 3, prot, [0] * 3, [0, 1, 2], 300, 144000)
        b.run(700, seed=9, nthreads=nt)
        runs.append((b.counts(), b.sub.copy(), b.st.copy()))
    assert (runs[0][0] == runs[1][0]).all()
    assert np.array_equal(runs[0][1], runs[1][1]) and np.array_equal(runs[0][2], runs[1][2])


def test_oracle_instrumentation_counters_are_consistent():
    """The oracle's instrumentation (common/instr.h:5-21): counting leaves the
    simulation unchanged; kills <= decisions; each histogram holds one entry
    per decision; the mean death probability lies in [0, 1]; no dose, no
    irradiation events; divisions equal the list growth between rebuilds."""
    import oracle as O
    from conftest import golden
    v = O.read_params_json(golden("default-parameters.json"))
    s = O.read_state_text(golden("tumor-lib", "tumor-3.txt"), v)
    n0 = int((s["st"] == O.OCCUPIED).sum())
    doses = np.zeros((2, 480), np.float32)
    doses[1, [0, 2]] = [3.0, 2.0]
    a = O.OracleBatch(v, [s, s], doses, [0, 1], [0, 1], 300, 144000)
    b = O.OracleBatch(v, [s, s], doses, [0, 1], [0, 1], 300, 144000)
    a.run(127, seed=9, instr=True)
    b.run(127, seed=9)
    np.testing.assert_array_equal(a.st, b.st)
    np.testing.assert_array_equal(a.sub, b.sub)
    r0, r1 = a.instr[0], a.instr[1]
    assert r0[2] == 0 and r0[0] == 0          # no dose: no irradiation, no repair
    assert int(r0[3]) == int(a.nlist[0]) - n0   # one list append per division
    assert 0 < r1[2] <= 2 * n0
    a.run(5000, seed=9, instr=True)   # repair delay 3.34 exp(0.149 R) h
    r1 = a.instr[1]
    assert 0 < r1[1] <= r1[0]
    assert r1[5:305].sum() == r1[0] == r1[305:].sum()
    assert 0.0 <= r1[4] / 2 ** 32 / r1[0] <= 1.0
