"""Seeded synthetic tumours: an occupied disk on a bordered lattice.

Cells get phases uniform over G1/S/G2/M, cycle ends drawn like createCell
(simulation/cell-division.h:10-19: cumulative normal phase lengths from the
parameter set), a proliferation clock uniform inside the current phase, no
irradiation, and every site holds the external substrate levels.
"""
import numpy as np

G1, S, G2, M, D = range(5)


def disk_arrays(values, H, W, radius, seed, center=None):
    """Per-site arrays of a synthetic tumour.

    values: flat parameter vector (40 floats, include/emt6ro_amd.h order).
    Returns a dict of (H, W) arrays keyed by field name (see
    emt6ro._native.FIELDS)."""
    v = np.asarray(values, dtype=np.float32)
    ext = v[19:22]
    mean = v[25:35:2].astype(np.float64)
    sd = v[26:35:2].astype(np.float64)
    rng = np.random.Generator(np.random.PCG64(int(seed)))
    rc, cc = center if center is not None else (H // 2, W // 2)
    rr, cc_ = np.mgrid[0:H, 0:W]
    inside = ((rr - rc) ** 2 + (cc_ - cc) ** 2 <= radius * radius)
    inside &= (rr >= 1) & (rr <= H - 2) & (cc_ >= 1) & (cc_ <= W - 2)
    n = int(inside.sum())
    f = {k: np.zeros((H, W), np.float32) for k in (
        "time_in_repair", "irradiation", "repair_delay", "proliferation_time",
        "cycle_g1", "cycle_s", "cycle_g2", "cycle_m", "cycle_d")}
    for k, name in enumerate(("cho", "ox", "gi")):
        f[name] = np.full((H, W), ext[k], np.float32)
    occ = np.zeros((H, W), np.uint8)
    occ[0, :] = occ[-1, :] = occ[:, 0] = occ[:, -1] = 2
    occ[inside] = 1
    lengths = rng.normal(mean, sd, size=(n, 5))
    ends = np.cumsum(lengths.astype(np.float32), axis=1, dtype=np.float32)
    phase = rng.integers(0, 4, size=n)
    start = np.where(phase == 0, 0.0, ends[np.arange(n), np.maximum(phase - 1, 0)])
    stop = ends[np.arange(n), phase]
    prolif = start + rng.random(n) * np.maximum(stop - start, 0.0)
    names = ("cycle_g1", "cycle_s", "cycle_g2", "cycle_m", "cycle_d")
    for k in range(5):
        f[names[k]][inside] = ends[:, k]
    f["proliferation_time"][inside] = prolif.astype(np.float32)
    f["occupancy"] = occ
    f["mode"] = np.zeros((H, W), np.uint8)
    ph = np.zeros((H, W), np.uint8)
    ph[inside] = phase.astype(np.uint8)
    f["phase"] = ph
    return f


def disk_state(params, H, W, radius, seed, center=None):
    """A TumorState holding disk_arrays(...)."""
    from .simulation import TumorState
    arrays = disk_arrays(params.values(), H, W, radius, seed, center)
    st = TumorState.empty(params, H, W)
    for name, arr in arrays.items():
        st.set_field(name, arr)
    return st, arrays
