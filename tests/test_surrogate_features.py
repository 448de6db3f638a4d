"""Protocol features of the fitness surrogate (emt6ro.surrogate), on hand-worked
dose protocols.  The damage history follows the repair rules of
cell.cu:84-100 and cell.h:104-133 with the timer started at each dose; the
expected values below are worked out by hand from those rules."""
import math

import numpy as np
import pytest

from conftest import golden

S = pytest.importorskip("emt6ro.surrogate")

DC, DE, SC, SE, HALF = S.DEFAULT_REPAIR


def delay(R):
    return DC * math.exp(DE * R)


def protocol(doses):
    """144 slots; doses = {hour: Gy} on the half-hour grid."""
    g = np.zeros(S.N_DOSE_SLOTS)
    for h, d in doses.items():
        g[int(round(2 * h))] = d
    return g


def test_repair_constants_come_from_the_parameter_vector(default_params_path):
    sim = pytest.importorskip("emt6ro.simulation")
    p = sim.load_parameters(default_params_path)
    got = S.repair_constants(p)
    assert got == pytest.approx(S.DEFAULT_REPAIR, rel=1e-6)
    assert S.repair_constants(None) == S.DEFAULT_REPAIR
    v = p.values().copy()
    v[39] = 2.0
    assert S.repair_constants(v)[4] == 2.0


def test_no_dose():
    f = S.protocol_features(np.zeros((1, S.N_DOSE_SLOTS)))
    assert f.shape == (1, S.N_PROTOCOL_FEATURES) and not f.any()


def test_single_dose():
    f = S.protocol_features(protocol({2.0: 2.0}))[0]
    end = 2.0 + delay(2.0)                      # 6.503 h
    assert f[0] == 2.0 and f[1] == 1 and f[2] == 2.0 and f[3] == 2.0 and f[4] == 2.0
    assert list(f[5:11]) == [2.0, 0, 0, 0, 0, 0]
    assert f[11] == pytest.approx(SE * 2.0)     # log-survival, surv_coeff 1
    assert f[12] == pytest.approx((1 - math.exp(SE * 2.0)) * (144.0 - end))
    assert f[13] == 1 and f[14] == 2.0
    assert f[15] == pytest.approx(delay(2.0))
    # Snapshots at 6, 12, ..., 72 h: the repair ends at 6.5 h.
    assert f[16] == 0.0 and f[17:28] == pytest.approx([SE * 2.0] * 11)


def test_dose_during_repair_accumulates_with_decay():
    f = S.protocol_features(protocol({0.0: 1.0, 1.0: 1.0}))[0]
    assert delay(1.0) > 1.0                     # still repairing at the second dose
    R = 1.0 / (1.0 + 1.0 / HALF) + 1.0          # 1.444 Gy
    assert f[14] == pytest.approx(R)
    assert f[13] == 1
    assert f[11] == pytest.approx(SE * R)
    assert f[15] == pytest.approx(1.0 + delay(R))   # in repair from 0 h to 1 h + delay(R)
    assert f[12] == pytest.approx((1 - math.exp(SE * R)) * (144.0 - (1.0 + delay(R))))


def test_separate_repairs_multiply_survival():
    f = S.protocol_features(protocol({0.0: 1.0, 10.0: 1.0}))[0]
    assert delay(1.0) < 10.0
    assert f[13] == 2 and f[14] == 1.0
    assert f[11] == pytest.approx(2 * SE)
    assert f[15] == pytest.approx(2 * delay(1.0))
    assert f[16] == pytest.approx(SE)           # first repair ends at 3.9 h, before 6 h
    assert f[17] == pytest.approx(SE)           # second one ends at 13.9 h, after 12 h
    assert f[18] == pytest.approx(2 * SE)
    assert f[2] == pytest.approx(5.0) and f[3] == 0.0 and f[4] == 10.0


def test_rows_and_row_mapping():
    g = np.stack([protocol({0.0: 1.0}), protocol({3.0: 2.5}), protocol({})])
    desc = np.arange(3 * S.N_DESCRIPTORS, dtype=np.float32).reshape(3, S.N_DESCRIPTORS)
    pi = np.array([2, 0, 1, 1])
    ti = np.array([0, 2, 1, 0])
    X = S.features(g, desc, pi, ti)
    assert X.shape == (4, S.N_FEATURES) and X.dtype == np.float32
    pf = S.protocol_features(g).astype(np.float32)
    for r in range(4):
        assert np.array_equal(X[r, :S.N_DESCRIPTORS], desc[ti[r]])
        assert np.array_equal(X[r, S.N_DESCRIPTORS:], pf[pi[r]])
    # Genomes shorter than 144 slots are zero-padded; longer ones cut.
    short = S.protocol_features(g[:, :100])
    assert np.array_equal(short, S.protocol_features(np.pad(g[:, :100], ((0, 0), (0, 44)))))
    longer = np.pad(g, ((0, 0), (0, 10)))
    longer[:, -1] = 3.0
    assert np.array_equal(S.protocol_features(longer), S.protocol_features(g))


def waits_block(f, w):
    """The five summary columns of the history with restart wait w."""
    k = 28 + 5 * S.REPAIR_WAITS.index(w)
    return f[k:k + 5]


def test_waits_delay_the_repair_and_keep_doses_undecayed():
    assert S.N_PROTOCOL_FEATURES == 28 + 5 * len(S.REPAIR_WAITS)
    f = S.protocol_features(protocol({2.0: 2.0}))[0]
    log_s, kill_w, n_rep, max_r, t_rep = waits_block(f, 4.0)
    end = 2.0 + 4.0 + delay(2.0)
    assert log_s == pytest.approx(SE * 2.0) and n_rep == 1 and max_r == 2.0
    assert kill_w == pytest.approx((1 - math.exp(SE * 2.0)) * (144.0 - end))
    assert t_rep == pytest.approx(4.0 + delay(2.0))
    # A second dose while the timer waits adds in full: R = 1 + 1.
    f = S.protocol_features(protocol({0.0: 1.0, 1.0: 1.0}))[0]
    for w in (1.0, 2.0):
        log_s, kill_w, n_rep, max_r, t_rep = waits_block(f, w)
        assert max_r == 2.0 and n_rep == 1
        assert t_rep == pytest.approx(1.0 + w + delay(2.0))
    # Once the timer runs, the earlier damage decays: 1 h wait, dose at 2 h.
    f = S.protocol_features(protocol({0.0: 1.0, 2.0: 1.0}))[0]
    R = 1.0 / (1.0 + 1.0 / HALF) + 1.0
    assert waits_block(f, 1.0)[3] == pytest.approx(R)
    assert waits_block(f, 2.0)[3] == 2.0
