"""Learned fitness surrogate: an MLP from (dose protocol, tumour descriptor)
to the mean final living-cell count, trained and evaluated on the GPU by the
library's BF16 matrix-core kernels (include/emt6ro_amd.h, emt6ro_mlp_*).

The reference has no surrogate of its own that runs offline: its GA example
loads an external pretrained forecasting model (python_ga/example_run_model.py:
24-184).  Here the simulator makes the training data and has the last word: the
GA screens candidates with the surrogate and verifies the best of them by
simulation (emt6ro.ga).

Input features (71): 8 descriptors of the initial tumour (living cells,
radius, fractions of living cells in G1, S, G2 and M (D counted with M),
mean oxygen and glucose over the living cells), then 63 features of the
dose protocol (doses in Gy on 144 half-hour slots, the first 72 h):

- total dose, number of doses, dose-weighted mean time, first and last dose
  time, and the dose in each 12 h window (11 features);
- the radiation-damage history of one cell under the model's repair rules
  (cell.cu:84-100, cell.h:104-133), in which a dose resets the repair timer
  and the timer restarts only at the cell's next cycle change: with the
  timer restarting at once, log-survival over all repair completions, kill
  probability weighted by the time left to 144 h, number of repair
  completions, largest accumulated dose, total time in repair, and the
  log-survival at every 6 h up to 72 h (17 features); then the first five
  of these for restart waits of 1, 2, 3, 4, 6, 8 and 12 h (35 features).

Raw per-slot doses are not inputs: with about a thousand training
protocols they let the network fit noise (held-out error 27.6 vs 18.5
counts in a CPU comparison on the same data), and the waits matter (16.4
vs 11.3 counts with 3,600 protocols).
"""
import ctypes as C

import numpy as np

from ._native import lib, check_mlp

N_DOSE_SLOTS = 144
N_DESCRIPTORS = 8
# Waits (h) from a dose to the start of the repair timer in the extra damage
# histories.
REPAIR_WAITS = (1.0, 2.0, 3.0, 4.0, 6.0, 8.0, 12.0)
N_PROTOCOL_FEATURES = 11 + 17 + 5 * len(REPAIR_WAITS)
N_FEATURES = N_DESCRIPTORS + N_PROTOCOL_FEATURES
SLOT_HOURS = 0.5
HORIZON_HOURS = 144.0

# Repair constants in the flat parameter vector of include/emt6ro_amd.h
# (delay coefficient and exponent, survival coefficient and exponent, repair
# half time), and their values in data/default-parameters.json.
_REPAIR_INDEX = (35, 36, 37, 38, 39)
DEFAULT_REPAIR = (3.3414, 0.1492, 1.0, -0.4993, 0.8)


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def tumour_descriptor(state):
    """The 8 descriptors of a TumorState (float32)."""
    occ = state.occupancy() == 1
    n = int(occ.sum())
    out = np.zeros(N_DESCRIPTORS, np.float32)
    out[0] = n
    if n == 0:
        return out
    rr, cc = np.nonzero(occ)
    r0, c0 = rr.mean(), cc.mean()
    out[1] = np.sqrt(((rr - r0) ** 2 + (cc - c0) ** 2).max())
    ph = state.phase()[occ]
    for k in range(4):
        out[2 + k] = np.count_nonzero(ph == k) / n
    out[5] += np.count_nonzero(ph == 4) / n
    out[6] = state.ox()[occ].mean()
    out[7] = state.cho()[occ].mean()
    return out


def repair_constants(params=None):
    """(delay coeff, delay exp, survival coeff, survival exp, repair half
    time) of a Parameters object or flat parameter vector; the defaults
    without one."""
    if params is None:
        return DEFAULT_REPAIR
    v = params.values() if hasattr(params, "values") and callable(params.values) else params
    v = np.asarray(v, np.float64)
    return tuple(float(v[i]) for i in _REPAIR_INDEX)


def _damage_history(G, wait, constants, snapshots):
    """Damage history of one cell for protocols G (P, 144 slots): columns
    log-survival, time-weighted kill probability, repair completions,
    largest accumulated dose, time in repair, then (with `snapshots`) the
    log-survival at every 6 h up to 72 h.

    A dose d at time t gives R = R / (1 + t_rep / half_time) + d, where
    t_rep is the time the repair timer has run since the previous dose, and
    resets the timer, which restarts `wait` hours later (in the model, at
    the cell's next cycle change); repair ends when the timer reaches
    coeff * exp(exp_coeff * R), and the cell then dies with probability
    1 - surv_coeff * exp(surv_exp * R) or starts clean (R = 0)."""
    dc, de, sc, se, half = constants
    P = G.shape[0]
    t = np.arange(N_DOSE_SLOTS) * SLOT_HOURS
    R = np.zeros(P)
    t_last = np.zeros(P)      # last dose
    ep0 = np.zeros(P)         # first dose of the current repair episode
    rep = np.zeros(P, bool)
    log_s = np.zeros(P)
    kill_w = np.zeros(P)
    n_rep = np.zeros(P)
    max_r = np.zeros(P)
    t_rep = np.zeros(P)
    snaps = []

    def complete(before):
        nonlocal R, rep, log_s, kill_w, n_rep, t_rep
        end = t_last + wait + dc * np.exp(de * R)
        done = rep & (end <= before)
        if done.any():
            surv = np.clip(sc * np.exp(se * R), 1e-12, 1.0)
            log_s = np.where(done, log_s + np.log(surv), log_s)
            kill_w = np.where(done, kill_w + (1.0 - surv) * np.maximum(HORIZON_HOURS - end, 0.0),
                              kill_w)
            n_rep = n_rep + done
            t_rep = np.where(done, t_rep + (end - ep0), t_rep)
            R = np.where(done, 0.0, R)
            rep = rep & ~done

    for s in range(N_DOSE_SLOTS):
        ts = t[s]
        complete(ts)
        d = G[:, s]
        hit = d > 0
        ran = np.maximum(ts - (t_last + wait), 0.0)
        R = np.where(hit, np.where(rep, R / (1.0 + ran / half), 0.0) + d, R)
        ep0 = np.where(hit & ~rep, ts, ep0)
        t_last = np.where(hit, ts, t_last)
        rep = rep | hit
        max_r = np.maximum(max_r, R)
        if snapshots and (s + 1) % 12 == 0:
            complete(ts + SLOT_HOURS)
            snaps.append(log_s.copy())
    complete(np.inf)
    return np.stack([log_s, kill_w, n_rep, max_r, t_rep] + snaps, axis=1)


def protocol_features(dose_rows, params=None):
    """(P, N_PROTOCOL_FEATURES) features of P dose protocols (see the
    module docstring)."""
    constants = repair_constants(params)
    G = np.asarray(dose_rows, np.float64)
    if G.ndim == 1:
        G = G[None]
    if G.shape[1] < N_DOSE_SLOTS:
        G = np.pad(G, ((0, 0), (0, N_DOSE_SLOTS - G.shape[1])))
    G = G[:, :N_DOSE_SLOTS]
    P = G.shape[0]
    t = np.arange(N_DOSE_SLOTS) * SLOT_HOURS
    given = G > 0
    tot = G.sum(1)
    stats = np.zeros((P, 11))
    stats[:, 0] = tot
    stats[:, 1] = given.sum(1)
    stats[:, 2] = (G * t).sum(1) / np.maximum(tot, 1e-12)
    any_dose = given.any(1)
    stats[:, 3] = np.where(any_dose, t[np.argmax(given, 1)], 0.0)
    stats[:, 4] = np.where(any_dose, t[N_DOSE_SLOTS - 1 - np.argmax(given[:, ::-1], 1)], 0.0)
    stats[:, 5:11] = G.reshape(P, 6, 24).sum(2)
    parts = [stats, _damage_history(G, 0.0, constants, True)]
    parts += [_damage_history(G, w, constants, False) for w in REPAIR_WAITS]
    return np.concatenate(parts, axis=1)


def features(dose_rows, descriptors, index_dose, index_tumour, params=None):
    """Rows [descriptors[index_tumour[i]], protocol_features(dose_rows)[index_dose[i]]]."""
    f = protocol_features(dose_rows, params).astype(np.float32)
    t = np.asarray(descriptors, np.float32).reshape(-1, N_DESCRIPTORS)
    return np.ascontiguousarray(
        np.concatenate([t[np.asarray(index_tumour)], f[np.asarray(index_dose)]], axis=1))


class MLP:
    """in -> hidden -> hidden -> hidden -> 1 with GELU (tanh form).

    BF16 weights and activations on the matrix cores, FP32 accumulation, FP32
    master weights and Adam (beta 0.9 / 0.999, eps 1e-8).  Parameters are
    flat float32 in the order W1 [hidden, in], b1, W2, b2, W3, b3, w4 [hidden],
    b4 [1]; initialisation U(-1/sqrt(fan_in), 1/sqrt(fan_in)) from `seed`."""

    def __init__(self, in_features=N_FEATURES, hidden=512, seed=0, gpu_id=0):
        h = C.c_void_p()
        check_mlp(lib.emt6ro_mlp_create(int(in_features), int(hidden), int(seed), int(gpu_id),
                                        C.byref(h)))
        self._h = h
        self.in_features, self.hidden = int(in_features), int(hidden)

    @property
    def num_params(self):
        return int(lib.emt6ro_mlp_num_params(self._h))

    def get_params(self):
        out = np.zeros(self.num_params, np.float32)
        check_mlp(lib.emt6ro_mlp_get_params(self._h, _fp(out)))
        return out

    def set_params(self, flat):
        """Replace the parameters (and zero the Adam moments)."""
        p = np.ascontiguousarray(flat, np.float32)
        if p.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {p.size}")
        check_mlp(lib.emt6ro_mlp_set_params(self._h, _fp(p)))

    def _x(self, X):
        x = np.ascontiguousarray(X, np.float32)
        if x.ndim != 2 or x.shape[1] != self.in_features:
            raise ValueError(f"X must be (n, {self.in_features})")
        return x

    def set_data(self, X, y):
        """Training set, held on the device in BF16 (targets FP32)."""
        x = self._x(X)
        t = np.ascontiguousarray(y, np.float32).reshape(-1)
        if t.size != x.shape[0]:
            raise ValueError("X and y differ in length")
        check_mlp(lib.emt6ro_mlp_set_data(self._h, _fp(x), _fp(t), x.shape[0]))

    def train(self, steps, batch=4096, lr=1e-3, seed=0):
        """`steps` Adam steps on minibatches drawn on the device; returns each
        step's mean squared error."""
        losses = np.zeros(max(int(steps), 1), np.float32)
        check_mlp(lib.emt6ro_mlp_train(self._h, int(steps), int(batch), float(lr), int(seed),
                                       _fp(losses)))
        return losses[:int(steps)]

    def set_train_graph(self, steps_per_graph):
        """Steps per replayed HIP graph in train() (0: launch each kernel
        directly); the results do not change."""
        check_mlp(lib.emt6ro_mlp_set_train_graph(self._h, int(steps_per_graph)))

    def predict(self, X):
        x = self._x(X)
        y = np.zeros(x.shape[0], np.float32)
        if x.shape[0]:
            check_mlp(lib.emt6ro_mlp_predict(self._h, _fp(x), x.shape[0], _fp(y)))
        return y

    def grads(self, X, y):
        """(loss, gradient) of loss = sum((f(X) - y)^2) / (2 n), no update."""
        x = self._x(X)
        t = np.ascontiguousarray(y, np.float32).reshape(-1)
        if t.size != x.shape[0]:
            raise ValueError("X and y differ in length")
        g = np.zeros(self.num_params, np.float32)
        loss = C.c_float()
        check_mlp(lib.emt6ro_mlp_grads(self._h, _fp(x), _fp(t), x.shape[0], _fp(g),
                                       C.byref(loss)))
        return float(loss.value), g

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.emt6ro_mlp_destroy(self._h)
            self._h = C.c_void_p()


def train_many(nets, steps, batch=4096, lr=1e-3, seeds=None):
    """`steps` Adam steps of each network in `nets` (distinct MLPs on one
    GPU), interleaved so that the GPU overlaps them; network i draws its
    minibatches with seeds[i].  Each ends as net.train(steps, batch, lr,
    seeds[i]) would leave it.  Returns the (len(nets), steps) losses."""
    nets = list(nets)
    seeds = np.ascontiguousarray(seeds if seeds is not None else np.arange(len(nets)), np.uint64)
    if len(seeds) != len(nets):
        raise ValueError("one seed per network")
    losses = np.zeros((len(nets), max(int(steps), 1)), np.float32)
    handles = (C.c_void_p * len(nets))(*[n._h.value for n in nets])
    check_mlp(lib.emt6ro_mlp_train_many(C.cast(handles, C.c_void_p), len(nets), int(steps),
                                        int(batch), float(lr),
                                        seeds.ctypes.data_as(C.c_void_p), _fp(losses)))
    return losses[:, :int(steps)]


class Surrogate:
    """An ensemble of `members` MLPs (different initialisations and
    minibatch streams) with per-feature input standardisation and a
    standardised target, fitted on (features, mean final living count)
    pairs; predictions are the members' mean."""

    def __init__(self, in_features=N_FEATURES, hidden=512, seed=0, gpu_id=0, members=1,
                 params=None):
        if members < 1:
            raise ValueError("members must be >= 1")
        self.params = params   # repair constants for features(); None: the defaults
        self.nets = [MLP(in_features, hidden, seed + 7919 * m, gpu_id) for m in range(members)]
        self.x_mu = np.zeros(in_features, np.float32)
        self.x_sd = np.ones(in_features, np.float32)
        self.y_mu, self.y_sd = 0.0, 1.0

    @property
    def net(self):
        return self.nets[0]

    def features(self, dose_rows, descriptors, index_dose, index_tumour):
        """The input rows of this surrogate (module-level `features` with its
        parameters)."""
        return features(dose_rows, descriptors, index_dose, index_tumour, self.params)

    def _nx(self, X):
        return (np.asarray(X, np.float32) - self.x_mu) / self.x_sd

    def _train_all(self, Xn, yn, steps, batch, lr, seed):
        # The members train together (train_many), each on its own minibatch
        # stream; the result is the same as training them one by one.
        for net in self.nets:
            net.set_data(Xn, yn)
        seeds = [seed + 104729 * m for m in range(len(self.nets))]
        return train_many(self.nets, steps, batch=batch, lr=lr, seeds=seeds).mean(axis=0)

    def fit(self, X, y, steps=300, batch=2048, lr=1e-3, seed=0):
        """Returns the members' mean training loss per step."""
        X = np.asarray(X, np.float32)
        y = np.asarray(y, np.float64)
        self.x_mu = X.mean(axis=0).astype(np.float32)
        sd = X.std(axis=0)
        self.x_sd = np.where(sd > 1e-6, sd, 1.0).astype(np.float32)
        self.y_mu, self.y_sd = float(y.mean()), float(max(y.std(), 1e-6))
        self._base = (X, y)
        return self._train_all(self._nx(X), ((y - self.y_mu) / self.y_sd).astype(np.float32),
                               steps, min(batch, len(y)), lr, seed)

    def refit(self, X, y, steps, batch=1024, lr=3e-4, seed=1):
        """Fine-tune on new samples mixed with the original training set (the
        new rows repeated to make up at least a quarter of the mix); the
        standardisation of the first fit is kept."""
        X = np.asarray(X, np.float32)
        y = np.asarray(y, np.float64)
        bx, by = self._base
        rep = max(1, int(np.ceil(len(by) / (3.0 * len(y)))))
        Xa = np.concatenate([bx, np.tile(X, (rep, 1))])
        ya = np.concatenate([by, np.tile(y, rep)])
        self._refits = getattr(self, "_refits", 0) + 1
        return self._train_all(self._nx(Xa), ((ya - self.y_mu) / self.y_sd).astype(np.float32),
                               steps, min(batch, len(ya)), lr, seed + self._refits)

    def predict(self, X):
        xn = self._nx(X)
        out = np.mean([net.predict(xn).astype(np.float64) for net in self.nets], axis=0)
        return out * self.y_sd + self.y_mu
