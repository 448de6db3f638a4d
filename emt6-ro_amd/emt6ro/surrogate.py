"""Learned fitness surrogate: an MLP from (dose-time vector, tumour descriptor)
to the mean final living-cell count, trained and evaluated on the GPU by the
library's BF16 matrix-core kernels (include/emt6ro_amd.h, emt6ro_mlp_*).

The reference has no surrogate of its own that runs offline: its GA example
loads an external pretrained forecasting model (python_ga/example_run_model.py:
24-184).  Here the simulator makes the training data and stays the judge: the
GA screens candidates with the surrogate and verifies the best of them by
simulation (emt6ro.ga).

Input features (152 by default): doses in Gy on 144 half-hour slots (the
first 72 h), then 8 descriptors of the initial tumour: living cells, radius,
fractions of living cells in G1, S, G2 and M (D counted with M), and mean
oxygen and glucose over the living cells.
"""
import ctypes as C

import numpy as np

from ._native import lib, check_mlp

N_DOSE_SLOTS = 144
N_DESCRIPTORS = 8
N_FEATURES = N_DOSE_SLOTS + N_DESCRIPTORS


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


def features(dose_rows, descriptors, index_dose, index_tumour):
    """Rows [dose_rows[index_dose[i]] (first 144 slots), descriptors[index_tumour[i]]]."""
    d = np.asarray(dose_rows, np.float32)
    if d.shape[1] < N_DOSE_SLOTS:
        d = np.pad(d, ((0, 0), (0, N_DOSE_SLOTS - d.shape[1])))
    d = d[:, :N_DOSE_SLOTS]
    t = np.asarray(descriptors, np.float32).reshape(-1, N_DESCRIPTORS)
    return np.ascontiguousarray(
        np.concatenate([d[np.asarray(index_dose)], t[np.asarray(index_tumour)]], axis=1))


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
        g = np.zeros(self.num_params, np.float32)
        loss = C.c_float()
        check_mlp(lib.emt6ro_mlp_grads(self._h, _fp(x), _fp(t), x.shape[0], _fp(g),
                                       C.byref(loss)))
        return float(loss.value), g

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.emt6ro_mlp_destroy(self._h)
            self._h = C.c_void_p()


class Surrogate:
    """The MLP with per-feature input standardisation and a standardised
    target, fitted on (features, mean final living count) pairs."""

    def __init__(self, in_features=N_FEATURES, hidden=512, seed=0, gpu_id=0):
        self.net = MLP(in_features, hidden, seed, gpu_id)
        self.x_mu = np.zeros(in_features, np.float32)
        self.x_sd = np.ones(in_features, np.float32)
        self.y_mu, self.y_sd = 0.0, 1.0

    def _nx(self, X):
        return (np.asarray(X, np.float32) - self.x_mu) / self.x_sd

    def fit(self, X, y, steps=2000, batch=4096, lr=1e-3, seed=0):
        X = np.asarray(X, np.float32)
        y = np.asarray(y, np.float64)
        self.x_mu = X.mean(axis=0).astype(np.float32)
        sd = X.std(axis=0)
        self.x_sd = np.where(sd > 1e-6, sd, 1.0).astype(np.float32)
        self.y_mu, self.y_sd = float(y.mean()), float(max(y.std(), 1e-6))
        self._base = (X, y)
        self.net.set_data(self._nx(X), ((y - self.y_mu) / self.y_sd).astype(np.float32))
        return self.net.train(steps, batch=min(batch, len(y)), lr=lr, seed=seed)

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
        self.net.set_data(self._nx(Xa), ((ya - self.y_mu) / self.y_sd).astype(np.float32))
        self._refits = getattr(self, "_refits", 0) + 1
        return self.net.train(steps, batch=min(batch, len(ya)), lr=lr, seed=seed + self._refits)

    def predict(self, X):
        return self.net.predict(self._nx(X)).astype(np.float64) * self.y_sd + self.y_mu
