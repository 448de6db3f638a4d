"""Batched simulations of uneven lattice sizes and horizons in one launch
(include/emt6ro_amd.h, emt6ro_batch_*).  The reference runs one lattice size
and one horizon per Simulation (simulation/simulation.cu:123-137)."""
import ctypes as C

import numpy as np

from ._native import lib, check, PATH_PERSISTENT, PATH_STEP, PATH_STEP_GRAPH
from .simulation import TumorState

PATHS = {"persistent": PATH_PERSISTENT, "step": PATH_STEP, "graph": PATH_STEP_GRAPH}


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Batch:
    """Simulation i runs tumors[tumor_index[i]] under doses[prot_index[i]]
    (a row of slot doses; a dose fires at steps s < prot_length with
    s % prot_res == 0, protocol.h:9-14) until step horizon[i]."""

    def __init__(self, params, tumors, tumor_index, prot_index, horizon, doses, prot_res,
                 prot_length, sim_id=None, gpu_id=0, seed=0):
        self._params, self._tumors = params, list(tumors)
        ti = np.ascontiguousarray(tumor_index, np.int32)
        pi = np.ascontiguousarray(prot_index, np.int32)
        hz = np.ascontiguousarray(horizon, np.uint32)
        n = len(ti)
        sid = np.ascontiguousarray(np.arange(n) if sim_id is None else sim_id, np.uint32)
        d = np.ascontiguousarray(doses, np.float32)
        if d.ndim != 2:
            raise ValueError("doses must be (n_protocols, n_slots)")
        self._dose_shape = d.shape
        if not (ti.ndim == pi.ndim == hz.ndim == sid.ndim == 1) or \
                not (len(pi) == len(hz) == len(sid) == n):
            raise ValueError("tumor_index, prot_index, horizon and sim_id must be 1-D arrays "
                             "of one length")
        arr = (C.c_void_p * len(self._tumors))(*[t._h.value for t in self._tumors])
        h = C.c_void_p()
        check(lib.emt6ro_batch_create(params._h, arr, len(self._tumors), n,
                                      _p(ti, C.c_int32), _p(pi, C.c_int32), _p(hz, C.c_uint32),
                                      _p(sid, C.c_uint32), d.shape[0], d.shape[1],
                                      _p(d, C.c_float), int(prot_res), int(prot_length),
                                      int(gpu_id), int(seed), C.byref(h)))
        self._h = h
        self.n = n
        self.window = -1

    def run(self, nsteps, path="persistent", variant=-1, steps_per_launch=0):
        """Advance every simulation nsteps (capped by its horizon); returns the
        device time of the run in ms (HIP events on the batch's stream)."""
        check(lib.emt6ro_batch_run(self._h, int(nsteps), PATHS[path], int(variant),
                                   int(steps_per_launch)))
        return self.last_run()[0]

    def last_run(self):
        ms, launches = C.c_double(), C.c_int64()
        check(lib.emt6ro_batch_last_run(self._h, C.byref(ms), C.byref(launches)))
        return ms.value, launches.value

    def results(self):
        out = np.zeros(self.n, np.uint32)
        check(lib.emt6ro_batch_results(self._h, _p(out, C.c_uint32)))
        return out

    def counters(self):
        out = np.zeros((self.n, 5), np.uint64)
        check(lib.emt6ro_batch_counters(self._h, _p(out, C.c_uint64)))
        return {"step": out[:, 0], "n_list": out[:, 1], "divisions": out[:, 2],
                "status": out[:, 3], "site_substeps": out[:, 4]}

    def state(self, i):
        h = C.c_void_p()
        check(lib.emt6ro_batch_state(self._h, int(i), C.byref(h)))
        return TumorState(h.value)

    def reset(self, epoch=0):
        check(lib.emt6ro_batch_reset(self._h, int(epoch)))

    def set_variant(self, rs, variant):
        """Kernel variant for the persistent launches of lattices needing LDS
        row stride rs (36, 56 or 88); -1 restores the default."""
        check(lib.emt6ro_batch_set_variant(self._h, int(rs), int(variant)))

    def set_window(self, rows=-1):
        """Window pass for lattices wider than 34 sites: simulations whose ROI
        plus two rings fits `rows` x 36 padded sites (and, on lattices wider
        than 54 sites with rows=-1, a 56-column window) run first in a kernel
        holding only that window in LDS.  -1: automatic, 0: off.  Results do
        not depend on it.  The setting is kept in `self.window`."""
        check(lib.emt6ro_batch_set_window(self._h, int(rows)))
        self.window = int(rows)

    def set_doses(self, doses):
        """Replace the protocol table; same shape as the one the batch was
        made with."""
        d = np.ascontiguousarray(doses, np.float32)
        if d.shape != self._dose_shape:
            raise ValueError(f"doses must be {self._dose_shape}, got {d.shape}")
        check(lib.emt6ro_batch_set_doses(self._h, _p(d, C.c_float)))

    def __del__(self):
        if getattr(self, "_h", None) and self._h.value:
            lib.emt6ro_batch_destroy(self._h)
            self._h = C.c_void_p()


def variants():
    """Compiled kernel shapes: dicts of id, kind (0 banded in place, 1
    ping-pong site-parallel, 2 banded with the band height chosen per ROI, 3
    as 2 with horizontal neighbours from adjacent lanes), nt (workgroup
    size), l (band rows),
    kmax (items or Jacobi sites per thread), kf (final sites per thread),
    ps (ring slots per thread), rs (LDS row stride)."""
    n = lib.emt6ro_variant_info(-1, None)
    out = []
    for i in range(n):
        info = (C.c_int32 * 8)()
        check(lib.emt6ro_variant_info(i, info))
        out.append(dict(zip(("id", "kind", "nt", "l", "kmax", "kf", "ps", "rs"), list(info))))
    return out


def fits(v, H, W):
    """Whether variant v (a dict from variants()) takes bordered H x W lattices."""
    if H + 2 > v["rs"] or W + 2 > v["rs"] or 4 * (H + W) > v["ps"] * v["nt"]:
        return False
    if v["kind"] == 0:
        return -(-H // v["l"]) * W <= v["kmax"] * v["nt"]
    if v["kind"] >= 2:
        return True
    return H * W <= v["kmax"] * v["nt"] and (H - 2) * (W - 2) <= v["kf"] * v["nt"]
