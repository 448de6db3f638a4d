"""ctypes wrapper of the CPU oracle (liboracle: emt6ro_oracle.c) -- TEST
INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module, and only as the checker or the timed CPU baseline.  It
reads parameters (JSON) and tumour states (text) itself, independently of the
product's loaders (parameters/parameters.cc:66-80, state/state.cc:9-53).
"""
import ctypes as C
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libemt6ro_oracle.so")

VACANT, OCCUPIED, MOCKED = 0, 1, 2
# Step-rule switches of eo_run_rules (emt6ro_oracle.h): the reference's literal
# duplicate-entry stepping and list-order vacancy reads, for A/B statistics.
RULE_Q3_EVERY_ENTRY, RULE_Q4_LIST_ORDER = 1, 2
# Instrumentation record (emt6ro_oracle.h, EO_INSTR_*): 5 counters, then two
# 300-bin histograms.
INSTR_N = 605


class EoParams(C.Structure):
    _fields_ = [
        ("H", C.c_int32), ("W", C.c_int32), ("time_step", C.c_float),
        ("diff_coeffs", C.c_float * 3), ("diff_time_step", C.c_float),
        ("metab", (C.c_float * 3) * 4), ("ext", C.c_float * 3),
        ("death_gi", C.c_float), ("quiescence_gi", C.c_float),
        ("death_gi_production", C.c_float),
        ("cycle_mean", C.c_float * 5), ("cycle_sd", C.c_float * 5),
        ("delay_coeff", C.c_float), ("delay_exp", C.c_float), ("surv_coeff", C.c_float),
        ("surv_exp", C.c_float), ("repair_half_time", C.c_float),
    ]


CELL_DTYPE = np.dtype([("t_rep", "<f4"), ("R", "<f4"), ("delay", "<f4"), ("prolif", "<f4"),
                       ("times", "<f4", (5,)), ("mode", "u1"), ("phase", "u1"),
                       ("pad0", "<u2"), ("pad1", "<u4", (2,))])
assert CELL_DTYPE.itemsize == 48


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    u32, i32, P = C.c_uint32, C.c_int32, C.c_void_p
    lib.eo_philox4x32_10.argtypes = [C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
    lib.eo_u01.argtypes = [u32]
    lib.eo_u01.restype = C.c_float
    lib.eo_expf.argtypes = [C.c_float]
    lib.eo_expf.restype = C.c_float
    lib.eo_logf.argtypes = [C.c_float]
    lib.eo_logf.restype = C.c_float
    lib.eo_sincos2pi.argtypes = [C.c_float, C.POINTER(C.c_float), C.POINTER(C.c_float)]
    lib.eo_normals5.argtypes = [u32, u32, u32, u32, u32, C.POINTER(C.c_float)]
    lib.eo_uniform.argtypes = [u32, u32, u32, u32, u32, u32]
    lib.eo_uniform.restype = C.c_float
    lib.eo_update_state.argtypes = [P, C.POINTER(C.c_float), C.POINTER(EoParams), C.c_int]
    lib.eo_progress_clock.argtypes = [P, C.c_float]
    lib.eo_metabolise.argtypes = [P, C.POINTER(C.c_float), C.POINTER(EoParams)]
    lib.eo_irradiate.argtypes = [P, C.c_float, C.POINTER(EoParams)]
    lib.eo_delay_time.argtypes = [C.c_float, C.POINTER(EoParams)]
    lib.eo_delay_time.restype = C.c_float
    lib.eo_try_repair.argtypes = [P, C.c_int, C.POINTER(EoParams), C.c_float]
    lib.eo_choose_neighbour.argtypes = [C.c_int, C.c_int, C.c_float, C.POINTER(C.c_int),
                                        C.POINTER(C.c_int)]
    lib.eo_arbitration_key.argtypes = [u32]
    lib.eo_arbitration_key.restype = u32
    lib.eo_find_roi.argtypes = [C.c_int, C.c_int, P, C.POINTER(i32)]
    lib.eo_mask.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
    lib.eo_diffuse_f64.argtypes = [C.POINTER(EoParams), C.c_int, C.c_int, P, C.POINTER(i32), P]
    lib.eo_diffuse_f32.argtypes = [C.POINTER(EoParams), C.c_int, C.c_int, P, C.POINTER(i32)]
    lib.eo_run.argtypes = [C.POINTER(EoParams), C.c_int, P, P, P, P, P, i32, P, P, P, P, P,
                           i32, u32, u32, u32, u32, u32, C.c_int]
    lib.eo_run.restype = C.c_int
    lib.eo_run_instr.argtypes = lib.eo_run.argtypes + [P]
    lib.eo_run_instr.restype = C.c_int
    lib.eo_run_rules.argtypes = lib.eo_run_instr.argtypes + [C.c_int]
    lib.eo_run_rules.restype = C.c_int
    lib.eo_count_living.argtypes = [C.c_int, C.c_int, P]
    lib.eo_count_living.restype = u32
    return lib


lib = _load()


def ptr(a):
    return a.ctypes.data_as(C.c_void_p)


# ---- input readers (independent of the product) ---------------------------

def read_params_json(path):
    """Flat parameter vector in the C ABI order, read with the json module."""
    with open(path) as f:
        o = json.load(f)
    m = o["metabolism"]
    subs = lambda d: [d["cho"], d["ox"], d["gi"]]  # noqa: E731
    v = [o["lattice_dims"]["height"], o["lattice_dims"]["width"], o["time_step"]]
    v += subs(o["diffusion_params"]["coeffs"]) + [o["diffusion_params"]["time_step"]]
    for k in ("aerobic_proliferation", "anaerobic_proliferation", "aerobic_quiescence",
              "anaerobic_quiescence"):
        v += subs(m[k])
    v += subs(o["external_levels"])
    v += [o["death_gi"], o["quiescence_gi"], o["death_gi_production"]]
    for k in ("g1", "s", "g2", "m", "d"):
        v += [o["cycle_times"][k]["mean"], o["cycle_times"][k]["stddev"]]
    cr = o["cell_repair"]
    v += [cr["delay_time"]["coeff"], cr["delay_time"]["exp_coeff"],
          cr["survival_prob"]["coeff"], cr["survival_prob"]["exp_coeff"],
          cr["repair_half_time"]]
    return np.array(v, dtype=np.float64).astype(np.float32)


def make_params(values, H, W):
    v = [float(x) for x in np.asarray(values, dtype=np.float32)]
    p = EoParams()
    p.H, p.W = int(H), int(W)
    p.time_step = v[2]
    for s in range(3):
        p.diff_coeffs[s] = v[3 + s]
    p.diff_time_step = v[6]
    for mm in range(4):
        for s in range(3):
            p.metab[mm][s] = v[7 + 3 * mm + s]
    for s in range(3):
        p.ext[s] = v[19 + s]
    p.death_gi, p.quiescence_gi, p.death_gi_production = v[22], v[23], v[24]
    for k in range(5):
        p.cycle_mean[k] = v[25 + 2 * k]
        p.cycle_sd[k] = v[26 + 2 * k]
    p.delay_coeff, p.delay_exp, p.surv_coeff, p.surv_exp, p.repair_half_time = v[35:40]
    return p


def empty_state(values, H, W):
    """Per-site arrays of an all-vacant bordered lattice (MOCKED border)."""
    ext = np.asarray(values, dtype=np.float32)[19:22]
    s = {"sub": np.empty((3, H, W), np.float32), "st": np.zeros((H, W), np.uint8),
         "cells": np.zeros(H * W, CELL_DTYPE)}
    for k in range(3):
        s["sub"][k] = ext[k]
    s["st"][0, :] = s["st"][-1, :] = s["st"][:, 0] = s["st"][:, -1] = MOCKED
    return s


def read_state_text(path, values):
    """Tumour text format of state/state.cc:9-53."""
    with open(path) as f:
        tok = f.read().split()
    h, w = int(tok[0]), int(tok[1])
    s = empty_state(values, h + 2, w + 2)
    p = make_params(values, h + 2, w + 2)
    i = 2
    for r in range(1, h + 1):
        for c in range(1, w + 1):
            occ = int(tok[i]); i += 1
            for k in range(3):
                s["sub"][k, r, c] = np.float32(tok[i]); i += 1
            if occ:
                s["st"][r, c] = OCCUPIED
                cell = s["cells"][r * (w + 2) + c]
                cell["t_rep"] = np.float32(tok[i]); i += 1
                R = np.float32(tok[i]); i += 1
                cell["R"] = R
                cell["delay"] = lib.eo_delay_time(float(R), C.byref(p)) if R > 0 else 0.0
                cell["prolif"] = np.float32(tok[i]); i += 1
                cell["times"] = [np.float32(tok[i + k]) for k in range(5)]; i += 5
                cell["mode"] = int(tok[i]); i += 1
                cell["phase"] = int(tok[i]); i += 1
            else:
                s["st"][r, c] = VACANT
    return s


class OracleBatch:
    """B simulations of one lattice size, run by eo_run.

    states: per-simulation dicts from read_state_text / empty_state (copied).
    prot: (n_prot, n_slots) doses; prot_index, sim_id: per simulation.
    """

    def __init__(self, values, states, prot, prot_index, sim_id, prot_res, prot_length):
        H, W = states[0]["st"].shape
        self.H, self.W, self.B = H, W, len(states)
        self.p = make_params(values, H, W)
        self.cap = (H - 2) * (W - 2) + 128
        self.sub = np.ascontiguousarray(np.stack([s["sub"] for s in states]), np.float32)
        self.st = np.ascontiguousarray(np.stack([s["st"] for s in states]), np.uint8)
        self.cells = np.ascontiguousarray(np.stack([s["cells"] for s in states]))
        self.list = np.zeros((self.B, self.cap), np.uint16)
        self.nlist = np.zeros(self.B, np.int32)
        self.roi = np.zeros((self.B, 5), np.int32)
        self.step = np.zeros(self.B, np.uint32)
        self.sim_id = np.ascontiguousarray(sim_id, np.int32)
        self.prot_index = np.ascontiguousarray(prot_index, np.int32)
        self.prot = np.ascontiguousarray(prot, np.float32)
        self.prot_res, self.prot_length = int(prot_res), int(prot_length)

    def run(self, nsteps, seed, epoch=0, nthreads=1, instr=False, rules=0):
        """With instr=True, also accumulate the instrumentation counters of
        common/instr.h into self.instr (B x INSTR_N uint64, see
        emt6ro_oracle.h).  rules: RULE_* bits (0 = the product's rules)."""
        if instr and getattr(self, "instr", None) is None:
            self.instr = np.zeros((self.B, INSTR_N), np.uint64)
        rc = lib.eo_run_rules(C.byref(self.p), self.B, ptr(self.sub), ptr(self.st),
                              ptr(self.cells), ptr(self.list), ptr(self.nlist), self.cap,
                              ptr(self.roi), ptr(self.step), ptr(self.sim_id),
                              ptr(self.prot_index), ptr(self.prot), self.prot.shape[1],
                              self.prot_res, self.prot_length, int(seed), int(epoch),
                              int(nsteps), int(nthreads),
                              ptr(self.instr) if instr else None, int(rules))
        if rc != 0:
            raise RuntimeError("oracle: occupied list overflow")

    def counts(self):
        return (self.st == OCCUPIED).reshape(self.B, -1).sum(1).astype(np.int64)
