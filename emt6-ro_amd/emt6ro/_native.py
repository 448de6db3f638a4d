"""ctypes binding of libemt6ro_amd.so (the C ABI in include/emt6ro_amd.h).

The library is built in-tree by ``make -C emt6-ro_amd``; importing this module
without it raises ImportError: there is no CPU fallback.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# EMT6RO_LIB selects another build of the same C ABI (e.g. a debug build).
LIB_PATH = os.environ.get("EMT6RO_LIB") or os.path.join(os.path.dirname(_HERE), "lib",
                                                        "libemt6ro_amd.so")

NUM_PARAMS = 40
FIELDS = {
    "cho": 0, "ox": 1, "gi": 2, "occupancy": 3, "time_in_repair": 4,
    "irradiation": 5, "repair_delay": 6, "proliferation_time": 7,
    "cycle_g1": 8, "cycle_s": 9, "cycle_g2": 10, "cycle_m": 11, "cycle_d": 12,
    "mode": 13, "phase": 14,
}
BYTE_FIELDS = {3, 13, 14}
PATH_PERSISTENT, PATH_STEP, PATH_STEP_GRAPH = 0, 1, 2


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `make -C emt6-ro_amd` "
            "(or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    i32, u32, u64, f32p = C.c_int32, C.c_uint32, C.c_uint64, C.POINTER(C.c_float)
    sig = {
        "emt6ro_last_error": ([], C.c_char_p),
        "emt6ro_version": ([], C.c_char_p),
        "emt6ro_device_count": ([], C.c_int),
        "emt6ro_params_load": ([C.c_char_p, C.POINTER(P)], C.c_int),
        "emt6ro_params_from_values": ([f32p, C.c_int, C.POINTER(P)], C.c_int),
        "emt6ro_params_values": ([P, f32p, C.c_int], C.c_int),
        "emt6ro_params_free": ([P], None),
        "emt6ro_state_load": ([C.c_char_p, P, C.POINTER(P)], C.c_int),
        "emt6ro_state_save": ([P, C.c_char_p], C.c_int),
        "emt6ro_state_create": ([i32, i32, P, C.POINTER(P)], C.c_int),
        "emt6ro_state_dims": ([P, C.POINTER(i32), C.POINTER(i32)], C.c_int),
        "emt6ro_state_get_field": ([P, C.c_int, P], C.c_int),
        "emt6ro_state_set_field": ([P, C.c_int, P], C.c_int),
        "emt6ro_state_free": ([P], None),
        "emt6ro_experiment_create": ([P, C.POINTER(P), i32, i32, i32, u32, u32, i32, u64,
                                      C.POINTER(P)], C.c_int),
        "emt6ro_experiment_add_irradiations": ([P, i32, C.POINTER(i32), C.POINTER(i32), f32p],
                                               C.c_int),
        "emt6ro_experiment_run": ([P, u32], C.c_int),
        "emt6ro_experiment_results": ([P, C.POINTER(u32)], C.c_int),
        "emt6ro_experiment_reset": ([P], C.c_int),
        "emt6ro_experiment_state": ([P, i32, C.POINTER(P)], C.c_int),
        "emt6ro_experiment_batch_size": ([P], C.c_int),
        "emt6ro_experiment_set_path": ([P, i32], C.c_int),
        "emt6ro_experiment_destroy": ([P], None),
        "emt6ro_batch_create": ([P, C.POINTER(P), i32, i32, C.POINTER(i32), C.POINTER(i32),
                                 C.POINTER(u32), C.POINTER(u32), i32, i32, f32p, u32, u32, i32,
                                 u64, C.POINTER(P)], C.c_int),
        "emt6ro_batch_run": ([P, u32, i32, i32, u32], C.c_int),
        "emt6ro_batch_results": ([P, C.POINTER(u32)], C.c_int),
        "emt6ro_batch_counters": ([P, C.POINTER(u64)], C.c_int),
        "emt6ro_batch_state": ([P, i32, C.POINTER(P)], C.c_int),
        "emt6ro_batch_reset": ([P, u32], C.c_int),
        "emt6ro_batch_last_run": ([P, C.POINTER(C.c_double), C.POINTER(C.c_int64)], C.c_int),
        "emt6ro_batch_destroy": ([P], None),
        "emt6ro_variant_info": ([i32, C.POINTER(i32)], C.c_int),
        "emt6ro_variant_occupancy": ([i32, i32, i32], C.c_int),
        "emt6ro_batch_set_variant": ([P, i32, i32], C.c_int),
        "emt6ro_batch_set_window": ([P, i32], C.c_int),
        "emt6ro_batch_set_doses": ([P, f32p], C.c_int),
        "emt6ro_instr_read": ([C.POINTER(u64), C.POINTER(C.c_double)], C.c_int),
        "emt6ro_instr_reset": ([], C.c_int),
        "emt6ro_instr_histograms": ([C.POINTER(u64), i32], C.c_int),
        "emt6ro_instr_enable": ([i32], C.c_int),
        "emt6ro_instr_enabled": ([], C.c_int),
        "emt6ro_mlp_last_error": ([], C.c_char_p),
        "emt6ro_mlp_create": ([i32, i32, u64, i32, C.POINTER(P)], C.c_int),
        "emt6ro_mlp_num_params": ([P], C.c_int64),
        "emt6ro_mlp_get_params": ([P, f32p], C.c_int),
        "emt6ro_mlp_set_params": ([P, f32p], C.c_int),
        "emt6ro_mlp_set_data": ([P, f32p, f32p, i32], C.c_int),
        "emt6ro_mlp_train": ([P, i32, i32, C.c_float, u64, f32p], C.c_int),
        "emt6ro_mlp_set_train_graph": ([P, i32], C.c_int),
        "emt6ro_mlp_train_many": ([P, i32, i32, i32, C.c_float, P, f32p], C.c_int),
        "emt6ro_mlp_predict": ([P, f32p, i32, f32p], C.c_int),
        "emt6ro_mlp_predict_noalloc": ([P, P, i32, P], C.c_int),
        "emt6ro_mlp_grads": ([P, f32p, f32p, i32, f32p, f32p], C.c_int),
        "emt6ro_mlp_destroy": ([P], None),
        "emt6ro_probe_lds": ([i32, i32, C.POINTER(C.c_double)], C.c_int),
        "emt6ro_probe_grid_barrier": ([i32, i32, i32, C.POINTER(C.c_double)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


lib = _load()


def check(rc):
    """Raise the library's last error as RuntimeError (pybind11 surfaces the
    reference's std::runtime_error the same way)."""
    if rc != 0:
        raise RuntimeError(lib.emt6ro_last_error().decode())
    return rc


def probe_lds(width_bytes=16, device=0):
    """LDS read bandwidth of the device (emt6ro_probe_lds): bytes/s, bytes
    per shader clock per CU, shader clock (Hz), CUs."""
    out = (C.c_double * 4)()
    rc = lib.emt6ro_probe_lds(int(device), int(width_bytes), out)
    if rc != 0:
        raise RuntimeError(f"emt6ro_probe_lds failed ({rc})")
    return {"bytes_per_s": out[0], "bytes_per_clk_cu": out[1], "clock_hz": out[2],
            "cus": int(out[3])}


def probe_grid_barrier(mode=1, words=0, device=0):
    """Microseconds per grid-wide barrier inside a kernel (one workgroup per
    CU) and per back-to-back kernel launch doing the same work
    (emt6ro_probe_grid_barrier)."""
    out = (C.c_double * 2)()
    rc = lib.emt6ro_probe_grid_barrier(int(device), int(mode), int(words), out)
    if rc != 0:
        raise RuntimeError(f"emt6ro_probe_grid_barrier failed ({rc})")
    return {"barrier_us": out[0], "launch_us": out[1]}


def check_mlp(rc):
    """`check` for the emt6ro_mlp_* calls, which keep their own last error."""
    if rc != 0:
        raise RuntimeError(lib.emt6ro_mlp_last_error().decode())
    return rc
