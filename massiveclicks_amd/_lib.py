"""ctypes binding of include/mclicks.h (libmclicks.so, built in-tree)."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmclicks.so")

_lib = None

i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p

_SIGS = {
    "mc_dataset_parse": (C.c_int, [C.c_char_p, C.c_int, C.c_float, C.POINTER(vp)]),
    "mc_dataset_from_sessions": (C.c_int, [C.c_int64, i32p, i32p, u8p, C.c_float, C.POINTER(vp)]),
    "mc_dataset_from_sessions_len": (C.c_int, [C.c_int64, i32p, i32p, u8p, u8p, C.c_float, C.POINTER(vp)]),
    "mc_dataset_sizes": (C.c_int, [vp, i64p, i64p, i64p, i64p]),
    "mc_dataset_export_len": (C.c_int, [vp, C.c_int, u8p]),
    "mc_dataset_export": (C.c_int, [vp, C.c_int, i32p, i32p, u8p, i32p]),
    "mc_dataset_free": (None, [vp]),
    "mc_model_create": (C.c_int, [C.c_int, C.c_int, C.POINTER(vp)]),
    "mc_model_set_option": (C.c_int, [vp, C.c_int, C.c_int64]),
    "mc_init_parameters": (C.c_int, [vp, C.c_int64, i32p, i32p, u8p, u8p, C.c_int64]),
    "mc_init_from_dataset": (C.c_int, [vp, vp]),
    "mc_process_session": (C.c_int, [vp]),
    "mc_update_parameters": (C.c_int, [vp]),
    "mc_em_iterations": (C.c_int, [vp, C.c_int, C.c_int]),
    "mc_em_iterations_timed": (C.c_int, [vp, C.c_int, f64p, f64p, f64p]),
    "mc_em_graph_timed": (C.c_int, [vp, C.c_int, f64p]),
    "mc_synchronize": (C.c_int, [vp]),
    "mc_get_n_tables": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "mc_get_table_info": (C.c_int, [vp, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_int), i64p]),
    "mc_get_parameters": (C.c_int, [vp, C.c_int, f32p, f32p]),
    "mc_set_parameters": (C.c_int, [vp, C.c_int, f32p, f32p]),
    "mc_reset_parameters": (C.c_int, [vp]),
    "mc_evaluate": (C.c_int, [vp, C.c_int64, i32p, u8p, u8p, f32p, f32p, f32p, f64p, f64p, i64p]),
    "mc_serp_probability": (C.c_int, [vp, C.c_int64, i32p, u8p, u8p, f32p]),
    "mc_set_test_set": (C.c_int, [vp, C.c_int64, i32p, u8p, u8p]),
    "mc_evaluate_test_set": (C.c_int, [vp, f32p, f32p, f32p, f64p, f64p, i64p, f64p]),
    "mc_evaluate_device": (C.c_int, [vp, C.c_int64, vp, vp, vp, vp, vp, vp, vp]),
    "mc_get_memory_usage": (C.c_int, [vp, C.POINTER(C.c_size_t)]),
    "mc_compute_memory_footprint": (C.c_int, [C.c_int, C.c_int64, C.c_int64, C.POINTER(C.c_size_t)]),
    "mc_get_layout_info": (C.c_int, [vp, i64p, i64p, i64p, i64p]),
    "mc_get_stream_info": (C.c_int, [vp, i64p, i64p]),
    "mc_get_session_formats": (C.c_int, [vp, i64p, i64p, i64p]),
    "mc_get_model_type": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "mc_autotune": (C.c_int, [C.c_int, C.c_int, C.c_int64, i32p, i32p, u8p, u8p, C.c_int64, C.c_int, i64p, f64p,
                              C.POINTER(C.c_int)]),
    "mc_grad_init": (C.c_int, [vp]),
    "mc_loglik_gradient": (C.c_int, [vp, C.c_int, f64p, C.POINTER(f32p)]),
    "mc_adam_configure": (C.c_int, [vp, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int]),
    "mc_adam_steps": (C.c_int, [vp, C.c_int, C.c_int]),
    "mc_adam_steps_timed": (C.c_int, [vp, C.c_int, f64p]),
    "mc_get_logits": (C.c_int, [vp, C.c_int, f32p]),
    "mc_set_logits": (C.c_int, [vp, C.c_int, f32p]),
    "mc_output_parameters": (C.c_int, [vp, vp, C.c_char_p]),
    "mc_write_parameters_csv": (C.c_int, [vp, C.c_char_p, C.c_int, C.POINTER(f32p), C.POINTER(f32p)]),
    "mc_save_parameters": (C.c_int, [vp, C.c_char_p]),
    "mc_load_parameters": (C.c_int, [vp, C.c_char_p]),
    "mc_model_destroy": (None, [vp]),
    "mc_last_error": (C.c_char_p, []),
}

EXPORTED = tuple(_SIGS)


class MCError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"mclicks status {status}: {msg}")
        self.status = status


def lib():
    """Loads libmclicks.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP engine first "
                "(python -c 'import __graft_entry__ as g; g.build()')")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def check(status: int) -> None:
    if status != 0:
        raise MCError(status, lib().mc_last_error().decode())


def ptr(a, typ):
    """ctypes pointer to a numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(typ)
