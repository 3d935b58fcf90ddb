"""CPU oracle of the EM path (test infrastructure only).

Loads oracle/_build/libmc_oracle.so, a C++ restatement of the reference's CPU
EM path (see mc_oracle.cpp for the per-function citations and parity pins).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it;
the product library never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libmc_oracle.so")
MAX_SERP = 10

_lib = None
i32p = C.POINTER(C.c_int32)
i64p = C.POINTER(C.c_int64)
u8p = C.POINTER(C.c_uint8)
f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
vp = C.c_void_p


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        h = C.CDLL(LIB_PATH)
        sigs = {
            "orc_parse_raw": (vp, [C.c_char_p, C.c_int, C.c_float, C.c_char_p, C.c_int]),
            "orc_from_sessions": (vp, [C.c_int64, i32p, i32p, u8p, C.c_float]),
            "orc_dataset_sizes": (None, [vp, i64p, i64p, i64p, i64p]),
            "orc_dataset_export": (None, [vp, C.c_int, i32p, i32p, u8p, i32p]),
            "orc_dataset_free": (None, [vp]),
            "orc_model_create": (vp, [C.c_int, C.c_int64, i32p, i32p, u8p, u8p, C.c_int64, C.c_int]),
            "orc_model_set_options": (None, [vp, C.c_int, C.c_int]),
            "orc_model_threads": (C.c_int, [vp]),
            "orc_estep": (None, [vp]),
            "orc_mstep": (None, [vp]),
            "orc_iterate": (None, [vp, C.c_int]),
            "orc_n_tables": (C.c_int, [vp]),
            "orc_table_size": (C.c_int64, [vp, C.c_int]),
            "orc_tmp_size": (C.c_int64, [vp, C.c_int]),
            "orc_table_name": (C.c_char_p, [vp, C.c_int]),
            "orc_get_table": (None, [vp, C.c_int, f32p, f32p, f32p]),
            "orc_set_table": (None, [vp, C.c_int, f32p, f32p]),
            "orc_get_tmp": (None, [vp, C.c_int, f32p, f32p]),
            "orc_eval": (None, [vp, C.c_int64, i32p, u8p, u8p, f32p, f32p, f32p, f32p, f64p, f32p, f64p, i64p]),
            "orc_serp_probability": (None, [vp, C.c_int64, i32p, u8p, u8p, f32p]),
            "orc_model_free": (None, [vp]),
        }
        for name, (res, args) in sigs.items():
            fn = getattr(h, name)
            fn.restype = res
            fn.argtypes = args
        _lib = h
    return _lib


def _p(a, t):
    return None if a is None else a.ctypes.data_as(t)


def _export_dataset(h):
    L = lib()
    n = [C.c_int64() for _ in range(4)]
    L.orc_dataset_sizes(h, *[C.byref(x) for x in n])
    n_parsed, n_train, n_test, n_qd = (x.value for x in n)
    out = {"n_parsed": n_parsed, "n_qd": n_qd}
    for which, key, cnt in ((0, "train", n_train), (1, "test", n_test)):
        q = np.empty(cnt, np.int32)
        d = np.empty((cnt, MAX_SERP), np.int32)
        c = np.empty((cnt, MAX_SERP), np.uint8)
        p = np.empty((cnt, MAX_SERP), np.int32)
        if cnt:
            L.orc_dataset_export(h, which, _p(q, i32p), _p(d, i32p), _p(c, u8p), _p(p, i32p))
        out[key] = {"query": q, "docs": d, "clicks": c, "pidx": p}
    L.orc_dataset_free(h)
    return out


def parse_raw(path: str, max_sessions: int = -1, test_share: float = 0.2):
    err = C.create_string_buffer(512)
    h = lib().orc_parse_raw(path.encode(), int(max_sessions), float(test_share), err, 512)
    if not h:
        raise ValueError(err.value.decode())
    return _export_dataset(h)


def from_sessions(query, docs, clicks, test_share: float = 0.2):
    q = np.ascontiguousarray(query, np.int32)
    d = np.ascontiguousarray(docs, np.int32)
    c = np.ascontiguousarray(clicks, np.uint8)
    h = lib().orc_from_sessions(len(q), _p(q, i32p), _p(d, i32p), _p(c, u8p), float(test_share))
    if not h:
        raise ValueError("invalid split")
    return _export_dataset(h)


class OracleModel:
    """CPU EM of one click model (exec mode 1 of the reference)."""

    def __init__(self, model_type, query, pidx, clicks, n_qd, length=None, n_threads=1,
                 accum_fp64=False, copy_dataset=False):
        self._q = np.ascontiguousarray(query, np.int32)
        self._p = np.ascontiguousarray(pidx, np.int32)
        self._c = np.ascontiguousarray(clicks, np.uint8)
        self._l = None if length is None else np.ascontiguousarray(length, np.uint8)
        self.h = lib().orc_model_create(int(model_type), len(self._q), _p(self._q, i32p), _p(self._p, i32p),
                                        _p(self._c, u8p), _p(self._l, u8p), int(n_qd), int(n_threads))
        if not self.h:
            raise ValueError("invalid oracle model arguments")
        lib().orc_model_set_options(self.h, int(accum_fp64), int(copy_dataset))

    def close(self):
        if getattr(self, "h", None):
            lib().orc_model_free(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @property
    def threads(self):
        return lib().orc_model_threads(self.h)

    def estep(self):
        lib().orc_estep(self.h)

    def mstep(self):
        lib().orc_mstep(self.h)

    def iterate(self, n):
        lib().orc_iterate(self.h, int(n))

    def n_tables(self):
        return lib().orc_n_tables(self.h)

    def table_name(self, t):
        return lib().orc_table_name(self.h, t).decode()

    def get_table(self, t):
        n = lib().orc_table_size(self.h, t)
        num, den, val = (np.empty(n, np.float32) for _ in range(3))
        lib().orc_get_table(self.h, t, _p(num, f32p), _p(den, f32p), _p(val, f32p))
        return num, den, val

    def set_table(self, t, num, den):
        num = np.ascontiguousarray(num, np.float32)
        den = np.ascontiguousarray(den, np.float32)
        lib().orc_set_table(self.h, t, _p(num, f32p), _p(den, f32p))

    def get_tmp(self, t):
        n = lib().orc_tmp_size(self.h, t)
        num, den = np.empty(n, np.float32), np.empty(n, np.float32)
        lib().orc_get_tmp(self.h, t, _p(num, f32p), _p(den, f32p))
        return num, den

    def serp_probability(self, pidx, clicks, length=None):
        p = np.ascontiguousarray(pidx, np.int32)
        c = np.ascontiguousarray(clicks, np.uint8)
        ln = None if length is None else np.ascontiguousarray(length, np.uint8)
        out = np.empty((len(p), MAX_SERP), np.float32)
        lib().orc_serp_probability(self.h, len(p), _p(p, i32p), _p(c, u8p), _p(ln, u8p), _p(out, f32p))
        return out

    def evaluate(self, pidx, clicks, length=None):
        p = np.ascontiguousarray(pidx, np.int32)
        c = np.ascontiguousarray(clicks, np.uint8)
        ln = None if length is None else np.ascontiguousarray(length, np.uint8)
        n = len(p)
        logp = np.empty((n, MAX_SERP), np.float32)
        full = np.empty((n, MAX_SERP), np.float32)
        llh = np.empty(n, np.float32)
        t32 = C.c_float()
        t64 = C.c_double()
        pr32 = np.empty(MAX_SERP, np.float32)
        pr64 = np.empty(MAX_SERP, np.float64)
        cnt = np.empty(MAX_SERP, np.int64)
        lib().orc_eval(self.h, n, _p(p, i32p), _p(c, u8p), _p(ln, u8p), _p(logp, f32p), _p(full, f32p),
                       _p(llh, f32p), C.byref(t32), C.byref(t64), _p(pr32, f32p), _p(pr64, f64p), _p(cnt, i64p))
        return {"logp": logp, "full": full, "llh": llh, "llh_sum_f32": t32.value, "llh_sum_f64": t64.value,
                "ppl_log2_sum_f32": pr32, "ppl_log2_sum_f64": pr64, "rank_count": cnt}
