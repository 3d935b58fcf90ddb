"""The C-ABI library loads on a machine without a GPU and exports every
entry point include/mclicks.h declares; host-only entry points work and the
device ones fail with a status instead of aborting."""
import ctypes as C
import os
import re

import numpy as np

from massiveclicks_amd import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
HEADER = os.path.join(os.path.dirname(HERE), "include", "mclicks.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mc_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) >= 25
    so = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing
    assert sorted(_lib.EXPORTED) == names


def test_memory_footprint_and_errors_without_device():
    L = _lib.lib()
    b = C.c_size_t()
    assert L.mc_compute_memory_footprint(2, 8_000_000, 6_800_000, C.byref(b)) == 0
    assert b.value > 8_000_000 * 42
    assert L.mc_compute_memory_footprint(7, 1, 1, C.byref(b)) == 1        # bad model id
    h = C.c_void_p()
    assert L.mc_model_create(9, 0, C.byref(h)) == 1                       # bad model id
    assert L.mc_last_error()
    assert L.mc_process_session(None) == 1                                # null model
    ds = C.c_void_p()
    q = np.zeros(1, np.int32)
    st = L.mc_dataset_from_sessions(0, _lib.ptr(q, _lib.i32p), _lib.ptr(q, _lib.i32p), None, C.c_float(0.2),
                                    C.byref(ds))
    assert st == 1


def test_session_arrays_are_shape_checked_before_native_calls():
    """The Python mirror rejects session arrays whose shapes differ from the
    element counts it passes to the C ABI (which trusts them)."""
    import pytest
    from massiveclicks_amd.model import SessionSet, _sessions_arrays
    n = 4
    ok = SessionSet(np.zeros(n, np.int32), np.zeros((n, 10), np.int32), np.zeros((n, 10), np.uint8),
                    np.zeros((n, 10), np.int32), np.full(n, 10, np.uint8))
    p, c, ln = _sessions_arrays(ok, "test")
    assert p.shape == (n, 10) and c.shape == (n, 10) and ln.shape == (n,)
    bad = [
        SessionSet(ok.query, ok.docs, ok.clicks, np.zeros((n, 9), np.int32)),
        SessionSet(ok.query, ok.docs, np.zeros((n - 1, 10), np.uint8), ok.pidx),
        SessionSet(ok.query, ok.docs, ok.clicks, ok.pidx, np.full(n + 1, 10, np.uint8)),
    ]
    for b in bad:
        with pytest.raises(ValueError):
            _sessions_arrays(b, "test")
