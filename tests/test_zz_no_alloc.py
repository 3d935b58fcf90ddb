"""mc_evaluate_device, the evaluation entry point that allocates nothing,
against mc_evaluate on the same sessions: two calls in a row on one model,
the second on fewer sessions, for every model.  Inputs and outputs are
device buffers the test allocates itself (hipMalloc through ctypes); the
entry point's FP64 partials live in the library's fixed scratch array.

Bar: the same kernel arithmetic and grid as mc_evaluate, so per-session
values and the 21 totals are bit-identical."""
import ctypes as C

import numpy as np
import pytest

from massiveclicks_amd import CCM, DBN, PBM, UBM, ClickModel, Dataset, synth
from massiveclicks_amd import _lib

pytestmark = pytest.mark.gpu

NAMES = {PBM: "pbm", UBM: "ubm", DBN: "dbn", CCM: "ccm"}


class Hip:
    """The few HIP runtime calls the test needs for its own device buffers."""

    def __init__(self):
        self.h = C.CDLL("libamdhip64.so")
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(max(a.nbytes, 16))) == 0
        self.bufs.append(p)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0   # H2D
        return p

    def empty(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(max(nbytes, 16))) == 0
        self.bufs.append(p)
        return p

    def download(self, p, shape, dtype):
        out = np.empty(shape, dtype)
        assert self.h.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0   # D2H
        return out

    def free_memory(self):
        free, total = C.c_size_t(), C.c_size_t()
        assert self.h.hipMemGetInfo(C.byref(free), C.byref(total)) == 0
        return free.value

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


def device_evaluate(hip, cm, pidx, clicks, length):
    n = len(pidx)
    d_p = hip.upload(pidx.astype(np.int32))
    d_c = hip.upload(clicks.astype(np.uint8))
    d_l = hip.upload(length.astype(np.uint8)) if length is not None else None
    d_lp, d_fp, d_llh = hip.empty(n * 40), hip.empty(n * 40), hip.empty(n * 4)
    d_tot = hip.empty(21 * 8)
    before = hip.free_memory()
    _lib.check(_lib.lib().mc_evaluate_device(cm._h, n, d_p, d_c, d_l, d_lp, d_fp, d_llh, d_tot))
    _lib.check(_lib.lib().mc_synchronize(cm._h))
    after = hip.free_memory()
    out = {
        "logp": hip.download(d_lp, (n, 10), np.float32),
        "full": hip.download(d_fp, (n, 10), np.float32),
        "llh": hip.download(d_llh, n, np.float32),
        "totals": hip.download(d_tot, 21, np.float64),
    }
    return out, before - after


@pytest.mark.parametrize("uneven", [False, True], ids=["ten", "uneven"])
@pytest.mark.parametrize("model", [PBM, UBM, DBN, CCM], ids=["PBM", "UBM", "DBN", "CCM"])
def test_evaluate_device_matches_evaluate(model, uneven):
    s = synth.generate(synth.SynthConfig(n_sessions=30000, n_queries=2000, vocab=80000, model=NAMES[model],
                                         variable_length=uneven, seed=800 + model + 10 * uneven))
    ds = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2, length=s.length if uneven else None)
    te = ds.test
    print(f"case: no-alloc evaluation model={model} uneven={uneven} n_test={len(te)}", flush=True)
    hip = Hip()
    try:
        with ClickModel(model) as cm:
            cm.init_from_dataset(ds)
            cm.em(10)
            for k, n in enumerate((len(te), len(te) // 3)):
                ln = te.length[:n] if te.length is not None else None
                got, used = device_evaluate(hip, cm, te.pidx[:n], te.clicks[:n], ln)
                sub = type(te)(te.query[:n], te.docs[:n], te.clicks[:n], te.pidx[:n], ln)
                ref = cm.evaluate(sub, per_session=True)
                for key in ("logp", "full", "llh"):
                    np.testing.assert_array_equal(got[key], ref[key], err_msg=f"call {k}: {key}")
                tot = got["totals"]
                assert tot[0] == ref["llh_sum"], (k, tot[0], ref["llh_sum"])
                np.testing.assert_array_equal(tot[1:11], ref["ppl_log2_sum"])
                np.testing.assert_array_equal(tot[11:21].astype(np.int64), ref["rank_count"])
                if k == 1:
                    assert used == 0, f"device memory changed by {used} bytes during the call"
            # Sizes beyond the maximum are rejected before anything runs.
            d = hip.empty(64)
            st = _lib.lib().mc_evaluate_device(cm._h, 1 << 31, d, d, None, None, None, None, d)
            assert st == 1 and b"out of range" in _lib.lib().mc_last_error()
    finally:
        hip.close()
