"""Parity at the configurations' full size: 10M synthetic sessions (8M
train, 6.8M query-document entries) per model, 50 EM iterations on the GPU
against 50 iterations of the CPU oracle on the same prepared sessions.

The oracle runs with FP64 M-step sums (its E-step arithmetic unchanged) on
up to 16 host threads: the E-step temporaries do not depend on the thread
count and the FP64 M-step is sequential, so the run is reproducible; it
takes 20-40 s per model on the GPU host.  Bars as in the smaller parity
tests: every table within 1e-5 absolute, held-out log-likelihood and
per-rank perplexity sums within 1e-6 relative of the FP64 total of the
oracle's per-session FP32 terms, per-qd denominators exactly equal.

The DBN case also checks properties that need no oracle: two segmented runs
are bit-identical, the per-qd denominators equal their closed forms, and the
engine's FP64 held-out total equals a float64 recomputation of the
conditional chain from the trained values.
"""
import gc
import os
import time

import numpy as np
import pytest

import oracle
from llh64 import loglik64
from massiveclicks_amd import CCM, DBN, PBM, UBM, REDUCE_ATOMIC, REDUCE_SEGMENTED, ClickModel, Dataset, synth

pytestmark = pytest.mark.gpu

TOL_PARAM = 1e-5
TOL_LL = 1e-6
ITERS = 50
NAMES = {PBM: "pbm", UBM: "ubm", DBN: "dbn", CCM: "ccm"}

_cache = {}


def full_workload(model):
    """The 10M-session workload of `model` (the bench's generator, seed 3,
    each model's own ground truth), prepared by the engine, and the
    oracle's 50-iteration tables and held-out evaluation on it."""
    if _cache.get("model") != model:
        _cache.clear()
        gc.collect()
        t0 = time.perf_counter()
        s = synth.generate(synth.SynthConfig(n_sessions=10_000_000, n_queries=1_000_000, vocab=20_000_000,
                                             zipf_s=1.1, model=NAMES[model], seed=3))
        ds = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2)
        del s
        tr, te = ds.train, ds.test
        threads = max(1, min(os.cpu_count() or 1, 16))
        om = oracle.OracleModel(model, tr.query, tr.pidx, tr.clicks, ds.n_qd, n_threads=threads, accum_fp64=True)
        om.iterate(ITERS)
        tabs = [om.get_table(t) for t in range(om.n_tables())]
        oev = om.evaluate(te.pidx, te.clicks)
        om.close()
        print(f"case: fullsize {NAMES[model]} n_train={ds.n_train} n_qd={ds.n_qd} oracle {threads} threads, "
              f"{time.perf_counter() - t0:.1f} s", flush=True)
        _cache.update(model=model, ds=ds, tabs=tabs, oev=oev)
    return _cache["ds"], _cache["tabs"], _cache["oev"]


def check_against_oracle(cm, ds, tabs, oev):
    info = cm.get_parameter_information()
    for t, (onum, oden, oval) in enumerate(tabs):
        err = float(np.max(np.abs(cm.get_values(t) - oval)))
        assert err < TOL_PARAM, f"table {t} ({info[t][0]}): max |value - oracle| = {err}"
        if not info[t][1]:
            _, den = cm.get_parameters(t)
            np.testing.assert_array_equal(den, oden)
    ev = cm.evaluate(ds.test)
    rel = abs(ev["llh_sum"] - oev["llh_sum_f64"]) / abs(oev["llh_sum_f64"])
    assert rel < TOL_LL, f"LL rel diff {rel}"
    prel = np.abs(ev["ppl_log2_sum"] - oev["ppl_log2_sum_f64"]) / np.abs(oev["ppl_log2_sum_f64"])
    assert prel.max() < TOL_LL, f"perplexity rel diff {prel.max()}"
    np.testing.assert_array_equal(ev["rank_count"], oev["rank_count"])
    return ev


def test_dbn_10m_matches_oracle_and_properties():
    ds, tabs, oev = full_workload(DBN)
    tr, te = ds.train, ds.test
    runs = []
    for k in range(2):
        with ClickModel(DBN) as cm:
            cm.init_from_dataset(ds)
            cm.em(ITERS)
            runs.append([cm.get_parameters(t) for t in range(3)])
            if k == 0:
                check_against_oracle(cm, ds, tabs, oev)
                vals = [cm.get_values(t).astype(np.float64) for t in range(3)]
                ev = cm.evaluate(te)
    # deterministic, bit for bit
    for (n0, d0), (n1, d1) in zip(*runs):
        np.testing.assert_array_equal(n0, n1)
        np.testing.assert_array_equal(d0, d1)
    # denominators are the closed forms of the occurrence counts
    occ = np.bincount(tr.pidx.reshape(-1), minlength=ds.n_qd).astype(np.float64)
    clk = np.bincount(tr.pidx.reshape(-1), weights=tr.clicks.reshape(-1), minlength=ds.n_qd)
    np.testing.assert_array_equal(runs[0][0][1].astype(np.float64), 2 + occ)
    np.testing.assert_array_equal(runs[0][1][1].astype(np.float64), 2 + 2 * occ - clk)
    # the float-atomic mode agrees within the parameter bar
    with ClickModel(DBN, reduce_mode=REDUCE_ATOMIC) as cm:
        cm.init_from_dataset(ds)
        cm.em(ITERS)
        for t in range(3):
            err = float(np.max(np.abs(cm.get_values(t) - tabs[t][2])))
            assert err < TOL_PARAM, (t, err)
    # held-out LL: FP64 total of the device's per-session terms vs a float64
    # recomputation of the conditional chain from the trained values
    # (unseen pairs valued 1/2), within 1e-6 relative
    v = [np.append(vals[0], 0.5), np.append(vals[1], 0.5)]
    pidx = np.where(te.pidx < 0, ds.n_qd, te.pidx)
    n = len(te.query)
    total = 0.0
    for a in range(0, n, 500_000):
        total += loglik64(DBN, v, vals[2], pidx[a:a + 500_000], te.clicks[a:a + 500_000]) / 10.0
    assert abs(ev["llh_sum"] - total) <= 1e-6 * abs(total), (ev["llh_sum"], total)


def test_ccm_10m_matches_oracle():
    ds, tabs, oev = full_workload(CCM)
    with ClickModel(CCM) as cm:
        cm.init_from_dataset(ds)
        cm.em(ITERS)
        check_against_oracle(cm, ds, tabs, oev)


@pytest.mark.parametrize("mode", [REDUCE_SEGMENTED, REDUCE_ATOMIC], ids=["segmented", "atomic"])
def test_ubm_10m_matches_oracle(mode):
    """The UBM configuration compares the deterministic segmented reduction
    with the atomic one: both against the oracle."""
    ds, tabs, oev = full_workload(UBM)
    with ClickModel(UBM, reduce_mode=mode) as cm:
        cm.init_from_dataset(ds)
        cm.em(ITERS)
        check_against_oracle(cm, ds, tabs, oev)


def test_pbm_10m_matches_oracle():
    ds, tabs, oev = full_workload(PBM)
    with ClickModel(PBM) as cm:
        cm.init_from_dataset(ds)
        cm.em(ITERS)
        check_against_oracle(cm, ds, tabs, oev)
    _cache.clear()
