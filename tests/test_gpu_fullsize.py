"""Properties at the headline size (DBN, 10M sessions, 8M train, 6.8M
entries), where the oracle is too slow to run 50 iterations: determinism of
the segmented mode, agreement of the two reduction modes, exact
denominators, and the FP64 held-out log-likelihood against a float64
recomputation from the trained values."""
import numpy as np
import pytest

from llh64 import loglik64
from massiveclicks_amd import DBN, REDUCE_ATOMIC, ClickModel, Dataset, synth

pytestmark = pytest.mark.gpu

_ds = {}


def dataset():
    if "ds" not in _ds:
        s = synth.generate(synth.SynthConfig(n_sessions=10_000_000, n_queries=1_000_000, vocab=20_000_000,
                                             zipf_s=1.1, model="dbn", seed=3))
        _ds["ds"] = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2)
    return _ds["ds"]


def test_dbn_10m_properties():
    ds = dataset()
    tr, te = ds.train, ds.test
    print(f"case: fullsize DBN n_train={ds.n_train} n_qd={ds.n_qd}", flush=True)
    runs = []
    for _ in range(2):
        with ClickModel(DBN) as cm:
            cm.init_from_dataset(ds)
            cm.em(50)
            runs.append([cm.get_parameters(t) for t in range(3)])
            if len(runs) == 1:
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
        cm.em(50)
        for t in range(3):
            err = float(np.max(np.abs(cm.get_values(t) - vals[t])))
            assert err < 1e-5, (t, err)
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
