"""GPU evaluation paths against the CPU oracle on identical parameters: the
resident held-out set (mc_set_test_set / mc_evaluate_test_set, sessions
reordered by query on the device, outputs back in the caller's order) and
get_serp_probability (mc_serp_probability).

Bar: per-session values within a few float ulps of the oracle's (only the
libm log differs: glibc logf/log2f against the device's), FP64 totals within
1e-7 relative; the resident path's per-session values equal mc_evaluate's
bit for bit (same arithmetic, other order of the FP64 sums)."""
import numpy as np
import pytest

import oracle
from massiveclicks_amd import CCM, DBN, PBM, UBM, ClickModel, Dataset, synth

pytestmark = pytest.mark.gpu

NAMES = {PBM: "pbm", UBM: "ubm", DBN: "dbn", CCM: "ccm"}
_cache = {}


def trained(model, uneven):
    """A 20-iteration GPU fit on 8,000 sessions, its held-out set, and an
    oracle model holding the GPU's tables."""
    key = (model, uneven)
    if key not in _cache:
        s = synth.generate(synth.SynthConfig(n_sessions=8000, n_queries=700, vocab=28000, model=NAMES[model],
                                             variable_length=uneven, seed=500 + model + 10 * uneven))
        ds = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2, length=s.length if uneven else None)
        cm = ClickModel(model)
        cm.init_from_dataset(ds)
        cm.em(20)
        tr = ds.train
        om = oracle.OracleModel(model, tr.query, tr.pidx, tr.clicks, ds.n_qd, length=tr.length)
        for t in range(cm.n_tables()):
            om.set_table(t, *cm.get_parameters(t))
        _cache[key] = (ds, cm, om)
    return _cache[key]


@pytest.mark.parametrize("uneven", [False, True], ids=["ten", "uneven"])
@pytest.mark.parametrize("model", [PBM, UBM, DBN, CCM], ids=["PBM", "UBM", "DBN", "CCM"])
def test_resident_test_set_matches_oracle(model, uneven):
    ds, cm, om = trained(model, uneven)
    te = ds.test
    print(f"case: resident test set model={model} uneven={uneven} n_test={len(te)}", flush=True)
    cm.set_test_set(te)
    ev = cm.evaluate_test_set(per_session=True, timed=True)
    assert ev["kernel_ms"] > 0
    oev = om.evaluate(te.pidx, te.clicks, te.length)
    np.testing.assert_allclose(ev["logp"], oev["logp"], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(ev["full"], oev["full"], rtol=2e-6, atol=1e-7)
    np.testing.assert_allclose(ev["llh"], oev["llh"], rtol=2e-6, atol=1e-7)
    assert abs(ev["llh_sum"] - oev["llh_sum_f64"]) <= 1e-7 * abs(oev["llh_sum_f64"])
    np.testing.assert_allclose(ev["ppl_log2_sum"], oev["ppl_log2_sum_f64"], rtol=1e-7)
    np.testing.assert_array_equal(ev["rank_count"], oev["rank_count"])
    # The caller-array path: same per-session values bit for bit.
    ref = cm.evaluate(te, per_session=True)
    for k in ("logp", "full", "llh"):
        np.testing.assert_array_equal(ev[k], ref[k])
    assert ev["llh_sum"] == pytest.approx(ref["llh_sum"], rel=1e-12)


@pytest.mark.parametrize("pool", [20, 400], ids=["packed", "rank-major"])
def test_resident_layout_follows_the_pair_spans(pool):
    """The resident set is stored as 16-byte records when every test
    session's seen pairs lie within 255 consecutive indices (pools of 20
    documents), else rank-major (pools of 400: a session can show pairs
    further apart); both match the oracle."""
    s = synth.generate(synth.SynthConfig(n_sessions=6000, n_queries=100, vocab=400000, pool=pool, model="dbn",
                                         seed=777 + pool))
    ds = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2)
    te = ds.test
    p = te.pidx.astype(np.int64)
    seen = p >= 0
    lo = np.where(seen, p, np.iinfo(np.int64).max).min(1)
    span = np.where(seen, p - lo[:, None], 0).max(1)
    packable = bool((span < 255).all())
    assert packable == (pool == 20), span.max()
    with ClickModel(DBN) as cm:
        cm.init_from_dataset(ds)
        cm.em(10)
        tr = ds.train
        om = oracle.OracleModel(DBN, tr.query, tr.pidx, tr.clicks, ds.n_qd)
        for t in range(cm.n_tables()):
            om.set_table(t, *cm.get_parameters(t))
        before = cm.get_memory_usage()
        cm.set_test_set(te)
        per_session = (cm.get_memory_usage() - before) / len(te)
        print(f"case resident layout pool={pool} n_test={len(te)} bytes/session={per_session:.1f}", flush=True)
        assert (per_session < 32) == packable, per_session
        ev = cm.evaluate_test_set(per_session=True)
        oev = om.evaluate(te.pidx, te.clicks, te.length)
        np.testing.assert_allclose(ev["llh"], oev["llh"], rtol=2e-6, atol=1e-7)
        np.testing.assert_allclose(ev["logp"], oev["logp"], rtol=2e-6, atol=1e-7)
        assert abs(ev["llh_sum"] - oev["llh_sum_f64"]) <= 1e-7 * abs(oev["llh_sum_f64"])
        np.testing.assert_array_equal(ev["rank_count"], oev["rank_count"])


def test_resident_test_set_is_replaced_and_dropped_by_init():
    """A second mc_set_test_set replaces the first (here a smaller subset in
    another order); mc_init_parameters drops the set and its memory."""
    from massiveclicks_amd._lib import MCError
    ds, cm, om = trained(DBN, False)
    te = ds.test
    perm = np.random.default_rng(1).permutation(len(te))[: len(te) // 3]
    sub = type(te)(te.query[perm], te.docs[perm], te.clicks[perm], te.pidx[perm], None)
    with ClickModel(DBN) as other:
        other.init_from_dataset(ds)
        fresh = other.get_memory_usage()
        for t in range(other.n_tables()):
            other.set_parameters(t, *cm.get_parameters(t))
        base = other.get_memory_usage()
        other.set_test_set(te)
        full_set = other.get_memory_usage()
        assert full_set > base
        other.set_test_set(sub)
        assert base < other.get_memory_usage() < full_set
        ev = other.evaluate_test_set(per_session=True)
        oev = om.evaluate(sub.pidx, sub.clicks)
        np.testing.assert_allclose(ev["llh"], oev["llh"], rtol=2e-6, atol=1e-7)
        assert abs(ev["llh_sum"] - oev["llh_sum_f64"]) <= 1e-7 * abs(oev["llh_sum_f64"])
        other.init_from_dataset(ds)
        assert other.get_memory_usage() == fresh
        with pytest.raises(MCError, match="no test set"):
            other.evaluate_test_set()


@pytest.mark.parametrize("uneven", [False, True], ids=["ten", "uneven"])
@pytest.mark.parametrize("model", [PBM, UBM, DBN, CCM], ids=["PBM", "UBM", "DBN", "CCM"])
def test_serp_probability_matches_oracle(model, uneven):
    """get_serp_probability (base.cuh:50) on the held-out sessions: within a
    float ulp of the oracle's (no transcendental on this path)."""
    ds, cm, om = trained(model, uneven)
    te = ds.test
    got = cm.get_serp_probability(te)
    want = om.serp_probability(te.pidx, te.clicks, te.length)
    np.testing.assert_allclose(got, want, rtol=2e-7, atol=1e-12)
    if model in (DBN, CCM):
        # The observed outcome's probability: exp of the conditional log.
        lp = cm.evaluate(te, per_session=True)["logp"]
        L = te.length if te.length is not None else np.full(len(te), 10)
        shown = np.arange(10)[None, :] < L[:, None]
        np.testing.assert_allclose(np.log(got[shown]), lp[shown], rtol=1e-5, atol=1e-6)
