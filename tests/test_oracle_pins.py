"""Pins of the CPU oracle against enumeration of hidden states and other
facts the mathematics fixes (the reference has no tests or golden vectors of
its own).  Each test names the oracle function it pins.

The DBN gamma and CCM tau factor updates (dbn.cu:436-467, ccm.cu:387-467
with factor.cu:72-142, 206-274) are not the textbook posterior, but each of
their three departures is a fact the enumeration can state:
  * the no-click branch multiplies every (x, y, z) by (1 - a), so the
    factor is the posterior over (E_r, S_r, E_{r+1}) jointly with A_r = c_r;
  * the tail chains score the clicks after rank r with the parameters of
    ranks 0, 1, ... (a fresh list), not of ranks r+1, r+2, ...;
  * a session without clicks has last-click rank 10, which masks every
    z = 0 branch, so each rank contributes exactly (1, 1).
The tests below pin the oracle against enumeration under exactly these
statements, and against the plain posterior when all ranks share their
parameters (where the fresh-list tail is the true tail).
"""
import numpy as np
import pytest

import oracle
from bruteforce import (atr_posterior_no_later_click, conditional_click_probs, conditional_from_patterns,
                        factor_posteriors, factor_posteriors_fresh_tail, independent_joint, marginal_click_probs,
                        pattern_probs, pbm_posteriors, sat_posterior_last_click, ubm_joint,
                        ubm_marginal_click_probs)

PBM, CCM, DBN, UBM = 0, 1, 2, 3


def one_session_model(model, L, clicks, n_tables_vals):
    """An oracle model over a single session of L ranks with distinct pairs."""
    pidx = np.zeros((1, 10), np.int32)
    pidx[0, :L] = np.arange(L)
    c = np.zeros((1, 10), np.uint8)
    c[0, :L] = clicks
    om = oracle.OracleModel(model, np.zeros(1, np.int32), pidx, c, L, length=np.array([L], np.uint8))
    for t, vals in n_tables_vals.items():
        vals = np.asarray(vals, np.float32)
        om.set_table(t, vals, np.ones_like(vals))
    return om, pidx, c


def test_param_value_clamp():
    """Param::value (param.cu:75-87): num/den, clamped at 1 - 1e-6."""
    om, _, _ = one_session_model(PBM, 3, [0, 0, 0], {})
    num = np.array([1, 3, 5, 7, 1, 2, 4, 8, 16, 0], np.float32)
    den = np.array([2, 4, 5, 6, 3, 2, 8, 8, 17, 1], np.float32)
    om.set_table(1, num, den)
    _, _, val = om.get_table(1)
    clamp = np.float32(1.0) - np.float32(1e-6)
    expect = np.where(num / den < clamp, num / den, clamp).astype(np.float32)
    np.testing.assert_array_equal(val, expect)
    assert val[2] == clamp and val[3] == clamp and val[5] == clamp


@pytest.mark.parametrize("seed", range(5))
def test_pbm_estep_is_the_posterior(seed):
    """PBM_Hst::process_session (pbm.cu:191-231): unclicked ranks carry
    P(A=1 | C=0) and P(E=1 | C=0); clicked ranks carry (1, 1)."""
    rng = np.random.default_rng(seed)
    a = rng.uniform(0.05, 0.95, 10)
    e = rng.uniform(0.05, 0.95, 10)
    clicks = rng.integers(0, 2, 10)
    om, _, _ = one_session_model(PBM, 10, clicks, {0: a, 1: e})
    om.estep()
    an, ad = om.get_tmp(0)
    en, ed = om.get_tmp(1)
    for r in range(10):
        pa, pe = pbm_posteriors(np.float32(a[r]), np.float32(e[r]))
        if clicks[r]:
            assert (an[r], ad[r], en[r], ed[r]) == (1, 1, 1, 1)
        else:
            assert an[r] == pytest.approx(pa, rel=2e-6) and ad[r] == 1
            assert en[r] == pytest.approx(pe, rel=2e-6) and ed[r] == 1


def ubm_expected_slot(r, prev):
    """Examination slot of rank r with previous click `prev` (-1 = none).
    Row r of the triangular table starts at r(r+1)/2 (ubm.cu:570-589); the
    'no earlier click' column is 9, so it lands in another row's range."""
    no_click = {0: 9, 1: 10, 2: 12, 3: 15, 4: 19, 5: 24, 6: 30, 7: 37, 8: 45, 9: 54}
    return no_click[r] if prev == -1 else r * (r + 1) // 2 + prev


def test_ubm_slots_and_last_writer_over_all_click_masks():
    """UBM_Hst::process_session (ubm.cu:185-223) over all 1024 click
    patterns: the slot each rank writes, the overwrite of an aliased slot by
    the later rank, the unwritten slots left at (1, 2) and the posterior."""
    n = 1024
    clicks = np.array([[(m >> r) & 1 for r in range(10)] for m in range(n)], np.uint8)
    pidx = np.tile(np.arange(10, dtype=np.int32), (n, 1))
    om = oracle.OracleModel(UBM, np.zeros(n, np.int32), pidx, clicks, 10)
    a = np.linspace(0.1, 0.8, 10).astype(np.float32)
    e = ((np.arange(55) + 1) / 60.0).astype(np.float32)
    om.set_table(0, a, np.ones(10, np.float32))
    om.set_table(1, e, np.ones(55, np.float32))
    om.estep()
    en, ed = om.get_tmp(1)
    en = en.reshape(55, n)
    ed = ed.reshape(55, n)
    for s in range(n):
        prev, written = -1, {}
        for r in range(10):
            written[ubm_expected_slot(r, prev)] = r       # later ranks overwrite
            if clicks[s, r]:
                prev = r
        for k in range(55):
            if k not in written:
                assert (en[k, s], ed[k, s]) == (1, 2), (s, k)
                continue
            r = written[k]
            if clicks[s, r]:
                assert (en[k, s], ed[k, s]) == (1, 1)
            else:
                _, pe = pbm_posteriors(a[r], e[k])
                assert en[k, s] == pytest.approx(pe, rel=2e-6) and ed[k, s] == 1


@pytest.mark.parametrize("model", [DBN, CCM], ids=["DBN", "CCM"])
@pytest.mark.parametrize("L", [1, 2, 3, 4, 5])
def test_conditional_and_full_click_probs_match_enumeration(model, L):
    """get_log_conditional_click_probs / get_full_click_probs (dbn.cu:611-683,
    ccm.cu:604-671) against enumeration of hidden states (length-L lists)."""
    rng = np.random.default_rng(10 * model + L)
    for trial in range(4):
        a = rng.uniform(0.05, 0.9, L).astype(np.float32)
        s = rng.uniform(0.05, 0.9, L).astype(np.float32)
        g = np.float32(rng.uniform(0.3, 0.95))
        taus = rng.uniform(0.2, 0.95, 3).astype(np.float32)
        clicks = rng.integers(0, 2, L)
        if model == DBN:
            om, pidx, c = one_session_model(DBN, L, clicks, {0: a, 1: s, 2: [g]})
            kw = dict(a=a.astype(np.float64), s=s.astype(np.float64), g=float(g))
            name = "dbn"
        else:
            om, pidx, c = one_session_model(CCM, L, clicks, {0: a, 1: taus})
            kw = dict(a=a.astype(np.float64), taus=taus.astype(np.float64))
            name = "ccm"
        ev = om.evaluate(pidx, c, np.array([L], np.uint8))
        cond = conditional_click_probs(name, clicks, **kw)
        np.testing.assert_allclose(np.exp(ev["logp"][0, :L]), cond, rtol=2e-5)
        marg = marginal_click_probs(name, **kw)
        full = np.where(clicks == 1, marg, 1 - marg)
        np.testing.assert_allclose(ev["full"][0, :L], full, rtol=2e-5, atol=1e-7)


def test_ubm_full_click_probs_match_enumeration():
    """UBM_Hst::get_full_click_probs (ubm.cu:412-468): marginal click
    probability over the previous-click rank."""
    rng = np.random.default_rng(3)
    a = rng.uniform(0.05, 0.9, 10).astype(np.float32)
    e = rng.uniform(0.05, 0.95, 55).astype(np.float32)
    om, pidx, c = one_session_model(UBM, 10, np.zeros(10, int), {0: a, 1: e})
    ev = om.evaluate(pidx, c)
    pc = [(lambda prev, r=r: float(a[r]) * float(e[ubm_expected_slot(r, prev)])) for r in range(10)]
    marg = ubm_marginal_click_probs(pc)
    np.testing.assert_allclose(1 - ev["full"][0], marg, rtol=1e-5, atol=1e-6)


def all_masks_model(model, L, tables):
    """An oracle model whose 2^L sessions show the same L pairs (indices
    0..L-1) with every click pattern: session m clicks rank r iff bit r of m
    is set.  Tables are set to the given values with denominator 1."""
    n = 1 << L
    clicks = np.zeros((n, 10), np.uint8)
    clicks[:, :L] = (np.arange(n)[:, None] >> np.arange(L)[None, :]) & 1
    pidx = np.zeros((n, 10), np.int32)
    pidx[:, :L] = np.arange(L)
    length = np.full(n, L, np.uint8)
    om = oracle.OracleModel(model, np.zeros(n, np.int32), pidx, clicks, L, length=length)
    for t, vals in tables.items():
        vals = np.asarray(vals, np.float32)
        om.set_table(t, vals, np.ones_like(vals))
    return om, pidx, clicks, length


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5, 10])
def test_pbm_conditional_and_full_click_probs_match_enumeration(L):
    """PBM_Hst::get_log_conditional_click_probs and get_full_click_probs
    (pbm.cu:376-439) over every click pattern, against enumeration of the
    hidden (A_r, E_r): under the PBM the ranks are independent, so the
    conditional and the full click probability of rank r are both
    a_r e_r (click) or 1 - a_r e_r (no click), with e indexed by rank."""
    rng = np.random.default_rng(40 + L)
    a = rng.uniform(0.05, 0.95, L).astype(np.float32)
    e = rng.uniform(0.05, 0.95, 10).astype(np.float32)
    om, pidx, clicks, length = all_masks_model(PBM, L, {0: a, 1: e})
    ev = om.evaluate(pidx, clicks, length)
    C, p = independent_joint(a.astype(np.float64), e[:L].astype(np.float64))
    cond = conditional_from_patterns(pattern_probs(C, p), L)
    marg = (C * p[:, None]).sum(axis=0)
    c = clicks[:, :L] == 1
    np.testing.assert_allclose(np.exp(ev["logp"][:, :L].astype(np.float64)), cond, rtol=2e-6)
    np.testing.assert_allclose(ev["full"][:, :L], np.where(c, marg[None, :], 1 - marg[None, :]), rtol=2e-6)
    # A pair unseen in training (pidx -1) is valued PARAM_DEF_NUM/DENOM = 1/2.
    p1 = pidx[:1].copy()
    p1[0, L - 1] = -1
    ev1 = om.evaluate(p1, clicks[:1], length[:1])
    assert ev1["full"][0, L - 1] == pytest.approx(1 - 0.5 * float(e[L - 1]), rel=1e-6)


def ubm_conditional_expected(a, e, slot_of, masks):
    """a_r * e[slot_of(r, prev)] for the observed clicks (prev: last click
    before r, -1 = none), as P(C_r = c_r | C_<r): [n_masks, 10]."""
    out = np.zeros((len(masks), 10))
    for i, m in enumerate(masks):
        prev = -1
        for r in range(10):
            c = (m >> r) & 1
            q = float(a[r]) * float(e[slot_of(r, prev, c)])
            out[i, r] = q if c else 1 - q
            if c:
                prev = r
    return out


def test_ubm_conditional_click_probs_over_all_click_masks():
    """UBM_Hst::get_log_conditional_click_probs (ubm.cu:373-401) on all 1024
    click patterns of a ten-rank list, against enumeration of the hidden
    (A_r, E_r) when E_r ~ Bern(e[slot]) with the slot of the explicit table
    (row r(r+1)/2, column = previous click, 9 when there is none).  Two
    mistaken slot rules fail this comparison, which shows that it tells
    them apart: the previous click taken at or before rank r, and column 0
    for "no earlier click"."""
    rng = np.random.default_rng(11)
    a = rng.uniform(0.05, 0.9, 10).astype(np.float32)
    e = rng.uniform(0.05, 0.95, 55).astype(np.float32)
    om, pidx, clicks, length = all_masks_model(UBM, 10, {0: a, 1: e})
    ev = om.evaluate(pidx, clicks, length)
    got = np.exp(ev["logp"].astype(np.float64))
    ef = e.astype(np.float64)
    C, p = ubm_joint(a.astype(np.float64), lambda r, prev: ef[ubm_expected_slot(r, prev)])
    cond = conditional_from_patterns(pattern_probs(C, p), 10)
    np.testing.assert_allclose(got, cond, rtol=2e-6)
    masks = np.arange(1024)
    right = ubm_conditional_expected(a, e, lambda r, prev, c: ubm_expected_slot(r, prev), masks)
    np.testing.assert_allclose(right, cond, rtol=1e-12)
    inclusive = ubm_conditional_expected(
        a, e, lambda r, prev, c: r * (r + 1) // 2 + (r if c else (9 if prev == -1 else prev)), masks)
    column0 = ubm_conditional_expected(
        a, e, lambda r, prev, c: r * (r + 1) // 2 + (0 if prev == -1 else prev), masks)
    for wrong in (inclusive, column0):
        bad = np.abs(wrong - got) > 1e-3 * got
        assert bad.any(axis=1).mean() > 0.5


def test_ubm_full_click_probs_over_all_click_masks():
    """UBM_Hst::get_full_click_probs (ubm.cu:412-468) on all 1024 click
    patterns: P(C_r = 1) marginalised over every click history (from the
    enumeration above), or one minus it for an unclicked rank."""
    rng = np.random.default_rng(12)
    a = rng.uniform(0.05, 0.9, 10).astype(np.float32)
    e = rng.uniform(0.05, 0.95, 55).astype(np.float32)
    om, pidx, clicks, length = all_masks_model(UBM, 10, {0: a, 1: e})
    ev = om.evaluate(pidx, clicks, length)
    ef = e.astype(np.float64)
    C, p = ubm_joint(a.astype(np.float64), lambda r, prev: ef[ubm_expected_slot(r, prev)])
    marg = (C * p[:, None]).sum(axis=0)
    c = clicks == 1
    np.testing.assert_allclose(ev["full"], np.where(c, marg[None, :], 1 - marg[None, :]), rtol=1e-5, atol=1e-6)


@pytest.mark.parametrize("model", [PBM, UBM, DBN, CCM], ids=["PBM", "UBM", "DBN", "CCM"])
def test_serp_probability_matches_enumeration(model):
    """get_serp_probability (pbm.cu:351-365, ubm.cu:345-362, dbn.cu:572-600,
    ccm.cu:566-593) on every click pattern, against enumeration of hidden
    states: for the PBM and UBM it is P(C_r = 1 | the earlier clicks)
    whatever rank r's own click (ten ranks, 1024 patterns); for the DBN and
    CCM it is P(C_r = c_r | the earlier clicks), the observed outcome's
    conditional probability (five ranks, 32 patterns)."""
    from bruteforce import cascade_joint
    rng = np.random.default_rng(60 + model)
    if model in (PBM, UBM):
        L = 10
        a = rng.uniform(0.05, 0.9, L).astype(np.float32)
        e = rng.uniform(0.05, 0.95, 10 if model == PBM else 55).astype(np.float32)
        ef = e.astype(np.float64)
        if model == PBM:
            C, p = independent_joint(a.astype(np.float64), ef)
        else:
            C, p = ubm_joint(a.astype(np.float64), lambda r, prev: ef[ubm_expected_slot(r, prev)])
        cond = conditional_from_patterns(pattern_probs(C, p), L)
        masks = np.arange(1 << L)
        want = np.stack([cond[masks | (1 << r), r] for r in range(L)], axis=1)
        tables = {0: a, 1: e}
    else:
        L = 5
        a = rng.uniform(0.05, 0.9, L).astype(np.float32)
        s = rng.uniform(0.05, 0.9, L).astype(np.float32)
        g = np.float32(rng.uniform(0.3, 0.95))
        taus = rng.uniform(0.2, 0.95, 3).astype(np.float32)
        if model == DBN:
            C, _, _, _, p = cascade_joint("dbn", a=a.astype(np.float64), s=s.astype(np.float64), g=float(g))
            tables = {0: a, 1: s, 2: [g]}
        else:
            C, _, _, _, p = cascade_joint("ccm", a=a.astype(np.float64), taus=taus.astype(np.float64))
            tables = {0: a, 1: taus}
        want = conditional_from_patterns(pattern_probs(C, p), L)
    om, pidx, clicks, length = all_masks_model(model, L, tables)
    got = om.serp_probability(pidx, clicks, length)
    np.testing.assert_allclose(got[:, :L], want, rtol=1e-5)
    assert np.all(got[:, L:] == 0)


@pytest.mark.parametrize("model", [PBM, UBM, DBN, CCM], ids=["PBM", "UBM", "DBN", "CCM"])
def test_session_mean_and_totals_follow_evaluation(model):
    """LogLikelihood::evaluate and Perplexity::evaluate
    (evaluation.cpp:25-40, 49-66): a session's log-likelihood is the float
    sum of its ranks' log conditional probabilities, from 0.f in rank order,
    divided by its length (10 in the reference); the FP32 total adds the
    sessions in order in float, the FP64 total in double; the perplexity
    sums add log2 of the full click probability per rank.  Checked against
    a recomputation from the per-rank outputs, for uneven lengths and
    unseen pairs."""
    rng = np.random.default_rng(50 + model)
    n, n_qd = 400, 50
    length = rng.integers(1, 11, n).astype(np.uint8)
    pidx = rng.integers(0, n_qd, (n, 10)).astype(np.int32)
    pidx[rng.random((n, 10)) < 0.1] = -1
    clicks = (rng.random((n, 10)) < 0.3).astype(np.uint8)
    for i in range(n):
        clicks[i, length[i]:] = 0
    tables = {0: rng.uniform(0.05, 0.9, n_qd)}
    K = {PBM: 10, UBM: 55, DBN: 1, CCM: 3}[model]
    if model == DBN:
        tables[1] = rng.uniform(0.05, 0.9, n_qd)
    tables[1 if model != DBN else 2] = rng.uniform(0.2, 0.95, K)
    om = oracle.OracleModel(model, np.zeros(1, np.int32), np.zeros((1, 10), np.int32), np.zeros((1, 10), np.uint8),
                            n_qd)
    for t, v in tables.items():
        v = np.asarray(v, np.float32)
        om.set_table(t, v, np.ones_like(v))
    ev = om.evaluate(pidx, clicks, length)
    tot32 = np.float32(0)
    tot64 = 0.0
    pr64 = np.zeros(10)
    for i in range(n):
        L = int(length[i])
        acc = np.float32(0)
        for r in range(L):
            acc = np.float32(acc + ev["logp"][i, r])
        s = np.float32(acc / np.float32(L))
        assert ev["llh"][i] == s, i
        tot32 = np.float32(tot32 + s)
        tot64 += float(s)
        pr64[:L] += np.log2(ev["full"][i, :L].astype(np.float64))
        assert np.all(ev["logp"][i, L:] == 0) and np.all(ev["full"][i, L:] == 0)
    assert ev["llh_sum_f32"] == tot32
    assert ev["llh_sum_f64"] == tot64
    np.testing.assert_allclose(ev["ppl_log2_sum_f64"], pr64, rtol=1e-6)
    np.testing.assert_array_equal(ev["rank_count"], [(length > r).sum() for r in range(10)])


def _random_clicks_with_last(rng, L, lcr):
    c = np.zeros(L, int)
    c[:lcr] = rng.integers(0, 2, lcr)
    c[lcr] = 1
    return c


@pytest.mark.parametrize("model", [DBN, CCM], ids=["DBN", "CCM"])
@pytest.mark.parametrize("L", [2, 3, 4, 5])
def test_atr_after_last_click_is_posterior_given_no_later_click(model, L):
    """compute_dbn_atr (dbn.cu:312-336) / compute_ccm_atr (ccm.cu:295-326)
    for unclicked ranks after the last click: a(1-exam)/(1-exam car) is
    P(A_r = 1 | no click at ranks >= r), the earlier ranks marginalised."""
    rng = np.random.default_rng(100 * model + L)
    for trial in range(3):
        lcr = int(rng.integers(0, L - 1))
        clicks = _random_clicks_with_last(rng, L, lcr)
        a = rng.uniform(0.05, 0.9, L).astype(np.float32)
        s = rng.uniform(0.05, 0.9, L).astype(np.float32)
        g = np.float32(rng.uniform(0.3, 0.95))
        taus = rng.uniform(0.2, 0.95, 3).astype(np.float32)
        if model == DBN:
            om, _, _ = one_session_model(DBN, L, clicks, {0: a, 1: s, 2: [g]})
            kw = dict(a=a.astype(np.float64), s=s.astype(np.float64), g=float(g))
            name = "dbn"
        else:
            om, _, _ = one_session_model(CCM, L, clicks, {0: a, 1: taus})
            kw = dict(a=a.astype(np.float64), taus=taus.astype(np.float64))
            name = "ccm"
        om.estep()
        an, ad = om.get_tmp(0)
        for r in range(lcr + 1, L):
            expect = atr_posterior_no_later_click(name, r, **kw)
            assert an[r] == pytest.approx(expect, rel=2e-5), (r, lcr)
            assert ad[r] == 1
        for r in range(lcr):        # before the last click: (c_r, 1 + c_r for CCM)
            assert an[r] == (1 if clicks[r] else 0)


@pytest.mark.parametrize("L", [1, 2, 3, 4, 5])
def test_dbn_satisfaction_at_last_click_is_posterior(L):
    """compute_dbn_sat (dbn.cu:350-378): at the last click s/(1-(1-s)g car) =
    P(S = 1 | the observed clicks); other clicked ranks carry (0, 1) and
    unclicked ranks are never written (they keep (1, 2))."""
    rng = np.random.default_rng(7 + L)
    for trial in range(3):
        lcr = int(rng.integers(0, L))
        clicks = _random_clicks_with_last(rng, L, lcr)
        a = rng.uniform(0.05, 0.9, L).astype(np.float32)
        s = rng.uniform(0.05, 0.9, L).astype(np.float32)
        g = np.float32(rng.uniform(0.3, 0.95))
        om, _, _ = one_session_model(DBN, L, clicks, {0: a, 1: s, 2: [g]})
        om.estep()
        sn, sd = om.get_tmp(1)
        expect = sat_posterior_last_click(lcr, clicks, a=a.astype(np.float64), s=s.astype(np.float64), g=float(g))
        assert sn[lcr] == pytest.approx(expect, rel=2e-5)
        for r in range(L):
            if r == lcr:
                continue
            assert (sn[r], sd[r]) == ((0, 1) if clicks[r] else (1, 2))


def _shared_factor_tmp(model, L, clicks, a, s, g, taus):
    """Runs one E-step of the oracle on one session; returns the session's
    gamma (DBN) or tau1..tau3 (CCM) temporaries as [(num, den), ...]."""
    if model == DBN:
        om, _, _ = one_session_model(DBN, L, clicks, {0: a, 1: s, 2: [g]})
        om.estep()
        n, d = om.get_tmp(2)
        return [(n[0], d[0])]
    om, _, _ = one_session_model(CCM, L, clicks, {0: a, 1: taus})
    om.estep()
    n, d = om.get_tmp(1)
    return [(n[k], d[k]) for k in range(3)]


def _expected_from_weights(model, clicks, w):
    """Sums the factor posteriors as compute_gamma / compute_taus do."""
    if model == DBN:
        num = sum(w[r][1, 0, 1] for r in range(len(clicks)))
        den = sum(w[r][1, 0, 0] + w[r][1, 0, 1] for r in range(len(clicks)))
        return [(num, den)]
    t = np.zeros((3, 2))
    for r, c in enumerate(clicks):
        if c == 0:
            t[0] += (w[r][1, :, 1].sum(), w[r][1].sum())
        else:
            t[1] += (w[r][1, 0, 1], w[r][1, 0].sum())
            t[2] += (w[r][1, 1, 1], w[r][1, 1].sum())
    return [tuple(x) for x in t]


def _draw(rng, L, uniform):
    k = 1 if uniform else L
    a = np.resize(rng.uniform(0.05, 0.9, k), L).astype(np.float32)
    s = np.resize(rng.uniform(0.05, 0.9, k), L).astype(np.float32)
    g = np.float32(rng.uniform(0.3, 0.95))
    taus = rng.uniform(0.2, 0.95, 3).astype(np.float32)
    return a, s, g, taus


def _kw(model, a, s, g, taus):
    if model == DBN:
        return dict(a=a.astype(np.float64), s=s.astype(np.float64), g=float(g))
    return dict(a=a.astype(np.float64), taus=taus.astype(np.float64))


@pytest.mark.parametrize("model", [DBN, CCM], ids=["DBN", "CCM"])
@pytest.mark.parametrize("L", [1, 2, 3, 4, 5])
def test_shared_factor_update_is_posterior_when_ranks_share_parameters(model, L):
    """compute_gamma (dbn.cu:436-467) / compute_taus (ccm.cu:387-467) with
    DBNFactor / CCMFactor (factor.cu:206-274, 72-142): with one attractiveness
    (and satisfaction) value at every rank, the update summed over ranks is
    the posterior of (E_r, S_r, E_{r+1}) given the clicks and A_r = c_r, over
    every click pattern with at least one click."""
    name = "dbn" if model == DBN else "ccm"
    rng = np.random.default_rng(31 * model + L)
    for pattern in range(1, 2 ** L):
        clicks = np.array([(pattern >> r) & 1 for r in range(L)])
        a, s, g, taus = _draw(rng, L, uniform=True)
        got = _shared_factor_tmp(model, L, clicks, a, s, g, taus)
        want = _expected_from_weights(model, clicks, factor_posteriors(name, clicks, **_kw(model, a, s, g, taus)))
        for (gn, gd), (wn, wd) in zip(got, want):
            assert gn == pytest.approx(wn, rel=3e-5, abs=3e-6), (pattern, got, want)
            assert gd == pytest.approx(wd, rel=3e-5, abs=3e-6), (pattern, got, want)


@pytest.mark.parametrize("model", [DBN, CCM], ids=["DBN", "CCM"])
@pytest.mark.parametrize("L", [2, 3, 4, 5])
def test_shared_factor_update_scores_the_tail_as_a_fresh_list(model, L):
    """The same update with a different value at every rank: the clicks
    after rank r are scored with the parameters of ranks 0, 1, ...
    (get_tail_clicks, dbn.cu:390-423, ccm.cu:338-374).  An oracle that used
    the parameters of ranks r+1, r+2, ... fails this test."""
    name = "dbn" if model == DBN else "ccm"
    rng = np.random.default_rng(57 * model + L)
    for pattern in range(1, 2 ** L):
        clicks = np.array([(pattern >> r) & 1 for r in range(L)])
        a, s, g, taus = _draw(rng, L, uniform=False)
        got = _shared_factor_tmp(model, L, clicks, a, s, g, taus)
        w = factor_posteriors_fresh_tail(name, clicks, **_kw(model, a, s, g, taus))
        want = _expected_from_weights(model, clicks, w)
        for (gn, gd), (wn, wd) in zip(got, want):
            assert gn == pytest.approx(wn, rel=3e-5, abs=3e-6), (pattern, got, want)
            assert gd == pytest.approx(wd, rel=3e-5, abs=3e-6), (pattern, got, want)


@pytest.mark.parametrize("model", [DBN, CCM], ids=["DBN", "CCM"])
@pytest.mark.parametrize("L", [1, 4, 10])
def test_shared_factor_update_of_a_session_without_clicks(model, L):
    """No click: last-click rank 10 masks every z = 0 branch, so each rank
    adds exactly (1, 1) to gamma / tau1 and nothing to tau2, tau3."""
    rng = np.random.default_rng(L)
    a, s, g, taus = _draw(rng, L, uniform=False)
    got = _shared_factor_tmp(model, L, np.zeros(L, int), a, s, g, taus)
    want = [(L, L)] if model == DBN else [(L, L), (0, 0), (0, 0)]
    for (gn, gd), (wn, wd) in zip(got, want):
        assert gn == pytest.approx(wn, rel=1e-6) and gd == pytest.approx(wd, rel=1e-6)


@pytest.mark.parametrize("model", [PBM, UBM, DBN, CCM], ids=["PBM", "UBM", "DBN", "CCM"])
def test_mstep_is_prior_plus_sum_of_temporaries(model):
    """update_unique/shared_parameters_hst (common.cu:114-192) after
    reset_parameters (common.cu:71-82): theta = (1, 2) + sum of the E-step's
    temporary pairs (per-qd: over occurrences; shared: over all K slots of
    every session, including never-written (1, 2) slots)."""
    from massiveclicks_amd import synth
    name = {PBM: "pbm", UBM: "ubm", DBN: "dbn", CCM: "ccm"}[model]
    s = synth.generate(synth.SynthConfig(n_sessions=600, n_queries=60, vocab=2000, model=name, seed=5))
    ds = oracle.from_sessions(s.query, s.docs, s.clicks, 0.2)
    tr = ds["train"]
    N = len(tr["query"])
    om = oracle.OracleModel(model, tr["query"], tr["pidx"], tr["clicks"], ds["n_qd"], accum_fp64=True)
    om.iterate(2)
    om.estep()
    tmps = [om.get_tmp(t) for t in range(om.n_tables())]
    om.mstep()
    for t in range(om.n_tables()):
        num, den, _ = om.get_table(t)
        tn, td = (x.astype(np.float64) for x in tmps[t])
        K = len(num)
        if K == ds["n_qd"]:
            en = np.ones(K) + np.bincount(tr["pidx"].ravel(), weights=tn.reshape(10, N).T.ravel(), minlength=K)
            ed = 2 * np.ones(K) + np.bincount(tr["pidx"].ravel(), weights=td.reshape(10, N).T.ravel(), minlength=K)
        else:
            en = 1 + tn.reshape(K, N).sum(axis=1)
            ed = 2 + td.reshape(K, N).sum(axis=1)
        np.testing.assert_allclose(num, en, rtol=1e-6)
        np.testing.assert_allclose(den, ed, rtol=1e-6)


@pytest.mark.parametrize("model", [PBM, DBN], ids=["PBM", "DBN"])
def test_query_aligned_threads_give_identical_per_qd_tables(model):
    """The CPU path's threads own whole query groups in the M-step
    (parallel_em.cu:27-197), so per-qd sums keep their session order: the
    per-qd tables of one iteration are bit-identical for any thread count;
    the shared table differs only by the merge order of the thread partials
    (which then feeds every later E-step)."""
    from massiveclicks_amd import synth
    name = {PBM: "pbm", DBN: "dbn"}[model]
    s = synth.generate(synth.SynthConfig(n_sessions=3000, n_queries=300, vocab=9000, model=name, seed=6))
    ds = oracle.from_sessions(s.query, s.docs, s.clicks, 0.2)
    tr = ds["train"]
    outs = []
    for th in (1, 4):
        om = oracle.OracleModel(model, tr["query"], tr["pidx"], tr["clicks"], ds["n_qd"], n_threads=th)
        om.iterate(1)
        outs.append([om.get_table(t) for t in range(om.n_tables())])
    for t in range(len(outs[0]) - 1):
        np.testing.assert_array_equal(outs[0][t][0], outs[1][t][0])
    # The merge order of FP32 partials alone moves the shared table by about
    # 1.5e-5 relative at 2,400 sessions: the reference's own run-to-run spread.
    np.testing.assert_allclose(outs[0][-1][2], outs[1][-1][2], rtol=1e-4)


@pytest.mark.parametrize("name", ["dbn", "ccm"])
def test_fisher_terms_helper_is_the_logit_gradient(name):
    """The enumeration helper used to check the GPU gradient: its expected
    counts minus expectations equal float64 central differences of
    log P(clicks) with respect to each logit."""
    from bruteforce import cascade_joint, fisher_terms, pattern_prob
    rng = np.random.default_rng(9)
    L = 4
    a = rng.uniform(0.1, 0.9, L)
    s = rng.uniform(0.1, 0.9, L)
    g = 0.7
    taus = np.array([0.8, 0.5, 0.3])
    clicks = np.array([0, 1, 0, 1])

    def logp(a, s, g, taus):
        C, _, _, _, p = cascade_joint(name, a=a, s=s, g=g, taus=taus)
        return np.log(pattern_prob(C, p, clicks))

    def logit_shift(v, d):
        x = np.log(v / (1 - v)) + d
        return 1 / (1 + np.exp(-x))

    da, dsat, dsh = fisher_terms(name, clicks, a=a, s=s, g=g, taus=taus)
    eps = 1e-5
    for r in range(L):
        for arr, want in ((a, da[r]), (s, dsat[r])):
            hi, lo = arr.copy(), arr.copy()
            hi[r] = logit_shift(arr[r], eps)
            lo[r] = logit_shift(arr[r], -eps)
            args_hi = (hi, s, g, taus) if arr is a else (a, hi, g, taus)
            args_lo = (lo, s, g, taus) if arr is a else (a, lo, g, taus)
            fd = (logp(*args_hi) - logp(*args_lo)) / (2 * eps)
            if name == "ccm" and arr is s:
                continue
            assert fd == pytest.approx(want, abs=1e-6), (r, arr is a)
    if name == "dbn":
        fd = (logp(a, s, logit_shift(g, eps), taus) - logp(a, s, logit_shift(g, -eps), taus)) / (2 * eps)
        assert fd == pytest.approx(dsh[0], abs=1e-6)
    else:
        for k in range(3):
            hi, lo = taus.copy(), taus.copy()
            hi[k] = logit_shift(taus[k], eps)
            lo[k] = logit_shift(taus[k], -eps)
            fd = (logp(a, s, g, hi) - logp(a, s, g, lo)) / (2 * eps)
            assert fd == pytest.approx(dsh[k], abs=1e-6), k
