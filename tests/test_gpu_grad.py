"""Gradient path on the GPU: the log-likelihood gradient kernels against
float64 central finite differences (eps 1e-4) through a torch.autograd
wrapper, against the Fisher identity (gradient = E-step expected counts
minus model expectations, from one EM step at the same parameters), over
minibatches, and the Adam fit."""
import numpy as np
import pytest
import torch

from llh64 import loglik64, sigmoid64
from massiveclicks_amd import CCM, DBN, PBM, UBM, ClickModel, Dataset, synth

pytestmark = pytest.mark.gpu

NAMES = {PBM: "pbm", UBM: "ubm", DBN: "dbn", CCM: "ccm"}
_ds = {}


def dataset(model, n=5000, varlen=False):
    key = (model, n, varlen)
    if key not in _ds:
        s = synth.generate(synth.SynthConfig(n_sessions=n, n_queries=n // 10, vocab=20 * n, model=NAMES[model],
                                             variable_length=varlen, seed=500 + model))
        _ds[key] = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2, length=s.length if varlen else None)
    return _ds[key]


class LogLik(torch.autograd.Function):
    """Training log-likelihood as a function of the logit tables; forward
    and backward run the engine's gradient kernels (tests only)."""

    @staticmethod
    def forward(ctx, cm, *logits):
        for t, x in enumerate(logits):
            cm.set_logits(t, x.detach().cpu().numpy())
        ll, grads = cm.loglik_gradient()
        ctx.grads = [torch.from_numpy(g.astype(np.float64)) for g in grads]
        return torch.tensor(ll, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        return (None, *[g * x for x in ctx.grads])


def random_logits(cm, rng):
    return [np.clip(rng.normal(0.0, 1.2, size), -6, 6).astype(np.float32)
            for _, _, size in cm.get_parameter_information()]


@pytest.mark.parametrize("model", [PBM, UBM, DBN, CCM], ids=["PBM", "UBM", "DBN", "CCM"])
@pytest.mark.parametrize("varlen", [False, True], ids=["L10", "L1-10"])
def test_gradient_matches_float64_central_differences(model, varlen):
    ds = dataset(model, varlen=varlen)
    tr = ds.train
    rng = np.random.default_rng(model * 7 + varlen)
    print(f"case grad-fd model={model} varlen={varlen} n={ds.n_train} n_qd={ds.n_qd}", flush=True)
    with ClickModel(model) as cm:
        cm.init_from_dataset(ds)
        cm.grad_init()
        logits = random_logits(cm, rng)
        xs = [torch.tensor(x, dtype=torch.float32, requires_grad=True) for x in logits]
        ll = LogLik.apply(cm, *xs)
        ll.backward()
        got = [x.grad.numpy() for x in xs]
        vals32 = [cm.get_values(t).astype(np.float64) for t in range(len(logits))]

    T = len(logits) - 1
    v64 = [sigmoid64(x) for x in logits]
    ll64 = loglik64(model, v64[:T], v64[T], tr.pidx, tr.clicks, tr.length)
    assert abs(ll.item() - ll64) <= 1e-5 * abs(ll64), (ll.item(), ll64)

    # sessions of each entry
    flat = tr.pidx.reshape(-1)
    sess = np.repeat(np.arange(len(tr.query)), 10)
    ok = flat >= 0
    order = np.argsort(flat[ok], kind="stable")
    ent, ss = flat[ok][order], sess[ok][order]
    starts = np.searchsorted(ent, np.arange(ds.n_qd + 1))
    eps = 1e-4
    n_fd = 1000 if model == DBN else 300
    checks = [(T, k) for k in range(len(logits[T]))]
    per_qd = [(t, int(p)) for t in range(T) for p in rng.choice(ds.n_qd, n_fd // T, replace=False)]
    worst = 0.0
    for t, p in checks + per_qd:
        if t == T:
            idx = np.arange(len(tr.query))
        else:
            idx = np.unique(ss[starts[p]:starts[p + 1]])
        sub = (tr.pidx[idx], tr.clicks[idx], None if tr.length is None else tr.length[idx])
        f = []
        for d in (eps, -eps):
            x = [v.copy() for v in logits]
            x64 = [xx.astype(np.float64) for xx in x]
            x64[t][p] += d
            v = [sigmoid64(xx) for xx in x64]
            f.append(loglik64(model, v[:T], v[T], *sub))
        fd = (f[0] - f[1]) / (2 * eps)
        g = float(got[t][p])
        err = abs(g - fd) / (abs(fd) + 1e-2 * np.sqrt(len(idx)) + 1e-3)
        worst = max(worst, err)
        assert err < 2e-3, (t, p, g, fd, len(idx))
    print(f"  worst scaled |grad - fd| = {worst:.2e}", flush=True)


def test_fisher_identity_against_one_em_step_pbm():
    """d LL / d logit(theta) = sum over occurrences of (E[hidden | clicks] -
    theta): the E-step's expected counts minus the model expectations.  The
    PBM E-step (pbm.cu:191-231) is the exact posterior, so one EM step from
    the same values gives the counts as num - 1 and the occurrences as
    den - 2."""
    model = PBM
    ds = dataset(model)
    rng = np.random.default_rng(3)
    with ClickModel(model) as cm:
        cm.init_from_dataset(ds)
        cm.grad_init()
        logits = random_logits(cm, rng)
        for t, x in enumerate(logits):
            cm.set_logits(t, x)
        theta = [cm.get_values(t).astype(np.float64) for t in range(len(logits))]
        _, grads = cm.loglik_gradient()
        cm.process_session()
        cm.update_parameters()
        nd = [cm.get_parameters(t) for t in range(len(logits))]
    for t in range(len(logits)):
        num, den = (x.astype(np.float64) for x in nd[t])
        want = (num - 1) - theta[t] * (den - 2)
        tol = 2e-5 * np.maximum(den, 1) + 1e-4
        bad = np.abs(grads[t] - want) > tol
        assert not bad.any(), (t, np.flatnonzero(bad)[:5], grads[t][bad][:5], want[bad][:5])


@pytest.mark.parametrize("model", [DBN, CCM], ids=["DBN", "CCM"])
def test_fisher_identity_against_enumerated_posteriors(model):
    """DBN/CCM: the gradient equals the exact posterior expected counts minus
    the model expectations, enumerated over every hidden (A, S, G)
    assignment of each session (lists of 1-5 ranks).  The reference's EM
    E-step is not this posterior (its examination probabilities ignore the
    clicks before a rank), so the identity is checked against enumeration."""
    from bruteforce import fisher_terms
    rng = np.random.default_rng(40 + model)
    n, nq = 400, 40
    q = rng.integers(0, nq, n).astype(np.int32)
    shown = np.array([rng.permutation(8)[:5] for _ in range(n)])      # 5 distinct pool slots
    docs = np.concatenate([q[:, None] * 8 + shown, np.full((n, 5), -1)], axis=1).astype(np.int32)
    length = rng.integers(1, 6, n).astype(np.uint8)
    clicks = (rng.random((n, 10)) < 0.3).astype(np.uint8) * (np.arange(10)[None, :] < length[:, None])
    ds = Dataset.from_sessions(q, docs, clicks.astype(np.uint8), 0.0, length=length)
    tr = ds.train
    with ClickModel(model) as cm:
        cm.init_from_dataset(ds)
        cm.grad_init()
        logits = random_logits(cm, rng)
        for t, x in enumerate(logits):
            cm.set_logits(t, x)
        vals = [cm.get_values(t).astype(np.float64) for t in range(len(logits))]
        _, grads = cm.loglik_gradient()
    T = len(logits) - 1
    want = [np.zeros(ds.n_qd) for _ in range(T)] + [np.zeros(len(vals[T]))]
    for i in range(len(tr.query)):
        L = int(tr.length[i])
        p = tr.pidx[i, :L]
        c = tr.clicks[i, :L]
        if model == DBN:
            da, dsat, dsh = fisher_terms("dbn", c, a=vals[0][p], s=vals[1][p], g=float(vals[2][0]))
            np.add.at(want[1], p, dsat)
        else:
            da, _, dsh = fisher_terms("ccm", c, a=vals[0][p], taus=vals[1])
        np.add.at(want[0], p, da)
        want[T] += dsh
    for t in range(T + 1):
        np.testing.assert_allclose(grads[t], want[t], rtol=1e-4, atol=2e-4, err_msg=f"table {t}")


@pytest.mark.parametrize("model", [PBM, UBM, DBN, CCM], ids=["PBM", "UBM", "DBN", "CCM"])
def test_minibatch_gradients_add_up_to_the_full_gradient(model):
    ds = dataset(model)
    with ClickModel(model, batches=4) as cm:
        cm.init_from_dataset(ds)
        cm.grad_init()
        rng = np.random.default_rng(1)
        for t, x in enumerate(random_logits(cm, rng)):
            cm.set_logits(t, x)
        ll, full = cm.loglik_gradient(-1)
        parts = [cm.loglik_gradient(b) for b in range(4)]
    assert abs(sum(p[0] for p in parts) - ll) <= 1e-9 * abs(ll)
    T = len(full) - 1
    for t in range(T):
        s = sum(p[1][t] for p in parts)
        np.testing.assert_array_equal(s, full[t])     # each entry belongs to one batch
        assert all(((p[1][t] != 0).astype(int)).sum() <= (full[t] != 0).sum() for p in parts)
    np.testing.assert_allclose(sum(p[1][T].astype(np.float64) for p in parts), full[T], rtol=1e-5, atol=1e-3)


@pytest.mark.parametrize("batches", [1, 3])
def test_adam_increases_the_objective_and_is_deterministic(batches):
    """DBN: Adam from the (1,2) start raises the training log-likelihood
    plus log-prior; graph replay and eager steps give identical tables, and
    so do two runs."""
    model = DBN
    ds = dataset(model)
    tr = ds.train
    occ = np.bincount(tr.pidx.reshape(-1), minlength=ds.n_qd).astype(np.float64)
    clk = np.bincount(tr.pidx.reshape(-1), weights=tr.clicks.reshape(-1), minlength=ds.n_qd)

    def objective(vals):
        ll = loglik64(model, vals[:2], vals[2], tr.pidx, tr.clicks)
        w = [np.ones(ds.n_qd), 1 + occ - clk, np.ones(1)]
        return ll + sum(float(np.sum(wi * (np.log(v) + np.log(1 - v)))) for wi, v in zip(w, vals))

    runs = []
    for use_graph in (True, True, False):
        with ClickModel(model, batches=batches) as cm:
            cm.init_from_dataset(ds)
            cm.grad_init()
            cm.adam_configure(lr=0.05)
            v0 = [cm.get_values(t).astype(np.float64) for t in range(3)]
            cm.adam_steps(60 * batches, use_graph=use_graph)
            runs.append([cm.get_logits(t) for t in range(3)])
            v1 = [cm.get_values(t).astype(np.float64) for t in range(3)]
    for other in runs[1:]:
        for a, b in zip(runs[0], other):
            np.testing.assert_array_equal(a, b)
    f0, f1 = objective(v0), objective(v1)
    print(f"objective {f0:.3f} -> {f1:.3f}", flush=True)
    assert f1 > f0 + 0.05 * abs(f0)
