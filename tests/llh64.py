"""Float64 click log-likelihood of a set of sessions, the function the
gradient kernels differentiate: the sum over sessions and shown ranks of the
log conditional click probabilities of the reference's evaluation
(pbm.cu:376-401, ubm.cu:373-401, dbn.cu:611-639, ccm.cu:604-630), with the
tables given as values.  Test infrastructure for finite differences."""
import numpy as np

PBM, CCM, DBN, UBM = 0, 1, 2, 3


def loglik64(model, vals, shared, pidx, clicks, length=None):
    """vals: per-qd value tables (float64); shared: the rank-indexed table.
    Returns the total log-likelihood (float64)."""
    n = pidx.shape[0]
    L = np.full(n, 10) if length is None else length.astype(np.int64)
    ll = np.zeros(n)
    c = clicks.astype(bool)
    if model in (PBM, UBM):
        prev = np.full(n, 9)
        for r in range(10):
            act = r < L
            p = np.where(act, pidx[:, r], 0)
            a = vals[0][p]
            e = shared[r] if model == PBM else shared[r * (r + 1) // 2 + prev]
            pr = np.where(c[:, r], a * e, 1 - a * e)
            ll += np.where(act, np.log(np.where(act, pr, 1.0)), 0.0)
            prev = np.where(act & c[:, r], r, prev)
        return float(ll.sum())
    ex = np.ones(n)
    for r in range(10):
        act = r < L
        p = np.where(act, pidx[:, r], 0)
        a = vals[0][p]
        pr = np.where(c[:, r], a * ex, 1 - a * ex)
        if model == DBN:
            s = vals[1][p]
            nxt = np.where(c[:, r], shared[0] * (1 - s), ex * shared[0] * (1 - a) / pr)
        else:
            nxt = np.where(c[:, r], shared[1] * (1 - a) + shared[2] * a, ex * shared[0] * (1 - a) / pr)
        ll += np.where(act, np.log(np.where(act, pr, 1.0)), 0.0)
        ex = np.where(act, nxt, ex)
    return float(ll.sum())


def sigmoid64(x):
    return 1.0 / (1.0 + np.exp(-np.asarray(x, np.float64)))
