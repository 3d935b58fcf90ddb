"""Brute-force enumeration of click-model hidden states (float64).

Each function enumerates every assignment of the hidden binary variables of
a short result list (L <= 5) under the model's generative story and sums
probabilities exactly.  These are independent of the oracle's recursions and
are used to pin them.

Generative stories (one query session, ranks 0..L-1):
  PBM/UBM  C_r = A_r * E_r with A_r ~ Bern(a_r), E_r ~ Bern(e_r) independent.
  DBN      E_0 = 1; C_r = E_r A_r; after a click S_r ~ Bern(s_r) and
           E_{r+1} = (1 - S_r) G_r, otherwise E_{r+1} = E_r G_r, G_r ~ Bern(g).
  CCM      E_0 = 1; C_r = E_r A_r; S_r ~ Bern(a_r) (satisfaction, used only
           after a click); after a click E_{r+1} = G_r with G_r ~ Bern(t3) if
           S_r else Bern(t2), otherwise E_{r+1} = E_r G_r with G_r ~ Bern(t1).
"""
import itertools

import numpy as np


def _assignments(n_bits):
    """Every 0/1 vector of n_bits entries, in itertools.product order (the
    last entry varies fastest)."""
    i = np.arange(1 << n_bits, dtype=np.int64)[:, None]
    return ((i >> (n_bits - 1 - np.arange(n_bits))[None, :]) & 1).astype(np.int8)


def cascade_joint(model, a, s=None, g=None, taus=None):
    """Enumerates hidden states; returns (clicks [M, L], A [M, L], E [M, L+1], S [M, L], prob [M])."""
    L = len(a)
    H = _assignments(3 * L)          # per rank: A, S (DBN) / unused (CCM), coin
    A = H[:, 0:L]
    Sx = H[:, L:2 * L]
    G = H[:, 2 * L:3 * L]
    M = H.shape[0]
    E = np.zeros((M, L + 1), np.int8)
    C = np.zeros((M, L), np.int8)
    p = np.ones(M)
    E[:, 0] = 1
    for r in range(L):
        p *= np.where(A[:, r] == 1, a[r], 1 - a[r])
        C[:, r] = E[:, r] * A[:, r]
        if model == "dbn":
            p *= np.where(Sx[:, r] == 1, s[r], 1 - s[r])
            p *= np.where(G[:, r] == 1, g, 1 - g)
            E[:, r + 1] = np.where(C[:, r] == 1, (1 - Sx[:, r]) * G[:, r], E[:, r] * G[:, r])
        else:
            t1, t2, t3 = taus
            p *= np.where(Sx[:, r] == 1, a[r], 1 - a[r])
            h = np.where(Sx[:, r] == 1, t3, t2)
            cont = np.where(C[:, r] == 1, np.where(G[:, r] == 1, h, 1 - h), np.where(G[:, r] == 1, t1, 1 - t1))
            p *= cont
            E[:, r + 1] = np.where(C[:, r] == 1, G[:, r], E[:, r] * G[:, r])
    return C, A, E, Sx, p


def pattern_prob(C, p, clicks, upto=None):
    L = C.shape[1] if upto is None else upto
    m = np.all(C[:, :L] == np.asarray(clicks[:L])[None, :], axis=1)
    return p[m].sum()


def conditional_click_probs(model, clicks, **kw):
    """P(C_r = c_r | C_<r = c_<r) for each rank."""
    C, _, _, _, p = cascade_joint(model, **kw)
    out = []
    for r in range(len(clicks)):
        out.append(pattern_prob(C, p, clicks, r + 1) / pattern_prob(C, p, clicks, r))
    return np.array(out)


def marginal_click_probs(model, **kw):
    """P(C_r = 1) for each rank."""
    C, _, _, _, p = cascade_joint(model, **kw)
    return (C * p[:, None]).sum(axis=0)


def atr_posterior_no_later_click(model, r, **kw):
    """P(A_r = 1 | C_j = 0 for every j >= r), marginalising over ranks < r."""
    C, A, _, _, p = cascade_joint(model, **kw)
    m = np.all(C[:, r:] == 0, axis=1)
    return p[m & (A[:, r] == 1)].sum() / p[m].sum()


def sat_posterior_last_click(r, clicks, **kw):
    """DBN: P(S_r = 1 | clicks as given), r = the last clicked rank."""
    C, _, _, S, p = cascade_joint("dbn", **kw)
    m = np.all(C == np.asarray(clicks)[None, :], axis=1)
    return p[m & (S[:, r] == 1)].sum() / p[m].sum()


def ubm_marginal_click_probs(p_click):
    """UBM: p_click(r, prev) is P(C_r = 1 | last click before r at prev, -1 = none).
    Returns P(C_r = 1) by enumerating all click patterns."""
    L = len(p_click)
    out = np.zeros(L)
    for pat in itertools.product((0, 1), repeat=L):
        pr = 1.0
        prev = -1
        for r in range(L):
            q = p_click[r](prev)
            pr *= q if pat[r] else 1 - q
            if pat[r]:
                prev = r
        out += pr * np.array(pat)
    return out


def pbm_posteriors(a, e):
    """(P(A=1 | C=0), P(E=1 | C=0)) with C = A E, A, E independent."""
    num_a = num_e = z = 0.0
    for A in (0, 1):
        for E in (0, 1):
            pr = (a if A else 1 - a) * (e if E else 1 - e)
            if A * E == 0:
                z += pr
                num_a += pr * A
                num_e += pr * E
    return num_a / z, num_e / z


def _factor_weights(C, A, E, S, p, clicks, r):
    """w[x, y, z] = P(E_r = x, S'_r = y, E_{r+1} = z, C = clicks, A_r = c_r)
    with S'_r = S_r C_r (satisfaction exists only after a click)."""
    c = np.asarray(clicks)
    m = np.all(C == c[None, :C.shape[1]], axis=1) & (A[:, r] == c[r])
    Sp = S[:, r] * C[:, r]
    w = np.zeros((2, 2, 2))
    for x in (0, 1):
        for y in (0, 1):
            for z in (0, 1):
                w[x, y, z] = p[m & (E[:, r] == x) & (Sp == y) & (E[:, r + 1] == z)].sum()
    return w


def factor_posteriors(model, clicks, **kw):
    """Per rank r of the whole list: the posterior over (E_r, S'_r, E_{r+1})
    given every click and A_r = c_r, as an array [L, 2, 2, 2]."""
    C, A, E, S, p = cascade_joint(model, **kw)
    out = []
    for r in range(len(clicks)):
        w = _factor_weights(C, A, E, S, p, clicks, r)
        out.append(w / w.sum())
    return np.array(out)


def factor_posteriors_fresh_tail(model, clicks, a, s=None, g=None, taus=None):
    """As factor_posteriors, except that the likelihood of the clicks after
    rank r given E_{r+1} = 1 is that of a fresh list started at rank 0: the
    clicks of ranks r+1 .. L-1 scored with the parameters of ranks
    0 .. L-r-2.  Ranks 0..r are enumerated with their own parameters."""
    L = len(clicks)
    c = np.asarray(clicks)
    out = []
    for r in range(L):
        kw = dict(a=a[:r + 1], s=None if s is None else s[:r + 1], g=g, taus=taus)
        C, A, E, S, p = cascade_joint(model, **kw)
        w = _factor_weights(C, A, E, S, p, c[:r + 1], r)
        if r + 1 < L:
            n = L - r - 1
            kt = dict(a=a[:n], s=None if s is None else s[:n], g=g, taus=taus)
            Ct, _, _, _, pt = cascade_joint(model, **kt)
            t1 = pattern_prob(Ct, pt, c[r + 1:])
            t0 = 0.0 if c[r + 1:].any() else 1.0
            w[:, :, 1] *= t1
            w[:, :, 0] *= t0
        out.append(w / w.sum())
    return np.array(out)


def fisher_terms(model, clicks, a, s=None, g=None, taus=None):
    """Per-rank E[X - p | clicks] for every Bernoulli variable X ~ Bern(p)
    of the generative story, i.e. the gradient of log P(clicks) with
    respect to each parameter's logit (Fisher's identity).  Returns
    (d_attr[L], d_sat[L], d_shared) with d_shared = [g] (DBN) or
    [t1, t2, t3] (CCM); for the CCM the satisfaction draws S_r ~ Bern(a_r)
    count towards d_attr."""
    C, A, E, S, p = cascade_joint(model, a=a, s=s, g=g, taus=taus)
    L = len(a)
    c = np.asarray(clicks)
    m = np.all(C == c[None, :], axis=1)
    w = p[m] / p[m].sum()
    A, S, E, C = A[m], S[m], E[m], C[m]
    H = _assignments(3 * L)[m]
    G = H[:, 2 * L:3 * L]
    d_attr = np.array([(w * (A[:, r] - a[r])).sum() for r in range(L)])
    if model == "dbn":
        d_sat = np.array([(w * (S[:, r] - s[r])).sum() for r in range(L)])
        d_sh = np.array([sum((w * (G[:, r] - g)).sum() for r in range(L))])
        return d_attr, d_sat, d_sh
    t1, t2, t3 = taus
    d_attr = d_attr + np.array([(w * (S[:, r] - a[r])).sum() for r in range(L)])
    d_sh = np.zeros(3)
    for r in range(L):
        clicked = C[:, r] == 1
        d_sh[0] += (w * np.where(~clicked, G[:, r] - t1, 0)).sum()
        d_sh[1] += (w * np.where(clicked & (S[:, r] == 0), G[:, r] - t2, 0)).sum()
        d_sh[2] += (w * np.where(clicked & (S[:, r] == 1), G[:, r] - t3, 0)).sum()
    return d_attr, np.zeros(L), d_sh


def independent_joint(a, e):
    """PBM: every assignment of (A_r, E_r) for L ranks with A_r ~ Bern(a_r),
    E_r ~ Bern(e_r) all independent.  Returns (clicks [M, L], prob [M])."""
    L = len(a)
    H = _assignments(2 * L)
    A, E = H[:, :L], H[:, L:]
    p = np.prod(np.where(A == 1, a[None, :], 1 - a[None, :]), axis=1)
    p = p * np.prod(np.where(E == 1, e[None, :], 1 - e[None, :]), axis=1)
    return (A * E).astype(np.int8), p


def ubm_joint(a, exam_of):
    """UBM: every assignment of (A_r, E_r) for L ranks generated rank by
    rank, A_r ~ Bern(a_r) and E_r ~ Bern(exam_of(r, prev)) where prev is the
    last clicked rank before r (-1 = none) of that assignment's own clicks.
    Returns (clicks [M, L], prob [M])."""
    L = len(a)
    H = _assignments(2 * L)
    A, E = H[:, :L], H[:, L:]
    M = H.shape[0]
    C = np.zeros((M, L), np.int8)
    p = np.ones(M)
    prev = np.full(M, -1)
    for r in range(L):
        e = np.array([exam_of(r, int(q)) for q in range(-1, r)])[prev + 1]
        p *= np.where(A[:, r] == 1, a[r], 1 - a[r]) * np.where(E[:, r] == 1, e, 1 - e)
        C[:, r] = A[:, r] * E[:, r]
        prev = np.where(C[:, r] == 1, r, prev)
    return C, p


def pattern_probs(C, p):
    """P(C = pattern) for every pattern, indexed by the bit mask sum_r c_r 2^r."""
    L = C.shape[1]
    code = (C.astype(np.int64) << np.arange(L)[None, :]).sum(axis=1)
    return np.bincount(code, weights=p, minlength=1 << L)


def conditional_from_patterns(pp, L):
    """P(C_r = c_r | C_<r = c_<r) for every mask m (bit r = c_r) and rank r,
    from the pattern probabilities pp: out[m, r]."""
    masks = np.arange(1 << L)
    out = np.zeros((1 << L, L))
    for r in range(L):
        upto = masks & ((1 << (r + 1)) - 1)
        before = masks & ((1 << r) - 1)
        p_upto = np.bincount(upto, weights=pp, minlength=1 << L)
        p_before = np.bincount(before, weights=pp, minlength=1 << L)
        out[:, r] = p_upto[upto] / p_before[before]
    return out
