"""Seeded synthetic click logs for the PBM, UBM, CCM and DBN workloads.

Queries are drawn from a Zipf law p_k ~ k^-s over n_queries ids; every query
owns a pool of `pool` distinct documents drawn uniformly from a vocabulary of
`vocab` ids; a session shows L documents sampled from its query's pool
without replacement, in random order.  Clicks come from a ground-truth click
model:

* PBM: P(click at r) = alpha_qd * gamma_r, gamma_r = 1 / (r+1)^0.8.
* UBM: P(click at r) = alpha_qd * gamma[r][d], d = r - (last clicked rank
  before r), d = r + 1 when nothing was clicked before; gamma[r][d] =
  0.9 exp(-0.15 r - 0.1 d).
* DBN: cascade; examined rank clicked w.p. alpha; after a click the user is
  satisfied w.p. sigma and otherwise continues w.p. gamma = 0.9; after a
  non-click continues w.p. gamma.
* CCM: cascade; continue after a non-click w.p. tau1 = 0.95, after a click
  w.p. tau2 (1 - alpha) + tau3 alpha with tau2 = 0.6, tau3 = 0.4.

alpha ~ Beta(2, 5) and sigma ~ Beta(2, 8) per (query, pool slot).

The raw-text writer emits the reference's log format: a 15-token query line
"<sid> 0 Q <query> 0 <doc_1> ... <doc_10>" followed by one 4-token click line
"<sid> 0 C <doc>" per click, and one trailing sentinel query line, because
the reference parser only emits a session when the next query line arrives.
"""
from __future__ import annotations

import dataclasses
import numpy as np

MAX_SERP = 10


@dataclasses.dataclass
class SynthConfig:
    n_sessions: int
    n_queries: int
    vocab: int
    zipf_s: float = 1.1
    zipf_shift: float = 0.0      # Zipf-Mandelbrot shift q: p_k ~ (k + q)^-s
    pool: int = 20
    model: str = "dbn"           # ground truth: pbm, ubm, dbn, ccm
    variable_length: bool = False
    seed: int = 1234


@dataclasses.dataclass
class Sessions:
    query: np.ndarray    # int32 [n]
    docs: np.ndarray     # int32 [n, 10] (ranks >= len hold -1)
    clicks: np.ndarray   # uint8 [n, 10]
    length: np.ndarray   # uint8 [n]
    stats: dict


def _zipf_queries(rng, n, n_queries, s, shift):
    k = np.arange(1, n_queries + 1, dtype=np.float64)
    w = (k + shift) ** (-s)
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    u = rng.random(n)
    return np.minimum(np.searchsorted(cdf, u, side="right"), n_queries - 1).astype(np.int64), w / w.sum()


def _pools(rng, qids, pool, vocab):
    """Distinct document ids per query (rows of `pool`)."""
    docs = rng.integers(0, vocab, size=(len(qids), pool), dtype=np.int64)
    for _ in range(100):
        srt = np.sort(docs, axis=1)
        dup = (srt[:, 1:] == srt[:, :-1]).any(axis=1)
        if not dup.any():
            break
        docs[dup] = rng.integers(0, vocab, size=(int(dup.sum()), pool), dtype=np.int64)
    return docs.astype(np.int32)


def generate(cfg: SynthConfig, chunk: int = 1 << 20) -> Sessions:
    rng = np.random.Generator(np.random.PCG64(cfg.seed))
    qidx, pmf = _zipf_queries(rng, cfg.n_sessions, cfg.n_queries, cfg.zipf_s, cfg.zipf_shift)
    used, inv = np.unique(qidx, return_inverse=True)
    pools = _pools(rng, used, cfg.pool, cfg.vocab)
    alpha = rng.beta(2.0, 5.0, size=(len(used), cfg.pool)).astype(np.float64)
    sigma = rng.beta(2.0, 8.0, size=(len(used), cfg.pool)).astype(np.float64)

    n = cfg.n_sessions
    docs = np.full((n, MAX_SERP), -1, dtype=np.int32)
    clicks = np.zeros((n, MAX_SERP), dtype=np.uint8)
    if cfg.variable_length:
        length = rng.integers(1, MAX_SERP + 1, size=n).astype(np.uint8)
    else:
        length = np.full(n, MAX_SERP, dtype=np.uint8)

    ranks = np.arange(MAX_SERP)
    pbm_gamma = 1.0 / (ranks + 1.0) ** 0.8
    for a in range(0, n, chunk):
        b = min(n, a + chunk)
        m = b - a
        g = inv[a:b]
        # L documents from the pool without replacement, random order.
        keys = rng.random((m, cfg.pool), dtype=np.float32)
        slot = np.argsort(keys, axis=1)[:, :MAX_SERP]
        d = pools[g[:, None], slot]
        al = alpha[g[:, None], slot]
        si = sigma[g[:, None], slot]
        u = rng.random((m, MAX_SERP))
        u2 = rng.random((m, MAX_SERP))
        u3 = rng.random((m, MAX_SERP))
        L = length[a:b]
        shown = ranks[None, :] < L[:, None]
        c = np.zeros((m, MAX_SERP), dtype=bool)
        if cfg.model == "pbm":
            c = u < al * pbm_gamma[None, :]
        elif cfg.model == "ubm":
            last = np.full(m, -1)
            for r in range(MAX_SERP):
                dist = r - last
                gam = 0.9 * np.exp(-0.15 * r - 0.1 * dist)
                c[:, r] = u[:, r] < al[:, r] * np.minimum(gam, 1.0)
                last = np.where(c[:, r], r, last)
        elif cfg.model in ("dbn", "ccm"):
            exam = np.ones(m, dtype=bool)
            for r in range(MAX_SERP):
                cl = exam & (u[:, r] < al[:, r])
                c[:, r] = cl
                if cfg.model == "dbn":
                    sat = cl & (u2[:, r] < si[:, r])
                    cont = u3[:, r] < 0.9
                    exam = exam & ~sat & cont
                else:
                    pc = np.where(cl, 0.6 * (1 - al[:, r]) + 0.4 * al[:, r], 0.95)
                    exam = exam & (u3[:, r] < pc)
        else:
            raise ValueError(f"unknown model {cfg.model}")
        c &= shown
        docs[a:b] = np.where(shown, d, -1)
        clicks[a:b] = c.astype(np.uint8)

    query = used[inv].astype(np.int32)
    counts = np.bincount(inv, minlength=len(used))
    stats = {
        "n_sessions": int(n),
        "distinct_queries": int(len(used)),
        "top_query_share": float(counts.max() / n),
        "mean_length": float(length.mean()),
        "click_rate": float(clicks.sum() / max(1, int(length.sum()))),
        "sessions_without_click": float((clicks.sum(axis=1) == 0).mean()),
    }
    return Sessions(query=query, docs=docs, clicks=clicks, length=length, stats=stats)


def write_raw_log(sessions: Sessions, path: str, sentinel: bool = True) -> None:
    """Writes full-length sessions in the reference's raw text format."""
    if (sessions.length != MAX_SERP).any():
        raise ValueError("the raw format only holds 10-document sessions")
    n = len(sessions.query)
    with open(path, "w") as f:
        for i in range(n):
            q = int(sessions.query[i])
            f.write(f"{i} 0 Q {q} 0 " + " ".join(str(int(x)) for x in sessions.docs[i]) + "\n")
            for r in np.nonzero(sessions.clicks[i])[0]:
                f.write(f"{i} 0 C {int(sessions.docs[i, r])}\n")
        if sentinel:
            f.write(f"{n} 0 Q 0 0 " + " ".join(["0"] * MAX_SERP) + "\n")


# Workloads used by the tests and the benchmark (sizes as stated in each name).
def config_pbm_40k(seed: int = 1) -> SynthConfig:
    return SynthConfig(n_sessions=40_000, n_queries=2_000, vocab=50_000, zipf_s=1.1, model="pbm", seed=seed)


def config_ubm_10m(seed: int = 2) -> SynthConfig:
    return SynthConfig(n_sessions=10_000_000, n_queries=1_000_000, vocab=20_000_000, zipf_s=1.1, model="ubm", seed=seed)


def config_dbn_10m(seed: int = 3) -> SynthConfig:
    return SynthConfig(n_sessions=10_000_000, n_queries=1_000_000, vocab=20_000_000, zipf_s=1.1, model="dbn", seed=seed)


def config_ccm_10m(seed: int = 4) -> SynthConfig:
    return SynthConfig(n_sessions=10_000_000, n_queries=1_000_000, vocab=20_000_000, zipf_s=1.1, model="ccm", seed=seed)
