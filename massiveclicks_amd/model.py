"""Host-side interface of the EM engine, mirroring the reference's
ClickModel_Hst (click_models/base.cuh:30-53), Dataset (data/dataset.h) and the
single-worker EM driver em_parallel (parallel_em/parallel_em.cu:349-746).

Everything runs through libmclicks.so (the HIP engine); there is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, ptr, i32p, i64p, u8p, f32p, f64p

PBM, CCM, DBN, UBM = 0, 1, 2, 3          # main.cpp:178-183
MODEL_NAMES = {PBM: "PBM", CCM: "CCM", DBN: "DBN", UBM: "UBM"}
MAX_SERP = 10

(OPT_REDUCE_MODE, OPT_ITEM_SESSIONS, OPT_ITEM_WINDOW, OPT_SORT_BY_LCR, OPT_BLOCK, OPT_DIRECT, OPT_PERSISTENT,
 OPT_BATCHES, OPT_PACKED, OPT_SKIP_CONSTANT, OPT_FUSED_MERGE) = range(11)
REDUCE_SEGMENTED, REDUCE_ATOMIC = 0, 1


@dataclass
class SessionSet:
    """Sessions in the engine's host layout (session-major, 10 ranks)."""
    query: np.ndarray      # int32 [n]
    docs: np.ndarray       # int32 [n, 10]
    clicks: np.ndarray     # uint8 [n, 10]
    pidx: np.ndarray       # int32 [n, 10], -1 = pair unseen in training
    length: np.ndarray | None = None   # uint8 [n] or None (= 10)

    def __len__(self):
        return int(self.query.shape[0])


class Dataset:
    """Parsed, split, sorted and indexed click log (dataset.cpp:421-455)."""

    def __init__(self, handle, has_length=False):
        self._h = handle
        self.has_length = has_length
        L = _lib.lib()
        n = [C.c_int64() for _ in range(4)]
        check(L.mc_dataset_sizes(self._h, *[C.byref(x) for x in n]))
        self.n_parsed, self.n_train, self.n_test, self.n_qd = (x.value for x in n)
        self.train = self._export(0, self.n_train)
        self.test = self._export(1, self.n_test)

    @classmethod
    def parse(cls, path: str, max_sessions: int = -1, test_share: float = 0.2) -> "Dataset":
        h = C.c_void_p()
        check(_lib.lib().mc_dataset_parse(path.encode(), int(max_sessions), float(test_share), C.byref(h)))
        return cls(h)

    @classmethod
    def from_sessions(cls, query, docs, clicks, test_share: float = 0.2, length=None) -> "Dataset":
        """In-memory sessions in log order; `length` (ranks shown, 1..10)
        makes ranks past each session's length absent."""
        q = np.ascontiguousarray(query, dtype=np.int32)
        d = np.ascontiguousarray(docs, dtype=np.int32)
        c = np.ascontiguousarray(clicks, dtype=np.uint8)
        ln = None if length is None else np.ascontiguousarray(length, dtype=np.uint8)
        h = C.c_void_p()
        check(_lib.lib().mc_dataset_from_sessions_len(len(q), ptr(q, i32p), ptr(d, i32p), ptr(c, u8p),
                                                       ptr(ln, u8p), float(test_share), C.byref(h)))
        return cls(h, has_length=ln is not None)

    def _export(self, which, n):
        q = np.empty(n, np.int32)
        d = np.empty((n, MAX_SERP), np.int32)
        c = np.empty((n, MAX_SERP), np.uint8)
        p = np.empty((n, MAX_SERP), np.int32)
        ln = None
        if n:
            check(_lib.lib().mc_dataset_export(self._h, which, ptr(q, i32p), ptr(d, i32p), ptr(c, u8p), ptr(p, i32p)))
            if self.has_length:
                ln = np.empty(n, np.uint8)
                check(_lib.lib().mc_dataset_export_len(self._h, which, ptr(ln, u8p)))
        return SessionSet(q, d, c, p, ln)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.lib().mc_dataset_free(self._h)
            self._h = None


def _sessions_arrays(sessions: SessionSet, what: str):
    """Contiguous (pidx [n, 10] int32, clicks [n, 10] uint8, length [n] uint8
    or None) of a session set, with the shapes the C ABI reads checked
    first: the native side trusts the element counts it is given."""
    n = len(sessions)
    p = np.ascontiguousarray(sessions.pidx, np.int32)
    c = np.ascontiguousarray(sessions.clicks, np.uint8)
    if p.shape != (n, MAX_SERP):
        raise ValueError(f"{what}: pidx must have shape ({n}, {MAX_SERP}), got {p.shape}")
    if c.shape != (n, MAX_SERP):
        raise ValueError(f"{what}: clicks must have shape ({n}, {MAX_SERP}), got {c.shape}")
    ln = None
    if sessions.length is not None:
        ln = np.ascontiguousarray(sessions.length, np.uint8)
        if ln.shape != (n,):
            raise ValueError(f"{what}: length must have shape ({n},), got {ln.shape}")
    return p, c, ln


def _eval_summary(n, llh_sum, ppl, cnt, per_session):
    """LL and perplexity from the FP64 totals (parallel_em.cu:697-724)."""
    ppl_rank = [2.0 ** (-ppl[r] / cnt[r]) if cnt[r] else float("nan") for r in range(MAX_SERP)]
    out = {
        "llh_sum": llh_sum,
        "log_likelihood": llh_sum / n,                 # parallel_em.cu:704
        "ppl_log2_sum": ppl,
        "rank_count": cnt,
        "perplexity_rank": ppl_rank,                  # parallel_em.cu:719-721
        "perplexity": float(np.mean(ppl_rank)),        # parallel_em.cu:723
    }
    if per_session is not None:
        out.update(logp=per_session[0], full=per_session[1], llh=per_session[2])
    return out


class ClickModel:
    """One click model on one GPU (ClickModel_Hst, base.cuh:30-53)."""

    def __init__(self, model_type: int, device: int = 0, reduce_mode: int = REDUCE_SEGMENTED,
                 item_sessions: int | None = None, item_window: int | None = None,
                 sort_by_lcr: bool = True, block: int | None = None, direct: bool = False,
                 persistent: bool = True, batches: int = 1, packed: bool = True, skip_constant: bool = True,
                 fused_merge: bool = True):
        self.model_type = int(model_type)
        self._h = C.c_void_p()
        L = _lib.lib()
        check(L.mc_model_create(self.model_type, int(device), C.byref(self._h)))
        check(L.mc_model_set_option(self._h, OPT_REDUCE_MODE, int(reduce_mode)))
        check(L.mc_model_set_option(self._h, OPT_SORT_BY_LCR, int(bool(sort_by_lcr))))
        check(L.mc_model_set_option(self._h, OPT_DIRECT, int(bool(direct))))
        check(L.mc_model_set_option(self._h, OPT_PERSISTENT, int(bool(persistent))))
        check(L.mc_model_set_option(self._h, OPT_BATCHES, int(batches)))
        check(L.mc_model_set_option(self._h, OPT_PACKED, int(bool(packed))))
        check(L.mc_model_set_option(self._h, OPT_SKIP_CONSTANT, int(bool(skip_constant))))
        check(L.mc_model_set_option(self._h, OPT_FUSED_MERGE, int(bool(fused_merge))))
        if item_sessions is not None:
            check(L.mc_model_set_option(self._h, OPT_ITEM_SESSIONS, int(item_sessions)))
        if item_window is not None:
            check(L.mc_model_set_option(self._h, OPT_ITEM_WINDOW, int(item_window)))
        if block is not None:
            check(L.mc_model_set_option(self._h, OPT_BLOCK, int(block)))
        self.n_qd = 0

    # -- lifecycle ------------------------------------------------------------
    def close(self):
        if getattr(self, "_h", None):
            _lib.lib().mc_model_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- init_parameters (common.cu:31-61) --------------------------------------
    def init_parameters(self, train: SessionSet, n_qd: int):
        q = np.ascontiguousarray(train.query, np.int32)
        if q.ndim != 1:
            raise ValueError(f"init_parameters: query must be one-dimensional, got shape {q.shape}")
        p, c, ln = _sessions_arrays(train, "init_parameters")
        check(_lib.lib().mc_init_parameters(self._h, len(q), ptr(q, i32p), ptr(p, i32p), ptr(c, u8p),
                                            ptr(ln, u8p), int(n_qd)))
        self.n_qd = int(n_qd)

    def init_from_dataset(self, ds: Dataset):
        check(_lib.lib().mc_init_from_dataset(self._h, ds._h))
        self.n_qd = ds.n_qd

    # -- EM ---------------------------------------------------------------------
    def process_session(self):
        check(_lib.lib().mc_process_session(self._h))

    def update_parameters(self):
        check(_lib.lib().mc_update_parameters(self._h))

    def em(self, n_iter: int, use_graph: bool = True):
        check(_lib.lib().mc_em_iterations(self._h, int(n_iter), int(bool(use_graph))))

    def em_timed(self, n_iter: int):
        t = [C.c_double() for _ in range(3)]
        check(_lib.lib().mc_em_iterations_timed(self._h, int(n_iter), *[C.byref(x) for x in t]))
        return {"ms_total": t[0].value, "ms_estep": t[1].value, "ms_mstep": t[2].value}

    def em_graph_timed(self, n_iter: int) -> float:
        t = C.c_double()
        check(_lib.lib().mc_em_graph_timed(self._h, int(n_iter), C.byref(t)))
        return t.value

    def reset_parameters(self):
        check(_lib.lib().mc_reset_parameters(self._h))

    # -- gradient path ----------------------------------------------------------
    def grad_init(self):
        """Logits from the current values, Adam state to zero."""
        check(_lib.lib().mc_grad_init(self._h))

    def loglik_gradient(self, batch: int = -1):
        """(training log-likelihood, [d LL / d logit per table]) at the
        current values, over minibatch `batch` (-1 = every session)."""
        info = self.get_parameter_information()
        grads = [np.empty(size, np.float32) for _, _, size in info]
        arr = (f32p * len(grads))(*[g.ctypes.data_as(f32p) for g in grads])
        ll = C.c_double()
        check(_lib.lib().mc_loglik_gradient(self._h, int(batch), C.byref(ll), arr))
        return ll.value, grads

    def adam_configure(self, lr=0.05, beta1=0.9, beta2=0.999, eps=1e-8, prior=True):
        check(_lib.lib().mc_adam_configure(self._h, float(lr), float(beta1), float(beta2), float(eps), int(bool(prior))))

    def adam_steps(self, n_steps: int, use_graph: bool = True):
        check(_lib.lib().mc_adam_steps(self._h, int(n_steps), int(bool(use_graph))))

    def adam_steps_timed(self, n_steps: int) -> float:
        ms = C.c_double()
        check(_lib.lib().mc_adam_steps_timed(self._h, int(n_steps), C.byref(ms)))
        return ms.value

    def get_logits(self, t: int):
        out = np.empty(self.get_parameter_information()[t][2], np.float32)
        check(_lib.lib().mc_get_logits(self._h, int(t), ptr(out, f32p)))
        return out

    def set_logits(self, t: int, logits):
        x = np.ascontiguousarray(logits, np.float32)
        size = self.get_parameter_information()[int(t)][2]
        if x.shape != (size,):
            raise ValueError(f"set_logits: table {t} takes logits of shape ({size},), got {x.shape}")
        check(_lib.lib().mc_set_logits(self._h, int(t), ptr(x, f32p)))

    # -- tables -----------------------------------------------------------------
    def n_tables(self) -> int:
        n = C.c_int()
        check(_lib.lib().mc_get_n_tables(self._h, C.byref(n)))
        return n.value

    def get_parameter_information(self):
        """[(name, is_public, size)] per table (get_parameter_information, base.cuh:33)."""
        out = []
        for t in range(self.n_tables()):
            name, pub, size = C.c_char_p(), C.c_int(), C.c_int64()
            check(_lib.lib().mc_get_table_info(self._h, t, C.byref(name), C.byref(pub), C.byref(size)))
            out.append((name.value.decode(), bool(pub.value), size.value))
        return out

    def get_parameters(self, t: int):
        """(numerator, denominator) float32 arrays of table t."""
        info = self.get_parameter_information()[t]
        num = np.empty(info[2], np.float32)
        den = np.empty(info[2], np.float32)
        check(_lib.lib().mc_get_parameters(self._h, int(t), ptr(num, f32p), ptr(den, f32p)))
        return num, den

    def get_values(self, t: int):
        """Param::value() of table t (param.cu:75-87)."""
        num, den = self.get_parameters(t)
        v = num / den
        return np.where(v < np.float32(1.0) - np.float32(1e-6), v, np.float32(1.0) - np.float32(1e-6)).astype(np.float32)

    def set_parameters(self, t: int, num, den):
        info = self.get_parameter_information()
        if not 0 <= int(t) < len(info):
            raise ValueError(f"set_parameters: no table {t}")
        size = info[int(t)][2]
        num = np.ascontiguousarray(num, np.float32)
        den = np.ascontiguousarray(den, np.float32)
        if num.shape != (size,) or den.shape != (size,):
            raise ValueError(f"set_parameters: table {t} takes num and den of shape ({size},), "
                             f"got {num.shape} and {den.shape}")
        check(_lib.lib().mc_set_parameters(self._h, int(t), ptr(num, f32p), ptr(den, f32p)))

    # -- evaluation (evaluation.cpp:25-66) --------------------------------------
    def evaluate(self, test: SessionSet, per_session: bool = False):
        n = len(test)
        p, c, ln = _sessions_arrays(test, "evaluate")
        logp = np.empty((n, MAX_SERP), np.float32) if per_session else None
        full = np.empty((n, MAX_SERP), np.float32) if per_session else None
        llh = np.empty(n, np.float32) if per_session else None
        llh_sum = C.c_double()
        ppl = np.empty(MAX_SERP, np.float64)
        cnt = np.empty(MAX_SERP, np.int64)
        check(_lib.lib().mc_evaluate(self._h, n, ptr(p, i32p), ptr(c, u8p), ptr(ln, u8p), ptr(logp, f32p),
                                     ptr(full, f32p), ptr(llh, f32p), C.byref(llh_sum), ptr(ppl, f64p),
                                     ptr(cnt, i64p)))
        return _eval_summary(n, llh_sum.value, ppl, cnt, (logp, full, llh) if per_session else None)

    def set_test_set(self, test: SessionSet):
        """Keeps the held-out set on the device (mc_set_test_set) for
        evaluate_test_set."""
        p, c, ln = _sessions_arrays(test, "set_test_set")
        check(_lib.lib().mc_set_test_set(self._h, len(test), ptr(p, i32p), ptr(c, u8p), ptr(ln, u8p)))
        self._n_test = len(test)

    def evaluate_test_set(self, per_session: bool = False, timed: bool = False):
        """evaluate() on the resident held-out set; per-session outputs in
        the order the set was given.  timed adds the evaluation kernel's
        duration in ms ("kernel_ms")."""
        n = getattr(self, "_n_test", 0)
        logp = np.empty((n, MAX_SERP), np.float32) if per_session else None
        full = np.empty((n, MAX_SERP), np.float32) if per_session else None
        llh = np.empty(n, np.float32) if per_session else None
        llh_sum = C.c_double()
        ppl = np.empty(MAX_SERP, np.float64)
        cnt = np.empty(MAX_SERP, np.int64)
        ms = C.c_double()
        check(_lib.lib().mc_evaluate_test_set(self._h, ptr(logp, f32p), ptr(full, f32p), ptr(llh, f32p),
                                              C.byref(llh_sum), ptr(ppl, f64p), ptr(cnt, i64p),
                                              C.byref(ms) if timed else None))
        out = _eval_summary(n, llh_sum.value, ppl, cnt, (logp, full, llh) if per_session else None)
        if timed:
            out["kernel_ms"] = ms.value
        return out

    def get_log_conditional_click_probs(self, test: SessionSet):
        return self.evaluate(test, per_session=True)["logp"]

    def get_full_click_probs(self, test: SessionSet):
        return self.evaluate(test, per_session=True)["full"]

    def get_serp_probability(self, test: SessionSet):
        """get_serp_probability (base.cuh:50) of every session: [n, 10]."""
        p, c, ln = _sessions_arrays(test, "get_serp_probability")
        out = np.empty((len(test), MAX_SERP), np.float32)
        check(_lib.lib().mc_serp_probability(self._h, len(test), ptr(p, i32p), ptr(c, u8p), ptr(ln, u8p),
                                             ptr(out, f32p)))
        return out

    # -- bookkeeping -------------------------------------------------------------
    def get_memory_usage(self) -> int:
        b = C.c_size_t()
        check(_lib.lib().mc_get_memory_usage(self._h, C.byref(b)))
        return b.value

    @staticmethod
    def compute_memory_footprint(model_type: int, n: int, n_qd: int) -> int:
        b = C.c_size_t()
        check(_lib.lib().mc_compute_memory_footprint(int(model_type), int(n), int(n_qd), C.byref(b)))
        return b.value

    def layout_info(self):
        a, b, c, d = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check(_lib.lib().mc_get_layout_info(self._h, C.byref(a), C.byref(b), C.byref(c), C.byref(d)))
        return {"n_items": a.value, "n_shared_entries": b.value, "max_item_sessions": c.value, "n_em_sessions": d.value}

    def stream_info(self):
        """Sessions of window items (window-relative indices) and the device
        bytes of all session indices (packed 16-byte records counted whole),
        with the sessions per stored format."""
        a, b = C.c_int64(), C.c_int64()
        L = _lib.lib()
        check(L.mc_get_stream_info(self._h, C.byref(a), C.byref(b)))
        p, w, g = C.c_int64(), C.c_int64(), C.c_int64()
        check(L.mc_get_session_formats(self._h, C.byref(p), C.byref(w), C.byref(g)))
        return {"n_window_sessions": a.value, "index_bytes": b.value, "n_packed": p.value, "n_window16": w.value,
                "n_absolute": g.value}


OPTION_NAMES = ("reduce_mode", "item_sessions", "item_window", "sort_by_lcr", "block", "direct", "persistent",
                "batches", "packed", "skip_constant", "fused_merge")
DEFAULT_OPTIONS = {"reduce_mode": REDUCE_SEGMENTED, "item_sessions": 8192, "item_window": 1280, "sort_by_lcr": 1,
                   "block": 256, "direct": 0, "persistent": 1, "batches": 1, "packed": 1, "skip_constant": 1, "fused_merge": 1}


def autotune(model_type: int, train: SessionSet, n_qd: int, device: int = 0, iterations: int = 5,
             start: dict | None = None, cache: str | None = None) -> dict:
    """Fastest EM options for this model and training set (mc_autotune).
    Returns (options, info): options are keyword arguments of ClickModel,
    info = {ms_per_iter, trials, cached}.  With `cache` (a JSON file) results
    are stored per (model, sessions, n_qd, uneven lengths) and reused."""
    import json
    import os
    key = f"{MODEL_NAMES[model_type]}:{len(train)}:{n_qd}:{int(train.length is not None)}"
    if cache and os.path.exists(cache):
        with open(cache) as f:
            known = json.load(f)
        if key in known:
            return dict(known[key]["options"]), {"ms_per_iter": known[key]["ms_per_iter"],
                                                 "trials": known[key]["trials"], "cached": True}
    opts = dict(DEFAULT_OPTIONS, **(start or {}))
    arr = np.array([int(opts[k]) for k in OPTION_NAMES], np.int64)
    q = np.ascontiguousarray(train.query, np.int32)
    p = np.ascontiguousarray(train.pidx, np.int32)
    c = np.ascontiguousarray(train.clicks, np.uint8)
    ln = None if train.length is None else np.ascontiguousarray(train.length, np.uint8)
    ms = C.c_double()
    n_trials = C.c_int()
    check(_lib.lib().mc_autotune(int(model_type), int(device), len(q), ptr(q, i32p), ptr(p, i32p), ptr(c, u8p),
                                 ptr(ln, u8p), int(n_qd), int(iterations), ptr(arr, i64p), C.byref(ms),
                                 C.byref(n_trials)))
    best = {k: int(v) for k, v in zip(OPTION_NAMES, arr)}
    best["sort_by_lcr"] = bool(best["sort_by_lcr"])
    best["direct"] = bool(best["direct"])
    best["persistent"] = bool(best["persistent"])
    best["packed"] = bool(best["packed"])
    best["skip_constant"] = bool(best["skip_constant"])
    best["fused_merge"] = bool(best["fused_merge"])
    if cache:
        known = {}
        if os.path.exists(cache):
            with open(cache) as f:
                known = json.load(f)
        known[key] = {"options": best, "ms_per_iter": ms.value, "trials": n_trials.value}
        with open(cache, "w") as f:
            json.dump(known, f, indent=1, sort_keys=True)
    return best, {"ms_per_iter": ms.value, "trials": n_trials.value, "cached": False}


def em_train(model_type: int, ds: Dataset, iterations: int = 50, device: int = 0, use_graph: bool = True,
             **options):
    """Single-worker em_parallel (parallel_em.cu:349-746): train, then evaluate
    held-out log-likelihood and perplexity."""
    cm = ClickModel(model_type, device=device, **options)
    cm.init_from_dataset(ds)
    cm.em(iterations, use_graph=use_graph)
    ev = cm.evaluate(ds.test) if len(ds.test) else None
    return cm, ev
