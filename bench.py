"""Benchmark: plain-EM DBN on 10M synthetic sessions (8M train), one MI355X.

One "step" is one EM iteration (E-step + M-step) over the whole training set,
replayed from a captured HIP graph with the sessions and tables resident in
HBM.  Prints one JSON line:

  value       sessions/s per iteration = N_train / mean iteration time
  roofline    the E/M item kernel (the dominant kernel): algorithmic bytes per
              launch / its mean duration (HIP events on the engine's stream)
              against the 8 TB/s HBM peak
  heldout     held-out LL and perplexity from the resident test set, with
              the evaluation kernel's duration and algorithmic bytes
  cpu_baseline  the CPU oracle (the reference's threading) on the whole
              training set, on 16 host threads (one GPU's share of the
              host), without and with the reference's per-thread copy of
              the training set
  other_models  CCM, UBM and PBM on their own 10M-session workloads: time
              per iteration and the window kernel's roofline fraction
  time_to_target  DBN on 20M sessions: wall time of plain EM's 50
              iterations (and of its first iteration in the target band),
              and of Adam on the log-likelihood gradient (lr 0.05,
              minibatches of ~1M sessions) to reach EM's 50-iteration
              held-out log-likelihood within 1e-4 relative

Algorithmic bytes (per iteration): 42 B per training session (ten int32
query-document indices + a 16-bit click mask) + 16 B per table entry and
per-qd table (read and write a numerator/denominator pair), DBN has two
per-qd tables.  The item kernel owns the entries of query groups it holds
whole; entries of groups split across items are finalised by the merge
kernel and are excluded from the item kernel's bytes.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12
METRIC = "EM sessions/sec/iter (DBN, 10M sess) + time-to-target held-out LL vs plain EM"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_model_name():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


# Published CPU comparator of the reference (DBN, 10M sessions, 48 threads,
# EPYC 7402P): experiments/data/model_2/EPYC 7402P/nodes_1/
# sessions_10000000/partitioning_2/iteration_50.txt:105.
PUBLISHED_CPU_S_PER_ITER = 17.52
PUBLISHED_CPU_THREADS = 48
SERP_HST_BYTES = 124     # one SERP_Hst: int query + 10 x (click, doc, param index)


def cpu_baseline(ds, model, iters, copy_iters, threads):
    """Times the CPU oracle in the reference's threading (even E-step split,
    query-aligned M-step ranges, mutex merge of the shared tables;
    parallel_em.cu:27-197) on the whole training set: `iters` iterations as
    is, then `copy_iters` iterations that also replay the DBN/CCM E-step's
    copy of the training set into every thread (std::thread takes the train
    set by value, dbn.cu:244, ccm.cu:223)."""
    import oracle
    tr = ds.train
    n = len(tr)
    om = oracle.OracleModel(model, tr.query, tr.pidx, tr.clicks, ds.n_qd, length=tr.length, n_threads=threads)

    def timed(k):
        te = tm = 0.0
        for _ in range(k):
            t0 = time.perf_counter()
            om.estep()
            t1 = time.perf_counter()
            om.mstep()
            t2 = time.perf_counter()
            te += t1 - t0
            tm += t2 - t1
        return te / k, tm / k

    om.iterate(1)   # warm
    e, m = timed(iters)
    dt = e + m
    om.close()
    out = {"value": n / dt, "unit": "sessions/s/iter", "cores": threads, "kind": "port",
           "cpu": cpu_model_name(), "host_cpus": os.cpu_count(),
           "sample": f"all {n} training sessions x {iters} iterations, {threads} threads: {dt:.3f} s/iter "
                     f"(E {e:.3f} s, M {m:.3f} s)"}
    if copy_iters and model in (1, 2):
        om = oracle.OracleModel(model, tr.query, tr.pidx, tr.clicks, ds.n_qd, length=tr.length, n_threads=threads,
                                copy_dataset=True)
        ec, mc = timed(copy_iters)
        om.close()
        dc = ec + mc
        per_copy = max(0.0, dc - dt) / threads
        # The same run on the published comparator's 48 threads: the
        # E-step's compute spread over 48 threads, 48 serial copies.
        est48 = e * threads / PUBLISHED_CPU_THREADS + m + PUBLISHED_CPU_THREADS * per_copy
        out.update({
            "value_with_copy": n / dc,
            "with_copy": f"{copy_iters} iterations with the per-thread train-set copy "
                         f"({threads} x {n * SERP_HST_BYTES / 1e9:.2f} GB per iteration): {dc:.3f} s/iter "
                         f"(E {ec:.3f} s, M {mc:.3f} s), {per_copy:.3f} s per copy",
            "published_reference": f"{PUBLISHED_CPU_S_PER_ITER} s/iter on {PUBLISHED_CPU_THREADS} threads "
                                   f"(EPYC 7402P, Yandex 10M); this host's figures with {PUBLISHED_CPU_THREADS} "
                                   f"copies per iteration: about {est48:.2f} s/iter",
        })
    return out


ADAM_ITEM_SESSIONS = 4096


def time_to_target(n_sessions, batch_sessions, lr, max_epochs, seed=5):
    """Wall time for plain EM and for Adam on the log-likelihood gradient to
    reach the held-out log-likelihood of EM's 50th iteration (within 1e-4
    relative), DBN.  EM: 50 iterations replayed from one graph and timed
    with HIP events; its held-out LL is also evaluated after every single
    iteration (untimed), which gives the first iteration that reaches the
    target band and its share of the 50-iteration time.  Adam: epochs
    (every minibatch once) replayed from one graph; held-out evaluations
    between epochs are not timed."""
    from massiveclicks_amd import Dataset, ClickModel, synth
    from massiveclicks_amd.model import DBN
    t0 = time.perf_counter()
    s = synth.generate(synth.SynthConfig(n_sessions=n_sessions, n_queries=1_000_000, vocab=20_000_000, zipf_s=1.1,
                                         model="dbn", seed=seed))
    ds = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2)
    del s
    log(f"time-to-target data: train {ds.n_train}, test {ds.n_test}, n_qd {ds.n_qd} in {time.perf_counter() - t0:.1f} s")
    test = ds.test
    with ClickModel(DBN) as cm:
        cm.init_from_dataset(ds)
        cm.set_test_set(test)
        em_ms = cm.em_graph_timed(50)      # one warm replay of 50 iterations from the start, then 50 timed
        cm.reset_parameters()
        em_curve = []
        for it in range(1, 51):
            cm.em(1)
            em_curve.append(cm.evaluate_test_set()["log_likelihood"])
        target = em_curve[-1]
    thr = target - 1e-4 * abs(target)
    em_first = next(i + 1 for i, ll in enumerate(em_curve) if ll >= thr)
    em_first_ms = em_ms * em_first / 50
    batches = max(1, round(ds.n_train / batch_sessions))
    curve = []
    reached = None
    # Minibatch steps have fewer sessions to spread over the chip than a full
    # EM pass: 4096-session work items were the fastest of 1024-8192 for them
    # (8192, the default, for EM).
    with ClickModel(DBN, batches=batches, item_sessions=ADAM_ITEM_SESSIONS) as cm:
        cm.init_from_dataset(ds)
        cm.set_test_set(test)
        cm.grad_init()
        cm.adam_configure(lr=lr)
        total = 0.0
        for ep in range(1, max_epochs + 1):
            total += cm.adam_steps_timed(batches)
            ll = cm.evaluate_test_set()["log_likelihood"]
            curve.append([ep, round(total, 3), round(ll, 6)])
            if ll >= thr:
                reached = (ep, total, ll)
                break
    log(f"time-to-target: EM 50 it {em_ms:.2f} ms -> LL {target:.6f}, first in band at iteration {em_first} "
        f"({em_first_ms:.3f} ms); Adam {curve[-1]}")
    # Algorithmic bytes of an Adam epoch (every session and entry once): per
    # session the 32-bit index of each shown rank and the 2-byte mask; per
    # entry and per-qd table the logit and both moments read, and the logit,
    # both moments and the value written (28 B).
    epoch_bytes = ds.n_train * (4 * 10 + 2) + 28 * 2 * ds.n_qd
    adam_frac = (epoch_bytes * reached[0] / (reached[1] * 1e-3) / HBM_PEAK) if reached else None
    return {
        "workload": f"DBN, {n_sessions} synthetic sessions ({ds.n_train} train), 1M Zipf(1.1) queries; Adam lr {lr}, "
                    f"{batches} minibatches of whole query groups (~{batch_sessions} sessions), log-prior on, "
                    f"{ADAM_ITEM_SESSIONS}-session work items (EM: 8192)",
        "target_heldout_ll": target,
        "threshold": "held-out LL at least EM's 50-iteration LL minus 1e-4 relative",
        "em_ms": em_ms,
        "em_first_iteration_in_band": em_first,
        "em_first_ms": em_first_ms,
        "em_curve": [round(x, 6) for x in em_curve],
        "adam_ms": reached[1] if reached else None,
        "adam_epochs": reached[0] if reached else None,
        "adam_steps": reached[0] * batches if reached else None,
        "adam_ll": reached[2] if reached else curve[-1][2],
        "speedup_vs_em": (em_ms / reached[1]) if reached else None,
        "speedup_vs_em_first": (em_first_ms / reached[1]) if reached else None,
        "adam_epoch_bytes": epoch_bytes,
        "adam_frac": adam_frac,
        "adam_curve": curve,
    }


def other_models(models, n_sessions, steps, warmup):
    """The other click models on their 10M-session workloads (same
    generator, each model's own ground truth, segmented reduction): time per
    graph-replayed iteration, the window kernel's event-timed duration and
    its fraction of the HBM roofline (algorithmic bytes as for the DBN, one
    per-qd table)."""
    from massiveclicks_amd import Dataset, ClickModel, synth
    from massiveclicks_amd.model import PBM, UBM, CCM
    ids = {"pbm": PBM, "ubm": UBM, "ccm": CCM}
    out = {}
    for name in models:
        t0 = time.perf_counter()
        s = synth.generate(synth.SynthConfig(n_sessions=n_sessions, n_queries=1_000_000, vocab=20_000_000,
                                             zipf_s=1.1, model=name, seed=3))
        ds = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2)
        del s
        with ClickModel(ids[name]) as cm:
            cm.init_from_dataset(ds)
            cm.em(warmup, use_graph=True)
            ms_iter = cm.em_graph_timed(steps) / steps
            kt = cm.em_timed(min(steps, 10))
            n_shared = cm.layout_info()["n_shared_entries"]
        item_bytes = ds.n_train * 42 + 16 * (ds.n_qd - n_shared)
        out[name] = {"ms_per_step": ms_iter, "value": ds.n_train / (ms_iter * 1e-3), "kernel_ms": kt["ms_estep"],
                     "bytes_per_launch": item_bytes, "frac": item_bytes / (kt["ms_estep"] * 1e-3) / HBM_PEAK,
                     "train_sessions": ds.n_train, "n_qd": ds.n_qd}
        log(f"{name.upper()} 10M: {ms_iter:.4f} ms/iter, kernel {kt['ms_estep']:.4f} ms, "
            f"frac {out[name]['frac']:.3f} ({time.perf_counter() - t0:.1f} s)")
        del ds
    return out


def calibration_copy(mib, reps=1):
    """Device-to-device copies of `mib` MiB: a plain copy kernel of known size
    (for scaling the FETCH_SIZE / WRITE_SIZE counters) and the copy bandwidth
    this device reaches, in GB/s counting bytes read plus bytes written (the
    best of `reps` event-timed copies after one untimed one)."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    a, b = C.c_void_p(), C.c_void_p()
    n = mib << 20
    nbytes = C.c_size_t(n)
    assert hip.hipMalloc(C.byref(a), nbytes) == 0 and hip.hipMalloc(C.byref(b), nbytes) == 0
    assert hip.hipMemset(a, 1, nbytes) == 0
    assert hip.hipMemcpy(b, a, nbytes, 3) == 0      # hipMemcpyDeviceToDevice
    e0, e1 = C.c_void_p(), C.c_void_p()
    assert hip.hipEventCreate(C.byref(e0)) == 0 and hip.hipEventCreate(C.byref(e1)) == 0
    best = None
    for _ in range(reps):
        assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpyAsync(b, a, nbytes, 3, None) == 0
        assert hip.hipEventRecord(e1, None) == 0
        assert hip.hipEventSynchronize(e1) == 0
        ms = C.c_float()
        assert hip.hipEventElapsedTime(C.byref(ms), e0, e1) == 0
        best = ms.value if best is None else min(best, ms.value)
    assert hip.hipDeviceSynchronize() == 0
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    return 2.0 * n / (best * 1e-3) / 1e9 if best else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--config", default="dbn10m", choices=["dbn10m", "batched"],
                    help="dbn10m: the headline plain-EM workload; batched: 100M sessions of 1-10 ranks, "
                         "10M Zipf(1.3) queries, autotuned")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--sessions", type=int, default=None)
    ap.add_argument("--queries", type=int, default=None)
    ap.add_argument("--zipf", type=float, default=None)
    ap.add_argument("--autotune", action="store_true", help="pick the launch options with mc_autotune")
    ap.add_argument("--no-autotune", action="store_true", help="batched configuration with the given options")
    ap.add_argument("--persistent", type=int, default=None, choices=[0, 1])
    ap.add_argument("--item-window", type=int, default=None)
    ap.add_argument("--model", default="dbn", choices=["pbm", "ubm", "dbn", "ccm"])
    ap.add_argument("--mode", default="segmented", choices=["segmented", "atomic"])
    # About 5 s of CPU work on 16 threads without the copy, 10 s with it.
    ap.add_argument("--cpu-iters", type=int, default=10)
    ap.add_argument("--cpu-copy-iters", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--item-sessions", type=int, default=None)
    ap.add_argument("--block", type=int, default=None)
    ap.add_argument("--calib-mib", type=int, default=0, help="also run a D2D copy of this size")
    ap.add_argument("--other-models", default="ccm,ubm,pbm",
                    help="also time these models on their 10M workloads ('' = none)")
    ap.add_argument("--ttt-sessions", type=int, default=20_000_000, help="time-to-target workload (0 = skip)")
    ap.add_argument("--ttt-batch", type=int, default=1_000_000)
    ap.add_argument("--ttt-lr", type=float, default=0.05)
    ap.add_argument("--ttt-max-epochs", type=int, default=40)
    args = ap.parse_args()
    batched = args.config == "batched"
    if batched:
        args.autotune = not args.no_autotune
        args.ttt_sessions = 0
    args.steps = args.steps or (20 if batched else 50)
    args.sessions = args.sessions or (100_000_000 if batched else 10_000_000)
    args.queries = args.queries or (10_000_000 if batched else 1_000_000)
    args.zipf = args.zipf or (1.3 if batched else 1.1)

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "query-sharded design is not built.", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from massiveclicks_amd import Dataset, ClickModel, synth
    from massiveclicks_amd.model import PBM, UBM, DBN, CCM, REDUCE_ATOMIC, REDUCE_SEGMENTED

    model = {"pbm": PBM, "ubm": UBM, "dbn": DBN, "ccm": CCM}[args.model]
    n_tab = 2 if model == DBN else 1

    t0 = time.perf_counter()
    cfg = synth.SynthConfig(n_sessions=args.sessions, n_queries=args.queries, vocab=20_000_000,
                            zipf_s=args.zipf, model=args.model, variable_length=batched, seed=3)
    s = synth.generate(cfg)
    stats = s.stats
    log(f"generated {args.sessions} sessions in {time.perf_counter() - t0:.1f} s: {stats}")
    t0 = time.perf_counter()
    ds = Dataset.from_sessions(s.query, s.docs, s.clicks, 0.2, length=s.length if batched else None)
    del s
    log(f"prepared: train {ds.n_train}, test {ds.n_test}, n_qd {ds.n_qd} in {time.perf_counter() - t0:.1f} s")

    opts = {"reduce_mode": REDUCE_ATOMIC if args.mode == "atomic" else REDUCE_SEGMENTED}
    if args.item_sessions:
        opts["item_sessions"] = args.item_sessions
    if args.block:
        opts["block"] = args.block
    if args.persistent is not None:
        opts["persistent"] = bool(args.persistent)
    if args.item_window:
        opts["item_window"] = args.item_window
    tune = None
    if args.autotune:
        from massiveclicks_amd import autotune
        t0 = time.perf_counter()
        opts, tune = autotune(model, ds.train, ds.n_qd, iterations=3, start=opts)
        tune["options"] = dict(opts)
        tune["seconds"] = round(time.perf_counter() - t0, 1)
        log(f"autotune: {tune}")
    cm = ClickModel(model, device=0, **opts)
    t0 = time.perf_counter()
    cm.init_from_dataset(ds)
    lay = cm.layout_info()
    log(f"init (layout + H2D) {time.perf_counter() - t0:.1f} s, layout {lay}")

    # Warmup, then the timed region: K iterations replayed from one graph.
    cm.em(args.warmup, use_graph=True)
    ms_total = cm.em_graph_timed(args.steps)
    ms_iter = ms_total / args.steps
    # Per-kernel durations from events around each launch (eager launches).
    kt = cm.em_timed(min(args.steps, 10))
    log(f"graph: {ms_iter:.4f} ms/iter; eager: item kernel {kt['ms_estep']:.4f} ms, merge {kt['ms_mstep']:.4f} ms")

    n_train = ds.n_train
    value = n_train / (ms_iter * 1e-3)
    n_shared = lay["n_shared_entries"]
    # Algorithmic bytes of the roofline: a 32-bit index per shown rank + 2 B
    # click mask (+ 1 B length) per session, 16 B per entry and table.
    mean_len = float(ds.train.length.mean()) if ds.train.length is not None else 10.0
    sess_bytes = n_train * (4 * mean_len + 2 + (1 if batched else 0))
    segmented = opts.get("reduce_mode", REDUCE_SEGMENTED) == REDUCE_SEGMENTED
    item_bytes = sess_bytes + 16 * n_tab * (ds.n_qd - n_shared) if segmented else sess_bytes
    iter_bytes = sess_bytes + 16 * n_tab * ds.n_qd
    achieved = item_bytes / (kt["ms_estep"] * 1e-3)
    # What this engine's own layout moves per launch: 16 B per session of a
    # packed window item (offsets, base, clicks and length in one record),
    # 2 B per shown rank for sessions of the other window items (16-bit
    # window-relative indices), 4 B for the rest, with the mask (and length),
    # over the sessions the kernels visit (DBN and CCM skip the sessions
    # without a click, whose counts are constants); and per entry an item owns
    # the value in and out of each table plus its 2-byte packed counts.
    si = cm.stream_info()
    n_w = si["n_window_sessions"]
    n_pk = si["n_packed"]
    extra = 2 + (1 if batched else 0)
    layout_sess = (n_pk * 16 + (n_w - n_pk) * (2 * mean_len + extra) +
                   (n_train - n_w) * (4 * mean_len + extra)) * (lay["n_em_sessions"] / n_train)
    layout_bytes = layout_sess + (8 * n_tab + 2) * (ds.n_qd - n_shared) if segmented else layout_sess

    # Held-out evaluation on the resident test set: the evaluation kernel's
    # duration (HIP events) against its algorithmic bytes: per session the
    # 32-bit index of every shown rank, the 2-byte click mask (+1 B length)
    # and one 4-byte value gather per shown rank and per-qd table.
    cm.set_test_set(ds.test)
    cm.evaluate_test_set()
    ev = min((cm.evaluate_test_set(timed=True) for _ in range(5)), key=lambda x: x["kernel_ms"])
    te_len = float(ds.test.length.mean()) if ds.test.length is not None else 10.0
    ev_bytes = ds.n_test * (4 * te_len + 2 + (1 if batched else 0) + 4 * n_tab * te_len)
    log(f"held-out LL {ev['log_likelihood']:.6f}, perplexity {ev['perplexity']:.6f}, "
        f"eval kernel {ev['kernel_ms']:.4f} ms")

    # HBM traffic per launch from the calibrated FETCH_SIZE / WRITE_SIZE
    # counter runs of this command (rocprofv3 cannot run inside it); the
    # profile's name travels with the figure.
    traffic, traffic_src = None, None
    tname = f"{args.model}_{args.sessions // 1_000_000}m_traffic.json"
    tfile = os.path.join(ROOT, "profiles", tname)
    if os.path.exists(tfile) and args.mode == "segmented" and not batched and not args.autotune:
        with open(tfile) as f:
            tj = json.load(f)
        traffic = tj.get("hbm_bytes_per_launch")
        traffic_src = f"profiles/{tname} ({tj.get('kernel')}, measured {tj.get('measured', 'see file')})"

    cpu = None
    if not args.no_cpu:
        # The GPU box's CPU share for one GPU is 16 threads.
        threads = max(1, min(os.cpu_count() or 1, 16))
        cpu = cpu_baseline(ds, model, args.cpu_iters, args.cpu_copy_iters, threads)
        log(f"cpu baseline: {cpu}")

    out = {
        "metric": METRIC if (args.model == "dbn" and not batched) else
        (f"EM sessions/sec/iter (batched uneven {args.model.upper()})" if batched
         else f"EM sessions/sec/iter ({args.model.upper()})"),
        "value": value,
        "unit": "sessions/s/iter",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_iter,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.model.upper()} plain EM, {args.sessions} synthetic sessions "
                               f"({n_train} train), {args.queries} Zipf({args.zipf}) queries, pools of 20, "
                               + ("1-10 ranks per session (uniform)" + (", autotuned options" if args.autotune else "")
                                  if batched else
                                  f"{args.mode} reduction" + (", autotuned options" if args.autotune else "")),
                   "sessions": args.sessions, "train_sessions": n_train, "n_qd": ds.n_qd,
                   "mean_ranks": mean_len, "top_query_share": stats.get("top_query_share"),
                   "iterations": args.steps, "options": {k: v for k, v in opts.items()},
                   "work_items": lay["n_items"], "autotune": tune},
        "roofline": {"bound": "hbm", "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": ("em_window_persistent_kernel" if segmented and opts.get("persistent", True)
                                and opts.get("block", 256) == 256 else "em_item_kernel"),
                     "bytes_per_launch": item_bytes,
                     "layout_bytes_per_launch": layout_bytes,
                     "layout_frac": layout_bytes / (kt["ms_estep"] * 1e-3) / HBM_PEAK,
                     "window_sessions": n_w, "packed_sessions": n_pk, "em_sessions": lay["n_em_sessions"],
                     "kernel_ms": kt["ms_estep"], "merge_kernel_ms": kt["ms_mstep"],
                     "iteration_bytes": iter_bytes,
                     "iteration_frac": iter_bytes / (ms_iter * 1e-3) / HBM_PEAK},
        "cpu_baseline": cpu,
        "heldout": {"log_likelihood": ev["log_likelihood"], "perplexity": ev["perplexity"],
                    "test_sessions": ds.n_test,
                    "eval_kernel": "eval_kernel (resident test set)", "eval_kernel_ms": ev["kernel_ms"],
                    "eval_bytes": ev_bytes,
                    "eval_frac": ev_bytes / (ev["kernel_ms"] * 1e-3) / HBM_PEAK},
    }
    cm.close()
    del ds
    if args.other_models and args.model == "dbn" and not batched:
        out["other_models"] = other_models(args.other_models.split(","), args.sessions, args.steps, args.warmup)
    if args.ttt_sessions and args.model == "dbn":
        out["time_to_target"] = time_to_target(args.ttt_sessions, args.ttt_batch, args.ttt_lr, args.ttt_max_epochs)
    if args.calib_mib:
        calibration_copy(args.calib_mib)
    # The copy bandwidth this device reaches (1 GiB device-to-device copy),
    # the practical ceiling beside the 8 TB/s peak.
    out["roofline"]["measured_copy_GBps"] = calibration_copy(1024, reps=5)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
