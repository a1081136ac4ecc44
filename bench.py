#!/usr/bin/env python3
"""Benchmark: train ms/tree at 1M x 100 (binary logistic, 10% NaN, depth 8,
256 bins) with the batched engine vs plain per-node launches, plus tree-ensemble
prediction rows/s and p50/p99 latency for the serving loop vs one plain launch
per request (500 trees, depth 8, 100 features).

One step = one boosting iteration (one tree) over the fixed 1M x 100 matrix,
inputs resident in HBM before the timed region.  The timed region is one
training call of the workload's 100 trees (--steps); ms/tree is its wall time
divided by the trees, so the call's fixed cost (base score, work-list set-up,
graph capture, the first round) is included.  Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import math
import os
import sys
import threading
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from boosters_amd import _abi as A  # noqa: E402

HBM_PEAK_GBPS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def gen_binary_nan(n, F, seed, w=None):
    """1M x 100 N(0,1), 10% NaN, label ~ Bernoulli(sigmoid(w . x[0:10])) on pre-mask x.

    `w` given: rows from the same labelling model (a held-out set of another
    seed); otherwise the seed's own draw is used."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((F, n), dtype=np.float32)
    mask = rng.random((F, n), dtype=np.float32) < 0.1
    w_seed = rng.standard_normal(10)
    w = w_seed if w is None else w
    score = w @ X[:10].astype(np.float64)
    y = (rng.random(n) < 1.0 / (1.0 + np.exp(-score))).astype(np.float32)
    X[mask] = np.nan
    return X, y, w


def cfg_c2(n_trees, depth):
    c = A.GBDTConfigC()
    A.lib().bst_gbdt_config_default(C.byref(c))
    c.objective = A.OBJ_LOGISTIC
    c.n_trees = n_trees
    c.max_depth = depth
    c.growth = A.GROW_DEPTHWISE
    return c


# (features per block, rows per block, blocks per CU, sub-histograms 1 LDS / 2 global);
# 0 = autotuned
HIST_SHAPE = (0, 0, 0, 0)


def ext(schedule):
    e = A.ExtConfig()
    A.lib().bst_ext_config_default(C.byref(e))
    e.schedule = schedule
    e.hist_fg, e.hist_rows_min, e.hist_blocks_per_cu, e.hist_priv = HIST_SHAPE
    return e


class Dev:
    """A device buffer allocated through the engine's ABI."""

    def __init__(self, ctx, arr):
        arr = np.ascontiguousarray(arr)
        self.ctx = ctx
        self.p = C.c_void_p()
        A.check(A.lib().bst_device_alloc(ctx, arr.nbytes, C.byref(self.p)))
        A.check(A.lib().bst_memcpy_h2d(ctx, self.p, arr.ctypes.data_as(C.c_void_p), arr.nbytes))

    @staticmethod
    def empty(ctx, nbytes):
        d = Dev.__new__(Dev)
        d.ctx = ctx
        d.p = C.c_void_p()
        A.check(A.lib().bst_device_alloc(ctx, nbytes, C.byref(d.p)))
        return d

    def free(self):
        A.lib().bst_device_free(self.ctx, self.p)


def train_timed(ctx, ds, y_dev, n, cfg, e):
    A.check(A.lib().bst_ctx_sync(ctx))
    h = C.c_void_p()
    t0 = time.perf_counter()
    A.check(A.lib().bst_train_device(ctx, ds, C.cast(y_dev.p, C.POINTER(C.c_float)), n, None,
                                     C.byref(cfg), C.byref(e), C.byref(h)))
    A.check(A.lib().bst_ctx_sync(ctx))
    return h, (time.perf_counter() - t0) * 1e3


PROFILE_C2 = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                          "k_hist_logistic_1Mx100_depth8.md")


def hist_traffic_per_launch():
    """HBM bytes per k_hist launch measured by rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE (calibrated against a copy kernel) in the committed C2 profile:
    the launch-weighted mean over its root and gathered histogram launches."""
    try:
        text = open(PROFILE_C2).read()
    except OSError:
        return None
    section = text.split("## HBM traffic", 1)[-1].split("\n## ", 1)[0]
    rows = [l for l in section.splitlines() if l.startswith("| `k_hist<")]
    n = tot = 0.0
    for l in rows:
        cells = [c.strip() for c in l.strip().strip("|").split("|")]
        launches, fetch_mb, write_mb = int(cells[1]), float(cells[2]), float(cells[3])
        n += launches
        tot += launches * (fetch_mb + write_mb) * 1e6
    return round(tot / n) if n else None


def tuning(ctx):
    """Histogram launch shape of the last train (autotuned or fixed)."""
    v = [C.c_int32() for _ in range(5)]
    ms = C.c_double()
    A.check(A.lib().bst_last_train_tuning(ctx, *[C.byref(x) for x in v], C.byref(ms)))
    return {"hist_fg": v[0].value, "hist_rows_min": v[1].value, "hist_blocks_per_cu": v[2].value,
            "hist_priv": "global" if v[3].value == 2 else "lds", "autotuned": bool(v[4].value),
            "tune_ms": round(ms.value, 1)}


def kernel_stats(ctx, kind):
    la, ms, by = C.c_int64(), C.c_double(), C.c_double()
    A.check(A.lib().bst_ctx_kernel_stats(ctx, kind, C.byref(la), C.byref(ms), C.byref(by)))
    return la.value, ms.value, by.value


def random_forest(n_trees, depth, F, seed):
    """Complete depth-`depth` trees: thresholds U[-5,5], leaves U[-1,1] (the
    shape of the reference's bench_medium model, benches/common/models.rs:14-66)."""
    rng = np.random.default_rng(seed)
    nn = (1 << (depth + 1)) - 1
    n_inner = (1 << depth) - 1
    info = A.ModelInfo()
    info.n_trees, info.n_groups, info.n_features = n_trees, 1, F
    info.total_nodes, info.total_cat_words = n_trees * nn, 0
    info.output_transform, info.objective = A.TRANSFORM_IDENTITY, A.OBJ_SQUARED
    a = A.alloc_forest(info)
    ids = np.arange(nn)
    for t in range(n_trees):
        s = slice(t * nn, (t + 1) * nn)
        leaf = ids >= n_inner
        a["is_leaf"][s] = leaf
        a["split_index"][s] = np.where(leaf, 0, rng.integers(0, F, nn))
        a["split_threshold"][s] = np.where(leaf, 0, rng.uniform(-5, 5, nn)).astype(np.float32)
        a["default_left"][s] = rng.integers(0, 2, nn)
        a["left"][s] = np.where(leaf, 0, 2 * ids + 1)
        a["right"][s] = np.where(leaf, 0, 2 * ids + 2)
        a["leaf_value"][s] = np.where(leaf, rng.uniform(-1, 1, nn), 0).astype(np.float32)
    a["tree_offset"][:] = np.arange(n_trees + 1) * nn
    return info, a


class HipEvents:
    """Kernel time on a given stream, from HIP events recorded on it."""

    def __init__(self):
        self.hip = C.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [C.POINTER(C.c_void_p)]
        self.hip.hipEventRecord.argtypes = [C.c_void_p, C.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [C.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [C.POINTER(C.c_float), C.c_void_p, C.c_void_p]
        self.hip.hipEventDestroy.argtypes = [C.c_void_p]
        self.a, self.b = C.c_void_p(), C.c_void_p()
        assert self.hip.hipEventCreate(C.byref(self.a)) == 0
        assert self.hip.hipEventCreate(C.byref(self.b)) == 0

    def time_ms(self, stream, fn):
        assert self.hip.hipEventRecord(self.a, stream) == 0
        fn()
        assert self.hip.hipEventRecord(self.b, stream) == 0
        assert self.hip.hipEventSynchronize(self.b) == 0
        ms = C.c_float()
        assert self.hip.hipEventElapsedTime(C.byref(ms), self.a, self.b) == 0
        return ms.value

    def close(self):
        self.hip.hipEventDestroy(self.a)
        self.hip.hipEventDestroy(self.b)


def predict_roofline(ctx, model, F, K, n_trees, depth, n_rows=1 << 20, reps=5):
    """bst_predict_device on n_rows device-resident rows, timed with HIP
    events on the context's stream: rows/s, node visits per clock per CU and
    the HBM fraction of the algorithmic bytes (F x 4 B in + K x 4 B out per
    row)."""
    L = A.lib()
    Xb = np.random.default_rng(5).standard_normal((F, n_rows), dtype=np.float32)
    xd = Dev(ctx, Xb)
    od = Dev.empty(ctx, n_rows * 4 * K)
    run = lambda: A.check(L.bst_predict_device(ctx, model, xd.p, n_rows, F, 1, od.p, None))  # noqa: E731
    run()
    A.check(L.bst_ctx_sync(ctx))
    ev = HipEvents()
    stream = L.bst_ctx_stream(ctx)
    ms = ev.time_ms(stream, lambda: [run() for _ in range(reps)]) / reps
    ev.close()
    xd.free()
    od.free()
    rows_s = n_rows / (ms * 1e-3)
    alg = n_rows * (F + K) * 4.0
    achieved = alg / (ms * 1e-3) / 1e9
    return {
        "kernel": "k_predict_lds", "rows": n_rows, "avg_launch_us": round(ms * 1e3, 1),
        "rows_per_s": rows_s, "node_visits_per_clk_per_cu": round(rows_s * n_trees * depth / (256 * 2.4e9), 3),
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 4),
    }


def serving_requests(F, n_req, seed):
    """Request row counts log-uniform in [1, 100000], features N(0,1)."""
    rng = np.random.default_rng(seed)
    rows = np.floor(np.exp(rng.uniform(0, math.log(100001), n_req))).astype(np.int64)
    reqs = [rng.standard_normal((F, int(r)), dtype=np.float32) for r in rows]
    return rows, reqs


def serving_bench(ctx, model, F, n_req, qps, seed, budget, inflight, reqs=None, plain=True):
    """Poisson arrivals at `qps`, rows log-uniform in [1, 100000]."""
    if reqs is None:
        rows, reqs = serving_requests(F, n_req, seed)
    else:
        rows, reqs = reqs
    rng = np.random.default_rng(seed + 1)
    gaps = rng.exponential(1.0 / qps, n_req)
    arrivals = np.cumsum(gaps) - gaps[0]
    outs = [np.zeros(int(r), np.float32) for r in rows]
    total_rows = int(rows.sum())

    # Plain: one synchronous predict launch per request, in arrival order.
    lat_plain = np.zeros(n_req)
    span_plain = float("nan")
    plain_outs = None
    if plain:
        t0 = time.perf_counter()
        for i, X in enumerate(reqs):
            now = time.perf_counter() - t0
            if now < arrivals[i]:
                time.sleep(arrivals[i] - now)
            A.check(A.lib().bst_predict(ctx, model, A.ptr(X), X.shape[1], F, 1, A.ptr(outs[i]), None))
            lat_plain[i] = time.perf_counter() - t0 - arrivals[i]
        span_plain = time.perf_counter() - t0
        plain_outs = [o.copy() for o in outs]

    # Server: continuous batching.
    srv = C.c_void_p()
    A.check(A.lib().bst_server_create(ctx, model, int(budget), int(inflight), C.byref(srv)))
    tickets = [None] * n_req
    lat_srv = np.zeros(n_req)
    done = threading.Event()
    t_end = [0.0]

    def waiter():
        for i in range(n_req):
            while tickets[i] is None:
                time.sleep(1e-5)
            ln = C.c_int64()
            A.check(A.lib().bst_server_wait(srv, C.c_uint64(tickets[i]), C.byref(ln)))
            lat_srv[i] = ln.value * 1e-9
        t_end[0] = time.perf_counter()
        done.set()

    th = threading.Thread(target=waiter)
    for o in outs:
        o[:] = 0
    th.start()
    t0 = time.perf_counter()
    for i, X in enumerate(reqs):
        now = time.perf_counter() - t0
        if now < arrivals[i]:
            time.sleep(arrivals[i] - now)
        t = C.c_uint64()
        A.check(A.lib().bst_server_submit(srv, A.ptr(X), X.shape[1], F, A.ptr(outs[i]), C.byref(t)))
        tickets[i] = t.value
    done.wait()
    th.join()
    span_srv = t_end[0] - t0
    A.check(A.lib().bst_server_destroy(srv))
    res = {
        "requests": n_req,
        "qps_offered": qps,
        "rows_total": total_rows,
        "rows_per_s_offered": total_rows / max(arrivals[-1], 1e-9),
        "server": {
            "rows_per_s": total_rows / span_srv,
            "p50_ms": float(np.percentile(lat_srv, 50) * 1e3),
            "p99_ms": float(np.percentile(lat_srv, 99) * 1e3),
        },
    }
    if plain:
        res["plain"] = {
            "rows_per_s": total_rows / span_plain,
            "p50_ms": float(np.percentile(lat_plain, 50) * 1e3),
            "p99_ms": float(np.percentile(lat_plain, 99) * 1e3),
        }
        res["results_identical"] = all(np.array_equal(a, b) for a, b in zip(plain_outs, outs))
    return res


def serving_sweep(ctx, model, F, n_req, qps_list, seed, budget, inflight, slo_ms):
    """The same requests offered at increasing Poisson rates.  Capacity of a
    mode = the highest delivered rows/s among the rates at which its p99
    latency stays within slo_ms (the same latency bound for both modes)."""
    reqs = serving_requests(F, n_req, seed)
    points = []
    for q in qps_list:
        r = serving_bench(ctx, model, F, n_req, q, seed, budget, inflight, reqs=reqs)
        log(f"serving at {q:.0f} QPS: server {r['server']['rows_per_s'] / 1e6:.1f} M rows/s "
            f"p99 {r['server']['p99_ms']:.1f} ms | plain {r['plain']['rows_per_s'] / 1e6:.1f} M rows/s "
            f"p99 {r['plain']['p99_ms']:.1f} ms")
        points.append(r)
    cap = {}
    for mode in ("server", "plain"):
        ok = [p[mode]["rows_per_s"] for p in points if p[mode]["p99_ms"] <= slo_ms]
        cap[mode] = max(ok) if ok else None
    return {
        "slo_p99_ms": slo_ms,
        "capacity_rows_per_s": cap,
        "points": [{"qps": p["qps_offered"], "server": p["server"], "plain": p["plain"],
                    "identical": p["results_identical"]} for p in points],
    }


def cpu_setup():
    """The CPU baseline's oracle build and thread count on this host.

    The oracle (a C++ restatement of the reference CPU path, the parity
    checker) is recompiled here with -O3 -march=native for timing; the
    threads are this process's CPU share (OMP_NUM_THREADS when set, else the
    affinity mask), and the host's core count and CPU model are recorded."""
    from oracle import oracle as O

    info = {"nproc": os.cpu_count()}
    try:
        info["affinity_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["affinity_cpus"] = os.cpu_count()
    try:
        with open("/proc/cpuinfo") as f:
            info["cpu_model"] = next((ln.split(":", 1)[1].strip() for ln in f
                                      if ln.startswith("model name")), "unknown")
    except OSError:
        info["cpu_model"] = "unknown"
    try:
        import tempfile
        O.use_library(O.build_native(tempfile.mkdtemp(prefix="bst_oracle_")))
        info["build"] = "g++ -O3 -march=native -fopenmp"
    except Exception as e:  # noqa: BLE001 - fall back to the shipped checker build
        log(f"native oracle build failed ({e}); timing the -O2 build")
        info["build"] = "g++ -O2 -fopenmp (native build failed)"
    share = int(os.environ.get("OMP_NUM_THREADS") or 0) or info["affinity_cpus"]
    threads = max(1, min(share, info["affinity_cpus"]))
    return O, threads, info


def cpu_baseline(X, y, n_trees, depth):
    """The oracle on the same data: ms/tree with the CPU share's threads
    (n_trees trees) and with one thread (one tree); binning timed apart."""
    O, threads, info = cpu_setup()
    O.set_threads(threads)
    t0 = time.perf_counter()
    b = O.Binned(X)
    t_bin = time.perf_counter() - t0
    m = O.train(b, y, cfg_c2(n_trees, depth))
    _, ms = m.stats()
    O.set_threads(1)
    m1 = O.train(b, y, cfg_c2(1, depth))
    _, ms1 = m1.stats()
    O.set_threads(threads)
    return ms / n_trees, ms1, t_bin, threads, info


# ---------------------------------------------------------------------------
# Other workloads (C1 regression, C3 softmax with categoricals, C4 ranking).
# Same contract: one JSON line; `value` = train ms per tree (graph schedule).
# ---------------------------------------------------------------------------
def gen_c1(n, F, seed):
    """X ~ N(0,1); y = x0 + 0.5 x1^2 + N(0, 0.1^2)."""
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((F, n), dtype=np.float32)
    y = (X[0] + 0.5 * X[1].astype(np.float64) ** 2 + rng.normal(0, 0.1, n)).astype(np.float32)
    return X, y, None, None


def gen_c3(n, n_num, n_cat, K, card, seed):
    """Numeric N(0,1) + categorical ints {0..card-1}; class = argmax_k W_k x + sum_j C_kj[x_j] + Gumbel."""
    rng = np.random.default_rng(seed)
    Xn = rng.standard_normal((n_num, n), dtype=np.float32)
    Xc = rng.integers(0, card, size=(n_cat, n), dtype=np.int32)
    W = rng.normal(0, 1.0 / math.sqrt(n_num), size=(K, n_num))
    Cm = rng.normal(0, 1.0 / math.sqrt(n_cat), size=(K, n_cat, card))
    z = W @ Xn.astype(np.float64)
    for j in range(n_cat):
        z += Cm[:, j, Xc[j]]
    z += rng.gumbel(size=(K, n))
    y = np.argmax(z, axis=0).astype(np.float32)
    X = np.concatenate([Xn, Xc.astype(np.float32)], axis=0)
    is_cat = np.zeros(n_num + n_cat, np.uint8)
    is_cat[n_num:] = 1
    return X, y, is_cat, None


def gen_c4(n, F, seed, zero_p=0.3):
    """X ~ U[0,1] with exact zeros (p 0.3); query sizes U{5..500}; label = quintile of
    s = sum_f a_f [x_f > t_f]."""
    rng = np.random.default_rng(seed)
    X = rng.random((F, n), dtype=np.float32)
    X[rng.random((F, n), dtype=np.float32) < zero_p] = 0.0
    sizes = []
    tot = 0
    while tot < n:
        sz = min(int(rng.integers(5, 501)), n - tot)
        sizes.append(sz)
        tot += sz
    a = rng.standard_normal(F).astype(np.float32)
    t = rng.random(F, dtype=np.float32)
    s = np.zeros(n, np.float32)
    for f in range(F):
        s += a[f] * (X[f] > t[f])
    q = np.quantile(s, [0.2, 0.4, 0.6, 0.8])
    y = np.searchsorted(q, s, side="right").astype(np.float32)
    return X, y, None, np.array(sizes, np.int32)


WORKLOADS = {
    # name: (description, generator, config overrides, ext overrides, default trees)
    "c1": ("regression 5000 x 100, depth-wise 6, lr 0.1",
           lambda: gen_c1(5000, 100, 0),
           dict(objective=A.OBJ_SQUARED, growth=A.GROW_DEPTHWISE, max_depth=6, learning_rate=0.1),
           {}, 50),
    "c3": ("softmax K=10, 4000000 x (180 numeric + 20 categorical), leaf-wise 255",
           lambda: gen_c3(4_000_000, 180, 20, 10, 32, 0),
           dict(objective=A.OBJ_SOFTMAX, n_classes=10, growth=A.GROW_LEAFWISE, max_leaves=255),
           {}, 5),
    "c4": ("LambdaRank 2000000 x 136, GOSS 0.2/0.1, colsample 0.5, leaf-wise 127",
           lambda: gen_c4(2_000_000, 136, 0),
           dict(objective=A.OBJ_LAMBDARANK, growth=A.GROW_LEAFWISE, max_leaves=127,
                colsample_bytree=0.5),
           dict(goss_top_rate=0.2, goss_other_rate=0.1), 10),
}


def run_workload(args, L, ctx, name=None, extra=False):
    """Train one workload of WORKLOADS on the graph schedule and return its
    bench line.  extra: a side measurement of the default run (its default
    tree count, no CPU baseline)."""
    name = name or args.workload
    desc, gen, cfg_kw, ext_kw, default_steps = WORKLOADS[name]
    steps = args.steps if (args.steps_given and not extra) else default_steps
    log(f"{name}: generating {desc}")
    X, y, is_cat, sizes = gen()
    F, n = X.shape
    t0 = time.perf_counter()
    ds = C.c_void_p()
    A.check(L.bst_dataset_create(ctx, A.ptr(X), n, F, A.ptr(is_cat), None, C.byref(ds)))
    bin_ms = (time.perf_counter() - t0) * 1e3
    y_dev = Dev(ctx, y)
    log(f"binned in {bin_ms:.0f} ms")

    def cfg_of(n_trees):
        c = A.GBDTConfigC()
        L.bst_gbdt_config_default(C.byref(c))
        for k, v in cfg_kw.items():
            setattr(c, k, v)
        c.n_trees = n_trees
        return c

    def ext_of():
        e = ext(A.SCHED_GRAPH)
        for k, v in ext_kw.items():
            setattr(e, k, v)
        if sizes is not None:
            e.query_group_sizes = A.ptr(sizes)
            e.n_query_groups = len(sizes)
        return e

    K = cfg_kw.get("n_classes", 1) if cfg_kw["objective"] == A.OBJ_SOFTMAX else 1
    rounds = max(1, steps)
    if args.warmup > 0:
        h, _ = train_timed(ctx, ds, y_dev, n, cfg_of(args.warmup), ext_of())
        L.bst_model_destroy(h)
    h_model, wall_ms = train_timed(ctx, ds, y_dev, n, cfg_of(rounds), ext_of())
    tune = tuning(ctx)
    n_trees = rounds * K
    ms_tree = wall_ms / n_trees
    log(f"graph: {ms_tree:.3f} ms/tree ({n_trees} trees in {wall_ms:.1f} ms)")
    if args.graph_only and not extra:
        L.bst_model_destroy(h_model)
        L.bst_dataset_destroy(ds)
        return {"metric": f"train ms/tree ({desc}, graph schedule only)",
                "value": round(ms_tree, 4), "unit": "ms/tree", "steps": rounds,
                "hist_tuning": tune}
    A.check(L.bst_ctx_set_profiling(ctx, 1))
    h_prof, _ = train_timed(ctx, ds, y_dev, n, cfg_of(rounds), ext_of())
    hist = kernel_stats(ctx, 0)
    split = kernel_stats(ctx, 1)
    part = kernel_stats(ctx, 2)
    A.check(L.bst_ctx_set_profiling(ctx, 0))
    L.bst_model_destroy(h_prof)
    launches, hms, hbytes = hist
    achieved = hbytes / (hms * 1e-3) / 1e9 if launches else 0.0
    result = {
        "metric": f"train ms/tree ({desc})",
        "value": round(ms_tree, 4),
        "unit": "ms/tree",
        "n_gpus": 1,
        "steps": rounds,
        "warmup": args.warmup,
        "ms_per_step": round(wall_ms / rounds, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 grad/hess, int64 fixed-point histograms, u8 bins",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "schedule": "graph",
            "hist_tuning": tune,
            "trees_per_step": K,
            "binning_ms": round(bin_ms, 1),
            "kernel_ms_per_tree": {
                "hist": round(hms / n_trees, 4),
                "split": round(split[1] / n_trees, 4),
                "partition": round(part[1] / n_trees, 4),
            },
        },
        "roofline": {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None, "kernel": "k_hist",
            "launches": launches,
            "avg_launch_us": round(hms / launches * 1e3, 2) if launches else None,
        },
    }
    if not args.no_cpu_baseline and not extra:
        O, threads, cinfo = cpu_setup()
        O.set_threads(threads)
        m = min(n, args.cpu_rows)
        log(f"cpu baseline: oracle, {args.cpu_trees} round(s) on the first {m} rows, {threads} threads")
        Xs = np.ascontiguousarray(X[:, :m])
        ys = y[:m]
        b = O.Binned(Xs, is_cat)
        e = ext_of()
        e.schedule = 0
        if sizes is not None:
            # whole query groups covering the first m rows
            cs = np.cumsum(sizes)
            ng = int(np.searchsorted(cs, m, side="left")) + 1
            ss = sizes[:ng].copy()
            ss[-1] -= int(cs[ng - 1] - m)
            ss = ss[ss > 0]
            e.query_group_sizes = A.ptr(ss)
            e.n_query_groups = len(ss)
        cm = O.train(b, ys, cfg_of(args.cpu_trees), e)
        _, ms = cm.stats()
        result["cpu_baseline"] = {
            "value": round(ms / (args.cpu_trees * K), 2),
            "unit": "ms/tree",
            "cores": threads,
            "kind": "port",
            "sample": f"first {m} of {n} rows, {args.cpu_trees} round(s) ({args.cpu_trees * K} trees)",
            **cinfo,
        }
    L.bst_model_destroy(h_model)
    L.bst_dataset_destroy(ds)
    y_dev.free()
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100,
                    help="trees of the timed training call (the workload trains 100)")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=1_000_000)
    ap.add_argument("--features", type=int, default=100)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--plain-steps", type=int, default=3)
    ap.add_argument("--cpu-trees", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--requests", type=int, default=300)
    ap.add_argument("--qps", type=float, default=1000.0)
    ap.add_argument("--qps-sweep", default="500,1000,2000,2500,3000,3500,4000,6000,8000",
                    help="offered request rates of the serving capacity sweep ('' = none)")
    ap.add_argument("--slo-ms", type=float, default=50.0,
                    help="p99 latency bound at which serving capacity is compared")
    ap.add_argument("--workload", choices=["c2", "c1", "c3", "c4"], default="c2",
                    help="c2 is the headline workload; the others print their own line")
    ap.add_argument("--cpu-rows", type=int, default=200_000,
                    help="rows of the CPU baseline sample for c1/c3/c4")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run the plain copy kernel 5x on this many bytes (HBM counter calibration)")
    ap.add_argument("--predict-only", action="store_true",
                    help="only the device-batch prediction of the 500-tree depth-8 model (the run "
                         "a kernel trace of the traversal kernel is taken on)")
    ap.add_argument("--no-leafwise", action="store_true",
                    help="skip the leaf-wise workloads (softmax K=10 255 leaves, LambdaRank 127 "
                         "leaves) measured beside the headline line")
    ap.add_argument("--graph-only", action="store_true",
                    help="only the warmup and timed graph-schedule runs (no per-kernel event run, "
                         "no batched or plain schedule): the run a kernel trace of the graph "
                         "schedule is taken on")
    ap.add_argument("--hist-shape", default="",
                    help="fixed histogram launch shape FG,ROWS_MIN,BLOCKS_PER_CU[,PRIV] instead of "
                         "the autotuner's choice (e.g. 32,2048,1; PRIV 1 = LDS sub-histograms, "
                         "2 = global memory, default 1)")
    args = ap.parse_args()
    if args.hist_shape:
        global HIST_SHAPE
        v = [int(x) for x in args.hist_shape.split(",")]
        HIST_SHAPE = tuple(v + [1] * (4 - len(v)))
    args.steps_given = any(a == "--steps" or a.startswith("--steps=") for a in sys.argv[1:])
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (row sharding "
              "across GPUs is designed, not built)", file=sys.stderr)
        sys.exit(2)

    L = A.lib()
    ctx = C.c_void_p()
    A.check(L.bst_ctx_create(0, C.byref(ctx)))
    if args.copy_bytes:
        nbytes = args.copy_bytes // 16 * 16
        src, dst = Dev.empty(ctx, nbytes), Dev.empty(ctx, nbytes)
        for _ in range(5):
            A.check(L.bst_kernel_copy(ctx, dst.p, src.p, nbytes))
        print(json.dumps({"copy_bytes": nbytes, "launches": 5}), flush=True)
        return
    if args.workload != "c2":
        print(json.dumps(run_workload(args, L, ctx)), flush=True)
        return
    if args.predict_only:
        info, arrays = random_forest(500, 8, 100, seed=0xB002)
        fa = A.forest_struct(arrays)
        mh = C.c_void_p()
        A.check(L.bst_model_import(ctx, C.byref(info), C.byref(fa), C.byref(mh)))
        pr = predict_roofline(ctx, mh, 100, 1, 500, 8)
        L.bst_model_destroy(mh)
        print(json.dumps({"metric": "device-batch prediction rows/s (500 trees, depth 8, 100 features)",
                          "value": round(pr["rows_per_s"]), "unit": "rows/s", "roofline": pr}), flush=True)
        return
    n, F, depth = args.rows, args.features, args.depth
    log(f"generating {n} x {F} binary-logistic data with 10% NaN")
    X, y, w_true = gen_binary_nan(n, F, seed=0)
    t0 = time.perf_counter()
    ds = C.c_void_p()
    A.check(L.bst_dataset_create(ctx, A.ptr(X), n, F, None, None, C.byref(ds)))
    bin_ms = (time.perf_counter() - t0) * 1e3
    y_dev = Dev(ctx, y)
    log(f"binned in {bin_ms:.0f} ms")

    # Warmup (untimed), then exactly `steps` timed boosting iterations on the
    # graph schedule (device-side growth, one hipGraph replay per iteration).
    if args.warmup > 0:
        h, _ = train_timed(ctx, ds, y_dev, n, cfg_c2(args.warmup, depth), ext(A.SCHED_GRAPH))
        L.bst_model_destroy(h)
    h_model, wall_ms = train_timed(ctx, ds, y_dev, n, cfg_c2(args.steps, depth), ext(A.SCHED_GRAPH))
    tune = tuning(ctx)
    ms_tree = wall_ms / args.steps
    log(f"graph: {ms_tree:.3f} ms/tree ({args.steps} trees in {wall_ms:.1f} ms)")

    if args.graph_only:
        L.bst_model_destroy(h_model)
        print(json.dumps({"metric": "train ms/tree (graph schedule only)", "value": round(ms_tree, 4),
                          "unit": "ms/tree", "steps": args.steps, "warmup": args.warmup,
                          "hist_tuning": tune}), flush=True)
        return

    # Kernel statistics: the same schedule launched eagerly with HIP events
    # around each kernel on the training stream (a separate, untimed run).
    A.check(L.bst_ctx_set_profiling(ctx, 1))
    h_prof, _ = train_timed(ctx, ds, y_dev, n, cfg_c2(args.steps, depth), ext(A.SCHED_GRAPH))
    hist = kernel_stats(ctx, 0)
    split = kernel_stats(ctx, 1)
    part = kernel_stats(ctx, 2)
    A.check(L.bst_ctx_set_profiling(ctx, 0))
    L.bst_model_destroy(h_prof)

    h_b, batched_ms = train_timed(ctx, ds, y_dev, n, cfg_c2(args.steps, depth), ext(A.SCHED_BATCHED))
    L.bst_model_destroy(h_b)
    batched_ms_tree = batched_ms / args.steps
    log(f"batched: {batched_ms_tree:.3f} ms/tree")
    plain_ms_tree = None
    if args.plain_steps > 0:
        h_plain, plain_ms = train_timed(ctx, ds, y_dev, n, cfg_c2(args.plain_steps, depth), ext(A.SCHED_PLAIN))
        L.bst_model_destroy(h_plain)
        plain_ms_tree = plain_ms / args.plain_steps
        log(f"plain: {plain_ms_tree:.3f} ms/tree")

    # Held-out validation loss of the timed model: n/4 new rows (another seed)
    # labelled by the training data's model.
    Xv, yv, _ = gen_binary_nan(n // 4, F, seed=1, w=w_true)

    # The same training with that validation set (per-round validation
    # prediction and metric inside the round's graph), from host targets.
    A.check(L.bst_ctx_sync(ctx))
    hv = C.c_void_p()
    t0 = time.perf_counter()
    A.check(L.bst_train(ctx, ds, A.ptr(y), n, None, A.ptr(Xv), A.ptr(yv), len(yv),
                        C.byref(cfg_c2(args.steps, depth)), C.byref(ext(A.SCHED_GRAPH)), C.byref(hv)))
    A.check(L.bst_ctx_sync(ctx))
    valid_ms_tree = (time.perf_counter() - t0) * 1e3 / args.steps
    L.bst_model_destroy(hv)
    log(f"graph with a {len(yv)}-row validation set: {valid_ms_tree:.3f} ms/tree")
    pv = np.zeros(len(yv), np.float32)
    A.check(L.bst_predict(ctx, h_model, A.ptr(Xv), len(yv), F, 0, A.ptr(pv), None))
    p = np.clip(pv.astype(np.float64), 1e-15, 1 - 1e-15)
    val_logloss = float(-np.mean(yv * np.log(p) + (1 - yv) * np.log(1 - p)))

    launches, hms, hbytes = hist
    roofline = None
    if launches:
        achieved = hbytes / (hms * 1e-3) / 1e9
        roofline = {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBPS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBPS, 4),
            "traffic": hist_traffic_per_launch(),
            "traffic_source": "profiles/k_hist_logistic_1Mx100_depth8.md (PMC FETCH_SIZE + WRITE_SIZE)",
            "kernel": "k_hist",
            "launches": launches,
            "avg_launch_us": round(hms / launches * 1e3, 2),
            "alg_bytes_per_launch": round(hbytes / launches),
        }

    result = {
        "metric": "train ms/tree at 1M x 100 (binary logistic, 10% NaN, depth 8, 256 bins)",
        "value": round(ms_tree, 4),
        "unit": "ms/tree",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_tree, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 grad/hess, int64 fixed-point histograms, u8 bins",
        "data": "synthetic",
        "config": {
            "workload": f"binary logistic {n} x {F}, 10% NaN, depth-wise {depth}, 256 bins, lr 0.3",
            "schedule": "graph (device-side growth, one hipGraph replay per boosting iteration)",
            "hist_tuning": tune,
            "batched_ms_per_tree": round(batched_ms_tree, 4),
            "plain_ms_per_tree": None if plain_ms_tree is None else round(plain_ms_tree, 4),
            "graph_speedup_vs_plain": None if plain_ms_tree is None else round(plain_ms_tree / ms_tree, 2),
            "binning_ms": round(bin_ms, 1),
            "val_logloss": round(val_logloss, 6),
            "with_validation_ms_per_tree": round(valid_ms_tree, 4),
            "validation_rows": len(yv),
            "kernel_ms_per_tree": {
                "hist": round(hms / args.steps, 4),
                "split": round(split[1] / args.steps, 4),
                "partition": round(part[1] / args.steps, 4),
            },
        },
        "roofline": roofline,
    }

    if not args.no_serving:
        log("serving: 500 trees x depth 8 x 100 features")
        info, arrays = random_forest(500, 8, 100, seed=0xB002)
        fa = A.forest_struct(arrays)
        mh = C.c_void_p()
        A.check(L.bst_model_import(ctx, C.byref(info), C.byref(fa), C.byref(mh)))
        # Device-resident batch throughput, timed with HIP events.
        pr = predict_roofline(ctx, mh, 100, 1, 500, 8)
        log(f"predict: {pr['rows_per_s'] / 1e6:.1f} M rows/s, "
            f"{pr['node_visits_per_clk_per_cu']} node visits/clk/CU")
        sv = serving_bench(ctx, mh, 100, args.requests, args.qps, seed=7, budget=4 << 30, inflight=4)
        sv["device_batch_rows_per_s"] = pr["rows_per_s"]
        sv["roofline"] = pr
        if args.qps_sweep:
            sv["sweep"] = serving_sweep(ctx, mh, 100, args.requests,
                                        [float(q) for q in args.qps_sweep.split(",")], seed=7,
                                        budget=4 << 30, inflight=4, slo_ms=args.slo_ms)
        result["predict"] = sv
        L.bst_model_destroy(mh)

    if not args.no_cpu_baseline:
        log(f"cpu baseline: oracle, {args.cpu_trees} trees on the same data (+1 tree on one thread)")
        cpu_ms_tree, cpu_ms1, cpu_bin_s, threads, cinfo = cpu_baseline(X, y, args.cpu_trees, depth)
        result["cpu_baseline"] = {
            "value": round(cpu_ms_tree, 2),
            "unit": "ms/tree",
            "cores": threads,
            "kind": "port",
            "sample": f"same {n} x {F} matrix, {args.cpu_trees} trees on {threads} threads and 1 tree "
                      f"on 1 thread (binning {cpu_bin_s:.1f} s excluded)",
            "single_thread_ms_per_tree": round(cpu_ms1, 2),
            **cinfo,
        }
    L.bst_model_destroy(h_model)
    if not args.no_leafwise:
        # The leaf-wise workloads, measured the same way beside the headline
        # line (their own data, graph schedule, per-kernel event run).
        L.bst_dataset_destroy(ds)
        ds = None
        y_dev.free()
        lw = {}
        for w in ("c4", "c3"):
            r = run_workload(args, L, ctx, name=w, extra=True)
            lw[w] = {"workload": r["config"]["workload"], "ms_per_tree": r["value"],
                     "trees": r["steps"] * r["config"]["trees_per_step"],
                     "kernel_ms_per_tree": r["config"]["kernel_ms_per_tree"],
                     "hist_tuning": r["config"]["hist_tuning"], "hist_frac": r["roofline"]["frac"]}
        result["config"]["leafwise"] = lw
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
