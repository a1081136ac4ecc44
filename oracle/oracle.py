"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the parity checker; the engine never imports this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from boosters_amd import _abi as A

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libbst_oracle.so")
_LIB = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build()
        _LIB = _load(LIB_PATH)
    return _LIB


def build_native(out_dir="/tmp"):
    """Compile the -O3 -march=native timing build of the same restatement on
    this machine (for the CPU baseline); returns its path."""
    subprocess.run(["make", "-s", "-C", HERE, "native", f"OUT_DIR={out_dir}"], check=True)
    return os.path.join(out_dir, "libbst_oracle_native.so")


def use_library(path):
    """Route every oracle call through another build of bst_oracle.cpp."""
    global _LIB
    _LIB = _load(path)


def _load(path):
    L = C.CDLL(path)
    P = C.POINTER
    vp = C.c_void_p
    sig = {
        "ora_last_error": (C.c_char_p, []),
        "ora_num_threads": (C.c_int, []),
        "ora_set_num_threads": (None, [C.c_int]),
        "ora_set_exp_mode": (None, [C.c_int]),
        "ora_bin": (C.c_int, [P(C.c_float), C.c_int64, C.c_int32, P(C.c_uint8),
                              P(A.BinningConfig), P(vp)]),
        "ora_binned_free": (None, [vp]),
        "ora_feature_info": (C.c_int, [vp, C.c_int32, P(C.c_uint32), P(C.c_int32),
                                       P(C.c_uint32), P(C.c_int32), P(C.c_int32)]),
        "ora_bin_bounds": (C.c_int, [vp, C.c_int32, P(C.c_double), C.c_int64]),
        "ora_get_bins": (C.c_int, [vp, P(C.c_uint16)]),
        "ora_value_to_bin": (C.c_uint32, [vp, C.c_int32, C.c_double]),
        "ora_leaf_weight": (C.c_float, [P(A.GBDTConfigC), C.c_double, C.c_double]),
        "ora_hist_layout": (C.c_int, [vp, P(C.c_int32), P(C.c_uint32), P(C.c_uint32),
                                      P(C.c_uint32)]),
        "ora_gradients": (C.c_int, [C.c_int32, C.c_int32, P(C.c_float), P(C.c_float),
                                    P(C.c_float), C.c_int64, P(C.c_float), P(C.c_float)]),
        "ora_lambdarank_gradients": (C.c_int, [P(C.c_float), P(C.c_float), P(C.c_float),
                                               C.c_int64, P(C.c_int32), C.c_int64,
                                               P(C.c_float), P(C.c_float)]),
        "ora_base_score": (C.c_int, [C.c_int32, C.c_int32, P(C.c_float), P(C.c_float),
                                     C.c_int64, P(C.c_float)]),
        "ora_histogram": (C.c_int, [vp, P(C.c_float), P(C.c_float), P(C.c_uint32),
                                    C.c_int64, P(C.c_double)]),
        "ora_find_split": (C.c_int, [vp, P(C.c_double), C.c_double, C.c_double, C.c_uint32,
                                     P(A.GBDTConfigC), C.c_int32, P(C.c_uint32), C.c_int32,
                                     P(A.SplitInfoC)]),
        "ora_partition": (C.c_int, [vp, P(C.c_uint32), C.c_int64, P(A.SplitInfoC),
                                    P(C.c_int64)]),
        "ora_train": (C.c_int, [vp, P(C.c_float), P(C.c_float), C.c_int64, P(C.c_float),
                                P(C.c_float), P(C.c_float), C.c_int64, P(A.GBDTConfigC),
                                P(A.ExtConfig), P(vp)]),
        "ora_model_info": (C.c_int, [vp, P(A.ModelInfo)]),
        "ora_model_export": (C.c_int, [vp, P(A.ForestArrays)]),
        "ora_model_metric": (C.c_int, [vp, P(C.c_double), C.c_int64, P(C.c_int64)]),
        "ora_model_stats": (C.c_int, [vp, P(C.c_int64), P(C.c_double)]),
        "ora_model_free": (None, [vp]),
        "ora_predict": (C.c_int, [P(A.ModelInfo), P(A.ForestArrays), P(C.c_float),
                                  C.c_int64, C.c_int32, C.c_int32, P(C.c_float),
                                  P(C.c_uint32)]),
        "ora_metric": (C.c_double, [C.c_int32, P(C.c_float), C.c_int32, P(C.c_float),
                                    P(C.c_float), C.c_int64]),
        "ora_splitmix64_stream": (None, [C.c_uint64, C.c_int64, P(C.c_uint64)]),
        "ora_partition_predict": (C.c_int, [vp, P(A.ModelInfo), P(A.ForestArrays),
                                            P(C.c_uint8), P(C.c_float), P(C.c_float)]),
        "ora_binning_sample": (C.c_int64, [C.c_int64, C.c_uint64, P(C.c_uint32)]),
        "ora_xoshiro256pp": (None, [P(C.c_uint64), C.c_uint64, C.c_int64, P(C.c_uint64)]),
        "ora_sample_inclusive": (C.c_int, [P(C.c_uint64), C.c_int64, C.c_uint64, C.c_uint64,
                                           P(C.c_uint64), P(C.c_int64)]),
        "ora_efb_sample_rows": (C.c_int64, [C.c_int64, C.c_int64, C.c_uint64, P(C.c_int64),
                                            C.c_int64]),
        "ora_bundles": (C.c_int32, [vp, P(C.c_int32), P(C.c_uint32), C.c_int64]),
        "ora_sample_rows": (C.c_int, [P(C.c_float), P(C.c_float), C.c_int64, P(A.GBDTConfigC),
                                      P(A.ExtConfig), C.c_uint32, C.c_uint32, P(C.c_float),
                                      P(C.c_float), P(C.c_uint32), P(C.c_int64)]),
        "ora_sample_cols": (C.c_int, [C.c_uint32, P(A.GBDTConfigC), C.c_uint32, C.c_uint32,
                                      P(C.c_uint32), P(C.c_int32)]),
        "ora_gradients_cfg": (C.c_int, [P(A.GBDTConfigC), P(C.c_float), P(C.c_float),
                                        P(C.c_float), C.c_int64, P(C.c_float), P(C.c_float)]),
        "ora_base_score_cfg": (C.c_int, [P(A.GBDTConfigC), P(C.c_float), P(C.c_float),
                                         C.c_int64, P(C.c_float)]),
        "ora_weighted_quantile": (C.c_float, [P(C.c_float), P(C.c_float), C.c_int64,
                                              C.c_float]),
        "ora_quantile_indexed": (C.c_float, [P(C.c_uint32), C.c_uint32, P(C.c_float),
                                             P(C.c_float), C.c_float]),
        "ora_metric_cfg": (C.c_double, [P(A.GBDTConfigC), C.c_int32, P(C.c_float), C.c_int32,
                                        P(C.c_float), P(C.c_float), C.c_int64]),
        "ora_default_metric": (C.c_int, [C.c_int32]),
        "ora_sample_level": (C.c_int, [P(C.c_uint32), C.c_uint32, P(A.GBDTConfigC), C.c_uint32,
                                       C.c_uint32, C.c_uint32, P(C.c_uint32), P(C.c_int32)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    return L


def _check(code):
    if code != 0:
        raise A.BoostersError(code, lib().ora_last_error().decode())


def set_exp_mode(mode: int):
    """0: glibc expf as the reference; 1: f64 exp rounded to f32 as the engine."""
    lib().ora_set_exp_mode(int(mode))


def set_threads(n: int):
    lib().ora_set_num_threads(int(n))


class Binned:
    def __init__(self, X, is_cat=None, binning=None):
        X = np.ascontiguousarray(X, dtype=np.float32)
        self.n_feat, self.n_rows = X.shape
        self.X = X
        h = C.c_void_p()
        cat = None if is_cat is None else np.ascontiguousarray(is_cat, dtype=np.uint8)
        _check(lib().ora_bin(A.ptr(X), self.n_rows, self.n_feat, A.ptr(cat),
                             C.byref(binning) if binning is not None else None, C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_binned_free(self.h)
            self.h = None

    def bundles(self):
        """The exclusive feature bundles (original feature indices, bundle order)."""
        n = lib().ora_bundles(self.h, None, None, 0)
        sizes = np.zeros(max(n, 1), np.int32)
        feats = np.zeros(self.n_feat, np.uint32)
        lib().ora_bundles(self.h, A.ptr(sizes), A.ptr(feats), len(feats))
        out, k = [], 0
        for s in sizes[:n]:
            out.append(feats[k:k + s].tolist())
            k += s
        return out

    def feature_info(self, f):
        nb, mt, db, cat, tr = C.c_uint32(), C.c_int32(), C.c_uint32(), C.c_int32(), C.c_int32()
        _check(lib().ora_feature_info(self.h, f, C.byref(nb), C.byref(mt), C.byref(db),
                                      C.byref(cat), C.byref(tr)))
        return dict(n_bins=nb.value, missing=mt.value, default_bin=db.value,
                    categorical=bool(cat.value), trivial=bool(tr.value))

    def bounds(self, f):
        out = np.zeros(512, dtype=np.float64)
        k = lib().ora_bin_bounds(self.h, f, A.ptr(out), 512)
        return out[:k].copy()

    def value_to_bin(self, f, v):
        return lib().ora_value_to_bin(self.h, f, float(v))

    def bins(self):
        out = np.zeros((self.n_feat, self.n_rows), dtype=np.uint16)
        _check(lib().ora_get_bins(self.h, A.ptr(out)))
        return out

    def layout(self):
        ne = C.c_int32()
        e2o = np.zeros(self.n_feat, np.uint32)
        nb = np.zeros(self.n_feat, np.uint32)
        off = np.zeros(self.n_feat, np.uint32)
        total = lib().ora_hist_layout(self.h, C.byref(ne), A.ptr(e2o), A.ptr(nb), A.ptr(off))
        e = ne.value
        return dict(n_eff=e, eff_to_orig=e2o[:e].copy(), n_bins=nb[:e].copy(),
                    offset=off[:e].copy(), total_bins=total)

    def histogram(self, grad, hess, rows=None):
        lay = self.layout()
        out = np.zeros(lay["total_bins"] * 2, dtype=np.float64)
        grad = np.ascontiguousarray(grad, np.float32)
        hess = np.ascontiguousarray(hess, np.float32)
        r = None if rows is None else np.ascontiguousarray(rows, np.uint32)
        _check(lib().ora_histogram(self.h, A.ptr(grad), A.ptr(hess), A.ptr(r),
                                   0 if r is None else len(r), A.ptr(out)))
        return out.reshape(-1, 2)

    def find_split(self, hist_gh, pg, ph, pc, cfg, n_threads=0, features=None):
        out = A.SplitInfoC()
        h = np.ascontiguousarray(hist_gh, np.float64).reshape(-1)
        fs = None if features is None else np.ascontiguousarray(features, np.uint32)
        _check(lib().ora_find_split(self.h, A.ptr(h), pg, ph, pc, C.byref(cfg), n_threads,
                                    A.ptr(fs), 0 if fs is None else len(fs), C.byref(out)))
        return out

    def training_predictions(self, info, arrays, by_threshold=None):
        """The reference trainer's training predictions of a forest over these
        rows: partition routing on the bins, or threshold traversal of the raw
        features for trees flagged in `by_threshold`."""
        out = np.zeros((info.n_groups, self.n_rows), np.float32)
        bt = None if by_threshold is None else np.ascontiguousarray(by_threshold, np.uint8)
        fa = A.forest_struct(arrays)
        _check(lib().ora_partition_predict(self.h, C.byref(info), C.byref(fa), A.ptr(bt),
                                           A.ptr(self.X), A.ptr(out)))
        return out

    def partition(self, rows, split):
        rows = np.ascontiguousarray(rows, np.uint32).copy()
        lc = C.c_int64()
        _check(lib().ora_partition(self.h, A.ptr(rows), len(rows), C.byref(split), C.byref(lc)))
        return rows, lc.value


def gradients(objective, preds, targets, weights=None):
    preds = np.ascontiguousarray(preds, np.float32)
    K, n = preds.shape
    g = np.zeros_like(preds)
    h = np.zeros_like(preds)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    _check(lib().ora_gradients(objective, K, A.ptr(preds),
                               A.ptr(np.ascontiguousarray(targets, np.float32)), A.ptr(w), n,
                               A.ptr(g), A.ptr(h)))
    return g, h


def gradients_cfg(cfg, preds, targets, weights=None):
    """Gradients with the objective's payload (pinball alphas, Huber delta)."""
    preds = np.ascontiguousarray(preds, np.float32)
    g = np.zeros_like(preds)
    h = np.zeros_like(preds)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    _check(lib().ora_gradients_cfg(C.byref(cfg), A.ptr(preds),
                                   A.ptr(np.ascontiguousarray(targets, np.float32)), A.ptr(w),
                                   preds.shape[-1], A.ptr(g), A.ptr(h)))
    return g, h


def base_score_cfg(cfg, targets, n_outputs=1, weights=None):
    out = np.zeros(n_outputs, np.float32)
    t = np.ascontiguousarray(targets, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    _check(lib().ora_base_score_cfg(C.byref(cfg), A.ptr(t), A.ptr(w), len(t), A.ptr(out)))
    return out


def weighted_quantile(values, weights, alpha):
    v = np.ascontiguousarray(values, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    return lib().ora_weighted_quantile(A.ptr(v), A.ptr(w), len(v), float(alpha))


def quantile_indexed(indices, values, alpha, weights=None):
    ind = np.ascontiguousarray(indices, np.uint32)
    v = np.ascontiguousarray(values, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    return lib().ora_quantile_indexed(A.ptr(ind), len(ind), A.ptr(v), A.ptr(w), float(alpha))


def metric_cfg(cfg, metric_id, preds, targets, weights=None):
    preds = np.ascontiguousarray(preds, np.float32)
    K, n = preds.shape
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    return lib().ora_metric_cfg(C.byref(cfg), metric_id, A.ptr(preds), K,
                                A.ptr(np.ascontiguousarray(targets, np.float32)), A.ptr(w), n)


def sample_level(tree_feats, cfg, round_, output, depth):
    t = np.ascontiguousarray(tree_feats, np.uint32)
    out = np.zeros(max(len(t), 1), np.uint32)
    k = C.c_int32()
    _check(lib().ora_sample_level(A.ptr(t), len(t), C.byref(cfg), round_, output, depth, A.ptr(out),
                                  C.byref(k)))
    return out[: k.value]


def default_metric(objective):
    return lib().ora_default_metric(objective)


def lambdarank_gradients(preds, targets, group_sizes, weights=None):
    preds = np.ascontiguousarray(preds, np.float32).reshape(-1)
    n = len(preds)
    g = np.zeros(n, np.float32)
    h = np.zeros(n, np.float32)
    sizes = np.ascontiguousarray(group_sizes, np.int32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    _check(lib().ora_lambdarank_gradients(A.ptr(preds), A.ptr(np.ascontiguousarray(targets, np.float32)),
                                          A.ptr(w), n, A.ptr(sizes), len(sizes), A.ptr(g), A.ptr(h)))
    return g, h


def base_score(objective, targets, n_outputs=1, weights=None):
    out = np.zeros(n_outputs, np.float32)
    t = np.ascontiguousarray(targets, np.float32)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    _check(lib().ora_base_score(objective, n_outputs, A.ptr(t), A.ptr(w), len(t), A.ptr(out)))
    return out


class Model:
    """An oracle-trained forest in exchange format."""

    def __init__(self, handle):
        self.h = handle
        info = A.ModelInfo()
        _check(lib().ora_model_info(self.h, C.byref(info)))
        self.info = info
        self.arrays = A.alloc_forest(info)
        fa = A.forest_struct(self.arrays)
        _check(lib().ora_model_export(self.h, C.byref(fa)))

    def __del__(self):
        if getattr(self, "h", None):
            lib().ora_model_free(self.h)
            self.h = None

    def metric_history(self):
        out = np.zeros(100000, np.float64)
        nr = C.c_int64()
        _check(lib().ora_model_metric(self.h, A.ptr(out), len(out), C.byref(nr)))
        return out[: nr.value].copy()

    def stats(self):
        hr, ms = C.c_int64(), C.c_double()
        _check(lib().ora_model_stats(self.h, C.byref(hr), C.byref(ms)))
        return hr.value, ms.value


def train(binned: Binned, targets, cfg, ext=None, weights=None, valid_X=None, valid_y=None):
    t = np.ascontiguousarray(targets, np.float32).reshape(-1)
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    vX = None if valid_X is None else np.ascontiguousarray(valid_X, np.float32)
    vy = None if valid_y is None else np.ascontiguousarray(valid_y, np.float32)
    h = C.c_void_p()
    _check(lib().ora_train(binned.h, A.ptr(binned.X), A.ptr(t), len(t), A.ptr(w), A.ptr(vX),
                           A.ptr(vy), 0 if vy is None else len(vy), C.byref(cfg),
                           C.byref(ext) if ext is not None else None, C.byref(h)))
    return Model(h)


def predict(info, arrays, X, raw=False, leaf_idx=False):
    X = np.ascontiguousarray(X, np.float32)
    F, n = X.shape
    out = np.zeros((info.n_groups, n), np.float32)
    li = np.zeros((info.n_trees, n), np.uint32) if leaf_idx else None
    fa = A.forest_struct(arrays)
    _check(lib().ora_predict(C.byref(info), C.byref(fa), A.ptr(X), n, F, int(raw), A.ptr(out),
                             A.ptr(li)))
    return (out, li) if leaf_idx else out


def leaf_weight(cfg, g, h):
    return lib().ora_leaf_weight(C.byref(cfg), float(g), float(h))


def metric(metric_id, preds, targets, weights=None):
    preds = np.ascontiguousarray(preds, np.float32)
    K, n = preds.shape
    w = None if weights is None else np.ascontiguousarray(weights, np.float32)
    return lib().ora_metric(metric_id, A.ptr(preds), K,
                            A.ptr(np.ascontiguousarray(targets, np.float32)), A.ptr(w), n)


def splitmix64_stream(state, count):
    out = np.zeros(count, np.uint64)
    lib().ora_splitmix64_stream(C.c_uint64(state), count, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def xoshiro256pp(count, seed=0, state=None):
    """xoshiro256++ outputs from `state` (4 words) or from seed_from_u64(seed)."""
    out = np.zeros(count, np.uint64)
    st = None if state is None else np.ascontiguousarray(state, np.uint64)
    lib().ora_xoshiro256pp(None if st is None else st.ctypes.data_as(C.POINTER(C.c_uint64)),
                           C.c_uint64(seed), count, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    return out


def sample_inclusive(draws, low, high):
    """rand's gen_range(low..=high) over the scripted u64 draws: (value, draws used)."""
    d = np.ascontiguousarray(draws, np.uint64)
    v, used = C.c_uint64(), C.c_int64()
    _check(lib().ora_sample_inclusive(d.ctypes.data_as(C.POINTER(C.c_uint64)), len(d),
                                      C.c_uint64(low), C.c_uint64(high), C.byref(v), C.byref(used)))
    return v.value, used.value


def efb_sample_rows(n, max_samples, seed=42):
    out = np.zeros(max(min(int(n), int(max_samples)), 1), np.int64)
    k = lib().ora_efb_sample_rows(int(n), int(max_samples), C.c_uint64(seed),
                                  out.ctypes.data_as(C.POINTER(C.c_int64)), len(out))
    return out[:k].copy()


def binning_sample(n, cnt):
    out = np.zeros(max(min(int(n), int(cnt)), 1), np.uint32)
    k = lib().ora_binning_sample(int(n), int(cnt), A.ptr(out))
    return out[:k].copy()


def sample_rows(grad, hess, cfg, ext=None, round_=0, output=0):
    """The row sample of tree (round, output): (rows or None, grad, hess)."""
    g = np.ascontiguousarray(grad, np.float32)
    h = np.ascontiguousarray(hess, np.float32)
    n = len(g)
    go = np.zeros(n, np.float32)
    ho = np.zeros(n, np.float32)
    rows = np.zeros(n, np.uint32)
    nr = C.c_int64()
    _check(lib().ora_sample_rows(A.ptr(g), A.ptr(h), n, C.byref(cfg),
                                 C.byref(ext) if ext is not None else None, round_, output,
                                 A.ptr(go), A.ptr(ho), A.ptr(rows), C.byref(nr)))
    return (None if nr.value < 0 else rows[: nr.value].copy()), go, ho


def sample_cols(n_feat, cfg, round_=0, output=0):
    out = np.zeros(n_feat, np.uint32)
    k = C.c_int32()
    _check(lib().ora_sample_cols(n_feat, C.byref(cfg), round_, output, A.ptr(out), C.byref(k)))
    return out[: k.value].copy()
