"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  The oracle is a plain C++ restatement of the reference's
CPU algorithms (each function in oracle.cpp cites the reference file:line it follows).
It is the checker for the HIP path: only the tests, the smoke check and the benchmark's
CPU-baseline leg load it.  Nothing in the bonsai_amd package imports it, and the
product path never falls back to it.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
LIB = ORACLE_DIR / "liboracle.so"

NODE_DTYPE = np.dtype([
    ("feature", np.int32), ("bin", np.uint32), ("default_left", np.uint32),
    ("left", np.uint32), ("right", np.uint32), ("value", np.float32),
    ("cover", np.float32), ("gain", np.float32),
])

HIST_F32_SERIAL, HIST_F32_CHUNKED, HIST_INT, HIST_F32_OMP, HIST_F32_FIXED = 0, 1, 2, 3, 4
# Rows per device histogram work unit: the FP32 chunked order the device sums in.
DEVICE_CHUNK = 8192
CHUNK_ADAPTIVE = 0x80000000  # | unit target T: leaf-sized FP32 partials (oracle.h)


class TreeCfg(C.Structure):
    _fields_ = [("min_child_hess", C.c_float), ("min_gain_to_split", C.c_float),
                ("lambda_l2", C.c_float), ("lambda_l1", C.c_float), ("max_depth", C.c_int32),
                ("min_data_in_leaf", C.c_int32), ("max_leaves", C.c_int32), ("_pad", C.c_int32)]


class Split(C.Structure):
    _fields_ = [("gain", C.c_double), ("feature", C.c_uint32), ("bin", C.c_uint32),
                ("default_left", C.c_uint32), ("valid", C.c_uint32)]


def cfg(min_child_hess=1.0, min_gain=0.0, l2=1.0, l1=0.0, max_depth=6, min_data=20,
        max_leaves=31) -> TreeCfg:
    return TreeCfg(min_child_hess, min_gain, l2, l1, max_depth, min_data, max_leaves, 0)


def P(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class Oracle:
    def __init__(self, lib: C.CDLL):
        self.L = lib
        for name in dir(lib):
            pass
        L = lib
        L.orc_fit_cuts.restype = C.c_int32
        L.orc_init_score.restype = C.c_float
        L.orc_logloss_eval.restype = C.c_float
        L.orc_leaf_value.restype = C.c_float
        L.orc_bernoulli_rows.restype = C.c_uint64
        L.orc_leaf_value.argtypes = [C.c_double, C.c_double, C.c_void_p]
        L.orc_partition.restype = C.c_uint64
        L.orc_grow_depthwise.restype = C.c_int32
        L.orc_grow_leafwise.restype = C.c_int32
        L.orc_grow_levelwise.restype = C.c_int32
        L.orc_train_depthwise.restype = C.c_int32
        L.orc_tree_expected_value.restype = C.c_double
        L.orc_cond_expectation.restype = C.c_double
        L.orc_pack_paths.restype = C.c_int32
        L.orc_auc.restype = C.c_float
        L.orc_auc.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64]

    # -- binning --
    def fit_cuts(self, col, max_bin):
        col = np.ascontiguousarray(col, np.float32)
        out = np.zeros(max_bin + 4, np.float32)
        n = self.L.orc_fit_cuts(P(col), C.c_uint64(col.size), C.c_uint64(1), C.c_int32(max_bin), P(out),
                                C.c_int32(out.size))
        assert n > 0
        return out[:n].copy()

    def transform(self, cuts, x):
        cuts = np.ascontiguousarray(cuts, np.float32)
        x = np.ascontiguousarray(x, np.float32)
        out = np.zeros(x.size, np.uint8)
        self.L.orc_transform(P(cuts), C.c_int32(cuts.size), P(x), C.c_uint64(x.size), C.c_uint64(1), P(out),
                             C.c_uint64(1))
        return out

    # -- objective --
    def logloss_grad(self, s, y):
        s = np.ascontiguousarray(s, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros_like(s)
        h = np.zeros_like(s)
        self.L.orc_logloss_grad(P(s), P(y), C.c_uint64(s.size), P(g), P(h))
        return g, h

    def mse_grad(self, s, y):
        s = np.ascontiguousarray(s, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros_like(s)
        h = np.zeros_like(s)
        self.L.orc_mse_grad(P(s), P(y), C.c_uint64(s.size), P(g), P(h))
        return g, h

    def poisson_grad(self, s, y):
        s = np.ascontiguousarray(s, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        g = np.zeros_like(s)
        h = np.zeros_like(s)
        self.L.orc_poisson_grad(P(s), P(y), C.c_uint64(s.size), P(g), P(h))
        return g, h

    def init_score(self, objective, y):
        y = np.ascontiguousarray(y, np.float32)
        return self.L.orc_init_score(C.c_int32(objective), P(y), C.c_uint64(y.size))

    def logloss_eval(self, s, y):
        s = np.ascontiguousarray(s, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        return self.L.orc_logloss_eval(P(s), P(y), C.c_uint64(s.size))

    def quantize(self, g, h, seed, it):
        g = np.ascontiguousarray(g, np.float32)
        h = np.ascontiguousarray(h, np.float32)
        gq = np.zeros(g.size, np.int8)
        hq = np.zeros(g.size, np.int16)
        mg = C.c_float(0)
        mh = C.c_float(0)
        self.L.orc_quantize(P(g), P(h), C.c_uint64(g.size), C.c_uint64(seed), C.c_uint32(it), P(gq), P(hq),
                            C.byref(mg), C.byref(mh))
        return gq, hq, np.array([np.float64(mg.value) / 127.0, np.float64(mh.value) / 32767.0])

    # -- histograms --
    def bernoulli_rows(self, n, subsample, seed, it):
        """Rows kept by the Bernoulli sampler at iteration `it` (ascending)."""
        if subsample >= 1.0:
            return np.arange(n, dtype=np.uint32)
        thr = min(int(np.floor(np.float64(np.float32(subsample)) * 4294967296.0)), 4294967295)
        out = np.zeros(max(n, 1), np.uint32)
        k = self.L.orc_bernoulli_rows(C.c_uint64(n), C.c_uint32(thr), C.c_uint64(seed & ((1 << 64) - 1)),
                                      C.c_uint32(it), P(out))
        return out[:k].copy()

    def set_feature_mask(self, mask):
        """Restrict the following finders and growers to mask[f] != 0 (None: all)."""
        if mask is None:
            self.L.orc_set_feature_mask(None, C.c_uint32(0))
            return
        m = np.ascontiguousarray(mask, np.uint8)
        self.L.orc_set_feature_mask(P(m), C.c_uint32(m.size))

    def set_monotone(self, dirs):
        """Monotone directions for the following depthwise/leafwise growers (None: none)."""
        if dirs is None:
            self.L.orc_set_monotone(None, C.c_uint32(0))
            return
        d = np.ascontiguousarray(dirs, np.int8)
        self.L.orc_set_monotone(P(d), C.c_uint32(d.size))

    def set_interaction(self, groups):
        """Interaction groups (lists of feature ids) for the following depthwise/leafwise
        growers (None or []: none)."""
        groups = [list(g) for g in (groups or []) if len(g)]
        if not groups:
            self.L.orc_set_interaction(None, None, C.c_uint32(0))
            return
        m = np.ascontiguousarray(np.concatenate([np.asarray(g, np.uint32) for g in groups]), np.uint32)
        s = np.array([len(g) for g in groups], np.uint32)
        self.L.orc_set_interaction(P(m), P(s), C.c_uint32(s.size))

    def feature_draw(self, n_features, fraction, seed, it):
        """The device's per-iteration feature subsample as a 0/1 mask."""
        out = np.zeros(n_features, np.uint8)
        self.L.orc_feature_draw(C.c_uint32(n_features), C.c_float(fraction), C.c_uint64(seed & ((1 << 64) - 1)),
                                C.c_uint32(it), P(out))
        return out

    def hist_f64(self, bins, rows, g, h, bin_stride):
        n, f = bins.shape
        out = np.zeros((f, bin_stride, 2), np.float64)
        rows = np.ascontiguousarray(rows, np.uint32)
        self.L.orc_hist_f64(P(bins), C.c_uint64(f), C.c_uint32(f), C.c_uint32(bin_stride), P(rows),
                            C.c_uint64(rows.size), P(np.ascontiguousarray(g, np.float32)),
                            P(np.ascontiguousarray(h, np.float32)), P(out))
        return out

    def fx_shift(self, v) -> int:
        """Fixed-point grid exponent of HIST_F32_FIXED for the values v (orc_fx_shift)."""
        v = np.ascontiguousarray(v, np.float32)
        return int(self.L.orc_fx_shift(P(v), C.c_uint64(v.size)))

    def hist_f32(self, bins, rows, g, h, bin_stride, mode=HIST_F32_SERIAL, chunk=DEVICE_CHUNK):
        n, f = bins.shape
        out = np.zeros((f, bin_stride, 2), np.float32)
        rows = np.ascontiguousarray(rows, np.uint32)
        if mode == HIST_F32_FIXED:
            # grids from every row's gradient, as the device takes them
            sg, sh = self.fx_shift(g), self.fx_shift(h)
            self.L.orc_hist_f32_fixed(P(bins), C.c_uint64(f), C.c_uint32(f), C.c_uint32(bin_stride), P(rows),
                                      C.c_uint64(rows.size), P(np.ascontiguousarray(g, np.float32)),
                                      P(np.ascontiguousarray(h, np.float32)), C.c_int32(sg), C.c_int32(sh),
                                      C.c_uint32(chunk), P(out))
            return out
        self.L.orc_hist_f32(P(bins), C.c_uint64(f), C.c_uint32(f), C.c_uint32(bin_stride), P(rows),
                            C.c_uint64(rows.size), P(np.ascontiguousarray(g, np.float32)),
                            P(np.ascontiguousarray(h, np.float32)), C.c_int32(mode), C.c_uint32(chunk), P(out))
        return out

    def hist_int(self, bins, rows, gq, hq, bin_stride):
        n, f = bins.shape
        out = np.zeros((f, bin_stride, 2), np.int64)
        rows = np.ascontiguousarray(rows, np.uint32)
        self.L.orc_hist_int(P(bins), C.c_uint64(f), C.c_uint32(f), C.c_uint32(bin_stride), P(rows),
                            C.c_uint64(rows.size), P(np.ascontiguousarray(gq, np.int8)),
                            P(np.ascontiguousarray(hq, np.int16)), P(out))
        return out

    # -- split find --
    def find_node_f32(self, hist, n_bins, tot_g, tot_h, n_rows, c):
        hist = np.ascontiguousarray(hist, np.float32)
        nb = np.ascontiguousarray(n_bins, np.uint16)
        out = Split()
        self.L.orc_find_node_f32(P(hist), P(nb), C.c_uint32(hist.shape[0]), C.c_uint32(hist.shape[1]),
                                 C.c_float(tot_g), C.c_float(tot_h), C.c_uint64(n_rows), C.byref(c),
                                 C.byref(out))
        return out

    def find_node_int(self, hist, n_bins, tot_g, tot_h, sg, sh, n_rows, c):
        hist = np.ascontiguousarray(hist, np.int64)
        nb = np.ascontiguousarray(n_bins, np.uint16)
        out = Split()
        self.L.orc_find_node_int(P(hist), P(nb), C.c_uint32(hist.shape[0]), C.c_uint32(hist.shape[1]),
                                 C.c_int64(int(tot_g)), C.c_int64(int(tot_h)), C.c_double(sg), C.c_double(sh),
                                 C.c_uint64(n_rows), C.byref(c), C.byref(out))
        return out

    def find_level_f32(self, hists, totals, n_bins, c):
        hists = np.ascontiguousarray(hists, np.float32)
        totals = np.ascontiguousarray(totals, np.float32)
        nb = np.ascontiguousarray(n_bins, np.uint16)
        out = Split()
        self.L.orc_find_level_f32(P(hists), P(totals), C.c_uint32(hists.shape[0]), P(nb),
                                  C.c_uint32(hists.shape[1]), C.c_uint32(hists.shape[2]), C.byref(c),
                                  C.byref(out))
        return out

    def find_level_int(self, hists, totals, n_bins, sg, sh, c):
        hists = np.ascontiguousarray(hists, np.int64)
        totals = np.ascontiguousarray(totals, np.int64)
        nb = np.ascontiguousarray(n_bins, np.uint16)
        out = Split()
        self.L.orc_find_level_int(P(hists), P(totals), C.c_uint32(hists.shape[0]), P(nb),
                                  C.c_uint32(hists.shape[1]), C.c_uint32(hists.shape[2]), C.c_double(sg),
                                  C.c_double(sh), C.byref(c), C.byref(out))
        return out

    def totals_f32(self, hist_f0, nb0):
        tg, th = C.c_float(0), C.c_float(0)
        self.L.orc_totals_f32(P(np.ascontiguousarray(hist_f0, np.float32)), C.c_uint32(nb0), C.byref(tg),
                              C.byref(th))
        return tg.value, th.value

    def leaf_value(self, g, h, c):
        return self.L.orc_leaf_value(g, h, C.byref(c))

    # -- partition --
    def partition(self, bins, rows, feature, bin_, dl, last_bin):
        rows = np.ascontiguousarray(rows, np.uint32)
        out = np.zeros(rows.size, np.uint32)
        nl = self.L.orc_partition(P(bins), C.c_uint64(bins.shape[1]), P(rows), C.c_uint64(rows.size),
                                  C.c_uint32(feature), C.c_uint32(bin_), C.c_uint32(int(dl)),
                                  C.c_uint32(last_bin), P(out))
        return out, nl

    # -- growers --
    def _grow(self, fn, bins, n_bins, g, h, gq, hq, scales, c, mode, chunk=DEVICE_CHUNK, cap=1 << 14):
        bins = np.ascontiguousarray(bins, np.uint8)
        nb = np.ascontiguousarray(n_bins, np.uint16)
        nodes = np.zeros(cap, NODE_DTYPE)
        lor = np.zeros(bins.shape[0], np.uint32)
        sg, sh = (0.0, 0.0) if scales is None else scales
        n = fn(P(bins), C.c_uint64(bins.shape[1]), C.c_uint64(bins.shape[0]), C.c_uint32(bins.shape[1]), P(nb),
               P(g), P(h), P(gq), P(hq), C.c_double(sg), C.c_double(sh), C.byref(c), C.c_int32(mode),
               C.c_uint32(chunk), P(nodes), C.c_int32(cap), P(lor))
        assert n > 0
        return nodes[:n].copy(), lor

    def grow_depthwise(self, bins, n_bins, g=None, h=None, gq=None, hq=None, scales=None, c=None,
                       mode=HIST_F32_SERIAL, chunk=DEVICE_CHUNK):
        return self._grow(self.L.orc_grow_depthwise, bins, n_bins, _f(g), _f(h), _i8(gq), _i16(hq), scales, c,
                          mode, chunk)

    def grow_leafwise(self, bins, n_bins, g=None, h=None, gq=None, hq=None, scales=None, c=None,
                      mode=HIST_F32_SERIAL, chunk=DEVICE_CHUNK):
        return self._grow(self.L.orc_grow_leafwise, bins, n_bins, _f(g), _f(h), _i8(gq), _i16(hq), scales, c,
                          mode, chunk)

    def grow_levelwise(self, bins, n_bins, g=None, h=None, gq=None, hq=None, scales=None, c=None,
                       mode=HIST_F32_SERIAL, chunk=DEVICE_CHUNK):
        bins = np.ascontiguousarray(bins, np.uint8)
        nb = np.ascontiguousarray(n_bins, np.uint16)
        D = c.max_depth
        levels = np.zeros((D, 3), np.uint32)
        leaves = np.zeros(1 << D, np.float32)
        covers = np.zeros(1 << D, np.float32)
        gains = np.zeros(D, np.float32)
        lor = np.zeros(bins.shape[0], np.uint32)
        sg, sh = (0.0, 0.0) if scales is None else scales
        d = self.L.orc_grow_levelwise(P(bins), C.c_uint64(bins.shape[1]), C.c_uint64(bins.shape[0]),
                                      C.c_uint32(bins.shape[1]), P(nb), P(_f(g)), P(_f(h)), P(_i8(gq)),
                                      P(_i16(hq)), C.c_double(sg), C.c_double(sh), C.byref(c), C.c_int32(mode),
                                      C.c_uint32(chunk), P(levels), P(leaves), P(covers), P(gains), P(lor))
        assert d >= 0
        return levels[:d].copy(), leaves[: 1 << d].copy(), covers[: 1 << d].copy(), gains[:d].copy(), lor

    def train_depthwise(self, bins, n_bins, labels, objective, n_iters, lr, c, mode=HIST_F32_OMP,
                        n_threads=0, max_nodes=1024):
        bins = np.ascontiguousarray(bins, np.uint8)
        nb = np.ascontiguousarray(n_bins, np.uint16)
        y = np.ascontiguousarray(labels, np.float32)
        trees = np.zeros((n_iters, max_nodes), NODE_DTYPE)
        sizes = np.zeros(n_iters, np.int32)
        init = C.c_float(0)
        scores = np.zeros(bins.shape[0], np.float32)
        r = self.L.orc_train_depthwise(P(bins), C.c_uint64(bins.shape[1]), C.c_uint64(bins.shape[0]),
                                       C.c_uint32(bins.shape[1]), P(nb), P(y), C.c_int32(objective),
                                       C.c_uint32(n_iters), C.c_float(lr), C.byref(c), C.c_int32(mode),
                                       C.c_int32(n_threads), P(trees), C.c_int32(max_nodes), P(sizes),
                                       C.byref(init), P(scores))
        assert r == n_iters
        return [trees[t, : sizes[t]].copy() for t in range(n_iters)], init.value, scores

    # -- predict / shap --
    def predict(self, trees, bins, n_bins, lr, init):
        offs = np.cumsum([0] + [len(t) for t in trees[:-1]]).astype(np.int32)
        nodes = np.ascontiguousarray(np.concatenate(trees), NODE_DTYPE)
        last = (np.asarray(n_bins, np.uint16) - 1).astype(np.uint16)
        bins = np.ascontiguousarray(bins, np.uint8)
        out = np.zeros(bins.shape[0], np.float32)
        self.L.orc_predict(P(nodes), P(offs), C.c_uint32(len(trees)), P(bins), C.c_uint64(bins.shape[1]),
                           C.c_uint64(bins.shape[0]), P(last), C.c_float(lr), C.c_float(init), P(out))
        return out

    def pred_contribs(self, trees, bins, n_bins, lr, init, max_depth):
        offs = np.cumsum([0] + [len(t) for t in trees[:-1]]).astype(np.int32)
        nodes = np.ascontiguousarray(np.concatenate(trees), NODE_DTYPE)
        last = (np.asarray(n_bins, np.uint16) - 1).astype(np.uint16)
        bins = np.ascontiguousarray(bins, np.uint8)
        F = bins.shape[1]
        out = np.zeros((bins.shape[0], F + 1), np.float64)
        self.L.orc_pred_contribs(P(nodes), P(offs), C.c_uint32(len(trees)), C.c_int32(max_depth), P(bins),
                                 C.c_uint64(F), C.c_uint64(bins.shape[0]), P(last), C.c_uint32(F), C.c_float(lr),
                                 C.c_float(init), P(out))
        return out

    def auc(self, p, y):
        p = np.ascontiguousarray(p, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        return self.L.orc_auc(P(p), P(y), C.c_uint64(p.size))


def _f(a):
    return None if a is None else np.ascontiguousarray(a, np.float32)


def _i8(a):
    return None if a is None else np.ascontiguousarray(a, np.int8)


def _i16(a):
    return None if a is None else np.ascontiguousarray(a, np.int16)


_OR = None


def load_oracle() -> Oracle:
    global _OR
    if _OR is None:
        if not LIB.exists() or LIB.stat().st_mtime < (ORACLE_DIR / "oracle.cpp").stat().st_mtime:
            subprocess.run(["make", "-C", str(ORACLE_DIR)], check=True, capture_output=True)
        _OR = Oracle(C.CDLL(str(LIB)))
    return _OR
