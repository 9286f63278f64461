"""bonsai_amd — MI355X-native GBDT hot path (histograms, split find, partition, growers,
predict, TreeSHAP) behind the C ABI of include/bonsai_hip.h.

This module is the Python host binding of libbonsai_hip.so (ctypes).  It mirrors the
reference's booster/engine vocabulary: a ``HipContext`` holds one resident training set
(the reference's CudaDeviceContext), ``FitConfig`` carries TreeConfig/BoosterConfig
fields plus the device knobs, growers are named ``hip_depthwise`` / ``hip_leafwise`` /
``hip_levelwise`` (the reference registers ``cuda_*`` under
include/bonsai/registry/names.hpp:81-94), and errors map to ``ConfigError`` (capacity
or configuration refusals) and ``DeviceError``.

The library must be built (``python -m bonsai_amd.build`` or
``__graft_entry__.build()``); importing the binding without it raises.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libbonsai_hip.so"


class ConfigError(ValueError):
    """A configuration or capacity refusal (the reference's bonsai::ConfigError)."""


class DeviceError(RuntimeError):
    """A HIP runtime failure."""


class LibraryMissing(ImportError):
    """libbonsai_hip.so has not been built: there is no CPU fallback."""


GROWERS = {"hip_depthwise": 0, "hip_leafwise": 1, "hip_levelwise": 2}
OBJECTIVES = {"mse": 0, "logloss": 1, "poisson": 2}


class FitConfigC(C.Structure):
    _fields_ = [
        ("grower", C.c_int32), ("objective", C.c_int32), ("quantized", C.c_int32),
        ("use_graph", C.c_int32), ("host_sync", C.c_int32), ("max_depth", C.c_int32),
        ("min_data_in_leaf", C.c_int32), ("max_leaves", C.c_int32),
        ("min_child_hess", C.c_float), ("min_gain_to_split", C.c_float),
        ("lambda_l2", C.c_float), ("lambda_l1", C.c_float), ("learning_rate", C.c_float),
        ("_pad", C.c_uint32), ("seed", C.c_uint64),
    ]


class TuningC(C.Structure):
    _fields_ = [("scatter_lds_bytes", C.c_uint32), ("find_blocks", C.c_uint32), ("hist_threads", C.c_uint32),
                ("part_grid", C.c_uint32), ("hist_layout", C.c_uint32), ("pair_levels", C.c_uint32),
                ("hist_group", C.c_uint32)]


class StepNodeC(C.Structure):
    _fields_ = [("depth", C.c_uint32), ("_pad", C.c_uint32), ("rows", C.c_uint64), ("sum_g", C.c_double),
                ("sum_h", C.c_double), ("isum_g", C.c_int64), ("isum_h", C.c_int64)]


class LeafRoundC(C.Structure):
    _fields_ = [("left_slot", C.c_uint32), ("right_slot", C.c_uint32), ("left_count", C.c_uint32),
                ("right_count", C.c_uint32)]


class ServeConfigC(C.Structure):
    _fields_ = [("n_slots", C.c_uint32), ("do_shap", C.c_uint32), ("max_batch_rows", C.c_uint64),
                ("hbm_budget", C.c_uint64)]


class ServeStatsC(C.Structure):
    _fields_ = [("wall_s", C.c_double), ("rows_per_s", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("mean_ms", C.c_double), ("rows", C.c_uint64),
                ("requests", C.c_uint32), ("max_inflight", C.c_uint32),
                ("max_device_concurrency", C.c_uint32), ("_pad", C.c_uint32), ("device_overlap_ms", C.c_double)]


NODE_DTYPE = np.dtype([
    ("feature", np.int32), ("bin", np.uint32), ("default_left", np.uint32),
    ("left", np.uint32), ("right", np.uint32), ("value", np.float32),
    ("cover", np.float32), ("gain", np.float32),
])
SPLIT_DTYPE = np.dtype([
    ("gain", np.float64), ("feature", np.uint32), ("bin", np.uint32),
    ("default_left", np.uint32), ("valid", np.uint32),
    ("gL", np.float64), ("hL", np.float64), ("gR", np.float64), ("hR", np.float64),
    ("igL", np.int64), ("ihL", np.int64), ("igR", np.int64), ("ihR", np.int64),
    ("n_left", np.uint32), ("n_right", np.uint32),
])
assert NODE_DTYPE.itemsize == 32 and SPLIT_DTYPE.itemsize == 96


@dataclass
class FitConfig:
    """TreeConfig / BoosterConfig fields the hot path reads, with the reference's
    defaults (include/bonsai/config/tree_config.hpp:9-29,
    include/bonsai/config/booster_config.hpp:8-24), plus device knobs."""
    grower: str = "hip_depthwise"
    objective: str = "logloss"
    quantized: bool = False
    use_graph: bool = True
    host_sync: bool = False
    max_depth: int = 6
    min_data_in_leaf: int = 20
    max_leaves: int = 31
    min_child_hess: float = 1.0
    min_gain_to_split: float = 0.0
    lambda_l2: float = 1.0
    lambda_l1: float = 0.0
    learning_rate: float = 0.05
    seed: int = 42
    # TreeConfig's sampling and constraint fields, applied by HipContext.configure
    feature_fraction: float = 1.0         # per-tree feature draw on the device; 1 = all
    feature_seed: int = 2
    monotone_constraints: tuple = ()      # +1 / -1 / 0 per feature, trailing entries free
    interaction_constraints: tuple = ()   # "0,1" / "0+1" strings or id lists
    # SamplerConfig.subsample (include/bonsai/config/sampler_config.hpp:8): the device
    # Bernoulli row draw per iteration, seeded by `seed`; 1 = every row
    subsample: float = 1.0

    def to_c(self) -> FitConfigC:
        if self.grower not in GROWERS:
            raise ConfigError(f"unknown grower_name {self.grower!r}; known: {sorted(GROWERS)}")
        if self.objective not in OBJECTIVES:
            raise ConfigError(f"objective {self.objective!r} has no device gradient")
        return FitConfigC(
            GROWERS[self.grower], OBJECTIVES[self.objective], int(self.quantized),
            int(self.use_graph), int(self.host_sync), self.max_depth, self.min_data_in_leaf,
            self.max_leaves, self.min_child_hess, self.min_gain_to_split, self.lambda_l2,
            self.lambda_l1, self.learning_rate, 0, self.seed)


_lib = None
_SIGS: dict = {}     # bound C-ABI functions -> ctypes argument types


def lib() -> C.CDLL:
    """Load libbonsai_hip.so once; raises LibraryMissing if it was never built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise LibraryMissing(f"{LIB_PATH} not built; run python -m bonsai_amd.build")
    L = C.CDLL(str(LIB_PATH))
    P, u32, u64, i32 = C.c_void_p, C.c_uint32, C.c_uint64, C.c_int32
    sig = {
        "bh_device_count": [C.POINTER(C.c_int)],
        "bh_synchronize": [P],
        "bh_copy_calibrate": [C.c_int, u64, u32, C.POINTER(C.c_float)],
        "bh_stream_calibrate": [C.c_int, u64, u32, u32, u32, C.POINTER(C.c_float)],
        "bh_create": [C.c_int, C.POINTER(P)],
        "bh_destroy": [P],
        "bh_upload_bins": [P, P, u64, u32, u64, P],
        "bh_upload_labels": [P, P],
        "bh_upload_weights": [P, P],
        "bh_upload_features": [P, P, u64, u32, P, P],
        "bh_upload_feature_columns": [P, P, u64, u32, P, P],
        "bh_upload_features_device": [P, P, u64, u32, P, P],
        "bh_gather_rows": [C.c_int, P, u64, u32, P, u64, P],
        "bh_download_bins": [P, P, u64],
        "bh_configure": [P, C.POINTER(FitConfigC)],
        "bh_set_scores": [P, P],
        "bh_get_scores": [P, P],
        "bh_train": [P, u32],
        "bh_set_validation": [P, P, u64, u64, P, C.c_float, u32],
        "bh_get_eval": [P, P, u32, C.POINTER(u32), C.POINTER(u32), C.POINTER(C.c_int32)],
        "bh_set_row_subset": [P, P, u64],
        "bh_set_bernoulli": [P, C.c_float, u64],
        "bh_set_feature_mask": [P, P, u32],
        "bh_set_feature_fraction": [P, C.c_float, u64],
        "bh_set_monotone": [P, P, u32],
        "bh_set_interaction": [P, P, P, u32],
        "bh_get_trees": [P, P, u32, P, u32, C.POINTER(u32)],
        "bh_max_nodes": [P, C.POINTER(u32)],
        "bh_set_gradients": [P, P, P, u32, P, P, P],
        "bh_get_gradients": [P, P, P, P, P, P],
        "bh_grow_tree": [P, P, u32, C.POINTER(u32), P],
        "bh_histogram": [P, P, u64, P, P],
        "bh_histogram_noalloc": [P, P, u64, P, P],
        "bh_find_split": [P, P, P, u64, u32, i32, C.c_double, C.c_double, P, P],
        "bh_partition": [P, P, u64, u32, u32, u32, P, C.POINTER(u64)],
        "bh_time_root_histogram": [P, u32, C.POINTER(C.c_float)],
        "bh_begin_root": [P, P, C.POINTER(u64)],
        "bh_partition_level": [P, P, u32, P],
        "bh_advance_level": [P, P, u32],
        "bh_advance_level_per_node": [P, P, u32],
        "bh_advance_layout_only": [P],
        "bh_find_splits_many": [P, P, u32, P],
        "bh_find_level_split": [P, P, u32, P],
        "bh_stamp_leaves": [P, P, u32],
        "bh_finalize_tree": [P, P, u32, P, P],
        "bh_leaf_begin_root": [P, P, C.POINTER(u64)],
        "bh_leaf_split": [P, u32, u32, u32, u32, C.POINTER(LeafRoundC)],
        "bh_leaf_build": [P, u32, u32],
        "bh_leaf_find": [P, P, P, u32, P],
        "bh_set_timing": [P, i32],
        "bh_profile_hist": [P, u32, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(u64),
                            C.POINTER(u64), C.POINTER(C.c_double), C.POINTER(u64)],
        "bh_get_timing": [P, P, u32],
        "bh_set_tuning": [P, C.POINTER(TuningC)],
        "bh_get_tuning": [P, C.POINTER(TuningC)],
        "bh_autotune": [P, u32, C.POINTER(TuningC), C.POINTER(C.c_float)],
        "bh_predictor_create": [C.c_int, P, P, u32, P, u32, C.c_float, C.c_float, C.POINTER(P)],
        "bh_predictor_destroy": [P],
        "bh_predict": [P, P, u64, u64, P],
        "bh_shap": [P, P, u64, u64, P],
        "bh_predict_device": [P, P, u64, u64, P, P],
        "bh_shap_device": [P, P, u64, u64, P, P],
        "bh_serve": [P, P, u64, P, P, u32, C.POINTER(ServeConfigC), P, P, P, C.POINTER(ServeStatsC)],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = C.c_int
    L.bh_last_error.restype = C.c_char_p
    L.bh_last_error.argtypes = []
    _SIGS.update(sig)
    _SIGS["bh_last_error"] = []
    _lib = L
    return L


def _check(rc: int) -> None:
    if rc == 0:
        return
    msg = lib().bh_last_error().decode(errors="replace")
    if rc > 0:
        raise ConfigError(msg)
    raise DeviceError(msg)


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _cut_tables(cuts: list) -> tuple[np.ndarray, np.ndarray]:
    """Per-feature cut arrays as one flat float32 table and its u32 offsets."""
    flat = np.ascontiguousarray(np.concatenate(cuts).astype(np.float32))
    offs = np.zeros(len(cuts) + 1, np.uint32)
    offs[1:] = np.cumsum([c.size for c in cuts])
    return flat, offs


def _hip_runtime() -> C.CDLL:
    """The HIP runtime libbonsai_hip.so links (loaded with it, not a second copy)."""
    lib()
    return C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_GLOBAL)


class DeviceMatrix:
    """A caller-owned, row-major float32 matrix in device memory: the reference's
    DeviceMatrix (include/bonsai/cuda/histogram_engine.hpp:41-50), whose producer there
    is a DLPack capsule.  Built from a host array it is allocated and filled through the
    HIP runtime the library itself links, so the pointer is valid for the device ingest
    calls; ``wrap`` adopts a pointer allocated elsewhere in that runtime (not freed here).
    A torch tensor is not accepted: the torch wheel carries a HIP runtime of its own, and
    in a process that has loaded this library torch reports no HIP device."""

    def __init__(self, X: np.ndarray | None = None, device: int = 0):
        self.ptr, self.n_rows, self.n_features, self.device, self._owned = 0, 0, 0, device, False
        if X is None:
            return
        X = np.ascontiguousarray(X, dtype=np.float32)
        if X.ndim != 2:
            raise ConfigError("a device matrix is 2-d")
        rt = _hip_runtime()
        p = C.c_void_p()
        if rt.hipSetDevice(C.c_int(device)) or rt.hipMalloc(C.byref(p), C.c_size_t(max(X.nbytes, 4))):
            raise DeviceError("device matrix allocation failed")
        self.ptr, self._owned = p.value, True
        if rt.hipMemcpy(C.c_void_p(self.ptr), C.c_void_p(X.ctypes.data), C.c_size_t(X.nbytes), C.c_int(1)):
            self.close()
            raise DeviceError("device matrix upload failed")
        self.n_rows, self.n_features = X.shape

    @classmethod
    def wrap(cls, ptr: int, n_rows: int, n_features: int, device: int = 0) -> "DeviceMatrix":
        m = cls(None, device)
        m.ptr, m.n_rows, m.n_features = int(ptr), int(n_rows), int(n_features)
        return m

    def close(self) -> None:
        if self._owned and self.ptr:
            _hip_runtime().hipFree(C.c_void_p(self.ptr))
        self.ptr, self._owned = 0, False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gather_rows(X: DeviceMatrix, rows: np.ndarray | None) -> np.ndarray:
    """Rows of a device matrix copied to the host (cuda_gather_rows,
    src/cuda/histogram_engine.cu:473-493): X[rows] for the sample rows, or the whole
    matrix when rows is None or empty."""
    r = np.zeros(0, np.uint32) if rows is None else np.ascontiguousarray(rows, dtype=np.uint32)
    out = np.empty((r.size if r.size else X.n_rows, X.n_features), np.float32)
    _check(lib().bh_gather_rows(X.device, C.c_void_p(X.ptr), X.n_rows, X.n_features, _ptr(r) if r.size else None,
                                r.size, _ptr(out)))
    return out


def copy_calibrate(nbytes: int = 1 << 30, reps: int = 20, device: int = 0) -> float:
    """Mean time (ms) of a plain device copy of nbytes; bandwidth = 2 * nbytes / time."""
    ms = C.c_float(0)
    _check(lib().bh_copy_calibrate(device, nbytes, reps, C.byref(ms)))
    return ms.value


def stream_calibrate(nbytes: int, mode: int, grid: int, reps: int = 20, device: int = 0) -> float:
    """Mean time (ms) of a bandwidth probe over nbytes: mode 1 a wide copy (traffic
    2 * nbytes), mode 2 a wide read (traffic nbytes), over grid workgroups."""
    ms = C.c_float(0)
    _check(lib().bh_stream_calibrate(device, nbytes, reps, mode, grid, C.byref(ms)))
    return ms.value


def synchronize(ctx: "HipContext") -> None:
    """Wait for the context's stream."""
    _check(lib().bh_synchronize(ctx._h))


def device_count() -> int:
    n = C.c_int(0)
    rc = lib().bh_device_count(C.byref(n))
    return n.value if rc == 0 else 0


def parse_interaction_groups(entries) -> list[list[int]]:
    """parse_interaction_groups (src/grower_impl.hpp:84-107): each entry a string of ids
    separated by ',' or '+' (empty pieces skipped), or a sequence of ids; each group
    sorted; empty groups dropped."""
    out = []
    for e in entries:
        if isinstance(e, str):
            ids = [int(p) for p in e.replace("+", ",").split(",") if p != ""]
        else:
            ids = [int(v) for v in e]
        if ids:
            out.append(sorted(ids))
    return out


class HipContext:
    """One resident training set on one GPU (the reference's CudaDeviceContext)."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        _check(lib().bh_create(device, C.byref(self._h)))
        self.n_rows = 0
        self.n_features = 0
        self.n_bins = None
        self.cfg: FitConfig | None = None
        self.n_trained = 0

    def close(self) -> None:
        if self._h:
            lib().bh_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # -- data plane --
    def upload_bins(self, bins: np.ndarray, n_bins: np.ndarray) -> None:
        bins = np.ascontiguousarray(bins, dtype=np.uint8)
        n_bins = np.ascontiguousarray(n_bins, dtype=np.uint16)
        n, f = bins.shape
        _check(lib().bh_upload_bins(self._h, _ptr(bins), n, f, f, _ptr(n_bins)))
        self.n_rows, self.n_features, self.n_bins = n, f, n_bins.copy()

    def upload_features(self, X: np.ndarray, cuts: list) -> None:
        """Bin a row-major float matrix on the device against per-feature cut tables."""
        X = np.ascontiguousarray(X, dtype=np.float32)
        n, f = X.shape
        if len(cuts) != f:
            raise ConfigError("one cut table per feature")
        flat, offs = _cut_tables(cuts)
        _check(lib().bh_upload_features(self._h, _ptr(X), n, f, _ptr(flat), _ptr(offs)))
        self.n_rows, self.n_features = n, f
        self.n_bins = np.diff(offs).astype(np.uint16)

    def upload_feature_columns(self, columns: list, cuts: list) -> None:
        """Bin per-feature float columns (host arrays of equal length) on the device."""
        cols = [np.ascontiguousarray(c, dtype=np.float32) for c in columns]
        f = len(cols)
        if f == 0 or len(cuts) != f:
            raise ConfigError("one cut table per column")
        n = cols[0].size
        if any(c.ndim != 1 or c.size != n for c in cols):
            raise ConfigError("columns must be 1-d and of equal length")
        flat, offs = _cut_tables(cuts)
        ptrs = (C.c_void_p * f)(*[c.ctypes.data for c in cols])
        _check(lib().bh_upload_feature_columns(self._h, ptrs, n, f, _ptr(flat), _ptr(offs)))
        self.n_rows, self.n_features = n, f
        self.n_bins = np.diff(offs).astype(np.uint16)

    def upload_features_device(self, X: DeviceMatrix, cuts: list) -> None:
        """Bin a device-resident matrix where it lies (cuda_ingest_device)."""
        if len(cuts) != X.n_features:
            raise ConfigError("one cut table per feature")
        flat, offs = _cut_tables(cuts)
        _check(lib().bh_upload_features_device(self._h, C.c_void_p(X.ptr), X.n_rows, X.n_features, _ptr(flat),
                                               _ptr(offs)))
        self.n_rows, self.n_features = X.n_rows, X.n_features
        self.n_bins = np.diff(offs).astype(np.uint16)

    def download_bins(self) -> np.ndarray:
        out = np.empty((self.n_rows, self.n_features), np.uint8)
        _check(lib().bh_download_bins(self._h, _ptr(out), self.n_features))
        return out

    def upload_labels(self, labels: np.ndarray) -> None:
        labels = np.ascontiguousarray(labels, dtype=np.float32)
        _check(lib().bh_upload_labels(self._h, _ptr(labels)))

    def upload_weights(self, weights: np.ndarray | None) -> None:
        """Per-row weights scaling the device gradients (None: unweighted)."""
        if weights is None:
            _check(lib().bh_upload_weights(self._h, None))
            return
        w = np.ascontiguousarray(weights, dtype=np.float32)
        if w.size != self.n_rows:
            raise ConfigError("one weight per row")
        _check(lib().bh_upload_weights(self._h, _ptr(w)))

    def configure(self, cfg: FitConfig) -> None:
        """Configure the fit.  The FitConfig owns the sampling and constraint fields:
        whatever an earlier configure or setter left (row sampler, feature selection,
        monotone and interaction constraints) is cleared and cfg's are applied."""
        c = cfg.to_c()
        if cfg.grower == "hip_levelwise" and (any(int(v) for v in cfg.monotone_constraints)
                                               or parse_interaction_groups(cfg.interaction_constraints)):
            raise ConfigError("monotone / interaction constraints are not supported by the levelwise grower")
        self.set_monotone(None)
        self.set_interaction_constraints(None)
        _check(lib().bh_configure(self._h, C.byref(c)))
        self.set_bernoulli(cfg.subsample, cfg.seed)
        self.set_feature_fraction(cfg.feature_fraction, cfg.feature_seed)
        self.set_monotone(list(cfg.monotone_constraints) or None)
        self.set_interaction_constraints(list(cfg.interaction_constraints) or None)
        self.cfg = cfg

    @property
    def bin_stride(self) -> int:
        return int(self.n_bins.max())

    # -- whole fit --
    def set_scores(self, scores: np.ndarray) -> None:
        scores = np.ascontiguousarray(scores, dtype=np.float32)
        _check(lib().bh_set_scores(self._h, _ptr(scores)))
        self.n_trained = 0

    def get_scores(self) -> np.ndarray:
        out = np.empty(self.n_rows, np.float32)
        _check(lib().bh_get_scores(self._h, _ptr(out)))
        return out

    def train(self, n_iters: int) -> None:
        _check(lib().bh_train(self._h, n_iters))
        self.n_trained += n_iters

    def max_nodes(self) -> int:
        m = C.c_uint32(0)
        _check(lib().bh_max_nodes(self._h, C.byref(m)))
        return m.value

    def trees(self) -> list[np.ndarray]:
        """Trees grown since the last set_scores, in bin space (NODE_DTYPE)."""
        m = self.max_nodes()
        cap = max(1, self.n_trained)
        nodes = np.zeros((cap, m), NODE_DTYPE)
        sizes = np.zeros(cap, np.uint32)
        n = C.c_uint32(0)
        _check(lib().bh_get_trees(self._h, _ptr(nodes), m, _ptr(sizes), cap, C.byref(n)))
        return [nodes[t, : sizes[t]].copy() for t in range(n.value)]

    def set_timing(self, on: bool) -> None:
        _check(lib().bh_set_timing(self._h, int(on)))

    def timing(self, n: int) -> np.ndarray:
        out = np.zeros(n, np.float32)
        _check(lib().bh_get_timing(self._h, _ptr(out), n))
        return out

    # -- engine seam --
    def set_gradients(self, g: np.ndarray, h: np.ndarray, it: int = 0):
        g = np.ascontiguousarray(g, np.float32)
        h = np.ascontiguousarray(h, np.float32)
        gq = np.zeros(self.n_rows, np.int8)
        hq = np.zeros(self.n_rows, np.int16)
        sc = np.zeros(2, np.float64)
        _check(lib().bh_set_gradients(self._h, _ptr(g), _ptr(h), it, _ptr(gq), _ptr(hq), _ptr(sc)))
        return gq, hq, sc

    def get_gradients(self):
        """The current iteration's gradients in row order: (g, h) float32 for an FP32
        fit, (gq, hq, scales) for a quantized one."""
        if self.cfg.quantized:
            gq = np.zeros(self.n_rows, np.int8)
            hq = np.zeros(self.n_rows, np.int16)
            sc = np.zeros(2, np.float64)
            _check(lib().bh_get_gradients(self._h, None, None, _ptr(gq), _ptr(hq), _ptr(sc)))
            return gq, hq, sc
        g = np.zeros(self.n_rows, np.float32)
        h = np.zeros(self.n_rows, np.float32)
        _check(lib().bh_get_gradients(self._h, _ptr(g), _ptr(h), None, None, None))
        return g, h

    def grow_tree(self) -> tuple[np.ndarray, np.ndarray]:
        cap = self.max_nodes()
        nodes = np.zeros(cap, NODE_DTYPE)
        nn = C.c_uint32(0)
        lor = np.zeros(self.n_rows, np.uint32)
        _check(lib().bh_grow_tree(self._h, _ptr(nodes), cap, C.byref(nn), _ptr(lor)))
        return nodes[: nn.value].copy(), lor

    def histogram(self, rows: np.ndarray) -> np.ndarray:
        rows = np.ascontiguousarray(rows, np.uint32)
        shape = (self.n_features, self.bin_stride, 2)
        if self.cfg.quantized:
            out = np.zeros(shape, np.int64)
            _check(lib().bh_histogram(self._h, _ptr(rows), rows.size, None, _ptr(out)))
        else:
            out = np.zeros(shape, np.float32)
            _check(lib().bh_histogram(self._h, _ptr(rows), rows.size, _ptr(out), None))
        return out

    def find_split(self, hist: np.ndarray, n_rows: int, depth: int = 0, totals=None,
                   scales=None) -> np.ndarray:
        out = np.zeros(1, SPLIT_DTYPE)
        tg, th = (0.0, 0.0) if totals is None else totals
        sc = None if scales is None else np.ascontiguousarray(scales, np.float64)
        if self.cfg.quantized:
            h = np.ascontiguousarray(hist, np.int64)
            _check(lib().bh_find_split(self._h, None, _ptr(h), n_rows, depth, int(totals is not None),
                                       float(tg), float(th), _ptr(sc), _ptr(out)))
        else:
            h = np.ascontiguousarray(hist, np.float32)
            _check(lib().bh_find_split(self._h, _ptr(h), None, n_rows, depth, int(totals is not None),
                                       float(tg), float(th), _ptr(sc), _ptr(out)))
        return out[0]

    def partition(self, rows: np.ndarray, feature: int, bin_: int, default_left: bool):
        rows = np.ascontiguousarray(rows, np.uint32)
        out = np.zeros(rows.size, np.uint32)
        nl = C.c_uint64(0)
        _check(lib().bh_partition(self._h, _ptr(rows), rows.size, feature, bin_, int(default_left),
                                  _ptr(out), C.byref(nl)))
        return out, nl.value

    def tuning(self) -> dict:
        t = TuningC()
        _check(lib().bh_get_tuning(self._h, C.byref(t)))
        return {k: getattr(t, k) for k, _ in TuningC._fields_}

    def set_validation(self, bins: np.ndarray, labels: np.ndarray, init_score: float,
                       early_stopping_rounds: int = 0) -> None:
        """Arm the validation plane: every training round records its mean loss;
        early_stopping_rounds > 0 stops training after that many rounds without a new
        best and keeps the trees up to the best round."""
        bins = np.ascontiguousarray(bins, np.uint8)
        labels = np.ascontiguousarray(labels, np.float32)
        if bins.ndim != 2 or labels.shape != (bins.shape[0],):
            raise ConfigError("validation bins must be [n, F] with one label per row")
        _check(lib().bh_set_validation(self._h, _ptr(bins), bins.shape[0], bins.shape[1], _ptr(labels),
                                       float(init_score), int(early_stopping_rounds)))

    def set_row_subset(self, rows) -> None:
        """Grow every following tree on these rows only (ascending, unique, below n_rows;
        the row_indices a booster's sampler hands to grow); scores still cover every
        row.  None or an empty array: every row."""
        if rows is None or len(rows) == 0:
            _check(lib().bh_set_row_subset(self._h, None, 0))
            return
        rows = np.ascontiguousarray(rows, np.uint32)
        _check(lib().bh_set_row_subset(self._h, _ptr(rows), rows.size))

    def set_bernoulli(self, subsample: float, seed: int = 0) -> None:
        """Device Bernoulli row sampler: every iteration keeps each row with probability
        `subsample` (> 0; >= 1 keeps every row), drawn from (seed, iteration, row)."""
        _check(lib().bh_set_bernoulli(self._h, float(subsample), int(seed) & ((1 << 64) - 1)))

    def set_feature_mask(self, mask) -> None:
        """Grow every following tree on the features whose mask byte is non-zero (the
        `selected` list of a per-tree feature draw); a mask selecting nothing is refused.
        None: every feature."""
        if mask is None:
            _check(lib().bh_set_feature_mask(self._h, None, 0))
            return
        mask = np.ascontiguousarray(np.asarray(mask) != 0, np.uint8)
        _check(lib().bh_set_feature_mask(self._h, _ptr(mask), mask.size))

    def set_monotone(self, dirs) -> None:
        """Monotone constraints: dirs[f] > 0 increasing, < 0 decreasing, 0 free (features
        past the list are free).  None: no constraints.  Refused with the levelwise grower."""
        if dirs is None or len(dirs) == 0:
            _check(lib().bh_set_monotone(self._h, None, 0))
            return
        d = np.ascontiguousarray(np.sign(np.asarray(dirs)), np.int8)
        _check(lib().bh_set_monotone(self._h, _ptr(d), d.size))

    def set_interaction_constraints(self, groups) -> None:
        """Interaction constraints: a list of groups, each a list of feature ids or a
        string such as "0,1" / "0+1" (the config's form).  None or []: no constraints."""
        parsed = parse_interaction_groups(groups or [])
        if not parsed:
            _check(lib().bh_set_interaction(self._h, None, None, 0))
            return
        members = np.ascontiguousarray(np.concatenate([np.asarray(g, np.uint32) for g in parsed]), np.uint32)
        sizes = np.array([len(g) for g in parsed], np.uint32)
        _check(lib().bh_set_interaction(self._h, _ptr(members), _ptr(sizes), sizes.size))

    def set_feature_fraction(self, fraction: float, seed: int = 0) -> None:
        """Device per-tree feature draw: tree t selects max(1, ceil(fraction * F))
        features, keyed by (seed, t, feature); fraction >= 1 selects every feature."""
        _check(lib().bh_set_feature_fraction(self._h, float(fraction), int(seed) & ((1 << 64) - 1)))

    def eval_history(self) -> tuple[np.ndarray, int, bool]:
        """(losses so far, best round, stopped)."""
        n = C.c_uint32(0)
        best = C.c_uint32(0)
        stopped = C.c_int32(0)
        _check(lib().bh_get_eval(self._h, None, 0, C.byref(n), C.byref(best), C.byref(stopped)))
        out = np.zeros(n.value, np.float32)
        _check(lib().bh_get_eval(self._h, _ptr(out), n.value, C.byref(n), C.byref(best), C.byref(stopped)))
        return out, best.value, bool(stopped.value)

    def set_tuning(self, **kw) -> None:
        cur = self.tuning()
        cur.update(kw)
        t = TuningC(**cur)
        _check(lib().bh_set_tuning(self._h, C.byref(t)))

    def autotune(self, iters: int = 5) -> tuple[dict, float]:
        """Pick launch shapes by timing the configured fit; model state is unchanged."""
        t = TuningC()
        ms = C.c_float(0)
        _check(lib().bh_autotune(self._h, iters, C.byref(t), C.byref(ms)))
        return {k: getattr(t, k) for k, _ in TuningC._fields_}, ms.value

    def profile_hist(self, n_iters: int) -> dict:
        """Per-launch hipEvent timing of every histogram launch over n_iters iterations:
        kernel ms, launches, algorithmic bytes / built rows (each built node's rows once)
        and scanned bytes / rows (what the launches read)."""
        ms, by, sb = C.c_double(0), C.c_double(0), C.c_double(0)
        la, ro, sr = C.c_uint64(0), C.c_uint64(0), C.c_uint64(0)
        _check(lib().bh_profile_hist(self._h, n_iters, C.byref(ms), C.byref(by), C.byref(la), C.byref(ro),
                                     C.byref(sb), C.byref(sr)))
        self.n_trained += n_iters
        return {"ms": ms.value, "bytes": by.value, "launches": la.value, "rows": ro.value,
                "scanned_bytes": sb.value, "scanned_rows": sr.value}

    # -- engine steps (the reference's GPULevelEngine / GPULeafEngine methods) --
    def begin_root(self, leaf_plane: bool = False):
        """Root histogram into slot 0; returns ((sum_g, sum_h), rows)."""
        sums = np.zeros(2, np.float64)
        rows = C.c_uint64(0)
        fn = lib().bh_leaf_begin_root if leaf_plane else lib().bh_begin_root
        _check(fn(self._h, _ptr(sums), C.byref(rows)))
        return (float(sums[0]), float(sums[1])), rows.value

    def partition_level(self, ops) -> np.ndarray:
        """ops: (parent_slot, left_slot, right_slot, feature, bin, default_left) tuples;
        returns [n_left, n_right] per op."""
        a = np.ascontiguousarray(np.array(ops, np.uint32).reshape(-1, 6))
        counts = np.zeros(2 * len(a), np.uint32)
        _check(lib().bh_partition_level(self._h, _ptr(a), len(a), _ptr(counts)))
        return counts.reshape(-1, 2)

    def advance_level(self, ops, per_node: bool = False) -> None:
        """ops: (parent_slot, small_slot, large_slot) tuples; per_node launches one build
        and one subtraction per node (the reference's cadence) instead of one per level."""
        a = np.ascontiguousarray(np.array(ops, np.uint32).reshape(-1, 3))
        fn = lib().bh_advance_level_per_node if per_node else lib().bh_advance_level
        _check(fn(self._h, _ptr(a), len(a)))

    def advance_layout_only(self) -> None:
        _check(lib().bh_advance_layout_only(self._h))

    @staticmethod
    def _step_nodes(depths):
        arr = (StepNodeC * max(1, len(depths)))()
        for i, d in enumerate(depths):
            arr[i].depth = int(d)
        return arr

    @staticmethod
    def _step_out(arr, n):
        return [(arr[i].rows, arr[i].sum_g, arr[i].sum_h, arr[i].isum_g, arr[i].isum_h) for i in range(n)]

    def find_splits_many(self, depths):
        """Split of every slot of the current level (node i = slot i); returns (splits,
        [(rows, sum_g, sum_h, isum_g, isum_h)])."""
        nodes = self._step_nodes(depths)
        out = np.zeros(max(1, len(depths)), SPLIT_DTYPE)
        _check(lib().bh_find_splits_many(self._h, nodes, len(depths), _ptr(out)))
        return out[: len(depths)], self._step_out(nodes, len(depths))

    def find_level_split(self, depths):
        nodes = self._step_nodes(depths)
        out = np.zeros(max(1, len(depths)), SPLIT_DTYPE)
        _check(lib().bh_find_level_split(self._h, nodes, len(depths), _ptr(out)))
        return out[: len(depths)], self._step_out(nodes, len(depths))

    def stamp_leaves(self, stamps) -> None:
        a = np.ascontiguousarray(np.array(stamps, np.uint32).reshape(-1, 2))
        _check(lib().bh_stamp_leaves(self._h, _ptr(a), len(a)))

    def finalize_tree(self, node_values):
        v = np.ascontiguousarray(node_values, np.float32)
        values = np.zeros(self.n_rows, np.float32)
        ids = np.zeros(self.n_rows, np.uint32)
        _check(lib().bh_finalize_tree(self._h, _ptr(v), v.size, _ptr(values), _ptr(ids)))
        return values, ids

    def leaf_split(self, parent_slot, feature, bin_, default_left):
        r = LeafRoundC()
        _check(lib().bh_leaf_split(self._h, parent_slot, feature, bin_, int(default_left), C.byref(r)))
        return r.left_slot, r.right_slot, r.left_count, r.right_count

    def leaf_build(self, small_slot, large_slot) -> None:
        _check(lib().bh_leaf_build(self._h, small_slot, large_slot))

    def leaf_find(self, depths, slots):
        nodes = self._step_nodes(depths)
        sl = np.ascontiguousarray(np.array(slots, np.uint32))
        out = np.zeros(max(1, len(depths)), SPLIT_DTYPE)
        _check(lib().bh_leaf_find(self._h, nodes, _ptr(sl), len(depths), _ptr(out)))
        return out[: len(depths)], self._step_out(nodes, len(depths))

    def time_root_histogram(self, reps: int = 10) -> float:
        ms = C.c_float(0)
        _check(lib().bh_time_root_histogram(self._h, reps, C.byref(ms)))
        return ms.value


class Predictor:
    """A resident ensemble for predict / TreeSHAP (the reference's CudaPredictPlan)."""

    def __init__(self, trees: list[np.ndarray], n_bins: np.ndarray, learning_rate: float,
                 init_score: float, device: int = 0):
        offsets = np.zeros(len(trees), np.uint32)
        pos = 0
        for i, t in enumerate(trees):
            offsets[i] = pos
            pos += len(t)
        nodes = np.concatenate(trees) if trees else np.zeros(0, NODE_DTYPE)
        nodes = np.ascontiguousarray(nodes, NODE_DTYPE)
        last = np.ascontiguousarray(np.asarray(n_bins, np.uint16) - 1, np.uint16)
        self.n_features = len(last)
        self._h = C.c_void_p()
        _check(lib().bh_predictor_create(device, _ptr(nodes), _ptr(offsets), len(trees), _ptr(last),
                                         self.n_features, learning_rate, init_score, C.byref(self._h)))

    def close(self):
        if self._h:
            lib().bh_predictor_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def predict(self, bins: np.ndarray) -> np.ndarray:
        bins = np.ascontiguousarray(bins, np.uint8)
        out = np.zeros(bins.shape[0], np.float32)
        _check(lib().bh_predict(self._h, _ptr(bins), bins.shape[0], bins.shape[1], _ptr(out)))
        return out

    def shap(self, bins: np.ndarray) -> np.ndarray:
        bins = np.ascontiguousarray(bins, np.uint8)
        out = np.zeros((bins.shape[0], self.n_features + 1), np.float64)
        _check(lib().bh_shap(self._h, _ptr(bins), bins.shape[0], bins.shape[1], _ptr(out)))
        return out

    def serve(self, rows: np.ndarray, req_offsets: np.ndarray, arrival_s: np.ndarray, n_slots: int = 4,
              do_shap: bool = True, max_batch_rows: int = 65536, hbm_budget: int = 1 << 30):
        """Continuous-batching request loop (bh_serve): returns (pred, shap, latency_ms, stats)."""
        rows = np.ascontiguousarray(rows, np.uint8)
        offs = np.ascontiguousarray(req_offsets, np.uint64)
        arr = np.ascontiguousarray(arrival_s, np.float64)
        n_req = arr.size
        pred = np.zeros(rows.shape[0], np.float32)
        shap = np.zeros((rows.shape[0], self.n_features + 1), np.float64) if do_shap else None
        lat = np.zeros(n_req, np.float64)
        cfg = ServeConfigC(n_slots, int(do_shap), max_batch_rows, hbm_budget)
        st = ServeStatsC()
        _check(lib().bh_serve(self._h, _ptr(rows), rows.shape[1], _ptr(offs), _ptr(arr), n_req, C.byref(cfg),
                              _ptr(pred), _ptr(shap), _ptr(lat), C.byref(st)))
        stats = {k: getattr(st, k) for k, _ in ServeStatsC._fields_}
        return pred, shap, lat, stats
