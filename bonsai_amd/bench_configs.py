"""Measurements of the other workload configurations (one JSON line each).

    python -m bonsai_amd.bench_configs --config {0,2,3,4} [--iters K]

0  100,000 x 28 regression, MSE, depthwise depth 6, 256 bins, 100 iterations, lr 0.1:
   full fits, quantized and FP32, with training RMSE.
2  16M x 28 binary, hip_leafwise 255 leaves, depth 12, quantized (int8 g / int16 h,
   stochastic rounding): K iterations timed (seconds per 500 extrapolated), held-out AUC
   against the FP32 leafwise grower after the same K iterations, and the quantized root
   histogram's error against an FP64 histogram of the unquantized gradients.
3  1,000,000 x 2,000 u8 bins (64 values; 90% sparse features), regression, hip_levelwise
   depth 10: K iterations, one hipGraph per iteration against eager launches and against
   a host round-trip after every level; and one tree grown three ways (the device grower,
   the host-driven engine steps with one histogram launch per level, and the steps with
   one launch per node, the reference's cadence).
4  Serving: 1,000 random depth-8 trees over 28 features (leaf values N(0, 0.01^2)),
   20,000 predict + TreeSHAP requests, lognormal sizes (median 256, sigma 1, clipped to
   [1, 65536]), Poisson arrivals at a fixed rate, admitted under an HBM budget:
   p50 / p99 latency, rows/s, and SHAP additivity on every row.

Synthetic data throughout (bonsai_amd.synth); nothing here reads the reference.
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

import bonsai_amd as B
from bonsai_amd import binning, synth


def log(msg: str) -> None:
    print(f"[bench_configs] {msg}", file=sys.stderr, flush=True)


def logloss_init(y: np.ndarray) -> float:
    m = np.float32(y.sum(dtype=np.float64) / y.size)
    return float(np.float32(np.log(m / (np.float32(1) - m))))


def timed_train(ctx, iters: int) -> float:
    B.synchronize(ctx)
    t0 = time.perf_counter()
    ctx.train(iters)
    B.synchronize(ctx)
    return time.perf_counter() - t0


def config0(args) -> dict:
    n = 100_000
    X = synth.tall_features(n, config_id=0)
    y = synth.regression_target(X)
    cuts = binning.fit_cuts(X, max_bin=256, n_samples=n)
    out = {"config": 0, "workload": "100k x 28 regression, MSE, depthwise depth 6, 256 bins, 100 iterations"}
    for quantized in (True, False):
        ctx = B.HipContext(0)
        ctx.upload_features(X, cuts)
        ctx.upload_labels(y)
        ctx.configure(B.FitConfig(grower="hip_depthwise", objective="mse", quantized=quantized, max_depth=6,
                                  learning_rate=0.1, lambda_l2=1.0))
        init = float(np.float32(y.sum(dtype=np.float64) / n))
        ctx.set_scores(np.full(n, init, np.float32))
        ctx.train(2)
        ctx.set_scores(np.full(n, init, np.float32))
        dt = timed_train(ctx, 100)
        rmse = float(np.sqrt(np.mean((ctx.get_scores().astype(np.float64) - y) ** 2)))
        ctx.close()
        out["quantized" if quantized else "fp32"] = {"fit_s": round(dt, 4), "ms_per_iter": round(dt * 10, 4),
                                                     "train_rmse": round(rmse, 6)}
    return out


def config2(args) -> dict:
    n, n_held = 16_000_000, 1_600_000
    X = synth.tall_features(n, config_id=1)
    y = synth.binary_label(X, config_id=1)
    Xh = synth.tall_features(n_held, config_id=1, stream=1)
    yh = synth.binary_label(Xh, config_id=1, stream=1)
    cuts = binning.fit_cuts(X, max_bin=255, n_samples=200_000)
    nb = binning.n_bins_of(cuts)
    bins_h = binning.transform(Xh, cuts)
    init = logloss_init(y)
    ctx = B.HipContext(0)
    ctx.upload_features(X, cuts)
    ctx.upload_labels(y)
    res = {"config": 2, "workload": "16M x 28 binary, hip_leafwise 255 leaves, depth 12", "iters": args.iters}
    if args.find_blocks:
        # workgroups per node of the round's split find, quantized fit
        sweep = {}
        for fb in (int(v) for v in args.find_blocks.split(",")):
            ctx.configure(B.FitConfig(grower="hip_leafwise", objective="logloss", quantized=True, max_depth=12,
                                      max_leaves=255, learning_rate=0.1, lambda_l2=1.0))
            ctx.set_tuning(find_blocks=fb)
            ctx.set_scores(np.full(n, init, np.float32))
            ctx.train(1)
            ctx.set_scores(np.full(n, init, np.float32))
            sweep[fb] = round(timed_train(ctx, args.iters) / args.iters * 1e3, 3)
        res["find_blocks_ms_per_iter"] = sweep
    for quantized in (True, False):
        ctx.configure(B.FitConfig(grower="hip_leafwise", objective="logloss", quantized=quantized, max_depth=12,
                                  max_leaves=255, learning_rate=0.1, lambda_l2=1.0))
        ctx.set_scores(np.full(n, init, np.float32))
        ctx.train(1)
        ctx.set_scores(np.full(n, init, np.float32))
        dt = timed_train(ctx, args.iters)
        p = B.Predictor(ctx.trees(), nb, 0.1, init)
        auc = synth.auc(p.predict(bins_h), yh)
        p.close()
        res["quantized" if quantized else "fp32"] = {"ms_per_iter": round(dt / args.iters * 1e3, 3),
                                                     "fit_s_500_extrapolated": round(dt / args.iters * 500, 3),
                                                     "held_out_auc": round(auc, 6)}
    res["auc_delta_quantized"] = round(res["quantized"]["held_out_auc"] - res["fp32"]["held_out_auc"], 6)
    # Quantized root histogram against an FP64 histogram of the unquantized gradients
    # (1M-row sample, score 0: g = 0.5 - y, h = 0.25 exactly in float).
    m = 1_000_000
    g = (np.float32(0.5) - y[:m]).astype(np.float32)
    h = np.full(m, 0.25, np.float32)
    rng = np.random.default_rng(7)
    g = g + rng.normal(0, 0.05, m).astype(np.float32)          # non-dyadic values
    h = h * rng.uniform(0.5, 1.0, m).astype(np.float32)
    ctx_small = B.HipContext(0)
    ctx_small.upload_features(X[:m], cuts)
    ctx_small.configure(B.FitConfig(grower="hip_depthwise", quantized=True, max_depth=6))
    gq, hq, sc = ctx_small.set_gradients(g, h)
    hist_q = ctx_small.histogram(np.arange(m, dtype=np.uint32)).astype(np.float64)
    bins = ctx_small.download_bins()
    ctx_small.close()
    ctx.close()
    SB = hist_q.shape[1]
    ref = np.zeros((bins.shape[1], SB, 2), np.float64)
    for f in range(bins.shape[1]):
        ref[f, :, 0] = np.bincount(bins[:, f], weights=g.astype(np.float64), minlength=SB)[:SB]
        ref[f, :, 1] = np.bincount(bins[:, f], weights=h.astype(np.float64), minlength=SB)[:SB]
    rec = hist_q * np.array([sc[0], sc[1]])
    err = np.abs(rec - ref)
    scale_g = np.abs(ref[..., 0]).sum(axis=1).max()
    res["hist_error_vs_fp64"] = {
        "rows": m, "max_abs_g": float(err[..., 0].max()), "max_abs_h": float(err[..., 1].max()),
        "max_abs_g_over_quantum": float(err[..., 0].max() / sc[0]),
        "rel_l1_g": float(err[..., 0].sum() / np.abs(ref[..., 0]).sum()),
        "rel_l1_h": float(err[..., 1].sum() / np.abs(ref[..., 1]).sum()),
        "scale_g": float(sc[0]), "scale_h": float(sc[1]), "feature_g_mass": float(scale_g)}
    return res


def wide_u8(n: int, f: int = 2000):
    """Config 3 data directly as u8 bins (the raw values are the 64 bin indices)."""
    rng = np.random.default_rng(np.random.SeedSequence([42, 3, n, f, 0]))
    n_sparse = int(round(0.9 * f))
    bins = np.empty((n, f), np.uint8)
    for j in range(f):
        if j < n_sparse:
            col = np.zeros(n, np.uint8)
            nz = rng.random(n, dtype=np.float32) >= np.float32(0.95)
            col[nz] = rng.integers(1, 64, int(nz.sum()), dtype=np.uint8)
        else:
            col = rng.integers(0, 64, n, dtype=np.uint8)
        bins[:, j] = col
    trng = np.random.default_rng(np.random.SeedSequence([42, 3, 50, 3, 7]))
    y = np.zeros(n, np.float32)
    for _ in range(50):
        feats = trng.choice(f, 3, replace=False)
        thr = trng.integers(1, 63, 3)
        w = trng.standard_normal(8).astype(np.float32)
        idx = np.zeros(n, np.int64)
        for k in range(3):
            idx = idx * 2 + (bins[:, feats[k]] > thr[k])
        y += w[idx]
    y += np.float32(0.1) * rng.standard_normal(n, dtype=np.float32)
    return bins, y


def config3(args) -> dict:
    n, f = 1_000_000, 2000
    t = time.perf_counter()
    bins, y = wide_u8(n, f)
    log(f"wide data {n} x {f} in {time.perf_counter() - t:.1f}s")
    nb = np.full(f, 66, np.uint16)   # 64 values + closer + missing sentinel, as binning would give
    ctx = B.HipContext(0)
    ctx.upload_bins(bins, nb)
    ctx.upload_labels(y)
    init = float(np.float32(y.sum(dtype=np.float64) / n))
    res = {"config": 3, "workload": "1M x 2000 u8 (64 bins), regression, hip_levelwise depth 10", "iters": args.iters}
    for mode, kw in (("graph", dict(use_graph=True)), ("eager", dict(use_graph=False)),
                     ("host_sync_per_level", dict(use_graph=False, host_sync=True))):
        for quantized in (True, False):
            ctx.configure(B.FitConfig(grower="hip_levelwise", objective="mse", quantized=quantized, max_depth=10,
                                      learning_rate=0.1, lambda_l2=1.0, **kw))
            ctx.set_scores(np.full(n, init, np.float32))
            ctx.train(1)
            ctx.set_scores(np.full(n, init, np.float32))
            dt = timed_train(ctx, args.iters)
            res.setdefault(mode, {})["quantized" if quantized else "fp32"] = {
                "ms_per_iter": round(dt / args.iters * 1e3, 3),
                "fit_s_300_extrapolated": round(dt / args.iters * 300, 3)}
    # One tree from the same gradients three ways: the device-resident grower (every
    # node of a level in one launch, no host round trip), the engine steps driven from
    # the host with one launch per level, and the steps with one build and one
    # subtraction launch per node (the reference's cadence).
    g = (np.full(n, init, np.float32) - y).astype(np.float32)
    h = np.ones(n, np.float32)
    trees = {}
    for quantized in (True, False):
        ctx.configure(B.FitConfig(grower="hip_levelwise", objective="mse", quantized=quantized, max_depth=10,
                                  learning_rate=0.1, lambda_l2=1.0))
        ctx.set_gradients(g, h)
        out = {}
        for name, fn in (("device_grower", lambda: ctx.grow_tree()),
                         ("steps_one_launch_per_level", lambda: levelwise_steps_tree(ctx, 10, False)),
                         ("steps_one_launch_per_node", lambda: levelwise_steps_tree(ctx, 10, True))):
            fn()
            t = time.perf_counter()
            reps = 3
            for _ in range(reps):
                fn()
            out[name] = round((time.perf_counter() - t) / reps * 1e3, 3)
        trees["quantized" if quantized else "fp32"] = out
    res["one_tree_ms"] = trees
    ctx.close()
    return res


def levelwise_steps_tree(ctx, depth: int, per_node: bool) -> None:
    """An oblivious tree through the engine steps (find_level_split, partition_level,
    advance_level), one histogram launch per level or per node."""
    ctx.begin_root()
    n = 1
    for d in range(depth):
        splits, _ = ctx.find_level_split([d] * n)
        if not splits[0]["valid"]:
            break
        s0 = splits[0]
        counts = ctx.partition_level([(i, 2 * i, 2 * i + 1, s0["feature"], s0["bin"], s0["default_left"])
                                      for i in range(n)])
        if d + 1 == depth:
            ctx.advance_layout_only()
        else:
            ctx.advance_level([(i, 2 * i + (0 if counts[i][0] <= counts[i][1] else 1),
                                2 * i + (1 if counts[i][0] <= counts[i][1] else 0)) for i in range(n)],
                              per_node=per_node)
        n *= 2
    ctx.stamp_leaves([(i, i) for i in range(n)])
    ctx.finalize_tree(np.zeros(n, np.float32))


def random_ensemble(n_trees: int, depth: int, n_features: int, n_bins: np.ndarray, seed: int = 11):
    rng = np.random.default_rng(seed)
    trees = []
    n_int = (1 << depth) - 1
    for _ in range(n_trees):
        t = np.zeros((2 << depth) - 1, B.NODE_DTYPE)
        cover = np.zeros(t.size)
        cover[0] = 1e6
        for i in range(n_int):
            f = int(rng.integers(0, n_features))
            b = int(rng.integers(0, n_bins[f] - 2)) if n_bins[f] > 2 else 0
            frac = rng.uniform(0.2, 0.8)
            cover[2 * i + 1] = cover[i] * frac
            cover[2 * i + 2] = cover[i] - cover[2 * i + 1]
            t[i] = (f, b, int(rng.integers(0, 2)), 2 * i + 1, 2 * i + 2, 0.0, cover[i], 1.0)
        for i in range(n_int, t.size):
            t[i] = (-1, 0, 0, 0, 0, np.float32(rng.normal(0, 0.01)), cover[i], 0.0)
        trees.append(t)
    return trees


def shap_path_work(trees):
    """(paths, Σ_paths K(K+1)/2 + 3K) with K = distinct features on the root-to-leaf path + 1."""
    n_paths, work = 0, 0
    for t in trees:
        stack = [(0, frozenset())]
        while stack:
            i, feats = stack.pop()
            f = int(t["feature"][i])
            if f < 0:
                k = len(feats) + 1
                n_paths += 1
                work += k * (k + 1) // 2 + 3 * k
                continue
            nf = feats | {f}
            stack.append((int(t["left"][i]), nf))
            stack.append((int(t["right"][i]), nf))
    return n_paths, work


def config4(args) -> dict:
    X = synth.tall_features(200_000, config_id=0, stream=2)
    cuts = binning.fit_cuts(X, max_bin=256, n_samples=200_000)
    nb = binning.n_bins_of(cuts)
    pool = binning.transform(X, cuts)
    trees = random_ensemble(1000, 8, 28, nb)
    p = B.Predictor(trees, nb, 1.0, 0.0)
    rng = np.random.default_rng(5)
    n_req = args.requests
    sizes = np.clip(np.round(rng.lognormal(np.log(256), 1.0, n_req)), 1, 65536).astype(np.int64)
    rows = pool[rng.integers(0, pool.shape[0], int(sizes.sum()))]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    arrival = np.cumsum(rng.exponential(1.0 / args.qps, n_req))
    log(f"{n_req} requests, {rows.shape[0]} rows, offered {args.qps} req/s over {arrival[-1]:.2f}s")
    out = {"config": 4, "workload": "1,000 depth-8 trees, 28 features; 20,000 predict + TreeSHAP requests",
           "requests": n_req, "rows": int(rows.shape[0]), "offered_qps": args.qps,
           "size_median": float(np.median(sizes)), "size_max": int(sizes.max())}
    # the TreeSHAP kernel alone on one 65,536-row batch (rocprof times k_shap itself):
    # work = Σ over paths of K(K+1)/2 + 3K fused multiply-adds per row, K = the path's
    # elements (its distinct features + 1)
    n_paths, fma_per_row = shap_path_work(trees)
    batch = pool[rng.integers(0, pool.shape[0], 65536)]
    p.shap(batch)
    t0 = time.perf_counter()
    for _ in range(5):
        p.shap(batch)
    dt = (time.perf_counter() - t0) / 5
    out["shap_batch"] = {"rows": 65536, "paths": n_paths, "fma_per_row": fma_per_row,
                         "flop_per_launch": 2 * fma_per_row * 65536, "wall_ms_incl_copies": round(dt * 1e3, 3)}
    for do_shap in (False, True):
        pred, shap, lat, st = p.serve(rows, offs, arrival, n_slots=args.slots, do_shap=do_shap,
                                      max_batch_rows=65536, hbm_budget=args.hbm_budget)
        key = "predict_shap" if do_shap else "predict"
        out[key] = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in st.items()}
        if do_shap:
            margin = pred.astype(np.float64)
            rel = np.abs(shap.sum(axis=1) - margin) / np.maximum(1.0, np.abs(margin))
            out[key]["additivity_max_rel_err"] = float(rel.max())
            out[key]["additivity_ok_1e-5"] = bool(rel.max() <= 1e-5)
    p.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--config", type=int, required=True, choices=[0, 2, 3, 4])
    ap.add_argument("--iters", type=int, default=20)
    ap.add_argument("--requests", type=int, default=20_000)
    ap.add_argument("--qps", type=float, default=2000.0)
    ap.add_argument("--slots", type=int, default=4)
    ap.add_argument("--hbm-budget", type=int, default=2 << 30)
    ap.add_argument("--find-blocks", default="", help="leafwise: comma list of find_blocks values to time (config 2)")
    a = ap.parse_args()
    if B.device_count() < 1:
        print("needs a HIP device", file=sys.stderr)
        return 1
    fn = {0: config0, 2: config2, 3: config3, 4: config4}[a.config]
    print(json.dumps(fn(a)), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
