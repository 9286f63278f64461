#!/usr/bin/env python3
"""Fit-time benchmark of the depthwise hot path at 16,000,000 x 28, depth 8.

Workload: the tall synthetic binary task (features 0-9 N(0,1), 10-17 U(-1,1), 18-23
lognormal, 24-27 integers in [0,16), 5% NaN in 18-27; Bernoulli(sigmoid(2(sin x0 +
x1 x2 - 0.5 x10^2 + 0.3 x24/16))) labels), logistic objective, hip_depthwise grower,
max_depth 8, 255 bins, learning rate 0.1, lambda 1.  Cuts are fitted on a 200,000-row
sample (the reference's default n_samples), the plane is binned on the device.

One step = one boosting iteration (gradient, quantization, every level's histogram /
split find / partition, leaf values, score update), captured as one hipGraph.  The
headline value is the quantized (int8 g / int16 h, stochastic rounding) grower's fit time
for 500 iterations; the FP32 fixed-order grower runs the same iterations beside it and
both models are scored by held-out AUC on 1.6M rows drawn from the same distribution.

roofline: the histogram kernels, every launch of a few eager iterations timed with
hipEvents on the engine stream.  achieved = algorithmic bytes / kernel time, where the
algorithmic bytes of a launch are every built node's rows read once, rows x (28 bin
bytes + 4 gradient-pair bytes), plus 8 bytes per built (node, feature, bin) cell; the
rate of the bytes the launches actually scan (levels whose rows were not moved read
each parent's whole segment) is reported beside it as scanned_GBps.
Equal AUC: after the timed region (untimed), both precisions fit 500 iterations and
their held-out AUCs are compared; the root histogram of both precisions is compared with
an FP64 histogram of the same float gradients.
cpu_baseline: the C++/OpenMP CPU restatement of the reference's depthwise grower on the
full 16M-row plane for a few iterations, scaled to 500 iterations.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "fit s/500 iters @16M×28 d8 (FP32 vs quantized, equal AUC); hist HBM GB/s"
# Histogram HBM traffic per launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE counter runs
# (calibrated against a copy of known size), with the algorithmic bytes of the same launches.
TRAFFIC = ROOT / "profiles" / "hist_quantized_16Mx28_traffic.json"
HBM_PEAK_GBPS = 8000.0
FULL_ITERS = 500


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def host_cores() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    try:
        q, p = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        if q != "max":
            n = min(n, max(1, int(int(q) / int(p))))
    except (OSError, ValueError):
        pass
    return max(1, min(n, 16))


def make_data(n_rows: int, n_held: int):
    from bonsai_amd import binning, synth
    t = time.perf_counter()
    X = synth.tall_features(n_rows, config_id=1, stream=0)
    y = synth.binary_label(X, config_id=1, stream=0)
    Xh = synth.tall_features(n_held, config_id=1, stream=1)
    yh = synth.binary_label(Xh, config_id=1, stream=1)
    cuts = binning.fit_cuts(X, max_bin=255, n_samples=200_000, seed=0)
    log(f"data {n_rows}x28 + {n_held} held out, cuts fitted in {time.perf_counter() - t:.1f}s")
    return X, y, Xh, yh, cuts


def fit(B, ctx, quantized: bool, steps: int, warmup: int, init: float, n_rows: int, autotune: int,
        shapes: dict | None = None):
    ctx.configure(B.FitConfig(grower="hip_depthwise", objective="logloss", quantized=quantized,
                              use_graph=True, max_depth=8, learning_rate=0.1, lambda_l2=1.0))
    start = np.full(n_rows, init, np.float32)
    tuned = None
    if shapes:
        ctx.set_tuning(**shapes)
        tuned = {"shapes": ctx.tuning(), "fixed": True}
    if autotune:
        ctx.set_scores(start)
        t0 = time.perf_counter()
        shapes, ms = ctx.autotune(autotune)
        tuned = {"shapes": shapes, "ms_per_iter": round(ms, 4), "seconds": round(time.perf_counter() - t0, 2)}
        log(f"autotune ({'quantized' if quantized else 'fp32'}): {shapes} -> {ms:.3f} ms/iter")
    if warmup:
        ctx.set_scores(start)
        ctx.train(warmup)
    ctx.set_scores(start)
    B.synchronize(ctx)
    t0 = time.perf_counter()
    ctx.train(steps)            # returns after the engine stream has drained
    B.synchronize(ctx)
    dt = time.perf_counter() - t0
    return dt, ctx.trees(), tuned


def held_out_auc(B, trees, n_bins, init, bins_h, yh):
    from bonsai_amd import synth
    p = B.Predictor(trees, n_bins, 0.1, init)
    margin = p.predict(bins_h)
    p.close()
    return synth.auc(margin, yh)


def cpu_info() -> dict:
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    quota = None
    try:
        q, per = Path("/sys/fs/cgroup/cpu.max").read_text().split()
        quota = None if q == "max" else round(int(q) / int(per), 2)
    except (OSError, ValueError):
        pass
    return {"model": model, "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)),
            "cgroup_quota_cpus": quota}


def cpu_baseline(bins, nb, y, iters: int):
    """The oracle's OpenMP depthwise grower (the reference's row-block fill) on the whole
    plane for `iters` iterations, on min(cores, quota, 16) threads."""
    from oracle import HIST_F32_OMP, cfg, load_oracle
    orc = load_oracle()
    cores = host_cores()
    info = cpu_info()
    log(f"cpu: {info['model']}, nproc {info['nproc']}, affinity {info['affinity']}, "
        f"quota {info['cgroup_quota_cpus']}, using {cores} threads")
    t0 = time.perf_counter()
    orc.train_depthwise(bins, nb, y, 1, iters, 0.1, cfg(max_depth=8), mode=HIST_F32_OMP, n_threads=cores)
    dt = time.perf_counter() - t0
    return {"value": round(dt / iters * FULL_ITERS, 3), "unit": "s (fit of 500 iterations, extrapolated)",
            "cores": cores, "kind": "port", "cpu": info,
            "sample": f"{iters} iterations on all {bins.shape[0]} rows ({dt:.1f}s, {dt / iters:.2f} s per "
                      f"iteration), scaled linearly to {FULL_ITERS} iterations"}


def hist_error(B, ctx, scores, y, nb, bins, what: str):
    """Root histogram of the logistic gradients at `scores` in both precisions against an
    FP64 histogram of the same float gradients."""
    from oracle import load_oracle
    orc = load_oracle()
    X_rows = scores.size
    g, h = orc.logloss_grad(scores, y)
    SB = int(nb.max())
    ref = np.zeros((bins.shape[1], SB, 2), np.float64)
    for f in range(bins.shape[1]):
        ref[f, :, 0] = np.bincount(bins[:, f], weights=g.astype(np.float64), minlength=SB)[:SB]
        ref[f, :, 1] = np.bincount(bins[:, f], weights=h.astype(np.float64), minlength=SB)[:SB]
    rows = np.arange(X_rows, dtype=np.uint32)
    out = {"rows": X_rows, "gradients": what}
    for quantized in (True, False):
        ctx.configure(B.FitConfig(grower="hip_depthwise", objective="logloss", quantized=quantized, max_depth=8))
        if quantized:
            gq, hq, sc = ctx.set_gradients(g, h)
            got = ctx.histogram(rows).astype(np.float64) * np.array([sc[0], sc[1]])
        else:
            ctx.set_gradients(g, h)
            got = ctx.histogram(rows).astype(np.float64)
        err = np.abs(got - ref)
        key = "quantized" if quantized else "fp32"
        out[key] = {"max_abs_g": float(err[..., 0].max()), "max_abs_h": float(err[..., 1].max()),
                    "rel_l1_g": float(err[..., 0].sum() / np.abs(ref[..., 0]).sum()),
                    "rel_l1_h": float(err[..., 1].sum() / np.abs(ref[..., 1]).sum())}
        if quantized:
            out[key]["scales"] = [float(sc[0]), float(sc[1])]
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=FULL_ITERS)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=16_000_000)
    ap.add_argument("--held-out", type=int, default=1_600_000)
    ap.add_argument("--prof-iters", type=int, default=3, help="eager iterations timed per histogram launch")
    ap.add_argument("--autotune", type=int, default=5, help="iterations per autotuner trial (0: default shapes)")
    ap.add_argument("--tuning", default="", help="fixed launch shapes, e.g. scatter_lds_bytes=49152,find_blocks=4")
    ap.add_argument("--skip-fp32", action="store_true")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--cpu-iters", type=int, default=3, help="CPU baseline iterations on the full plane")
    ap.add_argument("--auc-iters", type=int, default=FULL_ITERS,
                    help="untimed fit length of the equal-AUC check (0: skip)")
    ap.add_argument("--skip-hist-error", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the 1 GiB calibration copy kernel (for HBM counter scaling) and exit")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        return 2

    import bonsai_amd as B
    if B.device_count() < 1:
        print("bench.py needs a HIP device", file=sys.stderr)
        return 1
    from bonsai_amd import binning

    copy_ms = B.copy_calibrate(1 << 30, 20)
    copy_gbps = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    log(f"calibration copy of 1 GiB: {copy_ms * 1e3:.1f} us, {copy_gbps:.0f} GB/s")
    # streaming ceilings: the best grid for a wide copy and a wide read of 1 GiB
    wide = {}
    for mode, name, traffic_x in ((1, "copy", 2), (2, "read", 1)):
        best = max((traffic_x * (1 << 30) / (B.stream_calibrate(1 << 30, mode, g, 10) * 1e-3) / 1e9, g)
                   for g in (1024, 2048, 4096, 8192, 16384))
        wide[name] = best
        log(f"wide {name} of 1 GiB: {best[0]:.0f} GB/s at {best[1]} workgroups")
    if a.copy_calibration:
        print(json.dumps({"copy_bytes": 1 << 30, "copy_ms": copy_ms, "copy_GBps": round(copy_gbps, 1),
                          "wide_copy_GBps": round(wide["copy"][0], 1), "wide_read_GBps": round(wide["read"][0], 1)}))
        return 0

    X, y, Xh, yh, cuts = make_data(a.rows, a.held_out)
    nb = binning.n_bins_of(cuts)
    bins_h = binning.transform(Xh, cuts)
    init = float(np.float32(np.log(np.float32(y.sum(dtype=np.float64) / y.size)
                                   / (np.float32(1) - np.float32(y.sum(dtype=np.float64) / y.size)))))

    ctx = B.HipContext(0)
    t = time.perf_counter()
    ctx.upload_features(X, cuts)
    ctx.upload_labels(y)
    log(f"plane binned on the device in {time.perf_counter() - t:.2f}s")

    shapes = {k: int(v) for k, v in (kv.split("=") for kv in a.tuning.split(",") if kv)} or None
    dt_q, trees_q, tuned_q = fit(B, ctx, True, a.steps, a.warmup, init, a.rows, 0 if shapes else a.autotune, shapes)
    auc_q = held_out_auc(B, trees_q, nb, init, bins_h, yh)
    log(f"quantized: {a.steps} iters in {dt_q:.3f}s, held-out AUC {auc_q:.5f}")
    prof = ctx.profile_hist(a.prof_iters)
    hist_gbps = prof["bytes"] / (prof["ms"] * 1e-3) / 1e9
    scanned_gbps = prof["scanned_bytes"] / (prof["ms"] * 1e-3) / 1e9
    per_launch_bytes = prof["bytes"] / prof["launches"]
    per_launch_us = prof["ms"] * 1e3 / prof["launches"]
    log(f"histogram: {prof['launches']} launches, {per_launch_us:.1f} us avg, {hist_gbps:.0f} GB/s algorithmic "
        f"({prof['rows']} built rows), {scanned_gbps:.0f} GB/s scanned ({prof['scanned_rows']} rows)")

    fp32 = None
    if not a.skip_fp32:
        dt_f, trees_f, tuned_f = fit(B, ctx, False, a.steps, a.warmup, init, a.rows, 0 if shapes else a.autotune,
                                     shapes)
        auc_f = held_out_auc(B, trees_f, nb, init, bins_h, yh)
        log(f"fp32: {a.steps} iters in {dt_f:.3f}s, held-out AUC {auc_f:.5f}")
        # the FP32 histogram kernel against the same roofline (built-node bytes, 8-byte g/h)
        prof_f = ctx.profile_hist(a.prof_iters)
        hist_f_gbps = prof_f["bytes"] / (prof_f["ms"] * 1e-3) / 1e9
        log(f"fp32 histogram: {prof_f['launches']} launches, {prof_f['ms'] * 1e3 / prof_f['launches']:.1f} us avg, "
            f"{hist_f_gbps:.0f} GB/s algorithmic")
        fp32 = {"fit_s_500": round(dt_f / a.steps * FULL_ITERS, 3), "ms_per_step": round(dt_f / a.steps * 1e3, 3),
                "held_out_auc": round(auc_f, 6), "speedup_quantized": round(dt_f / dt_q, 3),
                "auc_delta_quantized": round(auc_q - auc_f, 6), "tuning": tuned_f,
                "hist_roofline": {"kernel": "k_hist_fx", "achieved": round(hist_f_gbps, 1), "peak": HBM_PEAK_GBPS,
                                  "unit": "GB/s", "frac": round(hist_f_gbps / HBM_PEAK_GBPS, 4),
                                  "avg_launch_us": round(prof_f["ms"] * 1e3 / prof_f["launches"], 2),
                                  "bytes_basis": f"built-node rows x ({X.shape[1]} + 8) + 8 B per built cell"}}

    # Equal AUC at the full fit length (untimed; the launch shapes stay as tuned).
    auc_full = None
    if a.auc_iters and not a.skip_fp32:
        res = {}
        for quantized, tuned in ((True, tuned_q), (False, fp32["tuning"])):
            ctx.configure(B.FitConfig(grower="hip_depthwise", objective="logloss", quantized=quantized,
                                      use_graph=True, max_depth=8, learning_rate=0.1, lambda_l2=1.0))
            if tuned and tuned.get("shapes"):
                ctx.set_tuning(**tuned["shapes"])
            ctx.set_scores(np.full(a.rows, init, np.float32))
            t0 = time.perf_counter()
            ctx.train(a.auc_iters)
            res["quantized" if quantized else "fp32"] = {
                "held_out_auc": round(held_out_auc(B, ctx.trees(), nb, init, bins_h, yh), 6),
                "fit_s": round(time.perf_counter() - t0, 3)}
        auc_full = {"iterations": a.auc_iters, **res,
                    "auc_delta": round(res["quantized"]["held_out_auc"] - res["fp32"]["held_out_auc"], 6)}
        auc_full["within_0.001"] = abs(auc_full["auc_delta"]) <= 0.001
        log(f"equal-AUC check at {a.auc_iters} iterations: {auc_full}")

    bins = ctx.download_bins()
    herr = None
    if not a.skip_hist_error:
        if auc_full:   # the FP32 fit's final scores: gradients with a realistic spread
            scores, what = ctx.get_scores(), f"after {a.auc_iters} FP32 iterations"
        else:
            scores, what = np.full(a.rows, init, np.float32), "iteration 0 (score = init score)"
        herr = hist_error(B, ctx, scores, y, nb, bins, what)
        log(f"root histogram error vs FP64: {herr}")
    ctx.close()

    cpu = None
    if not a.skip_cpu:
        cpu = cpu_baseline(bins, nb, y, a.cpu_iters)
        log(f"cpu baseline: {cpu['value']} s extrapolated on {cpu['cores']} threads")

    # HBM traffic per histogram launch from the counter runs, against the algorithmic
    # bytes of the same launches (built-node and scanned bases)
    traffic = traffic_ratio = traffic_scanned_ratio = None
    if TRAFFIC.exists() and a.rows == 16_000_000:
        t = json.loads(TRAFFIC.read_text())
        traffic = t["traffic_bytes_per_launch"]
        traffic_ratio = round(traffic / t["built_bytes_per_launch"], 3)
        traffic_scanned_ratio = round(traffic / t["scanned_bytes_per_launch"], 3)

    out = {
        "metric": METRIC,
        "value": round(dt_q / a.steps * FULL_ITERS, 4),
        "unit": "s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt_q / a.steps * 1e3, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int64",
        "data": "synthetic",
        "config": {
            "workload": "tall binary 16M x 28, logistic, hip_depthwise depth 8, 255 bins, lr 0.1, "
                        "quantized int8 g / int16 h (stochastic rounding), one hipGraph per iteration",
            "rows": a.rows, "features": 28, "held_out_rows": a.held_out,
            "extrapolated_to_500": a.steps != FULL_ITERS,
            "held_out_auc": round(auc_q, 6),
            "fp32": fp32,
            "hist_GBps": round(hist_gbps, 1),
            "auc_full_fit": auc_full,
            "hist_error_vs_fp64": herr,
            "tuning": tuned_q,
            "rows_per_s": round(a.rows * a.steps / dt_q, 1),
        },
        "roofline": {
            "bound": "hbm", "kernel": "k_hist_q / k_hist_q3",
            "achieved": round(hist_gbps, 1), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
            "frac": round(hist_gbps / HBM_PEAK_GBPS, 4),
            "bytes_per_launch": round(per_launch_bytes), "avg_launch_us": round(per_launch_us, 2),
            "launches_timed": prof["launches"],
            "bytes_basis": "built-node rows x (28 + 4) + 8 B per built cell",
            "scanned_GBps": round(scanned_gbps, 1),
            "scanned_bytes_per_launch": round(prof["scanned_bytes"] / prof["launches"]),
            "copy_GBps": round(copy_gbps, 1),
            "wide_copy_GBps": round(wide["copy"][0], 1),
            "wide_read_GBps": round(wide["read"][0], 1),
            "traffic": traffic,
            "traffic_over_algorithmic": traffic_ratio,
            "traffic_over_scanned": traffic_scanned_ratio,
            "traffic_source": "profiles/hist_quantized_16Mx28_counters.md" if traffic else None,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
