"""bh_histogram_noalloc: the histogram of device-listed rows into a device output with
no allocation (work list and unit partials in the library's fixed per-GPU scratch).

Checked through the C ABI against bh_histogram (the allocating entry point) on the same
rows, bit for bit, in both precisions, on two calls in a row, the second with fewer
rows (the scratch keeps the first call's contents, so a stale partial or count would
show), and the capacity refusal.  Device buffers come from the HIP runtime through
ctypes (plumbing only).
"""
import ctypes as C

import numpy as np
import pytest

from test_gpu_depthwise import grad_of, tall_data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tall():
    return tall_data(120_000, seed_cfg=8)


def hip_runtime():
    hip_lib = __import__("bonsai_amd").lib()          # loads the HIP runtime with the library
    del hip_lib
    try:
        rt = C.CDLL("libamdhip64.so.7")
    except OSError:
        rt = C.CDLL("libamdhip64.so")
    rt.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    rt.hipFree.argtypes = [C.c_void_p]
    rt.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    for f in (rt.hipMalloc, rt.hipFree, rt.hipMemcpy):
        f.restype = C.c_int
    return rt


def noalloc(hip, ctx, rows, dtype):
    """bh_histogram_noalloc on device copies of rows; the output read back."""
    rt = hip_runtime()
    H2D, D2H = 1, 2
    F, SB = ctx.n_features, ctx.bin_stride
    out = np.full((F, SB, 2), -7, dtype)
    d_rows, d_out = C.c_void_p(), C.c_void_p()
    assert rt.hipMalloc(C.byref(d_rows), max(4, rows.nbytes)) == 0
    assert rt.hipMalloc(C.byref(d_out), out.nbytes) == 0
    try:
        r32 = np.ascontiguousarray(rows, np.uint32)
        if r32.size:
            assert rt.hipMemcpy(d_rows, r32.ctypes.data, r32.nbytes, H2D) == 0
        assert rt.hipMemcpy(d_out, out.ctypes.data, out.nbytes, H2D) == 0
        rc = hip.lib().bh_histogram_noalloc(ctx._h, d_rows, rows.size, d_out, None)
        if rc != 0:
            msg = hip.lib().bh_last_error().decode()
            raise hip.ConfigError(msg) if rc > 0 else hip.DeviceError(msg)
        hip.synchronize(ctx)
        assert rt.hipMemcpy(out.ctypes.data, d_out, out.nbytes, D2H) == 0
    finally:
        rt.hipFree(d_rows)
        rt.hipFree(d_out)
    return out


@pytest.mark.parametrize("quantized", [True, False])
def test_noalloc_histogram_equals_bh_histogram(hip, orc, tall, quantized):
    bins, nb, X, y = tall
    g, h = grad_of(orc, y, seed=11)
    ctx = hip.HipContext(0)
    ctx.upload_bins(bins, nb)
    ctx.configure(hip.FitConfig(quantized=quantized, max_depth=6))
    ctx.set_gradients(g, h)
    rng = np.random.default_rng(3)
    dtype = np.int64 if quantized else np.float32
    # two calls in a row, the second smaller; then one single-unit call
    for frac in (0.8, 0.1, 0.02):
        rows = np.flatnonzero(rng.random(y.size) < frac).astype(np.uint32)
        got = noalloc(hip, ctx, rows, dtype)
        want = ctx.histogram(rows)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), frac
    empty = noalloc(hip, ctx, np.zeros(0, np.uint32), dtype)
    assert not np.any(empty)
    ctx.close()


def test_noalloc_refuses_more_rows_than_its_scratch(hip):
    """2,000 features x 256 bins: a quantized unit partial is 6 MB, so the 32 MiB scratch
    holds five 8,192-row units (40,960 rows)."""
    rng = np.random.default_rng(4)
    n, F = 50_000, 2000
    bins = rng.integers(0, 255, (n, F)).astype(np.uint8)
    nb = np.full(F, 256, np.uint16)
    ctx = hip.HipContext(0)
    ctx.upload_bins(bins, nb)
    ctx.configure(hip.FitConfig(quantized=True, max_depth=4))
    ctx.set_gradients(rng.normal(0, 1, n).astype(np.float32), np.ones(n, np.float32))
    rows = np.arange(45_000, dtype=np.uint32)
    with pytest.raises(hip.ConfigError, match="scratch"):
        noalloc(hip, ctx, rows, np.int64)
    rows = np.arange(0, 40_000, dtype=np.uint32)
    assert np.array_equal(noalloc(hip, ctx, rows, np.int64), ctx.histogram(rows))
    ctx.close()
