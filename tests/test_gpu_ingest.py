"""Device bin transform (the reference's transform_bin, src/cuda/detail/ingest_kernels.cuh:27-50)
is bit-identical to the oracle's BinMapper::transform (src/bin_mapper.cpp:284-291) on the
same cuts: lower_bound, NaN -> last bin, values equal to a cut land in that cut's bin,
+inf in the sentinel bin (test_bin_mapper.cpp:338-387 invariants)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def oracle_bins(orc, X, cuts):
    out = np.empty(X.shape, np.uint8)
    for j, c in enumerate(cuts):
        out[:, j] = orc.transform(c, X[:, j])
    return out


@pytest.mark.parametrize("n,f", [(1, 1), (1000, 3), (70001, 28), (5000, 130)])
def test_device_transform_bit_exact(hip, orc, n, f):
    from bonsai_amd import binning
    rng = np.random.default_rng(n + f)
    X = rng.standard_normal((n, f)).astype(np.float32)
    X[rng.random((n, f)) < 0.05] = np.nan
    if n > 10:
        X[0, :] = np.inf
        X[1, :] = -np.inf
        X[2, :] = -0.0
    max_bin = 255 if f < 100 else 64
    cuts = binning.fit_cuts(X, max_bin=max_bin, n_samples=n)
    # exact cut values land in their own bin (right-inclusive)
    if n > 10:
        for j in range(f):
            X[3, j] = cuts[j][len(cuts[j]) // 2]
    ctx = hip.HipContext(0)
    ctx.upload_features(X, cuts)
    dev = ctx.download_bins()
    ctx.close()
    ref = oracle_bins(orc, X, cuts)
    assert np.array_equal(dev, ref)
    assert np.array_equal(dev, binning.transform(X, cuts))


def test_device_transform_many_cuts_global_path(hip, orc):
    """2,000 features x 64 bins: the cut tables exceed LDS and are read from global memory."""
    from bonsai_amd import binning
    rng = np.random.default_rng(5)
    X = rng.integers(0, 64, (3000, 2000)).astype(np.float32)
    cuts = binning.fit_cuts(X, max_bin=255, n_samples=3000)
    ctx = hip.HipContext(0)
    ctx.upload_features(X, cuts)
    dev = ctx.download_bins()
    ctx.close()
    assert np.array_equal(dev, oracle_bins(orc, X, cuts))
