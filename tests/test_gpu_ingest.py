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


def ingest_case(n, f, seed):
    from bonsai_amd import binning
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, f)).astype(np.float32)
    X[rng.random((n, f)) < 0.05] = np.nan
    X[0, :] = np.inf
    X[1, :] = -np.inf
    cuts = binning.fit_cuts(X, max_bin=255 if f < 100 else 64, n_samples=n)
    for j in range(f):
        X[2, j] = cuts[j][len(cuts[j]) // 2]
    return X, cuts


@pytest.mark.parametrize("n,f", [(70001, 28), (3000, 130), (20, 5)])
def test_column_ingest_bit_exact(hip, orc, n, f):
    """Column-batch ingest (cuda_ingest(ColumnBatch), four columns per staged group, the
    last group short when f % 4 != 0) gives the row-major transform's plane."""
    X, cuts = ingest_case(n, f, seed=n + f)
    ctx = hip.HipContext(0)
    ctx.upload_feature_columns([X[:, j].copy() for j in range(f)], cuts)
    dev = ctx.download_bins()
    ctx.close()
    assert np.array_equal(dev, oracle_bins(orc, X, cuts))


def test_device_resident_ingest_and_sample_gather(hip, orc):
    """A device-resident matrix is binned in place (cuda_ingest_device), its cut-fitting
    sample is gathered on the device (cuda_gather_rows) and gives the host fit's cuts."""
    from bonsai_amd import binning
    X, _ = ingest_case(50000, 28, seed=3)
    with hip.DeviceMatrix(X) as Xd:
        rows = binning.sample_rows(X.shape[0], 7000, seed=11)
        got = hip.gather_rows(Xd, rows)
        assert np.array_equal(got.view(np.uint32), X[rows].view(np.uint32))
        whole = hip.gather_rows(Xd, None)
        assert np.array_equal(whole.view(np.uint32), X.view(np.uint32))
        cuts_dev = binning.fit_cuts_device(Xd, max_bin=255, n_samples=7000, seed=11)
        cuts_host = binning.fit_cuts(X, max_bin=255, n_samples=7000, seed=11)
        assert all(np.array_equal(a.view(np.uint32), b.view(np.uint32)) for a, b in zip(cuts_dev, cuts_host))
        ctx = hip.HipContext(0)
        ctx.upload_features_device(Xd, cuts_dev)
        dev = ctx.download_bins()
        ctx.close()
    assert np.array_equal(dev, oracle_bins(orc, X, cuts_dev))


def test_ingest_refusals(hip):
    X, cuts = ingest_case(100, 3, seed=1)
    with hip.DeviceMatrix(X) as Xd:
        with pytest.raises(hip.ConfigError):
            hip.gather_rows(Xd, np.array([0, 100], np.uint32))       # row out of range
    ctx = hip.HipContext(0)
    with pytest.raises(hip.ConfigError):
        ctx.upload_feature_columns([X[:, 0].copy(), X[:50, 1].copy(), X[:, 2].copy()], cuts)
    # a host pointer is not device memory of the context's device
    with pytest.raises(hip.ConfigError, match="device memory"):
        ctx.upload_features_device(hip.DeviceMatrix.wrap(X.ctypes.data, 100, 3), cuts)
    ctx.close()
