"""The C ABI library and the host-side package, without a GPU.

The library must load and export every function include/bonsai_hip.h declares; the
host mirror (binning, synthetic workloads, configuration checks) must agree with the
oracle.  No compute call reaches a device here.
"""
import ctypes as C
import re
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
HEADER = ROOT / "include" / "bonsai_hip.h"


def header_functions():
    text = re.sub(r"/\*.*?\*/", "", HEADER.read_text(), flags=re.S)
    return sorted(set(re.findall(r"\b(bh_[a-z_0-9]+)\s*\(", text)))


@pytest.fixture(scope="module")
def libpath():
    import bonsai_amd.build as b
    if not b.LIB.exists():
        b.build(verbose=False)
    return b.LIB


def test_library_exports_every_header_function(libpath):
    names = header_functions()
    assert len(names) >= 30
    lib = C.CDLL(str(libpath))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    out = subprocess.run(["nm", "-D", "--defined-only", str(libpath)], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (bh_[a-z_0-9]+)\b", out))
    assert set(names) <= exported


def test_binding_covers_the_header(libpath):
    import bonsai_amd
    bonsai_amd.lib()        # sets argtypes for every bound symbol, raises if one is missing
    bound = set(bonsai_amd._SIGS)
    assert set(header_functions()) <= bound


def test_null_context_is_a_config_error(libpath):
    import bonsai_amd
    L = bonsai_amd.lib()
    assert L.bh_destroy(None) == 0 or L.bh_destroy(None) > 0
    rc = L.bh_upload_labels(None, None)
    assert rc > 0
    assert b"context" in L.bh_last_error() or b"bins" in L.bh_last_error()
    # the ingest arms refuse missing inputs before any device call
    assert L.bh_upload_feature_columns(None, None, 10, 3, None, None) > 0
    assert b"columns" in L.bh_last_error()
    assert L.bh_upload_features_device(None, None, 10, 3, None, None) > 0
    assert L.bh_gather_rows(0, None, 10, 3, None, 0, None) > 0
    assert b"null" in L.bh_last_error()


def test_fit_config_validation():
    import bonsai_amd
    with pytest.raises(bonsai_amd.ConfigError):
        bonsai_amd.FitConfig(grower="cuda_depthwise").to_c()
    with pytest.raises(bonsai_amd.ConfigError):
        bonsai_amd.FitConfig(objective="quantile").to_c()
    assert bonsai_amd.FitConfig(objective="poisson").to_c().objective == 2
    c = bonsai_amd.FitConfig(grower="hip_levelwise", quantized=True).to_c()
    assert c.grower == 2 and c.quantized == 1


def test_host_binning_matches_oracle(orc):
    from bonsai_amd import binning
    rng = np.random.default_rng(0)
    for n, max_bin in ((50, 255), (5000, 255), (5000, 16), (20000, 64)):
        col = np.concatenate([rng.normal(0, 1, n // 2), np.zeros(n // 4), rng.integers(0, 5, n - n // 2 - n // 4)])
        col = col.astype(np.float32)
        col[rng.random(n) < 0.03] = np.nan
        a = orc.fit_cuts(col, max_bin)
        b = binning.create_cuts(col[~np.isnan(col)], max_bin)
        assert np.array_equal(a, b)
        x = np.concatenate([col, np.array([np.inf, -np.inf, -0.0, 0.0], np.float32)])
        assert np.array_equal(orc.transform(a, x), binning.transform(x[:, None], [a])[:, 0])


def test_synthetic_workloads_shapes_and_rates():
    from bonsai_amd import synth
    X = synth.tall_features(20000, config_id=0)
    assert X.shape == (20000, 28) and X.dtype == np.float32
    nan = np.isnan(X)
    assert not nan[:, :18].any()
    assert abs(nan[:, 18:].mean() - 0.05) < 0.01
    assert np.all(X[:, 10:18] >= -1) and np.all(X[:, 10:18] < 1)
    v = X[:, 24:][~nan[:, 24:]]
    assert np.array_equal(np.unique(v), np.arange(16, dtype=np.float32))
    y = synth.binary_label(X)
    assert set(np.unique(y)) == {0.0, 1.0}
    assert np.array_equal(synth.tall_features(1000, config_id=1), synth.tall_features(1000, config_id=1),
                          equal_nan=True)
    Xw, yw = synth.wide_bins(2000, 100)
    assert Xw.shape == (2000, 100)
    sparse = Xw[:, :90]
    assert abs((sparse == 0).mean() - 0.95) < 0.02
    assert Xw.max() <= 63


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
