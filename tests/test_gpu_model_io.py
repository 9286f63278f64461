"""The reference's committed v7 ensembles (tests/golden/fixture_v7_*.bonsai, real trees
and cuts) on the device: predict bit-equal to the oracle's bin-space walk per class, and
TreeSHAP within 1e-5 of the oracle's Algorithm 2 with additivity, on rows binned through
the fixtures' own cuts (random values, exact cut values and NaNs)."""
import pathlib

import numpy as np
import pytest

from bonsai_amd import binning, model_io
from test_model_io import probe_rows

pytestmark = pytest.mark.gpu
GOLDEN = pathlib.Path(__file__).parent / "golden"


@pytest.mark.parametrize("name", ["fixture_v7_regressor.bonsai", "fixture_v7_softmax.bonsai"])
def test_fixture_predict_and_shap(hip, orc, name):
    m = model_io.load(GOLDEN / name)
    X = probe_rows(m, 2000, seed=11)
    bins = binning.transform(X, m.cuts)
    nb = m.n_bins()
    lr = float(np.float32(m.learning_rate))
    inits = m.init_scores if m.init_scores is not None else [m.init_score]
    for cls in range(m.n_classes):
        trees = m.bin_space(cls)
        init = float(np.float32(inits[cls]))
        p = hip.Predictor(trees, nb, lr, init)
        dev = p.predict(bins)
        ref = orc.predict(trees, bins, nb, lr, init)
        assert np.array_equal(dev.view(np.uint32), ref.view(np.uint32))
        rows = bins[:300]
        phi = p.shap(rows)
        depth = max(int(t.depth) for i, t in enumerate(m.trees) if i % m.n_classes == cls)
        want = orc.pred_contribs(trees, rows, nb, lr, init, depth)
        scale = np.abs(want).sum(axis=1, keepdims=True)
        assert np.max(np.abs(phi - want) / scale) <= 1e-5
        margin = dev[:300].astype(np.float64)
        assert np.max(np.abs(phi.sum(axis=1) - margin) / np.maximum(1.0, np.abs(margin))) <= 1e-5
        p.close()
