"""Model file v7 (bonsai_amd.model_io) against the reference's committed fixtures.

The reference's own contract (tests/unit/test_model_io.cpp:270-297): a loaded fixture
saves back to the same bytes.  The fixtures (tests/data/fixture_v7_regressor.bonsai and
fixture_v7_softmax.bonsai in the reference, copied as data into tests/golden/) are real
ensembles with real cuts; their trees also pin the threshold -> bin-space conversion
that feeds the device predictor: a threshold walk on raw values (x <= threshold goes
left, NaN follows default_left) must reach the same leaf as the oracle's bin-space walk
on the transformed rows.
"""
import pathlib

import numpy as np
import pytest

from bonsai_amd import binning, model_io
from oracle_lib import load_oracle

GOLDEN = pathlib.Path(__file__).parent / "golden"
FIXTURES = ["fixture_v7_regressor.bonsai", "fixture_v7_softmax.bonsai"]


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_saves_back_to_the_same_bytes(name, tmp_path):
    raw = (GOLDEN / name).read_bytes()
    m = model_io.load(GOLDEN / name)
    out = tmp_path / "resaved.bonsai"
    model_io.save(m, out)
    assert out.read_bytes() == raw


@pytest.mark.parametrize("name", FIXTURES)
def test_codec_round_trip_and_encoding_choices(name):
    raw = (GOLDEN / name).read_bytes()
    doc = model_io.unpack(raw)
    assert model_io.pack(doc) == raw
    # the +inf sentinel is the only float64 in the file; every other float is float32
    assert raw.count(b"\xcb\x7f\xf0\x00\x00\x00\x00\x00\x00") == len(doc["bin_mappers"])


def test_fixture_contents():
    m = model_io.load(GOLDEN / "fixture_v7_regressor.bonsai")
    assert m.n_classes == 1 and len(m.trees) == 5 and len(m.cuts) == 6
    assert m.config["dispatch"]["objective_name"] == "mse"
    assert np.float32(m.learning_rate) == np.float32(0.1)
    for c in m.cuts:  # FLT_MAX closer and +inf sentinel close every cut table
        assert c[-1] == np.inf and c[-2] == np.finfo(np.float32).max and np.all(np.diff(c) > 0)
    s = model_io.load(GOLDEN / "fixture_v7_softmax.bonsai")
    assert s.n_classes == 3 and len(s.trees) == 15 and s.init_score is None


def test_bad_magic_and_version_raise(tmp_path):
    m = model_io.load(GOLDEN / "fixture_v7_regressor.bonsai")
    doc = m.document()
    doc["magic"] = "bonsai02"
    with pytest.raises(model_io.ModelFormatError):
        model_io.loads(model_io.pack(doc))
    doc["magic"] = model_io.MAGIC
    doc["version"] = 6
    with pytest.raises(model_io.ModelFormatError):
        model_io.loads(model_io.pack(doc))
    with pytest.raises(model_io.ModelFormatError):
        model_io.loads(b"not-a-bonsai-model")


def threshold_leaf(tree: model_io.DenseTree, x: np.ndarray) -> float:
    i = 0
    while tree.feature_id[i] != model_io.LEAF_FLAG:
        v = x[tree.feature_id[i]]
        go_left = bool(tree.default_left[i]) if np.isnan(v) else v <= tree.value[i]
        i = tree.left[i] if go_left else tree.right[i]
    return float(tree.value[i])


def probe_rows(m: model_io.Model, n: int, seed: int) -> np.ndarray:
    """Random rows plus exact cut values, values just above cuts, and NaNs."""
    rng = np.random.default_rng(seed)
    F = len(m.cuts)
    X = np.empty((n, F), np.float32)
    for f, c in enumerate(m.cuts):
        inner = c[:-2]
        pick = rng.integers(0, len(inner), n)
        kind = rng.integers(0, 4, n)
        X[:, f] = np.where(kind == 0, inner[pick],
                           np.where(kind == 1, np.nextafter(inner[pick], np.float32(np.inf)),
                                    np.where(kind == 2, np.float32(np.nan),
                                             rng.uniform(inner[0] - 1, inner[-1] + 1, n).astype(np.float32))))
    return X


@pytest.mark.parametrize("name", FIXTURES)
def test_bin_space_reaches_the_threshold_walk_leaf(name):
    orc = load_oracle()
    m = model_io.load(GOLDEN / name)
    X = probe_rows(m, 400, seed=7)
    bins = binning.transform(X, m.cuts)
    nb = m.n_bins()
    for cls in range(m.n_classes):
        trees = [t for i, t in enumerate(m.trees) if i % m.n_classes == cls]
        for tree, nodes in zip(trees, m.bin_space(cls)):
            got = orc.predict([nodes], bins, nb, 1.0, 0.0)  # init 0 + 1 * leaf = the leaf
            want = np.array([threshold_leaf(tree, x) for x in X], np.float32)
            assert np.array_equal(got, want)


def test_own_fit_saves_and_loads_back(tmp_path):
    """A bin-space ensemble written as v7 and read back gives the same nodes."""
    orc = load_oracle()
    m = model_io.load(GOLDEN / "fixture_v7_regressor.bonsai")
    trees = m.bin_space()
    cfg = m.config
    mine = model_io.from_fit(trees, m.cuts, cfg, init_score=0.25)
    path = tmp_path / "mine.bonsai"
    model_io.save(mine, path)
    back = model_io.load(path)
    for a, b in zip(trees, back.bin_space()):
        for k in ("feature", "bin", "default_left", "left", "right", "value", "cover"):
            leafmask = a["feature"] < 0
            if k in ("bin", "default_left", "left", "right"):
                assert np.array_equal(a[k][~leafmask], b[k][~leafmask])
            else:
                assert np.array_equal(a[k], b[k])
    X = probe_rows(m, 300, seed=3)
    bins = binning.transform(X, m.cuts)
    lr = np.float32(m.learning_rate)
    assert np.array_equal(orc.predict(trees, bins, m.n_bins(), lr, 0.25),
                          orc.predict(back.bin_space(), bins, back.n_bins(), lr, float(back.init_score)))
