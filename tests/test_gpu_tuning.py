"""Launch shapes never change results, and the autotuner leaves the model state as it
found it (scores, trees, iteration counter)."""
import numpy as np
import pytest

from test_gpu_depthwise import ctx_for, tall_data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tall():
    return tall_data(50000, seed_cfg=5)


def fit(hip, orc, bins, nb, y, quantized, tuning=None, autotune=False, iters=4):
    ctx = ctx_for(hip, bins, nb, quantized=quantized, max_depth=6, learning_rate=0.1)
    ctx.upload_labels(y)
    ctx.set_scores(np.full(y.size, orc.init_score(1, y), np.float32))
    if tuning:
        ctx.set_tuning(**tuning)
    if autotune:
        ctx.train(1)
        before = ctx.get_scores()
        chosen, ms = ctx.autotune(2)
        assert ms > 0
        assert np.array_equal(ctx.get_scores().view(np.uint32), before.view(np.uint32))
        assert ctx.tuning() == chosen
        ctx.train(iters - 1)
    else:
        ctx.train(iters)
    out = ctx.trees(), ctx.get_scores()
    ctx.close()
    return out


SHAPES = [
    dict(scatter_lds_bytes=4096, find_blocks=1, hist_threads=256, part_grid=64, hist_layout=0, pair_levels=1),
    dict(scatter_lds_bytes=49152, find_blocks=3, hist_threads=512, part_grid=4096, hist_layout=1, pair_levels=1),
    dict(scatter_lds_bytes=147456, find_blocks=8, hist_threads=1024, part_grid=2048, hist_layout=0,
         pair_levels=0),
    dict(scatter_lds_bytes=20480, find_blocks=8, hist_threads=1024, part_grid=2048, hist_layout=1, pair_levels=0),
    dict(scatter_lds_bytes=32768, find_blocks=4, hist_threads=512, part_grid=1024, hist_layout=2, pair_levels=1,
         hist_group=4),
    dict(scatter_lds_bytes=32768, find_blocks=4, hist_threads=1024, part_grid=1024, hist_layout=2, pair_levels=0,
         hist_group=3),
    dict(scatter_lds_bytes=20480, find_blocks=2, hist_threads=256, part_grid=2048, hist_layout=0, pair_levels=1,
         hist_group=1),
    # rotated row-in-lane histogram with transposed unit partials, paired and unpaired
    dict(scatter_lds_bytes=32768, find_blocks=4, hist_threads=512, part_grid=4096, hist_layout=3, pair_levels=1,
         hist_group=1),
    dict(scatter_lds_bytes=32768, find_blocks=4, hist_threads=1024, part_grid=4096, hist_layout=3, pair_levels=0,
         hist_group=1),
    dict(scatter_lds_bytes=32768, find_blocks=4, hist_threads=256, part_grid=4096, hist_layout=3, pair_levels=1,
         hist_group=2),
    # row-dword lanes on the routed levels too
    dict(scatter_lds_bytes=32768, find_blocks=8, hist_threads=512, part_grid=4096, hist_layout=4, pair_levels=1,
         hist_group=1),
    dict(scatter_lds_bytes=32768, find_blocks=8, hist_threads=1024, part_grid=4096, hist_layout=4, pair_levels=1,
         hist_group=2),
    # rotated row-dword lanes, feature-minor cells, transposed partials
    dict(scatter_lds_bytes=32768, find_blocks=8, hist_threads=512, part_grid=4096, hist_layout=5, pair_levels=1,
         hist_group=1),
    dict(scatter_lds_bytes=32768, find_blocks=8, hist_threads=256, part_grid=4096, hist_layout=5, pair_levels=0,
         hist_group=1),
    dict(scatter_lds_bytes=32768, find_blocks=8, hist_threads=1024, part_grid=4096, hist_layout=5, pair_levels=1,
         hist_group=2),
]


@pytest.mark.parametrize("quantized", [True, False])
def test_launch_shapes_do_not_change_results(hip, orc, tall, quantized):
    bins, nb, X, y = tall
    ref_trees, ref_scores = fit(hip, orc, bins, nb, y, quantized)
    for shape in SHAPES:
        trees, scores = fit(hip, orc, bins, nb, y, quantized, tuning=shape)
        assert len(trees) == len(ref_trees)
        for a, b in zip(trees, ref_trees):
            assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
        assert np.array_equal(scores.view(np.uint32), ref_scores.view(np.uint32))


def test_autotune_keeps_model_state(hip, orc, tall):
    bins, nb, X, y = tall
    ref_trees, ref_scores = fit(hip, orc, bins, nb, y, True)
    trees, scores = fit(hip, orc, bins, nb, y, True, autotune=True)
    assert len(trees) == len(ref_trees)
    for a, b in zip(trees, ref_trees):
        assert np.array_equal(a.view(np.uint8), b.view(np.uint8))
    assert np.array_equal(scores.view(np.uint32), ref_scores.view(np.uint32))


def test_tuning_validation(hip, tall):
    bins, nb, X, y = tall
    ctx = ctx_for(hip, bins[:1000], nb, quantized=True, max_depth=3)
    with pytest.raises(hip.ConfigError):
        ctx.set_tuning(hist_threads=300)
    with pytest.raises(hip.ConfigError):
        ctx.set_tuning(find_blocks=9)
    with pytest.raises(hip.ConfigError):
        ctx.set_tuning(pair_levels=2)
    with pytest.raises(hip.ConfigError):
        ctx.set_tuning(hist_layout=6)
    with pytest.raises(hip.ConfigError):
        ctx.set_tuning(hist_group=5)
    ctx.close()
