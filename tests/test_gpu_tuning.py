"""Launch tuning (hd_tuning_set) and the autotuner.

Bar: bit-exact. trace_tile, top_pairs and small_level give the same roots, pool words, color pool
and pixels as the default set on plain edits and a primary + shadow frame. batch_pass_edits splits
batched passes, which renumbers nodes (the batched path's own bar): the DAG up to renumbering and
the pixels of a frame without LOD are unchanged (a LOD cut-off shows the color stored at its cube's
corner voxel, which may be empty, and empty voxels' colors depend on the pass split). The autotuner's search (small terrain, few edits) returns a valid set
and applies it."""
import json

import numpy as np
import pytest

from test_gpu_async import assert_same_device_pools
from test_oracle_pins import small_cfg
from vkhashdag_amd import workloads as W

pytestmark = pytest.mark.gpu


def _capi():
    from vkhashdag_amd import capi
    return capi


@pytest.fixture
def default_tuning():
    capi = _capi()
    old = capi.get_tuning()
    yield
    capi.set_tuning(**old)


DEFAULT = dict(trace_tile=16, top_pairs=4096, small_level=1024, batch_pass_edits=0)


def _run(capi, res, h, tuning, batched):
    capi.set_tuning(**tuning)
    cfg = small_cfg(res, 9, 9, 2, 3)
    p = capi.DevicePool(cfg, heights=h)
    p.edit(W.terrain_edit(res, seed=4, skin=6))
    p.edit(W.terrain_sphere_edits(24, res, h, 2, 24, seed=5))
    if batched:
        p.edit(W.mixed_batch(160, res, h, 2, 20, seed=6), capi.MODE_BATCHED)
    vb = W.orbit_view((res / 2, res / 2, res / 3), res * 0.9, 3, 72, 40, res=res, lod=not batched)
    img = p.render(vb, 72, 40, shadow=True)
    return cfg, p, img


@pytest.mark.parametrize("tuning", [
    dict(trace_tile=8, top_pairs=64, small_level=64),
    dict(trace_tile=16, top_pairs=1024, small_level=256),
    dict(trace_tile=8, top_pairs=4096, small_level=1024),
])
def test_launch_tunings_give_the_same_words_and_pixels(default_tuning, tuning):
    capi = _capi()
    res = 128
    h = W.value_noise_heightmap(res, 4, 4, res // 6, res * 3 // 4)
    _, a, img_a = _run(capi, res, h, DEFAULT, False)
    _, b, img_b = _run(capi, res, h, dict(DEFAULT, **tuning), False)
    assert_same_device_pools(a, b, str(tuning))
    assert np.array_equal(img_a, img_b)


@pytest.mark.parametrize("q", [1, 7, 64])
def test_batch_pass_size_keeps_the_dag_and_pixels(default_tuning, q):
    import _oracle as O
    capi = _capi()
    res = 128
    h = W.value_noise_heightmap(res, 4, 4, res // 6, res * 3 // 4)
    cfg, a, img_a = _run(capi, res, h, DEFAULT, True)
    _, b, img_b = _run(capi, res, h, dict(DEFAULT, trace_tile=8, top_pairs=256, small_level=128,
                                          batch_pass_edits=q), True)
    assert a.stats().overflow == 0 and b.stats().overflow == 0
    L = cfg.node_levels
    ca = O.canonical_words(a.full_words(), L, np.array([a.root], np.uint32))
    cb = O.canonical_words(b.full_words(), L, np.array([b.root], np.uint32))
    assert np.array_equal(ca, cb)
    assert np.array_equal(img_a, img_b)


def test_autotuner_small_search(default_tuning, tmp_path):
    from vkhashdag_amd import autotune
    capi = _capi()
    doc = autotune.tune(res=256, plain_edits=8, batch_edits=64, width=128, height=96, repeats=1,
                        candidates={"top_pairs": [1024, 4096], "small_level": [256, 1024],
                                    "batch_pass_edits": [0, 16]})
    t = doc["tuning"]
    assert t["trace_tile"] in (8, 16) and t["top_pairs"] in (1024, 4096) and t["small_level"] in (256, 1024)
    assert t["batch_pass_edits"] in (0, 16)
    assert capi.get_tuning() == t
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps(doc))
    capi.set_tuning(trace_tile=16, top_pairs=4096, small_level=1024, batch_pass_edits=0)
    assert autotune.load(str(path)) == t and capi.get_tuning() == t
