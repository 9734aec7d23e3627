"""Device edit + traversal parity (HIP path through the C-ABI) against the CPU oracle.

Bar: bit-exact. After every plain edit the device pool's live words, per-bucket word counters,
roots, filled chain, color octree nodes and color leaf words equal the oracle's (which equals the
reference's single-threaded Edit, see test_oracle_pins.py). Rendering: RGBA8 pixels and per-pixel
hit records (voxel, size, hit position, normal, iterations, shadow) equal the oracle's
transliteration of trace.frag."""
import os
import sys

import numpy as np
import pytest

import _oracle as O
from test_oracle_pins import CASES, random_edits, small_cfg
from vkhashdag_amd import workloads as W

pytestmark = pytest.mark.gpu


def _capi():
    from vkhashdag_amd import capi
    return capi


def assert_same_pool(dev, ora, where=""):
    # both report the same bucket-full fallbacks (the reference's silent drop, NodePool.hpp:195,210)
    assert (dev.stats().overflow & 1) == (1 if ora.overflows() else 0), where
    assert dev.root == ora.root, (where, hex(dev.root), hex(ora.root))
    assert dev.color_root == ora.color_root, (where, hex(dev.color_root), hex(ora.color_root))
    bd, wd = dev.live_words()
    bo, wo = ora.live_words()
    assert np.array_equal(bd, bo), where
    assert np.array_equal(wd, wo), where
    assert np.array_equal(dev.color_nodes(), ora.color_nodes()), where
    assert np.array_equal(dev.color_leaves(), ora.color_leaves()), where


@pytest.mark.parametrize("case", CASES, ids=[f"res{c['res']}_b{c['bits']}_s{c['seed']}" for c in CASES])
def test_plain_edits_bit_exact(case):
    print(f"case {case}", flush=True)
    capi = _capi()
    cfg = small_cfg(case["res"], case["bits"], case["wbpp"], case["pbpb"], case["depth"])
    e = random_edits(case["n"], case["res"], case["seed"], case["colored"])
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    for i in range(len(e)):
        dev.edit(e[i:i + 1])
        ora.edit(e[i:i + 1])
        assert_same_pool(dev, ora, f"edit {i} kind {int(e[i]['kind'])}")
    assert np.array_equal(dev.filled(), ora.filled())


@pytest.mark.parametrize("bits", [2, 0])
def test_plain_walkers_of_every_size_bit_exact(bits):
    # levels of more than 64 entries leave the one-workgroup kernel (small_level 64) for the bucket
    # walkers; with 4 buckets per level (bits 2) the large edits give 16..31 new leaves per bucket (a
    # whole wavefront walks them), with 1 bucket (bits 0) up to 93 (the chain of minimum searches),
    # and the small edits 1..4 (one thread): every placement must stay the reference's, word for word
    capi = _capi()
    cfg = small_cfg(64, bits, 9, 2, 3)
    e = random_edits(24, 64, 9, True)
    old = capi.set_tuning(small_level=64)
    try:
        dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
        for i in range(len(e)):
            dev.edit(e[i:i + 1])
            ora.edit(e[i:i + 1])
            assert_same_pool(dev, ora, f"bits {bits} edit {i} kind {int(e[i]['kind'])}")
        assert ora.overflows() == 0
    finally:
        capi.set_tuning(**old)


def test_fragmented_color_leaf_bit_exact():
    # 1500 one-voxel colored fills in one 16^3 color leaf: a color change at almost every voxel makes
    # the encoded chunk (2 header words per change) outgrow the encoder's 4096-word staging slot, so
    # the later passes encode the chunk again from the dense colors instead of copying it
    capi = _capi()
    cfg = small_cfg(64, 8, 9, 2, 4)
    rng = np.random.default_rng(7)
    pts = rng.choice(16 ** 3, 1500, replace=False)
    e = W.edits([W.edit(W.SPHERE_FILL, (16 + int(p) % 16, 16 + int(p) // 16 % 16, 16 + int(p) // 256), r2=0,
                        color=int(rng.integers(1 << 24))) for p in pts])
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    for i in range(0, len(e), 250):
        dev.edit(e[i:i + 250], capi.MODE_PLAIN)
        ora.edit(e[i:i + 250])
        assert_same_pool(dev, ora, f"after {i + 250} edits")
    assert dev.stats().color_leaf_words > 4096


def test_terrain_column_fill_bit_exact():
    capi = _capi()
    res = 128
    h = W.value_noise_heightmap(res, 4, 5, 20, 100)
    cfg = small_cfg(res, 8, 9, 2, 3)
    dev, ora = capi.DevicePool(cfg, heights=h), O.OraclePool(cfg, heights=h)
    e = W.edits([W.terrain_edit(res, seed=3, skin=6)] + list(W.terrain_sphere_edits(12, res, h, 2, 24, seed=4)))
    for i in range(len(e)):
        dev.edit(e[i:i + 1])
        ora.edit(e[i:i + 1])
        assert_same_pool(dev, ora, f"edit {i}")


def test_multi_edit_call_equals_one_by_one():
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)
    e = random_edits(16, 64, 33, True)
    a, b = capi.DevicePool(cfg), capi.DevicePool(cfg)
    a.edit(e)
    for i in range(len(e)):
        b.edit(e[i:i + 1])
    assert (a.root, a.color_root) == (b.root, b.color_root)
    assert np.array_equal(a.live_words()[1], b.live_words()[1])


def test_edge_cases():
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    cases = [
        W.edit(W.SPHERE_DIG, (10, 10, 10), r2=25),                   # dig on an empty DAG
        W.edit(W.SPHERE_PAINT, (10, 10, 10), r2=25, color=0xFF),     # paint on an empty DAG
        W.edit(W.AABB_FILL, (0, 0, 0), (64, 64, 64), color=0x123),   # fill everything -> filled root
        W.edit(W.AABB_FILL, (0, 0, 0), (64, 64, 64), color=0x123),   # idempotent
        W.edit(W.SPHERE_PAINT, (32, 32, 32), r2=100, color=0x456),
        W.edit(W.AABB_DIG, (0, 0, 0), (64, 64, 64)),                 # clear everything -> null root
        W.edit(W.AABB_FILL, (63, 63, 63), (64, 64, 64), color=0x9),  # one voxel at the far corner
        W.edit(W.SPHERE_FILL, (0, 0, 0), r2=0, color=0x7),           # r = 0: one voxel
        W.edit(W.AABB_FILL, (5, 5, 5), (5, 9, 9), color=0x1),        # empty box
        W.edit(W.SPHERE_FILL, (100, 100, 100), r2=10, color=0x2),    # outside the grid
    ]
    for i, e in enumerate(cases):
        dev.edit(e)
        ora.edit(e)
        assert_same_pool(dev, ora, f"edge {i}")
    assert ora.root != O.HD_NULL


def _render_both(dev, ora, cfg, view, Wd, Ht, shadow=True):
    rgba_d, rec_d = dev.render(view, Wd, Ht, shadow=shadow, records=True)
    words = ora.full_words()
    rgba_o, rec_o = O.render(words, cfg.node_levels, ora.root, ora.color_nodes(), ora.color_leaves(), ora.color_root,
                             cfg.color_leaf_level, view, Wd, Ht, shadow=shadow)
    return rgba_d, rec_d, rgba_o, rec_o


@pytest.mark.parametrize("lod", [False, True])
def test_render_bit_exact(lod):
    capi = _capi()
    res = 128
    h = W.value_noise_heightmap(res, 4, 9, 20, 90)
    cfg = small_cfg(res, 8, 9, 2, 3)
    dev, ora = capi.DevicePool(cfg, heights=h), O.OraclePool(cfg, heights=h)
    e = W.edits([W.terrain_edit(res, seed=3, skin=6)] + list(W.terrain_sphere_edits(10, res, h, 3, 20, seed=8)))
    dev.edit(e)
    ora.edit(e)
    assert (dev.root, dev.color_root) == (ora.root, ora.color_root)
    for seed in range(3):
        view = W.orbit_view((64, 64, 40), 150, seed, 160, 96, res=res, lod=lod)
        rgba_d, rec_d, rgba_o, rec_o = _render_both(dev, ora, cfg, view, 160, 96)
        hit = (rec_o[:, 3] >> 31).astype(bool)
        assert hit.mean() > 0.2
        assert np.array_equal(rec_d, rec_o), np.nonzero(np.any(rec_d != rec_o, axis=1))[0][:8]
        assert np.array_equal(rgba_d, rgba_o)


def test_traversal_pick_matches_reference_golden():
    capi = _capi()
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "traversal64.npz"))
    cfg = small_cfg(64, 6, 5, 2, 3)
    dev = capi.DevicePool(cfg)
    words = np.ascontiguousarray(g["words"], np.uint32)
    capi.check(capi.lib().hd_pool_write_words(dev.h, 0, len(words), words.ctypes.data))
    out = dev.traversal(g["o"], g["d"], root=int(g["root"]))
    ref = g["hits"]
    assert np.array_equal(out[:, 3], ref[:, 3])
    hit = ref[:, 3] > 0
    # NodePoolTraversal::Traversal<float> computed on the device with the same fp32 order
    assert np.array_equal(out[hit, :3], ref[hit, :3])


def test_render_beam_bit_exact():
    """Beam prepass (beam.frag) + beam-started primary rays: RGBA8 and hit records equal the
    oracle's restatement; LOD off (no block projection factor) is rejected."""
    capi = _capi()
    res = 256
    h = W.value_noise_heightmap(res, 5, 4, 40, 200)
    cfg = small_cfg(res, 10, 9, 2, 4)
    dev, ora = capi.DevicePool(cfg, heights=h), O.OraclePool(cfg, heights=h)
    e = W.edits([W.terrain_edit(res, seed=4, skin=8)] + list(W.terrain_sphere_edits(12, res, h, 3, 30, seed=2)))
    dev.edit(e)
    ora.edit(e)
    words = ora.full_words()
    for seed, (Wd, Ht) in enumerate([(160, 96), (123, 77)]):
        view = W.orbit_view((res / 2, res / 2, res * 0.3), res * 0.9, seed, Wd, Ht, res=res, elevation=(0.3, 0.6))
        rgba_d, rec_d = dev.render(view, Wd, Ht, shadow=True, records=True, beam=True)
        rgba_o, rec_o = O.render(words, cfg.node_levels, ora.root, ora.color_nodes(), ora.color_leaves(),
                                 ora.color_root, cfg.color_leaf_level, view, Wd, Ht, shadow=True, beam=True)
        assert (rec_o[:, 3] >> 31).mean() > 0.3
        assert np.array_equal(rec_d, rec_o), np.nonzero(np.any(rec_d != rec_o, axis=1))[0][:8]
        assert np.array_equal(rgba_d, rgba_o)
    view = W.orbit_view((res / 2, res / 2, res * 0.3), res * 0.9, 0, 32, 32, res=res, lod=False)
    with pytest.raises(capi.HDError, match="finite proj_factor"):
        dev.render(view, 32, 32, beam=True)
