"""Batched device edits equal sequential application.

Bar: the DAG reachable from the batched root equals, up to node renumbering, the DAG the oracle
(= the reference's Edit) builds by applying the same edits one at a time; every occupied voxel
decodes to the same color. Empty-voxel colors and the octree shape are history-dependent in the
reference too (VBREditorWrapper colors empty voxels with the edit color), so they are not
compared. Colors are RGB8, for which FillNode is exact."""
import numpy as np
import pytest

import _oracle as O
from test_oracle_pins import random_edits, small_cfg
from vkhashdag_amd import workloads as W

pytestmark = pytest.mark.gpu


def _capi():
    from vkhashdag_amd import capi
    return capi


def assert_equivalent(dev, ora, cfg):
    # a full bucket silently drops an edit in the reference (NodePool.hpp:138-139,195,210); the
    # batched and sequential layouts differ, so equivalence is only defined without overflow
    assert ora.overflows() == 0 and dev.stats().overflow == 0, (ora.overflows(), dev.stats().overflow)
    L = cfg.node_levels
    wd = dev.full_words()
    cd = O.canonical_words(wd, L, np.array([dev.root], np.uint32))
    co = ora.canonical()
    assert np.array_equal(cd, co), "canonical DAGs differ"
    grid = O.decode_voxels(wd, L, dev.root)
    vox = np.ascontiguousarray(np.argwhere(grid)[:, ::-1].astype(np.uint32))
    if len(vox) == 0:
        return
    got = O.color_fetch(dev.color_nodes(), dev.color_leaves(), dev.color_root, cfg.voxel_level,
                        cfg.color_leaf_level, vox)
    want = ora.colors(vox)
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert len(bad) == 0, (len(bad), vox[bad[:5]], got[bad[:5]], want[bad[:5]])


@pytest.mark.parametrize("seed,res,n", [(1, 64, 40), (2, 64, 80), (3, 128, 60), (4, 32, 50)])
def test_batched_random_edits_equal_sequential(seed, res, n):
    capi = _capi()
    cfg = small_cfg(res, 8, 9 if res > 64 else 6, 2, 3)
    e = random_edits(n, res, seed, True)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    dev.edit(e, capi.MODE_BATCHED)
    print("groups", dev.stats().last_groups, "of", n, flush=True)
    ora.edit(e)
    assert_equivalent(dev, ora, cfg)


def test_batched_root_color_leaf_equals_sequential():
    # a color leaf of 32^3 voxels (more than a workgroup rebuilds in LDS): the dense per-voxel path
    capi = _capi()
    cfg = small_cfg(32, 6, 5, 2, 5)
    e = random_edits(30, 32, 13, True)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    dev.edit(e, capi.MODE_BATCHED)
    ora.edit(e)
    assert_equivalent(dev, ora, cfg)


def test_batched_terrain_mixed_batch_equals_sequential():
    capi = _capi()
    res = 128
    h = W.value_noise_heightmap(res, 4, 2, 20, 100)
    cfg = small_cfg(res, 8, 9, 2, 3)
    dev, ora = capi.DevicePool(cfg, heights=h), O.OraclePool(cfg, heights=h)
    t = W.terrain_edit(res, seed=3, skin=6)
    dev.edit(t)
    ora.edit(t)
    e = W.mixed_batch(200, res, h, 2, 24, seed=5)
    dev.edit(e, capi.MODE_BATCHED)
    groups = dev.stats().last_groups
    ora.edit(e)
    assert 1 <= groups < len(e)
    assert_equivalent(dev, ora, cfg)


def test_batched_terrain_build_inside_batch():
    # the column fill overlaps everything: it gets a group of its own, in order
    capi = _capi()
    res = 64
    h = W.value_noise_heightmap(res, 3, 4, 10, 50)
    cfg = small_cfg(res, 8, 6, 2, 3)
    dev, ora = capi.DevicePool(cfg, heights=h), O.OraclePool(cfg, heights=h)
    e = W.edits(list(random_edits(10, res, 7, True)) + [W.terrain_edit(res, seed=2, skin=3)]
                + list(W.mixed_batch(30, res, h, 2, 12, seed=9)))
    dev.edit(e, capi.MODE_BATCHED)
    ora.edit(e)
    assert_equivalent(dev, ora, cfg)


def test_single_edit_batched_equals_plain_bit_exact():
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)
    e = random_edits(12, 64, 11, True)
    a, b = capi.DevicePool(cfg), capi.DevicePool(cfg)
    for i in range(len(e)):
        a.edit(e[i:i + 1], capi.MODE_BATCHED)
        b.edit(e[i:i + 1], capi.MODE_PLAIN)
    assert (a.root, a.color_root) == (b.root, b.color_root)
    assert np.array_equal(a.live_words()[1], b.live_words()[1])
    assert np.array_equal(a.color_leaves(), b.color_leaves())


def test_disjoint_edits_one_group():
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)
    e = W.edits([W.edit(W.SPHERE_FILL, (8 + 16 * (i % 4), 8 + 16 * (i // 4 % 4), 32), r2=36, color=0x10101 * i)
                 for i in range(16)])
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    dev.edit(e, capi.MODE_BATCHED)
    assert dev.stats().last_groups == 1
    ora.edit(e)
    assert_equivalent(dev, ora, cfg)


@pytest.mark.parametrize("top_pairs", [4096, 64])
def test_one_pass_long_edit_lists(top_pairs):
    # 432 pairwise-disjoint edits form one pass: the root's list holds all of them and the upper
    # levels' lists stay longer than a wavefront splits (16), in the one-workgroup top-down
    # (top_pairs 4096) and in the grid kernels (top_pairs 64 sends every level after the root there)
    capi = _capi()
    old = capi.get_tuning()
    try:
        capi.set_tuning(**dict(old, top_pairs=top_pairs))
        res = 256
        cfg = small_cfg(res, 8, 9, 2, 3)
        rng = np.random.default_rng(21)
        ed = []
        for k in range(432):
            x, y, z = 12 + 28 * (k % 9), 12 + 28 * (k // 9 % 8), 12 + 40 * (k // 72)
            if k % 3 == 2:  # a paint box inside the cell
                ed.append(W.edit(W.AABB_PAINT, (x - 9, y - 9, z - 9), (x + 9, y + 9, z + 9),
                                 color=int(rng.integers(1 << 24))))
            else:
                ed.append(W.edit(W.SPHERE_FILL, (x, y, z), r2=int(rng.integers(4, 120)),
                                 color=int(rng.integers(1 << 24))))
        e = W.edits(ed)
        dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
        # a filled slab first, so the paints and fills meet existing voxels and colors
        base = W.edits([W.edit(W.AABB_FILL, (0, 0, 0), (res, res, 100), color=0x304050)])
        dev.edit(base)
        ora.edit(base)
        dev.edit(e, capi.MODE_BATCHED)
        assert dev.stats().last_groups == 1
        ora.edit(e)
        assert_equivalent(dev, ora, cfg)
    finally:
        capi.set_tuning(**old)


def test_reserve_changes_nothing_and_edits_match():
    # hd_edit_reserve allocates the passes' buffers only: the pool is unchanged, and the edit that
    # follows gives the same words as one without the reservation
    capi = _capi()
    cfg = small_cfg(128, 8, 9, 2, 3)
    e = random_edits(60, 128, 17, True)
    a, b = capi.DevicePool(cfg), capi.DevicePool(cfg)
    before = a.live_words()
    a.reserve(e, capi.MODE_BATCHED)
    a.reserve(e[:5], capi.MODE_PLAIN)
    after = a.live_words()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    assert a.stats().used_words == b.stats().used_words
    a.edit(e, capi.MODE_BATCHED)
    b.edit(e, capi.MODE_BATCHED)
    assert (a.root, a.color_root) == (b.root, b.color_root)
    assert np.array_equal(a.live_words()[1], b.live_words()[1])
    assert np.array_equal(a.color_nodes(), b.color_nodes())
    assert np.array_equal(a.color_leaves(), b.color_leaves())


@pytest.mark.parametrize("depth", [3, 4])
@pytest.mark.parametrize("words", [0, 24])
def test_unstaged_color_jobs_match_staged(depth, words, monkeypatch):
    # A color job whose encoded chunk does not fit the staging area is rebuilt and encoded straight
    # into the color leaves after the append scan. With no staging (every job) or a little (some
    # jobs) the edits give the same roots, node words, color nodes and color leaves as with the
    # default staging. depth 3: 512-voxel color leaves; depth 4: 4096-voxel ones.
    capi = _capi()
    res = 128
    h = W.value_noise_heightmap(res, 4, 2, 20, 100)
    cfg = small_cfg(res, 8, 9, 2, depth)
    t = W.terrain_edit(res, seed=3, skin=6)
    e = W.mixed_batch(200, res, h, 2, 24, seed=5)
    pools = []
    for env in (None, str(words)):
        if env is None:
            monkeypatch.delenv("HD_EDIT_STAGING_WORDS", raising=False)
        else:
            monkeypatch.setenv("HD_EDIT_STAGING_WORDS", env)
        p = capi.DevicePool(cfg, heights=h)
        p.edit(t)
        p.edit(e, capi.MODE_BATCHED)
        pools.append(p)
    monkeypatch.delenv("HD_EDIT_STAGING_WORDS", raising=False)
    a, b = pools
    assert (a.root, a.color_root) == (b.root, b.color_root)
    assert np.array_equal(a.live_words()[1], b.live_words()[1])
    assert np.array_equal(a.color_nodes(), b.color_nodes())
    assert np.array_equal(a.color_leaves(), b.color_leaves())
