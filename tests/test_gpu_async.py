"""The sync-free edit pipeline (default) against the host-synchronized one (HD_EDIT_PATH=sync) and
the oracle.

Bar: bit-exact for plain edits. Both device pipelines follow the reference's single-threaded Edit,
so after the same plain edit calls their pools agree word for word: roots, per-bucket word
counters, live words, color octree nodes and color leaf words. A batched call folds overlapping
edits into one pass on the sync-free pipeline and runs dependency groups on the synchronized one,
so after one the bar is the batched one: the same reachable DAG up to node renumbering and the
same colors on every occupied voxel. The cases push the sync-free pipeline's level bounds (spheres
and boxes from one voxel to the whole grid, bucket-full fallbacks, batched passes) and mix in
column fills, which take the synchronized pipeline in the middle of a call."""
import numpy as np
import pytest

import _oracle as O
from test_gpu_plain import assert_same_pool
from test_oracle_pins import random_edits, small_cfg
from vkhashdag_amd import workloads as W

pytestmark = pytest.mark.gpu


def _capi():
    from vkhashdag_amd import capi
    return capi


def assert_same_device_pools(a, b, where=""):
    assert (a.root, a.color_root) == (b.root, b.color_root), (where, hex(a.root), hex(b.root))
    assert a.stats().overflow == b.stats().overflow, where
    ba, wa = a.live_words()
    bb, wb = b.live_words()
    assert np.array_equal(ba, bb), where
    assert np.array_equal(wa, wb), where
    assert np.array_equal(a.color_nodes(), b.color_nodes()), where
    assert np.array_equal(a.color_leaves(), b.color_leaves()), where


def assert_equivalent_device_pools(a, b, cfg, where=""):
    assert a.stats().overflow == 0 and b.stats().overflow == 0, where
    L = cfg.node_levels
    wa, wb = a.full_words(), b.full_words()
    ca = O.canonical_words(wa, L, np.array([a.root], np.uint32))
    cb = O.canonical_words(wb, L, np.array([b.root], np.uint32))
    assert np.array_equal(ca, cb), (where, "canonical DAGs differ")
    grid = O.decode_voxels(wa, L, a.root)
    vox = np.ascontiguousarray(np.argwhere(grid)[:, ::-1].astype(np.uint32))
    if len(vox):
        ga = O.color_fetch(a.color_nodes(), a.color_leaves(), a.color_root, cfg.voxel_level, cfg.color_leaf_level, vox)
        gb = O.color_fetch(b.color_nodes(), b.color_leaves(), b.color_root, cfg.voxel_level, cfg.color_leaf_level, vox)
        bad = np.nonzero(np.any(ga != gb, axis=1))[0]
        assert len(bad) == 0, (where, len(bad), vox[bad[:5]], ga[bad[:5]], gb[bad[:5]])


def _run(cfg, calls, monkeypatch, sync, heights=None):
    capi = _capi()
    if sync:
        monkeypatch.setenv("HD_EDIT_PATH", "sync")
    else:
        monkeypatch.delenv("HD_EDIT_PATH", raising=False)
    dev = capi.DevicePool(cfg, heights=heights)
    for edits, mode in calls:
        dev.edit(edits, capi.MODE_BATCHED if mode == "batched" else capi.MODE_PLAIN)
    monkeypatch.delenv("HD_EDIT_PATH", raising=False)
    return dev


def _compare(cfg, calls, monkeypatch, heights=None):
    a = _run(cfg, calls, monkeypatch, False, heights)
    b = _run(cfg, calls, monkeypatch, True, heights)
    if any(mode == "batched" for _, mode in calls):
        assert_equivalent_device_pools(a, b, cfg)
    else:
        assert_same_device_pools(a, b)
    return a


def _extreme_edits(res, seed, n):
    """Spheres from r = 0 to beyond the grid and boxes from one voxel to the whole grid."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        k = int(rng.integers(0, 6))
        col = int(rng.integers(0, 1 << 24))
        c = tuple(int(v) for v in rng.integers(-res // 4, res + res // 4, 3).clip(0, (1 << 31) - 1))
        if k < 3:
            r = int(rng.choice([0, 1, 2, 3, res // 8, res // 2, res, 2 * res]))
            out.append(W.edit([W.SPHERE_FILL, W.SPHERE_DIG, W.SPHERE_PAINT][k], c, r2=r * r, color=col))
        else:
            lo = rng.integers(0, res, 3)
            ext = rng.choice([1, 2, res // 4, res], 3)
            hi = np.minimum(lo + ext, res)
            out.append(W.edit([W.AABB_FILL, W.AABB_DIG, W.AABB_PAINT][k - 3], tuple(int(v) for v in lo),
                              tuple(int(v) for v in hi), color=col))
    return W.edits(out)


@pytest.mark.parametrize("res,bits,wbpp,depth,seed", [(64, 6, 5, 3, 1), (64, 3, 5, 2, 2), (128, 8, 9, 3, 3),
                                                      (32, 2, 4, 1, 4), (64, 6, 5, 0, 5)])
def test_sync_free_equals_synchronized_plain(res, bits, wbpp, depth, seed, monkeypatch):
    print(f"res {res} bits {bits} wbpp {wbpp} depth {depth} seed {seed}", flush=True)
    cfg = small_cfg(res, bits, wbpp, 2, depth)
    e = random_edits(40, res, seed, True)
    x = _extreme_edits(res, seed + 50, 20)
    _compare(cfg, [(e, "plain"), (x, "plain")], monkeypatch)


@pytest.mark.parametrize("seed", [1, 2])
def test_sync_free_equals_oracle_extreme_edits(seed):
    capi = _capi()
    res = 64
    cfg = small_cfg(res, 6, 5, 2, 3)
    e = _extreme_edits(res, seed, 40)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    for i in range(len(e)):
        dev.edit(e[i:i + 1])
        ora.edit(e[i:i + 1])
        assert_same_pool(dev, ora, f"edit {i} kind {int(e[i]['kind'])}")


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_sync_free_equals_synchronized_batched(seed, monkeypatch):
    res = 128
    cfg = small_cfg(res, 8, 9, 2, 3)
    e = random_edits(120, res, seed, True)
    x = _extreme_edits(res, seed + 70, 30)
    _compare(cfg, [(e, "batched"), (x, "batched"), (e[:7], "plain")], monkeypatch)


def test_sync_free_terrain_mixed_with_column_fill(monkeypatch):
    res = 256
    h = W.value_noise_heightmap(res, 5, 3, 40, 200)
    cfg = small_cfg(res, 10, 9, 2, 4)
    terrain = W.edits([W.terrain_edit(res, seed=2, skin=8)])
    sph = W.terrain_sphere_edits(48, res, h, 2, 40, seed=6)
    mix = W.edits(list(sph[:10]) + [W.terrain_edit(res, seed=5, skin=4)] + list(sph[10:20]))
    batch = W.mixed_batch(300, res, h, 2, 48, seed=9)
    a = _compare(cfg, [(terrain, "plain"), (sph, "plain"), (mix, "plain"), (batch, "batched")], monkeypatch,
                 heights=h)
    assert a.stats().overflow == 0


def test_sync_free_terrain_vs_oracle_256():
    capi = _capi()
    res = 256
    h = W.value_noise_heightmap(res, 5, 4, 40, 200)
    cfg = small_cfg(res, 10, 9, 2, 4)
    dev, ora = capi.DevicePool(cfg, heights=h), O.OraclePool(cfg, heights=h)
    t = W.edits([W.terrain_edit(res, seed=4, skin=8)])
    dev.edit(t)
    ora.edit(t)
    e = W.terrain_sphere_edits(32, res, h, 2, 60, seed=11)
    dev.edit(e)
    ora.edit(e)
    assert_same_pool(dev, ora, "terrain 256 + 32 spheres")
