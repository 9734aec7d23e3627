"""Device GC parity: the compacted pool equals the oracle's GC (= the reference's ThreadedGC on one
worker, test_oracle_pins.py) word for word, roots are remapped identically, the reachable DAG and
colors are unchanged, and edits after GC stay bit-exact."""
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


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gc_bit_exact(seed):
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)
    e = random_edits(30, 64, seed, True)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    dev.edit(e)
    ora.edit(e)
    before = ora.canonical()
    dev.gc()
    ora.gc()
    assert_same_pool(dev, ora, "after gc")
    assert np.array_equal(dev.filled(), ora.filled())
    assert np.array_equal(ora.canonical(), before)
    # editing continues bit-exactly on the compacted pools
    e2 = random_edits(10, 64, seed + 100, True)
    for i in range(len(e2)):
        dev.edit(e2[i:i + 1])
        ora.edit(e2[i:i + 1])
        assert_same_pool(dev, ora, f"post-gc edit {i}")


def test_gc_multiple_roots_and_null():
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    e = random_edits(20, 64, 5, True)
    dev.edit(e[:10])
    ora.edit(e[:10])
    keep_d, keep_o = dev.root, ora.root
    dev.edit(e[10:])
    ora.edit(e[10:])
    rd = dev.gc(extra_roots=(keep_d, O.HD_NULL))
    ro = ora.gc(extra_roots=(keep_o, O.HD_NULL))
    assert list(rd) == list(ro)
    assert rd[2] == O.HD_NULL
    assert_same_pool(dev, ora, "multi-root gc")


def test_gc_after_batched_keeps_dag_and_colors():
    capi = _capi()
    res = 128
    h = W.value_noise_heightmap(res, 4, 2, 20, 100)
    cfg = small_cfg(res, 8, 9, 2, 3)
    dev = capi.DevicePool(cfg, heights=h)
    dev.edit(W.terrain_edit(res, seed=3, skin=6))
    dev.edit(W.mixed_batch(150, res, h, 2, 24, seed=5), capi.MODE_BATCHED)
    L = cfg.node_levels
    w0 = dev.full_words()
    c0 = O.canonical_words(w0, L, np.array([dev.root], np.uint32))
    used0 = dev.stats().used_words
    dev.gc()
    w1 = dev.full_words()
    assert np.array_equal(O.canonical_words(w1, L, np.array([dev.root], np.uint32)), c0)
    assert dev.stats().used_words < used0
    assert dev.stats().overflow == 0


def test_gc_empty_pool():
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    dev.edit(W.edit(W.SPHERE_DIG, (1, 1, 1), r2=4))
    ora.edit(W.edit(W.SPHERE_DIG, (1, 1, 1), r2=4))
    dev.gc()
    ora.gc()
    assert_same_pool(dev, ora, "empty gc")


def test_gc_twice_leaves_the_pool_unchanged():
    """A second GC finds every live node already packed in pointer order in its bucket: pool words,
    bucket counters and roots stay as they are (so repeating GC on a compacted pool, as the
    benchmark's steady-state GC does, redoes exactly the first call's work)."""
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    e = random_edits(30, 64, 7, True)
    dev.edit(e)
    ora.edit(e)
    dev.gc()
    ora.gc()
    w1, r1 = dev.full_words().copy(), dev.root
    dev.gc()
    ora.gc()
    assert dev.root == r1
    assert np.array_equal(dev.full_words(), w1)
    assert_same_pool(dev, ora, "second gc")

