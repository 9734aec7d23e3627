"""Checkpoint dump / load (hd_pool_save / hd_pool_load).

Bar: bit-exact. A pool loaded from a file equals the saved pool (roots, counters, used words, color
pool, overflow flags); edits, a GC with extra roots and a render continued on the loaded pool give
the same words and pixels as on the original. Malformed files are rejected with an error."""
import numpy as np
import pytest

from test_gpu_async import assert_same_device_pools
from test_oracle_pins import small_cfg
from vkhashdag_amd import workloads as W

pytestmark = pytest.mark.gpu


def _capi():
    from vkhashdag_amd import capi
    return capi


@pytest.mark.parametrize("keep_history", [False, True])
def test_save_load_continue_bit_exact(tmp_path, keep_history):
    capi = _capi()
    res, seed = 128, 5
    h = W.value_noise_heightmap(res, 4, seed, res // 6, res * 3 // 4)
    cfg = small_cfg(res, 9, 9, 2, 3)
    cfg.keep_history = keep_history
    a = capi.DevicePool(cfg, heights=h)
    a.edit(W.terrain_edit(res, seed=seed, skin=6))
    a.edit(W.mixed_batch(40, res, h, 2, 14, seed=11), capi.MODE_BATCHED)
    old = (a.root, a.color_root)
    a.edit(W.terrain_sphere_edits(8, res, h, 3, 12, seed=12))
    path = tmp_path / "pool.hdag"
    a.save(path, extra_roots=[old[0]])
    b, extra = capi.DevicePool.load(path, heights=h)
    assert extra == [old[0]]
    assert b.cfg.total_words == cfg.total_words and b.cfg.keep_history == keep_history
    assert_same_device_pools(a, b, "after load")

    e = W.mixed_batch(24, res, h, 2, 16, seed=13)
    for p in (a, b):
        p.edit(e, capi.MODE_BATCHED)
        p.edit(W.terrain_sphere_edits(4, res, h, 2, 10, seed=14))
    assert_same_device_pools(a, b, "after continued edits")
    ra = a.gc(extra_roots=[old[0]])
    rb = b.gc(extra_roots=extra)
    assert np.array_equal(ra, rb)
    assert_same_device_pools(a, b, "after GC")
    vb = W.orbit_view((res / 2, res / 2, res / 3), res * 0.9, 3, 64, 48, res=res)
    assert np.array_equal(a.render(vb, 64, 48), b.render(vb, 64, 48))


def test_load_rejects_bad_files(tmp_path):
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)
    a = capi.DevicePool(cfg)
    a.edit(W.edits([W.edit(W.SPHERE_FILL, (32, 32, 32), r2=100, color=0x123456)]))
    good = tmp_path / "good.hdag"
    a.save(good)
    data = good.read_bytes()
    bad = tmp_path / "bad.hdag"
    bad.write_bytes(b"NOTADAG!" + data[8:])
    with pytest.raises(capi.HDError, match="checkpoint"):
        capi.DevicePool.load(bad)
    bad.write_bytes(data[: len(data) - 8])
    with pytest.raises(capi.HDError, match="truncated"):
        capi.DevicePool.load(bad)
    with pytest.raises(capi.HDError, match="cannot open"):
        capi.DevicePool.load(tmp_path / "missing.hdag")
    b, extra = capi.DevicePool.load(good)
    assert extra == []
    assert_same_device_pools(a, b, "small pool")


def test_save_load_then_gc_keeps_the_filled_chain(tmp_path):
    # the filled chain is a GC root once created: a loaded pool knows it was created, so a GC right
    # after the load compacts exactly as on the original
    capi = _capi()
    res = 64
    cfg = small_cfg(res, 6, 5, 2, 3)
    a = capi.DevicePool(cfg)
    a.edit(W.edits([W.edit(W.AABB_FILL, (0, 0, 0), (64, 64, 40), color=0x3355AA),
                    W.edit(W.SPHERE_DIG, (30, 30, 38), r2=60)]))
    assert len(a.filled()) == cfg.node_levels
    path = tmp_path / "pool.hdag"
    a.save(path)
    b, _ = capi.DevicePool.load(path)
    assert_same_device_pools(a, b, "after load")
    a.gc()
    b.gc()
    assert_same_device_pools(a, b, "GC right after the load")
    assert np.array_equal(a.filled(), b.filled())


def test_interop_rejects_bucket_counters_over_capacity():
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)
    a = capi.DevicePool(cfg)
    a.edit(W.edits([W.edit(W.SPHERE_FILL, (32, 32, 32), r2=100, color=0x123456)]))
    bw, live = a.live_words()
    before = (a.bucket_words().copy(), a.live_words()[1].copy())
    bad = bw.copy()
    b = int(np.nonzero(bw)[0][0])
    bad[b] = cfg.words_per_bucket + 1
    with pytest.raises(capi.HDError, match="holds"):
        a.write_live_words(bad, np.zeros(int(bad.sum()), np.uint32))
    with pytest.raises(capi.HDError, match="holds"):
        capi.check(capi.lib().hd_pool_write_bucket_words(a.h, b, 1, bad[b:b + 1].ctypes.data))
    with pytest.raises(capi.HDError, match="counters sum"):
        a.write_live_words(bw, live[:-1])
    # nothing was written by the rejected calls
    assert np.array_equal(a.bucket_words(), before[0])
    assert np.array_equal(a.live_words()[1], before[1])
