"""Edit -> render frames captured as hipGraphs (hd_frame_*).

Bar: a one-edit frame leaves the same pool words, roots and color pool as hd_edit(...,
HD_MODE_PLAIN), bit-exact. A frame's passes are groups of pairwise-disjoint edits, while
hd_edit(..., HD_MODE_BATCHED) folds overlapping edits into one pass, so a multi-edit frame meets the
batched bar against it: the same reachable DAG up to node renumbering and the same colors on every
occupied voxel. Each view equals hd_trace of that view alone on the frame's new roots, bit-exact.
Frames queue back to back, survive a GC (roots handed back), and reject edits outside their
budget."""
import numpy as np
import pytest

from test_gpu_async import assert_equivalent_device_pools, assert_same_device_pools
from test_oracle_pins import small_cfg
from vkhashdag_amd import workloads as W

pytestmark = pytest.mark.gpu


def _capi():
    from vkhashdag_amd import capi
    return capi


def _terrain(capi, res, seed=3):
    h = W.value_noise_heightmap(res, 4, seed, res // 6, res * 3 // 4)
    cfg = small_cfg(res, 9, 9, 2, 3)
    pools = [capi.DevicePool(cfg, heights=h) for _ in range(2)]
    for p in pools:
        p.edit(W.terrain_edit(res, seed=seed, skin=6))
    return cfg, h, pools


def _views(res, seed, n=2):
    out = []
    for k in range(n):
        w, hh = (64, 48) if k == 0 else (40, 30)
        vb = W.orbit_view((res / 2, res / 2, res / 3), res * 0.9, seed * 7 + k, w, hh, res=res)
        out.append((vb, w, hh, 1))
    return out


def test_frames_equal_batched_edits_and_lone_renders():
    capi = _capi()
    res = 128
    cfg, h, (a, b) = _terrain(capi, res)
    f = capi.Frame(a, max_edits=16, max_passes=6, max_radius=20, n_views=2, max_width=64, max_height=48)
    for k in range(5):
        e = W.mixed_batch(12, res, h, 2, 16, seed=100 + k)
        views = _views(res, k)
        f.set(e, views)
        f.launch()
        f.wait()
        b.edit(e, capi.MODE_BATCHED)
        assert_equivalent_device_pools(a, b, cfg, f"frame {k}")
        imgs = f.images()
        for (vb, w, hh, fl), img in zip(views, imgs):
            assert np.array_equal(img, a.render(vb, w, hh, shadow=True)), f"frame {k}: view differs"
    f.close()


def test_one_edit_frames_equal_plain_edits():
    capi = _capi()
    res = 128
    cfg, h, (a, b) = _terrain(capi, res, seed=5)
    f = capi.Frame(a, max_edits=4, max_passes=2, max_radius=24, n_views=1, max_width=32, max_height=32)
    e = W.terrain_sphere_edits(10, res, h, 3, 20, seed=7)
    for k in range(len(e)):
        vb = W.orbit_view((res / 2, res / 2, res / 3), res * 0.9, k, 32, 32, res=res)
        f.set(e[k:k + 1], [(vb, 32, 32, 1)])
        f.launch()
    f.wait()
    b.edit(e, capi.MODE_PLAIN)
    assert_same_device_pools(a, b, "ten queued one-edit frames")
    f.close()


def test_frames_continue_after_gc_and_check_budgets():
    capi = _capi()
    res = 128
    cfg, h, (a, b) = _terrain(capi, res, seed=6)
    f = capi.Frame(a, max_edits=8, max_passes=3, max_radius=12, n_views=1, max_width=32, max_height=32)
    e = W.mixed_batch(8, res, h, 2, 10, seed=11)
    f.set(e, [])
    f.launch()
    f.wait()
    b.edit(e, capi.MODE_BATCHED)
    a.gc()
    b.gc()
    f.set_roots(a.root, a.color_root)
    e2 = W.mixed_batch(8, res, h, 2, 10, seed=12)
    f.set(e2, [])
    f.launch()
    f.wait()
    b.edit(e2, capi.MODE_BATCHED)
    assert_equivalent_device_pools(a, b, cfg, "after gc")
    with pytest.raises(capi.HDError, match="radius budget"):
        f.set(W.edits([W.edit(W.SPHERE_FILL, (64, 64, 64), r2=40 * 40, color=1)]), [])
    with pytest.raises(capi.HDError, match="column fills"):
        f.set(W.edits([W.terrain_edit(res, seed=1)]), [])
    f.close()


def test_frames_with_gc_in_the_graph_equal_edit_then_gc():
    # HD_FRAME_GC: edit -> GC -> render captured once; every replay equals hd_edit (batched) then
    # hd_gc of the new root on a twin pool (the GC runs inside the graph), and each view equals a
    # lone render of the compacted root
    capi = _capi()
    res = 128
    cfg, h, (a, b) = _terrain(capi, res)
    f = capi.Frame(a, max_edits=16, max_passes=6, max_radius=20, n_views=2, max_width=64, max_height=48, gc=True)
    for k in range(4):
        e = W.mixed_batch(12, res, h, 2, 16, seed=300 + k)
        views = _views(res, 20 + k)
        f.set(e, views)
        f.launch()
        f.wait()
        b.edit(e, capi.MODE_BATCHED)
        b.gc()
        assert_equivalent_device_pools(a, b, cfg, f"frame {k} with GC")
        for (vb, w, hh, fl), img in zip(views, f.images()):
            assert np.array_equal(img, a.render(vb, w, hh, shadow=True)), f"frame {k}: view differs"
    f.close()
    # the pool goes on as a plain pool afterwards
    e = W.mixed_batch(6, res, h, 2, 12, seed=400)
    a.edit(e, capi.MODE_BATCHED)
    b.edit(e, capi.MODE_BATCHED)
    assert_equivalent_device_pools(a, b, cfg, "after the frames")
