"""Generic host editors composed with device edits on one pool (hd_edit_host).

The reference's StatelessEditor concept (Editor.hpp:36-40) admits any functor; hd_edit_host runs
one on the host through the reference's single-threaded recursion (NodePool.hpp:319-417) against
the device pool's words and uploads the appended ranges and counters (DAGNodePool::Flush's dirty
ranges, src/DAGNodePool.cpp:56-85).

Bar: bit-exact. A host functor restating SphereEditor<kDig> / AABBEditor (main.cpp:32-150),
between device edits, leaves the same pool words, counters and roots as the oracle applying the
equivalent descriptor edits in the same order; a functor with no descriptor counterpart (a
checkerboard in a box) leaves the voxel set a brute-force grid predicts."""
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


def _bounds(level, x, y, z, voxel_level):
    b = voxel_level - level
    return [x << b, y << b, z << b], [(x + 1) << b, (y + 1) << b, (z + 1) << b]


def sphere_dig(c, r2, voxel_level, capi):
    """SphereEditor<kDig> (main.cpp:73-150) as a host functor."""
    def edit_node(level, x, y, z, _ptr):
        lb, ub = _bounds(level, x, y, z, voxel_level)
        ld = [lb[i] - c[i] for i in range(3)]
        ud = [ub[i] - c[i] for i in range(3)]
        if sum(max(ld[i] * ld[i], ud[i] * ud[i]) for i in range(3)) <= r2:
            return capi.EDIT_TYPE_CLEAR
        mn = sum((ld[i] * ld[i] if ld[i] > 0 else 0) + (ud[i] * ud[i] if ud[i] < 0 else 0) for i in range(3))
        return capi.EDIT_TYPE_NOT_AFFECTED if mn > r2 else capi.EDIT_TYPE_PROCEED

    def edit_voxel(x, y, z, v):
        return v and not ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= r2)
    return edit_node, edit_voxel


def box_fill(lo, hi, voxel_level, capi, voxel=None):
    """AABBEditor (main.cpp:32-70), stateless, as a host functor; `voxel` replaces the in-box rule
    (then no node is filled whole)."""
    def edit_node(level, x, y, z, _ptr):
        lb, ub = _bounds(level, x, y, z, voxel_level)
        if any(ub[i] <= lo[i] for i in range(3)) or any(lb[i] >= hi[i] for i in range(3)):
            return capi.EDIT_TYPE_NOT_AFFECTED
        if voxel is None and all(lb[i] >= lo[i] for i in range(3)) and all(ub[i] <= hi[i] for i in range(3)):
            return capi.EDIT_TYPE_FILL
        return capi.EDIT_TYPE_PROCEED

    def edit_voxel(x, y, z, v):
        inside = lo[0] <= x < hi[0] and lo[1] <= y < hi[1] and lo[2] <= z < hi[2]
        return v or (inside and (voxel is None or voxel(x, y, z)))
    return edit_node, edit_voxel


def test_host_functors_between_device_edits_bit_exact():
    capi = _capi()
    res = 64
    cfg = small_cfg(res, 6, 5, 2, 3)
    vl = cfg.voxel_level
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    e1 = random_edits(8, res, 21, True)
    dev.edit(e1)
    ora.edit(e1)
    print("host sphere dig", flush=True)
    dev.edit_host(*sphere_dig((30, 34, 28), 15 * 15, vl, capi))
    ora.edit(W.edit(W.SPHERE_DIG, (30, 34, 28), r2=15 * 15))
    assert_same_pool(dev, ora, "after the host dig")
    e2 = random_edits(8, res, 22, True)
    dev.edit(e2)
    ora.edit(e2)
    assert_same_pool(dev, ora, "device edits after the host dig")
    print("host box fill", flush=True)
    dev.edit_host(*box_fill((8, 0, 16), (40, 64, 48), vl, capi))
    ora.edit(W.edit(W.AABB_FILL, (8, 0, 16), (40, 64, 48), colored=False))
    assert_same_pool(dev, ora, "after the host box fill")
    dev.edit(e1[:4], capi.MODE_BATCHED)  # batched device edits after host edits too
    ora.edit(e1[:4])
    mirror = O.OraclePool(cfg)
    mirror.load_state(dev.state())
    assert np.array_equal(mirror.canonical(), ora.canonical())


def test_host_functor_without_a_descriptor_equals_brute_force():
    capi = _capi()
    res = 64
    cfg = small_cfg(res, 6, 5, 2, 3)
    dev = capi.DevicePool(cfg)
    e = random_edits(6, res, 31, True)
    dev.edit(e)
    mirror = O.OraclePool(cfg)
    mirror.load_state(dev.state())
    before = mirror.voxels().copy()
    lo, hi = (5, 9, 3), (50, 41, 60)
    dev.edit_host(*box_fill(lo, hi, cfg.voxel_level, capi, voxel=lambda x, y, z: (x + y + z) % 3 == 0))
    z, y, x = np.meshgrid(np.arange(res), np.arange(res), np.arange(res), indexing="ij")
    want = before | ((x >= lo[0]) & (x < hi[0]) & (y >= lo[1]) & (y < hi[1]) & (z >= lo[2]) & (z < hi[2])
                     & ((x + y + z) % 3 == 0))
    mirror = O.OraclePool(cfg)
    mirror.load_state(dev.state())
    assert np.array_equal(mirror.voxels(), want)
    # an editor that changes nothing keeps the root and appends nothing
    r, used = dev.root, dev.stats().used_words
    dev.edit_host(lambda *a: capi.EDIT_TYPE_PROCEED, lambda x, y, z, v: v)
    assert dev.root == r and dev.stats().used_words == used
