"""Weighted VBR colors on the device: R5G6B5 endpoints with 1-3-bit weights.

A VBRColor with bits_per_weight > 0 holds two R5G6B5 endpoints (low and high 16 bits) and a
weight; its color is the fp32 mix of the endpoints at weight / (2^bpw - 1) (VBRColor.hpp:19-44).
Color leaves store such runs as headers {colors, voxel offset, bpw, weight offset} plus packed
weight bits (VBRBitset / VBRChunkWriter, VBRColor.hpp:65-227, :368-487), and the renderer decodes
them (Color_GetWeight / Color_UnpackRGB565, trace.frag:287-349).

Bar: bit-exact. Plain edits leave the oracle's color node and leaf words (and every node-pool
word); frames equal the oracle's transliteration; batched edits decode to the oracle's colors on
every occupied voxel."""
import numpy as np
import pytest

import _oracle as O
from test_gpu_batched import assert_equivalent
from test_gpu_plain import assert_same_pool
from test_oracle_pins import random_edits, small_cfg
from vkhashdag_amd import workloads as W

pytestmark = pytest.mark.gpu


def _capi():
    from vkhashdag_amd import capi
    return capi


def weighted_edits(n, res, seed, bpws=(0, 1, 2, 3)):
    """random_edits with each colored edit given a random bits_per_weight from bpws: for bpw > 0 the
    color word is two random R5G6B5 endpoints and the weight is uniform in [0, 2^bpw)."""
    e = random_edits(n, res, seed, True)
    rng = np.random.default_rng(seed + 1000)
    for r in e:
        if r["colored"]:
            bpw = int(bpws[int(rng.integers(len(bpws)))])
            if bpw:
                r["colors"] = int(rng.integers(0, 1 << 32, dtype=np.uint64))
                r["bpw"] = bpw
                r["weight"] = int(rng.integers(0, 1 << bpw))
    return e


@pytest.mark.parametrize("seed,res,depth", [(1, 64, 3), (2, 64, 4), (3, 128, 3), (4, 32, 5)])
def test_weighted_colors_plain_bit_exact(seed, res, depth):
    capi = _capi()
    cfg = small_cfg(res, 8, 9 if res > 64 else 6, 2, depth)
    e = weighted_edits(30, res, seed)
    assert (e["bpw"] > 0).sum() >= 5
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    for i in range(len(e)):
        dev.edit(e[i:i + 1])
        ora.edit(e[i:i + 1])
        assert_same_pool(dev, ora, f"edit {i} kind {int(e[i]['kind'])} bpw {int(e[i]['bpw'])}")
    view = W.orbit_view((res / 2, res / 2, res / 2), res * 1.6, seed, 96, 72, res=res, lod=False)
    rgba_d, rec_d = dev.render(view, 96, 72, shadow=True, records=True)
    rgba_o, rec_o = ora.render(view, 96, 72, shadow=True)
    assert np.array_equal(rec_d, rec_o)
    assert np.array_equal(rgba_d, rgba_o)


def test_weighted_colors_every_bpw_in_one_leaf():
    # one 16^3 color leaf painted in slabs with bpw 1, 2, 3 and RGB8 runs between them, so the
    # writer's bitset copies start at every bit alignment
    capi = _capi()
    cfg = small_cfg(64, 8, 6, 2, 4)
    rng = np.random.default_rng(3)
    recs = [W.edit(W.AABB_FILL, (16, 16, 16), (32, 32, 32), color=0x808080)]
    for z in range(16, 32):
        bpw = z % 4
        e = W.edit(W.AABB_PAINT, (16, 16 + (z % 3), z), (32, 32, z + 1), color=int(rng.integers(1 << 24)))
        if bpw:
            e["colors"] = int(rng.integers(0, 1 << 32, dtype=np.uint64))
            e["bpw"] = bpw
            e["weight"] = int(rng.integers(0, 1 << bpw))
        recs.append(e)
    e = W.edits(recs)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    for i in range(len(e)):
        dev.edit(e[i:i + 1])
        ora.edit(e[i:i + 1])
        assert_same_pool(dev, ora, f"slab {i}")
    vox = np.argwhere(np.ones((16, 16, 16)))[:, ::-1].astype(np.uint32) + 16
    got = O.color_fetch(dev.color_nodes(), dev.color_leaves(), dev.color_root, cfg.voxel_level,
                        cfg.color_leaf_level, np.ascontiguousarray(vox))
    assert len(np.unique(got.round(4), axis=0)) > 8


@pytest.mark.parametrize("seed", [5, 6])
def test_weighted_colors_batched_equal_sequential(seed):
    capi = _capi()
    cfg = small_cfg(64, 8, 6, 2, 3)
    e = weighted_edits(60, 64, seed)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    dev.edit(e, capi.MODE_BATCHED)
    ora.edit(e)
    assert_equivalent(dev, ora, cfg)


def test_wide_setnode_levels_plain_bit_exact():
    # color leaves of 4^3 voxels under a 512^3 grid: a sphere of radius 200 has tens of thousands
    # of entries on the levels above the color leaves, so SetNode runs its lower levels as grid
    # launches (decisions before the workgroup, placement after it) and its upper ones in LDS
    capi = _capi()
    res = 512
    cfg = small_cfg(res, 10, 9, 2, 2)
    c = res // 2
    e = W.edits([
        W.edit(W.SPHERE_FILL, (c, c, c), r2=200 * 200, color=0x2050A0),
        W.edit(W.SPHERE_PAINT, (c + 60, c - 40, c + 30), r2=150 * 150, color=0xA03020),
        W.edit(W.AABB_FILL, (40, 60, 20), (300, 280, 140), color=0x10C040),
        W.edit(W.SPHERE_DIG, (c - 50, c + 20, c + 90), r2=120 * 120),
        W.edit(W.SPHERE_PAINT, (c, c, c - 40), r2=220 * 220, color=0x808080),
        W.edit(W.AABB_PAINT, (100, 100, 100), (420, 400, 380), color=0x0000F0),
    ])
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    for i in range(len(e)):
        dev.edit(e[i:i + 1])
        ora.edit(e[i:i + 1])
        assert_same_pool(dev, ora, f"edit {i} kind {int(e[i]['kind'])}")
