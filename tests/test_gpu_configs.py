"""The benchmark workloads at their stated sizes: the device path against the CPU oracle.

Bar: bit-exact where the reference's result is defined word for word, canonical otherwise.
  * 1024^3 (2^16 buckets per level, 512-word pages): a solid colored sphere of radius 400 at the
    centre, then 64 SphereEditor<kDig> edits one at a time (centres U[0,1024)^3, r ~ U[8,64],
    seed 0). After every edit the roots, bucket counters, used words and color arrays equal the
    oracle's (= the reference's NodePoolBase::Edit, NodePool.hpp:405-417; main.cpp:73-150). A
    1024x768 frame from a seeded radius-900 orbit camera (LOD on, 1 shadow ray per hit) equals the
    oracle's transliteration of trace.frag in RGBA8 and in every hit record.
  * 4096^3 terrain (6-octave value noise, seed 1, z in [512, 3072]) built by a column fill on both
    sides, then 256 sphere add/remove edits (r log-uniform 4..128, adds colored) through the plain
    path: bit-exact pools after the terrain and every 32 edits, then the 1920x1080 frame (LOD on,
    shadows) bit-exact.
  * 4096 mixed edits (40% sphere add, 40% sphere remove, 20% box paint, r log-uniform 2..256,
    seed 2) on that terrain, batched on the device, sequential in the oracle: equal canonical
    DAGs (the reduced DAG of the voxel set), equal occupancy and colors on a sample of voxels in
    every edit's footprint and on the surface, and an equal LOD-off frame (its hits are voxels,
    whose colors are defined; an LOD hit samples the color of a node's corner voxel, which may be
    empty and is then history-dependent in the reference too)."""
import numpy as np
import pytest

import _oracle as O
from test_gpu_plain import assert_same_pool
from vkhashdag_amd import workloads as W

pytestmark = pytest.mark.gpu


def _capi():
    from vkhashdag_amd import capi
    return capi


def cfg_4096():
    # the benchmark's pool: 2^16 buckets per level, 512-word pages, 4 pages per bucket, 16^3-voxel
    # color leaves
    return W.config_for_resolution(4096, bucket_bits=16, wbpp=9, pbpb=2, color_leaf_depth=4,
                                   color_node_capacity=1 << 24, color_leaf_capacity=1 << 28)


@pytest.fixture(scope="module")
def heights_4096():
    return W.value_noise_heightmap(4096, 6, 1, 4096 // 8, (4096 * 3) // 4)


def _render_compare(dev, ora, view, w, h):
    rgba_d, rec_d = dev.render(view, w, h, shadow=True, records=True)
    rgba_o, rec_o = ora.render(view, w, h, shadow=True)
    bad = np.nonzero(np.any(rec_d != rec_o, axis=1))[0]
    assert len(bad) == 0, (len(bad), bad[:8], rec_d[bad[:2]], rec_o[bad[:2]])
    assert np.array_equal(rgba_d, rgba_o)
    return rec_o


def test_config1_sphere_then_64_digs_1024():
    capi = _capi()
    cfg = W.config_for_resolution(1024, bucket_bits=16, wbpp=9, pbpb=2, color_leaf_depth=4)
    assert cfg.node_levels == 9 and cfg.words_per_bucket == 2048
    print("config 1: 1024^3, sphere r=400 + 64 digs (seed 0)", flush=True)
    dev, ora = capi.DevicePool(cfg), O.OraclePool(cfg)
    e = W.edits([W.edit(W.SPHERE_FILL, (512, 512, 512), r2=400 * 400, color=0x3366CC)]
                + list(W.sphere_digs(64, 1024, 8, 64, seed=0)))
    assert int(e[1]["kind"]) == W.SPHERE_DIG and len(e) == 65
    for i in range(len(e)):
        dev.edit(e[i:i + 1])
        ora.edit(e[i:i + 1])
        assert_same_pool(dev, ora, f"config 1 edit {i}")
    assert ora.overflows() == 0
    for seed in (0, 1):
        view = W.orbit_view((512, 512, 512), 900, seed, 1024, 768, res=1024)
        rec = _render_compare(dev, ora, view, 1024, 768)
        assert (rec[:, 3] >> 31).mean() > 0.2


def test_config2_terrain_then_256_spheres_4096(heights_4096):
    capi = _capi()
    res, h = 4096, heights_4096
    cfg = cfg_4096()
    print("config 2: 4096^3 terrain + 256 sphere edits + 1080p frame", flush=True)
    dev, ora = capi.DevicePool(cfg, heights=h), O.OraclePool(cfg, heights=h)
    t = W.terrain_edit(res, seed=1, skin=8)
    dev.edit(t)
    ora.edit(t)
    assert_same_pool(dev, ora, "terrain")
    batch = W.terrain_sphere_edits(256, res, h, 4, 128, seed=1000)
    for i in range(0, 256, 32):
        dev.edit(batch[i:i + 32], capi.MODE_PLAIN)
        ora.edit(batch[i:i + 32])
        assert_same_pool(dev, ora, f"after {i + 32} sphere edits")
    assert ora.overflows() == 0
    view = W.orbit_view((res / 2, res / 2, res * 0.3), 0.9 * res, 7, 1920, 1080, res=res, elevation=(0.35, 0.6))
    rec = _render_compare(dev, ora, view, 1920, 1080)
    assert (rec[:, 3] >> 31).mean() > 0.3 and rec[:, 11].sum() > 0


def test_config3_batched_4096_mixed_equals_sequential(heights_4096):
    capi = _capi()
    res, h = 4096, heights_4096
    cfg = cfg_4096()
    print("config 3: 4096 mixed edits, batched vs sequential", flush=True)
    dev, ora = capi.DevicePool(cfg, heights=h), O.OraclePool(cfg, heights=h)
    t = W.terrain_edit(res, seed=1, skin=8)
    dev.edit(t)
    ora.edit(t)
    batch = W.mixed_batch(4096, res, h, 2, 256, seed=2)
    dev.edit(batch, capi.MODE_BATCHED)
    assert 1 <= dev.stats().last_groups < len(batch)  # folded: far fewer passes than edits
    ora.edit(batch)
    assert ora.overflows() == 0 and dev.stats().overflow == 0
    mirror = O.OraclePool(cfg)  # the device's words, read by the oracle's checkers
    mirror.load_state(dev.state())
    cd, co = mirror.canonical(), ora.canonical()
    assert len(cd) == len(co) and np.array_equal(cd, co), "canonical DAGs differ"

    # occupancy and colors on voxels in every edit's footprint and on the terrain surface
    rng = np.random.default_rng(5)
    pts = []
    for r in batch:
        k = int(r["kind"])
        if k == W.AABB_PAINT:
            lo, hi = np.asarray(r["p0"], np.int64), np.asarray(r["p1"], np.int64)
            pts.append(rng.integers(lo, np.maximum(hi, lo + 1), size=(24, 3)))
        else:
            rad = int(np.sqrt(float(r["r2"]))) + 1
            pts.append(np.asarray(r["p0"], np.int64) + rng.integers(-rad, rad + 1, size=(24, 3)))
    xy = rng.integers(0, res, size=(40000, 2))
    z = h[xy[:, 1], xy[:, 0]].astype(np.int64)[:, None] + rng.integers(-3, 2, size=(40000, 1))
    pts.append(np.concatenate([xy, z], axis=1))
    vox = np.clip(np.concatenate(pts), 0, res - 1).astype(np.uint32)
    occ_d, occ_o = mirror.voxel_query(vox), ora.voxel_query(vox)
    assert np.array_equal(occ_d, occ_o)
    vox = np.ascontiguousarray(vox[occ_o.astype(bool)])
    assert len(vox) > 50000
    got = O.color_fetch(dev.color_nodes(), dev.color_leaves(), dev.color_root, cfg.voxel_level,
                        cfg.color_leaf_level, vox)
    want = ora.colors(vox)
    bad = np.nonzero(np.any(got != want, axis=1))[0]
    assert len(bad) == 0, (len(bad), vox[bad[:5]], got[bad[:5]], want[bad[:5]])

    view = W.orbit_view((res / 2, res / 2, res * 0.3), 0.9 * res, 11, 960, 540, res=res, lod=False,
                        elevation=(0.35, 0.6))
    rgba_d = dev.render(view, 960, 540, shadow=True)
    rgba_o, rec_o = ora.render(view, 960, 540, shadow=True)
    assert (rec_o[:, 3] >> 31).mean() > 0.3
    assert np.array_equal(rgba_d, rgba_o)


@pytest.mark.parametrize("res,octaves,seed,zlo,zhi", [(256, 4, 5, 20, 200), (1024, 7, 3, 100, 900),
                                                      (4096, 6, 1, 512, 3072)])
def test_device_heightmap_equals_host_generator(res, octaves, seed, zlo, zhi, heights_4096):
    # the larger terrains (16384^2, 32768^2 heightfields) are generated on the device; the generator
    # must give exactly the host recipe's heights (same float64 operations in the same order)
    capi = _capi()
    cfg = W.config_for_resolution(res, bucket_bits=8, wbpp=9, pbpb=2, color_leaf_depth=4)
    want = heights_4096 if res == 4096 else W.value_noise_heightmap(res, octaves, seed, zlo, zhi)
    dev = capi.DevicePool(cfg)
    got = dev.value_noise_heightmap(octaves, seed, zlo, zhi)
    bad = np.argwhere(got != want)
    assert len(bad) == 0, (len(bad), bad[:4], got[tuple(bad[0])], want[tuple(bad[0])])
    # and the column fill over it equals the fill over the uploaded heights
    if res <= 1024:
        ref = capi.DevicePool(cfg, heights=want)
        t = W.terrain_edit(res, seed=seed, skin=6)
        dev.edit(t)
        ref.edit(t)
        assert (dev.root, dev.color_root) == (ref.root, ref.color_root)
        assert np.array_equal(dev.live_words()[1], ref.live_words()[1])


def test_tiled_terrain_build_bit_exact():
    # the large terrains are built tile by tile (column rectangles); device == oracle word for word
    capi = _capi()
    res = 256
    h = W.value_noise_heightmap(res, 5, 7, 30, 200)
    cfg = W.config_for_resolution(res, bucket_bits=10, wbpp=9, pbpb=2, color_leaf_depth=4)
    dev, ora = capi.DevicePool(cfg, heights=h), O.OraclePool(cfg, heights=h)
    t = W.terrain_tiles(res, 5, 4, skin=8)
    for i in range(len(t)):
        dev.edit(t[i:i + 1])
        ora.edit(t[i:i + 1])
        assert_same_pool(dev, ora, f"tile {i}")
