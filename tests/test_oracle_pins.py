"""Pins of the CPU oracle (oracle/hd_oracle.hpp).

Each check is one of: the compiled reference (oracle/_ref, built from the VkHashDAG headers),
the committed golden fixtures those produced (tests/golden/*.npz), or what the mathematics fixes
(brute-force voxel predicates, the reduced-DAG invariants, known answers of the reference's own
test intents in test/test.cpp:118-209 and test/vbr_test.cpp:14-95)."""
import os

import numpy as np
import pytest

import _oracle as O
from vkhashdag_amd import workloads as W

needs_ref = pytest.mark.skipif(not O.have_ref(), reason="oracle/_ref not built (reference tree absent)")
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def small_cfg(res=64, bits=6, wbpp=5, pbpb=2, depth=3, **kw):
    return W.config_for_resolution(res, bucket_bits=bits, wbpp=wbpp, pbpb=pbpb, color_leaf_depth=depth, **kw)


def random_edits(n, res, seed, colored=True, kinds=None):
    rng = np.random.default_rng(seed)
    kinds = kinds or [W.SPHERE_FILL, W.SPHERE_DIG, W.SPHERE_PAINT, W.AABB_FILL, W.AABB_DIG, W.AABB_PAINT]
    out = []
    for _ in range(n):
        k = kinds[int(rng.integers(len(kinds)))]
        col = int(rng.integers(0, 1 << 24))
        if k in (W.SPHERE_FILL, W.SPHERE_DIG, W.SPHERE_PAINT):
            c = tuple(int(v) for v in rng.integers(0, res, 3))
            r = rng.uniform(1, res / 3)
            out.append(W.edit(k, c, r2=int(r * r), color=col, colored=colored and k != W.SPHERE_DIG))
        else:
            a = rng.integers(0, res, 3)
            b = np.minimum(a + rng.integers(1, res // 2, 3), res + 3)
            out.append(W.edit(k, tuple(int(v) for v in a), tuple(int(v) for v in b), color=col,
                              colored=colored and k != W.AABB_DIG))
    return W.edits(out)


def brute_force(e, res, heights=None):
    """Apply edits to a dense grid with the integer predicates of src/main.cpp:57-59,127-132."""
    g = np.zeros((res, res, res), bool)
    col = np.zeros((res, res, res), np.int64) - 1
    z, y, x = np.meshgrid(np.arange(res), np.arange(res), np.arange(res), indexing="ij")
    for r in np.atleast_1d(e):
        k = int(r["kind"])
        if k in (W.SPHERE_FILL, W.SPHERE_DIG, W.SPHERE_PAINT):
            c = [int(v) for v in r["p0"]]
            m = (x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2 <= int(r["r2"])
        elif k in (W.AABB_FILL, W.AABB_DIG, W.AABB_PAINT):
            a, b = [int(v) for v in r["p0"]], [int(v) for v in r["p1"]]
            m = (x >= a[0]) & (x < b[0]) & (y >= a[1]) & (y < b[1]) & (z >= a[2]) & (z < b[2])
        else:
            m = z < heights[y, x].astype(np.int64)
        if k in (W.SPHERE_DIG, W.AABB_DIG):
            g &= ~m
        elif k in (W.SPHERE_PAINT, W.AABB_PAINT):
            if r["colored"]:
                col[m & g] = int(r["colors"])
        else:
            g |= m
            if r["colored"] and k != W.COLUMN_FILL:
                col[m] = int(r["colors"])
    return g, col


# ------------------------------------------------------------------ hash
def test_hash_closed_form_known_answers():
    # fmix32 / fmix64 are bijections with fmix(0) == 0; an empty span hashes 0 ^ 0 then fmix32(0) = 0.
    assert O.olib().hdo_hash_leaf(0, 0) == 0
    assert O.olib().hdo_hash_span(np.zeros(1, np.uint32), 0) == 0
    g = np.load(os.path.join(GOLDEN, "hash.npz"))
    for i in range(len(g["leaf_in"])):
        assert O.olib().hdo_hash_leaf(int(g["leaf_in"][i, 0]), int(g["leaf_in"][i, 1])) == int(g["leaf_out"][i])
    for i, n in enumerate(g["span_len"]):
        w = np.ascontiguousarray(g["span_in"][i, :n])
        assert O.olib().hdo_hash_span(w if n else np.zeros(1, np.uint32), int(n)) == int(g["span_out"][i])


@needs_ref
def test_hash_matches_reference():
    rng = np.random.default_rng(7)
    for n in range(0, 10):
        for _ in range(200):
            w = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
            w = np.ascontiguousarray(w) if n else np.zeros(1, np.uint32)
            assert O.olib().hdo_hash_span(w, n) == O.rlib().hdr_hash_span(w, n)
    for _ in range(2000):
        a, b = (int(v) for v in rng.integers(0, 1 << 32, 2, dtype=np.uint64))
        assert O.olib().hdo_hash_leaf(a, b) == O.rlib().hdr_hash_leaf(a, b)


# ------------------------------------------------------------------ upsert known answers
def test_upsert_dedup_known_answers():
    cfg = small_cfg()
    p = O.OraclePool(cfg)
    n3 = np.array([0b101, 7, 9], np.uint32)
    a = p.upsert(2, n3)
    assert p.upsert(2, n3) == a  # re-insert -> same pointer (test.cpp:118-147)
    bw = p.bucket_words()
    assert bw.sum() == 3
    b = p.upsert(3, n3)  # same node, other level -> other pointer
    assert b != a
    n9 = np.array([0xFF] + list(range(8)), np.uint32)
    p.upsert(1, n9)
    assert p.bucket_words().sum() == 3 + 3 + 9
    p.upsert(cfg.node_levels - 1, np.array([1, 2], np.uint32), leaf=True)
    assert p.bucket_words().sum() == 3 + 3 + 9 + 2


def test_bucket_full_and_page_padding():
    # One bucket per level (bucket_bits = 0) puts every node in bucket 0 (the ZeroHasher case of
    # test.cpp:148-160): exactly floor(W/3) * pages 3-word nodes fit, each at ptr % W % 3 == 0.
    cfg = W.PoolConfig(4, 2, [0, 0, 0, 0], 1)
    p = O.OraclePool(cfg)
    wpp, pages = 16, 4
    ptrs = []
    for i in range(100):
        r = p.upsert(1, np.array([0b11, i, i + 1000], np.uint32), fallback=0xABCD)
        ptrs.append(r)
    fitted = [x for x in ptrs if x != 0xABCD]
    assert len(fitted) == (wpp // 3) * pages
    base = cfg.level_bases()[1] * cfg.words_per_bucket
    assert all(((x - base) % wpp) % 3 == 0 for x in fitted)
    assert all(x == 0xABCD for x in ptrs[len(fitted):])


# ------------------------------------------------------------------ edits vs reference
CASES = [
    dict(res=64, bits=6, wbpp=5, pbpb=2, depth=3, seed=1, n=24, colored=True),
    dict(res=64, bits=2, wbpp=4, pbpb=2, depth=2, seed=2, n=24, colored=True),   # heavy collisions, full buckets
    dict(res=128, bits=8, wbpp=9, pbpb=2, depth=4, seed=3, n=16, colored=True),
    dict(res=64, bits=6, wbpp=5, pbpb=2, depth=3, seed=4, n=24, colored=False),
    dict(res=32, bits=4, wbpp=4, pbpb=1, depth=5, seed=5, n=24, colored=True),   # color leaf at the root
]


@needs_ref
@pytest.mark.parametrize("case", CASES, ids=[f"res{c['res']}_b{c['bits']}_s{c['seed']}" for c in CASES])
def test_edit_sequence_bit_exact_vs_reference(case):
    cfg = small_cfg(case["res"], case["bits"], case["wbpp"], case["pbpb"], case["depth"])
    e = random_edits(case["n"], case["res"], case["seed"], case["colored"])
    o, r = O.OraclePool(cfg), O.RefPool(cfg)
    for i in range(len(e)):
        o.edit(e[i:i + 1])
        r.edit(e[i:i + 1])
        assert o.root == r.root and o.color_root == r.color_root, i
        bo, wo = o.live_words()
        br, wr = r.live_words()
        assert np.array_equal(bo, br) and np.array_equal(wo, wr), i
        assert np.array_equal(o.color_nodes(), r.color_nodes()), i
        assert np.array_equal(o.color_leaves(), r.color_leaves()), i
    assert np.array_equal(o.filled(), r.filled())


@needs_ref
def test_threaded_reference_edit_equals_oracle_canonically():
    cfg = small_cfg(128, 8, 9, 2, 4)
    e = random_edits(12, 128, 11, True)
    o, r = O.OraclePool(cfg), O.RefPool(cfg, workers=4)
    o.edit(e)
    r.edit(e, threaded=True)
    assert np.array_equal(o.canonical(), r.canonical())
    g = o.voxels()
    vox = np.argwhere(g)[:, ::-1].astype(np.uint32)
    assert np.array_equal(o.colors(vox), r.colors(vox))


@needs_ref
@pytest.mark.parametrize("seed", [1, 2])
def test_gc_bit_exact_vs_reference_one_worker(seed):
    cfg = small_cfg(64, 6, 5, 2, 3)
    e = random_edits(20, 64, seed, True)
    o, r = O.OraclePool(cfg), O.RefPool(cfg, workers=1)
    o.edit(e)
    r.edit(e)
    keep = o.root
    o.gc()
    r.gc()
    assert o.root == r.root
    bo, wo = o.live_words()
    br, wr = r.live_words()
    assert np.array_equal(bo, br) and np.array_equal(wo, wr)
    assert np.array_equal(o.filled(), r.filled())
    assert keep != O.HD_NULL


@needs_ref
def test_gc_reference_multiworker_canonical():
    cfg = small_cfg(64, 6, 5, 2, 3)
    e = random_edits(20, 64, 9, True)
    o, r = O.OraclePool(cfg), O.RefPool(cfg, workers=4)
    o.edit(e)
    r.edit(e)
    before = o.canonical()
    o.gc()
    r.gc()
    assert np.array_equal(o.canonical(), before)
    assert np.array_equal(r.canonical(), before)


@pytest.mark.parametrize("seed", [7, 8])
def test_gc_idempotent(seed):
    """GC of a compacted pool changes nothing: leaves are already packed in pointer order and inner
    nodes re-hash to the buckets and slots they occupy (a wrong processing order or padding rule
    would move them on the second pass)."""
    cfg = small_cfg(64, 6, 5, 2, 3)
    o = O.OraclePool(cfg)
    o.edit(random_edits(30, 64, seed, True))
    o.gc()
    w1, r1 = o.full_words().copy(), o.root
    o.gc()
    assert o.root == r1
    assert np.array_equal(o.full_words(), w1)


# ------------------------------------------------------------------ mathematics
@pytest.mark.parametrize("colored", [True, False])
def test_voxels_and_colors_equal_brute_force(colored):
    cfg = small_cfg(64, 6, 5, 2, 3)
    e = random_edits(30, 64, 21, colored)
    o = O.OraclePool(cfg)
    o.edit(e)
    g, col = brute_force(e, 64)
    assert np.array_equal(o.voxels().astype(bool), g)
    if colored:
        vox = np.argwhere(g)
        got = o.colors(vox[:, ::-1].astype(np.uint32))
        want = col[g]
        rgb = np.stack([(want >> s) & 0xFF for s in (0, 8, 16)], 1).astype(np.float32) / np.float32(255.0)
        known = want >= 0
        assert np.array_equal(got[known], rgb[known])


def test_reduced_dag_invariants():
    # Idempotence and no-op edits return the identical root (test.cpp:161-199 intent).
    cfg = small_cfg(64, 6, 5, 2, 3)
    o = O.OraclePool(cfg)
    box = W.edit(W.AABB_FILL, (3, 5, 7), (41, 33, 60), color=0x123456)
    r1 = o.edit(box)
    r2 = o.edit(box)
    assert r1 == r2
    noop = W.edit(W.SPHERE_DIG, (1000, 1000, 1000), r2=4)
    assert o.edit(noop) == r1
    # A full fill collapses onto the filled chain root.
    o2 = O.OraclePool(cfg)
    r = o2.edit(W.edit(W.AABB_FILL, (0, 0, 0), (64, 64, 64), color=0x1))
    assert r == o2.filled()[0]


def test_column_fill_equals_brute_force():
    res = 64
    h = W.value_noise_heightmap(res, 3, 5, 10, 50)
    cfg = small_cfg(res, 6, 5, 2, 3)
    o = O.OraclePool(cfg, heights=h)
    o.edit(W.terrain_edit(res, seed=3, skin=4))
    g, _ = brute_force(W.terrain_edit(res, seed=3, skin=4), res, h)
    assert np.array_equal(o.voxels().astype(bool), g)


@needs_ref
def test_column_fill_vs_reference():
    res = 128
    h = W.value_noise_heightmap(res, 4, 5, 20, 100)
    cfg = small_cfg(res, 8, 9, 2, 4)
    o, r = O.OraclePool(cfg, heights=h), O.RefPool(cfg, heights=h)
    e = W.edits([W.terrain_edit(res, seed=3, skin=6)]
                + list(W.terrain_sphere_edits(8, res, h, 2, 20, seed=4)))
    for i in range(len(e)):
        o.edit(e[i:i + 1])
        r.edit(e[i:i + 1])
        assert o.root == r.root and o.color_root == r.color_root
    assert np.array_equal(o.live_words()[1], r.live_words()[1])
    assert np.array_equal(o.color_leaves(), r.color_leaves())


# ------------------------------------------------------------------ VBR writer
def _random_ops(rng, n_vox, allow_weights):
    ops, left = [], n_vox
    while left:
        k = int(rng.integers(0, 3))
        cnt = int(min(left, rng.integers(1, 3000) if k < 2 else rng.integers(1, 40)))
        bpw = int(rng.integers(0, 4)) if allow_weights else 0
        col = int(rng.integers(0, 1 << 32)) if bpw else int(rng.integers(0, 4))
        w = int(rng.integers(0, 1 << bpw)) if bpw else 0
        has = int(rng.uniform() < 0.8)
        ops.append((k, cnt, col, w | (bpw << 8) | (has << 16)))
        left -= cnt
    return np.array(ops, np.uint32)


@needs_ref
@pytest.mark.parametrize("weights", [False, True])
def test_vbr_writer_bit_exact_vs_reference(weights):
    rng = np.random.default_rng(5 if weights else 6)
    for trial in range(40):
        n_vox = int(rng.choice([64, 512, 4096, 32768, 40000]))
        src = None
        if trial % 2:
            src = O.vbr_write(_random_ops(rng, n_vox, weights), None, "ref")
        ops = _random_ops(rng, n_vox, weights)
        a = O.vbr_write(ops, src, "oracle")
        b = O.vbr_write(ops, src, "ref")
        assert np.array_equal(a, b), trial


def test_vbr_writer_golden():
    g = np.load(os.path.join(GOLDEN, "vbr.npz"))
    for i in range(int(g["n"])):
        ops = g[f"ops{i}"]
        src = g[f"src{i}"] if f"src{i}" in g.files else None
        assert np.array_equal(O.vbr_write(ops, src, "oracle"), g[f"out{i}"]), i


def test_vbr_bitset_push_equals_single_pushes():
    # vbr_test.cpp:14-95 intent: Push(w, b, n) equals n single pushes (Push via one Push op vs
    # n one-voxel Edit ops that set the same color), for b in 1..3.
    for bpw in (1, 2, 3):
        for w in range(1 << bpw):
            for n in (1, 5, 31, 32, 33, 100, 1000):
                col = 0x12345678
                a = O.vbr_write(np.array([(1, n, col, w | (bpw << 8) | (1 << 16))], np.uint32))
                b = O.vbr_write(np.array([(2, n, col, w | (bpw << 8) | (1 << 16))], np.uint32))
                assert np.array_equal(a, b)


# ------------------------------------------------------------------ traversal
def _traversal_case():
    g = np.load(os.path.join(GOLDEN, "traversal64.npz"))
    return g["words"], int(g["root"]), g["o"], g["d"], g["hits"]


def test_march_no_lod_matches_reference_traversal_golden():
    # NodePoolTraversal::Traversal<float> (no LOD) hits inside the voxel DAG_RayMarch returns,
    # and both agree on hit / miss.
    words, root, o, d, ref = _traversal_case()
    L, res = 5, 64
    rec = O.march_rays(words, L, root, o, d)
    hit = (rec[:, 3] >> 31).astype(bool)
    assert np.array_equal(hit, ref[:, 3] > 0)
    vox = rec[hit, :3].astype(np.float64)
    size = (rec[hit, 3] & 0x7FFFFFFF).astype(np.float64)
    assert np.all(size == 1)
    p = ref[hit, :3].astype(np.float64) * res
    assert np.all(p >= vox - 1e-3) and np.all(p <= vox + size[:, None] + 1e-3)


def test_march_matches_dda():
    words, root, o, d, _ = _traversal_case()
    L, res = 5, 64
    grid = O.decode_voxels(words, L, root)
    rec = O.march_rays(words, L, root, o, d)
    dd = O.dda(grid, o, d)
    hit = (rec[:, 3] >> 31).astype(bool)
    agree_hit = hit == (dd[:, 3] > 0)
    same_vox = np.all(rec[:, :3] == dd[:, :3], axis=1) | ~hit
    ok = agree_hit & same_vox
    # fp32 ESVO vs fp64 DDA can only disagree on rays grazing a voxel edge; none of these do.
    assert ok.mean() == 1.0, np.nonzero(~ok)[0][:10]


@pytest.mark.parametrize("seed", [0, 1])
def test_render_frame_hits_match_dda(seed):
    # Full-frame transliteration (Camera_GenRay + DAG_RayMarch, LOD off) vs DDA on the decoded grid.
    res = 64
    cfg = small_cfg(res, 6, 5, 2, 3)
    p = O.OraclePool(cfg)
    p.edit(random_edits(20, res, 40 + seed, True))
    view = W.orbit_view((32, 32, 32), 60, seed, 96, 64, res=res, lod=False)
    words = p.full_words()
    rgba, rec = O.render(words, cfg.node_levels, p.root, p.color_nodes(), p.color_leaves(), p.color_root,
                         cfg.color_leaf_level, view, 96, 64)
    grid = O.decode_voxels(words, cfg.node_levels, p.root)
    # rebuild the same rays in float32 exactly as Camera_GenRay does
    W_, H_ = 96, 64
    px, py = np.meshgrid(np.arange(W_, dtype=np.float32), np.arange(H_, dtype=np.float32))
    cx = ((px + np.float32(0.5)) / np.float32(W_)) * np.float32(2) - np.float32(1)
    cy = ((py + np.float32(0.5)) / np.float32(H_)) * np.float32(2) - np.float32(1)
    look, side, up = view[3:6], view[6:9], view[9:12]
    d = np.stack([look[i] - side[i] * cx - up[i] * cy for i in range(3)], -1).reshape(-1, 3).astype(np.float32)
    d = d / np.sqrt((d * d).sum(1, keepdims=True, dtype=np.float32)).astype(np.float32)
    o = np.broadcast_to(view[:3], d.shape).astype(np.float32)
    dd = O.dda(grid, o, d)
    hit = (rec[:, 3] >> 31).astype(bool)
    ok = (hit == (dd[:, 3] > 0)) & (np.all(rec[:, :3] == dd[:, :3], axis=1) | ~hit)
    assert ok.mean() >= 0.999, ok.mean()


def test_edit_and_gc_golden_fixture():
    # The reference's single-threaded Edit and one-worker ThreadedGC outputs, committed.
    g = np.load(os.path.join(GOLDEN, "edits64.npz"))
    e = g["edits"].view(W.EDIT_DTYPE)
    cfg = small_cfg(64, 6, 5, 2, 3)
    o = O.OraclePool(cfg)
    for i in range(len(e)):
        o.edit(e[i:i + 1])
        assert (o.root, o.color_root) == tuple(int(v) for v in g["roots"][i])
    bw, lw = o.live_words()
    assert np.array_equal(bw, g["bucket_words"]) and np.array_equal(lw, g["live_words"])
    assert np.array_equal(o.color_nodes(), g["color_nodes"]) and np.array_equal(o.color_leaves(), g["color_leaves"])
    assert np.array_equal(o.filled(), g["filled"])
    o.gc()
    assert o.root == int(g["gc_root"])
    bw, lw = o.live_words()
    assert np.array_equal(bw, g["gc_bucket_words"]) and np.array_equal(lw, g["gc_live_words"])


def _np_block_hash(seed, bx, by, bz):
    M = np.uint64(0xFFFFFFFFFFFFFFFF)
    k = np.uint64(seed) ^ (bx.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ (
        by.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)) ^ (bz.astype(np.uint64) * np.uint64(0x165667B19E3779F9))
    k ^= k >> np.uint64(33)
    k = (k * np.uint64(0xff51afd7ed558ccd)) & M
    k ^= k >> np.uint64(33)
    k = (k * np.uint64(0xc4ceb9fe1a85ec53)) & M
    k ^= k >> np.uint64(33)
    return (k & np.uint64(0xFFFFFFFF)).astype(np.int64)


def test_terrain_colors_follow_the_block_rule():
    # Column fill on an empty pool: every occupied voxel decodes to the per-voxel rule --
    # deep color below the skin, else the block's height ramp + [-8, 8] noise.
    res, skin, seed, deep = 64, 4, 11, 0x405060
    h = W.value_noise_heightmap(res, 3, 3, 8, 56)
    cfg = small_cfg(res, 6, 5, 2, 3)
    o = O.OraclePool(cfg, heights=h)
    o.edit(W.terrain_edit(res, seed=seed, deep_color=deep, skin=skin))
    g = o.voxels().astype(bool)
    vz, vy, vx = np.nonzero(g)
    bx, by, bz = vx >> 2, vy >> 2, vz >> 2
    hmin_b = h.reshape(res // 4, 4, res // 4, 4).min(axis=(1, 3))[by, bx].astype(np.int64)
    is_deep = bz * 4 + 4 + skin <= hmin_b
    t = np.minimum((bz * 4 * 255) // (res - 1), 255)
    base = np.stack([50 + t * 170 // 255, 110 + t * 120 // 255, 40 + t * 200 // 255], 1)
    hh = _np_block_hash(seed, bx, by, bz)
    noise = np.stack([((hh >> (8 * c)) & 0xFF) % 17 - 8 for c in range(3)], 1)
    rgb = np.clip(base + noise, 0, 255)
    rgb = np.where(is_deep[:, None], np.array([deep & 0xFF, (deep >> 8) & 0xFF, (deep >> 16) & 0xFF]), rgb)
    got = o.colors(np.stack([vx, vy, vz], 1).astype(np.uint32))
    assert np.array_equal(got, rgb.astype(np.float32) / np.float32(255.0))
    assert len(np.unique(rgb, axis=0)) > 20


def _beam_scene():
    res = 256
    h = W.value_noise_heightmap(res, 5, 4, 40, 200)
    cfg = small_cfg(res, 10, 9, 2, 4)
    ora = O.OraclePool(cfg, heights=h)
    ora.edit(W.edits([W.terrain_edit(res, seed=4, skin=8)]))
    ora.edit(W.terrain_sphere_edits(12, res, h, 3, 30, seed=2))
    return res, cfg, ora


def test_beam_prepass_is_conservative():
    """beam.frag + trace.frag:384-390 restated: a beam-started ray never loses a hit the plain ray
    finds, hits the same voxel where both hit, and only adds hits past the plain ray's t <= 1 cap
    (the shifted origin moves the cap)."""
    res, cfg, ora = _beam_scene()
    words = ora.full_words()
    Wd, Ht = 120, 72
    for seed in range(2):
        view = W.orbit_view((res / 2, res / 2, res * 0.3), res * 0.9, seed, Wd, Ht, res=res, elevation=(0.3, 0.6))
        args = (words, cfg.node_levels, ora.root, ora.color_nodes(), ora.color_leaves(), ora.color_root,
                cfg.color_leaf_level, view, Wd, Ht)
        _, ra = O.render(*args, shadow=False)
        _, rb = O.render(*args, shadow=False, beam=True)
        ha, hb = (ra[:, 3] >> 31).astype(bool), (rb[:, 3] >> 31).astype(bool)
        assert not np.any(ha & ~hb), "the beam lost a hit"
        both = ha & hb
        same = np.all(ra[both, :3] == rb[both, :3], axis=1)
        assert same.mean() > 0.995, same.mean()
        extra = ~ha & hb
        # distance to the hit voxel's centre (the recorded hit position of a voxel on the cube's far
        # face is clamped to 0 by trace.frag's vox_max, whose bit OR overflows into the exponent)
        size = (rb[extra, 3] & 0x7FFFFFFF).astype(np.float64)
        centre = (rb[extra, :3] + size[:, None] / 2) / res
        assert np.all(np.linalg.norm(centre - view[:3], axis=1) > 0.95), "an extra beam hit inside the plain range"
        # the beam image: a finite block value never exceeds the plain hit distance of its pixels
        B = O.beam_image(words, cfg.node_levels, ora.root, view, Wd, Ht)
        assert np.all(B[np.isfinite(B)] >= 0.0)


# ------------------------------------------------------------------ shadow rays, LOD cut, shading
def _shading_scene(lod, seed=0, Wd=160, Ht=96):
    res = 128
    h = W.value_noise_heightmap(res, 4, 9, 20, 90)
    cfg = small_cfg(res, 8, 9, 2, 3)
    p = O.OraclePool(cfg, heights=h)
    p.edit(W.edits([W.terrain_edit(res, seed=3, skin=6)] + list(W.terrain_sphere_edits(10, res, h, 3, 20, seed=8))))
    # cameras on the light's side (+x, +y, above), so the frame holds lit and shadowed pixels
    theta, phi = (0.8, 0.35, 1.25)[seed], (0.7, 0.5, 0.9)[seed]
    c = np.array([64.0, 64.0, 40.0])
    pos = c + 150 * np.array([np.cos(theta) * np.cos(phi), np.sin(theta) * np.cos(phi), np.sin(phi)])
    look, side, up = W.look_side_up(c - pos, aspect=Wd / Ht)
    proj = Ht / (2.0 * np.tan(np.pi / 6)) if lod else float("inf")
    view = W.view_block(pos / res, look, side, up, proj)
    words = p.full_words()
    rgba, rec = O.render(words, cfg.node_levels, p.root, p.color_nodes(), p.color_leaves(), p.color_root,
                         cfg.color_leaf_level, view, Wd, Ht, shadow=True)
    return p, cfg, words, view, rgba, rec


def _light():
    light = np.array([4, 5, 3], np.float32)
    return (light / np.sqrt(np.float32((light * light).sum(dtype=np.float32)))).astype(np.float32)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_shadow_flags_match_dda(seed):
    # The shadow ray (LOD off) starts at the hit position trace.frag:248-258 computes -- epsilon off
    # the hit face -- toward normalize(4, 5, 3), the shading light of trace.frag:396. A brute-force
    # DDA on the decoded grid from the same point gives the same occluded / lit answer.
    p, cfg, words, view, rgba, rec = _shading_scene(False, seed)
    hit = (rec[:, 3] >> 31).astype(bool)
    o = np.ascontiguousarray(rec[hit, 4:7].view(np.float32))
    d = np.ascontiguousarray(np.broadcast_to(_light(), o.shape))
    grid = O.decode_voxels(words, cfg.node_levels, p.root)
    dd = O.dda(grid, o, d)
    shadow = rec[hit, 11].astype(bool)
    assert 0.05 < shadow.mean() < 0.95, shadow.mean()
    agree = shadow == (dd[:, 3] > 0)
    # fp32 march vs fp64 DDA can only differ on rays grazing a voxel edge
    assert agree.mean() >= 0.998, (agree.mean(), np.nonzero(~agree)[0][:8])
    # a wrong light (a transposed component) disagrees on many pixels: the check has teeth
    bad = O.dda(grid, o, np.ascontiguousarray(np.broadcast_to(_light()[[0, 2, 1]], o.shape)))
    assert (shadow == (bad[:, 3] > 0)).mean() < agree.mean() - 0.01


def _slab(o, d, lo, hi):
    """t range of the ray o + t d inside the box [lo, hi] (float64)."""
    with np.errstate(divide="ignore", invalid="ignore"):
        a = (lo - o) / d
        b = (hi - o) / d
    t0 = np.nanmax(np.minimum(a, b), axis=1)
    t1 = np.nanmin(np.maximum(a, b), axis=1)
    return t0, t1


def test_lod_stop_cube_properties():
    # With the LOD cut-off on (trace.frag:148), the march stops at a node cube the ray enters: that
    # cube holds at least one voxel of the decoded grid, the ray enters it no later than the first
    # voxel a brute-force DDA hits, and the stop rule held there: cube size x proj_factor is below
    # the ray's exit distance from the cube.
    # a small frame: a voxel projects to ~0.4 pixel, so the cut stops at 2- and 4-voxel nodes
    Wd, Ht = 96, 56
    p, cfg, words, view, rgba, rec = _shading_scene(True, 0, Wd, Ht)
    res = cfg.resolution
    hit = (rec[:, 3] >> 31).astype(bool)
    size = (rec[:, 3] & 0x7FFFFFFF).astype(np.int64)
    assert (size[hit] > 1).mean() > 0.05, "the LOD cut never engaged"
    grid = O.decode_voxels(words, cfg.node_levels, p.root)
    S = np.concatenate([[0], np.cumsum(np.cumsum(np.cumsum(grid.astype(np.int64), 0), 1), 2).ravel()])
    cs = np.zeros((res + 1,) * 3, np.int64)
    cs[1:, 1:, 1:] = S[1:].reshape(res, res, res)

    def occupied(v, s):  # voxels of the grid ([z, y, x]) inside the cube [v, v + s)
        x0, y0, z0 = v[:, 0], v[:, 1], v[:, 2]
        x1, y1, z1 = x0 + s, y0 + s, z0 + s
        return (cs[z1, y1, x1] - cs[z0, y1, x1] - cs[z1, y0, x1] - cs[z1, y1, x0] + cs[z0, y0, x1]
                + cs[z0, y1, x0] + cs[z1, y0, x0] - cs[z0, y0, x0])

    v = rec[hit, :3].astype(np.int64)
    s = size[hit]
    assert np.all(occupied(v, s) > 0)
    # rays as Camera_GenRay builds them (float32), in world units [0,1]^3
    px, py = np.meshgrid(np.arange(Wd, dtype=np.float32), np.arange(Ht, dtype=np.float32))
    cx = ((px + np.float32(0.5)) / np.float32(Wd)) * np.float32(2) - np.float32(1)
    cy = ((py + np.float32(0.5)) / np.float32(Ht)) * np.float32(2) - np.float32(1)
    look, side, up = view[3:6], view[6:9], view[9:12]
    d = np.stack([look[i] - side[i] * cx - up[i] * cy for i in range(3)], -1).reshape(-1, 3).astype(np.float32)
    d = d / np.sqrt((d * d).sum(1, keepdims=True, dtype=np.float32)).astype(np.float32)
    o = np.broadcast_to(view[:3], d.shape).astype(np.float32)
    t0, t1 = _slab(o[hit].astype(np.float64), d[hit].astype(np.float64), v / res, (v + s[:, None]) / res)
    assert np.all(t0 <= t1 + 1e-6)
    dd = O.dda(grid, np.ascontiguousarray(o[hit]), np.ascontiguousarray(d[hit]))
    g = dd[:, 3] > 0
    tv0, _ = _slab(o[hit][g].astype(np.float64), d[hit][g].astype(np.float64), dd[g, :3] / res,
                   (dd[g, :3] + 1) / res)
    assert np.all(t0[g] <= tv0 + 1e-5)
    # stop rule: scale_exp2 * proj_factor < tc_max (+ bias 0), scale in the [1,2] cube = size / res
    proj = np.float64(view[12])
    assert np.all((s / res) * proj < t1 * (1 + 1e-5))


@pytest.mark.parametrize("lod", [False, True])
def test_shading_equals_numpy_formula(lod):
    # trace.frag:395-396 restated in numpy from the hit records: diffuse = max(dot(n, L), 0) * .5 + .5
    # with L = normalize(4, 5, 3) (0.5 where the shadow ray is blocked), times Color_Fetch of the hit
    # voxel, clamped and rounded to RGBA8; misses are black with alpha 255.
    p, cfg, words, view, rgba, rec = _shading_scene(lod, 1)
    hit = (rec[:, 3] >> 31).astype(bool)
    L = _light()
    n = rec[:, 7:10].view(np.float32)
    nd = (n[:, 0] * L[0] + n[:, 1] * L[1]).astype(np.float32) + n[:, 2] * L[2]
    diffuse = np.where(rec[:, 11] > 0, np.float32(0.5),
                       np.maximum(nd, np.float32(0)) * np.float32(0.5) + np.float32(0.5)).astype(np.float32)
    col = np.zeros((len(rec), 3), np.float32)
    col[hit] = O.color_fetch(p.color_nodes(), p.color_leaves(), p.color_root, cfg.voxel_level, cfg.color_leaf_level,
                             np.ascontiguousarray(rec[hit, :3]))
    c = np.clip(diffuse[:, None] * col, 0, 1) * np.float32(255) + np.float32(0.5)
    q = np.where(hit[:, None], np.floor(c), 0).astype(np.uint32)
    want = q[:, 0] | (q[:, 1] << 8) | (q[:, 2] << 16) | np.uint32(255 << 24)
    assert hit.mean() > 0.2 and (rec[hit, 11] > 0).any()
    assert np.array_equal(rgba, want.astype(np.uint32))


def test_voxel_query_and_pool_render_equal_array_checkers():
    # the pool-handle checkers used at full size (voxel queries, threaded frames, bulk live words,
    # canonical from the pool) equal the array checkers on a small scene
    res = 64
    cfg = small_cfg(res, 6, 5, 2, 3)
    p = O.OraclePool(cfg)
    p.edit(random_edits(24, res, 77, True))
    words = p.full_words()
    grid = O.decode_voxels(words, cfg.node_levels, p.root)
    vox = np.argwhere(np.ones((res, res, res), bool))[:, ::-1].astype(np.uint32)
    assert np.array_equal(p.voxel_query(vox), grid.reshape(-1))
    view = W.orbit_view((32, 32, 32), 70, 2, 80, 60, res=res)
    a = O.render(words, cfg.node_levels, p.root, p.color_nodes(), p.color_leaves(), p.color_root,
                 cfg.color_leaf_level, view, 80, 60)
    b = p.render(view, 80, 60, threads=4)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    bw = p.bucket_words()
    parts = [p.read_words(int(i) * cfg.words_per_bucket, int(bw[i])) for i in np.nonzero(bw)[0]]
    assert np.array_equal(p.live_words()[1], np.concatenate(parts))
    assert np.array_equal(p.canonical(), O.canonical_words(words, cfg.node_levels, np.array([p.root], np.uint32)))


def test_edit_traffic_known_answers():
    # Closed forms on single-voxel edits (L node levels, the filled chain created beforehand):
    #  * one voxel into an empty pool: no input node is read; one new node per level (the 2-word
    #    leaf and L - 1 one-child inner nodes of 2 words), each hashed and written once, one
    #    counter word each;
    #  * the same edit again: the recursion reads the 2-word path node at every level and changes
    #    nothing (no upsert);
    #  * a second voxel of the same leaf as a batch of one: the path is read once and L new nodes
    #    appended (leaf 2 words, inner nodes still one child).
    cfg = small_cfg(64, 6, 5, 2, 3)
    L = cfg.node_levels
    p = O.OraclePool(cfg)
    O.olib().hdo_make_filled(p.h)
    one = W.edit(W.AABB_FILL, (9, 9, 9), (10, 10, 10))
    t = p.edit_traffic(one)
    assert (t["visited"], t["hashed"], t["written"], t["counters"], t["upserts"]) == (0, 2 * L, 2 * L, L, L)
    assert t["scanned"] <= 9 * L  # only the filled chain's nodes can share a bucket
    t = p.edit_traffic(one)
    assert (t["visited"], t["hashed"], t["scanned"], t["written"], t["counters"], t["upserts"]) == (2 * L, 0, 0, 0, 0, 0)
    t = p.edit_traffic(W.edit(W.AABB_FILL, (8, 9, 9), (9, 10, 10)), batch=True)
    assert (t["visited"], t["hashed"], t["written"], t["counters"], t["upserts"]) == (2 * L, 2 * L, 2 * L, L, L)
    # batch counting never exceeds the per-edit sums of the same edits from the same state
    q1, q2 = O.OraclePool(cfg), O.OraclePool(cfg)
    base = random_edits(6, 64, 3, True)
    q1.edit(base)
    q2.edit(base)
    more = random_edits(12, 64, 4, True)
    a, b = q1.edit_traffic(more), q2.edit_traffic(more, batch=True)
    assert q1.root == q2.root
    assert b["visited"] <= a["visited"] and b["written"] <= a["written"] and b["upserts"] <= a["upserts"]


@pytest.mark.parametrize("colored", [True, False])
def test_tiled_terrain_equals_whole_terrain(colored):
    # a column fill restricted to column rectangles [p0.xy, p1.xy): the tiles applied one after
    # another give the whole-grid fill's voxels (a reduced DAG: canonical equality) and colors
    res = 128
    h = W.value_noise_heightmap(res, 4, 6, 20, 100)
    cfg = small_cfg(res, 8, 9, 2, 3)
    whole, tiled = O.OraclePool(cfg, heights=h), O.OraclePool(cfg, heights=h)
    whole.edit(W.terrain_edit(res, seed=3, skin=6, colored=colored))
    t = W.terrain_tiles(res, 3, 4, skin=6, colored=colored)
    assert len(t) == 16 and int(t[5]["p0"][0]) == 32 and int(t[5]["p1"][1]) == 64
    for i in range(len(t)):
        tiled.edit(t[i:i + 1])
    assert np.array_equal(whole.canonical(), tiled.canonical())
    g = whole.voxels()
    assert g.sum() == sum(int(v) for v in h.astype(np.int64).ravel())
    if colored:
        vox = np.ascontiguousarray(np.argwhere(g)[:, ::-1].astype(np.uint32))
        assert np.array_equal(whole.colors(vox), tiled.colors(vox))
    # a rectangle covers only its columns
    part = O.OraclePool(cfg, heights=h)
    part.edit(W.terrain_edit(res, seed=3, skin=6, lo=(16, 40), hi=(80, 64), colored=colored))
    gp = part.voxels()
    z, y, x = np.meshgrid(np.arange(res), np.arange(res), np.arange(res), indexing="ij")
    want = (z < h[y, x].astype(np.int64)) & (x >= 16) & (x < 80) & (y >= 40) & (y < 64)
    assert np.array_equal(gp, want)
