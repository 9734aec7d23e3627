"""Serving loop (hd_serve_run): each result is the same alone and in the mix.

Bar: renders are bit-exact against rendering the same view alone on the DAG version the server
reports for it; after every tick the DAG equals, up to node renumbering, the oracle applying that
tick's edits one at a time in submission order on top of the previous tick (occupied-voxel colors
equal too). The one-request-at-a-time loop obeys the same bar, and both loops end on the same DAG.
"""
import numpy as np
import pytest

import _oracle as O
from test_oracle_pins import small_cfg
from vkhashdag_amd import workloads as W

pytestmark = pytest.mark.gpu


def _capi():
    from vkhashdag_amd import capi
    return capi


def _setup(res=128, seed=3):
    capi = _capi()
    h = W.value_noise_heightmap(res, 4, seed, res // 6, res * 3 // 4)
    cfg = small_cfg(res, 9, 9, 2, 3, keep_history=True)
    dev = capi.DevicePool(cfg, heights=h)
    dev.edit(W.terrain_edit(res, seed=seed, skin=6))
    return capi, cfg, h, dev


def _check_renders(capi, dev, reqs, res, buf):
    for i in np.nonzero(reqs["kind"] == W.REQ_RENDER)[0]:
        v = reqs[i]["view"]
        w, hh = int(v["width"]), int(v["height"])
        vb = np.concatenate([v["pos"], v["look"], v["side"], v["up"], [v["proj_factor"]]]).astype(np.float32)
        alone = dev.render(vb, w, hh, shadow=bool(v["flags"] & 1), root=int(res[i]["root"]),
                           color_root=int(res[i]["color_root"]))
        off = int(res[i]["pixel_offset"])
        served = buf.download(np.uint32, off + w * hh)[off:]
        assert np.array_equal(served, alone), f"render {i} differs from rendering it alone"


def _check_ticks(cfg, h, dev, reqs, res):
    """Replays the edits tick by tick on the oracle and compares canonical DAGs + colors."""
    ora = O.OraclePool(cfg, heights=h)
    ora.edit(W.edits([W.terrain_edit(cfg.resolution, seed=3, skin=6)]))
    edit_ids = np.nonzero(reqs["kind"] == W.REQ_EDIT)[0]
    words = dev.full_words()
    L = cfg.node_levels
    for t in sorted(set(int(res[i]["tick"]) for i in edit_ids)):
        ids = [i for i in edit_ids if int(res[i]["tick"]) == t]
        for i in ids:  # submission order within the tick
            ora.edit(reqs[i]["edit"][None])
        root = int(res[ids[0]]["root"])
        cd = O.canonical_words(words, L, np.array([root], np.uint32))
        assert np.array_equal(cd, ora.canonical()), f"tick {t}: DAG differs from sequential application"
        grid = O.decode_voxels(words, L, root)
        vox = np.ascontiguousarray(np.argwhere(grid)[:, ::-1].astype(np.uint32))
        if len(vox):
            got = O.color_fetch(dev.color_nodes(), dev.color_leaves(), int(res[ids[0]]["color_root"]),
                                cfg.voxel_level, cfg.color_leaf_level, vox)
            assert np.array_equal(got, ora.colors(vox)), f"tick {t}: occupied-voxel colors differ"
    return ora


@pytest.mark.parametrize("plain", [False, True])
def test_serve_results_same_alone_and_in_the_mix(plain):
    capi, cfg, h, dev = _setup()
    reqs = W.serving_trace(90, cfg.resolution, h, clients=16, rate_per_s=20000, seed=5, rmin=2, rmax=16,
                           px_min=(32, 24), px_max=(96, 64))
    res, st, buf = dev.serve(reqs, max_edits_per_tick=24, max_pixels_per_tick=1 << 16, plain=plain,
                             keep_images=True)
    print("ticks", st.ticks, "edits", st.edits, "renders", st.renders, "max edits/tick", st.max_edits_in_tick,
          flush=True)
    assert st.edits + st.renders == len(reqs)
    assert np.all(res["done_ms"] >= res["admit_ms"]) and np.all(res["admit_ms"] + 1e-9 >= reqs["arrival_ms"] - 1e-6)
    if not plain:
        assert st.max_edits_in_tick > 1, "the trace should batch several edits per tick"
    assert dev.stats().overflow == 0
    _check_renders(capi, dev, reqs, res, buf)
    _check_ticks(cfg, h, dev, reqs, res)


def test_serve_continuous_and_plain_end_on_the_same_dag():
    capi, cfg, h, a = _setup()
    _, _, _, b = _setup()
    reqs = W.serving_trace(60, cfg.resolution, h, clients=8, rate_per_s=20000, seed=9, rmin=2, rmax=16,
                           px_min=(16, 16), px_max=(48, 32))
    a.serve(reqs, max_edits_per_tick=16, max_pixels_per_tick=1 << 15)
    b.serve(reqs, plain=True)
    L = cfg.node_levels
    ca = O.canonical_words(a.full_words(), L, np.array([a.root], np.uint32))
    cb = O.canonical_words(b.full_words(), L, np.array([b.root], np.uint32))
    assert np.array_equal(ca, cb)


def _paint_trace(n, res, seed):
    """Edit-only trace of box paints: they recolor voxels but leave the DAG (and its pages) as is."""
    reqs = W.serving_trace(n, res, None, clients=4, rate_per_s=20000, seed=seed, edit_frac=1.0)
    rng = np.random.default_rng(seed)
    for i in range(n):
        lo = rng.integers(0, res - 8, 3)
        reqs[i]["edit"] = W.edit(W.AABB_PAINT, tuple(int(v) for v in lo), tuple(int(v) + 8 for v in lo),
                                 color=int(rng.integers(0, 1 << 24)))
    return reqs


def test_serve_admission_gate_runs_gc():
    capi, cfg, h, dev = _setup()
    churn = W.mixed_batch(30, cfg.resolution, h, 4, 20, seed=8)
    dev.edit(churn)
    dev.gc()
    dev.edit(churn[::-1].copy())  # leaves garbage behind
    free_c, _ = dev.free_pages()
    reqs = _paint_trace(24, cfg.resolution, seed=2)
    # one more free page than the pool has: the first admission must compact it
    res, st, _ = dev.serve(reqs, max_edits_per_tick=8, max_pixels_per_tick=1 << 12, min_free_pages=free_c + 1)
    assert st.gcs >= 1 and st.edits == len(reqs) and st.gated_ticks == 0
    assert st.min_free_pages >= free_c + 1


def test_serve_gate_beyond_capacity_is_an_error():
    capi, cfg, h, dev = _setup()
    free, total = dev.free_pages()
    with pytest.raises(capi.HDError, match="free pages"):
        dev.serve(_paint_trace(4, cfg.resolution, seed=3), min_free_pages=total + 1)


def test_serve_rejects_in_place_colors():
    capi = _capi()
    cfg = small_cfg(64, 6, 5, 2, 3)  # keeps no color history
    dev = capi.DevicePool(cfg)
    reqs = W.serving_trace(4, 64, None, clients=1, rate_per_s=1000, seed=1, rmax=8)
    with pytest.raises(capi.HDError, match="color_keep_history"):
        dev.serve(reqs)
