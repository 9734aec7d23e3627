"""Autotuner for the launch parameters of hd_tuning. trace_tile, top_pairs and small_level change
speed only (the same words, roots and pixels); batch_pass_edits also renumbers the nodes a batched
edit appends (the DAG up to renumbering, the voxels and the colors stay the same).

Parameters and the workload each is measured on (seeded synthetic inputs, the benchmark's shapes):
  trace_tile        8 or 16 pixel tiles for the frame kernel: a primary + shadow frame of the terrain;
  top_pairs, small_level
                    the levels the one-workgroup top-down / bottom-up kernels take: plain per-edit
                    passes (sphere add/remove, r log-uniform 4..128) from one checkpointed pool state;
  batch_pass_edits  the largest batched pass (queue depth): a batch of mixed edits from the same state.
The search is coordinate-wise in that order; each measurement is the best of `repeats` runs. Every
plain candidate must give the first candidate's roots, every batched one its rendered frame
(LOD off).

    python -m vkhashdag_amd.autotune --res 4096 --out vkhashdag_amd/tuning.json
    autotune.load("vkhashdag_amd/tuning.json")   # applies a stored set (hd_tuning_set)
"""
from __future__ import annotations

import argparse
import itertools
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_PATH = os.path.join(HERE, "tuning.json")

CANDIDATES = {
    "trace_tile": [8, 16],
    "top_pairs": [512, 1024, 2048, 4096],
    "small_level": [256, 512, 1024],
    "batch_pass_edits": [0, 512, 1024],
}


def load(path: str = DEFAULT_PATH) -> dict | None:
    """Applies the tuning stored at path (if the file exists); returns it."""
    from . import capi

    if not os.path.exists(path):
        return None
    with open(path) as f:
        doc = json.load(f)
    t = {k: int(v) for k, v in doc["tuning"].items() if k in capi.TUNING_FIELDS}
    capi.set_tuning(**t)
    return t


def _log(msg):
    print(f"[autotune] {msg}", file=sys.stderr, flush=True)


def _setup(res: int, seed: int):
    from . import capi
    from . import workloads as W

    cfg = W.config_for_resolution(res, bucket_bits=16, wbpp=9, pbpb=2, color_leaf_depth=4,
                                  color_node_capacity=1 << 24, color_leaf_capacity=1 << 28)
    heights = W.value_noise_heightmap(res, 6, seed, res // 8, (res * 3) // 4)
    pool = capi.DevicePool(cfg, heights=heights)
    pool.edit(W.terrain_edit(res, seed=seed, skin=8))
    return cfg, heights, pool


def _time_frame(capi, W, pool, res, width, height, repeats):
    vb = W.orbit_view((res / 2, res / 2, res * 0.3), 0.9 * res, 7, width, height, res=res, elevation=(0.35, 0.6))
    v = capi.make_view(vb, width, height, capi.TRACE_SHADOW)
    tiles = np.zeros(2, np.uint32)
    capi.check(capi.lib().hd_trace_tiles(capi.C.byref(v), 1, tiles.ctypes.data))  # at the current tile
    dview = capi.DeviceBuffer(72)
    dview.upload(np.frombuffer(bytes(v), np.uint8))
    dtiles = capi.DeviceBuffer(8)
    dtiles.upload(tiles)
    rgba = capi.DeviceBuffer(width * height * 4)
    best, img = float("inf"), None
    for _ in range(repeats + 1):  # the first launch warms up
        pool.timer_begin()
        capi.check(capi.lib().hd_trace_async(pool.h, pool.root, pool.color_root, dview.ptr, dtiles.ptr, 1,
                                             int(tiles[1]), rgba.ptr, None, None))
        best = min(best, pool.timer_end())
    img = rgba.download(np.uint32, width * height)
    return best, img


def _time_edits(capi, ckpt, heights, edits, mode, repeats, check_view=None):
    """Best time of `repeats` runs from the saved state; returns (ms, roots or the check frame)."""
    best, out = float("inf"), None
    for _ in range(repeats):
        pool, _ = capi.DevicePool.load(ckpt, heights=heights)
        pool.sync()
        t0 = time.perf_counter()
        pool.edit(edits, mode)
        pool.sync()
        best = min(best, time.perf_counter() - t0)
        if check_view is None:
            out = (pool.root, pool.color_root)
        else:
            vb, w, h = check_view
            out = pool.render(vb, w, h, shadow=True)
        pool.close()
    return best * 1e3, out


def tune(res: int = 4096, seed: int = 1, plain_edits: int = 64, batch_edits: int = 2048, width: int = 1920,
         height: int = 1080, repeats: int = 2, candidates: dict | None = None) -> dict:
    """Measures the candidates on the current device; returns {device, tuning, measurements}.
    Leaves the best set applied."""
    from . import capi
    from . import workloads as W

    cand = dict(CANDIDATES, **(candidates or {}))
    best = capi.get_tuning()
    meas: dict = {"trace_ms": {}, "plain_ms": {}, "batched_ms": {}}
    cfg, heights, pool = _setup(res, seed)
    _log(f"{res}^3 terrain built; used words {pool.stats().used_words}")

    # 1. frame kernel tile
    ref_img = None
    for t in cand["trace_tile"]:
        capi.set_tuning(trace_tile=t)
        ms, img = _time_frame(capi, W, pool, res, width, height, repeats)
        if ref_img is None:
            ref_img = img
        elif not np.array_equal(img, ref_img):
            raise capi.HDError(f"trace_tile={t} changed the image")
        meas["trace_ms"][str(t)] = ms
        _log(f"trace_tile {t}: {ms:.3f} ms")
    best["trace_tile"] = int(min(cand["trace_tile"], key=lambda t: meas["trace_ms"][str(t)]))
    capi.set_tuning(trace_tile=best["trace_tile"])

    # 2./3. edit passes from one saved state
    with tempfile.TemporaryDirectory() as tmp:
        ckpt = os.path.join(tmp, "pool.hdag")
        pool.save(ckpt)
        pool.close()
        plain = W.terrain_sphere_edits(plain_edits, res, heights, 4, 128, seed=1000 + seed)
        ref = None
        for tp, sl in itertools.product(cand["top_pairs"], cand["small_level"]):
            capi.set_tuning(top_pairs=tp, small_level=sl)
            ms, roots = _time_edits(capi, ckpt, heights, plain, capi.MODE_PLAIN, repeats)
            if ref is None:
                ref = roots
            elif roots != ref:
                raise capi.HDError(f"top_pairs={tp} small_level={sl} changed the roots")
            meas["plain_ms"][f"{tp},{sl}"] = ms
            _log(f"top_pairs {tp} small_level {sl}: {ms:.2f} ms for {plain_edits} plain edits")
        tp, sl = min(itertools.product(cand["top_pairs"], cand["small_level"]),
                     key=lambda k: meas["plain_ms"][f"{k[0]},{k[1]}"])
        best["top_pairs"], best["small_level"] = int(tp), int(sl)
        capi.set_tuning(top_pairs=tp, small_level=sl)

        batch = W.mixed_batch(batch_edits, res, heights, 2, 256, seed=2)
        # LOD off: a LOD cut-off shows the color stored at its cube's corner voxel, which may be
        # empty, and empty voxels' colors depend on how the edits were split into passes (the
        # batched bar is the DAG and the colors of occupied voxels)
        view = (W.orbit_view((res / 2, res / 2, res * 0.3), 0.9 * res, 11, 256, 144, res=res, lod=False), 256, 144)
        ref = None
        for q in cand["batch_pass_edits"]:
            capi.set_tuning(batch_pass_edits=q)
            ms, img = _time_edits(capi, ckpt, heights, batch, capi.MODE_BATCHED, repeats, view)
            if ref is None:
                ref = img
            elif not np.array_equal(img, ref):
                raise capi.HDError(f"batch_pass_edits={q} changed the rendered frame")
            meas["batched_ms"][str(q)] = ms
            _log(f"batch_pass_edits {q}: {ms:.2f} ms for {batch_edits} batched edits")
        best["batch_pass_edits"] = int(min(cand["batch_pass_edits"], key=lambda q: meas["batched_ms"][str(q)]))
    capi.set_tuning(**best)
    return {"device": _device_name(), "workload": {"res": res, "seed": seed, "plain_edits": plain_edits,
                                                   "batch_edits": batch_edits, "frame": [width, height]},
            "tuning": best, "measurements": meas}


def _device_name() -> str:
    from . import capi

    return capi.device_name(0)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--res", type=int, default=4096)
    ap.add_argument("--plain-edits", type=int, default=64)
    ap.add_argument("--batch-edits", type=int, default=2048)
    ap.add_argument("--repeats", type=int, default=2)
    ap.add_argument("--out", default=DEFAULT_PATH)
    args = ap.parse_args(argv)
    doc = tune(args.res, plain_edits=args.plain_edits, batch_edits=args.batch_edits, repeats=args.repeats)
    with open(args.out, "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps(doc["tuning"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
