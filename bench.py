"""Headline benchmark: edit -> render steps on a 4096^3 value-noise terrain on one MI355X.

A step = one batch of 256 sphere add/remove edits (radii log-uniform in [4, 128], centred on the
terrain surface, adds colored) applied through the plain per-edit device path, then one
1920x1080 frame with a primary ray per pixel and one shadow ray per hit. Every step uses its own
seeded batch, so the pool keeps growing as in interactive editing.

Prints ONE JSON line. value = edited voxels/s of the timed steps (the lattice points inside each
edit's shape, clipped to the grid, summed); rays/s is reported beside it.
The CPU baseline is the compiled reference's NodePoolThreadedEdit::ThreadedEdit (oracle/_ref,
built from the VkHashDAG headers) on the host's cores, started from the device pool's exact
state and fed the same edits (a bounded prefix of the first batch)."""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxel edits/s + Mrays/s (4096³ terrain, 1080p); batched vs plain; p99 latency"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--res", type=int, default=4096)
    ap.add_argument("--edits", type=int, default=256)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--mode", choices=["plain", "batched"], default="plain")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed runs of the CPU baseline (after a warm-up)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tuning", action="store_true", help="keep the default launch parameters")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--batch-edits", type=int, default=4096)
    ap.add_argument("--plain-prefix", type=int, default=512, help="edits of the batch also run plain, for the ratio")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--trace-only", action="store_true",
                    help="terrain + the edit batches of the warmup and timed steps (the scene of the last "
                         "timed frame), then --frames frames and a copy of --copy-mib MiB (the byte-count "
                         "calibration of the HBM counters); for rocprofv3 runs")
    ap.add_argument("--edit-only", action="store_true",
                    help="terrain + the warmup batches, then a 64 MiB copy, the timed edit work (--steps batches of "
                         "--mode, or with --mode batched the 4096-edit batch), then a --copy-mib copy: the copies "
                         "bracket the edit kernels in a profile and calibrate its byte counters; for rocprofv3 runs")
    ap.add_argument("--profile-gc", action="store_true", help="with --edit-only: the timed work is one GC")
    ap.add_argument("--frames", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--serve-requests", type=int, default=3000)
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="aggregate request arrivals per second")
    ap.add_argument("--serve-overload", type=float, default=10.0, help="rate multiple of the overload run")
    ap.add_argument("--no-config4", action="store_true", help="skip the 16384^3 multi-view frame section")
    ap.add_argument("--c4-frames", type=int, default=8, help="timed frames of the 16384^3 section")
    ap.add_argument("--no-config5", action="store_true", help="skip the 32768^3 serving section")
    ap.add_argument("--c5-requests", type=int, default=2000)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is designed, not built",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    from vkhashdag_amd import autotune
    from vkhashdag_amd import capi
    from vkhashdag_amd import workloads as W

    # launch parameters from the autotuner's stored set (speed only, never results)
    if not args.no_tuning and autotune.load() is not None:
        log(f"tuning {capi.get_tuning()} from {autotune.DEFAULT_PATH}")
    res = args.res
    cfg = W.config_for_resolution(res, bucket_bits=16, wbpp=9, pbpb=2, color_leaf_depth=4,
                                  color_node_capacity=1 << 24, color_leaf_capacity=1 << 28)
    t = time.time()
    heights = W.value_noise_heightmap(res, 6, 1, res // 8, (res * 3) // 4)
    log(f"heightmap {res}^2 in {time.time() - t:.1f}s")
    pool = capi.DevicePool(cfg, heights=heights)
    t = time.time()
    pool.edit(W.terrain_edit(res, seed=1, skin=8))
    st = pool.stats()
    log(f"terrain build {time.time() - t:.2f}s, used words {st.used_words}, color leaves {st.color_leaf_words}")
    state0 = None if args.no_cpu else pool.state()

    batches = [W.terrain_sphere_edits(args.edits, res, heights, 4, 128, seed=1000 + s)
               for s in range(args.warmup + args.steps)]
    vox = [W.edited_voxels(b, res) for b in batches]
    mode = capi.MODE_BATCHED if args.mode == "batched" else capi.MODE_PLAIN

    # camera: seeded orbit at 0.9 * res around the terrain centre, LOD on (TracePass.cpp:115-118)
    view_block = W.orbit_view((res / 2, res / 2, res * 0.3), 0.9 * res, 7, args.width, args.height, res=res,
                              elevation=(0.35, 0.6))
    v = capi.make_view(view_block, args.width, args.height, capi.TRACE_SHADOW)
    tiles = np.zeros(2, np.uint32)
    capi.check(capi.lib().hd_trace_tiles(C_ref(v), 1, tiles.ctypes.data))
    dview = capi.DeviceBuffer(72)
    dview.upload(np.frombuffer(bytes(v), np.uint8))
    dtiles = capi.DeviceBuffer(8)
    dtiles.upload(tiles)
    npx = args.width * args.height
    rgba = capi.DeviceBuffer(npx * 4)
    counters = capi.DeviceBuffer(32)

    def render(count=False):
        if count:
            counters.upload(np.zeros(4, np.uint64))
        pool.timer_begin()
        capi.check(capi.lib().hd_trace_async(pool.h, pool.root, pool.color_root, dview.ptr, dtiles.ptr, 1,
                                             int(tiles[1]), rgba.ptr, None, counters.ptr if count else None))
        return pool.timer_end()

    if args.trace_only:
        return trace_only(args, pool, batches, render, npx)
    if args.edit_only:
        return edit_only(args, pool, batches, heights, res, mode)

    edit_ms, edit_gpu_ms, trace_ms, step_ms = [], [], [], []
    state1 = None
    for s in range(args.warmup + args.steps):
        pool.sync()
        t0 = time.perf_counter()
        pool.timer_begin()
        pool.edit(batches[s], mode)
        ems = pool.timer_end()  # HIP events around the step's edit work on the pool's stream
        t1 = time.perf_counter()
        kms = render()
        pool.sync()
        t2 = time.perf_counter()
        if s >= args.warmup:
            edit_ms.append((t1 - t0) * 1e3)
            edit_gpu_ms.append(ems)
            trace_ms.append(kms)
            step_ms.append((t2 - t0) * 1e3)
        if s == 0 and state0 is not None:
            state1 = pool.state()  # after the timed region of this step: the CPU leg's result is checked against it
        log(f"step {s}: edit {1e3 * (t1 - t0):.1f} ms ({ems:.1f} ms on the device, {vox[s] / 1e6:.1f} M voxels), "
            f"trace kernel {kms:.3f} ms")

    # the same frame with the beam prepass (beam.frag): coarse 8x8-block march, then the trace
    vb_ = capi.make_view(view_block, args.width, args.height, capi.TRACE_SHADOW | capi.TRACE_BEAM)
    beams = np.zeros(2, np.uint32)
    capi.check(capi.lib().hd_trace_beam_sizes(C_ref(vb_), 1, beams.ctypes.data))
    dview_b = capi.DeviceBuffer(72)
    dview_b.upload(np.frombuffer(bytes(vb_), np.uint8))
    dbeams = capi.DeviceBuffer(8)
    dbeams.upload(beams)
    dbeam = capi.DeviceBuffer(int(beams[1]) * 4)
    beam_ms = []
    for _ in range(3):
        pool.timer_begin()
        capi.check(capi.lib().hd_trace_beam_async(pool.h, pool.root, pool.color_root, dview_b.ptr, dtiles.ptr,
                                                  dbeams.ptr, 1, int(tiles[1]), int(beams[1]), dbeam.ptr, rgba.ptr))
        beam_ms.append(pool.timer_end())
    beam_ms = float(np.mean(beam_ms[1:]))
    log(f"trace with beam prepass: {beam_ms:.3f} ms (beam image {int(beams[1])} blocks)")

    timed_vox = sum(vox[args.warmup:])
    edit_s = sum(edit_ms) / 1e3
    value = timed_vox / edit_s
    # instrumented frame (untimed): rays, iterations, pool words read
    render(count=True)
    c = counters.download(np.uint64, 4)
    primary, shadow, iters, words = (int(x) for x in c)
    rays = primary + shadow
    trace_avg_ms = float(np.mean(trace_ms))
    rays_per_s = rays / (trace_avg_ms / 1e3)
    bytes_frame = 4.0 * words + 4.0 * npx  # words read by the march + color fetch, RGBA8 output
    achieved = bytes_frame / (trace_avg_ms / 1e3) / 1e9
    st = pool.stats()
    log(f"edits: {value / 1e9:.3f} G voxels/s; trace: {rays_per_s / 1e9:.3f} G rays/s "
        f"({primary} primary + {shadow} shadow, {iters / max(rays, 1):.1f} iterations/ray, "
        f"{words / max(rays, 1):.1f} words/ray)")

    # GC of the edited pool (NodePoolThreadedGC::ThreadedGC)
    used_before = pool.stats().used_words
    pool.sync()
    t0 = time.perf_counter()
    pool.gc()
    gc_first_ms = (time.perf_counter() - t0) * 1e3  # includes the one-time allocation of its scratch
    used_after = pool.stats().used_words
    # again on the compacted pool: the same live nodes and sorts, scratch already allocated
    t0 = time.perf_counter()
    pool.gc()
    gc_ms = (time.perf_counter() - t0) * 1e3
    log(f"gc: {gc_ms:.1f} ms ({gc_first_ms:.1f} ms on the first call), used words {used_before} -> {used_after}")
    pool.close()

    batched = None if args.no_batched else batched_section(args, cfg, heights, res)
    serving = None if args.no_serve else serving_section(args, cfg, heights, res)
    config4 = None if args.no_config4 else guarded(config4_section, args)
    config5 = None if args.no_config5 else guarded(config5_section, args)

    cpu = None
    if state0 is not None:
        cpu = cpu_baseline(cfg, heights, state0, batches[0], res, args.cpu_runs, state1,
                           "the first batch of 256 sphere edits")
        if cpu is not None:
            cpu["gpu_plain_vs_cpu"] = value / cpu["value"]
        if batched is not None and batched.get("_state0") is not None:
            b3 = cpu_baseline(cfg, heights, batched.pop("_state0"), batched.pop("_batch"), res, args.cpu_runs,
                              batched.pop("_state1"), "the 4096 mixed edits")
            if b3 is not None:
                batched["cpu_reference"] = b3
                batched["cpu_reference_voxels_per_s"] = b3["value"]
                batched["batched_vs_cpu"] = batched["batched_voxels_per_s"] / b3["value"]
    if batched is not None:
        for k in ("_state0", "_batch", "_state1"):
            batched.pop(k, None)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "edited voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": float(np.mean(step_ms)),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"{res}^3 value-noise terrain (6 octaves, seed 1, z in [{res // 8},{(res * 3) // 4}]), "
                        f"{args.edits} sphere add/remove edits per step (r log-uniform 4..128, adds VBR-colored), "
                        f"{args.mode} device edit path, then a {args.width}x{args.height} frame "
                        f"(primary + 1 shadow ray per hit, LOD on)",
            "edits_per_step": args.edits,
            "resolution": res,
            "frame": [args.width, args.height],
            "edit_mode": args.mode,
            "tuning": capi.get_tuning(),
        },
        "edit_ms_per_step": float(np.mean(edit_ms)),
        "rays_per_s": rays_per_s,
        "mrays_per_s": rays_per_s / 1e6,
        "trace_ms": trace_avg_ms,
        "trace_beam_ms": beam_ms,
        "rays_per_frame": rays,
        "roofline": edit_roofline(args, edit_gpu_ms),
        "roofline_trace": {
            "kernel": "k_trace (DAG_RayMarch + shadow + Color_Fetch)",
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": profiled_traffic(args),
            "traffic_source": TRAFFIC_PROFILE if profiled_traffic(args) else None,
            "algorithmic_bytes_per_launch": bytes_frame,
            # what limits the kernel instead: vector-instruction issue, from the SQ counters
            "valu_issue_frac": profiled_valu_issue(args),
            "valu_issue_source": VALU_PROFILE if profiled_valu_issue(args) else None,
        },
        "cpu_baseline": cpu,
        "gc_ms": gc_ms,
        "gc_first_ms": gc_first_ms,
        "gc_used_words": [int(used_before), int(used_after)],
        "batched": batched,
        "serving": serving,
        "multiview_16384": config4,
        "serving_32768": config5,
        "pool": {"used_words": int(st.used_words), "color_nodes": int(st.color_nodes),
                 "color_leaf_words": int(st.color_leaf_words), "overflow": int(st.overflow)},
    }
    print(json.dumps(out))


def batched_section(args, cfg, heights, res):
    """Batched edits on the terrain: 4096 edits of uneven size (40% sphere add, 40% sphere remove,
    20% box paint, radii / half-extents log-uniform in [2, 256], centres on the surface, seed 2)
    applied in dependency-ordered passes; the plain path and the batched path also run the same
    prefix from the same state for the ratio."""
    from vkhashdag_amd import capi
    from vkhashdag_amd import workloads as W

    batch = W.mixed_batch(args.batch_edits, res, heights, 2, 256, seed=2)
    vox = W.edited_voxels(batch, res)
    prefix = batch[:args.plain_prefix]
    vox_prefix = W.edited_voxels(prefix, res)
    out = {"workload": f"{args.batch_edits} mixed edits (40% sphere add, 40% sphere remove, 20% box paint, "
                       f"r log-uniform 2..256, seed 2) on the same terrain",
           "edited_voxels": vox}

    def fresh():
        p = capi.DevicePool(cfg, heights=heights)
        p.edit(W.terrain_edit(res, seed=1, skin=8))
        p.sync()
        return p

    p = fresh()
    out["_state0"] = None if args.no_cpu else p.state()
    out["_batch"] = batch
    # the pass buffers are allocated before the clock starts (hd_edit_reserve), as a serving loop
    # would: the timed call plans, schedules and runs the passes
    t0 = time.perf_counter()
    p.reserve(batch, capi.MODE_BATCHED)
    out["reserve_s"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    p.timer_begin()
    p.edit(batch, capi.MODE_BATCHED)
    tb_gpu = p.timer_end() / 1e3
    tb = time.perf_counter() - t0
    st = p.stats()
    out["_state1"] = None if args.no_cpu else p.state()
    out.update(batched_s=tb, batched_device_s=tb_gpu, batched_voxels_per_s=vox / tb, groups=int(st.last_groups),
               overflow=int(st.overflow))
    eb = edit_bytes_profile(args)
    if eb is not None and "batched_bytes" in eb:
        achieved = eb["batched_bytes"] / tb_gpu / 1e9
        out["roofline"] = {"kernel": "batched edit passes (node levels + color stream)", "bound": "hbm",
                           "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                           "frac": achieved / HBM_PEAK_GBS, "algorithmic_bytes": eb["batched_bytes"],
                           "algorithmic_source": EDIT_BYTES_PROFILE}
    log(f"batched: {args.batch_edits} edits in {tb * 1e3:.1f} ms ({st.last_groups} passes), "
        f"{vox / tb / 1e9:.2f} G voxels/s")
    p.close()
    # same prefix, plain and batched, from the same state
    p = fresh()
    p.reserve(prefix, capi.MODE_PLAIN)
    t0 = time.perf_counter()
    p.edit(prefix, capi.MODE_PLAIN)
    tp = time.perf_counter() - t0
    p.close()
    p = fresh()
    p.reserve(prefix, capi.MODE_BATCHED)
    t0 = time.perf_counter()
    p.edit(prefix, capi.MODE_BATCHED)
    tbp = time.perf_counter() - t0
    p.close()
    out.update(prefix_edits=len(prefix), prefix_voxels=vox_prefix, plain_prefix_s=tp, batched_prefix_s=tbp,
               plain_voxels_per_s=vox_prefix / tp, batched_prefix_voxels_per_s=vox_prefix / tbp,
               batched_vs_plain=tp / tbp)
    log(f"prefix {len(prefix)} edits: plain {tp * 1e3:.1f} ms, batched {tbp * 1e3:.1f} ms -> {tp / tbp:.1f}x")
    return out


def guarded(section, args):
    """Runs a workload section; a library error (capacity, argument) is reported in the line
    instead of ending the run."""
    from vkhashdag_amd import capi
    try:
        return section(args)
    except capi.HDError as e:
        log(f"{section.__name__}: {e}")
        return {"error": str(e)}


def pool_bytes(cfg, st):
    """Node-pool bytes per level (used words x 4 B) against the pool's capacity and 288 GB of HBM."""
    from vkhashdag_amd import capi  # noqa: F401
    hbm = 288e9
    levels = []
    for l, bits in enumerate(cfg.bucket_bits):
        used = int(st.level_used_words[l]) * 4
        cap = (1 << bits) * cfg.words_per_bucket * 4
        levels.append({"level": l, "used_bytes": used, "capacity_bytes": cap, "used_of_capacity": used / cap})
    used = sum(x["used_bytes"] for x in levels)
    reserved = cfg.total_words * 4 + cfg.color_node_capacity * 32 + cfg.color_leaf_capacity * 4
    return {"levels": levels, "node_used_bytes": used, "color_leaf_bytes": int(st.color_leaf_words) * 4,
            "color_node_bytes": int(st.color_nodes) * 32, "reserved_bytes": reserved,
            "used_of_hbm": (used + int(st.color_leaf_words) * 4 + int(st.color_nodes) * 32) / hbm,
            "reserved_of_hbm": reserved / hbm}


def config4_section(args):
    """A 16384^3 terrain (the same value-noise recipe, 8 octaves, seed 3, z in [2048, 12288]) with VBR
    colors, generated and built on the device. Each frame is one hipGraph: up to 4 passes of small
    mixed edits (radius <= 64), a GC of the new root, then 16 views at uneven resolutions between
    640x360 and 1920x1080 (primary + 1 shadow ray per hit) in one launch. Launch parameters come from
    the autotuner's stored set. Node-pool bytes per level are reported against 288 GB."""
    from vkhashdag_amd import capi
    from vkhashdag_amd import workloads as W

    res = 16384
    cfg = W.config_for_resolution(res, bucket_bits=16, wbpp=9, pbpb=2, color_leaf_depth=4,
                                  color_node_capacity=1 << 26, color_leaf_capacity=1 << 30)
    pool = capi.DevicePool(cfg)
    t0 = time.perf_counter()
    heights = pool.value_noise_heightmap(8, 3, res // 8, (res * 3) // 4)
    t_hm = time.perf_counter() - t0
    t0 = time.perf_counter()
    # 16 column tiles: one whole-grid column fill would rebuild 4.5 M color leaves in one pass
    pool.edit(W.terrain_tiles(res, 3, 4, skin=8), capi.MODE_PLAIN)
    pool.sync()
    t_build = time.perf_counter() - t0
    st = pool.stats()
    log(f"16384^3 terrain: heightmap {t_hm:.2f}s, build {t_build:.2f}s, {st.used_words} words, "
        f"{st.color_leaf_words} color leaf words, overflow {st.overflow}")
    rng = np.random.default_rng(3)
    views = []
    for k in range(16):
        w, h = int(rng.integers(640, 1921)), int(rng.integers(360, 1081))
        vb = W.orbit_view((res / 2, res / 2, res * 0.3), 0.9 * res, 100 + k, w, h, res=res, elevation=(0.35, 0.6))
        views.append((vb, w, h, capi.TRACE_SHADOW))
    pixels = sum(w * h for _, w, h, _ in views)
    out = {"workload": "16384^3 value-noise terrain (8 octaves, seed 3, z in [2048,12288], VBR colors), frames of "
                       "24 mixed edits (r log-uniform 2..64) -> GC -> 16 views (640x360..1920x1080, primary + 1 "
                       "shadow ray per hit), each frame one hipGraph",
           "heightmap_s": t_hm, "terrain_build_s": t_build, "views": [[w, h] for _, w, h, _ in views],
           "pixels_per_frame": pixels, "pool_after_build": pool_bytes(cfg, st)}
    frame = capi.Frame(pool, max_edits=32, max_passes=4, max_radius=64, n_views=16, max_width=1920,
                       max_height=1080, colors=True, gc=True)
    ms, wall = [], []
    n_frames = max(1, args.c4_frames)
    for f in range(n_frames + 1):
        e = W.mixed_batch(24, res, heights, 2, 64, seed=500 + f)
        frame.set(e, views)
        t0 = time.perf_counter()
        pool.timer_begin()
        frame.launch()
        frame.wait()
        fm = pool.timer_end()
        if f:  # the first replay is a warm-up
            ms.append(fm)
            wall.append((time.perf_counter() - t0) * 1e3)
    frame.close()
    st = pool.stats()
    # rays of one frame's views on the final DAG (untimed, instrumented trace)
    vs = (capi.hd_view * len(views))()
    off = 0
    for i, (vb, w, h, fl) in enumerate(views):
        vs[i] = capi.make_view(vb, w, h, fl, off)
        off += w * h
    tiles = np.zeros(len(views) + 1, np.uint32)
    capi.check(capi.lib().hd_trace_tiles(vs, len(views), tiles.ctypes.data))
    dv, dt = capi.DeviceBuffer(72 * len(views)), capi.DeviceBuffer(4 * len(tiles))
    dv.upload(np.frombuffer(bytes(vs), np.uint8))
    dt.upload(tiles)
    rgba, counters = capi.DeviceBuffer(4 * pixels), capi.DeviceBuffer(32)
    counters.upload(np.zeros(4, np.uint64))
    capi.check(capi.lib().hd_trace_async(pool.h, pool.root, pool.color_root, dv.ptr, dt.ptr, len(views),
                                         int(tiles[-1]), rgba.ptr, None, counters.ptr))
    pool.sync()  # the trace runs on the pool's stream; the copy below does not wait for it
    c = counters.download(np.uint64, 4)
    pool.timer_begin()
    capi.check(capi.lib().hd_trace_async(pool.h, pool.root, pool.color_root, dv.ptr, dt.ptr, len(views),
                                         int(tiles[-1]), rgba.ptr, None, None))
    trace_ms = pool.timer_end()
    rays = int(c[0] + c[1])
    out.update(frame_ms=float(np.median(ms)), frame_ms_all=ms, frame_wall_ms=float(np.median(wall)),
               trace_only_ms=trace_ms, rays_per_frame=rays, rays_per_s_trace=rays / (trace_ms / 1e3),
               pool_after_frames=pool_bytes(cfg, st), overflow=int(st.overflow))
    log(f"16384^3 frames: {np.median(ms):.2f} ms per edit->GC->render frame ({rays / 1e6:.1f} M rays; the trace "
        f"alone {trace_ms:.2f} ms = {rays / trace_ms / 1e6:.2f} G rays/s)")
    pool.close()
    return out


def config5_section(args):
    """Serving on a 32768^3 terrain (the value-noise recipe, 9 octaves, seed 4, z in [4096, 24576]):
    64 clients, Poisson arrivals at 2000 req/s aggregate; 70% edits (40% sphere add, 40% sphere
    remove, 20% box paint, radii log-uniform in [2, 512], colored), 30% renders at uniform-random
    resolutions 256x256..1920x1080 from random orbit cameras; continuous batching vs one request at a
    time. The deepest levels get more buckets (within the 2^32-word pointer space) and the terrain
    carries no colors: its per-block colors would need about 1.2 G color-leaf words at this size, past
    the 30-bit color pointers (DAGColorPool.hpp:24); the edits' colors are stored."""
    from vkhashdag_amd import capi
    from vkhashdag_amd import workloads as W

    res = 32768
    bits = [16] * 10 + [18, 19, 19, 16]
    cfg = W.PoolConfig(9, 2, bits, 15 - 4, keep_history=True, color_node_capacity=1 << 24,
                       color_leaf_capacity=1 << 28)
    out = {"workload": f"{args.c5_requests} requests from 64 clients, Poisson at 2000/s aggregate: 70% edits "
                       f"(r log-uniform 2..512, colored) and 30% renders (256x256..1920x1080, primary + shadow) on "
                       f"a 32768^3 value-noise terrain (9 octaves, seed 4, z in [4096,24576]; terrain uncolored)",
           "bucket_bits": bits}
    heights = None
    for name, plain in (("continuous", False), ("plain", True)):
        pool = capi.DevicePool(cfg)
        t0 = time.perf_counter()
        h = pool.value_noise_heightmap(9, 4, res // 8, (res * 3) // 4, download=heights is None)
        heights = h if heights is None else heights
        pool.edit(W.terrain_tiles(res, 4, 4, skin=8, colored=False), capi.MODE_PLAIN)
        pool.sync()
        out.setdefault("terrain_build_s", time.perf_counter() - t0)
        if "pool_after_build" not in out:
            out["pool_after_build"] = pool_bytes(cfg, pool.stats())
            reqs = W.serving_trace(args.c5_requests, res, heights, clients=64, rate_per_s=2000.0, seed=4, rmin=2,
                                   rmax=512)
            out["edited_voxels"] = W.edited_voxels(reqs[reqs["kind"] == W.REQ_EDIT]["edit"], res)
        r, st, _ = pool.serve(reqs, max_edits_per_tick=256, max_pixels_per_tick=1 << 23, min_free_pages=1 << 16,
                              plain=plain)
        lat = r["done_ms"] - reqs["arrival_ms"]
        le, lr = lat[reqs["kind"] == W.REQ_EDIT], lat[reqs["kind"] == W.REQ_RENDER]
        wall = st.wall_ms / 1e3
        out[name] = {"wall_s": wall, "requests_per_s": len(reqs) / wall, "ticks": int(st.ticks), "gcs": int(st.gcs),
                     "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                                    "edit_p99": float(np.percentile(le, 99)),
                                    "render_p99": float(np.percentile(lr, 99))},
                     "overflow": int(pool.stats().overflow)}
        log(f"32768^3 serve {name}: {len(reqs)} requests in {wall:.2f}s, p50 {out[name]['latency_ms']['p50']:.2f} "
            f"ms p99 {out[name]['latency_ms']['p99']:.2f} ms")
        pool.close()
    return out


def edit_only(args, pool, batches, heights, res, mode):
    """Profiling driver for the edit kernels: the warmup batches, a 64 MiB device copy (a marker), the
    timed work, a --copy-mib copy (marker + the byte-count calibration: it reads and writes exactly
    that many bytes). The timed work is --steps plain batches, or the 4096-edit mixed batch
    (batched) on the terrain."""
    from vkhashdag_amd import capi
    from vkhashdag_amd import workloads as W
    for b in batches[:args.warmup + (args.steps if args.profile_gc else 0)]:
        pool.edit(b, capi.MODE_PLAIN)
    work = None
    if mode == capi.MODE_BATCHED and not args.profile_gc:
        work = W.mixed_batch(args.batch_edits, res, heights, 2, 256, seed=2)
        pool.reserve(work, capi.MODE_BATCHED)
    marker = 64 << 20
    m0, m1 = capi.DeviceBuffer(marker), capi.DeviceBuffer(marker)
    capi.check(capi.lib().hd_device_copy(m1.ptr, m0.ptr, marker, pool.h))
    pool.timer_begin()
    if args.profile_gc:
        pool.gc()
    elif mode == capi.MODE_BATCHED:
        pool.edit(work, capi.MODE_BATCHED)
    else:
        for b in batches[args.warmup:]:
            pool.edit(b, capi.MODE_PLAIN)
    ms = pool.timer_end()
    nbytes = args.copy_mib << 20
    src, dst = capi.DeviceBuffer(nbytes), capi.DeviceBuffer(nbytes)
    capi.check(capi.lib().hd_device_copy(dst.ptr, src.ptr, nbytes, pool.h))
    pool.sync()
    st = pool.stats()  # the work of the last hd_edit (the last plain step, or the whole batch)
    print(json.dumps({"mode": args.mode, "steps": args.steps, "device_ms": ms, "marker_bytes": marker,
                      "copy_bytes": nbytes, "last_entries": st.last_entries, "last_upserts": st.last_upserts,
                      "last_jobs": st.last_jobs, "last_passes": st.last_groups}))


def trace_only(args, pool, batches, render, npx):
    """Profiling driver: the edited terrain, --frames frames of k_trace, then one device copy of
    --copy-mib MiB (a kernel of known traffic: reads and writes exactly that many bytes)."""
    from vkhashdag_amd import capi
    for batch in batches:
        pool.edit(batch)
    ms = [render() for _ in range(args.frames)]
    nbytes = args.copy_mib << 20
    src, dst = capi.DeviceBuffer(nbytes), capi.DeviceBuffer(nbytes)
    capi.check(capi.lib().hd_device_copy(dst.ptr, src.ptr, nbytes, pool.h))
    pool.sync()
    print(json.dumps({"frames": args.frames, "trace_ms": ms, "pixels": npx, "copy_bytes": nbytes}))


def serving_section(args, cfg, heights, res):
    """64 clients, Poisson arrivals at an aggregate rate; 70% edits (40% sphere add, 40% sphere
    remove, 20% box paint, radii log-uniform in [2, 64] -- the [2, 512] range of a 32768^3 scene
    scaled to this grid), 30% renders at uniform-random resolutions 256x256..1920x1080 from random
    orbit cameras with 1 shadow ray per hit. The continuous-batching loop and the
    one-request-at-a-time loop each serve the same trace from the same terrain."""
    from dataclasses import replace

    from vkhashdag_amd import capi
    from vkhashdag_amd import workloads as W

    scfg = replace(cfg, keep_history=True)  # renders overlap edits: colors must be append-only
    out = {}
    for run, rate in (("offered", args.serve_rate), ("overload", args.serve_rate * args.serve_overload)):
        reqs = W.serving_trace(args.serve_requests, res, heights, clients=64, rate_per_s=rate, seed=4,
                               rmin=2, rmax=64)
        out[run] = serve_run(args, scfg, heights, res, reqs, rate)
    return out


def serve_run(args, scfg, heights, res, reqs, rate):
    from vkhashdag_amd import capi
    from vkhashdag_amd import workloads as W

    edits = reqs[reqs["kind"] == W.REQ_EDIT]["edit"]
    vox = W.edited_voxels(edits, res)
    rend = reqs[reqs["kind"] == W.REQ_RENDER]["view"]
    pixels = int(np.sum(rend["width"].astype(np.int64) * rend["height"]))
    out = {"workload": f"{len(reqs)} requests from 64 clients, Poisson at {rate:.0f}/s aggregate: "
                       f"{len(edits)} edits (r log-uniform 2..64) and {len(rend)} renders "
                       f"(256x256..1920x1080, primary + shadow) on the {res}^3 terrain",
           "edited_voxels": vox, "render_pixels": pixels}
    for name, plain in (("continuous", False), ("plain", True)):
        p = capi.DevicePool(scfg, heights=heights)
        p.edit(W.terrain_edit(res, seed=1, skin=8))
        p.sync()
        res_, st, _ = p.serve(reqs, max_edits_per_tick=256, max_pixels_per_tick=1 << 23,
                              min_free_pages=1 << 14, plain=plain)
        lat = res_["done_ms"] - reqs["arrival_ms"]
        le = lat[reqs["kind"] == W.REQ_EDIT]
        lr = lat[reqs["kind"] == W.REQ_RENDER]
        wall = st.wall_ms / 1e3
        out[name] = {
            "wall_s": wall, "requests_per_s": len(reqs) / wall, "edited_voxels_per_s": vox / wall,
            "pixels_per_s": pixels / wall, "ticks": int(st.ticks), "gcs": int(st.gcs),
            "max_edits_in_tick": int(st.max_edits_in_tick), "max_renders_in_tick": int(st.max_renders_in_tick),
            "latency_ms": {"p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
                           "edit_p50": float(np.percentile(le, 50)), "edit_p99": float(np.percentile(le, 99)),
                           "render_p50": float(np.percentile(lr, 50)), "render_p99": float(np.percentile(lr, 99))},
            "overflow": int(p.stats().overflow)}
        log(f"serve {name} @ {rate:.0f}/s: {len(reqs)} requests in {wall:.2f}s, {st.ticks} ticks, latency p50 "
            f"{out[name]['latency_ms']['p50']:.2f} ms p99 {out[name]['latency_ms']['p99']:.2f} ms")
        p.close()
    return out


EDIT_BYTES_PROFILE = "profiles/edit_algorithmic_bytes_4096_terrain.md"
EDIT_TRAFFIC_PROFILE = "profiles/edit_plain_4096_terrain_256_spheres_hbm.md"


def _profile_values(path):
    """`key: value` lines of a committed profile summary (values parsed as JSON)."""
    out = {}
    try:
        for line in open(os.path.join(ROOT, path)):
            if ":" in line and not line.startswith(("#", "|", " ")):
                k, v = line.split(":", 1)
                try:
                    out[k.strip()] = json.loads(v)
                except ValueError:
                    pass
    except OSError:
        return None
    return out


def edit_bytes_profile(args):
    """B_edit of the benchmark's workloads, counted by the oracle's instrumented edit recursion
    (oracle/edit_bytes.py wrote the table); only for the workload it was counted on."""
    if (args.res, args.edits) != (4096, 256):
        return None
    return _profile_values(EDIT_BYTES_PROFILE)


def edit_roofline(args, edit_gpu_ms):
    """The plain edit path against HBM: B_edit of the timed steps (existing node words read, words
    hashed, scanned in target buckets, appended, counter words) over their HIP-event-timed device
    time. traffic = the step's HBM bytes from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs over every
    edit kernel (scaled by a copy of known size), per 256-edit step."""
    eb = edit_bytes_profile(args)
    if eb is None or args.mode != "plain" or len(eb.get("plain_step_bytes", [])) < args.warmup + args.steps:
        return None
    b = eb["plain_step_bytes"][args.warmup:args.warmup + args.steps]
    per_step_bytes = float(np.mean(b))
    t = float(np.mean(edit_gpu_ms)) / 1e3
    achieved = per_step_bytes / t / 1e9
    tr = _profile_values(EDIT_TRAFFIC_PROFILE) or {}
    return {"kernel": f"plain edit pass x {args.edits} per step (a_top, a_count/a_emit, a_leaf, a_build/a_dedup/"
                      f"a_walk, a_bottom + the color stream)",
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": tr.get("traffic_bytes_per_step"),
            "traffic_source": EDIT_TRAFFIC_PROFILE if tr.get("traffic_bytes_per_step") else None,
            "algorithmic_bytes_per_step": per_step_bytes, "algorithmic_source": EDIT_BYTES_PROFILE,
            "device_ms_per_step": t * 1e3,
            "note": "latency-bound: ~10 dependent launches per edit, each a few memory round trips"}


TRAFFIC_PROFILE = "profiles/k_trace_4096_terrain_1080p.md"


def profiled_traffic(args):
    """HBM bytes per k_trace launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (scaled
    by a copy of known size); only for the workload they were taken on."""
    if (args.res, args.width, args.height) != (4096, 1920, 1080):
        return None
    try:
        for line in open(os.path.join(ROOT, TRAFFIC_PROFILE)):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        return None
    return None


VALU_PROFILE = "profiles/k_trace_4096_terrain_1080p_instructions.md"


def profiled_valu_issue(args):
    """Fraction of the vector-instruction issue rate k_trace reaches, from the committed rocprofv3
    SQ_INSTS_VALU run (wave instructions x 4 cycles over 1024 SIMDs at 2.4 GHz for the kernel-trace
    average duration); only for the workload it was taken on."""
    if (args.res, args.width, args.height) != (4096, 1920, 1080):
        return None
    try:
        for line in open(os.path.join(ROOT, VALU_PROFILE)):
            if line.startswith("valu_issue_frac:"):
                return float(line.split(":")[1])
    except OSError:
        return None
    return None


def C_ref(v):
    import ctypes
    return ctypes.byref(v)


def host_cpu():
    """The host's CPU: model name, online CPUs, the CPUs this process may run on, and its share: the
    GPU box sets OMP_NUM_THREADS to the cores one GPU's share holds (16), fewer than it may see."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    share = int(os.environ.get("OMP_NUM_THREADS") or 0) or affinity
    return {"model": model, "nproc": os.cpu_count(), "affinity": affinity, "workers": min(affinity, share)}


def cpu_baseline(cfg, heights, state0, batch, res, runs, check_state, what):
    """The reference's NodePoolThreadedEdit::ThreadedEdit (oracle/_ref: the VkHashDAG headers compiled
    with a libfork busy_pool over the host's cores, VBREditorWrapper for colored edits, max_task_level
    = the color leaf level as at main.cpp:216) on the same edits from the device pool's exact state.
    steady_clock around each ThreadedEdit call (main.cpp:152-157), summed over the edits; one warm-up
    run, then the median of `runs` runs, each from the same state. The reference's final DAG is
    checked against the device's (canonically: the threaded reference appends in a nondeterministic
    order)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O  # the checker (oracle/_ref; oracle/liboracle for the canonical compare)
    from vkhashdag_amd import workloads as W

    if not O.have_ref():
        log("oracle/_ref not built: no reference CPU baseline")
        return None
    cpu = host_cpu()
    workers = cpu["workers"]
    ref = O.RefPool(cfg, heights=heights, workers=workers)
    try:
        secs = []
        for r in range(runs + 1):
            ref.load_state(state0)
            t = ref.edit_timed(batch, threaded=True)
            if r:
                secs.append(t)
            log(f"cpu baseline ({what}): run {r}{' (warm-up)' if r == 0 else ''} {t * 1e3:.1f} ms")
        same = None
        if check_state is not None:
            a, b = O.OraclePool(cfg), O.OraclePool(cfg)
            a.load_state(ref.state())
            b.load_state(check_state)
            same = bool(np.array_equal(a.canonical(), b.canonical()))
            log(f"cpu baseline ({what}): reference DAG {'equals' if same else 'DIFFERS FROM'} the device DAG")
            del a, b
    finally:
        ref.close()  # joins the busy_pool's worker threads before the process exits
    vox = W.edited_voxels(batch, res)
    med = float(np.median(secs))
    return {"value": vox / med, "unit": "edited voxels/s", "cores": workers, "kind": "reference",
            "sample": f"{what} ({len(batch)} edits, {vox / 1e6:.1f} M voxels) from the device pool's state before "
                      f"them: NodePoolThreadedEdit::ThreadedEdit with VBREditorWrapper, libfork busy_pool({workers}), "
                      f"max_task_level = color leaf level; median of {runs} runs after a warm-up",
            "seconds": med, "seconds_runs": secs, "cpu_model": cpu["model"], "nproc": cpu["nproc"],
            "affinity_cpus": cpu["affinity"], "result_equals_device": same}


if __name__ == "__main__":
    main()
