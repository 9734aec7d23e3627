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
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample")
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
    ap.add_argument("--frames", type=int, default=3)
    ap.add_argument("--copy-mib", type=int, default=1024)
    ap.add_argument("--serve-requests", type=int, default=3000)
    ap.add_argument("--serve-rate", type=float, default=2000.0, help="aggregate request arrivals per second")
    ap.add_argument("--serve-overload", type=float, default=10.0, help="rate multiple of the overload run")
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

    edit_ms, trace_ms, step_ms = [], [], []
    for s in range(args.warmup + args.steps):
        pool.sync()
        t0 = time.perf_counter()
        pool.edit(batches[s], mode)
        t1 = time.perf_counter()
        kms = render()
        pool.sync()
        t2 = time.perf_counter()
        if s >= args.warmup:
            edit_ms.append((t1 - t0) * 1e3)
            trace_ms.append(kms)
            step_ms.append((t2 - t0) * 1e3)
        log(f"step {s}: edit {1e3 * (t1 - t0):.1f} ms ({vox[s] / 1e6:.1f} M voxels), trace kernel {kms:.3f} ms")

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

    cpu = None
    if state0 is not None:
        cpu = cpu_baseline(cfg, heights, state0, batches[0], args.cpu_seconds, res)
        if batched is not None:
            b3 = cpu_baseline(cfg, heights, batched.pop("_state0"), batched.pop("_batch"), args.cpu_seconds, res)
            batched["cpu_reference_voxels_per_s"] = b3["value"]
            batched["cpu_reference_sample"] = b3["sample"]
            batched["batched_vs_cpu"] = batched["batched_voxels_per_s"] / b3["value"]
    if batched is not None:
        batched.pop("_state0", None)
        batched.pop("_batch", None)

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
        "roofline": {
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
    t0 = time.perf_counter()
    p.edit(batch, capi.MODE_BATCHED)
    tb = time.perf_counter() - t0
    st = p.stats()
    out.update(batched_s=tb, batched_voxels_per_s=vox / tb, groups=int(st.last_groups),
               overflow=int(st.overflow))
    log(f"batched: {args.batch_edits} edits in {tb * 1e3:.1f} ms ({st.last_groups} passes), "
        f"{vox / tb / 1e9:.2f} G voxels/s")
    p.close()
    # same prefix, plain and batched, from the same state
    p = fresh()
    t0 = time.perf_counter()
    p.edit(prefix, capi.MODE_PLAIN)
    tp = time.perf_counter() - t0
    p.close()
    p = fresh()
    t0 = time.perf_counter()
    p.edit(prefix, capi.MODE_BATCHED)
    tbp = time.perf_counter() - t0
    p.close()
    out.update(prefix_edits=len(prefix), prefix_voxels=vox_prefix, plain_prefix_s=tp, batched_prefix_s=tbp,
               plain_voxels_per_s=vox_prefix / tp, batched_prefix_voxels_per_s=vox_prefix / tbp,
               batched_vs_plain=tp / tbp)
    log(f"prefix {len(prefix)} edits: plain {tp * 1e3:.1f} ms, batched {tbp * 1e3:.1f} ms -> {tp / tbp:.1f}x")
    return out


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


def cpu_baseline(cfg, heights, state0, batch, budget_s, res):
    """The reference's ThreadedEdit (libfork busy_pool over the host cores) from the same state."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O  # the checker (oracle/_ref)
    from vkhashdag_amd import workloads as W

    if not O.have_ref():
        log("oracle/_ref not built: no reference CPU baseline")
        return None
    workers = min(16, os.cpu_count() or 1)
    ref = O.RefPool(cfg, heights=heights, workers=workers)
    ref.load_state(state0)
    done, secs = 0, 0.0
    while done < len(batch) and secs < budget_s:
        chunk = batch[done:done + 8]
        secs += ref.edit_timed(chunk, threaded=True)
        done += len(chunk)
        log(f"cpu baseline: {done} edits in {secs:.1f}s")
    vox = W.edited_voxels(batch[:done], res)
    return {"value": vox / secs, "unit": "edited voxels/s", "cores": workers, "kind": "reference",
            "sample": f"first {done} edits of the first batch ({vox / 1e6:.1f} M voxels), "
                      f"NodePoolThreadedEdit::ThreadedEdit with VBREditorWrapper, libfork busy_pool({workers}), "
                      f"max_task_level = color leaf level, from the device pool's state after the terrain build",
            "seconds": secs}


if __name__ == "__main__":
    main()
