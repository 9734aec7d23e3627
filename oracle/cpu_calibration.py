"""Calibration of the CPU baseline: the compiled reference's ThreadedEdit timed on this host.

TEST INFRASTRUCTURE. bench.py times oracle/_ref (the VkHashDAG headers compiled with our in-memory
storage backends) on the GPU box's host cores. This script times the same thing on the build
container -- the benchmark's 4096^3 terrain, then its first batch of 256 sphere edits, one warm-up
and the median of 5 runs, each from the same state -- and the oracle's single-threaded restatement
beside it, and writes the table that ties the box's figure to a known host.

    python oracle/cpu_calibration.py [--out profiles/cpu_reference_calibration.md]
"""
from __future__ import annotations

import argparse
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import _oracle as O  # noqa: E402
from vkhashdag_amd import workloads as W  # noqa: E402


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--runs", type=int, default=5)
    ap.add_argument("--out", default=os.path.join(ROOT, "profiles", "cpu_reference_calibration.md"))
    args = ap.parse_args()
    if not O.have_ref():
        sys.exit("oracle/_ref is not built (make -C oracle ref, with the reference tree present)")
    res = 4096
    cfg = W.config_for_resolution(res, bucket_bits=16, wbpp=9, pbpb=2, color_leaf_depth=4,
                                  color_node_capacity=1 << 24, color_leaf_capacity=1 << 28)
    heights = W.value_noise_heightmap(res, 6, 1, res // 8, (res * 3) // 4)
    ora = O.OraclePool(cfg, heights=heights)
    ora.edit(W.terrain_edit(res, seed=1, skin=8))
    bw, live = ora.live_words()
    state = dict(bucket_words=bw, live=live, filled=ora.filled(), color_nodes=ora.color_nodes(),
                 color_leaves=ora.color_leaves(), root=ora.root, color_root=ora.color_root)
    batch = W.terrain_sphere_edits(256, res, heights, 4, 128, seed=1000)
    vox = W.edited_voxels(batch, res)
    workers = len(os.sched_getaffinity(0))
    ref = O.RefPool(cfg, heights=heights, workers=workers)
    secs = []
    for r in range(args.runs + 1):
        ref.load_state(state)
        t = ref.edit_timed(batch, threaded=True)
        if r:
            secs.append(t)
    ref.close()
    ones = []
    for r in range(3):
        o = O.OraclePool(cfg, heights=heights)
        o.load_state(state)
        t0 = time.perf_counter()
        o.edit(batch)
        ones.append(time.perf_counter() - t0)
    med = float(np.median(secs))
    with open(args.out, "w") as f:
        f.write("# CPU baseline calibration: the compiled reference on the build container\n\n")
        f.write("command: python oracle/cpu_calibration.py\n\n")
        f.write("Workload: the benchmark's 4096^3 terrain (6-octave value noise, seed 1), then its first batch of "
                "256 sphere add/remove edits (r log-uniform 4..128, seed 1000), NodePoolThreadedEdit::ThreadedEdit "
                "with VBREditorWrapper and max_task_level = the color leaf level, steady_clock around each call "
                "(main.cpp:152-157), one warm-up then the median of the runs, each from the same state.\n\n")
        f.write(f"host: {cpu_model()}, {os.cpu_count()} CPUs, {workers} workers\n\n")
        f.write("| runs (s) | median (s) | edited voxels | voxels/s |\n|---|---|---|---|\n")
        f.write(f"| {', '.join(f'{x:.4f}' for x in secs)} | {med:.4f} | {vox} | {vox / med:.4g} |\n\n")
        f.write(f"oracle (the single-threaded restatement, same edits): median {float(np.median(ones)):.3f} s "
                f"= {vox / float(np.median(ones)):.4g} voxels/s\n\n")
        f.write("Machine-readable:\n\n")
        f.write(f"container_reference_voxels_per_s: {vox / med:.6g}\n")
        f.write(f"container_workers: {workers}\n")
        f.write(f"container_oracle_voxels_per_s: {vox / float(np.median(ones)):.6g}\n")
    print(open(args.out).read())


if __name__ == "__main__":
    main()
