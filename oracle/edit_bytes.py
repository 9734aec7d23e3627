"""Algorithmic node-pool bytes of the benchmark's edit workloads (the edit path's roofline numerator).

TEST INFRASTRUCTURE: runs the CPU oracle's instrumented edit recursion (NodePool::Traffic in
oracle/hd_oracle.hpp) over the same seeded workloads bench.py times, and writes the per-step
counts as a text table that bench.py reads (bench.py itself never loads the oracle).

B_edit = 4 B x [ words of existing nodes the recursion reads (edit_leaf / edit_node unpack,
NodePool.hpp:319-396) + for each upsert (NodePool.hpp:159-211): words hashed + words scanned in
the target bucket + words appended (incl. page padding) + 1 bucket-counter word ].
Plain steps count every edit against the state it sees (sequential). The batched workload is
counted as one batch: an input node once however many edits read it, and the appended nodes are
the new nodes of the final DAG (each hashed, scanned against its bucket's input words, written,
counted once).

    python oracle/edit_bytes.py [--steps 32] [--out profiles/edit_algorithmic_bytes_4096_terrain.md]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

import _oracle as O  # noqa: E402
from vkhashdag_amd import workloads as W  # noqa: E402

FIELDS = O.OraclePool.TRAFFIC_FIELDS


def words_of(t):
    return t["visited"] + t["hashed"] + t["scanned"] + t["written"] + t["counters"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=32, help="plain steps (256 edits each) to count")
    ap.add_argument("--res", type=int, default=4096)
    ap.add_argument("--edits", type=int, default=256)
    ap.add_argument("--batch-edits", type=int, default=4096)
    ap.add_argument("--plain-prefix", type=int, default=512)
    ap.add_argument("--out", default=os.path.join(ROOT, "profiles", "edit_algorithmic_bytes_4096_terrain.md"))
    args = ap.parse_args()
    res = args.res
    # the benchmark's pool, terrain and edit lists (bench.py)
    cfg = W.config_for_resolution(res, bucket_bits=16, wbpp=9, pbpb=2, color_leaf_depth=4,
                                  color_node_capacity=1 << 24, color_leaf_capacity=1 << 28)
    heights = W.value_noise_heightmap(res, 6, 1, res // 8, (res * 3) // 4)
    t0 = time.time()
    pool = O.OraclePool(cfg, heights=heights)
    pool.edit(W.terrain_edit(res, seed=1, skin=8))
    steps = []
    for s in range(args.steps):
        batch = W.terrain_sphere_edits(args.edits, res, heights, 4, 128, seed=1000 + s)
        t = pool.edit_traffic(batch)
        t["edited_voxels"] = W.edited_voxels(batch, res)
        steps.append(t)
        print(f"plain step {s}: {4 * words_of(t) / 1e6:.1f} MB, {t['edited_voxels'] / 1e6:.1f} M voxels "
              f"({time.time() - t0:.0f}s)", file=sys.stderr, flush=True)

    batch = W.mixed_batch(args.batch_edits, res, heights, 2, 256, seed=2)
    fresh = O.OraclePool(cfg, heights=heights)
    fresh.edit(W.terrain_edit(res, seed=1, skin=8))
    tb = fresh.edit_traffic(batch, batch=True)
    prefix = O.OraclePool(cfg, heights=heights)
    prefix.edit(W.terrain_edit(res, seed=1, skin=8))
    tp = prefix.edit_traffic(batch[:args.plain_prefix])
    tpb = None
    if args.plain_prefix:
        q = O.OraclePool(cfg, heights=heights)
        q.edit(W.terrain_edit(res, seed=1, skin=8))
        tpb = q.edit_traffic(batch[:args.plain_prefix], batch=True)
    print(f"batched: {4 * words_of(tb) / 1e6:.1f} MB ({time.time() - t0:.0f}s)", file=sys.stderr)

    with open(args.out, "w") as f:
        f.write("# Algorithmic node-pool bytes of the benchmark's edits (4096^3 terrain)\n\n")
        f.write(f"command: python oracle/edit_bytes.py --steps {args.steps}\n\n")
        f.write("Counted by the CPU oracle's instrumented edit recursion over the benchmark's seeded workloads: "
                "the 4096^3 value-noise terrain (6 octaves, seed 1, z in [512, 3072]), then per plain step 256 "
                "sphere add/remove edits (r log-uniform 4..128, seed 1000 + step) applied one at a time; and "
                "the 4096 mixed edits (seed 2) on a fresh terrain counted as one batch. Bytes = 4 x words of: "
                "existing nodes read + words hashed + words scanned in target buckets + words appended + one "
                "counter word per append. Color-pool words are not part of this figure.\n\n")
        f.write("| step | edited voxels | visited | hashed | scanned | written | counters | upserts | bytes |\n")
        f.write("|---|---|---|---|---|---|---|---|---|\n")
        for i, t in enumerate(steps):
            f.write(f"| {i} | {t['edited_voxels']} | " + " | ".join(str(t[k]) for k in FIELDS)
                    + f" | {4 * words_of(t)} |\n")
        f.write(f"\nbatched (4096 mixed edits, one batch): " + ", ".join(f"{k} {tb[k]}" for k in FIELDS)
                + f", bytes {4 * words_of(tb)}\n")
        f.write(f"plain prefix ({args.plain_prefix} edits, sequential): bytes {4 * words_of(tp)}\n")
        if tpb:
            f.write(f"batched prefix ({args.plain_prefix} edits, one batch): bytes {4 * words_of(tpb)}\n")
        f.write("\nMachine-readable (read by bench.py):\n\n")
        f.write("plain_step_bytes: " + json.dumps([4 * words_of(t) for t in steps]) + "\n")
        f.write("plain_step_voxels: " + json.dumps([t["edited_voxels"] for t in steps]) + "\n")
        f.write(f"batched_bytes: {4 * words_of(tb)}\n")
        f.write(f"plain_prefix_bytes: {4 * words_of(tp)}\n")
        if tpb:
            f.write(f"batched_prefix_bytes: {4 * words_of(tpb)}\n")
    print(f"wrote {args.out}", file=sys.stderr)


if __name__ == "__main__":
    main()
