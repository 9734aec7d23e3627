"""Writes the golden fixtures of tests/golden/ from the compiled reference (oracle/_ref/libhdref.so,
built by oracle/Makefile from the VkHashDAG headers). Run in a tree where the reference exists:

    make -C oracle && python tests/golden/make_golden.py

Fixtures are inputs + the reference's outputs only (npz, no pickles)."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import _oracle as O  # noqa: E402
from vkhashdag_amd import workloads as W  # noqa: E402
from test_oracle_pins import _random_ops, random_edits, small_cfg  # noqa: E402


def main():
    assert O.have_ref(), "build oracle/_ref first"
    R = O.rlib()
    rng = np.random.default_rng(1234)
    # hash
    leaf_in = rng.integers(0, 1 << 32, (256, 2), dtype=np.uint64).astype(np.uint32)
    leaf_out = np.array([R.hdr_hash_leaf(int(a), int(b)) for a, b in leaf_in], np.uint32)
    span_len = rng.integers(0, 10, 256).astype(np.uint32)
    span_in = rng.integers(0, 1 << 32, (256, 9), dtype=np.uint64).astype(np.uint32)
    span_out = np.array([R.hdr_hash_span(np.ascontiguousarray(span_in[i, :n]) if n else np.zeros(1, np.uint32),
                                         int(n)) for i, n in enumerate(span_len)], np.uint32)
    np.savez_compressed(os.path.join(HERE, "hash.npz"), leaf_in=leaf_in, leaf_out=leaf_out, span_len=span_len,
                        span_in=span_in, span_out=span_out)
    # VBR writer streams
    vbr = {}
    n = 0
    for trial in range(12):
        weights = trial % 3 == 2
        n_vox = int(rng.choice([64, 4096, 32768]))
        src = O.vbr_write(_random_ops(rng, n_vox, weights), None, "ref") if trial % 2 else None
        ops = _random_ops(rng, n_vox, weights)
        vbr[f"ops{n}"] = ops
        if src is not None:
            vbr[f"src{n}"] = src
        vbr[f"out{n}"] = O.vbr_write(ops, src, "ref")
        n += 1
    np.savez_compressed(os.path.join(HERE, "vbr.npz"), n=n, **vbr)
    # Edit sequence + GC on a 64^3 pool (single-threaded reference Edit, one-worker ThreadedGC)
    cfg = small_cfg(64, 6, 5, 2, 3)
    e = random_edits(24, 64, 77, True)
    r = O.RefPool(cfg, workers=1)
    roots = []
    for i in range(len(e)):
        r.edit(e[i:i + 1])
        roots.append((r.root, r.color_root))
    bw, lw = r.live_words()
    cn, cl = r.color_nodes(), r.color_leaves()
    words = r.full_words()
    r.gc()
    gbw, glw = r.live_words()
    np.savez_compressed(os.path.join(HERE, "edits64.npz"), edits=e.view(np.uint8), roots=np.array(roots, np.uint32),
                        bucket_words=bw, live_words=lw, color_nodes=cn, color_leaves=cl, filled=r.filled(),
                        gc_root=np.uint32(r.root), gc_bucket_words=gbw, gc_live_words=glw)
    # Traversal: NodePoolTraversal::Traversal<float> on the pre-GC pool words
    ro = O.RefPool(cfg, workers=1)
    ro.edit(e)
    o = rng.uniform(-0.2, 1.2, (2000, 3)).astype(np.float32)
    tgt = rng.uniform(0.2, 0.8, (2000, 3)).astype(np.float32)
    d = tgt - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    hits = ro.traversal(o, d)
    np.savez_compressed(os.path.join(HERE, "traversal64.npz"), o=o, d=d, hits=hits, words=words,
                        root=np.uint32(ro.root))
    print("golden fixtures written")


if __name__ == "__main__":
    main()
