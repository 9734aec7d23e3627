"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the reference-backed checker
(oracle/_ref/libhdref.so). Checkers only: the product never imports this module."""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from vkhashdag_amd.workloads import EDIT_DTYPE, HD_COLOR_NULL, HD_NULL

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libhdref.so")

u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
editp = np.ctypeslib.ndpointer(EDIT_DTYPE, flags="C_CONTIGUOUS")
VP = C.c_void_p
U32, U64 = C.c_uint32, C.c_uint64


def _build_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"], check=True)


def _sig(lib, name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_olib = None
_rlib = None


def olib():
    global _olib
    if _olib is None:
        _build_oracle()
        L = C.CDLL(ORACLE_SO)
        _sig(L, "hdo_pool_create", VP, [U32, U32, U32, u32p, U32, U32])
        _sig(L, "hdo_pool_destroy", None, [VP])
        _sig(L, "hdo_set_heightmap", None, [VP, u32p, U32])
        _sig(L, "hdo_set_color_caps", None, [VP, U64, U64])
        _sig(L, "hdo_edit", U32, [VP, U32, editp, U32, C.POINTER(U32), VP])
        _sig(L, "hdo_edit_timed", C.c_double, [VP, U32, editp, U32, C.POINTER(U32), C.POINTER(U32)])
        _sig(L, "hdo_gc", None, [VP, u32p, U32])
        _sig(L, "hdo_total_words", U64, [VP])
        _sig(L, "hdo_overflows", U64, [VP])
        _sig(L, "hdo_total_buckets", U64, [VP])
        _sig(L, "hdo_read_words", None, [VP, U64, U64, u32p])
        _sig(L, "hdo_bucket_words", None, [VP, u32p])
        _sig(L, "hdo_filled", U32, [VP, u32p])
        _sig(L, "hdo_make_filled", None, [VP])
        _sig(L, "hdo_level_counts", None, [VP, u64p])
        _sig(L, "hdo_color_node_count", U64, [VP])
        _sig(L, "hdo_color_leaf_count", U64, [VP])
        _sig(L, "hdo_color_nodes", None, [VP, u32p])
        _sig(L, "hdo_color_leaves", None, [VP, u32p])
        _sig(L, "hdo_upsert", U32, [VP, U32, U32, u32p, U32, U32])
        _sig(L, "hdo_hash_span", U32, [u32p, U32])
        _sig(L, "hdo_hash_leaf", U32, [U32, U32])
        _sig(L, "hdo_vbr_write", U64, [VP, U32, u32p, U32, u32p, U64])
        _sig(L, "hdo_canonical_words", U64, [u32p, U64, U32, u32p, U32, u32p, U64])
        _sig(L, "hdo_canonical", U64, [VP, u32p, U32, u32p, U64])
        _sig(L, "hdo_decode_voxels_words", None, [u32p, U64, U32, U32, u8p])
        _sig(L, "hdo_decode_voxels", None, [VP, U32, u8p])
        _sig(L, "hdo_color_fetch_arrays", None, [u32p, U64, u32p, U64, U32, U32, U32, u32p, U64, f32p])
        _sig(L, "hdo_color_fetch", None, [VP, U32, u32p, U64, f32p])
        _sig(L, "hdo_march_rays_words", None, [u32p, U64, U32, U32, C.c_float, C.c_float, f32p, f32p, U64, u32p])
        _sig(L, "hdo_render_arrays", None,
             [u32p, U64, U32, U32, u32p, U64, u32p, U64, U32, U32, f32p, U32, U32, U32, u32p, VP])
        _sig(L, "hdo_dda", None, [u8p, U32, f32p, f32p, U64, u32p])
        _sig(L, "hdo_beam_image", None, [u32p, U64, U32, U32, f32p, U32, U32, f32p])
        _sig(L, "hdo_load_live", None, [VP, u32p, u32p, U64, u32p, U32, u32p, U64, u32p, U64])
        _sig(L, "hdo_render_pool", None, [VP, U32, U32, f32p, U32, U32, U32, u32p, VP, U32])
        _sig(L, "hdo_live_words", U64, [VP, VP, U64])
        _sig(L, "hdo_voxel_query", None, [VP, U32, u32p, U64, u8p])
        _sig(L, "hdo_traffic_begin", None, [VP, U32])
        _sig(L, "hdo_traffic_end", None, [VP, U32, u64p])
        _olib = L
    return _olib


def have_ref() -> bool:
    return os.path.exists(REF_SO)


def rlib():
    global _rlib
    if _rlib is None:
        L = C.CDLL(REF_SO)
        _sig(L, "hdr_pool_create", VP, [U32, U32, U32, u32p, U32, U32, U32, U64, U64])
        _sig(L, "hdr_pool_destroy", None, [VP])
        _sig(L, "hdr_set_heightmap", None, [VP, u32p, U32])
        _sig(L, "hdr_edit", U32, [VP, U32, editp, U32, C.POINTER(U32), U32, U32])
        _sig(L, "hdr_edit_timed", C.c_double, [VP, U32, editp, U32, C.POINTER(U32), U32, U32, C.POINTER(U32)])
        _sig(L, "hdr_gc", None, [VP, u32p, U32])
        _sig(L, "hdr_read_words", None, [VP, U64, U64, u32p])
        _sig(L, "hdr_bucket_words", None, [VP, u32p])
        _sig(L, "hdr_filled", U32, [VP, u32p])
        _sig(L, "hdr_color_node_count", U64, [VP])
        _sig(L, "hdr_color_leaf_count", U64, [VP])
        _sig(L, "hdr_color_nodes", None, [VP, u32p])
        _sig(L, "hdr_color_leaves", None, [VP, u32p])
        _sig(L, "hdr_traversal", None, [VP, U32, f32p, f32p, U64, f32p])
        _sig(L, "hdr_hash_span", U32, [u32p, U32])
        _sig(L, "hdr_hash_leaf", U32, [U32, U32])
        _sig(L, "hdr_vbr_write", U64, [VP, U32, u32p, U32, u32p, U64])
        _sig(L, "hdr_load_live", None, [VP, u32p, u32p, U64, u32p, U32, u32p, U64, u32p, U64])
        _sig(L, "hdr_live_words", U64, [VP, VP, U64])
        _rlib = L
    return _rlib


def _edits(e):
    return np.ascontiguousarray(np.atleast_1d(e).astype(EDIT_DTYPE))


class _PoolBase:
    """Common view of a CPU pool (oracle or reference): words, buckets, colors."""

    def __init__(self, cfg):
        self.cfg = cfg
        self.root = HD_NULL
        self.color_root = HD_COLOR_NULL

    def live_words(self):
        """(bucket_words, concatenated live words of every bucket)."""
        bw = self.bucket_words()
        wpb = self.cfg.words_per_bucket
        parts = []
        for b in np.nonzero(bw)[0]:
            parts.append(self.read_words(int(b) * wpb, int(bw[b])))
        return bw, (np.concatenate(parts) if parts else np.zeros(0, np.uint32))

    def full_words(self):
        return self.read_words(0, self.cfg.total_words)

    def load_state(self, st):
        """Continue from a state dict (vkhashdag_amd.capi.DevicePool.state())."""
        f = olib().hdo_load_live if isinstance(self, OraclePool) else rlib().hdr_load_live
        a = lambda x: np.ascontiguousarray(x if len(x) else np.zeros(1, np.uint32), np.uint32)  # noqa: E731
        cn, cl = st["color_nodes"], st["color_leaves"]
        f(self.h, a(st["bucket_words"]), a(st["live"]), len(st["live"]), a(st["filled"]), len(st["filled"]), a(cn),
          len(cn) // 8, a(cl), len(cl))
        self.root, self.color_root = int(st["root"]), int(st["color_root"])

    def canonical(self, roots=None):
        roots = np.asarray([self.root] if roots is None else roots, np.uint32)
        words = self.full_words()
        return canonical_words(words, self.cfg.node_levels, roots)

    def voxels(self, root=None):
        words = self.full_words()
        return decode_voxels(words, self.cfg.node_levels, self.root if root is None else root)

    def colors(self, vox, root=None):
        n, lv = self.color_nodes(), self.color_leaves()
        return color_fetch(n, lv, self.color_root if root is None else root, self.cfg.voxel_level,
                           self.cfg.color_leaf_level, vox)


class OraclePool(_PoolBase):
    def __init__(self, cfg, heights=None):
        super().__init__(cfg)
        L = olib()
        bits = np.asarray(cfg.bucket_bits, np.uint32)
        self.h = L.hdo_pool_create(cfg.word_bits_per_page, cfg.page_bits_per_bucket, cfg.node_levels, bits,
                                   cfg.color_leaf_level, int(cfg.keep_history))
        assert self.h, "invalid config"
        L.hdo_set_color_caps(self.h, cfg.color_node_capacity, cfg.color_leaf_capacity)
        if heights is not None:
            L.hdo_set_heightmap(self.h, np.ascontiguousarray(heights, np.uint32), heights.shape[0])

    def __del__(self):
        if getattr(self, "h", None):
            olib().hdo_pool_destroy(self.h)
            self.h = None

    def edit(self, e, per_edit=False):
        e = _edits(e)
        cr = U32(self.color_root)
        per = np.zeros(len(e), np.uint32) if per_edit else None
        self.root = olib().hdo_edit(self.h, self.root, e, len(e), C.byref(cr),
                                    per.ctypes.data if per is not None else None)
        self.color_root = cr.value
        return per if per_edit else self.root

    def edit_timed(self, e):
        e = _edits(e)
        cr, ro = U32(self.color_root), U32(0)
        t = olib().hdo_edit_timed(self.h, self.root, e, len(e), C.byref(cr), C.byref(ro))
        self.root, self.color_root = ro.value, cr.value
        return t

    def gc(self, extra_roots=()):
        roots = np.asarray([self.root, *extra_roots], np.uint32)
        olib().hdo_gc(self.h, roots, len(roots))
        self.root = int(roots[0])
        return roots

    def read_words(self, first, count):
        out = np.zeros(count, np.uint32)
        olib().hdo_read_words(self.h, first, count, out)
        return out

    def overflows(self):
        return int(olib().hdo_overflows(self.h))

    def live_words(self):
        """(bucket_words, used words of every bucket packed in bucket order), in one call."""
        n = int(olib().hdo_live_words(self.h, None, 0))
        out = np.zeros(max(1, n), np.uint32)
        olib().hdo_live_words(self.h, out.ctypes.data, n)
        return self.bucket_words(), out[:n]

    def render(self, view, W, H, shadow=True, beam=False, threads=None, root=None, color_root=None):
        """The frame hdo_render_arrays computes, read straight from this pool, over host threads."""
        rgba = np.zeros(W * H, np.uint32)
        rec = np.zeros((W * H, 12), np.uint32)
        flags = (1 if shadow else 0) | (4 if beam else 0)
        threads = threads or min(16, len(os.sched_getaffinity(0)))
        olib().hdo_render_pool(self.h, self.root if root is None else root,
                               self.color_root if color_root is None else color_root,
                               np.ascontiguousarray(view, np.float32), W, H, flags, rgba, rec.ctypes.data, threads)
        return rgba, rec

    def canonical(self, roots=None):
        """Canonical serialisation of the DAG below roots, read straight from this pool."""
        roots = np.ascontiguousarray([self.root] if roots is None else roots, np.uint32)
        cap = 1 << 20
        while True:
            out = np.zeros(cap, np.uint32)
            n = int(olib().hdo_canonical(self.h, roots, len(roots), out, cap))
            if n <= cap:
                return out[:n]
            cap = n

    TRAFFIC_FIELDS = ("visited", "hashed", "scanned", "written", "counters", "upserts")

    def edit_traffic(self, e, batch=False):
        """Applies edits like edit() and returns their algorithmic node-pool traffic in words
        (NodePool::Traffic): per-edit sums, or batch counting (input nodes once, appended nodes
        = the new nodes of the final DAG)."""
        olib().hdo_traffic_begin(self.h, int(batch))
        self.edit(e)
        out = np.zeros(6, np.uint64)
        olib().hdo_traffic_end(self.h, self.root, out)
        return {k: int(v) for k, v in zip(self.TRAFFIC_FIELDS, out)}

    def voxel_query(self, vox, root=None):
        """Occupancy (0/1) of voxels (x, y, z) of the DAG below root."""
        vox = np.ascontiguousarray(vox, np.uint32).reshape(-1, 3)
        out = np.zeros(len(vox), np.uint8)
        olib().hdo_voxel_query(self.h, self.root if root is None else root, vox, len(vox), out)
        return out

    def bucket_words(self):
        out = np.zeros(self.cfg.total_buckets, np.uint32)
        olib().hdo_bucket_words(self.h, out)
        return out

    def filled(self):
        out = np.zeros(64, np.uint32)
        n = olib().hdo_filled(self.h, out)
        return out[:n]

    def level_counts(self):
        out = np.zeros(self.cfg.node_levels, np.uint64)
        olib().hdo_level_counts(self.h, out)
        return out

    def color_nodes(self):
        n = olib().hdo_color_node_count(self.h)
        out = np.zeros(max(1, 8 * n), np.uint32)
        olib().hdo_color_nodes(self.h, out)
        return out[: 8 * n]

    def color_leaves(self):
        n = olib().hdo_color_leaf_count(self.h)
        out = np.zeros(max(1, n), np.uint32)
        olib().hdo_color_leaves(self.h, out)
        return out[:n]

    def upsert(self, level, node, leaf=False, fallback=HD_NULL):
        node = np.ascontiguousarray(node, np.uint32)
        return olib().hdo_upsert(self.h, level, int(leaf), node, len(node), fallback)


class RefPool(_PoolBase):
    """The reference's NodePoolBase/ThreadedEdit/ThreadedGC over in-memory storage."""

    def __init__(self, cfg, heights=None, workers=1):
        super().__init__(cfg)
        L = rlib()
        bits = np.asarray(cfg.bucket_bits, np.uint32)
        self.h = L.hdr_pool_create(cfg.word_bits_per_page, cfg.page_bits_per_bucket, cfg.node_levels, bits,
                                   cfg.color_leaf_level, int(cfg.keep_history), workers,
                                   cfg.color_node_capacity, cfg.color_leaf_capacity)
        assert self.h
        if heights is not None:
            L.hdr_set_heightmap(self.h, np.ascontiguousarray(heights, np.uint32), heights.shape[0])

    def close(self):
        """Destroys the pool and joins its worker threads."""
        if getattr(self, "h", None):
            rlib().hdr_pool_destroy(self.h)
            self.h = None

    __del__ = close

    def edit(self, e, threaded=False, max_task_level=None):
        e = _edits(e)
        cr = U32(self.color_root)
        mtl = self.cfg.color_leaf_level if max_task_level is None else max_task_level
        self.root = rlib().hdr_edit(self.h, self.root, e, len(e), C.byref(cr), int(threaded), mtl)
        self.color_root = cr.value
        return self.root

    def edit_timed(self, e, threaded=True, max_task_level=None):
        e = _edits(e)
        cr, ro = U32(self.color_root), U32(0)
        mtl = self.cfg.color_leaf_level if max_task_level is None else max_task_level
        t = rlib().hdr_edit_timed(self.h, self.root, e, len(e), C.byref(cr), int(threaded), mtl, C.byref(ro))
        self.root, self.color_root = ro.value, cr.value
        return t

    def gc(self, extra_roots=()):
        roots = np.asarray([self.root, *extra_roots], np.uint32)
        rlib().hdr_gc(self.h, roots, len(roots))
        self.root = int(roots[0])
        return roots

    def read_words(self, first, count):
        out = np.zeros(count, np.uint32)
        rlib().hdr_read_words(self.h, first, count, out)
        return out

    def bucket_words(self):
        out = np.zeros(self.cfg.total_buckets, np.uint32)
        rlib().hdr_bucket_words(self.h, out)
        return out

    def filled(self):
        out = np.zeros(64, np.uint32)
        n = rlib().hdr_filled(self.h, out)
        return out[:n]

    def color_nodes(self):
        n = rlib().hdr_color_node_count(self.h)
        out = np.zeros(max(1, 8 * n), np.uint32)
        rlib().hdr_color_nodes(self.h, out)
        return out[: 8 * n]

    def color_leaves(self):
        n = rlib().hdr_color_leaf_count(self.h)
        out = np.zeros(max(1, n), np.uint32)
        rlib().hdr_color_leaves(self.h, out)
        return out[:n]

    def live_words(self):
        n = int(rlib().hdr_live_words(self.h, None, 0))
        out = np.zeros(max(1, n), np.uint32)
        rlib().hdr_live_words(self.h, out.ctypes.data, n)
        return self.bucket_words(), out[:n]

    def state(self):
        """Everything a pool needs to continue from this state (DevicePool.state()'s layout)."""
        bw, live = self.live_words()
        return dict(bucket_words=bw, live=live, filled=self.filled(), color_nodes=self.color_nodes(),
                    color_leaves=self.color_leaves(), root=self.root, color_root=self.color_root)

    def traversal(self, o, d):
        o = np.ascontiguousarray(o, np.float32)
        d = np.ascontiguousarray(d, np.float32)
        out = np.zeros((len(o), 4), np.float32)
        rlib().hdr_traversal(self.h, self.root, o, d, len(o), out)
        return out


# ------------------------------------------------------------------ array checkers
def canonical_words(words, L, roots) -> np.ndarray:
    words = np.ascontiguousarray(words, np.uint32)
    roots = np.ascontiguousarray(roots, np.uint32)
    cap = 1 << 20
    while True:
        out = np.zeros(cap, np.uint32)
        n = olib().hdo_canonical_words(words, len(words), L, roots, len(roots), out, cap)
        if n <= cap:
            return out[:n]
        cap = int(n)


def decode_voxels(words, L, root) -> np.ndarray:
    res = 1 << (L + 1)
    g = np.zeros(res ** 3, np.uint8)
    olib().hdo_decode_voxels_words(np.ascontiguousarray(words, np.uint32), len(words), L, root, g)
    return g.reshape(res, res, res)  # [z, y, x]


def color_fetch(nodes, leaves, root, voxel_level, leaf_level, vox) -> np.ndarray:
    vox = np.ascontiguousarray(vox, np.uint32).reshape(-1, 3)
    out = np.zeros((len(vox), 3), np.float32)
    nodes = np.ascontiguousarray(nodes, np.uint32)
    leaves = np.ascontiguousarray(leaves, np.uint32)
    olib().hdo_color_fetch_arrays(nodes if len(nodes) else np.zeros(8, np.uint32), len(nodes) // 8,
                                  leaves if len(leaves) else np.zeros(1, np.uint32), len(leaves), root,
                                  voxel_level, leaf_level, vox, len(vox), out)
    return out


def march_rays(words, L, root, o, d, proj_factor=float("inf"), proj_bias=0.0):
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    rec = np.zeros((len(o), 12), np.uint32)
    olib().hdo_march_rays_words(np.ascontiguousarray(words, np.uint32), len(words), L, root, proj_factor,
                                proj_bias, o, d, len(o), rec)
    return rec


def render(words, L, root, cnodes, cleaves, color_root, color_leaf_level, view, W, H, shadow=True, beam=False):
    rgba = np.zeros(W * H, np.uint32)
    rec = np.zeros((W * H, 12), np.uint32)
    cn = np.ascontiguousarray(cnodes, np.uint32)
    cl = np.ascontiguousarray(cleaves, np.uint32)
    flags = (1 if shadow else 0) | (4 if beam else 0)
    olib().hdo_render_arrays(np.ascontiguousarray(words, np.uint32), len(words), L, root,
                             cn if len(cn) else np.zeros(8, np.uint32), len(cn) // 8,
                             cl if len(cl) else np.zeros(1, np.uint32), len(cl), color_root, color_leaf_level,
                             np.ascontiguousarray(view, np.float32), W, H, flags, rgba, rec.ctypes.data)
    return rgba, rec


def beam_image(words, L, root, view, W, H):
    """beam.frag's image for a W x H view: ceil(W/8) x ceil(H/8) start distances (+inf = miss)."""
    wb, hb = (W + 7) // 8, (H + 7) // 8
    out = np.zeros(wb * hb, np.float32)
    olib().hdo_beam_image(np.ascontiguousarray(words, np.uint32), len(words), L, root,
                          np.ascontiguousarray(view, np.float32), W, H, out)
    return out.reshape(hb, wb)


def dda(grid, o, d):
    res = grid.shape[0]
    o = np.ascontiguousarray(o, np.float32)
    d = np.ascontiguousarray(d, np.float32)
    out = np.zeros((len(o), 4), np.uint32)
    olib().hdo_dda(np.ascontiguousarray(grid.reshape(-1), np.uint8), res, o, d, len(o), out)
    return out


def vbr_write(ops, src=None, lib="oracle"):
    ops = np.ascontiguousarray(ops, np.uint32).reshape(-1, 4)
    src_arr = np.ascontiguousarray(src if src is not None else np.zeros(1, np.uint32), np.uint32)
    cap = 1 << 22
    out = np.zeros(cap, np.uint32)
    f = olib().hdo_vbr_write if lib == "oracle" else rlib().hdr_vbr_write
    n = f(src_arr.ctypes.data if src is not None else None, len(src_arr) if src is not None else 0, ops, len(ops),
          out, cap)
    return out[:n]
