"""ctypes binding of include/hd_capi.h and a thin host object mirroring the reference interface
(hashdag::NodePoolBase::Edit / NodePoolThreadedEdit::ThreadedEdit / NodePoolThreadedGC::ThreadedGC /
NodePoolTraversal::Traversal and the DAGColorPool root). The library is the HIP build
vkhashdag_amd/libhashdag_amd.so; there is no CPU fallback: loading fails loudly without it."""
from __future__ import annotations

import ctypes as C
import math
import os

import numpy as np

from .workloads import EDIT_DTYPE, HD_COLOR_NULL, HD_NULL, PoolConfig

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libhashdag_amd.so")
HD_MAX_LEVELS = 24
MODE_PLAIN, MODE_BATCHED = 0, 1
TRACE_SHADOW, TRACE_RECORDS, TRACE_BEAM = 1, 2, 4


class HDError(RuntimeError):
    pass


SERVE_PLAIN = 1


class hd_serve_config(C.Structure):
    _fields_ = [("max_edits_per_tick", C.c_uint32), ("flags", C.c_uint32), ("max_pixels_per_tick", C.c_uint64),
                ("min_free_pages", C.c_uint64)]


FRAME_COLORS, FRAME_GC = 1, 2


class hd_frame_config(C.Structure):
    _fields_ = [("max_edits", C.c_uint32), ("max_passes", C.c_uint32), ("max_radius", C.c_uint32),
                ("n_views", C.c_uint32), ("max_width", C.c_uint32), ("max_height", C.c_uint32),
                ("flags", C.c_uint32), ("pad_", C.c_uint32)]


class hd_serve_stats(C.Structure):
    _fields_ = [("ticks", C.c_uint32), ("edits", C.c_uint32), ("renders", C.c_uint32), ("gcs", C.c_uint32),
                ("gated_ticks", C.c_uint32), ("max_edits_in_tick", C.c_uint32),
                ("max_renders_in_tick", C.c_uint32), ("pad_", C.c_uint32), ("min_free_pages", C.c_uint64),
                ("wall_ms", C.c_double)]


class hd_config(C.Structure):
    _fields_ = [("word_bits_per_page", C.c_uint32), ("page_bits_per_bucket", C.c_uint32),
                ("node_levels", C.c_uint32), ("bucket_bits", C.c_uint32 * HD_MAX_LEVELS),
                ("color_leaf_level", C.c_uint32), ("color_keep_history", C.c_uint32),
                ("color_node_capacity", C.c_uint64), ("color_leaf_capacity", C.c_uint64)]


class hd_stats(C.Structure):
    _fields_ = [("total_words", C.c_uint64), ("total_buckets", C.c_uint64), ("used_words", C.c_uint64),
                ("color_nodes", C.c_uint64), ("color_leaf_words", C.c_uint64), ("overflow", C.c_uint32),
                ("node_levels", C.c_uint32), ("level_used_words", C.c_uint64 * HD_MAX_LEVELS),
                ("last_entries", C.c_uint64), ("last_upserts", C.c_uint64), ("last_jobs", C.c_uint64),
                ("last_ms", C.c_double), ("last_groups", C.c_uint32), ("pad_", C.c_uint32)]


class hd_tuning(C.Structure):
    _fields_ = [("trace_tile", C.c_uint32), ("top_pairs", C.c_uint32), ("small_level", C.c_uint32),
                ("batch_pass_edits", C.c_uint32)]


TUNING_FIELDS = tuple(f for f, _ in hd_tuning._fields_)


EDIT_NODE_FN = C.CFUNCTYPE(C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32)
EDIT_VOXEL_FN = C.CFUNCTYPE(C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32)
EDIT_TYPE_NOT_AFFECTED, EDIT_TYPE_PROCEED, EDIT_TYPE_FILL, EDIT_TYPE_CLEAR = range(4)


class hd_host_editor(C.Structure):
    _fields_ = [("user", C.c_void_p), ("edit_node", EDIT_NODE_FN), ("edit_voxel", EDIT_VOXEL_FN)]


class hd_view(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("look", C.c_float * 3), ("side", C.c_float * 3), ("up", C.c_float * 3),
                ("proj_factor", C.c_float), ("width", C.c_uint32), ("height", C.c_uint32), ("flags", C.c_uint32),
                ("out_offset", C.c_uint64)]


assert C.sizeof(hd_view) == 72

_lib = None
VP, U32, U64 = C.c_void_p, C.c_uint32, C.c_uint64
P32 = C.POINTER(C.c_uint32)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HDError(f"{LIB_PATH} is missing: build it with `python -m vkhashdag_amd.build`")
        L = C.CDLL(LIB_PATH)
        sigs = {
            "hd_last_error": (C.c_char_p, []),
            "hd_pool_create": (C.c_int, [C.POINTER(hd_config), C.POINTER(VP)]),
            "hd_pool_destroy": (C.c_int, [VP]),
            "hd_pool_sync": (C.c_int, [VP]),
            "hd_pool_stream": (VP, [VP]),
            "hd_pool_set_heightmap": (C.c_int, [VP, VP, U32]),
            "hd_pool_value_noise_heightmap": (C.c_int, [VP, U32, U64, U32, U32, U32, C.c_double, C.c_double, VP]),
            "hd_pool_read_words": (C.c_int, [VP, U64, U64, VP]),
            "hd_pool_write_words": (C.c_int, [VP, U64, U64, VP]),
            "hd_pool_read_bucket_words": (C.c_int, [VP, U64, U64, VP]),
            "hd_pool_write_bucket_words": (C.c_int, [VP, U64, U64, VP]),
            "hd_pool_read_color_nodes": (C.c_int, [VP, U64, U64, VP]),
            "hd_pool_read_color_leaves": (C.c_int, [VP, U64, U64, VP]),
            "hd_pool_filled": (C.c_int, [VP, VP, U32, P32]),
            "hd_pool_stats": (C.c_int, [VP, C.POINTER(hd_stats)]),
            "hd_pool_clear_overflow": (C.c_int, [VP]),
            "hd_edit": (C.c_int, [VP, U32, U32, VP, U32, U32, P32, P32]),
            "hd_edit_reserve": (C.c_int, [VP, VP, U32, U32]),
            "hd_gc": (C.c_int, [VP, VP, U32]),
            "hd_edit_host": (C.c_int, [VP, U32, C.POINTER(hd_host_editor), P32]),
            "hd_trace": (C.c_int, [VP, U32, U32, VP, U32, VP, VP]),
            "hd_trace_noalloc": (C.c_int, [VP, U32, U32, VP, U32, VP, VP]),
            "hd_trace_async": (C.c_int, [VP, U32, U32, VP, VP, U32, U32, VP, VP, VP]),
            "hd_timer_begin": (C.c_int, [VP]),
            "hd_timer_end": (C.c_int, [VP, C.POINTER(C.c_float)]),
            "hd_pool_read_live_words": (C.c_int, [VP, VP, U64, C.POINTER(U64)]),
            "hd_pool_write_live_words": (C.c_int, [VP, VP, VP, U64]),
            "hd_pool_save": (C.c_int, [VP, C.c_char_p, VP, U32]),
            "hd_pool_load": (C.c_int, [C.c_char_p, C.POINTER(VP), VP, U32, C.POINTER(U32)]),
            "hd_pool_get_config": (C.c_int, [VP, C.POINTER(hd_config)]),
            "hd_device_name": (C.c_int, [C.c_int, C.c_char_p, U32]),
            "hd_tuning_get": (C.c_int, [C.POINTER(hd_tuning)]),
            "hd_tuning_set": (C.c_int, [C.POINTER(hd_tuning)]),
            "hd_trace_tiles": (C.c_int, [VP, U32, VP]),
            "hd_trace_beam_sizes": (C.c_int, [VP, U32, VP]),
            "hd_trace_beam_async": (C.c_int, [VP, U32, U32, VP, VP, VP, U32, U32, U32, VP, VP]),
            "hd_traversal": (C.c_int, [VP, U32, VP, VP, U64, VP]),
            "hd_device_alloc": (C.c_int, [U64, C.POINTER(VP)]),
            "hd_device_free": (C.c_int, [VP]),
            "hd_memcpy_h2d": (C.c_int, [VP, VP, U64]),
            "hd_memcpy_d2h": (C.c_int, [VP, VP, U64]),
            "hd_device_count": (C.c_int, [C.POINTER(C.c_int)]),
            "hd_device_copy": (C.c_int, [VP, VP, U64, VP]),
            "hd_frame_create": (C.c_int, [VP, C.POINTER(hd_frame_config), U32, U32, C.POINTER(VP)]),
            "hd_frame_set": (C.c_int, [VP, VP, U32, VP, U32]),
            "hd_frame_set_roots": (C.c_int, [VP, U32, U32]),
            "hd_frame_launch": (C.c_int, [VP]),
            "hd_frame_wait": (C.c_int, [VP, P32, P32]),
            "hd_frame_rgba": (VP, [VP]),
            "hd_frame_destroy": (C.c_int, [VP]),
            "hd_serve_run": (C.c_int, [VP, C.POINTER(hd_serve_config), U32, U32, VP, U32, VP, U64, VP, P32, P32,
                                       C.POINTER(hd_serve_stats)]),
            "hd_pool_free_pages": (C.c_int, [VP, C.POINTER(U64), C.POINTER(U64)]),
        }
        for name, (res, args) in sigs.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise HDError(f"hd error {rc}: {lib().hd_last_error().decode()}")


class DeviceBuffer:
    """A device allocation (hd_device_alloc); freed with the object."""

    def __init__(self, nbytes):
        p = VP()
        check(lib().hd_device_alloc(max(4, int(nbytes)), C.byref(p)))
        self.ptr, self.nbytes = p.value, int(nbytes)

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes
        check(lib().hd_memcpy_h2d(self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, dtype, count):
        out = np.empty(count, dtype)
        check(lib().hd_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def __del__(self):
        if getattr(self, "ptr", None):
            lib().hd_device_free(self.ptr)
            self.ptr = None


def make_view(vb, width, height, flags=0, out_offset=0) -> hd_view:
    """vb: 13 floats (pos, look, side, up, proj_factor) as produced by workloads.orbit_view."""
    v = hd_view()
    vb = np.asarray(vb, np.float32)
    for i in range(3):
        v.pos[i], v.look[i], v.side[i], v.up[i] = vb[i], vb[3 + i], vb[6 + i], vb[9 + i]
    v.proj_factor = float(vb[12])
    v.width, v.height, v.flags, v.out_offset = width, height, flags, out_offset
    return v


class DevicePool:
    """A device-resident HashDAG node pool + color pool (DAGNodePool + DAGColorPool semantics)."""

    def __init__(self, cfg: PoolConfig, heights=None):
        self.cfg = cfg
        c = hd_config()
        c.word_bits_per_page = cfg.word_bits_per_page
        c.page_bits_per_bucket = cfg.page_bits_per_bucket
        c.node_levels = cfg.node_levels
        for i, b in enumerate(cfg.bucket_bits):
            c.bucket_bits[i] = b
        c.color_leaf_level = cfg.color_leaf_level
        c.color_keep_history = int(cfg.keep_history)
        c.color_node_capacity = cfg.color_node_capacity
        c.color_leaf_capacity = cfg.color_leaf_capacity
        h = VP()
        check(lib().hd_pool_create(C.byref(c), C.byref(h)))
        self.h = h.value
        self.root = HD_NULL
        self.color_root = HD_COLOR_NULL
        if heights is not None:
            self.set_heightmap(heights)

    # -- checkpoint dump / load (pool words + counters, color pool, roots)
    def save(self, path, extra_roots=()):
        """Write the pool and its roots (root, color_root, *extra_roots) to one file."""
        roots = np.asarray([self.root, self.color_root, *extra_roots], np.uint32)
        check(lib().hd_pool_save(self.h, os.fsencode(path), roots.ctypes.data, len(roots)))

    @classmethod
    def load(cls, path, heights=None, max_roots=64):
        """A new pool from a file written by save(); returns (pool, extra_roots)."""
        h, n = VP(), U32(0)
        roots = np.zeros(max_roots, np.uint32)
        check(lib().hd_pool_load(os.fsencode(path), C.byref(h), roots.ctypes.data, max_roots, C.byref(n)))
        c = hd_config()
        rc = lib().hd_pool_get_config(h.value, C.byref(c))
        if rc != 0:
            lib().hd_pool_destroy(h.value)
            check(rc)
        self = cls.__new__(cls)
        self.h = h.value
        self.cfg = PoolConfig(c.word_bits_per_page, c.page_bits_per_bucket, list(c.bucket_bits[: c.node_levels]),
                              c.color_leaf_level, bool(c.color_keep_history), c.color_node_capacity,
                              c.color_leaf_capacity)
        if n.value < 2:
            self.close()
            raise HDError("checkpoint holds fewer than two roots (root, color_root)")
        self.root, self.color_root = int(roots[0]), int(roots[1])
        if heights is not None:
            self.set_heightmap(heights)
        return self, [int(r) for r in roots[2: n.value]]

    def close(self):
        if getattr(self, "h", None):
            lib().hd_pool_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def value_noise_heightmap(self, octaves, seed, zlo, zhi, base_cells=4, lacunarity=2.0, gain=0.5,
                              download=True):
        """Generates workloads.value_noise_heightmap's terrain on the device (identical heights) and
        installs it as the column-fill input; returns the heights (host) when download."""
        res = self.cfg.resolution
        out = np.empty((res, res), np.uint32) if download else None
        check(lib().hd_pool_value_noise_heightmap(self.h, octaves, seed, zlo, zhi, base_cells, lacunarity, gain,
                                                  out.ctypes.data if download else None))
        return out

    def set_heightmap(self, heights):
        heights = np.ascontiguousarray(heights, np.uint32)
        check(lib().hd_pool_set_heightmap(self.h, heights.ctypes.data, heights.shape[0]))

    # -- NodePoolBase::Edit / ThreadedEdit (one device pass per edit) and the batched pass
    def edit(self, edits, mode=MODE_PLAIN):
        e = np.ascontiguousarray(np.atleast_1d(edits).astype(EDIT_DTYPE))
        r, cr = U32(0), U32(0)
        check(lib().hd_edit(self.h, self.root, self.color_root, e.ctypes.data, len(e), mode, C.byref(r),
                            C.byref(cr)))
        self.root, self.color_root = r.value, cr.value
        return self.root

    Edit = edit
    ThreadedEdit = edit

    def reserve(self, edits, mode=MODE_PLAIN):
        """hd_edit_reserve: allocate what edit(edits, mode) would, without editing."""
        e = np.ascontiguousarray(np.atleast_1d(edits).astype(EDIT_DTYPE))
        check(lib().hd_edit_reserve(self.h, e.ctypes.data, len(e), mode))

    def gc(self, extra_roots=()):
        roots = np.asarray([self.root, *extra_roots], np.uint32)
        check(lib().hd_gc(self.h, roots.ctypes.data, len(roots)))
        self.root = int(roots[0])
        return roots

    ThreadedGC = gc

    def edit_host(self, edit_node, edit_voxel):
        """NodePoolBase::Edit with a generic stateless editor evaluated on the host (hd_edit_host):
        edit_node(level, x, y, z, node_ptr) -> EDIT_TYPE_*, edit_voxel(x, y, z, voxel) -> bool."""
        fn = EDIT_NODE_FN(lambda _u, l, x, y, z, p: int(edit_node(l, x, y, z, p)))
        fv = EDIT_VOXEL_FN(lambda _u, x, y, z, v: 1 if edit_voxel(x, y, z, bool(v)) else 0)
        ed = hd_host_editor(None, fn, fv)
        r = U32(0)
        check(lib().hd_edit_host(self.h, self.root, C.byref(ed), C.byref(r)))
        self.root = r.value
        return self.root

    # -- host interop
    def read_words(self, first, count):
        out = np.empty(count, np.uint32)
        check(lib().hd_pool_read_words(self.h, first, count, out.ctypes.data))
        return out

    def full_words(self):
        return self.read_words(0, self.cfg.total_words)

    def bucket_words(self):
        out = np.empty(self.cfg.total_buckets, np.uint32)
        check(lib().hd_pool_read_bucket_words(self.h, 0, len(out), out.ctypes.data))
        return out

    def live_words(self):
        """(per-bucket word counters, used words of every bucket packed in bucket order)."""
        bw = self.bucket_words()
        n = U64(0)
        check(lib().hd_pool_read_live_words(self.h, None, 0, C.byref(n)))
        out = np.empty(max(1, n.value), np.uint32)
        check(lib().hd_pool_read_live_words(self.h, out.ctypes.data, len(out), C.byref(n)))
        return bw, out[: n.value]

    def write_live_words(self, bucket_words, live):
        bucket_words = np.ascontiguousarray(bucket_words, np.uint32)
        live = np.ascontiguousarray(live, np.uint32)
        check(lib().hd_pool_write_live_words(self.h, bucket_words.ctypes.data, live.ctypes.data, len(live)))

    def state(self):
        """Everything a CPU pool needs to continue from this state."""
        bw, live = self.live_words()
        return dict(bucket_words=bw, live=live, filled=self.filled(), color_nodes=self.color_nodes(),
                    color_leaves=self.color_leaves(), root=self.root, color_root=self.color_root)

    def timer_begin(self):
        check(lib().hd_timer_begin(self.h))

    def timer_end(self):
        ms = C.c_float(0)
        check(lib().hd_timer_end(self.h, C.byref(ms)))
        return ms.value

    def sync(self):
        check(lib().hd_pool_sync(self.h))

    def filled(self):
        out = np.zeros(64, np.uint32)
        n = U32(0)
        check(lib().hd_pool_filled(self.h, out.ctypes.data, 64, C.byref(n)))
        return out[: n.value]

    def stats(self):
        s = hd_stats()
        check(lib().hd_pool_stats(self.h, C.byref(s)))
        return s

    def color_nodes(self):
        n = self.stats().color_nodes
        out = np.zeros(max(8, 8 * n), np.uint32)
        if n:
            check(lib().hd_pool_read_color_nodes(self.h, 0, n, out.ctypes.data))
        return out[: 8 * n]

    def color_leaves(self):
        n = self.stats().color_leaf_words
        out = np.zeros(max(1, n), np.uint32)
        if n:
            check(lib().hd_pool_read_color_leaves(self.h, 0, n, out.ctypes.data))
        return out[:n]

    # -- rendering (trace.frag) and picking (NodePoolTraversal)
    def render(self, vb, width, height, shadow=True, records=False, root=None, color_root=None, beam=False):
        flags = (TRACE_SHADOW if shadow else 0) | (TRACE_RECORDS if records else 0) | (TRACE_BEAM if beam else 0)
        v = make_view(vb, width, height, flags)
        npx = width * height
        rgba = DeviceBuffer(npx * 4)
        rec = DeviceBuffer(npx * 48) if records else None
        check(lib().hd_trace(self.h, self.root if root is None else root,
                             self.color_root if color_root is None else color_root, C.byref(v), 1, rgba.ptr,
                             rec.ptr if rec else None))
        out = rgba.download(np.uint32, npx)
        if records:
            return out, rec.download(np.uint32, npx * 12).reshape(npx, 12)
        return out

    def traversal(self, o, d, root=None):
        o = np.ascontiguousarray(o, np.float32).reshape(-1, 3)
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        n = len(o)
        bo, bd, bout = DeviceBuffer(o.nbytes), DeviceBuffer(d.nbytes), DeviceBuffer(n * 16)
        bo.upload(o)
        bd.upload(d)
        check(lib().hd_traversal(self.h, self.root if root is None else root, bo.ptr, bd.ptr, n, bout.ptr))
        return bout.download(np.float32, 4 * n).reshape(n, 4)

    Traversal = traversal

    # -- serving: a request loop over this pool (hd_serve_run)
    def free_pages(self):
        f, t = U64(0), U64(0)
        check(lib().hd_pool_free_pages(self.h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def serve(self, requests, max_edits_per_tick=64, max_pixels_per_tick=1 << 24, min_free_pages=0, plain=False,
              keep_images=False):
        """Serves a request trace (workloads.serving_trace). Returns (results, stats, rgba buffer or
        None); results follow workloads.RESULT_DTYPE."""
        from .workloads import REQUEST_DTYPE, RESULT_DTYPE, REQ_RENDER
        reqs = np.ascontiguousarray(requests, REQUEST_DTYPE)
        res = np.zeros(len(reqs), RESULT_DTYPE)
        cfg = hd_serve_config(max_edits_per_tick, SERVE_PLAIN if plain else 0, max_pixels_per_tick, min_free_pages)
        px = int(sum(int(r["view"]["width"]) * int(r["view"]["height"]) for r in reqs if r["kind"] == REQ_RENDER))
        buf = DeviceBuffer(max(1, px) * 4) if keep_images else None
        st = hd_serve_stats()
        r, cr = U32(0), U32(0)
        check(lib().hd_serve_run(self.h, C.byref(cfg), self.root, self.color_root, reqs.ctypes.data, len(reqs),
                                 buf.ptr if buf else None, px if buf else 0, res.ctypes.data, C.byref(r),
                                 C.byref(cr), C.byref(st)))
        self.root, self.color_root = r.value, cr.value
        return res, st, buf


class Frame:
    """An edit -> render frame captured as a hipGraph (hd_frame_*): set edits + views, launch, wait."""

    def __init__(self, pool, max_edits, max_passes, max_radius, n_views, max_width, max_height, colors=True,
                 gc=False):
        self.pool = pool
        cfg = hd_frame_config(max_edits, max_passes, max_radius, n_views, max_width, max_height,
                              (FRAME_COLORS if colors else 0) | (FRAME_GC if gc else 0), 0)
        h = VP()
        check(lib().hd_frame_create(pool.h, C.byref(cfg), pool.root, pool.color_root, C.byref(h)))
        self.h = h.value
        self.views = []

    def set(self, edits, views):
        """edits: EDIT_DTYPE records; views: list of (13-float view block, width, height, flags)."""
        e = np.ascontiguousarray(np.atleast_1d(edits).astype(EDIT_DTYPE)) if len(edits) else np.zeros(0, EDIT_DTYPE)
        vs = (hd_view * max(1, len(views)))()
        for i, (vb, w, hh, fl) in enumerate(views):
            vs[i] = make_view(vb, w, hh, fl)
        self.views = [(w, hh) for (_, w, hh, _) in views]
        check(lib().hd_frame_set(self.h, e.ctypes.data if len(e) else None, len(e), vs, len(views)))

    def launch(self):
        check(lib().hd_frame_launch(self.h))

    def wait(self):
        r, cr = U32(0), U32(0)
        check(lib().hd_frame_wait(self.h, C.byref(r), C.byref(cr)))
        self.pool.root, self.pool.color_root = r.value, cr.value
        return r.value, cr.value

    def set_roots(self, root, color_root):
        check(lib().hd_frame_set_roots(self.h, root, color_root))

    def images(self):
        """RGBA8 pixels of the last frame, one array per view."""
        n = sum(w * h for w, h in self.views)
        out = np.empty(max(1, n), np.uint32)
        ptr = lib().hd_frame_rgba(self.h)
        check(lib().hd_memcpy_d2h(out.ctypes.data, ptr, out.nbytes))
        res, off = [], 0
        for w, h in self.views:
            res.append(out[off:off + w * h])
            off += w * h
        return res

    def close(self):
        if getattr(self, "h", None):
            lib().hd_frame_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def get_tuning() -> dict:
    """The launch parameters in effect (hd_tuning): speed only, never results."""
    t = hd_tuning()
    check(lib().hd_tuning_get(C.byref(t)))
    return {f: int(getattr(t, f)) for f in TUNING_FIELDS}


def set_tuning(**kw) -> dict:
    """Sets some launch parameters (the others keep their values); returns the previous set."""
    old = get_tuning()
    unknown = set(kw) - set(TUNING_FIELDS)
    if unknown:
        raise HDError(f"unknown tuning fields {sorted(unknown)}")
    t = hd_tuning(**{**old, **{k: int(v) for k, v in kw.items()}})
    check(lib().hd_tuning_set(C.byref(t)))
    return old


def device_name(dev: int = 0) -> str:
    buf = C.create_string_buffer(256)
    check(lib().hd_device_name(dev, buf, 256))
    return buf.value.decode()


def device_count() -> int:
    n = C.c_int(0)
    try:
        check(lib().hd_device_count(C.byref(n)))
    except HDError:
        return 0
    return n.value
