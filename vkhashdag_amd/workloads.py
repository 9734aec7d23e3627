"""Seeded synthetic workloads for the edit -> GC -> traversal path.

Everything here is plain numpy on the host: the same POD edit lists (``hd_edit_desc``) and the
same heightfield are fed to the device path, the CPU oracle and the reference-backed checker.

PRNG: numpy ``Generator(PCG64(seed))`` for edit lists and cameras; the value-noise lattice uses
an integer splitmix-style hash so the terrain does not depend on the numpy version.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

# hd_edit_desc (include/hd_capi.h), 56 bytes.
EDIT_DTYPE = np.dtype(
    [
        ("kind", "<u4"),
        ("colored", "<u4"),
        ("p0", "<u4", (3,)),
        ("p1", "<u4", (3,)),
        ("r2", "<u8"),
        ("colors", "<u4"),
        ("weight", "u1"),
        ("bpw", "u1"),
        ("has_value", "u1"),
        ("pad_", "u1"),
        ("skin", "<u4"),
        ("reserved", "<u4"),
    ]
)
assert EDIT_DTYPE.itemsize == 56

SPHERE_FILL, SPHERE_DIG, SPHERE_PAINT, AABB_FILL, AABB_DIG, AABB_PAINT, COLUMN_FILL = range(7)
HD_NULL = 0xFFFFFFFF
HD_COLOR_NULL = 0xC0000000


def edit(kind, p0=(0, 0, 0), p1=(0, 0, 0), r2=0, color=None, colored=None, skin=0):
    """One hd_edit_desc record. color is an RGB8 int (0x00BBGGRR) or None."""
    e = np.zeros((), EDIT_DTYPE)
    e["kind"] = kind
    if colored is None:
        colored = kind not in (SPHERE_DIG, AABB_DIG) and color is not None
    e["colored"] = 1 if colored else 0
    e["p0"] = p0
    e["p1"] = p1
    e["r2"] = r2
    if color is not None:
        e["colors"] = color & 0xFFFFFF
        e["has_value"] = 1
    e["skin"] = skin
    return e


def edits(records) -> np.ndarray:
    return np.array([r for r in records], dtype=EDIT_DTYPE)


@dataclass
class PoolConfig:
    """hashdag::Config (Config.hpp:15-57) + DAGColorPool::Config (DAGColorPool.hpp:36-41)."""

    word_bits_per_page: int
    page_bits_per_bucket: int
    bucket_bits: list
    color_leaf_level: int
    keep_history: bool = False
    color_node_capacity: int = 1 << 22
    color_leaf_capacity: int = 1 << 26

    @property
    def node_levels(self):
        return len(self.bucket_bits)

    @property
    def voxel_level(self):
        return self.node_levels + 1

    @property
    def resolution(self):
        return 1 << self.voxel_level

    @property
    def total_buckets(self):
        return sum(1 << b for b in self.bucket_bits)

    @property
    def words_per_bucket(self):
        return 1 << (self.word_bits_per_page + self.page_bits_per_bucket)

    @property
    def total_words(self):
        return self.total_buckets * self.words_per_bucket

    def level_bases(self):
        out, acc = [], 0
        for b in self.bucket_bits:
            out.append(acc)
            acc += 1 << b
        return out


def config_for_resolution(res: int, bucket_bits: int = 16, wbpp: int = 9, pbpb: int = 2,
                          color_leaf_depth: int = 5, **kw) -> PoolConfig:
    """Pool geometry for a res^3 grid: node levels = log2(res) - 1. The color leaf sits
    color_leaf_depth levels above the voxels (8^depth voxels per VBR chunk)."""
    vl = int(round(math.log2(res)))
    assert 1 << vl == res
    L = vl - 1
    return PoolConfig(wbpp, pbpb, [bucket_bits] * L, max(0, min(L - 1, vl - color_leaf_depth)), **kw)


# ------------------------------------------------------------------ terrain
def _lattice(ix, iy, octave, seed):
    """Value-noise lattice values in [0,1) from an integer hash (splitmix64 finaliser)."""
    k = (ix.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)) ^ (
        iy.astype(np.uint64) * np.uint64(0xC2B2AE3D27D4EB4F)) ^ np.uint64(
        (seed * 0x165667B19E3779F9 + octave * 0x27D4EB2F165667C5) & 0xFFFFFFFFFFFFFFFF)
    k ^= k >> np.uint64(30)
    k *= np.uint64(0xBF58476D1CE4E5B9)
    k ^= k >> np.uint64(27)
    k *= np.uint64(0x94D049BB133111EB)
    k ^= k >> np.uint64(31)
    return (k >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def value_noise_heightmap(res: int, octaves: int, seed: int, zlo: int, zhi: int,
                          base_cells: int = 4, lacunarity: float = 2.0, gain: float = 0.5) -> np.ndarray:
    """res x res heights (uint32, row-major [y, x], x fastest) in [zlo, zhi]: fractal value noise
    (smoothstep-interpolated hashed lattices, base frequency base_cells cells across the grid,
    each octave lacunarity x the frequency and gain x the amplitude), normalised to [0, 1]."""
    xs = (np.arange(res, dtype=np.float64) + 0.5) / res
    acc = np.zeros((res, res), np.float64)
    amp, freq, norm = 1.0, float(base_cells), 0.0
    for o in range(octaves):
        fx = xs * freq
        i0 = np.floor(fx).astype(np.int64)
        t = fx - i0
        t = t * t * (3 - 2 * t)
        n = int(i0.max()) + 2
        idx = np.arange(n, dtype=np.int64)
        lat = _lattice(idx[None, :], idx[:, None], o, seed)  # [iy, ix]
        top = lat[np.ix_(i0, i0)] * (1 - t)[None, :] + lat[np.ix_(i0, i0 + 1)] * t[None, :]
        bot = lat[np.ix_(i0 + 1, i0)] * (1 - t)[None, :] + lat[np.ix_(i0 + 1, i0 + 1)] * t[None, :]
        acc += amp * (top * (1 - t)[:, None] + bot * t[:, None])
        norm += amp
        amp *= gain
        freq *= lacunarity
    acc /= norm
    lo, hi = acc.min(), acc.max()
    acc = (acc - lo) / max(hi - lo, 1e-12)
    h = np.rint(zlo + acc * (zhi - zlo)).astype(np.int64)
    return np.clip(h, 0, res).astype(np.uint32)


def terrain_edit(res: int, seed: int, deep_color: int = 0x505A64, skin: int = 8, lo=(0, 0), hi=None,
                 colored=True):
    """The column-fill edit that builds the terrain (heights come from the pool's heightmap) over the
    columns [lo, hi) (default: the whole grid)."""
    hi = (res, res) if hi is None else hi
    return edit(COLUMN_FILL, (lo[0], lo[1], 0), (hi[0], hi[1], 0), r2=seed, color=deep_color if colored else None,
                colored=colored, skin=skin)


def terrain_tiles(res: int, seed: int, tiles: int, deep_color: int = 0x505A64, skin: int = 8, colored=True):
    """The terrain as tiles x tiles column fills over disjoint column rectangles, applied in order
    (the same voxels and colors as the whole-grid fill; each pass touches 1/tiles^2 of the surface)."""
    t = res // tiles
    return edits([terrain_edit(res, seed, deep_color, skin, (i * t, j * t), ((i + 1) * t, (j + 1) * t), colored)
                  for j in range(tiles) for i in range(tiles)])


# ------------------------------------------------------------------ edit lists
def random_color(rng) -> int:
    return int(rng.integers(0, 1 << 24))


def sphere_digs(n, res, rmin, rmax, seed):
    """n SphereEditor<kDig> edits, centres U[0,res)^3, r ~ U[rmin,rmax], r2 = floor(r^2)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        c = rng.integers(0, res, size=3)
        r = rng.uniform(rmin, rmax)
        out.append(edit(SPHERE_DIG, tuple(int(v) for v in c), r2=int(r * r)))
    return edits(out)


def log_uniform(rng, lo, hi):
    return float(math.exp(rng.uniform(math.log(lo), math.log(hi))))


def terrain_sphere_edits(n, res, heights, rmin, rmax, seed, colored=True):
    """n sphere add/remove edits (50/50) with log-uniform radii, centred on the terrain surface."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x, y = (int(v) for v in rng.integers(0, res, size=2))
        z = int(heights[y, x]) if heights is not None else int(rng.integers(0, res))
        r = log_uniform(rng, rmin, rmax)
        if rng.uniform() < 0.5:
            out.append(edit(SPHERE_FILL, (x, y, z), r2=int(r * r), color=random_color(rng), colored=colored))
        else:
            out.append(edit(SPHERE_DIG, (x, y, z), r2=int(r * r)))
    return edits(out)


def mixed_batch(n, res, heights, rmin, rmax, seed, colored=True):
    """40% sphere add, 40% sphere remove, 20% box paint; radii / half-extents log-uniform."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x, y = (int(v) for v in rng.integers(0, res, size=2))
        z = int(heights[y, x]) if heights is not None else int(rng.integers(0, res))
        r = log_uniform(rng, rmin, rmax)
        u = rng.uniform()
        if u < 0.4:
            out.append(edit(SPHERE_FILL, (x, y, z), r2=int(r * r), color=random_color(rng), colored=colored))
        elif u < 0.8:
            out.append(edit(SPHERE_DIG, (x, y, z), r2=int(r * r)))
        else:
            h = int(max(1, round(r)))
            lo = tuple(max(0, v - h) for v in (x, y, z))
            hi = tuple(min(res, v + h) for v in (x, y, z))
            out.append(edit(AABB_PAINT, lo, hi, color=random_color(rng), colored=colored))
    return edits(out)


def edited_voxels(e: np.ndarray, res: int, heights=None) -> int:
    """Lattice points inside each edit's shape, clipped to [0,res)^3, summed over the list."""
    total = 0
    for r in np.atleast_1d(e):
        k = int(r["kind"])
        if k in (SPHERE_FILL, SPHERE_DIG, SPHERE_PAINT):
            c = [int(v) for v in r["p0"]]
            r2 = int(r["r2"])
            rad = math.isqrt(r2)
            zs = np.arange(max(0, c[2] - rad), min(res, c[2] + rad + 1), dtype=np.int64)
            ys = np.arange(max(0, c[1] - rad), min(res, c[1] + rad + 1), dtype=np.int64)
            if zs.size == 0 or ys.size == 0:
                continue
            dz2 = (zs - c[2]) ** 2
            dy2 = (ys - c[1]) ** 2
            rem = r2 - dz2[:, None] - dy2[None, :]
            ok = rem >= 0
            remc = np.where(ok, rem, 0)
            hx = np.floor(np.sqrt(remc.astype(np.float64))).astype(np.int64)
            # fix floating sqrt to exact isqrt
            hx = np.where((hx + 1) ** 2 <= remc, hx + 1, hx)
            hx = np.where(hx ** 2 > remc, hx - 1, hx)
            lo = np.maximum(c[0] - hx, 0)
            hi = np.minimum(c[0] + hx, res - 1)
            total += int(np.sum(np.where(ok & (hi >= lo), hi - lo + 1, 0)))
        elif k in (AABB_FILL, AABB_DIG, AABB_PAINT):
            lo = [min(int(v), res) for v in r["p0"]]
            hi = [min(int(v), res) for v in r["p1"]]
            total += int(np.prod([max(0, b - a) for a, b in zip(lo, hi)]))
        else:
            total += int(heights.astype(np.int64).sum()) if heights is not None else 0
    return total


# ------------------------------------------------------------------ cameras
def look_side_up(look, fov=math.pi / 3, aspect=1.0, world_up=(0.0, 0.0, 1.0)):
    """Camera::GetLookSideUp (Camera.hpp:36-52) for an explicit look direction: side is
    horizontal, scaled by tan(fov/2)*aspect; up = normalize(cross(look, side)) * tan(fov/2)."""
    look = np.asarray(look, np.float64)
    look = look / np.linalg.norm(look)
    side = np.cross(np.asarray(world_up, np.float64), look)
    if np.linalg.norm(side) < 1e-9:
        side = np.array([1.0, 0.0, 0.0])
    side /= np.linalg.norm(side)
    tg = math.tan(fov * 0.5)
    side = side * tg * aspect
    up = np.cross(look, side)
    up = up / np.linalg.norm(up) * tg
    return look.astype(np.float32), side.astype(np.float32), up.astype(np.float32)


def orbit_view(center, radius, seed, width, height, fov=math.pi / 3, res=None, lod=True,
               elevation=(0.15, 0.6)):
    """A seeded orbit camera looking at center (voxel units). Returns the 13-float view block
    (pos, look, side, up, proj_factor) in [0,1]^3 world units. proj_factor follows
    TracePass.cpp:115-118 (H / (2 tan(fov/2))); lod=False gives +inf (no LOD cut-off)."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0, 2 * math.pi)
    phi = rng.uniform(*elevation)
    d = np.array([math.cos(theta) * math.cos(phi), math.sin(theta) * math.cos(phi), math.sin(phi)])
    center = np.asarray(center, np.float64)
    pos = center + radius * d
    look, side, up = look_side_up(center - pos, fov, width / height)
    scale = 1.0 / (res if res else 1.0)
    proj = height / (2.0 * math.tan(fov * 0.5)) if lod else float("inf")
    return np.concatenate([(pos * scale).astype(np.float32), look, side, up,
                           np.array([proj], np.float32)]).astype(np.float32)


def view_block(pos, look, side, up, proj_factor) -> np.ndarray:
    return np.concatenate([np.asarray(pos, np.float32), look, side, up,
                           np.array([proj_factor], np.float32)]).astype(np.float32)


# ------------------------------------------------------------------ serving traces
# hd_view / hd_request / hd_request_result (include/hd_capi.h)
VIEW_DTYPE = np.dtype([("pos", "<f4", (3,)), ("look", "<f4", (3,)), ("side", "<f4", (3,)), ("up", "<f4", (3,)),
                       ("proj_factor", "<f4"), ("width", "<u4"), ("height", "<u4"), ("flags", "<u4"),
                       ("out_offset", "<u8")])
assert VIEW_DTYPE.itemsize == 72
REQUEST_DTYPE = np.dtype([("kind", "<u4"), ("client", "<u4"), ("arrival_ms", "<f8"), ("edit", EDIT_DTYPE),
                          ("view", VIEW_DTYPE)])
assert REQUEST_DTYPE.itemsize == 144
RESULT_DTYPE = np.dtype([("admit_ms", "<f8"), ("done_ms", "<f8"), ("pixel_offset", "<u8"), ("tick", "<u4"),
                         ("root", "<u4"), ("color_root", "<u4"), ("pad_", "<u4")])
assert RESULT_DTYPE.itemsize == 40
REQ_EDIT, REQ_RENDER = 0, 1


def view_record(vb, width, height, flags) -> np.ndarray:
    v = np.zeros((), VIEW_DTYPE)
    vb = np.asarray(vb, np.float32)
    v["pos"], v["look"], v["side"], v["up"], v["proj_factor"] = vb[0:3], vb[3:6], vb[6:9], vb[9:12], vb[12]
    v["width"], v["height"], v["flags"] = width, height, flags
    return v


def serving_trace(n, res, heights, clients, rate_per_s, seed, rmin=2, rmax=64, edit_frac=0.7,
                  px_min=(256, 256), px_max=(1920, 1080), shadow=True):
    """Poisson arrivals at an aggregate rate from `clients` clients: edit_frac edits (40% sphere
    add, 40% sphere remove, 20% box paint; radii / half-extents log-uniform in [rmin, rmax], centred
    on the terrain surface) and renders at uniform-random resolutions in [px_min, px_max] from
    random orbit cameras (1 shadow ray per hit). Sorted by arrival (ms)."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1000.0 / rate_per_s, size=n)
    t = np.cumsum(gaps)
    out = np.zeros(n, REQUEST_DTYPE)
    # the orbit centre's height: the median of a strided sample (the whole field can be 2^30 heights)
    zc = float(np.median(heights[::max(1, res // 1024), ::max(1, res // 1024)])) if heights is not None else res / 2
    for i in range(n):
        out[i]["arrival_ms"] = t[i]
        out[i]["client"] = int(rng.integers(0, clients))
        if rng.uniform() < edit_frac:
            out[i]["kind"] = REQ_EDIT
            out[i]["edit"] = mixed_batch(1, res, heights, rmin, rmax, int(rng.integers(1 << 31)))[0]
        else:
            out[i]["kind"] = REQ_RENDER
            w = int(rng.integers(px_min[0], px_max[0] + 1))
            h = int(rng.integers(px_min[1], px_max[1] + 1))
            vb = orbit_view((res / 2, res / 2, zc), res * 0.8, int(rng.integers(1 << 31)), w, h, res=res)
            out[i]["view"] = view_record(vb, w, h, 1 if shadow else 0)
    return out
