// hd_device.hpp -- device-side building blocks shared by the edit, GC and traversal kernels:
// pool geometry, the node hash, the closed editor set and the VBR color rules.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/hd_capi.h"

namespace hd {

using u32 = uint32_t;
using u64 = uint64_t;
using i64 = int64_t;

constexpr u32 kNull = 0xFFFFFFFFu;
constexpr u32 kColorNull = 0xC0000000u;
constexpr u32 kTagNode = 0, kTagColor = 1, kTagLeaf = 2, kTagNull = 3;
// Marker for "a color node this edit appends" before its index is known (tag kNode, data 2^30-1;
// color-node capacities are kept below 2^30 - 1 so no real node has this index).
constexpr u32 kColorNewNode = 0x3FFFFFFFu;

// hashdag::Config<uint32_t> (Config.hpp:15-57) flattened for kernels.
struct DevConfig {
	u32 wbpp, pbpb, L, color_leaf, voxel_level, keep_history;
	u32 bucket_bits[HD_MAX_LEVELS];
	u32 level_base[HD_MAX_LEVELS]; // first bucket index of each level (Config.hpp:33-38)
	__host__ __device__ u32 words_per_page() const { return 1u << wbpp; }
	__host__ __device__ u32 wbpb() const { return wbpp + pbpb; }
	__host__ __device__ u32 words_per_bucket() const { return 1u << (wbpp + pbpb); }
};

// Read-only heightfield pyramid for the column-fill editor: level k holds min / max over
// aligned 2^k x 2^k footprints, (res >> k)^2 entries at off[k].
struct DevHeightmap {
	const u32 *mn = nullptr, *mx = nullptr;
	u64 off[HD_MAX_LEVELS + 2] = {};
	u32 res = 0;
	__device__ u32 fmin(u32 k, u32 x, u32 y) const { return mn[off[k] + (u64)y * (res >> k) + x]; }
	__device__ u32 fmax(u32 k, u32 x, u32 y) const { return mx[off[k] + (u64)y * (res >> k) + x]; }
	__device__ u32 h(u32 x, u32 y) const { return mn[(u64)y * res + x]; }
};

// ---------------------------------------------------------------- hash (Hasher.hpp:21-49)
__host__ __device__ inline u32 murmur_span(const u32 *w, u32 n) {
	u32 h = 0;
	for (u32 i = 0; i < n; ++i) {
		u32 k = w[i] * 0xcc9e2d51u;
		k = (k << 15) | (k >> 17);
		k *= 0x1b873593u;
		h ^= k;
		h = (h << 13) | (h >> 19);
		h = h * 5u + 0xe6546b64u;
	}
	h ^= n;
	h ^= h >> 16;
	h *= 0x85ebca6bu;
	h ^= h >> 13;
	h *= 0xc2b2ae35u;
	h ^= h >> 16;
	return h;
}
// murmur_span over a node held in a 9-word register array (static indices only, so the array
// never goes to scratch memory); n <= 9.
__device__ inline u32 murmur_regs(const u32 (&w)[9], u32 n) {
	u32 h = 0;
#pragma unroll
	for (u32 i = 0; i < 9; ++i) {
		if (i < n) {
			u32 k = w[i] * 0xcc9e2d51u;
			k = (k << 15) | (k >> 17);
			k *= 0x1b873593u;
			h ^= k;
			h = (h << 13) | (h >> 19);
			h = h * 5u + 0xe6546b64u;
		}
	}
	h ^= n;
	h ^= h >> 16;
	h *= 0x85ebca6bu;
	h ^= h >> 13;
	h *= 0xc2b2ae35u;
	h ^= h >> 16;
	return h;
}
// The packed inner node of child mask nm over the 8 slot pointers nk (NodePool.hpp:214-226 word
// layout: mask word, then the set children in slot order), compacted with static register
// indices: word j (1..8) is the child of the slot whose rank among the set slots is j.
__device__ inline void pack_inner(u32 nm, const u32 (&nk)[8], u32 (&c)[9]) {
	c[0] = nm;
#pragma unroll
	for (u32 j = 1; j < 9; ++j)
		c[j] = 0u;
#pragma unroll
	for (u32 i = 0; i < 8; ++i) {
		const bool set = (nm >> i) & 1u;
		const u32 pos = (u32)__builtin_popcount(nm & ((1u << i) - 1u)) + 1u;
#pragma unroll
		for (u32 j = 1; j <= i + 1; ++j)
			c[j] = (set && pos == j) ? nk[i] : c[j];
	}
}
__host__ __device__ inline u32 murmur_leaf(u32 w0, u32 w1) {
	u64 h = (u64)w0 | ((u64)w1 << 32);
	h ^= h >> 33;
	h *= 0xff51afd7ed558ccdull;
	h ^= h >> 33;
	h *= 0xc4ceb9fe1a85ec53ull;
	h ^= h >> 33;
	return (u32)h;
}
// Packed inner node size from its mask word; 0 marks page padding (NodePool.hpp:214-226).
__host__ __device__ inline u32 inner_words(u32 w0) {
	u32 m = w0 & 0xFFu;
	return m ? (u32)__builtin_popcount(m) + 1u : 0u;
}

// ---------------------------------------------------------------- colors
struct Color { // hashdag::VBRColor (VBRColor.hpp:19-44)
	u32 colors;
	u32 wb; // weight | bpw << 8 | has << 16
	__host__ __device__ bool has() const { return (wb >> 16) & 1u; }
	__host__ __device__ u32 bpw() const { return (wb >> 8) & 0xFFu; }
	__host__ __device__ u32 weight() const { return wb & 0xFFu; }
	__host__ __device__ bool operator==(const Color &o) const { return colors == o.colors && wb == o.wb; }
	__host__ __device__ static Color none() { return {0u, 0u}; }
	__host__ __device__ static Color rgb8(u32 c) { return {c & 0xFFFFFFu, 1u << 16}; }
};
__host__ __device__ inline Color from_hd(const hd_color &c) {
	if (!c.has_value)
		return Color::none();
	return {c.colors, (u32)c.weight | ((u32)c.bits_per_weight << 8) | (1u << 16)};
}
// VBRColor::Get then RGB8Color(RGBColor): what DAGColorPool::FillNode stores (DAGColorPool.hpp:164-166).
__host__ __device__ inline u32 color_to_rgb8(const Color &c) {
	float r, g, b;
	if (c.bpw() == 0) {
		r = float(c.colors & 0xFFu) / 255.0f;
		g = float((c.colors >> 8) & 0xFFu) / 255.0f;
		b = float((c.colors >> 16) & 0xFFu) / 255.0f;
	} else {
		u32 lo = c.colors & 0xFFFFu, hi = c.colors >> 16;
		float lr = float(lo & 0x1Fu) / 31.0f, lg = float((lo >> 5) & 0x3Fu) / 63.0f, lb = float((lo >> 11) & 0x1Fu) / 31.0f;
		float hr = float(hi & 0x1Fu) / 31.0f, hg = float((hi >> 5) & 0x3Fu) / 63.0f, hb = float((hi >> 11) & 0x1Fu) / 31.0f;
		float a = float(c.weight()) / float((1u << c.bpw()) - 1u);
		r = lr * (1.0f - a) + hr * a;
		g = lg * (1.0f - a) + hg * a;
		b = lb * (1.0f - a) + hb * a;
	}
	u32 ur = (u32)(uint8_t)(r * 255.0f), ug = (u32)(uint8_t)(g * 255.0f), ub = (u32)(uint8_t)(b * 255.0f);
	return ur | (ug << 8) | (ub << 16);
}
__host__ __device__ inline u32 ctag(u32 p) { return p >> 30; }
__host__ __device__ inline u32 cdata(u32 p) { return p & 0x3FFFFFFFu; }
__host__ __device__ inline u32 cmake(u32 t, u32 d) { return (t << 30) | d; }
// DAGColorPool::GetFill (DAGColorPool.hpp:144-146)
__host__ __device__ inline Color get_fill(u32 p) { return ctag(p) == kTagColor ? Color::rgb8(cdata(p)) : Color::none(); }

// Column-fill block color: seeded height ramp + uniform [-8, 8] noise per 4^3 block.
__host__ __device__ inline u32 block_hash(u64 seed, u32 bx, u32 by, u32 bz) {
	u64 k = seed ^ ((u64)bx * 0x9E3779B97F4A7C15ull) ^ ((u64)by * 0xC2B2AE3D27D4EB4Full) ^
	        ((u64)bz * 0x165667B19E3779F9ull);
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdull;
	k ^= k >> 33;
	k *= 0xc4ceb9fe1a85ec53ull;
	k ^= k >> 33;
	return (u32)k;
}
__host__ __device__ inline u32 terrain_block_rgb8(u64 seed, u32 bx, u32 by, u32 bz, u32 res) {
	u32 t = (u32)(((u64)(bz * 4u) * 255u) / (res - 1u));
	t = t > 255u ? 255u : t;
	int base[3] = {50 + (int)(t * 170u / 255u), 110 + (int)(t * 120u / 255u), 40 + (int)(t * 200u / 255u)};
	u32 h = block_hash(seed, bx, by, bz), out = 0;
	for (int c = 0; c < 3; ++c) {
		int v = base[c] + (int)(((h >> (8 * c)) & 0xFFu) % 17u) - 8;
		v = v < 0 ? 0 : (v > 255 ? 255 : v);
		out |= (u32)v << (8 * c);
	}
	return out;
}

// ---------------------------------------------------------------- editors
enum EditType : u32 { kNotAffected = 0, kProceed = 1, kFill = 2, kClear = 3 }; // Editor.hpp:18

// The closed editor set; integer predicates of src/main.cpp:32-150 (+ box dig / paint and the
// heightfield column fill). Identical semantics to the CPU oracle's Editor.
struct DevEditor {
	hd_edit_desc d;
	const DevHeightmap *hm; // device copy of the pool's heightmap view
	u32 voxel_level;

	// the kinds in enum order: sphere fill, dig, paint, box fill, dig, paint, column fill
	static_assert(HD_EDIT_SPHERE_FILL == 0 && HD_EDIT_SPHERE_DIG == 1 && HD_EDIT_SPHERE_PAINT == 2 &&
	                  HD_EDIT_AABB_FILL == 3 && HD_EDIT_AABB_DIG == 4 && HD_EDIT_AABB_PAINT == 5 &&
	                  HD_EDIT_COLUMN_FILL == 6,
	              "kind ranges");
	__device__ bool is_dig() const { return d.kind == HD_EDIT_SPHERE_DIG || d.kind == HD_EDIT_AABB_DIG; }
	__device__ bool is_paint() const { return d.kind == HD_EDIT_SPHERE_PAINT || d.kind == HD_EDIT_AABB_PAINT; }
	__device__ bool is_box() const { return d.kind - (u32)HD_EDIT_AABB_FILL <= 2u; }
	__device__ bool is_sphere() const { return d.kind <= (u32)HD_EDIT_SPHERE_PAINT; }
	__device__ bool stateful() const { return (d.colored || is_paint()) && !is_dig(); }
	__device__ Color color() const { return from_hd(d.color); }

	// AABBEditor::EditNode (main.cpp:35-46), u32 bounds as glm::u32vec3.
	__device__ EditType aabb_node(u32 level, u32 x, u32 y, u32 z) const {
		u32 bits = voxel_level - level;
		u32 p[3] = {x, y, z};
		bool in = true;
		for (int i = 0; i < 3; ++i) {
			u32 lb = p[i] << bits, ub = (p[i] + 1u) << bits;
			if (ub <= d.p0[i] || lb >= d.p1[i])
				return kNotAffected;
			in = in && lb >= d.p0[i] && ub <= d.p1[i];
		}
		return in ? kFill : kProceed;
	}
	__device__ bool aabb_voxel(u32 x, u32 y, u32 z) const {
		return x >= d.p0[0] && x < d.p1[0] && y >= d.p0[1] && y < d.p1[1] && z >= d.p0[2] && z < d.p1[2];
	}
	// SphereEditor::EditNode geometric part (main.cpp:77-106).
	__device__ EditType sphere_node(u32 level, u32 x, u32 y, u32 z) const {
		u32 bits = voxel_level - level;
		u32 p[3] = {x, y, z};
		u64 max_n2 = 0, min_n2 = 0;
		for (int i = 0; i < 3; ++i) {
			u32 lb = p[i] << bits, ub = (p[i] + 1u) << bits;
			i64 lbd = (i64)lb - (i64)d.p0[i], ubd = (i64)ub - (i64)d.p0[i];
			u64 lb2 = (u64)(lbd * lbd), ub2 = (u64)(ubd * ubd);
			max_n2 += lb2 > ub2 ? lb2 : ub2;
			if (lbd > 0)
				min_n2 += lb2;
			if (ubd < 0)
				min_n2 += ub2;
		}
		if (max_n2 <= d.r2)
			return kFill;
		return min_n2 > d.r2 ? kNotAffected : kProceed;
	}
	__device__ bool sphere_voxel(u32 x, u32 y, u32 z) const { // main.cpp:127-132
		i64 dx = (i64)x - (i64)d.p0[0], dy = (i64)y - (i64)d.p0[1], dz = (i64)z - (i64)d.p0[2];
		return (u64)(dx * dx + dy * dy + dz * dz) <= d.r2;
	}
	// The columns [p0.x, p1.x) x [p0.y, p1.y): 0 = outside, 1 = the node's footprint crosses the
	// rectangle's edge, 2 = inside (a terrain is built whole, or tile by tile)
	__device__ u32 column_rect(u32 k, u32 x, u32 y) const {
		const u64 x0 = (u64)x << k, x1 = (u64)(x + 1u) << k, y0 = (u64)y << k, y1 = (u64)(y + 1u) << k;
		if (x1 <= d.p0[0] || x0 >= d.p1[0] || y1 <= d.p0[1] || y0 >= d.p1[1])
			return 0u;
		return x0 >= d.p0[0] && x1 <= d.p1[0] && y0 >= d.p0[1] && y1 <= d.p1[1] ? 2u : 1u;
	}
	__device__ EditType column_node(u32 level, u32 x, u32 y, u32 z) const {
		u32 k = voxel_level - level;
		const u32 rect = column_rect(k, x, y);
		if (!rect)
			return kNotAffected;
		u32 hmin = hm->fmin(k, x, y), hmax = hm->fmax(k, x, y);
		u64 z0 = (u64)z << k, z1 = (u64)(z + 1u) << k;
		if (z0 >= hmax)
			return kNotAffected;
		if (z1 <= hmin)
			return rect == 2u ? kFill : kProceed;
		return kProceed;
	}
	__device__ bool column_voxel(u32 x, u32 y, u32 z) const {
		return x >= d.p0[0] && x < d.p1[0] && y >= d.p0[1] && y < d.p1[1] && z < hm->h(x, y);
	}
	__device__ u32 column_rgb8(u32 x, u32 y, u32 z) const {
		u32 bx = x >> 2, by = y >> 2, bz = z >> 2;
		if ((u64)bz * 4u + 4u + d.skin <= hm->fmin(2, bx, by))
			return d.color.colors & 0xFFFFFFu;
		return terrain_block_rgb8(d.r2, bx, by, bz, hm->res);
	}
	__device__ bool column_uniform(u32 level, u32 x, u32 y, u32 z, u32 &rgb) const {
		u32 k = voxel_level - level;
		if (column_rect(k, x, y) != 2u)
			return false;
		if (k <= 2) {
			rgb = column_rgb8(x << k, y << k, z << k);
			return true;
		}
		u64 z1 = (u64)(z + 1u) << k;
		if (z1 + d.skin <= hm->fmin(k, x, y)) {
			rgb = d.color.colors & 0xFFFFFFu;
			return true;
		}
		return false;
	}

	__device__ EditType geom(u32 level, u32 x, u32 y, u32 z) const {
		if (is_sphere())
			return sphere_node(level, x, y, z);
		if (is_box())
			return aabb_node(level, x, y, z);
		return column_node(level, x, y, z);
	}
	__device__ bool in_range(u32 x, u32 y, u32 z) const {
		if (is_sphere())
			return sphere_voxel(x, y, z);
		if (is_box())
			return aabb_voxel(x, y, z);
		return column_voxel(x, y, z);
	}
	// StatelessEditorWrapper path (Editor.hpp:42-55)
	__device__ EditType edit_node(u32 level, u32 x, u32 y, u32 z) const {
		EditType t = geom(level, x, y, z);
		if (is_dig())
			return t == kFill ? kClear : t;
		return t;
	}
	__device__ bool edit_voxel(u32 x, u32 y, u32 z, bool voxel) const {
		if (is_paint())
			return voxel;
		bool in = in_range(x, y, z);
		return is_dig() ? (voxel && !in) : (voxel || in);
	}
	// VBREditor path: EditNode with the final color (main.cpp:47-56, 107-126).
	__device__ EditType edit_node_vbr(u32 level, u32 x, u32 y, u32 z, bool has_ptr, Color &final_color) const {
		if (d.kind == HD_EDIT_COLUMN_FILL) {
			EditType t = column_node(level, x, y, z);
			u32 rgb;
			bool uni = column_uniform(level, x, y, z, rgb);
			Color u = Color::rgb8(rgb);
			if (t == kFill && !uni)
				t = kProceed;
			if (uni && (t == kFill || !has_ptr || final_color == u))
				final_color = u;
			else if (!uni && !has_ptr && t == kNotAffected) // untouched empty space: any color
				final_color = Color::rgb8(d.color.colors);
			else
				final_color = Color::none();
			return t;
		}
		Color c = color();
		EditType t = is_box() ? aabb_node(level, x, y, z) : sphere_node(level, x, y, z);
		if (t == kFill) {
			final_color = c;
			if (is_paint())
				t = kNotAffected;
		} else if (!has_ptr || final_color == c) {
			final_color = c;
		} else
			final_color = Color::none();
		if (is_paint() && !has_ptr)
			t = kNotAffected;
		return t;
	}
	// edit_voxel_vbr without the voxel's current color: asg is set when the edit assigns c (the color
	// stays the current one otherwise).
	__device__ bool edit_voxel_vbr_asg(u32 x, u32 y, u32 z, bool voxel, Color &c, bool &asg) const {
		bool in = in_range(x, y, z);
		if (d.kind == HD_EDIT_COLUMN_FILL) {
			if (in || !voxel) {
				c = Color::rgb8(column_rgb8(x, y, z));
				asg = true;
			}
			return voxel || in;
		}
		if (in || !voxel) {
			c = color();
			asg = true;
		}
		return is_paint() ? voxel : (voxel || in);
	}
	__device__ bool edit_voxel_vbr(u32 x, u32 y, u32 z, bool voxel, Color &c) const {
		bool in = in_range(x, y, z);
		if (d.kind == HD_EDIT_COLUMN_FILL) {
			if (in || !voxel)
				c = Color::rgb8(column_rgb8(x, y, z));
			return voxel || in;
		}
		c = (in || !voxel) ? color() : c;
		return is_paint() ? voxel : (voxel || in);
	}
};

// Morton index of a position inside a color leaf (x in bit 0), as Color_Morton32 (trace.frag:272-278).
__host__ __device__ inline u32 morton_part(u32 u) {
	u = (u | (u << 16u)) & 0x030000FFu;
	u = (u | (u << 8u)) & 0x0300F00Fu;
	u = (u | (u << 4u)) & 0x030C30C3u;
	u = (u | (u << 2u)) & 0x09249249u;
	return u;
}
__host__ __device__ inline u32 morton3(u32 x, u32 y, u32 z) { return morton_part(x) | (morton_part(y) << 1) | (morton_part(z) << 2); }
// The inverse of morton_part: bits 0, 3, 6, ... of m, packed (x of a Morton index m; y of m >> 1).
__host__ __device__ inline u32 morton_compact(u32 m) {
	m &= 0x09249249u;
	m = (m | (m >> 2u)) & 0x030C30C3u;
	m = (m | (m >> 4u)) & 0x0300F00Fu;
	m = (m | (m >> 8u)) & 0x030000FFu;
	m = (m | (m >> 16u)) & 0x000003FFu;
	return m;
}


// Loads and stores through a pointer read from memory (a struct field): the compiler cannot tell
// its address space and would emit flat instructions, which also count against the LDS wait
// counter (an LDS wait then waits for outstanding global loads too). These address global memory
// explicitly.
// (The host compilation pass only checks these device bodies; it sees plain accesses.)
template <typename T> __device__ inline T ldg(const T *p, u64 i) {
#if defined(__HIP_DEVICE_COMPILE__)
	return ((const __attribute__((address_space(1))) T *)p)[i];
#else
	return p[i];
#endif
}
template <typename T> __device__ inline void stg(T *p, u64 i, T v) {
#if defined(__HIP_DEVICE_COMPILE__)
	((__attribute__((address_space(1))) T *)p)[i] = v;
#else
	p[i] = v;
#endif
}

// ---------------------------------------------------------------- pool find / append
// find_node (NodePool.hpp:79-132): scan the bucket's used words node by node; inner-node size
// comes from the mask word, a zero word is page-tail padding (skip to the next page); leaf
// buckets hold 2-word leaves only.
// [from, bw) restricts the scan to the words appended since bucket word `from` (a node boundary).
__device__ inline u32 find_in_bucket_from(const u32 *__restrict__ words, const DevConfig &c, u32 bucket, u32 from,
                                          u32 bw, const u32 *node, u32 n, bool leaf) {
	const u64 base = (u64)bucket << c.wbpb();
	if (leaf) {
		for (u32 i = from; i + 2 <= bw; i += 2)
			if (ldg(words, base + i) == node[0] && ldg(words, base + i + 1) == node[1])
				return (u32)(base + i);
		return kNull;
	}
	u32 i = from;
	while (i < bw) {
		u32 w0 = ldg(words, base + i);
		u32 sz = inner_words(w0);
		if (sz == 0) {
			i = ((i >> c.wbpp) + 1u) << c.wbpp;
			continue;
		}
		if (sz == n && w0 == node[0]) {
			bool eq = true;
			for (u32 k = 1; k < n; ++k)
				eq = eq && ldg(words, base + i + k) == node[k];
			if (eq)
				return (u32)(base + i);
		}
		i += sz;
	}
	return kNull;
}
__device__ inline u32 find_in_bucket(const u32 *__restrict__ words, const DevConfig &c, u32 bucket, u32 bw,
                                     const u32 *node, u32 n, bool leaf) {
	return find_in_bucket_from(words, c, bucket, 0u, bw, node, n, leaf);
}
// find_in_bucket with the bucket's first 16 words fetched together with its word count, so a
// bucket of up to 16 words costs one memory round trip. bw_out: the bucket's word count.
__device__ inline u32 find_in_bucket_win(const u32 *__restrict__ words, const u32 *__restrict__ bucket_words,
                                         const DevConfig &c, u32 bucket, const u32 (&node)[9], u32 n, bool leaf,
                                         u32 &bw_out) {
	const u64 base = (u64)bucket << c.wbpb();
	u32 W[16];
	if (c.wbpb() >= 4) {
		const uint4 *q = reinterpret_cast<const uint4 *>(words + base);
#pragma unroll
		for (u32 k = 0; k < 4; ++k) {
			uint4 t = ldg(q, k);
			W[4 * k] = t.x;
			W[4 * k + 1] = t.y;
			W[4 * k + 2] = t.z;
			W[4 * k + 3] = t.w;
		}
	} else {
#pragma unroll
		for (u32 k = 0; k < 16; ++k)
			W[k] = ldg(words, base + k); // a bucket this small sits inside the padded pool
	}
	const u32 bw = ldg(bucket_words, bucket);
	bw_out = bw;
	auto get = [&](u32 i) -> u32 {
		if (i >= 16u)
			return ldg(words, base + i);
		u32 v = 0;
#pragma unroll
		for (u32 k = 0; k < 16; ++k)
			v = k == i ? W[k] : v;
		return v;
	};
	if (leaf) {
		for (u32 i = 0; i + 2 <= bw; i += 2)
			if (get(i) == node[0] && get(i + 1) == node[1])
				return (u32)(base + i);
		return kNull;
	}
	u32 i = 0;
	while (i < bw) {
		u32 w0 = get(i);
		u32 sz = inner_words(w0);
		if (sz == 0) {
			i = ((i >> c.wbpp) + 1u) << c.wbpp;
			continue;
		}
		if (sz == n && w0 == node[0]) {
			bool eq = true;
#pragma unroll
			for (u32 k = 1; k < 9; ++k)
				if (k < n)
					eq = eq && get(i + k) == node[k];
			if (eq)
				return (u32)(base + i);
		}
		i += sz;
	}
	return kNull;
}
// find_node over an inner-node bucket held in LDS (T[0, bw)); node in registers (static indices).
__device__ inline u32 find_in_words(const u32 *T, const DevConfig &c, u64 base, u32 bw, const u32 (&node)[9],
                                    u32 n) {
	u32 i = 0;
	while (i < bw) {
		const u32 w0 = T[i], sz = inner_words(w0);
		if (sz == 0) {
			i = ((i >> c.wbpp) + 1u) << c.wbpp;
			continue;
		}
		if (sz == n && w0 == node[0]) {
			bool eq = true;
#pragma unroll
			for (u32 k = 1; k < 9; ++k)
				if (k < n)
					eq = eq && T[i + k] == node[k];
			if (eq)
				return (u32)(base + i);
		}
		i += sz;
	}
	return kNull;
}
// The same scan over global memory (one word per dependent load).
__device__ inline u32 find_in_global(const u32 *__restrict__ words, const DevConfig &c, u64 base, u32 bw,
                                     const u32 (&node)[9], u32 n) {
	return find_in_words(words + base, c, base, bw, node, n);
}

// Scan of staged bucket words where every word equal to node[0] starts a node: inner-node headers
// are child masks (< 256), the level's child pointers are all >= 256 and page padding is 0 while
// node[0] != 0. The first position whose n words equal the node is then the first equal node in
// bucket order (find_node's answer), and the scan needs no parse chain: 16 words per step.
// Reads up to 15 words past bw (inside the LDS allocation; those positions are masked off).
__device__ inline u32 find_in_words_hv(const u32 *T, u64 base, u32 bw, const u32 (&node)[9], u32 n) {
	for (u32 i = 0; i < bw; i += 16) {
		const uint4 *T4 = reinterpret_cast<const uint4 *>(T + i);
		const uint4 a = T4[0], b = T4[1], c = T4[2], d = T4[3];
		const u32 w[16] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w, c.x, c.y, c.z, c.w, d.x, d.y, d.z, d.w};
		u32 cand = 0;
#pragma unroll
		for (u32 k = 0; k < 16; ++k)
			cand |= (w[k] == node[0] ? 1u : 0u) << k;
		if (bw - i < 16u)
			cand &= (1u << (bw - i)) - 1u;
		while (cand) {
			const u32 at = i + (u32)__builtin_ctz(cand);
			cand &= cand - 1u;
			bool eq = true;
#pragma unroll
			for (u32 k = 1; k < 9; ++k)
				if (k < n)
					eq = eq && T[at + k] == node[k];
			if (eq)
				return (u32)(base + at);
		}
	}
	return kNull;
}

// Status of find_in_bucket_staged when the workgroup's stage had no room left: the caller retries
// after the stage is reset (word indices of a pool stay below it).
constexpr u32 kFindRetry = 0xFFFFFFFEu;

// find_node for an inner node with an LDS staging area shared by a workgroup: the bucket's first
// 64 words come with its word count in one round trip, the rest (if any) in one more round trip
// of independent 64-word batches. A bucket of at most 16 words is scanned in registers; a longer
// one is copied to stage[*top ..] (bump-allocated, cap words) and scanned there -- by header value
// when hv (find_in_words_hv), else node by node. Returns kFindRetry when the stage is full (a
// bucket larger than the whole stage is scanned in global memory). Configurations with buckets
// of fewer than 64 words take find_in_bucket_win.
__device__ inline u32 find_in_bucket_staged(const u32 *__restrict__ words, const u32 *__restrict__ bucket_words,
                                            const DevConfig &c, u32 bucket, const u32 (&node)[9], u32 n, bool hv,
                                            u32 &bw_out, u32 *stage, u32 *top, u32 cap) {
	if (c.wbpb() < 6)
		return find_in_bucket_win(words, bucket_words, c, bucket, node, n, false, bw_out);
	const u64 base = (u64)bucket << c.wbpb();
	const uint4 *q = reinterpret_cast<const uint4 *>(words + base);
	uint4 v[16];
#pragma unroll
	for (u32 k = 0; k < 16; ++k)
		v[k] = ldg(q, k);
	const u32 bw = ldg(bucket_words, bucket);
	bw_out = bw;
	if (bw == 0u)
		return kNull;
	if (bw <= 16u && !hv) { // scan the first 16 words in registers (node by node)
		u32 W[16];
#pragma unroll
		for (u32 k = 0; k < 4; ++k) {
			W[4 * k] = v[k].x;
			W[4 * k + 1] = v[k].y;
			W[4 * k + 2] = v[k].z;
			W[4 * k + 3] = v[k].w;
		}
		auto get = [&](u32 i) -> u32 {
			u32 x = 0;
#pragma unroll
			for (u32 k = 0; k < 16; ++k)
				x = k == i ? W[k] : x;
			return x;
		};
		u32 i = 0;
		while (i < bw) {
			const u32 w0 = get(i), sz = inner_words(w0);
			if (sz == 0) { // page-tail padding: the next node starts the next page
				i = ((i >> c.wbpp) + 1u) << c.wbpp;
				continue;
			}
			if (sz == n && w0 == node[0]) {
				bool eq = true;
#pragma unroll
				for (u32 k = 1; k < 9; ++k)
					if (k < n)
						eq = eq && get(i + k) == node[k];
				if (eq)
					return (u32)(base + i);
			}
			i += sz;
		}
		return kNull;
	}
	const u32 len = (bw + 3u) & ~3u;
	if (len > cap)
		return find_in_global(words, c, base, bw, node, n);
	const u32 off = atomicAdd(top, len);
	if (off + len > cap)
		return kFindRetry;
	uint4 *T4 = reinterpret_cast<uint4 *>(stage + off);
#pragma unroll
	for (u32 k = 0; k < 16; ++k)
		if (4u * k < len)
			T4[k] = v[k];
	for (u32 i = 64; i < len; i += 64) { // whole 64-word batches stay inside the bucket (64 | its size)
		uint4 t[16];
#pragma unroll
		for (u32 k = 0; k < 16; ++k)
			t[k] = ldg(q, (i >> 2) + k);
#pragma unroll
		for (u32 k = 0; k < 16; ++k)
			if (i + 4u * k < len)
				T4[(i >> 2) + k] = t[k];
	}
	return hv ? find_in_words_hv(stage + off, base, bw, node, n) : find_in_words(stage + off, c, base, bw, node, n);
}
// Whether every child pointer stored at level l (pointers into level l + 1) is >= 256, so inner
// node headers (< 256) can be told apart by value (find_in_words_hv).
__device__ inline bool headers_by_value(const DevConfig &c, u32 l) {
	return l + 1u < c.L && ((u64)c.level_base[l + 1u] << c.wbpb()) >= 256u;
}

// append_node (NodePool.hpp:134-157): full -> kNull; a node that would straddle a page zero-fills
// the page tail and starts the next page. Updates bw.
__device__ inline u32 append_to_bucket(u32 *__restrict__ words, const DevConfig &c, u32 bucket, u32 &bw,
                                       const u32 *node, u32 n) {
	if (bw + n > c.words_per_bucket())
		return kNull;
	u32 slot = bw >> c.wbpp, off = bw & (c.words_per_page() - 1u);
	if (off + n > c.words_per_page()) {
		u64 pg = ((u64)bucket << c.wbpb()) + ((u64)slot << c.wbpp);
		for (u32 k = off; k < c.words_per_page(); ++k)
			stg(words, pg + k, 0u);
		++slot;
		off = 0;
	}
	u32 rel = (slot << c.wbpp) | off;
	u64 at = ((u64)bucket << c.wbpb()) + rel;
	for (u32 k = 0; k < n; ++k)
		stg(words, at + k, node[k]);
	bw = rel + n;
	return (u32)at;
}
// append_to_bucket for a node held in registers (static indices).
__device__ inline u32 append_regs(u32 *__restrict__ words, const DevConfig &c, u32 bucket, u32 &bw,
                                  const u32 (&node)[9], u32 n) {
	if (bw + n > c.words_per_bucket())
		return kNull;
	u32 slot = bw >> c.wbpp, off = bw & (c.words_per_page() - 1u);
	if (off + n > c.words_per_page()) {
		u64 pg = ((u64)bucket << c.wbpb()) + ((u64)slot << c.wbpp);
		for (u32 k = off; k < c.words_per_page(); ++k)
			stg(words, pg + k, 0u);
		++slot;
		off = 0;
	}
	u32 rel = (slot << c.wbpp) | off;
	u64 at = ((u64)bucket << c.wbpb()) + rel;
#pragma unroll
	for (u32 k = 0; k < 9; ++k)
		if (k < n)
			stg(words, at + k, node[k]);
	bw = rel + n;
	return (u32)at;
}
// bucket choice (NodePool.hpp:163-164)
__device__ inline u32 bucket_of(const DevConfig &c, u32 level, const u32 *node, u32 n, bool leaf) {
	u32 h = leaf ? murmur_leaf(node[0], node[1]) : murmur_span(node, n);
	return c.level_base[level] + (h & ((1u << c.bucket_bits[level]) - 1u));
}

} // namespace hd
