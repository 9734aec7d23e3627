// hd_oracle.hpp -- CPU restatement of the VkHashDAG edit -> GC -> traversal path.
//
// TEST INFRASTRUCTURE ONLY. Nothing in the product (vkhashdag_amd/, include/) may
// include, link or call this file; only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py use it, and only as the checker.
//
// Every function cites the reference file:line it restates (paths relative to the
// VkHashDAG tree). The restatement keeps the reference's recursion order, hash,
// bucket/page layout and VBR writer semantics, so a single-threaded Edit() here
// produces the same pool words as the reference's NodePoolBase::Edit.
//
// Parity pins (see tests/test_oracle_*.py):
//   * murmur_span / murmur_leaf: compiled reference (oracle/_ref) on random spans;
//     fmix64 closed form on known values.
//   * NodePool::edit / upsert / make_filled: identical pool words, bucket counts and
//     roots to the compiled reference on seeded edit sequences; brute-force voxel
//     sets; the reduced-DAG invariants (idempotent edit => same root).
//   * VBR writer / color pool: identical color node and leaf words to the compiled
//     reference (reference VBREditorWrapper + VBRChunkWriter driven over this
//     file's color-pool restatement is NOT used: the harness carries its own
//     DAGColorPool restatement, see oracle/ref_harness.cpp), plus decoded colors
//     equal to the brute-force "last colored edit" rule.
//   * gc: identical pool words to the compiled reference ThreadedGC on one worker.
//   * traversal: hit voxel equal to a brute-force DDA on the decoded grid; the
//     no-LOD march equals the compiled NodePoolTraversal::Traversal<float>.
#pragma once

#include "../include/hd_capi.h"

#include <algorithm>
#include <array>
#include <bit>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <optional>
#include <span>
#include <stdexcept>
#include <unordered_map>
#include <unordered_set>
#include <vector>

namespace hdo {

using u32 = uint32_t;
using u64 = uint64_t;
using i64 = int64_t;
constexpr u32 kNull = 0xFFFFFFFFu;

// ---------------------------------------------------------------- Config
// include/hashdag/Config.hpp:15-57
struct Config {
	u32 wbpp = 9, pbpb = 2;
	std::vector<u32> bucket_bits;

	u32 words_per_page() const { return 1u << wbpp; }
	u32 words_per_bucket() const { return 1u << (wbpp + pbpb); }
	u32 word_bits_per_bucket() const { return wbpp + pbpb; }
	u32 buckets_at(u32 l) const { return 1u << bucket_bits[l]; }
	u32 node_levels() const { return (u32)bucket_bits.size(); }
	u32 leaf_level() const { return node_levels(); }      // Config.hpp:30
	u32 voxel_level() const { return node_levels() + 1; } // Config.hpp:31
	u32 resolution() const { return 1u << voxel_level(); }
	std::vector<u32> level_bases() const { // Config.hpp:33-38
		std::vector<u32> b(bucket_bits.size(), 0);
		for (size_t i = 1; i < b.size(); ++i)
			b[i] = b[i - 1] + buckets_at((u32)i - 1);
		return b;
	}
	u64 total_buckets() const {
		u64 t = 0;
		for (u32 b : bucket_bits)
			t += 1ull << b;
		return t;
	}
	u64 total_words() const { return total_buckets() << (wbpp + pbpb); }
	u64 total_pages() const { return total_buckets() << pbpb; }
	static bool validate(const Config &c) { // Config.hpp:48-56
		if (c.wbpp < 4)
			return false;
		return c.total_words() - 1 <= u64(0xFFFFFFFEu);
	}
};

// ---------------------------------------------------------------- Hash
// include/hashdag/Hasher.hpp:22-39 (inner nodes: Murmur3 body, h ^= word count, fmix32)
inline u32 murmur_span(const u32 *w, size_t n) {
	u32 h = 0;
	for (size_t i = 0; i < n; ++i) {
		u32 k = w[i] * 0xcc9e2d51u;
		k = (k << 15) | (k >> 17);
		k *= 0x1b873593u;
		h ^= k;
		h = (h << 13) | (h >> 19);
		h = h * 5u + 0xe6546b64u;
	}
	h ^= (u32)n;
	h ^= h >> 16;
	h *= 0x85ebca6bu;
	h ^= h >> 13;
	h *= 0xc2b2ae35u;
	h ^= h >> 16;
	return h;
}
// include/hashdag/Hasher.hpp:40-48 (leaves: fmix64 of the 64-bit mask, truncated)
inline u32 murmur_leaf(u32 w0, u32 w1) {
	u64 h = (u64)w0 | ((u64)w1 << 32);
	h ^= h >> 33;
	h *= 0xff51afd7ed558ccdull;
	h ^= h >> 33;
	h *= 0xc4ceb9fe1a85ec53ull;
	h ^= h >> 33;
	return (u32)h;
}

// Size of a packed inner node from its mask word (NodePool.hpp:214-226): popcount+1, 0 for a zero word.
inline u32 inner_words(u32 w0) {
	u32 m = w0 & 0xFFu;
	return m ? (u32)std::popcount(m) + 1u : 0u;
}

// ---------------------------------------------------------------- Colors
enum class EditType { kNotAffected, kProceed, kFill, kClear }; // Editor.hpp:18

// VBRColor (VBRColor.hpp:19-44): RGB8 when bpw == 0, else a weighted mix of two R5G6B5.
struct VBRColor {
	u32 colors = 0;
	uint8_t weight = 0, bpw = 0;
	bool has = false;
	static VBRColor rgb8(u32 c) { return {c & 0xFFFFFFu, 0, 0, true}; }
	bool operator==(const VBRColor &) const = default;
};
inline VBRColor from_hd(const hd_color &c) {
	VBRColor v;
	if (!c.has_value)
		return v; // a VBRColor without value is default-constructed (all zero)
	v.colors = c.colors;
	v.weight = c.weight;
	v.bpw = c.bits_per_weight;
	v.has = c.has_value != 0;
	return v;
}
// VBRColor::Get (VBRColor.hpp:34-38) followed by RGB8Color(RGBColor) (Color.hpp:24): the value
// DAGColorPool::FillNode stores (DAGColorPool.hpp:164-166). fp32, glm::mix = x*(1-a)+y*a.
inline u32 vbr_to_rgb8(const VBRColor &c) {
	float r, g, b;
	if (c.bpw == 0) {
		r = float(c.colors & 0xFFu) / 255.0f;
		g = float((c.colors >> 8) & 0xFFu) / 255.0f;
		b = float((c.colors >> 16) & 0xFFu) / 255.0f;
	} else {
		u32 lo = c.colors & 0xFFFFu, hi = c.colors >> 16;
		float lr = float(lo & 0x1Fu) / 31.0f, lg = float((lo >> 5) & 0x3Fu) / 63.0f, lb = float((lo >> 11) & 0x1Fu) / 31.0f;
		float hr = float(hi & 0x1Fu) / 31.0f, hg = float((hi >> 5) & 0x3Fu) / 63.0f, hb = float((hi >> 11) & 0x1Fu) / 31.0f;
		float a = float(c.weight) / float((1u << c.bpw) - 1u);
		r = lr * (1.0f - a) + hr * a;
		g = lg * (1.0f - a) + hg * a;
		b = lb * (1.0f - a) + hb * a;
	}
	u32 ur = (u32)(uint8_t)(r * 255.0f), ug = (u32)(uint8_t)(g * 255.0f), ub = (u32)(uint8_t)(b * 255.0f);
	return ur | (ug << 8) | (ub << 16);
}

// ---------------------------------------------------------------- Heightmap (column fill)
// Own editor (no reference counterpart): a full-resolution heightfield with a min/max
// pyramid so a node's footprint bounds are O(1).
struct Heightmap {
	u32 res = 0;
	std::vector<std::vector<u32>> mn, mx; // level k: (res >> k)^2 entries
	void set(const u32 *h, u32 r) {
		res = r;
		mn.clear();
		mx.clear();
		mn.emplace_back(h, h + (size_t)r * r);
		mx.emplace_back(h, h + (size_t)r * r);
		for (u32 s = r >> 1, k = 1; s >= 1; s >>= 1, ++k) {
			std::vector<u32> a((size_t)s * s), b((size_t)s * s);
			u32 ps = s * 2;
			for (u32 y = 0; y < s; ++y)
				for (u32 x = 0; x < s; ++x) {
					const auto &pm = mn[k - 1], &px = mx[k - 1];
					size_t i0 = (size_t)(2 * y) * ps + 2 * x, i1 = i0 + ps;
					a[(size_t)y * s + x] = std::min({pm[i0], pm[i0 + 1], pm[i1], pm[i1 + 1]});
					b[(size_t)y * s + x] = std::max({px[i0], px[i0 + 1], px[i1], px[i1 + 1]});
				}
			mn.push_back(std::move(a));
			mx.push_back(std::move(b));
			if (s == 1)
				break;
		}
	}
	u32 h(u32 x, u32 y) const { return mn[0][(size_t)y * res + x]; }
	// min / max over the aligned footprint of side 2^k whose index is (fx, fy)
	u32 fmin(u32 k, u32 fx, u32 fy) const { return mn[k][(size_t)fy * (res >> k) + fx]; }
	u32 fmax(u32 k, u32 fx, u32 fy) const { return mx[k][(size_t)fy * (res >> k) + fx]; }
};

// Column-fill block color: seeded height ramp + uniform [-8, 8] noise per 4^3 block.
inline u32 block_hash(u64 seed, u32 bx, u32 by, u32 bz) {
	u64 k = seed ^ ((u64)bx * 0x9E3779B97F4A7C15ull) ^ ((u64)by * 0xC2B2AE3D27D4EB4Full) ^
	        ((u64)bz * 0x165667B19E3779F9ull);
	k ^= k >> 33;
	k *= 0xff51afd7ed558ccdull;
	k ^= k >> 33;
	k *= 0xc4ceb9fe1a85ec53ull;
	k ^= k >> 33;
	return (u32)k;
}
inline u32 terrain_block_rgb8(u64 seed, u32 bx, u32 by, u32 bz, u32 res) {
	u32 t = (u32)(((u64)(bz * 4u) * 255u) / (res - 1u));
	if (t > 255u)
		t = 255u;
	int base[3] = {50 + (int)(t * 170u / 255u), 110 + (int)(t * 120u / 255u), 40 + (int)(t * 200u / 255u)};
	u32 h = block_hash(seed, bx, by, bz);
	u32 out = 0;
	for (int c = 0; c < 3; ++c) {
		int v = base[c] + (int)(((h >> (8 * c)) & 0xFFu) % 17u) - 8;
		v = v < 0 ? 0 : (v > 255 ? 255 : v);
		out |= (u32)v << (8 * c);
	}
	return out;
}

// ---------------------------------------------------------------- Editors
// The closed editor set. Integer predicates restate src/main.cpp:32-150.
struct Editor {
	hd_edit_desc d{};
	const Heightmap *hm = nullptr;

	VBRColor color() const { return from_hd(d.color); }

	static void bounds(const Config &cfg, u32 level, const u32 pos[3], u64 lb[3], u64 ub[3]) {
		// NodeCoord::GetLower/UpperBoundAtLevel (NodeCoord.hpp:86-93)
		u32 bits = cfg.voxel_level() - level;
		for (int i = 0; i < 3; ++i) {
			lb[i] = (u64)pos[i] << bits;
			ub[i] = (u64)(pos[i] + 1u) << bits;
		}
	}
	// AABBEditor::EditNode (main.cpp:35-46). Upper bounds are computed in u32 like glm::u32vec3.
	EditType aabb_node(const Config &cfg, u32 level, const u32 pos[3]) const {
		u32 bits = cfg.voxel_level() - level;
		u32 lb[3], ub[3];
		for (int i = 0; i < 3; ++i) {
			lb[i] = pos[i] << bits;
			ub[i] = (pos[i] + 1u) << bits;
		}
		for (int i = 0; i < 3; ++i)
			if (ub[i] <= d.p0[i] || lb[i] >= d.p1[i])
				return EditType::kNotAffected;
		bool in = true;
		for (int i = 0; i < 3; ++i)
			in = in && lb[i] >= d.p0[i] && ub[i] <= d.p1[i];
		return in ? EditType::kFill : EditType::kProceed;
	}
	bool aabb_voxel(const u32 p[3]) const { // main.cpp:57-59
		for (int i = 0; i < 3; ++i)
			if (!(p[i] >= d.p0[i] && p[i] < d.p1[i]))
				return false;
		return true;
	}
	// SphereEditor::EditNode (main.cpp:77-106) geometric part; returns kFill when inside.
	EditType sphere_node(const Config &cfg, u32 level, const u32 pos[3]) const {
		u32 bits = cfg.voxel_level() - level;
		i64 lbd[3], ubd[3];
		u64 lb2[3], ub2[3];
		for (int i = 0; i < 3; ++i) {
			u32 lb = pos[i] << bits, ub = (pos[i] + 1u) << bits; // glm::u32vec3 arithmetic
			lbd[i] = (i64)lb - (i64)d.p0[i];
			ubd[i] = (i64)ub - (i64)d.p0[i];
			lb2[i] = (u64)(lbd[i] * lbd[i]);
			ub2[i] = (u64)(ubd[i] * ubd[i]);
		}
		u64 max_n2 = 0;
		for (int i = 0; i < 3; ++i)
			max_n2 += std::max(lb2[i], ub2[i]);
		if (max_n2 <= d.r2)
			return EditType::kFill;
		u64 min_n2 = 0;
		for (int i = 0; i < 3; ++i) {
			if (lbd[i] > 0)
				min_n2 += lb2[i];
			if (ubd[i] < 0)
				min_n2 += ub2[i];
		}
		return min_n2 > d.r2 ? EditType::kNotAffected : EditType::kProceed;
	}
	bool sphere_voxel(const u32 p[3]) const { // main.cpp:127-132
		i64 dx = (i64)p[0] - (i64)d.p0[0], dy = (i64)p[1] - (i64)d.p0[1], dz = (i64)p[2] - (i64)d.p0[2];
		u64 n2 = (u64)(dx * dx + dy * dy + dz * dz);
		return n2 <= d.r2;
	}
	// Column fill geometry: voxel (x,y,z) is inside iff z < h(x,y).
	// Column rectangle [p0.x, p1.x) x [p0.y, p1.y) of the node's footprint: 0 outside, 1 crossing, 2 inside.
	u32 column_rect(u32 k, u32 x, u32 y) const {
		const u64 x0 = (u64)x << k, x1 = (u64)(x + 1u) << k, y0 = (u64)y << k, y1 = (u64)(y + 1u) << k;
		if (x1 <= d.p0[0] || x0 >= d.p1[0] || y1 <= d.p0[1] || y0 >= d.p1[1])
			return 0u;
		return x0 >= d.p0[0] && x1 <= d.p1[0] && y0 >= d.p0[1] && y1 <= d.p1[1] ? 2u : 1u;
	}
	EditType column_node(const Config &cfg, u32 level, const u32 pos[3]) const {
		u32 k = cfg.voxel_level() - level;
		const u32 rect = column_rect(k, pos[0], pos[1]);
		if (!rect)
			return EditType::kNotAffected;
		u32 hmin = hm->fmin(k, pos[0], pos[1]), hmax = hm->fmax(k, pos[0], pos[1]);
		u64 z0 = (u64)pos[2] << k, z1 = (u64)(pos[2] + 1u) << k;
		if (z0 >= hmax)
			return EditType::kNotAffected;
		if (z1 <= hmin)
			return rect == 2u ? EditType::kFill : EditType::kProceed;
		return EditType::kProceed;
	}
	bool column_voxel(const u32 p[3]) const {
		return p[0] >= d.p0[0] && p[0] < d.p1[0] && p[1] >= d.p0[1] && p[1] < d.p1[1] && p[2] < hm->h(p[0], p[1]);
	}
	// Per-voxel column color: deep color below the skin, else the block's ramp+noise color.
	u32 column_rgb8(const u32 p[3]) const {
		u32 bx = p[0] >> 2, by = p[1] >> 2, bz = p[2] >> 2;
		u32 hmin_b = hm->fmin(2, bx, by);
		if ((u64)bz * 4u + 4u + d.skin <= hmin_b)
			return d.color.colors & 0xFFFFFFu;
		return terrain_block_rgb8(d.r2, bx, by, bz, hm->res);
	}
	// Uniform column color over a node, when the rule gives one.
	bool column_uniform(const Config &cfg, u32 level, const u32 pos[3], u32 &rgb) const {
		u32 k = cfg.voxel_level() - level;
		if (column_rect(k, pos[0], pos[1]) != 2u)
			return false;
		if (k <= 2) {
			u32 p[3] = {pos[0] << k, pos[1] << k, pos[2] << k};
			rgb = column_rgb8(p);
			return true;
		}
		u32 hmin = hm->fmin(k, pos[0], pos[1]);
		u64 z1 = (u64)(pos[2] + 1u) << k;
		if (z1 + d.skin <= hmin) {
			rgb = d.color.colors & 0xFFFFFFu;
			return true;
		}
		return false;
	}

	bool is_dig() const { return d.kind == HD_EDIT_SPHERE_DIG || d.kind == HD_EDIT_AABB_DIG; }
	bool is_paint() const { return d.kind == HD_EDIT_SPHERE_PAINT || d.kind == HD_EDIT_AABB_PAINT; }
	// Paint only has meaning through VBREditorWrapper: the reference's stateless kPaint EditNode
	// returns kFill inside the sphere (main.cpp:88-89), so paint edits always run with colors.
	bool stateful() const { return (d.colored || is_paint()) && !is_dig(); }

	EditType geom(const Config &cfg, u32 level, const u32 pos[3]) const {
		switch (d.kind) {
		case HD_EDIT_SPHERE_FILL:
		case HD_EDIT_SPHERE_DIG:
		case HD_EDIT_SPHERE_PAINT:
			return sphere_node(cfg, level, pos);
		case HD_EDIT_AABB_FILL:
		case HD_EDIT_AABB_DIG:
		case HD_EDIT_AABB_PAINT:
			return aabb_node(cfg, level, pos);
		default:
			return column_node(cfg, level, pos);
		}
	}
	bool in_range(const u32 p[3]) const {
		switch (d.kind) {
		case HD_EDIT_SPHERE_FILL:
		case HD_EDIT_SPHERE_DIG:
		case HD_EDIT_SPHERE_PAINT:
			return sphere_voxel(p);
		case HD_EDIT_AABB_FILL:
		case HD_EDIT_AABB_DIG:
		case HD_EDIT_AABB_PAINT:
			return aabb_voxel(p);
		default:
			return column_voxel(p);
		}
	}

	// Stateless EditNode: SphereEditor/AABBEditor::EditNode(config, coord, ptr) (main.cpp:35-46,77-106).
	EditType edit_node(const Config &cfg, u32 level, const u32 pos[3]) const {
		EditType t = geom(cfg, level, pos);
		if (is_dig())
			return t == EditType::kFill ? EditType::kClear : t; // main.cpp:88-89
		return t;
	}
	// Stateless EditVoxel (main.cpp:60-63, 133-142).
	bool edit_voxel(const u32 p[3], bool voxel) const {
		if (is_paint())
			return voxel;
		bool in = in_range(p);
		return is_dig() ? (voxel && !in) : (voxel || in);
	}
	// VBR EditNode with final color (main.cpp:47-56 for boxes, 107-126 for spheres; paint mirrors
	// the sphere kPaint rule for boxes; column fill: see column_uniform).
	EditType edit_node_vbr(const Config &cfg, u32 level, const u32 pos[3], bool has_ptr, VBRColor &final_color) const {
		VBRColor c = color();
		if (d.kind == HD_EDIT_COLUMN_FILL) {
			EditType t = column_node(cfg, level, pos);
			u32 rgb;
			bool uni = column_uniform(cfg, level, pos, rgb);
			VBRColor u = VBRColor::rgb8(rgb);
			if (t == EditType::kFill && !uni)
				t = EditType::kProceed;
			if (uni && (t == EditType::kFill || !has_ptr || final_color == u))
				final_color = u;
			else if (!uni && !has_ptr && t == EditType::kNotAffected) // untouched empty space: any color
				final_color = VBRColor::rgb8(d.color.colors);
			else
				final_color = {};
			return t;
		}
		bool box = d.kind == HD_EDIT_AABB_FILL || d.kind == HD_EDIT_AABB_PAINT;
		EditType t = box ? aabb_node(cfg, level, pos) : sphere_node(cfg, level, pos);
		if (t == EditType::kFill) {
			final_color = c;
			if (is_paint())
				t = EditType::kNotAffected;
		} else if (!has_ptr || final_color == c) {
			final_color = c;
		} else
			final_color = {};
		if (is_paint() && !has_ptr)
			t = EditType::kNotAffected;
		return t;
	}
	// VBR EditVoxel (main.cpp:64-69, 143-149).
	bool edit_voxel_vbr(const u32 p[3], bool voxel, VBRColor &color_io) const {
		bool in = in_range(p);
		if (d.kind == HD_EDIT_COLUMN_FILL) {
			if (in || !voxel)
				color_io = VBRColor::rgb8(column_rgb8(p));
			return voxel || in;
		}
		color_io = (in || !voxel) ? color() : color_io;
		return is_paint() ? voxel : (voxel || in);
	}
};

// ---------------------------------------------------------------- VBR chunk codec
// include/hashdag/VBRColor.hpp:65-487. A chunk is macro blocks {first_block, weight_start},
// block headers {colors, voxoff:14 | bpw:2 | woff:16} and a weight bitset.
struct Chunk {
	std::vector<std::array<u32, 2>> macros, headers;
	std::vector<u32> weights;
	bool empty() const { return macros.empty(); }
};

// Bit-append writer; its words equal VBRBitsetWriter's (VBRColor.hpp:96-227): bits are
// concatenated LSB-first and the tail of the last word is zero.
struct BitWriter {
	std::vector<u32> words;
	u64 count = 0;
	void push_bits(u32 v, u32 bits) {
		for (u32 i = 0; i < bits; ++i, ++count) {
			if ((count & 31u) == 0)
				words.push_back(0);
			if ((v >> i) & 1u)
				words.back() |= 1u << (count & 31u);
		}
	}
	void push(u32 v, u32 bits) {
		if (bits)
			push_bits(v, bits);
	}
	void push(u32 v, u32 bits, u64 n) {
		if (!bits)
			return;
		for (u64 i = 0; i < n; ++i)
			push_bits(v, bits);
	}
	void copy(const std::vector<u32> &src, u64 begin, u64 nbits) {
		for (u64 i = 0; i < nbits; ++i) {
			u64 b = begin + i;
			u32 v = (b >> 5) < src.size() ? (src[b >> 5] >> (b & 31u)) & 1u : 0u;
			push_bits(v, 1);
		}
	}
};
// VBRBitset::Get (VBRColor.hpp:83-93)
inline u32 bitset_get(const std::vector<u32> &w, u64 index, u32 bits) {
	u32 off = index & 31u;
	u32 w0 = w[index >> 5] >> off;
	if (off + bits <= 32)
		return w0 & ((1u << bits) - 1u);
	u32 w1 = w[(index >> 5) + 1] & ((1u << (bits + off - 32)) - 1u);
	return w0 | (w1 << (32 - off));
}

inline u32 hdr_voxoff(const std::array<u32, 2> &h) { return h[1] >> 18; }
inline u32 hdr_bpw(const std::array<u32, 2> &h) { return (h[1] >> 16) & 3u; }
inline u32 hdr_woff(const std::array<u32, 2> &h) { return h[1] & 0xFFFFu; }
inline std::array<u32, 2> make_hdr(u32 colors, u32 voxoff, u32 bpw, u32 woff) {
	return {colors, (voxoff << 18) | (bpw << 16) | woff};
}

// VBRChunkIterator (VBRColor.hpp:292-366)
struct ChunkIter {
	const Chunk *c = nullptr;
	u32 macro = 0, block = 0, off = 0;
	bool empty() const { return !c || c->empty(); }
	bool last() const { return block == c->headers.size() - 1; }
	u32 block_size() const {
		u32 nv = hdr_voxoff(c->headers[block + 1]);
		nv = nv == 0 ? (1u << 14) : nv;
		return nv - hdr_voxoff(c->headers[block]);
	}
	void next_block() {
		off = 0;
		++block;
		macro += hdr_voxoff(c->headers[block]) == 0;
	}
	u32 weight_index() const {
		const auto &h = c->headers[block];
		return c->macros[macro][1] + hdr_woff(h) + off * hdr_bpw(h);
	}
	VBRColor color() const {
		const auto &h = c->headers[block];
		u32 bpw = hdr_bpw(h);
		if (bpw == 0)
			return VBRColor::rgb8(h[0]);
		VBRColor v;
		v.colors = h[0];
		v.bpw = (uint8_t)bpw;
		v.weight = (uint8_t)bitset_get(c->weights, weight_index(), bpw);
		v.has = true;
		return v;
	}
	VBRColor next() {
		VBRColor r = color();
		++off;
		if (!last() && off == block_size())
			next_block();
		return r;
	}
	template <typename F> void jump(u32 count, F &&on_move) {
		if (count == 0)
			return;
		u32 step;
		while (!last() && count >= (step = block_size() - off)) {
			on_move(*this, step);
			count -= step;
			next_block();
		}
		if (count) {
			on_move(*this, count);
			off += count;
		}
	}
};

// VBRChunkWriter (VBRColor.hpp:368-487)
struct ChunkWriter {
	Chunk out;
	BitWriter bits;
	ChunkIter src;
	Chunk src_chunk;
	u32 voxel_count = 0;

	explicit ChunkWriter(Chunk source) : src_chunk(std::move(source)) { src.c = &src_chunk; }
	ChunkWriter(const ChunkWriter &) = delete;

	void append(u32 colors, u32 bpw, u32 n, u32 &wstart) { // VBRColor.hpp:384-411
		for (u32 i = 0; i < n;) {
			u32 vi = voxel_count + i, m = vi >> 14;
			if (m >= out.macros.size()) {
				out.macros.push_back({(u32)out.headers.size(), wstart});
				out.headers.push_back(make_hdr(colors, 0, bpw, 0));
			}
			if (colors != out.headers.back()[0] || bpw != hdr_bpw(out.headers.back()))
				out.headers.push_back(make_hdr(colors, vi & 0x3FFFu, bpw, wstart - out.macros.back()[1]));
			u32 k = std::min(n - i, ((m + 1u) << 14) - vi);
			wstart += k * bpw;
			i += k;
		}
		voxel_count += n;
	}
	void copy_n(u32 n) { // VBRColor.hpp:435-443
		u32 wstart = (u32)bits.count, prev = wstart, src_w = src.weight_index();
		src.jump(n, [&](const ChunkIter &it, u32 k) {
			const auto &h = it.c->headers[it.block];
			append(h[0], hdr_bpw(h), k, wstart);
		});
		bits.copy(src_chunk.weights, src_w, wstart - prev);
	}
	void push_n(const VBRColor &c, u32 n) { // VBRColor.hpp:444-448
		u32 wstart = (u32)bits.count;
		append(c.colors, c.bpw, n, wstart);
		bits.push(c.weight, c.bpw, n);
	}
	void Copy(u32 n, const VBRColor &empty_color) { // VBRColor.hpp:465-471
		if (src.empty()) {
			push_n(empty_color, n);
			return;
		}
		copy_n(n);
	}
	void Push(const VBRColor &c, u32 n) { // VBRColor.hpp:472-476
		push_n(c, n);
		if (!src.empty())
			src.jump(n, [](const ChunkIter &, u32) {});
	}
	template <typename F> bool Edit(F &&editor, const VBRColor &empty_color) { // VBRColor.hpp:477-486
		VBRColor c = src.empty() ? empty_color : src.next();
		bool r = editor(c);
		push_n(c, 1);
		return r;
	}
	Chunk flush() {
		out.weights = bits.words;
		return std::move(out);
	}
};

// ---------------------------------------------------------------- Color pool
// src/DAGColorPool.hpp:23-204 (tagged 30-bit pointers; nodes of 8 pointers; VBR leaves).
struct ColorPool {
	enum Tag : u32 { kNode = 0, kColor = 1, kLeaf = 2, kNullTag = 3 };
	static constexpr u32 kNullPtr = 0xC0000000u;
	u32 leaf_level = 0;
	bool keep_history = false;
	std::vector<std::array<u32, 8>> nodes;
	std::vector<u32> leaves;
	u64 node_cap = 1ull << 30, leaf_cap = 1ull << 30;

	static u32 tag(u32 p) { return p >> 30; }
	static u32 data(u32 p) { return p & 0x3FFFFFFFu; }
	static u32 make(u32 t, u32 d) { return (t << 30) | d; }

	u32 get_child(u32 p, u32 idx) const { // DAGColorPool.hpp:139-143
		return tag(p) == kNode ? nodes[data(p)][idx] : (tag(p) == kColor ? p : kNullPtr);
	}
	static VBRColor get_fill(u32 p) { // DAGColorPool.hpp:144-146
		return tag(p) == kColor ? VBRColor::rgb8(data(p)) : VBRColor{};
	}
	u32 set_node(u32 p, const std::array<u32, 8> &ch) { // DAGColorPool.hpp:147-162
		if (std::all_of(ch.begin(), ch.end(), [](u32 c) { return tag(c) == kNullTag; }))
			return kNullPtr;
		if (tag(ch[0]) == kColor && std::all_of(ch.begin() + 1, ch.end(), [&](u32 c) { return c == ch[0]; }))
			return ch[0];
		if (tag(p) == kNode && nodes[data(p)] == ch)
			return p;
		if (nodes.size() >= node_cap)
			return p;
		nodes.push_back(ch);
		return make(kNode, (u32)nodes.size() - 1);
	}
	static u32 clear_node(u32) { return kNullPtr; }
	static u32 fill_node(u32, const VBRColor &c) { return make(kColor, vbr_to_rgb8(c)); }
	Chunk get_leaf(u32 p) const { // DAGColorPool.hpp:98-113,168-171
		Chunk c;
		if (tag(p) != kLeaf)
			return c;
		size_t i = data(p) + 1;
		u32 M = leaves[i], B = leaves[i + 1], W = leaves[i + 2];
		i += 3;
		for (u32 k = 0; k < M; ++k, i += 2)
			c.macros.push_back({leaves[i], leaves[i + 1]});
		for (u32 k = 0; k < B; ++k, i += 2)
			c.headers.push_back({leaves[i], leaves[i + 1]});
		c.weights.assign(leaves.begin() + i, leaves.begin() + i + W);
		return c;
	}
	void write_chunk(size_t i, const Chunk &c) { // DAGColorPool.hpp:66-96
		leaves[i++] = (u32)c.macros.size();
		leaves[i++] = (u32)c.headers.size();
		leaves[i++] = (u32)c.weights.size();
		for (auto &m : c.macros)
			leaves[i++] = m[0], leaves[i++] = m[1];
		for (auto &h : c.headers)
			leaves[i++] = h[0], leaves[i++] = h[1];
		for (u32 w : c.weights)
			leaves[i++] = w;
	}
	u32 set_leaf(u32 p, const Chunk &c) { // DAGColorPool.hpp:172-204
		size_t data_size = c.macros.size() * 2 + c.headers.size() * 2 + c.weights.size() + 4;
		size_t append_size = (data_size & 1) ? data_size + 1 : data_size;
		if (!keep_history && tag(p) == kLeaf) {
			size_t idx = data(p), block = leaves[idx];
			if (data_size <= block) {
				write_chunk(idx + 1, c);
				return p;
			}
			append_size = std::max(block << 1, append_size);
		}
		if (leaves.size() + append_size > leaf_cap)
			return p;
		size_t idx = leaves.size();
		leaves.resize(idx + append_size, 0u);
		leaves[idx] = (u32)append_size;
		write_chunk(idx + 1, c);
		return make(kLeaf, (u32)idx);
	}
};

// ---------------------------------------------------------------- Node pool
// include/hashdag/NodePool.hpp:40-418 over an in-memory paged store (src/DAGNodePool.hpp:56-74).
struct NodePool {
	Config cfg;
	std::vector<u32> bases;
	std::vector<u32> bucket_words;
	std::vector<std::unique_ptr<u32[]>> pages;
	std::vector<u32> filled;
	u64 overflows = 0; // upserts that fell back because the bucket was full

	// Algorithmic node-pool traffic of edits, in words (the numerator of the edit path's roofline):
	// words of existing nodes the recursion reads (edit_leaf / edit_node unpack), words hashed and
	// words scanned in the target bucket by each upsert, words appended (incl. page padding) and one
	// bucket-counter word per append. In batch mode an input node is counted once however many
	// edits read it, and the appended nodes are those of the final DAG (see batch_traffic_end).
	struct Traffic {
		u64 visited = 0, hashed = 0, scanned = 0, written = 0, counters = 0, upserts = 0;
		bool batch = false;
		std::vector<u32> in_bw;         // batch mode: bucket counters of the input state
		std::unordered_set<u32> seen;   // batch mode: input nodes already counted
	};
	Traffic traffic;
	bool in_input(u32 ptr) const { // the node existed in the batch's input state
		return (ptr & (cfg.words_per_bucket() - 1u)) < traffic.in_bw[ptr >> cfg.word_bits_per_bucket()];
	}
	void count_visit(u32 ptr, u32 words) {
		if (ptr == kNull)
			return;
		if (traffic.batch && (!in_input(ptr) || !traffic.seen.insert(ptr).second))
			return;
		traffic.visited += words;
	}
	void traffic_begin(bool batch) {
		traffic = Traffic{};
		traffic.batch = batch;
		if (batch)
			traffic.in_bw = bucket_words;
	}
	// Batch mode: every node of the final DAG that is not in the input was appended once by the
	// batch: its words hashed, the bucket's input words scanned (a miss reads them all), its words
	// written and the counter bumped. The per-edit upsert counts of the sequential replay are dropped.
	void batch_traffic_end(u32 root) {
		traffic.hashed = traffic.scanned = traffic.written = traffic.counters = traffic.upserts = 0;
		std::unordered_set<u32> done;
		const u32 L = cfg.node_levels();
		std::vector<std::pair<u32, u32>> stack;
		if (root != kNull)
			stack.push_back({root, 0});
		while (!stack.empty()) {
			auto [ptr, l] = stack.back();
			stack.pop_back();
			if (in_input(ptr) || !done.insert(ptr).second)
				continue;
			const u32 *p = read_node(ptr);
			const u32 sz = l + 1 == L ? 2u : inner_words(p[0]);
			traffic.hashed += sz;
			traffic.scanned += traffic.in_bw[ptr >> cfg.word_bits_per_bucket()];
			traffic.written += sz;
			traffic.counters += 1;
			traffic.upserts += 1;
			if (l + 1 < L)
				for (u32 k = 1; k < sz; ++k)
					stack.push_back({p[k], l + 1});
		}
	}

	explicit NodePool(Config c) : cfg(std::move(c)) {
		if (!Config::validate(cfg))
			throw std::invalid_argument("invalid config");
		bases = cfg.level_bases();
		bucket_words.assign(cfg.total_buckets(), 0);
		pages.resize(cfg.total_pages());
	}

	u32 *page(u32 p) {
		if (!pages[p]) {
			pages[p] = std::make_unique<u32[]>(cfg.words_per_page());
			std::fill(pages[p].get(), pages[p].get() + cfg.words_per_page(), 0u);
		}
		return pages[p].get();
	}
	const u32 *read_page(u32 p) const { return pages[p].get(); }
	const u32 *read_node(u32 ptr) const { // NodePool.hpp:75-77
		return read_page(ptr >> cfg.wbpp) + (ptr & (cfg.words_per_page() - 1u));
	}
	u32 word(u64 w) const {
		const u32 *p = pages[w >> cfg.wbpp].get();
		return p ? p[w & (cfg.words_per_page() - 1u)] : 0u;
	}

	// find_node_in_span (NodePool.hpp:79-91)
	static u32 find_in_span(bool leaf, u32 base, const u32 *w, u32 n, const u32 *node, u32 nsz, u64 *scanned) {
		u32 i = 0;
		while (nsz <= n - i) {
			u32 sz = leaf ? 2u : inner_words(w[i]);
			if (sz == 0)
				break;
			if (sz == nsz && std::equal(node, node + nsz, w + i)) {
				*scanned += i + nsz;
				return base + i;
			}
			i += sz;
		}
		*scanned += i;
		return kNull;
	}
	// find_node (NodePool.hpp:93-132)
	u32 find(bool leaf, u32 bucket, u32 bwords, u32 boff, const u32 *node, u32 nsz, u64 *scanned) const {
		if (bwords <= boff)
			return kNull;
		u32 find_words = bwords - boff;
		u32 pidx = (bucket << cfg.pbpb) | (boff >> cfg.wbpp);
		u32 poff = boff & (cfg.words_per_page() - 1u);
		while (poff + find_words >= cfg.words_per_page()) {
			u32 r = find_in_span(leaf, (pidx << cfg.wbpp) | poff, read_page(pidx) + poff, cfg.words_per_page() - poff,
			                     node, nsz, scanned);
			if (r != kNull)
				return r;
			find_words -= cfg.words_per_page() - poff;
			poff = 0;
			++pidx;
		}
		if (find_words) {
			u32 r = find_in_span(leaf, (pidx << cfg.wbpp) | poff, read_page(pidx) + poff, find_words, node, nsz,
			                     scanned);
			if (r != kNull)
				return r;
		}
		return kNull;
	}
	// append_node (NodePool.hpp:134-157)
	std::pair<u32, u32> append(u32 bucket, u32 bwords, const u32 *node, u32 nsz) {
		if (bwords + nsz > cfg.words_per_bucket())
			return {kNull, 0u};
		u32 slot = bwords >> cfg.wbpp, off = bwords & (cfg.words_per_page() - 1u);
		u32 pidx = (bucket << cfg.pbpb) | slot;
		if (off + nsz > cfg.words_per_page()) {
			u32 *pg = page(pidx);
			std::fill(pg + off, pg + cfg.words_per_page(), 0u);
			++slot;
			++pidx;
			off = 0;
		}
		std::copy(node, node + nsz, page(pidx) + off);
		return {(pidx << cfg.wbpp) | off, ((slot << cfg.wbpp) | off) + nsz};
	}
	u32 bucket_of(u32 level, bool leaf, const u32 *node, u32 nsz) const { // NodePool.hpp:163-164
		u32 h = leaf ? murmur_leaf(node[0], node[1]) : murmur_span(node, nsz);
		return bases[level] + (h & (cfg.buckets_at(level) - 1u));
	}
	// upsert_node, single-threaded branch (NodePool.hpp:199-211)
	u32 upsert(u32 level, bool leaf, const u32 *node, u32 nsz, u32 fallback) {
		u32 b = bucket_of(level, leaf, node, nsz);
		u32 bw = bucket_words[b];
		traffic.hashed += nsz;
		traffic.upserts += 1;
		u32 f = find(leaf, b, bw, 0, node, nsz, &traffic.scanned);
		if (f != kNull)
			return f;
		auto [p, nw] = append(b, bw, node, nsz);
		if (p != kNull) {
			bucket_words[b] = nw;
			traffic.written += nw - bw;
			traffic.counters += 1;
			return p;
		}
		++overflows;
		return fallback;
	}
	// make_filled_node_pointers (NodePool.hpp:240-262)
	void make_filled() {
		if (!filled.empty())
			return;
		u32 L = cfg.node_levels();
		filled.assign(L, kNull);
		u32 leaf[2] = {~0u, ~0u};
		filled[L - 1] = upsert(L - 1, true, leaf, 2, kNull);
		if (L == 1)
			return;
		for (int l = (int)L - 2; l >= 0; --l) {
			u32 n[9];
			n[0] = 0xFFu;
			for (int i = 1; i < 9; ++i)
				n[i] = filled[l + 1];
			filled[l] = upsert((u32)l, false, n, 9, kNull);
		}
	}
	std::array<u32, 9> unpack(u32 ptr) const { // NodePool.hpp:278-309
		std::array<u32, 9> u;
		u.fill(kNull);
		u[0] = 0;
		if (ptr == kNull)
			return u;
		const u32 *p = read_node(ptr);
		u[0] = p[0];
		u32 k = 1;
		for (u32 i = 0; i < 8; ++i)
			if ((p[0] >> i) & 1u)
				u[1 + i] = p[k++];
		return u;
	}
	static u32 pack(std::array<u32, 9> &u) { // NodePool.hpp:272-277 (in place)
		u32 k = 1;
		for (u32 i = 0; i < 8; ++i)
			if ((u[0] >> i) & 1u)
				u[k++] = u[1 + i];
		return k;
	}

	// ---- Edit recursion with the VBREditorWrapper / StatelessEditorWrapper restated.
	struct State { // VBREditorWrapper::NodeState (VBREditor.hpp:28-32)
		ChunkWriter *writer = nullptr;
		u32 octree = ColorPool::kNullPtr;
		bool is_final = false;
	};
	struct Ctx {
		const Editor *ed;
		ColorPool *cp; // null => stateless
		u32 octree_root;
	};

	// VBREditorWrapper::EditNode (VBREditor.hpp:37-71) / StatelessEditorWrapper::EditNode
	EditType wrap_edit_node(const Ctx &x, u32 level, const u32 pos[3], u32 ptr, State &st, const State &ps) const {
		if (!x.cp)
			return x.ed->edit_node(cfg, level, pos);
		ColorPool &cp = *x.cp;
		u32 child_idx = (pos[0] & 1u) | ((pos[1] & 1u) << 1) | ((pos[2] & 1u) << 2);
		st.octree = level <= cp.leaf_level ? (level == 0 ? x.octree_root : cp.get_child(ps.octree, child_idx))
		                                   : ps.octree;
		st.writer = ps.writer;
		st.is_final = ps.is_final;
		VBRColor fill = ColorPool::get_fill(st.octree), color = fill;
		EditType t = x.ed->edit_node_vbr(cfg, level, pos, ptr != kNull, color);
		if (!st.is_final) {
			if (level <= cp.leaf_level) {
				if (color.has) {
					st.octree = ColorPool::fill_node(st.octree, color);
					st.is_final = true;
				} else if (t == EditType::kClear)
					st.octree = ColorPool::clear_node(st.octree);
				else if (t == EditType::kProceed && level == cp.leaf_level)
					st.writer = new ChunkWriter(cp.get_leaf(st.octree));
			} else if (st.writer) {
				u32 n = 1u << ((cfg.voxel_level() - level) * 3u);
				if (color.has) {
					st.writer->Push(color, n);
					st.is_final = true;
				} else if (t != EditType::kProceed)
					st.writer->Copy(n, fill);
			}
		}
		return t;
	}
	// VBREditorWrapper::EditVoxel (VBREditor.hpp:72-80)
	bool wrap_edit_voxel(const Ctx &x, const u32 p[3], bool voxel, State &st) const {
		if (!x.cp)
			return x.ed->edit_voxel(p, voxel);
		if (st.is_final) {
			VBRColor dummy{};
			return x.ed->edit_voxel_vbr(p, voxel, dummy);
		}
		return st.writer->Edit([&](VBRColor &c) { return x.ed->edit_voxel_vbr(p, voxel, c); },
		                       ColorPool::get_fill(st.octree));
	}
	// VBREditorWrapper::JoinNode / JoinLeaf (VBREditor.hpp:82-106)
	void wrap_join(const Ctx &x, u32 level, State &st, const std::array<State, 8> *children) const {
		if (!x.cp)
			return;
		ColorPool &cp = *x.cp;
		if (level == cp.leaf_level) {
			if (st.writer) {
				st.octree = cp.set_leaf(st.octree, st.writer->flush());
				delete st.writer;
				st.writer = nullptr;
			}
		} else if (level < cp.leaf_level && children && !st.is_final) {
			std::array<u32, 8> ch;
			for (int i = 0; i < 8; ++i)
				ch[i] = (*children)[i].octree;
			st.octree = cp.set_node(st.octree, ch);
		}
	}

	// edit_leaf (NodePool.hpp:319-343)
	u32 edit_leaf(const Ctx &x, u32 ptr, u32 level, const u32 pos[3], State &st) {
		u32 leaf[2] = {0, 0};
		if (ptr != kNull) {
			const u32 *p = read_node(ptr);
			leaf[0] = p[0];
			leaf[1] = p[1];
		}
		count_visit(ptr, 2);
		bool changed = false;
		for (u32 i = 0; i < 64; ++i) {
			bool v = (leaf[i >> 5] >> (i & 31u)) & 1u;
			u32 vp[3] = {(pos[0] << 2) | ((i >> 2) & 2u) | (i & 1u), (pos[1] << 2) | ((i >> 3) & 2u) | ((i >> 1) & 1u),
			             (pos[2] << 2) | ((i >> 4) & 2u) | ((i >> 2) & 1u)}; // NodeCoord::GetLeafCoord
			bool nv = wrap_edit_voxel(x, vp, v, st);
			changed |= nv != v;
			leaf[i >> 5] ^= (u32)(nv != v) << (i & 31u);
		}
		wrap_join(x, level, st, nullptr);
		if (!changed)
			return ptr;
		if (leaf[0] == 0 && leaf[1] == 0)
			return kNull;
		return upsert(level, true, leaf, 2, ptr);
	}
	// edit_switch (NodePool.hpp:345-360) + edit_node (NodePool.hpp:362-396)
	u32 edit_node(const Ctx &x, u32 ptr, u32 level, const u32 pos[3], State &st) {
		if (level == cfg.node_levels() - 1)
			return edit_leaf(x, ptr, level, pos, st);
		std::array<u32, 9> u = unpack(ptr);
		count_visit(ptr, inner_words(u[0]));
		std::array<State, 8> cs{};
		bool changed = false;
		for (u32 i = 0; i < 8; ++i) {
			u32 cpos[3] = {(pos[0] << 1) | (i & 1u), (pos[1] << 1) | ((i >> 1) & 1u), (pos[2] << 1) | ((i >> 2) & 1u)};
			u32 cptr = u[1 + i];
			EditType t = wrap_edit_node(x, level + 1, cpos, cptr, cs[i], st);
			u32 nptr;
			if (t == EditType::kClear)
				nptr = kNull;
			else if (t == EditType::kFill)
				nptr = filled[level + 1];
			else if (t == EditType::kNotAffected)
				nptr = cptr;
			else
				nptr = edit_node(x, cptr, level + 1, cpos, cs[i]);
			changed |= nptr != cptr;
			u[1 + i] = nptr;
			u[0] ^= (u32)((nptr != kNull) != (cptr != kNull)) << i;
		}
		wrap_join(x, level, st, &cs);
		if (!changed)
			return ptr;
		if (!u[0])
			return kNull;
		u32 n = pack(u);
		return upsert(level, false, u.data(), n, ptr);
	}
	// Edit (NodePool.hpp:405-417). Returns the new root; *color_root is updated for VBR edits.
	u32 edit(u32 root, const Editor &ed, ColorPool *cp, u32 *color_root) {
		make_filled();
		Ctx x{&ed, ed.stateful() ? cp : nullptr, color_root ? *color_root : ColorPool::kNullPtr};
		State st{}, ps{};
		u32 pos[3] = {0, 0, 0};
		EditType t = wrap_edit_node(x, 0, pos, root, st, ps);
		u32 r;
		if (t == EditType::kClear)
			r = kNull;
		else if (t == EditType::kFill)
			r = filled[0];
		else if (t == EditType::kNotAffected)
			r = root;
		else
			r = edit_node(x, root, 0, pos, st);
		if (x.cp && color_root)
			*color_root = st.octree;
		return r;
	}

	// ---- GC: NodePoolThreadedGC (NodePoolThreadedGC.hpp:29-403) on one worker. Live nodes are
	// processed in increasing old-pointer order; leaves are compacted in their bucket, inner
	// nodes are remapped, re-hashed and appended to their (possibly new) bucket.
	std::vector<u32> gc(std::vector<u32> roots) {
		u32 L = cfg.node_levels();
		std::vector<std::vector<u32>> live(L);
		if (!filled.empty() && filled[0] != kNull)
			live[0].push_back(filled[0]);
		for (u32 r : roots)
			if (r != kNull)
				live[0].push_back(r);
		for (u32 l = 0; l < L; ++l) { // forward pass (:72-103)
			auto &v = live[l];
			std::sort(v.begin(), v.end());
			v.erase(std::unique(v.begin(), v.end()), v.end());
			if (l + 1 < L)
				for (u32 n : v) {
					const u32 *p = read_node(n);
					u32 k = 1;
					for (u32 i = 0; i < 8; ++i)
						if ((p[0] >> i) & 1u)
							live[l + 1].push_back(p[k++]);
				}
		}
		std::unordered_map<u32, u32> child_map, cur_map;
		u32 wpb = cfg.words_per_bucket();
		for (int l = (int)L - 1; l >= 0; --l) { // backward pass (:276-348)
			cur_map.clear();
			u32 b0 = bases[l], nb = cfg.buckets_at((u32)l);
			std::vector<std::vector<u32>> cache(nb);
			if (l == (int)L - 1) { // leaf shrink (:229-274), leaves keep their bucket
				std::vector<std::array<u32, 2>> content(live[l].size());
				for (size_t i = 0; i < live[l].size(); ++i) {
					const u32 *p = read_node(live[l][i]);
					content[i] = {p[0], p[1]};
				}
				for (size_t i = 0; i < live[l].size(); ++i) {
					u32 b = live[l][i] >> cfg.word_bits_per_bucket();
					cache[b - b0].push_back(content[i][0]);
					cache[b - b0].push_back(content[i][1]);
					cur_map[live[l][i]] = (b << cfg.word_bits_per_bucket()) | (u32)(cache[b - b0].size() - 2);
				}
			} else { // inner shrink (:105-182)
				for (u32 n : live[l]) {
					const u32 *p = read_node(n);
					u32 sz = inner_words(p[0]);
					u32 buf[9];
					std::copy(p, p + sz, buf);
					for (u32 i = 1; i < sz; ++i)
						buf[i] = child_map.at(buf[i]);
					u32 b = b0 + (murmur_span(buf, sz) & (nb - 1u));
					auto &c = cache[b - b0];
					u32 off = (u32)c.size();
					u32 slot = off >> cfg.wbpp, poff = off & (cfg.words_per_page() - 1u);
					if (poff + sz > cfg.words_per_page()) {
						off = (slot + 1) << cfg.wbpp;
						c.resize(off, 0u);
					}
					c.insert(c.end(), buf, buf + sz);
					cur_map[n] = (b << cfg.word_bits_per_bucket()) | off;
				}
			}
			for (u32 i = 0; i < nb; ++i) { // flush (:198-227)
				auto &c = cache[i];
				u32 b = b0 + i;
				if (c.size() > wpb)
					throw std::runtime_error("gc bucket overflow");
				for (size_t w = 0; w < c.size(); ++w)
					page((b << cfg.pbpb) | (u32)(w >> cfg.wbpp))[w & (cfg.words_per_page() - 1u)] = c[w];
				bucket_words[b] = (u32)c.size();
			}
			std::swap(child_map, cur_map);
		}
		for (u32 &r : roots)
			if (r != kNull)
				r = child_map.at(r);
		if (!filled.empty()) { // re-initialise the filled chain (:378-381)
			filled.clear();
			make_filled();
		}
		return roots;
	}

	// Per-level node counts of everything stored (bucket walk; zero words are page padding).
	std::vector<u64> level_node_counts() const {
		u32 L = cfg.node_levels();
		std::vector<u64> out(L, 0);
		for (u32 l = 0; l < L; ++l)
			for (u32 b = bases[l]; b < bases[l] + cfg.buckets_at(l); ++b) {
				u32 bw = bucket_words[b];
				if (l == L - 1) {
					out[l] += bw / 2;
					continue;
				}
				u32 i = 0;
				while (i < bw) {
					u32 w = word(((u64)b << cfg.word_bits_per_bucket()) + i);
					u32 sz = inner_words(w);
					if (sz == 0) {
						i = ((i >> cfg.wbpp) + 1) << cfg.wbpp;
						continue;
					}
					++out[l];
					i += sz;
				}
			}
		return out;
	}
};

// ---------------------------------------------------------------- Canonical DAG
// Reachable DAG from the roots, nodes numbered per level in first-visit DFS order (children
// 0..7). Serialised as: per level, [count, then per node: leaf words | mask + child ids].
// Equal serialisations <=> equal DAGs up to node renumbering.
template <typename WordFn> std::vector<u32> canonical(u32 L, WordFn &&word, const std::vector<u32> &roots) {
	std::vector<std::unordered_map<u32, u32>> ids(L);
	std::vector<std::vector<u32>> body(L);
	struct Rec {
		static u32 visit(u32 L, WordFn &word, std::vector<std::unordered_map<u32, u32>> &ids,
		                 std::vector<std::vector<u32>> &body, u32 level, u32 ptr) {
			auto it = ids[level].find(ptr);
			if (it != ids[level].end())
				return it->second;
			std::vector<u32> rec;
			if (level == L - 1) {
				rec = {word(ptr), word(ptr + 1)};
			} else {
				u32 m = word(ptr) & 0xFFu;
				rec.push_back(m);
				u32 k = 1;
				for (u32 i = 0; i < 8; ++i)
					if ((m >> i) & 1u)
						rec.push_back(visit(L, word, ids, body, level + 1, word(ptr + k++)));
			}
			u32 id = (u32)ids[level].size();
			ids[level][ptr] = id;
			body[level].insert(body[level].end(), rec.begin(), rec.end());
			return id;
		}
	};
	std::vector<u32> out;
	std::vector<u32> root_ids;
	for (u32 r : roots)
		root_ids.push_back(r == kNull ? kNull : Rec::visit(L, word, ids, body, 0, r));
	out.insert(out.end(), root_ids.begin(), root_ids.end());
	for (u32 l = 0; l < L; ++l) {
		out.push_back((u32)ids[l].size());
		out.insert(out.end(), body[l].begin(), body[l].end());
	}
	return out;
}

// Decode the voxel set below a root into a dense bitset (res^3 bits, x fastest).
template <typename WordFn> void decode_voxels(u32 L, WordFn &&word, u32 root, std::vector<uint8_t> &grid) {
	u32 res = 1u << (L + 1);
	grid.assign((size_t)res * res * res, 0);
	struct Rec {
		static void go(u32 L, WordFn &word, std::vector<uint8_t> &g, u32 res, u32 level, u32 ptr, u32 x, u32 y, u32 z) {
			if (level == L - 1) {
				u32 w[2] = {word(ptr), word(ptr + 1)};
				for (u32 i = 0; i < 64; ++i)
					if ((w[i >> 5] >> (i & 31u)) & 1u) {
						u32 vx = (x << 2) | ((i >> 2) & 2u) | (i & 1u), vy = (y << 2) | ((i >> 3) & 2u) | ((i >> 1) & 1u),
						    vz = (z << 2) | ((i >> 4) & 2u) | ((i >> 2) & 1u);
						g[((size_t)vz * res + vy) * res + vx] = 1;
					}
				return;
			}
			u32 m = word(ptr) & 0xFFu, k = 1;
			for (u32 i = 0; i < 8; ++i)
				if ((m >> i) & 1u)
					go(L, word, g, res, level + 1, word(ptr + k++), (x << 1) | (i & 1u), (y << 1) | ((i >> 1) & 1u),
					   (z << 1) | ((i >> 2) & 1u));
		}
	};
	if (root != kNull)
		Rec::go(L, word, grid, res, 0, root, 0, 0, 0);
}

// ---------------------------------------------------------------- Traversal
inline u32 f2u(float f) {
	u32 u;
	std::memcpy(&u, &f, 4);
	return u;
}
inline float u2f(u32 u) {
	float f;
	std::memcpy(&f, &u, 4);
	return f;
}
inline int find_msb(u32 v) { return v ? 31 - std::countl_zero(v) : -1; } // glm::findMSB(0) = -1

struct V3 {
	float x, y, z;
};
inline V3 normalize3(V3 v) {
	float l = std::sqrt(v.x * v.x + v.y * v.y + v.z * v.z);
	return {v.x / l, v.y / l, v.z / l};
}

struct Hit {
	bool hit = false;
	u32 vox[3] = {0, 0, 0};
	u32 vox_size = 0;
	float hit_pos[3] = {0, 0, 0};
	float norm[3] = {0, 0, 0};
	u32 iter = 0;
	float t = 0, scale = 0; // t_min and scale_exp2 where the march stopped (beam.frag's outputs)
};

// DAG_GetLeafFirstChildBits (trace.frag:56-67)
template <typename WordFn> u32 leaf_first_child_bits(WordFn &word, u32 node) {
	u32 l0 = word(node), l1 = word(node + 1), r = 0;
	for (u32 i = 0; i < 4; ++i) {
		r |= ((l0 >> (8 * i)) & 0xFFu) ? (1u << i) : 0u;
		r |= ((l1 >> (8 * i)) & 0xFFu) ? (1u << (4 + i)) : 0u;
	}
	return r;
}

// DAG_RayMarch (trace.frag:69-270), same fp32 operation order; after Laine & Karras, "Efficient
// Sparse Voxel Octrees" (I3D 2010). trace.frag carries the following notice for this march:
/*
 *  Copyright (c) 2009-2011, NVIDIA Corporation
 *  All rights reserved.
 *
 *  Redistribution and use in source and binary forms, with or without
 *  modification, are permitted provided that the following conditions are met:
 *      * Redistributions of source code must retain the above copyright
 *        notice, this list of conditions and the following disclaimer.
 *      * Redistributions in binary form must reproduce the above copyright
 *        notice, this list of conditions and the following disclaimer in the
 *        documentation and/or other materials provided with the distribution.
 *      * Neither the name of NVIDIA Corporation nor the
 *        names of its contributors may be used to endorse or promote products
 *        derived from this software without specific prior written permission.
 *
 *  THIS SOFTWARE IS PROVIDED BY THE COPYRIGHT HOLDERS AND CONTRIBUTORS "AS IS" AND
 *  ANY EXPRESS OR IMPLIED WARRANTIES, INCLUDING, BUT NOT LIMITED TO, THE IMPLIED
 *  WARRANTIES OF MERCHANTABILITY AND FITNESS FOR A PARTICULAR PURPOSE ARE
 *  DISCLAIMED. IN NO EVENT SHALL <COPYRIGHT HOLDER> BE LIABLE FOR ANY
 *  DIRECT, INDIRECT, INCIDENTAL, SPECIAL, EXEMPLARY, OR CONSEQUENTIAL DAMAGES
 *  (INCLUDING, BUT NOT LIMITED TO, PROCUREMENT OF SUBSTITUTE GOODS OR SERVICES;
 *  LOSS OF USE, DATA, OR PROFITS; OR BUSINESS INTERRUPTION) HOWEVER CAUSED AND
 *  ON ANY THEORY OF LIABILITY, WHETHER IN CONTRACT, STRICT LIABILITY, OR TORT
 *  (INCLUDING NEGLIGENCE OR OTHERWISE) ARISING IN ANY WAY OUT OF THE USE OF THIS
 *  SOFTWARE, EVEN IF ADVISED OF THE POSSIBILITY OF SUCH DAMAGE.
 */
template <typename WordFn>
Hit ray_march(WordFn &&word, u32 root, u32 leaf_level, float proj_factor, float proj_bias, V3 o, V3 d) {
	constexpr u32 S = 23;
	Hit out;
	if (root == kNull)
		return out;
	u32 stack[S] = {};
	u32 iter = 0;
	o.x += 1.0f, o.y += 1.0f, o.z += 1.0f;
	const float eps = u2f((127u - S) << 23u);
	d.x = std::fabs(d.x) > eps ? d.x : (d.x >= 0 ? eps : -eps);
	d.y = std::fabs(d.y) > eps ? d.y : (d.y >= 0 ? eps : -eps);
	d.z = std::fabs(d.z) > eps ? d.z : (d.z >= 0 ? eps : -eps);
	V3 tc = {1.0f / -std::fabs(d.x), 1.0f / -std::fabs(d.y), 1.0f / -std::fabs(d.z)};
	V3 tb = {tc.x * o.x, tc.y * o.y, tc.z * o.z};
	u32 oct = 0;
	if (d.x > 0.0f)
		oct ^= 1u, tb.x = 3.0f * tc.x - tb.x;
	if (d.y > 0.0f)
		oct ^= 2u, tb.y = 3.0f * tc.y - tb.y;
	if (d.z > 0.0f)
		oct ^= 4u, tb.z = 3.0f * tc.z - tb.z;
	float t_min = std::max(std::max(2.0f * tc.x - tb.x, 2.0f * tc.y - tb.y), 2.0f * tc.z - tb.z);
	float t_max = std::min(std::min(tc.x - tb.x, tc.y - tb.y), tc.z - tb.z);
	float h = t_max;
	t_min = std::max(t_min, 0.0f);
	t_max = std::min(t_max, 1.0f);
	u32 parent = root, child_bits = 0;
	V3 pos = {1.0f, 1.0f, 1.0f};
	u32 idx = 0;
	if (1.5f * tc.x - tb.x > t_min)
		idx ^= 1u, pos.x = 1.5f;
	if (1.5f * tc.y - tb.y > t_min)
		idx ^= 2u, pos.y = 1.5f;
	if (1.5f * tc.z - tb.z > t_min)
		idx ^= 4u, pos.z = 1.5f;
	u32 scale = S - 1;
	float scale_exp2 = 0.5f;
	const u32 leaf_scale = S - leaf_level;
	for (;;) {
		++iter;
		if (child_bits == 0u)
			child_bits = scale > leaf_scale ? word(parent) : (scale == leaf_scale ? leaf_first_child_bits(word, parent) : parent);
		V3 t_corner = {pos.x * tc.x - tb.x, pos.y * tc.y - tb.y, pos.z * tc.z - tb.z};
		float tc_max = std::min(std::min(t_corner.x, t_corner.y), t_corner.z);
		u32 child_shift = idx ^ oct, child_mask = 1u << child_shift;
		if ((child_bits & child_mask) != 0 && t_min <= t_max) {
			float half = scale_exp2 * 0.5f;
			V3 t_center = {half * tc.x + t_corner.x, half * tc.y + t_corner.y, half * tc.z + t_corner.z};
			if (scale < leaf_scale || scale_exp2 * proj_factor < tc_max + proj_bias)
				break;
			if (tc_max < h)
				stack[scale] = parent;
			h = tc_max;
			parent = scale > leaf_scale ? word(parent + 1 + (u32)std::popcount(child_bits & (child_mask - 1)))
			                            : (word(parent + (child_shift >> 2u)) >> ((child_shift & 3u) << 3u)) & 0xFFu;
			idx = 0u;
			--scale;
			scale_exp2 = half;
			if (t_center.x > t_min)
				idx ^= 1u, pos.x += scale_exp2;
			if (t_center.y > t_min)
				idx ^= 2u, pos.y += scale_exp2;
			if (t_center.z > t_min)
				idx ^= 4u, pos.z += scale_exp2;
			child_bits = 0;
			continue;
		}
		u32 step = 0;
		if (t_corner.x <= tc_max)
			step ^= 1u, pos.x -= scale_exp2;
		if (t_corner.y <= tc_max)
			step ^= 2u, pos.y -= scale_exp2;
		if (t_corner.z <= tc_max)
			step ^= 4u, pos.z -= scale_exp2;
		t_min = tc_max;
		idx ^= step;
		if ((idx & step) != 0) {
			u32 diff = 0;
			if (step & 1u)
				diff |= f2u(pos.x) ^ f2u(pos.x + scale_exp2);
			if (step & 2u)
				diff |= f2u(pos.y) ^ f2u(pos.y + scale_exp2);
			if (step & 4u)
				diff |= f2u(pos.z) ^ f2u(pos.z + scale_exp2);
			scale = (u32)find_msb(diff);
			if (scale >= S)
				break;
			scale_exp2 = u2f((scale - S + 127u) << 23u);
			parent = stack[scale];
			u32 shx = f2u(pos.x) >> scale, shy = f2u(pos.y) >> scale, shz = f2u(pos.z) >> scale;
			pos.x = u2f(shx << scale);
			pos.y = u2f(shy << scale);
			pos.z = u2f(shz << scale);
			idx = (shx & 1u) | ((shy & 1u) << 1u) | ((shz & 1u) << 2u);
			h = 0.0f;
			child_bits = 0;
		}
	}
	V3 t_corner = {tc.x * (pos.x + scale_exp2) - tb.x, tc.y * (pos.y + scale_exp2) - tb.y, tc.z * (pos.z + scale_exp2) - tb.z};
	float n[3];
	if (t_corner.x > t_corner.y && t_corner.x > t_corner.z)
		n[0] = -1, n[1] = 0, n[2] = 0;
	else if (t_corner.y > t_corner.z)
		n[0] = 0, n[1] = -1, n[2] = 0;
	else
		n[0] = 0, n[1] = 0, n[2] = -1;
	if ((oct & 1u) == 0u)
		n[0] = -n[0];
	if ((oct & 2u) == 0u)
		n[1] = -n[1];
	if ((oct & 4u) == 0u)
		n[2] = -n[2];
	const u32 voxel_level = leaf_level + 1u, voxel_scale = S - voxel_level;
	out.iter = iter;
	out.t = t_min;
	out.scale = scale_exp2;
	if (!(scale < S && t_min <= t_max))
		return out; // a miss reports zeros (trace.frag leaves them unused)
	u32 vox_size = 1u << (scale - voxel_scale);
	u32 vp[3] = {(f2u(pos.x) & 0x7FFFFFu) >> voxel_scale, (f2u(pos.y) & 0x7FFFFFu) >> voxel_scale,
	             (f2u(pos.z) & 0x7FFFFFu) >> voxel_scale};
	for (int i = 0; i < 3; ++i)
		if (oct & (1u << i))
			vp[i] = (1u << voxel_level) - vox_size - vp[i];
	float vmin[3], vmax[3], hp[3];
	float of[3] = {o.x, o.y, o.z}, df[3] = {d.x, d.y, d.z};
	for (int i = 0; i < 3; ++i) {
		vmin[i] = u2f((vp[i] << voxel_scale) | (127u << 23u));
		vmax[i] = u2f(((vp[i] + vox_size) << voxel_scale) | (127u << 23u));
		float v = of[i] + t_min * df[i];
		hp[i] = std::min(std::max(v, vmin[i]), vmax[i]); // GLSL clamp = min(max(x, lo), hi)
		if (n[i] != 0)
			hp[i] = n[i] > 0 ? vmax[i] + eps * 2 : vmin[i] - eps;
		hp[i] -= 1.0f;
	}
	out.hit = scale < S && t_min <= t_max;
	for (int i = 0; i < 3; ++i) {
		out.vox[i] = vp[i];
		out.hit_pos[i] = hp[i];
		out.norm[i] = n[i];
	}
	out.vox_size = vox_size;
	out.iter = iter;
	return out;
}

// Color_Morton32 (trace.frag:272-278)
inline u32 morton32_1(u32 u) {
	u = (u | (u << 16u)) & 0x030000FFu;
	u = (u | (u << 8u)) & 0x0300F00Fu;
	u = (u | (u << 4u)) & 0x030C30C3u;
	u = (u | (u << 2u)) & 0x09249249u;
	return u;
}
inline u32 morton32(u32 x, u32 y, u32 z) { return morton32_1(x) | (morton32_1(y) << 1u) | (morton32_1(z) << 2u); }

// Color fetch result: the float RGB the shader computes (trace.frag:272-364).
struct RGBf {
	float r = 0, g = 0, b = 0;
};
inline RGBf unpack_unorm8(u32 c) {
	return {float(c & 0xFFu) / 255.0f, float((c >> 8) & 0xFFu) / 255.0f, float((c >> 16) & 0xFFu) / 255.0f};
}
inline RGBf unpack565(u32 c) {
	return {float(c & 0x1Fu) / 31.0f, float((c >> 5) & 0x3Fu) / 63.0f, float((c >> 11) & 0x1Fu) / 31.0f};
}
// Color_GetLeafColor (trace.frag:304-349)
inline RGBf leaf_color(const std::vector<u32> &L, u32 idx, const u32 vp[3]) {
	u32 macro_cnt = L[idx + 1u], block_cnt = L[idx + 2u];
	u32 macro_offset = idx + 4u, block_offset = macro_offset + (macro_cnt << 1u);
	u32 weight_offset = block_offset + (block_cnt << 1u);
	u32 vox_id = morton32(vp[0], vp[1], vp[2]);
	u32 macro_id = vox_id >> 14u;
	if (macro_id >= macro_cnt)
		return {};
	u32 mx = L[macro_offset + (macro_id << 1u)], my = L[(macro_offset + (macro_id << 1u)) | 1u];
	block_offset += mx << 1u;
	block_cnt = macro_id + 1u < macro_cnt ? L[macro_offset + ((macro_id + 1u) << 1u)] - mx : block_cnt - mx;
	vox_id &= 0x3FFFu;
	if (block_cnt == 0u)
		return {};
	for (u32 s = 0; s <= 14u && block_cnt != 0u; ++s) {
		u32 step = block_cnt >> 1u;
		if ((L[(block_offset + (step << 1u)) | 1u] >> 18u) <= vox_id) {
			block_cnt -= (step + 1u);
			block_offset += (step + 1u) << 1u;
		} else
			block_cnt = step;
	}
	block_offset -= 2u;
	u32 bx = L[block_offset], by = L[block_offset | 1u];
	u32 bpw = (by >> 16u) & 0x3u;
	if (bpw == 0u)
		return unpack_unorm8(bx);
	vox_id -= (by >> 18u);
	u32 bit_id = my + (by & 0xFFFFu) + vox_id * bpw;
	u32 bit_ofst = bit_id & 31u;
	u32 w = L[weight_offset + (bit_id >> 5u)] >> bit_ofst;
	if (bit_ofst + bpw <= 32)
		w &= (1u << bpw) - 1u;
	else {
		u32 w1 = L[weight_offset + (bit_id >> 5u) + 1u] & ((1u << (bit_ofst + bpw - 32u)) - 1u);
		w = w | (w1 << (32u - bit_ofst));
	}
	float a = float(w) / float((1u << bpw) - 1u);
	RGBf x = unpack565(bx), y = unpack565(bx >> 16u);
	return {x.r * (1.0f - a) + y.r * a, x.g * (1.0f - a) + y.g * a, x.b * (1.0f - a) + y.b * a};
}
// Color_Fetch (trace.frag:351-364)
inline RGBf color_fetch(const ColorPool &cp, u32 root, u32 voxel_level, u32 leaf_level, const u32 vp[3]) {
	u32 ptr = root;
	for (u32 l = 0; l < leaf_level; ++l) {
		u32 tag = ptr >> 30u, data = ptr & 0x3FFFFFFFu;
		if (tag != 0)
			return unpack_unorm8(data);
		u32 o[3] = {(vp[0] >> (voxel_level - 1u - l)) & 1u, (vp[1] >> (voxel_level - 1u - l)) & 1u,
		            (vp[2] >> (voxel_level - 1u - l)) & 1u};
		ptr = cp.nodes[ptr][o[0] | (o[1] << 1u) | (o[2] << 2u)];
	}
	u32 tag = ptr >> 30u, data = ptr & 0x3FFFFFFFu;
	if (tag == 2u) {
		u32 m = (1u << (voxel_level - leaf_level)) - 1u;
		u32 lp[3] = {vp[0] & m, vp[1] & m, vp[2] & m};
		return leaf_color(cp.leaves, data, lp);
	}
	return unpack_unorm8(data);
}

// Beam prepass (beam.frag:59-223, BeamPass.cpp:9-111): one coarse ray per 8x8-pixel block through
// the block's centre, Camera_GenRay over the ceil(W/8) x ceil(H/8) beam image, LOD with the
// projection factor of a block (the view's proj_factor / 8); the block's value is
// max(t - cube size, 0) where the march stopped, +inf on a miss.
constexpr u32 kBeamBlock = 8;
template <typename WordFn>
inline float beam_value(WordFn &word, u32 root, u32 L, float proj, V3 pos, V3 look, V3 side, V3 up, u32 bx, u32 by,
                        u32 Wb, u32 Hb) {
	float cx = (float(bx) + 0.5f) / float(Wb), cy = (float(by) + 0.5f) / float(Hb);
	cx = cx * 2.0f - 1.0f;
	cy = cy * 2.0f - 1.0f;
	V3 d = normalize3(V3{look.x - side.x * cx - up.x * cy, look.y - side.y * cx - up.y * cy,
	                     look.z - side.z * cx - up.z * cy});
	Hit h = ray_march(word, root, L, proj * (1.0f / float(kBeamBlock)), 0.0f, pos, d);
	return h.hit ? std::max(h.t - h.scale, 0.0f) : INFINITY;
}
// The trace's beam lookup (trace.frag:384-390): a MIN-reduction bilinear fetch, i.e. the smallest of
// the 2x2 beam texels around the pixel centre (clamped to the edge), scaled by 0.98.
inline float beam_sample(const float *B, u32 Wb, u32 Hb, float u, float v) {
	float x = u * float(Wb) - 0.5f, y = v * float(Hb) - 0.5f;
	int x0 = (int)std::floor(x), y0 = (int)std::floor(y);
	auto cl = [](int a, u32 n) { return (u32)std::min(std::max(a, 0), (int)n - 1); };
	u32 xa = cl(x0, Wb), xb = cl(x0 + 1, Wb), ya = cl(y0, Hb), yb = cl(y0 + 1, Hb);
	return std::min(std::min(B[ya * Wb + xa], B[ya * Wb + xb]), std::min(B[yb * Wb + xa], B[yb * Wb + xb])) * 0.98f;
}

// Brute-force DDA over a dense grid in double precision (Amanatides & Woo), with the
// same [0,1]^3 world, entry clamp t >= 0 and the t <= 1 cap of trace.frag:110-111.
inline bool dda(const std::vector<uint8_t> &g, u32 res, const double o[3], const double d[3], u32 out[3],
                double *t_hit) {
	double t0 = 0.0, t1 = 1.0;
	for (int i = 0; i < 3; ++i) {
		if (d[i] == 0.0) {
			if (o[i] < 0.0 || o[i] > 1.0)
				return false;
			continue;
		}
		double a = (0.0 - o[i]) / d[i], b = (1.0 - o[i]) / d[i];
		if (a > b)
			std::swap(a, b);
		t0 = std::max(t0, a);
		t1 = std::min(t1, b);
	}
	if (t0 > t1)
		return false;
	int v[3], step[3];
	double tmax[3], tdelta[3];
	for (int i = 0; i < 3; ++i) {
		double p = (o[i] + t0 * d[i]) * res;
		int c = (int)std::floor(p);
		if (d[i] < 0 && p == std::floor(p))
			c -= 1;
		c = std::min(std::max(c, 0), (int)res - 1);
		v[i] = c;
		step[i] = d[i] > 0 ? 1 : -1;
		if (d[i] == 0) {
			tmax[i] = INFINITY;
			tdelta[i] = INFINITY;
		} else {
			double nb = d[i] > 0 ? (c + 1.0) / res : (double)c / res;
			tmax[i] = (nb - o[i]) / d[i];
			tdelta[i] = 1.0 / (res * std::fabs(d[i]));
		}
	}
	double t = t0;
	while (t <= t1) {
		if (g[((size_t)v[2] * res + v[1]) * res + v[0]]) {
			for (int i = 0; i < 3; ++i)
				out[i] = (u32)v[i];
			*t_hit = t;
			return true;
		}
		int a = tmax[0] < tmax[1] ? (tmax[0] < tmax[2] ? 0 : 2) : (tmax[1] < tmax[2] ? 1 : 2);
		t = tmax[a];
		v[a] += step[a];
		if (v[a] < 0 || v[a] >= (int)res)
			return false;
		tmax[a] += tdelta[a];
	}
	return false;
}

} // namespace hdo
