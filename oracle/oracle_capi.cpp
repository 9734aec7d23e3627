// oracle_capi.cpp -- extern "C" surface of the CPU oracle (hd_oracle.hpp) for the test suite.
// TEST INFRASTRUCTURE ONLY: loaded by tests/ and bench.py's cpu_baseline leg, never by the product.
#include "hd_oracle.hpp"

#include <atomic>
#include <chrono>
#include <cstdio>
#include <thread>

using namespace hdo;

namespace {
struct OPool {
	NodePool pool;
	ColorPool color;
	Heightmap hm;
	explicit OPool(Config c) : pool(std::move(c)) {}
};

Config make_config(u32 wbpp, u32 pbpb, u32 L, const u32 *bits) {
	Config c;
	c.wbpp = wbpp;
	c.pbpb = pbpb;
	c.bucket_bits.assign(bits, bits + L);
	return c;
}

struct WordView {
	const u32 *w;
	u64 n;
	u32 operator()(u64 i) const { return i < n ? w[i] : 0u; }
};
} // namespace

extern "C" {

void *hdo_pool_create(u32 wbpp, u32 pbpb, u32 L, const u32 *bucket_bits, u32 color_leaf_level, u32 keep_history) {
	Config c = make_config(wbpp, pbpb, L, bucket_bits);
	if (!Config::validate(c) || color_leaf_level > L - 1)
		return nullptr;
	auto *p = new OPool(c);
	p->color.leaf_level = color_leaf_level;
	p->color.keep_history = keep_history != 0;
	return p;
}
void hdo_pool_destroy(void *p) { delete static_cast<OPool *>(p); }

void hdo_set_heightmap(void *p, const u32 *h, u32 res) { static_cast<OPool *>(p)->hm.set(h, res); }
void hdo_set_color_caps(void *p, u64 node_cap, u64 leaf_cap) {
	auto *o = static_cast<OPool *>(p);
	o->color.node_cap = node_cap;
	o->color.leaf_cap = leaf_cap;
}

// Sequential application of n edits (NodePoolBase::Edit each). Returns the final root;
// *color_root is threaded through the colored edits. If per_edit_roots != null it receives
// the root after each edit.
u32 hdo_edit(void *p, u32 root, const hd_edit_desc *edits, u32 n, u32 *color_root, u32 *per_edit_roots) {
	auto *o = static_cast<OPool *>(p);
	for (u32 i = 0; i < n; ++i) {
		Editor ed;
		ed.d = edits[i];
		ed.hm = &o->hm;
		root = o->pool.edit(root, ed, &o->color, color_root);
		if (per_edit_roots)
			per_edit_roots[i] = root;
	}
	return root;
}
// Same, timed: returns elapsed seconds of the edit calls (steady_clock around each, main.cpp:152-157).
double hdo_edit_timed(void *p, u32 root, const hd_edit_desc *edits, u32 n, u32 *color_root, u32 *root_out) {
	auto *o = static_cast<OPool *>(p);
	double total = 0;
	for (u32 i = 0; i < n; ++i) {
		Editor ed;
		ed.d = edits[i];
		ed.hm = &o->hm;
		auto t0 = std::chrono::steady_clock::now();
		root = o->pool.edit(root, ed, &o->color, color_root);
		total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	}
	*root_out = root;
	return total;
}

void hdo_gc(void *p, u32 *roots, u32 n) {
	auto *o = static_cast<OPool *>(p);
	std::vector<u32> r(roots, roots + n);
	r = o->pool.gc(r);
	std::copy(r.begin(), r.end(), roots);
}

u64 hdo_overflows(void *p) { return static_cast<OPool *>(p)->pool.overflows; }
u64 hdo_total_words(void *p) { return static_cast<OPool *>(p)->pool.cfg.total_words(); }
u64 hdo_total_buckets(void *p) { return static_cast<OPool *>(p)->pool.cfg.total_buckets(); }
void hdo_read_words(void *p, u64 first, u64 count, u32 *out) {
	auto *o = static_cast<OPool *>(p);
	for (u64 i = 0; i < count; ++i)
		out[i] = o->pool.word(first + i);
}
void hdo_bucket_words(void *p, u32 *out) {
	auto *o = static_cast<OPool *>(p);
	std::copy(o->pool.bucket_words.begin(), o->pool.bucket_words.end(), out);
}
u32 hdo_filled(void *p, u32 *out) {
	auto *o = static_cast<OPool *>(p);
	std::copy(o->pool.filled.begin(), o->pool.filled.end(), out);
	return (u32)o->pool.filled.size();
}
void hdo_make_filled(void *p) { static_cast<OPool *>(p)->pool.make_filled(); }
void hdo_level_counts(void *p, u64 *out) {
	auto c = static_cast<OPool *>(p)->pool.level_node_counts();
	std::copy(c.begin(), c.end(), out);
}
u64 hdo_color_node_count(void *p) { return static_cast<OPool *>(p)->color.nodes.size(); }
u64 hdo_color_leaf_count(void *p) { return static_cast<OPool *>(p)->color.leaves.size(); }
void hdo_color_nodes(void *p, u32 *out) {
	auto *o = static_cast<OPool *>(p);
	for (size_t i = 0; i < o->color.nodes.size(); ++i)
		std::copy(o->color.nodes[i].begin(), o->color.nodes[i].end(), out + 8 * i);
}
void hdo_color_leaves(void *p, u32 *out) {
	auto *o = static_cast<OPool *>(p);
	std::copy(o->color.leaves.begin(), o->color.leaves.end(), out);
}

// Upsert of raw nodes (the hash-consing known-answer tests, test.cpp:118-160 intent).
u32 hdo_upsert(void *p, u32 level, u32 leaf, const u32 *node, u32 nsz, u32 fallback) {
	auto *o = static_cast<OPool *>(p);
	return o->pool.upsert(level, leaf != 0, node, nsz, fallback);
}

u32 hdo_hash_span(const u32 *w, u32 n) { return murmur_span(w, n); }
u32 hdo_hash_leaf(u32 w0, u32 w1) { return murmur_leaf(w0, w1); }

// VBR writer driver: ops = (kind, count, colors, weight | bpw << 8 | has << 16).
// kind 0 = Copy(count, fill), 1 = Push(color, count), 2 = Edit (per voxel, color replaced
// when has != 0). The source chunk is given as leaf words at src (layout of
// DAGColorPool::write_leaf_chunk without the alloc word) or null.
u64 hdo_vbr_write(const u32 *src, u32 src_words, const u32 *ops, u32 nops, u32 *out, u64 cap) {
	Chunk sc;
	if (src && src_words) {
		u32 M = src[0], B = src[1], W = src[2], i = 3;
		for (u32 k = 0; k < M; ++k, i += 2)
			sc.macros.push_back({src[i], src[i + 1]});
		for (u32 k = 0; k < B; ++k, i += 2)
			sc.headers.push_back({src[i], src[i + 1]});
		sc.weights.assign(src + i, src + i + W);
	}
	ChunkWriter w(sc);
	for (u32 k = 0; k < nops; ++k) {
		const u32 *op = ops + 4 * k;
		VBRColor c;
		c.colors = op[2];
		c.weight = op[3] & 0xFFu;
		c.bpw = (op[3] >> 8) & 0xFFu;
		c.has = (op[3] >> 16) & 1u;
		if (!c.has)
			c = VBRColor{}; // an empty VBRColor is all-zero (VBRColor.hpp:29-31)
		if (op[0] == 0)
			w.Copy(op[1], c);
		else if (op[0] == 1)
			w.Push(c, op[1]);
		else
			for (u32 j = 0; j < op[1]; ++j)
				w.Edit([&](VBRColor &x) {
					if (c.has)
						x = c;
					return true;
				}, VBRColor{});
	}
	Chunk ch = w.flush();
	std::vector<u32> o = {(u32)ch.macros.size(), (u32)ch.headers.size(), (u32)ch.weights.size()};
	for (auto &m : ch.macros)
		o.push_back(m[0]), o.push_back(m[1]);
	for (auto &h : ch.headers)
		o.push_back(h[0]), o.push_back(h[1]);
	o.insert(o.end(), ch.weights.begin(), ch.weights.end());
	for (u64 i = 0; i < o.size() && i < cap; ++i)
		out[i] = o[i];
	return o.size();
}

// ---- Array-based checkers (also applied to pools downloaded from the device).

// Canonical serialisation of the DAG reachable from roots (hd_oracle.hpp canonical()).
u64 hdo_canonical_words(const u32 *words, u64 nwords, u32 L, const u32 *roots, u32 nroots, u32 *out, u64 cap) {
	WordView v{words, nwords};
	auto c = canonical(L, v, std::vector<u32>(roots, roots + nroots));
	for (u64 i = 0; i < c.size() && i < cap; ++i)
		out[i] = c[i];
	return c.size();
}
u64 hdo_canonical(void *p, const u32 *roots, u32 nroots, u32 *out, u64 cap) {
	auto *o = static_cast<OPool *>(p);
	auto fn = [o](u64 i) { return o->pool.word(i); };
	auto c = canonical(o->pool.cfg.node_levels(), fn, std::vector<u32>(roots, roots + nroots));
	for (u64 i = 0; i < c.size() && i < cap; ++i)
		out[i] = c[i];
	return c.size();
}
void hdo_decode_voxels_words(const u32 *words, u64 nwords, u32 L, u32 root, uint8_t *grid) {
	WordView v{words, nwords};
	std::vector<uint8_t> g;
	decode_voxels(L, v, root, g);
	std::copy(g.begin(), g.end(), grid);
}
void hdo_decode_voxels(void *p, u32 root, uint8_t *grid) {
	auto *o = static_cast<OPool *>(p);
	auto fn = [o](u64 i) { return o->pool.word(i); };
	std::vector<uint8_t> g;
	decode_voxels(o->pool.cfg.node_levels(), fn, root, g);
	std::copy(g.begin(), g.end(), grid);
}
// Color_Fetch of n voxels (trace.frag:351-364) from raw color arrays; out = 3 floats each.
void hdo_color_fetch_arrays(const u32 *nodes, u64 nnodes, const u32 *leaves, u64 nleaves, u32 root, u32 voxel_level,
                            u32 leaf_level, const u32 *vox, u64 n, float *out) {
	ColorPool cp;
	cp.nodes.resize(nnodes);
	for (u64 i = 0; i < nnodes; ++i)
		std::copy(nodes + 8 * i, nodes + 8 * i + 8, cp.nodes[i].begin());
	cp.leaves.assign(leaves, leaves + nleaves);
	for (u64 i = 0; i < n; ++i) {
		RGBf c = color_fetch(cp, root, voxel_level, leaf_level, vox + 3 * i);
		out[3 * i] = c.r, out[3 * i + 1] = c.g, out[3 * i + 2] = c.b;
	}
}
void hdo_color_fetch(void *p, u32 root, const u32 *vox, u64 n, float *out) {
	auto *o = static_cast<OPool *>(p);
	for (u64 i = 0; i < n; ++i) {
		RGBf c = color_fetch(o->color, root, o->pool.cfg.voxel_level(), o->color.leaf_level, vox + 3 * i);
		out[3 * i] = c.r, out[3 * i + 1] = c.g, out[3 * i + 2] = c.b;
	}
}

// Traversal of an explicit ray list with DAG_RayMarch (trace.frag:69-270).
// rec layout per ray (12 x 4 bytes): vox[3], vox_size | hit << 31, hit_pos[3] (float), norm[3] (float), iter, pad.
void hdo_march_rays_words(const u32 *words, u64 nwords, u32 L, u32 root, float proj_factor, float proj_bias,
                          const float *o, const float *d, u64 n, u32 *rec) {
	WordView v{words, nwords};
	for (u64 i = 0; i < n; ++i) {
		Hit h = ray_march(v, root, L, proj_factor, proj_bias, V3{o[3 * i], o[3 * i + 1], o[3 * i + 2]},
		                  V3{d[3 * i], d[3 * i + 1], d[3 * i + 2]});
		u32 *r = rec + 12 * i;
		r[0] = h.vox[0], r[1] = h.vox[1], r[2] = h.vox[2];
		r[3] = h.vox_size | (h.hit ? 0x80000000u : 0u);
		std::memcpy(r + 4, h.hit_pos, 12);
		std::memcpy(r + 7, h.norm, 12);
		r[10] = h.iter;
		r[11] = 0;
	}
}

// Full frame as the device renders it: Camera_GenRay (trace.frag:366-370, 375), DAG_RayMarch,
// optional shadow ray toward normalize(4,5,3) from the hit position (LOD off), Color_Fetch and
// the diffuse shading of trace.frag:395-397 (shadowed pixels keep the 0.5 ambient term).
// view: pos[3], look[3], side[3], up[3], proj_factor; flags bit0 = shadow.
// rgba: W*H u32 (RGBA8, round-to-nearest), rec: W*H*12 u32 (hdo_march_rays_words layout;
// rec[11] = shadow hit flag).
} // extern "C"

namespace {
// One frame over rows [y0, y1) (see hdo_render_arrays).
template <typename WordFn>
void render_rows(const WordFn &v, u32 L, u32 root, const ColorPool &cp, u32 color_root, u32 color_leaf_level,
                 const float *view, u32 W, u32 H, u32 flags, const std::vector<float> &B, u32 y0, u32 y1, u32 *rgba,
                 u32 *rec) {
	V3 pos{view[0], view[1], view[2]}, look{view[3], view[4], view[5]}, side{view[6], view[7], view[8]},
	    up{view[9], view[10], view[11]};
	float proj = view[12];
	const V3 light = normalize3(V3{4.0f, 5.0f, 3.0f});
	const bool beam = (flags & 4u) != 0;
	const u32 Wb = (W + kBeamBlock - 1) / kBeamBlock, Hb = (H + kBeamBlock - 1) / kBeamBlock;
	for (u32 py = y0; py < y1; ++py)
		for (u32 px = 0; px < W; ++px) {
			float u = (float(px) + 0.5f) / float(W), w = (float(py) + 0.5f) / float(H);
			float cx = u * 2.0f - 1.0f, cy = w * 2.0f - 1.0f;
			V3 d = normalize3(V3{look.x - side.x * cx - up.x * cy, look.y - side.y * cx - up.y * cy,
			                     look.z - side.z * cx - up.z * cy});
			Hit h;
			if (!beam)
				h = ray_march(v, root, L, proj, 0.0f, pos, d);
			else {
				float t0 = beam_sample(B.data(), Wb, Hb, u, w);
				if (!std::isinf(t0))
					h = ray_march(v, root, L, proj, t0, V3{pos.x + t0 * d.x, pos.y + t0 * d.y, pos.z + t0 * d.z}, d);
			}
			bool shadow = false;
			if (h.hit && (flags & 1u)) {
				Hit s = ray_march(v, root, L, INFINITY, 0.0f, V3{h.hit_pos[0], h.hit_pos[1], h.hit_pos[2]}, light);
				shadow = s.hit;
			}
			float r = 0, g = 0, b = 0;
			if (h.hit) {
				float nd = h.norm[0] * light.x + h.norm[1] * light.y + h.norm[2] * light.z;
				float diffuse = shadow ? 0.5f : std::max(nd, 0.0f) * 0.5f + 0.5f;
				RGBf c = color_fetch(cp, color_root, L + 1, color_leaf_level, h.vox);
				r = diffuse * c.r, g = diffuse * c.g, b = diffuse * c.b;
			}
			auto q = [](float x) { return (u32)(std::min(std::max(x, 0.0f), 1.0f) * 255.0f + 0.5f); };
			size_t i = (size_t)py * W + px;
			rgba[i] = q(r) | (q(g) << 8) | (q(b) << 16) | (255u << 24);
			if (rec) {
				u32 *o = rec + 12 * i;
				o[0] = h.vox[0], o[1] = h.vox[1], o[2] = h.vox[2];
				o[3] = h.vox_size | (h.hit ? 0x80000000u : 0u);
				std::memcpy(o + 4, h.hit_pos, 12);
				std::memcpy(o + 7, h.norm, 12);
				o[10] = h.iter;
				o[11] = shadow ? 1u : 0u;
			}
		}
}

// A frame split by rows over `threads` host threads (each pixel is independent, so the image does
// not depend on the split).
template <typename WordFn>
void render_frame(const WordFn &v, u32 L, u32 root, const ColorPool &cp, u32 color_root, u32 color_leaf_level,
                  const float *view, u32 W, u32 H, u32 flags, u32 *rgba, u32 *rec, u32 threads) {
	V3 pos{view[0], view[1], view[2]}, look{view[3], view[4], view[5]}, side{view[6], view[7], view[8]},
	    up{view[9], view[10], view[11]};
	const bool beam = (flags & 4u) != 0;
	const u32 Wb = (W + kBeamBlock - 1) / kBeamBlock, Hb = (H + kBeamBlock - 1) / kBeamBlock;
	std::vector<float> B;
	if (beam) {
		B.resize((size_t)Wb * Hb);
		for (u32 by = 0; by < Hb; ++by)
			for (u32 bx = 0; bx < Wb; ++bx)
				B[(size_t)by * Wb + bx] = beam_value(v, root, L, view[12], pos, look, side, up, bx, by, Wb, Hb);
	}
	threads = std::max(1u, std::min(threads, H));
	if (threads == 1) {
		render_rows(v, L, root, cp, color_root, color_leaf_level, view, W, H, flags, B, 0, H, rgba, rec);
		return;
	}
	std::vector<std::thread> pool;
	std::atomic<u32> next{0};
	for (u32 t = 0; t < threads; ++t)
		pool.emplace_back([&] {
			for (u32 y; (y = next.fetch_add(8)) < H;)
				render_rows(v, L, root, cp, color_root, color_leaf_level, view, W, H, flags, B, y, std::min(H, y + 8),
				            rgba, rec);
		});
	for (auto &t : pool)
		t.join();
}
} // namespace

extern "C" {

// Full frame as the device renders it: Camera_GenRay (trace.frag:366-370, 375), DAG_RayMarch,
// optional shadow ray toward normalize(4,5,3) from the hit position (LOD off), Color_Fetch and
// the diffuse shading of trace.frag:395-397 (shadowed pixels keep the 0.5 ambient term).
// view: pos[3], look[3], side[3], up[3], proj_factor; flags bit0 = shadow, bit2 = beam prepass
// (beam.frag) with beam-started primary rays (trace.frag:384-390).
// rgba: W*H u32 (RGBA8, round-to-nearest), rec: W*H*12 u32 (hdo_march_rays_words layout;
// rec[11] = shadow hit flag).
void hdo_render_arrays(const u32 *words, u64 nwords, u32 L, u32 root, const u32 *cnodes, u64 ncnodes,
                       const u32 *cleaves, u64 ncleaves, u32 color_root, u32 color_leaf_level, const float *view,
                       u32 W, u32 H, u32 flags, u32 *rgba, u32 *rec) {
	WordView v{words, nwords};
	ColorPool cp;
	cp.nodes.resize(ncnodes);
	for (u64 i = 0; i < ncnodes; ++i)
		std::copy(cnodes + 8 * i, cnodes + 8 * i + 8, cp.nodes[i].begin());
	cp.leaves.assign(cleaves, cleaves + ncleaves);
	render_frame(v, L, root, cp, color_root, color_leaf_level, view, W, H, flags, rgba, rec, 1);
}
// The same frame straight from an oracle pool (no flat copy of the node pool), over `threads` threads.
void hdo_render_pool(void *p, u32 root, u32 color_root, const float *view, u32 W, u32 H, u32 flags, u32 *rgba,
                     u32 *rec, u32 threads) {
	auto *o = static_cast<OPool *>(p);
	auto fn = [o](u64 i) { return o->pool.word(i); };
	render_frame(fn, o->pool.cfg.node_levels(), root, o->color, color_root, o->color.leaf_level, view, W, H, flags,
	             rgba, rec, threads);
}

// Edit traffic counting (NodePool::Traffic): begin, apply edits with hdo_edit, end. out = words
// [visited, hashed, scanned, written, counters, upserts]; batch != 0 counts the edits between
// begin and end as one batch (input nodes once, appended nodes = the new nodes of the final DAG
// below root).
void hdo_traffic_begin(void *p, u32 batch) { static_cast<OPool *>(p)->pool.traffic_begin(batch != 0); }
void hdo_traffic_end(void *p, u32 root, u64 *out) {
	auto &pool = static_cast<OPool *>(p)->pool;
	if (pool.traffic.batch)
		pool.batch_traffic_end(root);
	const auto &t = pool.traffic;
	out[0] = t.visited, out[1] = t.hashed, out[2] = t.scanned, out[3] = t.written, out[4] = t.counters;
	out[5] = t.upserts;
}

// Used words of every bucket packed in bucket order (hd_pool_read_live_words' layout); returns the count.
u64 hdo_live_words(void *p, u32 *out, u64 cap) {
	auto *o = static_cast<OPool *>(p);
	const auto &pool = o->pool;
	const auto &cfg = pool.cfg;
	u64 at = 0;
	for (u64 b = 0; b < cfg.total_buckets(); ++b) {
		u64 base = b << cfg.word_bits_per_bucket();
		for (u32 i = 0; i < pool.bucket_words[b]; ++i, ++at)
			if (at < cap)
				out[at] = pool.word(base + i);
	}
	return at;
}

// Occupancy of n voxels (x, y, z) of the DAG below root: descend node by node (NodeCoord child
// index x | y << 1 | z << 2, NodeCoord.hpp:18-30; leaf bit index NodeCoord.hpp:32-43).
void hdo_voxel_query(void *p, u32 root, const u32 *vox, u64 n, uint8_t *out) {
	auto *o = static_cast<OPool *>(p);
	const auto &pool = o->pool;
	const u32 L = pool.cfg.node_levels();
	for (u64 k = 0; k < n; ++k) {
		const u32 x = vox[3 * k], y = vox[3 * k + 1], z = vox[3 * k + 2];
		u32 ptr = root;
		bool occ = false;
		for (u32 l = 0; ptr != kNull; ++l) {
			if (l + 1 == L) { // leaf: 4^3 voxels, bit = (hi-bits xyz) << 3 | (lo-bits xyz)
				const u32 lx = x & 3u, ly = y & 3u, lz = z & 3u;
				const u32 bit = (((lx >> 1) | ((ly >> 1) << 1) | ((lz >> 1) << 2)) << 3) |
				                ((lx & 1u) | ((ly & 1u) << 1) | ((lz & 1u) << 2));
				occ = (pool.word((u64)ptr + (bit >> 5)) >> (bit & 31u)) & 1u;
				break;
			}
			const u32 sh = L - l; // child size at level l + 1 is 2^(L - l) voxels
			const u32 ci = ((x >> sh) & 1u) | (((y >> sh) & 1u) << 1) | (((z >> sh) & 1u) << 2);
			const u32 m = pool.word(ptr) & 0xFFu;
			if (!((m >> ci) & 1u))
				break;
			ptr = pool.word((u64)ptr + 1 + (u32)std::popcount(m & ((1u << ci) - 1u)));
		}
		out[k] = occ ? 1 : 0;
	}
}

// Brute-force DDA on a dense grid (res^3 bytes): out = vox[3] + hit flag per ray.
void hdo_dda(const uint8_t *grid, u32 res, const float *o, const float *d, u64 n, u32 *out) {
	std::vector<uint8_t> g(grid, grid + (size_t)res * res * res);
	for (u64 i = 0; i < n; ++i) {
		double oo[3] = {o[3 * i], o[3 * i + 1], o[3 * i + 2]}, dd[3] = {d[3 * i], d[3 * i + 1], d[3 * i + 2]};
		u32 v[3] = {0, 0, 0};
		double t;
		bool hit = dda(g, res, oo, dd, v, &t);
		out[4 * i] = v[0], out[4 * i + 1] = v[1], out[4 * i + 2] = v[2], out[4 * i + 3] = hit ? 1u : 0u;
	}
}

// Same state loader as hdr_load_live for the oracle.
void hdo_load_live(void *p, const u32 *bucket_words, const u32 *live, u64 nlive, const u32 *filled, u32 nfilled,
                   const u32 *cnodes, u64 ncnodes, const u32 *cleaves, u64 ncleaves) {
	auto *o = static_cast<OPool *>(p);
	auto &pool = o->pool;
	const auto &cfg = pool.cfg;
	u64 nb = cfg.total_buckets(), at = 0;
	for (u64 b = 0; b < nb; ++b) {
		u32 bw = bucket_words[b];
		pool.bucket_words[b] = bw;
		for (u32 i = 0; i < bw; ++i) {
			u64 w = (b << cfg.word_bits_per_bucket()) + i;
			pool.page((u32)(w >> cfg.wbpp))[w & (cfg.words_per_page() - 1u)] = live[at++];
		}
	}
	(void)nlive;
	pool.filled.assign(filled, filled + nfilled);
	o->color.nodes.resize(ncnodes);
	for (u64 i = 0; i < ncnodes; ++i)
		std::copy(cnodes + 8 * i, cnodes + 8 * i + 8, o->color.nodes[i].begin());
	o->color.leaves.assign(cleaves, cleaves + ncleaves);
}

} // extern "C"

extern "C" {
// The beam image of one view (beam.frag), ceil(W/8) x ceil(H/8) floats.
void hdo_beam_image(const u32 *words, u64 nwords, u32 L, u32 root, const float *view, u32 W, u32 H, float *out) {
	WordView v{words, nwords};
	V3 pos{view[0], view[1], view[2]}, look{view[3], view[4], view[5]}, side{view[6], view[7], view[8]},
	    up{view[9], view[10], view[11]};
	const u32 Wb = (W + kBeamBlock - 1) / kBeamBlock, Hb = (H + kBeamBlock - 1) / kBeamBlock;
	for (u32 by = 0; by < Hb; ++by)
		for (u32 bx = 0; bx < Wb; ++bx)
			out[(size_t)by * Wb + bx] = beam_value(v, root, L, view[12], pos, look, side, up, bx, by, Wb, Hb);
}
} // extern "C"
