// hd_trace.hip -- ESVO ray traversal of the device pool (shader/src/trace.frag:69-402 and
// include/hashdag/NodePoolTraversal.hpp:93-256).
//
// The march keeps trace.frag's fp32 operation order (the library is built with
// -ffp-contract=off and correctly rounded division / sqrt) so hits are bit-identical to the
// CPU transliteration. The ESVO march follows Laine & Karras, "Efficient Sparse Voxel Octrees"
// (I3D 2010), as restated in trace.frag, which carries the BSD-3 notice reproduced below; it
// applies to `march` in this file.
//
// Launch: one workgroup per pixel tile (8x8 or 16x16, from the tuning); the per-scale parent stack
// lives in LDS laid out [scale][thread] so a wavefront's stack accesses hit 64 consecutive banks.
// The march is bound by VALU issue (about 90 vector instructions per wave iteration, with the
// descend / advance / pop paths serialised under divergence), not by memory: fewer instructions
// per iteration and a shorter grid tail are what speed it up.
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

#include "hd_internal.hpp"

#include <cmath>
#include <mutex>

namespace hd {

constexpr u32 kStack = 23;  // STACK_SIZE, fraction bits of float (trace.frag:52-54)
constexpr u32 kThreads = 256; // workgroup of the beam and traversal kernels
// The frame kernel's pixel tile (workgroup) is kT x kT, kT in {8, 16}, picked by the tuning.
constexpr u32 kMaxIter = 1u << 16;

__device__ inline float fmx(float a, float b) { return a < b ? b : a; } // std::max / GLSL max order
__device__ inline float fmn(float a, float b) { return b < a ? b : a; }
__device__ inline u32 f2u(float f) { return __float_as_uint(f); }
__device__ inline float u2f(u32 u) { return __uint_as_float(u); }
// Bit i of the result = (byte i of x != 0), i < 4: bit 7 of each byte is set by the byte's top bit
// or by a carry out of its low 7 bits (no carry crosses bytes); the multiply by 2^0+2^7+2^14+2^21
// moves bits 7, 15, 23, 31 to 28..31 without overlapping partial products.
__device__ inline u32 nonzero_bytes4(u32 x) {
	const u32 t = (((x & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | x) & 0x80808080u;
	return (t * 0x00204081u) >> 28u;
}

struct MarchOut {
	bool hit;
	u32 vox[3], vox_size, iter;
	float hit_pos[3], norm[3];
	float t, scale; // t_min and scale_exp2 where the march stopped (beam.frag's outputs)
	// NodePoolTraversal result: clamp(o + t_min * d, cube) - 1 (valid when hit)
	float pick[3];
};

// DAG_RayMarch (trace.frag:69-270). stack: LDS base for this thread, stride = blockDim.
// kCount: also count the pool words the march reads (reads += ...), for the algorithmic-bytes pass.
template <bool kCount>
__device__ inline MarchOut march(const u32 *__restrict__ W, u32 root, u32 leaf_level, float proj_factor,
                                 float proj_bias, float ox, float oy, float oz, float dx, float dy, float dz,
                                 u32 *stack, u32 stride, u32 &reads) {
	MarchOut out;
	out.hit = false;
	out.iter = 0;
	out.vox_size = 0;
	out.t = 0.0f;
	out.scale = 0.0f;
	for (int i = 0; i < 3; ++i)
		out.vox[i] = 0, out.hit_pos[i] = 0.0f, out.norm[i] = 0.0f, out.pick[i] = 0.0f;
	if (root == kNull)
		return out;
	u32 iter = 0;
	ox += 1.0f, oy += 1.0f, oz += 1.0f;
	const float eps = u2f((127u - kStack) << 23u);
	dx = fabsf(dx) > eps ? dx : (dx >= 0 ? eps : -eps);
	dy = fabsf(dy) > eps ? dy : (dy >= 0 ? eps : -eps);
	dz = fabsf(dz) > eps ? dz : (dz >= 0 ? eps : -eps);
	float tcx = 1.0f / -fabsf(dx), tcy = 1.0f / -fabsf(dy), tcz = 1.0f / -fabsf(dz);
	float tbx = tcx * ox, tby = tcy * oy, tbz = tcz * oz;
	u32 oct = 0;
	if (dx > 0.0f)
		oct ^= 1u, tbx = 3.0f * tcx - tbx;
	if (dy > 0.0f)
		oct ^= 2u, tby = 3.0f * tcy - tby;
	if (dz > 0.0f)
		oct ^= 4u, tbz = 3.0f * tcz - tbz;
	float t_min = fmx(fmx(2.0f * tcx - tbx, 2.0f * tcy - tby), 2.0f * tcz - tbz);
	float t_max = fmn(fmn(tcx - tbx, tcy - tby), tcz - tbz);
	float h = t_max;
	t_min = fmx(t_min, 0.0f);
	t_max = fmn(t_max, 1.0f);
	u32 parent = root, child_bits = 0;
	float px = 1.0f, py = 1.0f, pz = 1.0f;
	u32 idx = 0;
	if (1.5f * tcx - tbx > t_min)
		idx ^= 1u, px = 1.5f;
	if (1.5f * tcy - tby > t_min)
		idx ^= 2u, py = 1.5f;
	if (1.5f * tcz - tbz > t_min)
		idx ^= 4u, pz = 1.5f;
	u32 scale = kStack - 1;
	float scale_exp2 = 0.5f;
	const u32 leaf_scale = kStack - leaf_level;
	for (;;) {
		// A valid ray takes at most a few thousand steps; the cap only guards the device against a
		// malformed pool (it is never reached by the parity cases).
		if (++iter > kMaxIter)
			return out;
		if (child_bits == 0u) {
			if (scale >= leaf_scale) { // one load path for inner nodes and leaves (lanes diverge less)
				const bool at_leaf = scale == leaf_scale;
				// word 1 exists for both (a leaf's second mask word, an inner node's first child)
				const u32 w0 = W[parent], w1 = W[parent + 1u];
				child_bits = w0;
				if (kCount)
					reads += at_leaf ? 2 : 1;
				if (at_leaf) // DAG_GetLeafFirstChildBits (trace.frag:56-67): bit i = byte i != 0
					child_bits = nonzero_bytes4(w0) | (nonzero_bytes4(w1) << 4u);
			} else
				child_bits = parent;
		}
		float tx = px * tcx - tbx, ty = py * tcy - tby, tz = pz * tcz - tbz;
		float tc_max = fmn(fmn(tx, ty), tz);
		u32 child_shift = idx ^ oct, child_mask = 1u << child_shift;
		if ((child_bits & child_mask) != 0 && t_min <= t_max) {
			float half = scale_exp2 * 0.5f;
			float cx = half * tcx + tx, cy = half * tcy + ty, cz = half * tcz + tz;
			if (scale < leaf_scale || scale_exp2 * proj_factor < tc_max + proj_bias)
				break;
			if (tc_max < h)
				stack[(scale - leaf_scale) * stride] = parent;
			h = tc_max;
			if (kCount)
				reads += 1;
			// one load for both node kinds: an inner node's child pointer, or the leaf word holding the
			// child's byte
			const bool inner = scale > leaf_scale;
			const u32 w = W[inner ? parent + 1u + (u32)__builtin_popcount(child_bits & (child_mask - 1u))
			                      : parent + (child_shift >> 2u)];
			parent = inner ? w : (w >> ((child_shift & 3u) << 3u)) & 0xFFu;
			idx = 0u;
			--scale;
			scale_exp2 = half;
			if (cx > t_min)
				idx ^= 1u, px += scale_exp2;
			if (cy > t_min)
				idx ^= 2u, py += scale_exp2;
			if (cz > t_min)
				idx ^= 4u, pz += scale_exp2;
			child_bits = 0;
			continue;
		}
		u32 step = 0;
		if (tx <= tc_max)
			step ^= 1u, px -= scale_exp2;
		if (ty <= tc_max)
			step ^= 2u, py -= scale_exp2;
		if (tz <= tc_max)
			step ^= 4u, pz -= scale_exp2;
		t_min = tc_max;
		idx ^= step;
		if ((idx & step) != 0) {
			u32 diff = 0;
			if (step & 1u)
				diff |= f2u(px) ^ f2u(px + scale_exp2);
			if (step & 2u)
				diff |= f2u(py) ^ f2u(py + scale_exp2);
			if (step & 4u)
				diff |= f2u(pz) ^ f2u(pz + scale_exp2);
			scale = diff ? 31u - (u32)__builtin_clz(diff) : 0xFFFFFFFFu; // findMSB(0) = -1
			if (scale >= kStack)
				break;
			scale_exp2 = u2f((scale - kStack + 127u) << 23u);
			parent = stack[(scale - leaf_scale) * stride];
			u32 shx = f2u(px) >> scale, shy = f2u(py) >> scale, shz = f2u(pz) >> scale;
			px = u2f(shx << scale);
			py = u2f(shy << scale);
			pz = u2f(shz << scale);
			idx = (shx & 1u) | ((shy & 1u) << 1u) | ((shz & 1u) << 2u);
			h = 0.0f;
			child_bits = 0;
		}
	}
	// normal / voxel / hit position (trace.frag:223-258)
	float ex = tcx * (px + scale_exp2) - tbx, ey = tcy * (py + scale_exp2) - tby, ez = tcz * (pz + scale_exp2) - tbz;
	float n[3];
	if (ex > ey && ex > ez)
		n[0] = -1, n[1] = 0, n[2] = 0;
	else if (ey > ez)
		n[0] = 0, n[1] = -1, n[2] = 0;
	else
		n[0] = 0, n[1] = 0, n[2] = -1;
	if ((oct & 1u) == 0u)
		n[0] = -n[0];
	if ((oct & 2u) == 0u)
		n[1] = -n[1];
	if ((oct & 4u) == 0u)
		n[2] = -n[2];
	const u32 voxel_level = leaf_level + 1u, voxel_scale = kStack - voxel_level;
	const bool ok = scale < kStack && t_min <= t_max;
	out.iter = iter;
	out.t = t_min;
	out.scale = scale_exp2;
	if (!ok) { // a miss reports zeros (trace.frag leaves them unused)
		for (int i = 0; i < 3; ++i)
			out.vox[i] = 0, out.hit_pos[i] = 0.0f, out.norm[i] = 0.0f, out.pick[i] = 0.0f;
		out.vox_size = 0;
		return out;
	}
	u32 vs = 1u << (scale - voxel_scale);
	float pp[3] = {px, py, pz}, oo[3] = {ox, oy, oz}, dd[3] = {dx, dy, dz};
	for (int i = 0; i < 3; ++i) {
		u32 v = (f2u(pp[i]) & 0x7FFFFFu) >> voxel_scale;
		if (oct & (1u << i))
			v = (1u << voxel_level) - vs - v;
		out.vox[i] = v;
		float vmin = u2f((v << voxel_scale) | (127u << 23u)), vmax = u2f(((v + vs) << voxel_scale) | (127u << 23u));
		float hp = fmn(fmx(oo[i] + t_min * dd[i], vmin), vmax);
		if (n[i] != 0)
			hp = n[i] > 0 ? vmax + eps * 2 : vmin - eps;
		out.hit_pos[i] = hp - 1.0f;
		out.norm[i] = n[i];
		// NodePoolTraversal: undo mirroring of pos, clamp(o + t d, pos, pos + scale) - 1
		float q = pp[i];
		if (oct & (1u << i))
			q = 3.0f - scale_exp2 - q;
		out.pick[i] = fmn(fmx(oo[i] + t_min * dd[i], q), q + scale_exp2) - 1.0f;
	}
	out.vox_size = vs;
	out.iter = iter;
	out.hit = ok;
	return out;
}

struct RGB {
	float r, g, b;
};
__device__ inline RGB unorm8(u32 c) {
	return {float(c & 0xFFu) / 255.0f, float((c >> 8) & 0xFFu) / 255.0f, float((c >> 16) & 0xFFu) / 255.0f};
}
__device__ inline RGB rgb565(u32 c) {
	return {float(c & 0x1Fu) / 31.0f, float((c >> 5) & 0x3Fu) / 63.0f, float((c >> 11) & 0x1Fu) / 31.0f};
}
// Color_GetLeafColor (trace.frag:304-349)
template <bool kCount>
__device__ inline RGB leaf_color(const u32 *__restrict__ L, u32 idx, u32 x, u32 y, u32 z, u32 &reads) {
	u32 macro_cnt = L[idx + 1u], block_cnt = L[idx + 2u];
	if (kCount)
		reads += 5; // counts, macro pair, next macro
	u32 macro_offset = idx + 4u, block_offset = macro_offset + (macro_cnt << 1u);
	u32 weight_offset = block_offset + (block_cnt << 1u);
	u32 vox_id = morton3(x, y, z);
	u32 macro_id = vox_id >> 14u;
	if (macro_id >= macro_cnt)
		return {0, 0, 0};
	u32 mx = L[macro_offset + (macro_id << 1u)], my = L[(macro_offset + (macro_id << 1u)) | 1u];
	block_offset += mx << 1u;
	block_cnt = macro_id + 1u < macro_cnt ? L[macro_offset + ((macro_id + 1u) << 1u)] - mx : block_cnt - mx;
	vox_id &= 0x3FFFu;
	if (block_cnt == 0u)
		return {0, 0, 0};
	for (u32 s = 0; s <= 14u && block_cnt != 0u; ++s) {
		u32 step = block_cnt >> 1u;
		if (kCount)
			reads += 1;
		if ((L[(block_offset + (step << 1u)) | 1u] >> 18u) <= vox_id) {
			block_cnt -= (step + 1u);
			block_offset += (step + 1u) << 1u;
		} else
			block_cnt = step;
	}
	block_offset -= 2u;
	u32 bx = L[block_offset], by = L[block_offset | 1u];
	u32 bpw = (by >> 16u) & 0x3u;
	if (kCount)
		reads += 2 + (bpw ? 2 : 0);
	if (bpw == 0u)
		return unorm8(bx);
	vox_id -= (by >> 18u);
	u32 bit_id = my + (by & 0xFFFFu) + vox_id * bpw;
	u32 bo = bit_id & 31u;
	u32 w = L[weight_offset + (bit_id >> 5u)] >> bo;
	if (bo + bpw <= 32u)
		w &= (1u << bpw) - 1u;
	else
		w = w | ((L[weight_offset + (bit_id >> 5u) + 1u] & ((1u << (bo + bpw - 32u)) - 1u)) << (32u - bo));
	float a = float(w) / float((1u << bpw) - 1u);
	RGB p = rgb565(bx), q = rgb565(bx >> 16u);
	return {p.r * (1.0f - a) + q.r * a, p.g * (1.0f - a) + q.g * a, p.b * (1.0f - a) + q.b * a};
}
// Color_Fetch (trace.frag:351-364)
template <bool kCount>
__device__ inline RGB color_fetch(const u32 *__restrict__ CN, const u32 *__restrict__ CL, u32 root, u32 voxel_level,
                                  u32 leaf_level, u32 x, u32 y, u32 z, u32 &reads) {
	u32 ptr = root;
	for (u32 l = 0; l < leaf_level; ++l) {
		u32 tag = ptr >> 30u, data = ptr & 0x3FFFFFFFu;
		if (tag != 0)
			return unorm8(data);
		u32 sh = voxel_level - 1u - l;
		if (kCount)
			reads += 1;
		ptr = CN[((u64)ptr << 3u) | ((x >> sh) & 1u) | (((y >> sh) & 1u) << 1u) | (((z >> sh) & 1u) << 2u)];
	}
	u32 tag = ptr >> 30u, data = ptr & 0x3FFFFFFFu;
	if (tag == 2u) {
		u32 m = (1u << (voxel_level - leaf_level)) - 1u;
		return leaf_color<kCount>(CL, data, x & m, y & m, z & m, reads);
	}
	return unorm8(data);
}

struct TraceArgs {
	const u32 *words, *cnodes, *cleaves;
	u32 root, color_root, leaf_level, color_leaf_level;
	const hd_view *views;
	const u32 *tile_start; // prefix of tiles per view (n_views + 1)
	u32 n_views;
	u32 *rgba, *records;
	unsigned long long *counters; // [primary rays, shadow rays, march iterations, words read] (may be null)
	const u32 *roots;             // if set: (root, color root) read from the device (graph replays)
	const float *beam;            // beam images of the HD_TRACE_BEAM views (may be null)
	const u32 *beam_start;        // prefix of beam texels per view (n_views + 1)
};

// Beam prepass (beam.frag:59-223, BeamPass.cpp:9-111): one coarse ray per 8x8-pixel block through
// the block's centre on the ceil(W/8) x ceil(H/8) beam image, LOD with a block's projection factor
// (the view's / 8); the block's value is max(t - cube size, 0) where the march stopped, +inf on a
// miss. The trace starts each pixel at the smallest of the 2x2 beam texels around it, times 0.98
// (a MIN-reduction bilinear fetch, trace.frag:384-390), with that distance as the LOD bias.
constexpr u32 kBeamBlock = 8;
__device__ inline u32 beam_dim(u32 n) { return (n + kBeamBlock - 1u) / kBeamBlock; }
__device__ inline float beam_sample(const float *__restrict__ B, u32 Wb, u32 Hb, float u, float v) {
	float x = u * float(Wb) - 0.5f, y = v * float(Hb) - 0.5f;
	int x0 = (int)floorf(x), y0 = (int)floorf(y);
	auto cl = [](int a, u32 n) { return (u32)min(max(a, 0), (int)n - 1); };
	u32 xa = cl(x0, Wb), xb = cl(x0 + 1, Wb), ya = cl(y0, Hb), yb = cl(y0 + 1, Hb);
	return fminf(fminf(B[ya * Wb + xa], B[ya * Wb + xb]), fminf(B[yb * Wb + xa], B[yb * Wb + xb])) * 0.98f;
}

__device__ inline u32 q8(float x) { return (u32)(fmn(fmx(x, 0.0f), 1.0f) * 255.0f + 0.5f); }

template <bool kCount, u32 kT> __global__ void __launch_bounds__(kT *kT) k_trace(TraceArgs A) {
	constexpr u32 kTT = kT * kT;
	extern __shared__ u32 stack_lds[]; // node levels x kTT words (trace_lds_bytes)
	// view of this tile (binary search over the tile prefix); a fixed grid may run past the last tile
	u32 tile = blockIdx.x, lo = 0, hi = A.n_views;
	if (tile >= A.tile_start[A.n_views])
		return;
	const u32 root = A.roots ? A.roots[0] : A.root, color_root = A.roots ? A.roots[1] : A.color_root;
	while (hi - lo > 1) {
		u32 mid = (lo + hi) >> 1;
		if (A.tile_start[mid] <= tile)
			lo = mid;
		else
			hi = mid;
	}
	const hd_view V = A.views[lo];
	// Tiles go out in reverse raster order, bottom row first: a view's lower rows usually see the
	// near ground, whose marches are the longest, so they start first and the short sky tiles fill
	// the end of the grid (longest-first order shortens its tail).
	u32 t = A.tile_start[lo + 1] - 1u - tile;
	u32 tiles_x = (V.width + kT - 1) / kT;
	u32 px = (t % tiles_x) * kT + (threadIdx.x % kT), py = (t / tiles_x) * kT + (threadIdx.x / kT);
	if (px >= V.width || py >= V.height)
		return;
	u32 *stack = stack_lds + threadIdx.x;
	// Camera_GenRay (trace.frag:366-370), gl_FragCoord at pixel centres, origin top-left
	const float u = (float(px) + 0.5f) / float(V.width), w = (float(py) + 0.5f) / float(V.height);
	const float cx = u * 2.0f - 1.0f, cy = w * 2.0f - 1.0f;
	float dx = V.look[0] - V.side[0] * cx - V.up[0] * cy, dy = V.look[1] - V.side[1] * cx - V.up[1] * cy,
	      dz = V.look[2] - V.side[2] * cx - V.up[2] * cy;
	float len = sqrtf(dx * dx + dy * dy + dz * dz);
	dx = dx / len, dy = dy / len, dz = dz / len;
	u32 reads = 0;
	MarchOut h;
	if ((V.flags & HD_TRACE_BEAM) && A.beam) {
		const float t0 = beam_sample(A.beam + A.beam_start[lo], beam_dim(V.width), beam_dim(V.height), u, w);
		if (isinf(t0)) { // the block's coarse ray missed: so does this pixel
			h.hit = false;
			h.iter = 0, h.vox_size = 0, h.t = 0.0f, h.scale = 0.0f;
			for (int i = 0; i < 3; ++i)
				h.vox[i] = 0, h.hit_pos[i] = 0.0f, h.norm[i] = 0.0f, h.pick[i] = 0.0f;
		} else
			h = march<kCount>(A.words, root, A.leaf_level, V.proj_factor, t0, V.pos[0] + t0 * dx, V.pos[1] + t0 * dy,
			                  V.pos[2] + t0 * dz, dx, dy, dz, stack, kTT, reads);
	} else
		h = march<kCount>(A.words, root, A.leaf_level, V.proj_factor, 0.0f, V.pos[0], V.pos[1], V.pos[2], dx, dy,
		                  dz, stack, kTT, reads);
	u32 iters = h.iter;
	float lx = 4.0f, ly = 5.0f, lz = 3.0f;
	float ll = sqrtf(lx * lx + ly * ly + lz * lz);
	lx = lx / ll, ly = ly / ll, lz = lz / ll;
	bool shadow = false;
	if (h.hit && (V.flags & HD_TRACE_SHADOW)) {
		MarchOut s = march<kCount>(A.words, root, A.leaf_level, __builtin_inff(), 0.0f, h.hit_pos[0], h.hit_pos[1],
		                           h.hit_pos[2], lx, ly, lz, stack, kTT, reads);
		shadow = s.hit;
		iters += s.iter;
	}
	float r = 0, g = 0, b = 0;
	if (h.hit) {
		float nd = h.norm[0] * lx + h.norm[1] * ly + h.norm[2] * lz;
		float diffuse = shadow ? 0.5f : fmx(nd, 0.0f) * 0.5f + 0.5f;
		RGB c = color_fetch<kCount>(A.cnodes, A.cleaves, color_root, A.leaf_level + 1u, A.color_leaf_level,
		                            h.vox[0], h.vox[1], h.vox[2], reads);
		r = diffuse * c.r, g = diffuse * c.g, b = diffuse * c.b;
	}
	u64 o = V.out_offset + (u64)py * V.width + px;
	A.rgba[o] = q8(r) | (q8(g) << 8) | (q8(b) << 16) | (255u << 24);
	if (A.records && (V.flags & HD_TRACE_RECORDS)) {
		u32 *rec = A.records + o * 12u;
		rec[0] = h.vox[0], rec[1] = h.vox[1], rec[2] = h.vox[2];
		rec[3] = h.vox_size | (h.hit ? 0x80000000u : 0u);
		rec[4] = f2u(h.hit_pos[0]), rec[5] = f2u(h.hit_pos[1]), rec[6] = f2u(h.hit_pos[2]);
		rec[7] = f2u(h.norm[0]), rec[8] = f2u(h.norm[1]), rec[9] = f2u(h.norm[2]);
		rec[10] = h.iter;
		rec[11] = shadow ? 1u : 0u;
	}
	if (kCount && A.counters) {
		atomicAdd(&A.counters[0], 1ull);
		atomicAdd(&A.counters[1], h.hit && (V.flags & HD_TRACE_SHADOW) ? 1ull : 0ull);
		atomicAdd(&A.counters[2], (unsigned long long)iters);
		atomicAdd(&A.counters[3], (unsigned long long)reads);
	}
}

// One thread per beam texel of every HD_TRACE_BEAM view (views without the flag have no texels).
__global__ void __launch_bounds__(kThreads) k_beam(TraceArgs A, float *__restrict__ beam) {
	extern __shared__ u32 stack_lds[];
	const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= A.beam_start[A.n_views])
		return;
	const u32 root = A.roots ? A.roots[0] : A.root;
	u32 lo = 0, hi = A.n_views;
	while (hi - lo > 1) {
		u32 mid = (lo + hi) >> 1;
		if (A.beam_start[mid] <= i)
			lo = mid;
		else
			hi = mid;
	}
	const hd_view V = A.views[lo];
	const u32 Wb = beam_dim(V.width), Hb = beam_dim(V.height), t = i - A.beam_start[lo];
	const u32 bx = t % Wb, by = t / Wb;
	float cx = (float(bx) + 0.5f) / float(Wb), cy = (float(by) + 0.5f) / float(Hb);
	cx = cx * 2.0f - 1.0f;
	cy = cy * 2.0f - 1.0f;
	float dx = V.look[0] - V.side[0] * cx - V.up[0] * cy, dy = V.look[1] - V.side[1] * cx - V.up[1] * cy,
	      dz = V.look[2] - V.side[2] * cx - V.up[2] * cy;
	float len = sqrtf(dx * dx + dy * dy + dz * dz);
	dx = dx / len, dy = dy / len, dz = dz / len;
	u32 reads = 0;
	MarchOut h = march<false>(A.words, root, A.leaf_level, V.proj_factor * (1.0f / float(kBeamBlock)), 0.0f, V.pos[0],
	                          V.pos[1], V.pos[2], dx, dy, dz, stack_lds + threadIdx.x, kThreads, reads);
	beam[i] = h.hit ? fmaxf(h.t - h.scale, 0.0f) : __builtin_inff();
}

__global__ void __launch_bounds__(kThreads) k_traversal(const u32 *__restrict__ W, u32 root, u32 leaf_level,
                                                         const float *__restrict__ o, const float *__restrict__ d,
                                                         u64 n, float *__restrict__ out) {
	extern __shared__ u32 stack_lds[];
	u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	u32 reads = 0;
	MarchOut h = march<false>(W, root, leaf_level, __builtin_inff(), 0.0f, o[3 * i], o[3 * i + 1], o[3 * i + 2],
	                          d[3 * i], d[3 * i + 1], d[3 * i + 2], stack_lds + threadIdx.x, kThreads, reads);
	out[4 * i] = h.hit ? h.pick[0] : 0.0f;
	out[4 * i + 1] = h.hit ? h.pick[1] : 0.0f;
	out[4 * i + 2] = h.hit ? h.pick[2] : 0.0f;
	out[4 * i + 3] = h.hit ? 1.0f : 0.0f;
}

// The frame kernel at the tuned tile (tiles counted with the same tile by trace_tile_count).
// The march keeps one stack word per node level and thread in LDS (pushes and pops stay at scales
// >= the leaf scale, so the slots are scale - leaf_scale in [0, levels)); sizing the stack to the
// pool's levels rather than the 23 float scales leaves room for more resident waves.
static u32 trace_lds_bytes(u32 levels, u32 threads) { return levels * threads * 4u; }
static void launch_trace(bool count, u32 n_tiles, hipStream_t s, const TraceArgs &A) {
	if (g_tuning.trace_tile == 8) {
		const u32 lds = trace_lds_bytes(A.leaf_level, 64);
		if (count)
			k_trace<true, 8><<<n_tiles, 64, lds, s>>>(A);
		else
			k_trace<false, 8><<<n_tiles, 64, lds, s>>>(A);
	} else {
		const u32 lds = trace_lds_bytes(A.leaf_level, 256);
		if (count)
			k_trace<true, 16><<<n_tiles, 256, lds, s>>>(A);
		else
			k_trace<false, 16><<<n_tiles, 256, lds, s>>>(A);
	}
}
static u32 trace_tile_count(const hd_view &v) {
	const u32 t = g_tuning.trace_tile;
	return div_up(v.width, t) * div_up(v.height, t);
}

int trace_on_stream(Pool &p, hipStream_t s, u32 root, u32 color_root, const hd_view *d_views, const u32 *d_tiles,
                    u32 n_views, u32 n_tiles, u32 *rgba, const u32 *d_roots) {
	TraceArgs A{p.words, p.cnodes, p.cleaves, root,    color_root, p.dc.L, p.dc.color_leaf,
	            d_views, d_tiles,  n_views,   rgba,    nullptr,    nullptr, d_roots, nullptr, nullptr};
	launch_trace(false, n_tiles, s, A);
	HD_TRY(hipGetLastError());
	return HD_OK;
}

} // namespace hd

using namespace hd;

extern "C" int hd_trace(hd_pool *h, uint32_t root, uint32_t color_root, const hd_view *views, uint32_t n_views,
                        uint32_t *rgba, uint32_t *records) {
	if (!h || !views || !n_views || !rgba)
		return set_error("hd_trace: null argument"), HD_ERR_ARG;
	Pool &p = h->p;
	std::vector<u32> tiles(n_views + 1, 0), beams(n_views + 1, 0);
	for (u32 v = 0; v < n_views; ++v) {
		if (!views[v].width || !views[v].height)
			return set_error("hd_trace: empty view"), HD_ERR_ARG;
		if ((views[v].flags & HD_TRACE_RECORDS) && !records)
			return set_error("hd_trace: records requested without a buffer"), HD_ERR_ARG;
		if ((views[v].flags & HD_TRACE_BEAM) && !std::isfinite(views[v].proj_factor))
			return set_error("hd_trace: the beam prepass needs a finite proj_factor (LOD on)"), HD_ERR_ARG;
		tiles[v + 1] = tiles[v] + trace_tile_count(views[v]);
	}
	HD_RET(hd_trace_beam_sizes(views, n_views, beams.data()));
	hd_view *dv;
	u32 *dt, *db = nullptr;
	float *beam = nullptr;
	HD_RET(p.scratch.get<hd_view>("trace_views", n_views, &dv));
	HD_RET(p.scratch.get<u32>("trace_tiles", n_views + 1, &dt));
	HD_TRY(hipMemcpyAsync(dv, views, n_views * sizeof(hd_view), hipMemcpyHostToDevice, p.stream));
	HD_TRY(hipMemcpyAsync(dt, tiles.data(), (n_views + 1) * 4, hipMemcpyHostToDevice, p.stream));
	if (beams[n_views]) {
		HD_RET(p.scratch.get<u32>("trace_beam_start", n_views + 1, &db));
		HD_RET(p.scratch.get<float>("trace_beam", beams[n_views], &beam));
		HD_TRY(hipMemcpyAsync(db, beams.data(), (n_views + 1) * 4, hipMemcpyHostToDevice, p.stream));
	}
	TraceArgs A{p.words, p.cnodes, p.cleaves, root,   color_root, p.dc.L, p.dc.color_leaf,
	            dv,      dt,       n_views,   rgba, records, nullptr, nullptr, beam, db};
	if (beams[n_views])
		k_beam<<<div_up(beams[n_views], kThreads), kThreads, trace_lds_bytes(p.dc.L, kThreads), p.stream>>>(A, beam);
	launch_trace(false, tiles[n_views], p.stream, A);
	HD_TRY(hipGetLastError());
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

// Asynchronous variant for timing loops and graph capture: views and the tile prefix already on the
// device (hd_trace_prepare), nothing copied, no synchronisation.
// counters (device u64[4], zeroed by the caller) selects the instrumented variant that also counts
// rays, march iterations and the pool words read (the algorithmic bytes of the frame).
extern "C" int hd_trace_async(hd_pool *h, uint32_t root, uint32_t color_root, const hd_view *d_views,
                              const uint32_t *d_tile_start, uint32_t n_views, uint32_t n_tiles, uint32_t *rgba,
                              uint32_t *records, uint64_t *counters) {
	Pool &p = h->p;
	TraceArgs A{p.words,      p.cnodes, p.cleaves, root, color_root, p.dc.L, p.dc.color_leaf, d_views, d_tile_start,
	            n_views,      rgba,     records,   (unsigned long long *)counters, nullptr, nullptr, nullptr};
	launch_trace(counters != nullptr, n_tiles, p.stream, A);
	HD_TRY(hipGetLastError());
	return HD_OK;
}

extern "C" int hd_trace_beam_sizes(const hd_view *views, uint32_t n_views, uint32_t *beam_start) {
	beam_start[0] = 0;
	for (u32 v = 0; v < n_views; ++v)
		beam_start[v + 1] = beam_start[v] + ((views[v].flags & HD_TRACE_BEAM)
		                                         ? div_up(views[v].width, kBeamBlock) * div_up(views[v].height, kBeamBlock)
		                                         : 0u);
	return HD_OK;
}

// Asynchronous beam prepass + trace: views, tile and beam-texel prefixes already on the device.
extern "C" int hd_trace_beam_async(hd_pool *h, uint32_t root, uint32_t color_root, const hd_view *d_views,
                                   const uint32_t *d_tile_start, const uint32_t *d_beam_start, uint32_t n_views,
                                   uint32_t n_tiles, uint32_t n_beam, float *d_beam, uint32_t *rgba) {
	Pool &p = h->p;
	TraceArgs A{p.words, p.cnodes, p.cleaves, root, color_root, p.dc.L, p.dc.color_leaf, d_views, d_tile_start,
	            n_views, rgba,     nullptr,   nullptr, nullptr, d_beam, d_beam_start};
	if (n_beam)
		k_beam<<<div_up(n_beam, kThreads), kThreads, trace_lds_bytes(p.dc.L, kThreads), p.stream>>>(A, d_beam);
	launch_trace(false, n_tiles, p.stream, A);
	HD_TRY(hipGetLastError());
	return HD_OK;
}

extern "C" int hd_trace_tiles(const hd_view *views, uint32_t n_views, uint32_t *tile_start) {
	tile_start[0] = 0;
	for (u32 v = 0; v < n_views; ++v)
		tile_start[v + 1] = tile_start[v] + trace_tile_count(views[v]);
	return HD_OK;
}

extern "C" int hd_traversal(hd_pool *h, uint32_t root, const float *o, const float *d, uint64_t n, float *out) {
	Pool &p = h->p;
	if (!n)
		return HD_OK;
	k_traversal<<<div_up(n, kThreads), kThreads, trace_lds_bytes(p.dc.L, kThreads), p.stream>>>(p.words, root, p.dc.L, o,
	                                                                                         d, n, out);
	HD_TRY(hipGetLastError());
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

// ---------------------------------------------------------------- trace without allocation
// hd_trace for callers that cannot allocate device memory: the staged views, the tile and beam
// prefixes and the beam prepass image live in one file-scope device array instead of the pool's
// scratch. Layout (u32 words): views [0, 64 x 18), tile prefix [1152, 1217), beam prefix
// [1217, 1282), beam image from word 1536 (2^20 floats, 256-byte aligned). Every call on a GPU
// shares that GPU's copy: overlapping calls on one GPU race.
namespace {
constexpr u32 kNoAllocViews = 64;
constexpr u32 kNoAllocBeam = 1u << 20;
constexpr u32 kNoAllocViewWords = kNoAllocViews * (u32)(sizeof(hd_view) / 4);
constexpr u32 kNoAllocBeamOff = 1536;
constexpr u64 kNoAllocWords = kNoAllocBeamOff + (u64)kNoAllocBeam;
static_assert(sizeof(hd_view) % 8 == 0 && kNoAllocViewWords + 2 * (kNoAllocViews + 1) <= kNoAllocBeamOff &&
              kNoAllocBeamOff % 64 == 0);
} // namespace

static __device__ __align__(256) u32 scratch[kNoAllocWords];

namespace {
constexpr int kNoAllocDevices = 64;
std::mutex g_scratch_mu;
u32 *g_scratch_ptr[kNoAllocDevices] = {};

// The device address of `scratch` on the current device, looked up once per device (each GPU holds
// its own copy at its own address; the host-side name is not a device address).
int scratch_address(u32 **out) {
	int dev = 0;
	HD_TRY(hipGetDevice(&dev));
	if (dev < 0 || dev >= kNoAllocDevices)
		return set_error("hd_trace_noalloc: device %d beyond the %d supported", dev, kNoAllocDevices), HD_ERR_DEVICE;
	std::lock_guard<std::mutex> lock(g_scratch_mu);
	if (!g_scratch_ptr[dev]) {
		void *ptr = nullptr;
		HD_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
		if (!ptr)
			return set_error("hd_trace_noalloc: no device address for the scratch array"), HD_ERR_DEVICE;
		g_scratch_ptr[dev] = static_cast<u32 *>(ptr);
	}
	*out = g_scratch_ptr[dev];
	return HD_OK;
}
} // namespace

extern "C" int hd_trace_noalloc(hd_pool *h, uint32_t root, uint32_t color_root, const hd_view *views,
                                uint32_t n_views, uint32_t *rgba, uint32_t *records) {
	if (!h || !views || !n_views || !rgba)
		return set_error("hd_trace_noalloc: null argument"), HD_ERR_ARG;
	if (n_views > kNoAllocViews)
		return set_error("hd_trace_noalloc: %u views, at most %u", n_views, kNoAllocViews), HD_ERR_CAPACITY;
	Pool &p = h->p;
	u32 tiles[kNoAllocViews + 1], beams[kNoAllocViews + 1];
	tiles[0] = 0;
	u64 n_beam = 0;
	for (u32 v = 0; v < n_views; ++v) {
		if (!views[v].width || !views[v].height)
			return set_error("hd_trace_noalloc: empty view"), HD_ERR_ARG;
		if ((views[v].flags & HD_TRACE_RECORDS) && !records)
			return set_error("hd_trace_noalloc: records requested without a buffer"), HD_ERR_ARG;
		if ((views[v].flags & HD_TRACE_BEAM) && !std::isfinite(views[v].proj_factor))
			return set_error("hd_trace_noalloc: the beam prepass needs a finite proj_factor (LOD on)"), HD_ERR_ARG;
		tiles[v + 1] = tiles[v] + trace_tile_count(views[v]);
		if (views[v].flags & HD_TRACE_BEAM)
			n_beam += (u64)div_up(views[v].width, kBeamBlock) * div_up(views[v].height, kBeamBlock);
	}
	if (n_beam > kNoAllocBeam)
		return set_error("hd_trace_noalloc: %llu beam texels, at most %u", (unsigned long long)n_beam, kNoAllocBeam),
		       HD_ERR_CAPACITY;
	HD_RET(hd_trace_beam_sizes(views, n_views, beams));
	u32 *base = nullptr;
	HD_RET(scratch_address(&base));
	hd_view *dv = reinterpret_cast<hd_view *>(base);
	u32 *dt = base + kNoAllocViewWords, *db = dt + kNoAllocViews + 1;
	float *beam = reinterpret_cast<float *>(base + kNoAllocBeamOff);
	HD_TRY(hipMemcpyAsync(dv, views, n_views * sizeof(hd_view), hipMemcpyHostToDevice, p.stream));
	HD_TRY(hipMemcpyAsync(dt, tiles, (n_views + 1) * 4, hipMemcpyHostToDevice, p.stream));
	if (n_beam)
		HD_TRY(hipMemcpyAsync(db, beams, (n_views + 1) * 4, hipMemcpyHostToDevice, p.stream));
	TraceArgs A{p.words, p.cnodes, p.cleaves, root,   color_root,          p.dc.L,         p.dc.color_leaf,
	            dv,      dt,       n_views,   rgba,   records,             nullptr,        nullptr,
	            n_beam ? beam : nullptr,     n_beam ? db : nullptr};
	if (n_beam)
		k_beam<<<div_up((u32)n_beam, kThreads), kThreads, trace_lds_bytes(p.dc.L, kThreads), p.stream>>>(A, beam);
	launch_trace(false, tiles[n_views], p.stream, A);
	HD_TRY(hipGetLastError());
	// the host arrays above are read by the copies: wait before they go out of scope
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

// Event timing on the pool stream: hd_timer_begin records a start event, hd_timer_end records an
// end event, waits for it and returns the elapsed milliseconds between the two.
namespace {
hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
}
extern "C" int hd_timer_begin(hd_pool *h) {
	if (!g_ev0) {
		HD_TRY(hipEventCreate(&g_ev0));
		HD_TRY(hipEventCreate(&g_ev1));
	}
	HD_TRY(hipEventRecord(g_ev0, h->p.stream));
	return HD_OK;
}
extern "C" int hd_timer_end(hd_pool *h, float *ms) {
	HD_TRY(hipEventRecord(g_ev1, h->p.stream));
	HD_TRY(hipEventSynchronize(g_ev1));
	HD_TRY(hipEventElapsedTime(ms, g_ev0, g_ev1));
	return HD_OK;
}
