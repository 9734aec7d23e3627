// hd_gc.hip -- device garbage collection of the node pool (NodePoolThreadedGC, NodePoolThreadedGC.hpp:29-403).
//
// forward  (top-down, :72-103): level by level the live set is the sorted, de-duplicated set of node
//          pointers reachable from the roots and the filled chain's root. It is held as a bitmap over
//          the level's words (bit o: a live node starts at word o), so "sorted and unique" is the
//          bitmap's ascending order, and a node's index in that order (its rank) is an exclusive
//          prefix of popcounts.
// backward (bottom-up, :105-348): leaves stay in their bucket and are packed in pointer order;
//          inner nodes get their children remapped (the child's rank in the level below indexes
//          the level's new pointers), are re-hashed into a possibly new bucket and appended there in
//          pointer order, with the page-tail padding of the reference's bucket caches; every bucket
//          counter of the level is rewritten. Roots are remapped at level 0 and the filled chain
//          is looked up again (:378-381).
// Processing order equals the reference run on one worker, so the compacted pool words are
// identical to its output.
//
// Every launch has a size fixed by the pool's configuration (counts live in device memory), so the
// whole collection needs no host round trip: hd_gc queues it and waits once, and a frame graph
// captures it between its edit passes and its trace (hd_edit.hip).
#include "hd_internal.hpp"

#include <algorithm>
#include <string>

namespace hd {
namespace {

constexpr u32 kGcThreads = 256;
constexpr u32 kScanBlock = 4096; // elements per block of the rank / bucket scans
constexpr u32 kGcGrid = 2048;    // workgroups of the grid-stride kernels (bounded dispatch)
constexpr u32 kGcSortCap = 1024; // nodes a bucket can receive without overflowing (2-word nodes)

__global__ void k_gc_fill(u32 *__restrict__ p, u64 n, u32 v) {
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
		p[i] = v;
}

__device__ inline u32 low_mask(u32 k) { return k >= 32u ? ~0u : ((1u << k) - 1u); }

// rank of bit o in a bitmap whose word prefix is rank[]: set bits before o
__device__ inline u32 bit_rank(const u32 *__restrict__ bits, const u32 *__restrict__ rank, u64 o) {
	return rank[o >> 5] + (u32)__builtin_popcount(bits[o >> 5] & low_mask((u32)(o & 31u)));
}

// ---- exclusive scans over device-sized arrays (three launches; sizes fixed by the config)
struct PopcountOf {
	const u32 *a;
	__device__ u32 operator()(u64 i) const { return (u32)__builtin_popcount(a[i]); }
};
struct ValueOf {
	const u32 *a;
	__device__ u32 operator()(u64 i) const { return a[i]; }
};
// out[i] = the block-local exclusive prefix of f(i); bsum[block] = the block's total
template <typename F> __global__ void __launch_bounds__(kGcThreads) k_scan_local(F f, u64 n, u32 *__restrict__ out,
                                                                               u32 *__restrict__ bsum) {
	__shared__ u32 wsum[kGcThreads / 64];
	constexpr u32 per = kScanBlock / kGcThreads;
	const u64 first = (u64)blockIdx.x * kScanBlock + (u64)threadIdx.x * per;
	u32 v[per], s = 0;
#pragma unroll
	for (u32 k = 0; k < per; ++k) {
		v[k] = first + k < n ? f(first + k) : 0u;
		s += v[k];
	}
	const u32 lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	u32 x = s;
#pragma unroll
	for (u32 d = 1; d < 64; d <<= 1) {
		const u32 y = __shfl_up(x, d);
		if (lane >= d)
			x += y;
	}
	if (lane == 63)
		wsum[w] = x;
	__syncthreads();
	u32 pre = 0, tot = 0;
#pragma unroll
	for (u32 k = 0; k < kGcThreads / 64; ++k) {
		pre += k < w ? wsum[k] : 0u;
		tot += wsum[k];
	}
	u32 o = pre + x - s;
#pragma unroll
	for (u32 k = 0; k < per; ++k) {
		if (first + k < n)
			out[first + k] = o;
		o += v[k];
	}
	if (threadIdx.x == 0)
		bsum[blockIdx.x] = tot;
}
// one block: exclusive scan of the nb block totals in place; bsum[nb] = the grand total
__global__ void __launch_bounds__(kGcThreads) k_scan_sums(u32 *__restrict__ bsum, u32 nb) {
	__shared__ u32 wsum[kGcThreads / 64];
	const u32 lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	u32 run = 0;
	for (u32 b = 0; b < nb; b += kGcThreads) {
		const u32 i = b + threadIdx.x;
		const u32 v = i < nb ? bsum[i] : 0u;
		u32 x = v;
#pragma unroll
		for (u32 d = 1; d < 64; d <<= 1) {
			const u32 y = __shfl_up(x, d);
			if (lane >= d)
				x += y;
		}
		if (lane == 63)
			wsum[w] = x;
		__syncthreads();
		u32 pre = 0, tot = 0;
#pragma unroll
		for (u32 k = 0; k < kGcThreads / 64; ++k) {
			pre += k < w ? wsum[k] : 0u;
			tot += wsum[k];
		}
		if (i < nb)
			bsum[i] = run + pre + x - v;
		run += tot;
		__syncthreads();
	}
	if (threadIdx.x == 0)
		bsum[nb] = run;
}
// out[i] += the block's offset; out[n] = the total
__global__ void k_scan_add(u32 *__restrict__ out, u64 n, const u32 *__restrict__ bsum, u32 nb) {
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i <= n; i += (u64)gridDim.x * blockDim.x) {
		if (i < n)
			out[i] += bsum[i / kScanBlock];
		else
			out[n] = bsum[nb];
	}
}

// ---- forward
// the roots (and the filled chain's root) mark level 0
__global__ void k_gc_mark_roots(const u32 *__restrict__ roots, u32 n, const u32 *__restrict__ filled, u32 has_filled,
                                u64 base, u64 words, u32 *__restrict__ bits) {
	const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i > n)
		return;
	const u32 r = i < n ? roots[i] : (has_filled ? filled[0] : kNull);
	if (r == kNull || (u64)r < base || (u64)r - base >= words)
		return;
	const u64 o = (u64)r - base;
	atomicOr(&bits[o >> 5], 1u << (o & 31u));
}
// every live node of level l marks its children in level l + 1 (one thread per bitmap word). A
// hash-consed child can have millions of parents (the filled and uniform subtrees above all): the
// bit is read first and the atomic issued only while it is still clear, so those parents read a
// cached word instead of queueing atomics on one address. A stale read only costs a redundant
// atomic: bits go from 0 to 1 only.
__global__ void k_gc_expand(const u32 *__restrict__ words, const u32 *__restrict__ bits, u64 nw, u64 base,
                            u64 cbase, u64 cwords, u32 *cbits) {
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (u64)gridDim.x * blockDim.x)
	for (u32 m = bits[i]; m; m &= m - 1u) {
		const u64 ptr = base + (i << 5) + (u32)__builtin_ctz(m);
		const u32 h = words[ptr] & 0xFFu;
		u32 k = 1;
		for (u32 c = 0; c < 8; ++c)
			if ((h >> c) & 1u) {
				const u64 o = (u64)words[ptr + k++] - cbase;
				if (o < cwords) {
					const u32 bit = 1u << (o & 31u);
					if (!(__hip_atomic_load(&cbits[o >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit))
						atomicOr(&cbits[o >> 5], bit);
				}
			}
	}
}

// ---- backward
// leaves keep their bucket (:229-274): each bucket's live leaves, in pointer order, are packed in
// place from its start (a node never moves up, so reads stay ahead of writes); one thread per bucket
__global__ void k_gc_pack_leaves(u32 *__restrict__ words, u32 *__restrict__ bucket_words, DevConfig c, u32 level,
                                 const u32 *__restrict__ bits, const u32 *__restrict__ rank, u32 *__restrict__ newp,
                                 u32 *__restrict__ meta) {
	const u32 b = blockIdx.x * blockDim.x + threadIdx.x;
	if (b >= (1u << c.bucket_bits[level]))
		return;
	const u32 gb = c.level_base[level] + b, wpb = c.words_per_bucket();
	const u64 wbase = (u64)gb << c.wbpb(), lo = (u64)b * wpb, hi = lo + wpb;
	u32 bw = 0, r = bit_rank(bits, rank, lo);
	for (u64 w = lo >> 5; w <= (hi - 1) >> 5; ++w) {
		u32 m = bits[w];
		if ((w << 5) < lo)
			m &= ~low_mask((u32)(lo - (w << 5)));
		if (((w + 1) << 5) > hi)
			m &= low_mask((u32)(hi - (w << 5)));
		for (; m; m &= m - 1u) {
			const u64 off = (w << 5) + (u32)__builtin_ctz(m) - lo;
			const u32 leaf[2] = {words[wbase + off], words[wbase + off + 1]};
			const u32 np = append_to_bucket(words, c, gb, bw, leaf, 2);
			if (np == kNull)
				atomicOr(&meta[kMetaOverflow], 4u);
			newp[r++] = np;
		}
	}
	bucket_words[gb] = bw;
}
// an inner level's used words copied out (the pack writes the pool's words in place)
__global__ void k_gc_stage(const u32 *__restrict__ words, const u32 *__restrict__ bucket_words, DevConfig c,
                           u32 level, u32 *__restrict__ stage) {
	const u32 lane = threadIdx.x & 63u, nb = 1u << c.bucket_bits[level];
	for (u32 b = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6); b < nb; b += gridDim.x * (blockDim.x / 64)) {
		const u32 gb = c.level_base[level] + b, n = min(bucket_words[gb], c.words_per_bucket());
		const u64 src = (u64)gb << c.wbpb(), dst = (u64)b << c.wbpb();
		for (u32 i = lane; i < n; i += 64)
			stage[dst + i] = words[src + i];
	}
}
// a live inner node's content with its children remapped (stage: the level's words, level-relative)
struct ChildMap {
	const u32 *bits, *rank, *newp;
	u64 cbase, cwords;
};
__device__ inline u32 remapped(const u32 *__restrict__ stage, u64 o, const ChildMap &M, u32 (&buf)[9]) {
	const u32 h = stage[o] & 0xFFu, k = (u32)__builtin_popcount(h);
	buf[0] = stage[o];
#pragma unroll
	for (u32 j = 1; j < 9; ++j)
		if (j <= k) {
			const u64 co = (u64)stage[o + j] - M.cbase;
			buf[j] = co < M.cwords ? M.newp[bit_rank(M.bits, M.rank, co)] : kNull;
		}
	return k + 1u;
}
// the new bucket of every live node (re-hash, :147-148) and the buckets' node counts
__global__ void k_gc_key(const u32 *__restrict__ stage, const u32 *__restrict__ bits, const u32 *__restrict__ rank,
                         u64 nw, ChildMap M, DevConfig c, u32 level, u32 *__restrict__ key, u32 *__restrict__ cnt) {
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (u64)gridDim.x * blockDim.x) {
	u32 m = bits[i];
	if (!m)
		continue;
	u32 r = rank[i];
	for (; m; m &= m - 1u, ++r) {
		u32 buf[9];
		const u32 n = remapped(stage, (i << 5) + (u32)__builtin_ctz(m), M, buf);
		const u32 kb = murmur_span(buf, n) & ((1u << c.bucket_bits[level]) - 1u);
		key[r] = kb;
		atomicAdd(&cnt[kb], 1u);
	}
	}
}
// each live node's old offset into its new bucket's list (unordered within the bucket)
__global__ void k_gc_scatter(const u32 *__restrict__ bits, const u32 *__restrict__ rank, u64 nw,
                             const u32 *__restrict__ key, const u32 *__restrict__ bstart, u32 *__restrict__ fill,
                             u32 *__restrict__ list) {
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += (u64)gridDim.x * blockDim.x) {
		u32 m = bits[i];
		if (!m)
			continue;
		u32 r = rank[i];
		for (; m; m &= m - 1u, ++r) {
			const u32 kb = key[r];
			list[bstart[kb] + atomicAdd(&fill[kb], 1u)] = (u32)((i << 5) + (u32)__builtin_ctz(m));
		}
	}
}
// one bucket of k_gc_pack_inner (one wavefront; so / sp: its LDS lists)
__device__ inline void gc_pack_bucket(u32 *__restrict__ words, u32 *__restrict__ bucket_words, const DevConfig &c,
                                      u32 level, const u32 *__restrict__ stage, const ChildMap &M,
                                      const u32 *__restrict__ bits, const u32 *__restrict__ rank,
                                      const u32 *__restrict__ cnt, const u32 *__restrict__ bstart,
                                      const u32 *__restrict__ list, u32 *__restrict__ newp, u32 *__restrict__ meta,
                                      u32 b, u32 lane, u32 *so, u32 *sp) {
	const u32 gb = c.level_base[level] + b;
	u32 m = cnt[b];
	const u32 s0 = bstart[b];
	if (m > kGcSortCap) { // more 2-word nodes than the bucket holds: the reference's cache overflows
		if (lane == 0)
			atomicOr(&meta[kMetaOverflow], 4u);
		for (u32 k = kGcSortCap + lane; k < m; k += 64)
			newp[bit_rank(bits, rank, list[s0 + k])] = kNull;
		m = kGcSortCap;
	}
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier(); // the previous bucket's readers of so / sp are done
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	u32 P = 1;
	while (P < m)
		P <<= 1;
	for (u32 k = lane; k < P; k += 64)
		so[k] = k < m ? list[s0 + k] : ~0u;
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	for (u32 size = 2; size <= P; size <<= 1) // bitonic sort, ascending
		for (u32 stride = size >> 1; stride > 0; stride >>= 1) {
			for (u32 k = lane; k < P; k += 64) {
				const u32 q = k ^ stride;
				if (q > k) {
					const bool up = (k & size) == 0;
					const u32 a = so[k], d = so[q];
					if ((a > d) == up) {
						so[k] = d;
						so[q] = a;
					}
				}
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		}
	for (u32 k = lane; k < m; k += 64)
		sp[k] = (u32)__builtin_popcount(stage[so[k]] & 0xFFu) + 1u; // node sizes
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	const u32 wpp = c.words_per_page(), wpb = c.words_per_bucket();
	const u64 wbase = (u64)gb << c.wbpb();
	u32 bw = 0;
	if (lane == 0) { // positions in pointer order; a page tail a node skips is zero (the cache's resize)
		for (u32 k = 0; k < m; ++k) {
			const u32 sz = sp[k];
			u32 at = bw;
			if ((at & (wpp - 1u)) + sz > wpp)
				at = (at / wpp + 1u) * wpp;
			if (at + sz > wpb) { // full: the node falls back (flagged), later ones may still fit
				sp[k] = ~0u;
				atomicOr(&meta[kMetaOverflow], 4u);
				continue;
			}
			for (u32 z = bw; z < at; ++z)
				words[wbase + z] = 0u;
			sp[k] = at;
			bw = at + sz;
		}
	}
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	bw = __builtin_amdgcn_readfirstlane(bw);
	for (u32 k = lane; k < m; k += 64) {
		const u32 o = so[k], at = sp[k];
		u32 buf[9];
		const u32 n = remapped(stage, o, M, buf);
		const u32 r = bit_rank(bits, rank, o);
		if (at == ~0u) {
			newp[r] = kNull;
			continue;
		}
#pragma unroll
		for (u32 j = 0; j < 9; ++j)
			if (j < n)
				words[wbase + at + j] = buf[j];
		newp[r] = (u32)(wbase + at);
	}
	if (lane == 0)
		bucket_words[gb] = bw;
}
// one wavefront per new bucket: its nodes sorted by old pointer (bitonic sort in LDS), placed one
// after another with page-tail padding (lf_gc_shrink_inner_bucket's cache, :105-182, and the flush,
// :198-227), then written in parallel; the bucket's counter is rewritten
__global__ void __launch_bounds__(kGcThreads) k_gc_pack_inner(u32 *__restrict__ words, u32 *__restrict__ bucket_words,
                                                              DevConfig c, u32 level, const u32 *__restrict__ stage,
                                                              ChildMap M, const u32 *__restrict__ bits,
                                                              const u32 *__restrict__ rank,
                                                              const u32 *__restrict__ cnt,
                                                              const u32 *__restrict__ bstart,
                                                              const u32 *__restrict__ list, u32 *__restrict__ newp,
                                                              u32 *__restrict__ meta) {
	__shared__ u32 s_off[kGcThreads / 64][kGcSortCap];
	__shared__ u32 s_pos[kGcThreads / 64][kGcSortCap];
	const u32 lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	u32 *so = s_off[w], *sp = s_pos[w];
	for (u32 b = blockIdx.x * (kGcThreads / 64) + w; b < (1u << c.bucket_bits[level]);
	     b += gridDim.x * (kGcThreads / 64))
		gc_pack_bucket(words, bucket_words, c, level, stage, M, bits, rank, cnt, bstart, list, newp, meta, b, lane, so,
		               sp);
}
// roots remapped through level 0's new pointers
__global__ void k_gc_roots(u32 *__restrict__ roots, u32 n, const u32 *__restrict__ bits, const u32 *__restrict__ rank,
                           const u32 *__restrict__ newp, u64 base, u64 words) {
	const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n || roots[i] == kNull || (u64)roots[i] < base || (u64)roots[i] - base >= words)
		return;
	roots[i] = newp[bit_rank(bits, rank, (u64)roots[i] - base)];
}

inline u32 grid(u64 n, u32 t = kGcThreads) { return (u32)std::max<u64>(1, (n + t - 1) / t); }

} // namespace

// Device buffers of a collection, sized by the pool's configuration (reused across calls).
int GcPlan::prepare(Pool &p, const char *prefix) {
	const u32 L = p.dc.L, wpb = p.dc.words_per_bucket();
	lv.assign(L, Level{});
	u64 max_words = 0, max_buckets = 0, max_scan = 0;
	std::string pre(prefix);
	for (u32 l = 0; l < L; ++l) {
		Level &V = lv[l];
		V.buckets = 1ull << p.dc.bucket_bits[l];
		V.words = V.buckets * wpb;
		V.base = (u64)p.dc.level_base[l] << p.dc.wbpb();
		V.nw = (V.words + 31) / 32;
		HD_RET(p.scratch.get<u32>(pre + "gc_bits" + std::to_string(l), V.nw + 1, &V.bits));
		HD_RET(p.scratch.get<u32>(pre + "gc_rank" + std::to_string(l), V.nw + 1, &V.rank));
		max_words = std::max(max_words, V.words);
		max_buckets = std::max(max_buckets, V.buckets);
		max_scan = std::max(max_scan, std::max(V.nw, V.buckets));
	}
	const u64 max_nodes = max_words / 2;
	HD_RET(p.scratch.get<u32>(pre + "gc_newp0", max_nodes + 1, &newp[0]));
	HD_RET(p.scratch.get<u32>(pre + "gc_newp1", max_nodes + 1, &newp[1]));
	HD_RET(p.scratch.get<u32>(pre + "gc_stage", max_words + 16, &stage));
	HD_RET(p.scratch.get<u32>(pre + "gc_key", max_nodes + 1, &key));
	HD_RET(p.scratch.get<u32>(pre + "gc_list", max_nodes + 1, &list));
	HD_RET(p.scratch.get<u32>(pre + "gc_cnt", max_buckets + 1, &cnt));
	HD_RET(p.scratch.get<u32>(pre + "gc_bstart", max_buckets + 1, &bstart));
	HD_RET(p.scratch.get<u32>(pre + "gc_fill", max_buckets + 1, &fill));
	HD_RET(p.scratch.get<u32>(pre + "gc_bsum", max_scan / kScanBlock + 2, &bsum));
	return HD_OK;
}

template <typename F> static void scan(hipStream_t s, F f, u64 n, u32 *out, u32 *bsum) {
	const u32 nb = (u32)((n + kScanBlock - 1) / kScanBlock);
	k_scan_local<<<std::max(nb, 1u), kGcThreads, 0, s>>>(f, n, out, bsum);
	k_scan_sums<<<1, kGcThreads, 0, s>>>(bsum, std::max(nb, 1u));
	k_scan_add<<<std::min(grid(n + 1), kGcGrid), kGcThreads, 0, s>>>(out, n, bsum, std::max(nb, 1u));
}

// Queues a whole collection on stream s: roots (device, n_roots) are remapped in place, the filled
// chain (when the pool has one) is a root and is found again afterwards in d_filled. No host
// synchronisation and no allocation (capture-safe once the plan is prepared).
int gc_queue(Pool &p, const GcPlan &G, hipStream_t s, u32 *d_roots, u32 n_roots, bool has_filled) {
	const u32 L = p.dc.L;
	for (u32 l = 0; l < L; ++l)
		k_gc_fill<<<std::min(grid(G.lv[l].nw + 1), kGcGrid), kGcThreads, 0, s>>>(G.lv[l].bits, G.lv[l].nw + 1, 0u);
	k_gc_mark_roots<<<grid(n_roots + 1), kGcThreads, 0, s>>>(d_roots, n_roots, p.d_filled, has_filled ? 1u : 0u,
	                                                          G.lv[0].base, G.lv[0].words, G.lv[0].bits);
	for (u32 l = 0; l < L; ++l) {
		const GcPlan::Level &V = G.lv[l];
		scan(s, PopcountOf{V.bits}, V.nw, V.rank, G.bsum);
		if (l + 1 < L)
			k_gc_expand<<<std::min(grid(V.nw), kGcGrid), kGcThreads, 0, s>>>(p.words, V.bits, V.nw, V.base, G.lv[l + 1].base,
			                                              G.lv[l + 1].words, G.lv[l + 1].bits);
	}
	for (int l = (int)L - 1; l >= 0; --l) {
		const GcPlan::Level &V = G.lv[l];
		u32 *newp = G.newp[l & 1];
		if (l == (int)L - 1) {
			k_gc_pack_leaves<<<grid(V.buckets), kGcThreads, 0, s>>>(p.words, p.bucket_words, p.dc, (u32)l, V.bits,
			                                                       V.rank, newp, p.meta);
			continue;
		}
		const GcPlan::Level &C = G.lv[l + 1];
		const ChildMap M{C.bits, C.rank, G.newp[(l + 1) & 1], C.base, C.words};
		k_gc_stage<<<std::min(grid(V.buckets * 64), kGcGrid), kGcThreads, 0, s>>>(p.words, p.bucket_words, p.dc, (u32)l, G.stage);
		k_gc_fill<<<std::min(grid(V.buckets), kGcGrid), kGcThreads, 0, s>>>(G.cnt, V.buckets, 0u);
		k_gc_fill<<<std::min(grid(V.buckets), kGcGrid), kGcThreads, 0, s>>>(G.fill, V.buckets, 0u);
		k_gc_key<<<std::min(grid(V.nw), kGcGrid), kGcThreads, 0, s>>>(G.stage, V.bits, V.rank, V.nw, M, p.dc, (u32)l, G.key, G.cnt);
		scan(s, ValueOf{G.cnt}, V.buckets, G.bstart, G.bsum);
		k_gc_scatter<<<std::min(grid(V.nw), kGcGrid), kGcThreads, 0, s>>>(V.bits, V.rank, V.nw, G.key, G.bstart, G.fill, G.list);
		k_gc_pack_inner<<<std::min(grid(V.buckets * 64), kGcGrid), kGcThreads, 0, s>>>(p.words, p.bucket_words, p.dc, (u32)l, G.stage, M,
		                                                           V.bits, V.rank, G.cnt, G.bstart, G.list, newp,
		                                                           p.meta);
	}
	if (n_roots)
		k_gc_roots<<<grid(n_roots), kGcThreads, 0, s>>>(d_roots, n_roots, G.lv[0].bits, G.lv[0].rank, G.newp[0],
		                                                G.lv[0].base, G.lv[0].words);
	if (has_filled) // re-initialise the filled chain (:378-381): its nodes are live, so this finds them
		launch_make_filled(p, s);
	HD_TRY(hipGetLastError());
	return HD_OK;
}

} // namespace hd

using namespace hd;

extern "C" int hd_gc(hd_pool *h, uint32_t *roots, uint32_t n) {
	if (!h || (!roots && n))
		return set_error("hd_gc: null argument"), HD_ERR_ARG;
	Pool &p = h->p;
	HD_RET(p.gc_plan.prepare(p, ""));
	u32 *droots;
	HD_RET(p.scratch.get<u32>("gc_roots", n ? n : 1, &droots));
	if (n)
		HD_TRY(hipMemcpyAsync(droots, roots, n * 4, hipMemcpyHostToDevice, p.stream));
	const bool has_filled = !p.filled.empty();
	HD_RET(gc_queue(p, p.gc_plan, p.stream, droots, n, has_filled));
	if (n)
		HD_TRY(hipMemcpyAsync(roots, droots, n * 4, hipMemcpyDeviceToHost, p.stream));
	if (has_filled)
		HD_TRY(hipMemcpyAsync(p.filled.data(), p.d_filled, p.dc.L * 4, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}
