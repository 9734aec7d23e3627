// hd_gc.hip -- device garbage collection of the node pool (NodePoolThreadedGC, NodePoolThreadedGC.hpp:29-403).
//
// forward  (top-down, :72-103): level by level the live set is the sorted, de-duplicated list of
//          node pointers reachable from the roots and the filled chain's root;
// backward (bottom-up, :229-348): leaves stay in their bucket and are packed in pointer order;
//          inner nodes get their children remapped (binary search in the child level's live list),
//          are re-hashed into a possibly new bucket and appended there in pointer order (stable
//          sort by new bucket), with the page-tail padding of the reference's bucket caches;
//          every bucket counter of the level is rewritten. Roots are remapped at level 0 and the
//          filled chain is looked up again (:378-381).
// Processing order equals the reference run on one worker, so the compacted pool words are
// identical to its output.
#include "hd_internal.hpp"

#include <rocprim/rocprim.hpp>

namespace hd {

// children of every live inner node, flattened (count pass then write pass)
__global__ void k_gc_child_count(const u32 *__restrict__ words, const u32 *__restrict__ live, u32 n,
                                 u32 *__restrict__ cnt) {
	u32 i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i > n)
		return;
	cnt[i] = i < n ? (u32)__builtin_popcount(words[live[i]] & 0xFFu) : 0u;
}
__global__ void k_gc_child_write(const u32 *__restrict__ words, const u32 *__restrict__ live, u32 n,
                                 const u32 *__restrict__ off, u32 *__restrict__ out) {
	u32 i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	u32 p = live[i], k = (u32)__builtin_popcount(words[p] & 0xFFu), o = off[i];
	for (u32 j = 0; j < k; ++j)
		out[o + j] = words[p + 1u + j];
}

__device__ inline u32 lower_bound(const u32 *a, u32 n, u32 x) {
	u32 lo = 0, cnt = n;
	while (cnt > 0) {
		u32 s = cnt >> 1;
		if (a[lo + s] < x)
			lo += s + 1, cnt -= s + 1;
		else
			cnt = s;
	}
	return lo;
}

// gather the content of each live node; inner nodes get their children remapped and a new bucket
__global__ void k_gc_gather(const u32 *__restrict__ words, const u32 *__restrict__ live, u32 n, bool leaf,
                            const u32 *__restrict__ child_live, u32 child_n, const u32 *__restrict__ child_new,
                            DevConfig c, u32 level, u32 *__restrict__ content, u32 *__restrict__ csize,
                            u32 *__restrict__ key) {
	u32 i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n)
		return;
	u32 p = live[i];
	if (leaf) {
		content[2 * (u64)i] = words[p];
		content[2 * (u64)i + 1] = words[p + 1u];
		key[i] = (p >> c.wbpb()) - c.level_base[level]; // leaves keep their bucket (:237)
		return;
	}
	u32 buf[9];
	u32 m = words[p] & 0xFFu, k = (u32)__builtin_popcount(m);
	buf[0] = words[p];
	for (u32 j = 0; j < k; ++j)
		buf[1 + j] = child_new[lower_bound(child_live, child_n, words[p + 1u + j])];
	for (u32 j = 0; j <= k; ++j)
		content[9 * (u64)i + j] = buf[j];
	csize[i] = k + 1u;
	key[i] = murmur_span(buf, k + 1u) & ((1u << c.bucket_bits[level]) - 1u); // re-hash (:147-148)
}

// one thread per bucket segment of the (stable) sorted list: pack the nodes in pointer order
__global__ void k_gc_pack(u32 *__restrict__ words, u32 *__restrict__ bucket_words, DevConfig c, u32 level, u32 m,
                          const u32 *__restrict__ keys, const u32 *__restrict__ idx, bool leaf,
                          const u32 *__restrict__ content, const u32 *__restrict__ csize, u32 *__restrict__ new_ptr,
                          u32 *__restrict__ meta) {
	u32 j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= m || (j > 0 && keys[j] == keys[j - 1]))
		return;
	u32 key = keys[j], b = key + c.level_base[level];
	u32 bw = 0;
	for (u32 k = j; k < m && keys[k] == key; ++k) {
		u32 i = idx[k];
		u32 nsz = leaf ? 2u : csize[i];
		u32 r = append_to_bucket(words, c, b, bw, content + (u64)i * (leaf ? 2u : 9u), nsz);
		if (r == kNull)
			atomicOr(&meta[kMetaOverflow], 4u);
		new_ptr[i] = r;
	}
	bucket_words[b] = bw;
}

__global__ void k_gc_roots(u32 *roots, u32 n, const u32 *__restrict__ live, u32 ln, const u32 *__restrict__ new_ptr) {
	u32 i = blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= n || roots[i] == kNull)
		return;
	roots[i] = new_ptr[lower_bound(live, ln, roots[i])];
}

static int sort_unique(Pool &p, u32 *in, u32 n, u32 *tmp_keys, u32 *out, u32 *out_n) {
	size_t tb = 0;
	HD_TRY(rocprim::radix_sort_keys(nullptr, tb, in, tmp_keys, n, 0, 32, p.stream));
	size_t tb2 = 0;
	u32 *cnt;
	HD_RET(p.scratch.get<u32>("gc_ucnt", 1, &cnt));
	HD_TRY(rocprim::unique(nullptr, tb2, tmp_keys, out, cnt, n, rocprim::equal_to<u32>(), p.stream));
	void *t;
	HD_RET(p.scratch.get<char>("gc_tmp", std::max(tb, tb2) + 16, (char **)&t));
	HD_TRY(rocprim::radix_sort_keys(t, tb, in, tmp_keys, n, 0, 32, p.stream));
	HD_TRY(rocprim::unique(t, tb2, tmp_keys, out, cnt, n, rocprim::equal_to<u32>(), p.stream));
	HD_TRY(hipMemcpyAsync(out_n, cnt, 4, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

static int gc(Pool &p, u32 *roots, u32 n_roots) {
	const u32 L = p.dc.L;
	// live set of level 0: the filled chain's root and the given roots
	std::vector<u32> r0;
	if (!p.filled.empty() && p.filled[0] != kNull)
		r0.push_back(p.filled[0]);
	for (u32 i = 0; i < n_roots; ++i)
		if (roots[i] != kNull)
			r0.push_back(roots[i]);
	std::vector<u32 *> live(L, nullptr);
	std::vector<u32> nlive(L, 0);
	char name[32];
	auto lbuf = [&](const char *s, u32 l, size_t cnt, u32 **out) {
		snprintf(name, sizeof(name), "%s%u", s, l);
		return p.scratch.get<u32>(name, cnt ? cnt : 1, out);
	};
	u32 *cand, *keys;
	u32 nc = (u32)r0.size();
	HD_RET(p.scratch.get<u32>("gc_cand", nc ? nc : 1, &cand));
	if (nc)
		HD_TRY(hipMemcpyAsync(cand, r0.data(), nc * 4, hipMemcpyHostToDevice, p.stream));
	// ---- forward
	for (u32 l = 0; l < L; ++l) {
		HD_RET(lbuf("gc_live", l, nc, &live[l]));
		HD_RET(p.scratch.get<u32>("gc_keys", nc ? nc : 1, &keys));
		if (nc)
			HD_RET(sort_unique(p, cand, nc, keys, live[l], &nlive[l]));
		else
			nlive[l] = 0;
		if (l + 1 == L)
			break;
		u32 n = nlive[l];
		u32 *cnt, *off;
		HD_RET(p.scratch.get<u32>("gc_cnt", n + 1, &cnt));
		HD_RET(p.scratch.get<u32>("gc_off", n + 1, &off));
		k_gc_child_count<<<div_up(n + 1, 256), 256, 0, p.stream>>>(p.words, live[l], n, cnt);
		size_t tb = 0;
		HD_TRY(rocprim::exclusive_scan(nullptr, tb, cnt, off, 0u, (size_t)n + 1, rocprim::plus<u32>(), p.stream));
		void *t;
		HD_RET(p.scratch.get<char>("gc_tmp", tb + 16, (char **)&t));
		HD_TRY(rocprim::exclusive_scan(t, tb, cnt, off, 0u, (size_t)n + 1, rocprim::plus<u32>(), p.stream));
		HD_TRY(hipMemcpyAsync(&nc, off + n, 4, hipMemcpyDeviceToHost, p.stream));
		HD_TRY(hipStreamSynchronize(p.stream));
		HD_RET(p.scratch.get<u32>("gc_cand", nc ? nc : 1, &cand));
		if (n)
			k_gc_child_write<<<div_up(n, 256), 256, 0, p.stream>>>(p.words, live[l], n, off, cand);
		HD_TRY(hipGetLastError());
	}
	// ---- backward
	u32 *child_new = nullptr;
	for (int l = (int)L - 1; l >= 0; --l) {
		const bool leaf = l == (int)L - 1;
		u32 n = nlive[l];
		u32 *content, *csize, *key, *idx, *keys2, *idx2, *new_ptr;
		HD_RET(lbuf("gc_content", 0, (size_t)n * (leaf ? 2 : 9), &content));
		HD_RET(lbuf("gc_csize", 0, n, &csize));
		HD_RET(lbuf("gc_key", 0, n, &key));
		HD_RET(lbuf("gc_idx", 0, n, &idx));
		HD_RET(lbuf("gc_key2", 0, n, &keys2));
		HD_RET(lbuf("gc_idx2", 0, n, &idx2));
		HD_RET(lbuf("gc_new", (u32)l, n, &new_ptr));
		if (n) {
			k_gc_gather<<<div_up(n, 256), 256, 0, p.stream>>>(p.words, live[l], n, leaf, leaf ? nullptr : live[l + 1],
			                                                  leaf ? 0u : nlive[l + 1], child_new, p.dc, (u32)l,
			                                                  content, csize, key);
			// index i (pointer order) as the sort value keeps each bucket's nodes in pointer order
			rocprim::transform(rocprim::counting_iterator<u32>(0), idx, n, [] __device__(u32 x) { return x; },
			                   p.stream);
			size_t tb = 0;
			u32 bits = std::max(1u, p.dc.bucket_bits[l]);
			HD_TRY(rocprim::radix_sort_pairs(nullptr, tb, key, keys2, idx, idx2, n, 0, bits, p.stream));
			void *t;
			HD_RET(p.scratch.get<char>("gc_tmp", tb + 16, (char **)&t));
			HD_TRY(rocprim::radix_sort_pairs(t, tb, key, keys2, idx, idx2, n, 0, bits, p.stream));
		}
		// every bucket of the level restarts empty (lf_gc_flush_inner_bucket / shrink_leaf_bucket)
		HD_TRY(hipMemsetAsync(p.bucket_words + p.dc.level_base[l], 0, (size_t)4 << p.dc.bucket_bits[l], p.stream));
		if (n)
			k_gc_pack<<<div_up(n, 256), 256, 0, p.stream>>>(p.words, p.bucket_words, p.dc, (u32)l, n, keys2, idx2, leaf,
			                                                content, csize, new_ptr, p.meta);
		HD_TRY(hipGetLastError());
		child_new = new_ptr;
	}
	// ---- roots
	u32 *droots;
	HD_RET(p.scratch.get<u32>("gc_roots", n_roots ? n_roots : 1, &droots));
	if (n_roots) {
		HD_TRY(hipMemcpyAsync(droots, roots, n_roots * 4, hipMemcpyHostToDevice, p.stream));
		k_gc_roots<<<div_up(n_roots, 256), 256, 0, p.stream>>>(droots, n_roots, live[0], nlive[0], child_new);
		HD_TRY(hipMemcpyAsync(roots, droots, n_roots * 4, hipMemcpyDeviceToHost, p.stream));
	}
	HD_TRY(hipStreamSynchronize(p.stream));
	if (!p.filled.empty()) { // re-initialise the filled chain (:378-381)
		p.filled.clear();
		HD_RET(make_filled(p));
	}
	return HD_OK;
}

} // namespace hd

using namespace hd;

extern "C" int hd_gc(hd_pool *h, uint32_t *roots, uint32_t n) {
	if (!h || (!roots && n))
		return set_error("hd_gc: null argument"), HD_ERR_ARG;
	return gc(h->p, roots, n);
}
