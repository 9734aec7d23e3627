// hd_host.hip -- generic functor edits on the host, composed with device edits on one pool.
//
// The reference's StatelessEditor concept (include/hashdag/Editor.hpp:36-40: EditNode(config,
// coord, node_ptr) -> EditType, EditVoxel(config, coord, voxel) -> bool) admits any C++ functor; a
// device pass can only evaluate the closed editor set of hd_edit_desc. hd_edit_host runs such an
// editor through the reference's single-threaded recursion (NodePoolBase::Edit with the
// StatelessEditorWrapper, NodePool.hpp:319-417) on the host, against a mirror of the device pool's
// used words, and uploads what changed the way DAGNodePool::Flush uploads dirty page ranges
// (src/DAGNodePool.cpp:56-85): the words appended to each bucket and its new counter, scattered
// by one kernel. Node pointers, bucket placement, page padding and the full-bucket fallback follow
// the reference, so host and device edits interleave word for word like a sequence of Edit calls.
#include "hd_internal.hpp"

#include <algorithm>
#include <unordered_map>
#include <vector>

namespace hd {
namespace {

// One bucket's appended range: destination word, source offset in the staged words, length, and
// the bucket's new counter.
struct DirtyRange {
	u64 dst;
	u32 src, len, bucket, count;
	u32 pad_[2];
};

__global__ void k_scatter_ranges(u32 *__restrict__ words, u32 *__restrict__ bucket_words,
                                 const DirtyRange *__restrict__ r, const u32 *__restrict__ staged, u32 n) {
	const u32 k = blockIdx.x;
	if (k >= n)
		return;
	const DirtyRange d = r[k];
	for (u32 i = threadIdx.x; i < d.len; i += blockDim.x)
		words[d.dst + i] = staged[d.src + i];
	if (threadIdx.x == 0)
		bucket_words[d.bucket] = d.count;
}

// The device pool's used words on the host: bucket b's words [0, counter) are live[off[b], ...);
// words appended by this edit go to `added`.
struct Mirror {
	const DevConfig &c;
	std::vector<u32> bw0, bw, live;
	std::vector<u64> off;
	std::unordered_map<u32, std::vector<u32>> added;
	u64 overflows = 0;
	explicit Mirror(const DevConfig &dc) : c(dc) {}

	u32 word(u32 bucket, u32 o) const {
		if (o < bw0[bucket])
			return live[off[bucket] + o];
		auto it = added.find(bucket);
		return it->second[o - bw0[bucket]];
	}
	u32 at(u32 ptr) const { return word(ptr >> c.wbpb(), ptr & (c.words_per_bucket() - 1u)); }
	void put(u32 bucket, u32 o, u32 v) {
		auto &a = added[bucket];
		const u32 i = o - bw0[bucket];
		if (a.size() <= i)
			a.resize(i + 1, 0u);
		a[i] = v;
	}
	// find_node (NodePool.hpp:79-132) from offset 0: page by page; inner nodes are sized by their
	// mask word, a zero word ends the page's nodes; leaves have a fixed stride of 2 words.
	u32 find(bool leaf, u32 bucket, const u32 *node, u32 nsz) const {
		const u32 n = bw[bucket], wpp = c.words_per_page();
		for (u32 page = 0; page * wpp < n; ++page) {
			const u32 start = page * wpp, span = std::min(wpp, n - start);
			for (u32 i = 0; nsz <= span - i;) {
				const u32 w0 = word(bucket, start + i);
				const u32 m = w0 & 0xFFu;
				const u32 sz = leaf ? 2u : (m ? (u32)__builtin_popcount(m) + 1u : 0u);
				if (!sz)
					break;
				if (sz == nsz) {
					bool eq = true;
					for (u32 k = 0; k < nsz && eq; ++k)
						eq = word(bucket, start + i + k) == node[k];
					if (eq)
						return (bucket << c.wbpb()) | (start + i);
				}
				i += sz;
			}
		}
		return kNull;
	}
	// upsert_node (NodePool.hpp:159-211, single-threaded): find, else append (append_node,
	// :134-157: a node never straddles a page, the page tail is zero-filled), else the fallback.
	u32 upsert(u32 level, bool leaf, const u32 *node, u32 nsz, u32 fallback) {
		const u32 h = leaf ? murmur_leaf(node[0], node[1]) : murmur_span(node, nsz);
		const u32 bucket = c.level_base[level] + (h & ((1u << c.bucket_bits[level]) - 1u));
		const u32 f = find(leaf, bucket, node, nsz);
		if (f != kNull)
			return f;
		u32 n = bw[bucket];
		if (n + nsz > c.words_per_bucket()) {
			++overflows;
			return fallback;
		}
		const u32 wpp = c.words_per_page();
		if ((n & (wpp - 1u)) + nsz > wpp) {
			const u32 next = (n / wpp + 1u) * wpp;
			for (; n < next; ++n)
				put(bucket, n, 0u);
		}
		const u32 ptr = (bucket << c.wbpb()) | n;
		for (u32 k = 0; k < nsz; ++k)
			put(bucket, n + k, node[k]);
		bw[bucket] = n + nsz;
		return ptr;
	}
};

struct HostEdit {
	Mirror &m;
	const hd_host_editor &ed;
	const std::vector<u32> &filled;
	u32 L;

	u32 classify(u32 level, u32 x, u32 y, u32 z, u32 ptr) const {
		return ed.edit_node(ed.user, level, x, y, z, ptr);
	}
	// edit_leaf (NodePool.hpp:319-343): the 64 voxels in index order (NodeCoord::GetLeafCoord)
	u32 edit_leaf(u32 ptr, u32 level, u32 x, u32 y, u32 z) {
		u32 leaf[2] = {0u, 0u};
		if (ptr != kNull)
			leaf[0] = m.at(ptr), leaf[1] = m.at(ptr + 1);
		bool changed = false;
		for (u32 i = 0; i < 64; ++i) {
			const u32 vx = (x << 2) | ((i >> 2) & 2u) | (i & 1u), vy = (y << 2) | ((i >> 3) & 2u) | ((i >> 1) & 1u),
			          vz = (z << 2) | ((i >> 4) & 2u) | ((i >> 2) & 1u);
			const bool v = (leaf[i >> 5] >> (i & 31u)) & 1u;
			const bool nv = ed.edit_voxel(ed.user, vx, vy, vz, v ? 1u : 0u) != 0u;
			changed |= nv != v;
			leaf[i >> 5] ^= (u32)(nv != v) << (i & 31u);
		}
		if (!changed)
			return ptr;
		if (!leaf[0] && !leaf[1])
			return kNull;
		return m.upsert(level, true, leaf, 2, ptr);
	}
	// edit_switch + edit_node (NodePool.hpp:345-396): children 0..7 in order
	u32 edit_node(u32 ptr, u32 level, u32 x, u32 y, u32 z) {
		if (level + 1 == L)
			return edit_leaf(ptr, level, x, y, z);
		u32 u[9];
		u[0] = 0;
		for (u32 i = 1; i < 9; ++i)
			u[i] = kNull;
		if (ptr != kNull) {
			u[0] = m.at(ptr) & 0xFFu;
			u32 k = 1;
			for (u32 i = 0; i < 8; ++i)
				if ((u[0] >> i) & 1u)
					u[1 + i] = m.at(ptr + k++);
		}
		bool changed = false;
		for (u32 i = 0; i < 8; ++i) {
			const u32 cx = (x << 1) | (i & 1u), cy = (y << 1) | ((i >> 1) & 1u), cz = (z << 1) | ((i >> 2) & 1u);
			const u32 cptr = u[1 + i];
			const u32 t = classify(level + 1, cx, cy, cz, cptr);
			u32 nptr;
			if (t == HD_EDIT_TYPE_CLEAR)
				nptr = kNull;
			else if (t == HD_EDIT_TYPE_FILL)
				nptr = filled[level + 1];
			else if (t == HD_EDIT_TYPE_PROCEED)
				nptr = edit_node(cptr, level + 1, cx, cy, cz);
			else
				nptr = cptr;
			changed |= nptr != cptr;
			u[1 + i] = nptr;
			u[0] ^= (u32)((nptr != kNull) != (cptr != kNull)) << i;
		}
		if (!changed)
			return ptr;
		if (!u[0])
			return kNull;
		u32 k = 1; // pack (NodePool.hpp:272-277)
		for (u32 i = 0; i < 8; ++i)
			if ((u[0] >> i) & 1u)
				u[k++] = u[1 + i];
		return m.upsert(level, false, u, k, ptr);
	}
};

} // namespace
} // namespace hd

using namespace hd;

extern "C" int hd_edit_host(hd_pool *h, uint32_t root, const hd_host_editor *editor, uint32_t *root_out) {
	if (!h || !editor || !editor->edit_node || !editor->edit_voxel || !root_out)
		return set_error("hd_edit_host: null argument"), HD_ERR_ARG;
	Pool &p = h->p;
	HD_RET(make_filled(p)); // NodePoolBase::Edit creates the chain first (NodePool.hpp:408)
	Mirror m(p.dc);
	m.bw0.resize(p.total_buckets);
	HD_RET(hd_pool_read_bucket_words(h, 0, p.total_buckets, m.bw0.data()));
	m.bw = m.bw0;
	m.off.resize(p.total_buckets + 1);
	m.off[0] = 0;
	for (u64 b = 0; b < p.total_buckets; ++b)
		m.off[b + 1] = m.off[b] + m.bw0[b];
	u64 n = 0;
	m.live.resize(std::max<u64>(m.off[p.total_buckets], 1));
	HD_RET(hd_pool_read_live_words(h, m.live.data(), m.live.size(), &n));
	if (n != m.off[p.total_buckets])
		return set_error("hd_edit_host: live words changed while reading"), HD_ERR_DEVICE;

	HostEdit E{m, *editor, p.filled, p.dc.L};
	// Edit (NodePool.hpp:405-417): the root's own classification, then the recursion
	const u32 t = editor->edit_node(editor->user, 0, 0, 0, 0, root);
	u32 r;
	if (t == HD_EDIT_TYPE_CLEAR)
		r = kNull;
	else if (t == HD_EDIT_TYPE_FILL)
		r = p.filled[0];
	else if (t == HD_EDIT_TYPE_PROCEED)
		r = E.edit_node(root, 0, 0, 0, 0);
	else
		r = root;

	// upload the appended words and counters of the dirty buckets (DAGNodePool::Flush's dirty ranges)
	std::vector<DirtyRange> ranges;
	std::vector<u32> staged;
	for (auto &kv : m.added) {
		const u32 b = kv.first, from = m.bw0[b], to = m.bw[b];
		if (to <= from)
			continue;
		DirtyRange d{};
		d.dst = ((u64)b << p.dc.wbpb()) + from;
		d.src = (u32)staged.size();
		d.len = to - from;
		d.bucket = b;
		d.count = to;
		staged.insert(staged.end(), kv.second.begin(), kv.second.begin() + d.len);
		ranges.push_back(d);
	}
	if (!ranges.empty()) {
		DirtyRange *dr;
		u32 *ds;
		HD_RET(p.scratch.get<DirtyRange>("host_ranges", ranges.size(), &dr));
		HD_RET(p.scratch.get<u32>("host_staged", staged.size(), &ds));
		HD_TRY(hipMemcpyAsync(dr, ranges.data(), ranges.size() * sizeof(DirtyRange), hipMemcpyHostToDevice, p.stream));
		HD_TRY(hipMemcpyAsync(ds, staged.data(), staged.size() * 4, hipMemcpyHostToDevice, p.stream));
		k_scatter_ranges<<<(u32)ranges.size(), 256, 0, p.stream>>>(p.words, p.bucket_words, dr, ds,
		                                                           (u32)ranges.size());
		HD_TRY(hipGetLastError());
	}
	if (m.overflows) { // the reference's silent drop, flagged (bit 0 of the overflow word)
		u32 meta[kMetaCount];
		HD_RET(read_meta(p, meta));
		meta[kMetaOverflow] |= 1u;
		HD_TRY(hipMemcpyAsync(p.meta + kMetaOverflow, &meta[kMetaOverflow], 4, hipMemcpyHostToDevice, p.stream));
	}
	HD_TRY(hipStreamSynchronize(p.stream));
	*root_out = r;
	return HD_OK;
}
