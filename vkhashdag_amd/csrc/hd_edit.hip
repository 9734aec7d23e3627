// hd_edit.hip -- level-synchronous device edit pipeline.
//
// Replaces the CPU recursion of NodePoolBase::Edit / NodePoolThreadedEdit::ThreadedEdit
// (NodePool.hpp:319-417, NodePoolThreadedEdit.hpp:27-126) and the VBREditorWrapper color path
// (VBREditor.hpp:26-107, VBRColor.hpp:368-487, DAGColorPool.hpp:139-204):
//
//   top-down   k_root, then per level k_expand (one lane per (node, child): classify with the
//              editor, fill / clear / keep / proceed) + stable compaction (k_emit) so every level's
//              frontier is in the reference's DFS order;
//   leaves     k_leaf: one wavefront per 4^3 leaf, one lane per voxel, __ballot builds the mask;
//   bottom-up  k_build packs changed inner nodes; k_dedup / k_find / stable sort by bucket /
//              k_append / k_resolve give find-or-insert whose appends land in each bucket in DFS
//              order -- the pool words equal the reference's single-threaded Edit;
//   colors     every color leaf a proceeding edit enters is decoded to a dense per-voxel array,
//              Push runs and per-voxel edits are applied, and the chunk is re-encoded with a
//              block-wide scan (the VBR encoding is canonical for a given color sequence);
//              SetLeaf appends in DFS order, SetNode appends in post-order (as the recursion).
#include "hd_internal.hpp"

#include <rocprim/rocprim.hpp>

#include <chrono>
#include <cmath>

namespace hd {

struct Entry {
	u32 x, y, z;  // node coordinate at this level
	u32 ptr;      // old node pointer (kNull if none)
	u32 parent;   // index of the parent entry
	u32 sf;       // bits 0-2 child slot, 3 is_final, 4 has a color writer, 6 color-active
	u32 octree;   // VBREditorWrapper::NodeState::octree_node
	u32 job;      // color leaf job (entry index at the color leaf level) or kNull
	u32 loff;     // edit list: the edit index when llen == 1, else offset into the level's list array
	u32 llen;
};
constexpr u32 kSfFinal = 8u, kSfWriter = 16u, kSfNewJob = 32u, kSfActive = 64u;

struct PushEvent {
	u32 job, offset, count, colors, wb, pad0, pad1, pad2;
};

// One pass applies a group of edits whose voxel AABBs are pairwise disjoint (so they commute):
// a node reached by one edit runs that edit's wrapper logic exactly as NodePoolBase::Edit would;
// a node reached by several proceeds with its ordered edit list and no color decision.
struct EditParams {
	DevConfig c;
	const DevHeightmap *hm;
	const hd_edit_desc *edits; // the pass's edits (device)
	u32 n_edits;
	u32 *words, *bucket_words;
	const u32 *filled;
	u32 *cnodes, *cleaves, *meta;
	u32 stateful; // some edit of the pass runs through VBREditorWrapper
	u32 octree_root;
	u32 leaf_bits; // voxel_level - color_leaf_level
	__device__ DevEditor editor(u32 k) const {
		DevEditor e;
		e.d = edits[k];
		e.hm = hm;
		e.voxel_level = c.voxel_level;
		return e;
	}
};

__device__ inline u32 color_child(const u32 *cnodes, u32 p, u32 i) { // DAGColorPool::GetChild (:139-143)
	u32 t = ctag(p);
	return t == kTagNode ? cnodes[(u64)cdata(p) * 8u + i] : (t == kTagColor ? p : kColorNull);
}

struct Classified {
	EditType t;
	u32 octree, sf;
	bool push;
	Color col;
};

// edit_switch + VBREditorWrapper::EditNode (NodePool.hpp:345-360, VBREditor.hpp:37-71) for one
// edit on the node (cl, cx, cy, cz). octree: the node's state octree (GetChild already applied
// for cl <= color leaf level); sf: slot | inherited is_final / writer bits.
__device__ inline Classified classify_single(const EditParams &P, const DevEditor &ed, u32 cl, u32 cx, u32 cy, u32 cz,
                                            u32 old, u32 octree, u32 sf) {
	Classified r{kNotAffected, octree, sf, false, Color::none()};
	if (!ed.stateful()) {
		r.t = ed.edit_node(cl, cx, cy, cz);
		return r;
	}
	r.sf |= kSfActive;
	const u32 leafL = P.c.color_leaf;
	Color col = get_fill(octree);
	r.t = ed.edit_node_vbr(cl, cx, cy, cz, old != kNull, col);
	if (!(r.sf & kSfFinal)) {
		if (cl <= leafL) {
			if (col.has()) {
				r.octree = cmake(kTagColor, color_to_rgb8(col));
				r.sf |= kSfFinal;
			} else if (r.t == kClear)
				r.octree = kColorNull;
			else if (r.t == kProceed && cl == leafL)
				r.sf |= kSfWriter | kSfNewJob;
		} else if (r.sf & kSfWriter) {
			if (col.has()) { // writer->Push(color, 8^(voxel_level - level)); Copy keeps the decoded colors
				r.push = true;
				r.col = col;
				r.sf |= kSfFinal;
			}
		}
	}
	return r;
}

// ---------------------------------------------------------------- top-down
// out: [0] = 1 if the root proceeds, [1] = root result otherwise, [2] = color root state
__global__ void k_root(EditParams P, u32 root, Entry *ent0, u32 *list0, u32 *out) {
	if (threadIdx.x | blockIdx.x)
		return;
	u32 cnt = 0, single = 0;
	bool active = false;
	if (P.n_edits == 1)
		cnt = 1;
	else
		for (u32 k = 0; k < P.n_edits; ++k) {
			DevEditor e = P.editor(k);
			if (e.geom(0, 0, 0, 0) != kNotAffected) {
				list0[cnt++] = k;
				single = k;
				active |= e.stateful();
			}
		}
	Classified c{kNotAffected, P.octree_root, 0u, false, Color::none()};
	if (cnt == 1)
		c = classify_single(P, P.editor(single), 0, 0, 0, 0, root, P.octree_root, 0u);
	else if (cnt > 1) {
		c.t = kProceed;
		if (active)
			c.sf |= kSfActive | (P.c.color_leaf == 0 ? (kSfWriter | kSfNewJob) : 0u);
	}
	if (c.t == kProceed) {
		ent0[0] = Entry{0, 0, 0, root, kNull, c.sf & ~kSfNewJob, c.octree, (c.sf & kSfWriter) ? 0u : kNull,
		                cnt == 1 ? single : 0u, cnt};
		out[0] = 1;
	} else {
		out[0] = 0;
		out[1] = c.t == kClear ? kNull : (c.t == kFill ? P.filled[0] : root);
		out[2] = c.octree;
	}
}

// One lane per (node, child). Writes the child's new pointer when it terminates, its color state,
// the proceed flag and the length of its filtered edit list (lists of 2+ edits are scanned).
__global__ void k_expand(EditParams P, u32 l, const Entry *__restrict__ ent, u32 n, const u32 *__restrict__ lists,
                         u32 *__restrict__ kids, u32 *__restrict__ okids, u32 *__restrict__ flag,
                         u32 *__restrict__ lcnt, uint4 *__restrict__ cstate, PushEvent *__restrict__ ev,
                         u32 *__restrict__ ev_count, u32 ev_cap) {
	u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
	u64 total = (u64)n * 8u;
	if (t >= total) {
		if (t == total)
			flag[t] = 0, lcnt[t] = 0;
		return;
	}
	u32 e = (u32)(t >> 3), i = (u32)(t & 7u);
	Entry E = ent[e];
	u32 old = kNull;
	if (E.ptr != kNull) {
		u32 m = P.words[E.ptr];
		if ((m >> i) & 1u)
			old = P.words[E.ptr + 1u + (u32)__builtin_popcount(m & ((1u << i) - 1u))];
	}
	const u32 cl = l + 1, leafL = P.c.color_leaf;
	const u32 cx = (E.x << 1) | (i & 1u), cy = (E.y << 1) | ((i >> 1) & 1u), cz = (E.z << 1) | ((i >> 2) & 1u);
	u32 octree = E.octree;
	if (P.stateful && cl <= leafL)
		octree = color_child(P.cnodes, E.octree, i);
	const u32 sf_in = i | (E.sf & (kSfFinal | kSfWriter));
	Classified c{kNotAffected, octree, sf_in, false, Color::none()};
	u32 cnt = 1, single = E.loff;
	if (E.llen == 1)
		c = classify_single(P, P.editor(E.loff), cl, cx, cy, cz, old, octree, sf_in);
	else {
		cnt = 0;
		bool active = false;
		for (u32 k = 0; k < E.llen; ++k) {
			u32 idx = lists[E.loff + k];
			DevEditor ed = P.editor(idx);
			if (ed.geom(cl, cx, cy, cz) != kNotAffected) {
				++cnt;
				single = idx;
				active |= ed.stateful();
			}
		}
		if (cnt == 1)
			c = classify_single(P, P.editor(single), cl, cx, cy, cz, old, octree, sf_in);
		else if (cnt > 1) {
			c.t = kProceed;
			if (active)
				c.sf |= kSfActive | (cl == leafL ? (kSfWriter | kSfNewJob) : 0u);
		}
	}
	if (c.push) {
		u32 d = cl - leafL, msk = (1u << d) - 1u, sh = 3u * (P.c.voxel_level - cl);
		u32 k = atomicAdd(ev_count, 1u);
		if (k < ev_cap)
			ev[k] = PushEvent{E.job, morton3(cx & msk, cy & msk, cz & msk) << sh, 1u << sh, c.col.colors, c.col.wb, 0, 0, 0};
	}
	kids[t] = c.t == kClear ? kNull : (c.t == kFill ? P.filled[cl] : old);
	if (okids)
		okids[t] = c.octree;
	bool proceed = c.t == kProceed;
	flag[t] = proceed ? 1u : 0u;
	lcnt[t] = proceed && cnt > 1 ? cnt : 0u;
	cstate[t] = make_uint4(c.octree, c.sf, single, cnt);
}

__global__ void k_emit(u32 l, const Entry *__restrict__ ent, u32 n, const u32 *__restrict__ lists,
                       const EditParams P, const u32 *__restrict__ kids, const u32 *__restrict__ flag,
                       const u32 *__restrict__ off, const u32 *__restrict__ loffs, const uint4 *__restrict__ cstate,
                       Entry *__restrict__ next, u32 *__restrict__ next_lists, u32 *__restrict__ first_child,
                       u32 *__restrict__ nchild) {
	u64 t = (u64)blockIdx.x * blockDim.x + threadIdx.x;
	if (t >= (u64)n * 8u)
		return;
	u32 e = (u32)(t >> 3), i = (u32)(t & 7u);
	if (i == 0 && first_child) {
		first_child[e] = off[t];
		nchild[e] = off[t + 8] - off[t];
	}
	if (!flag[t])
		return;
	Entry E = ent[e];
	uint4 cs = cstate[t];
	u32 idx = off[t], sf = cs.y;
	u32 job = (sf & kSfNewJob) ? idx : ((sf & kSfWriter) ? E.job : kNull);
	const u32 cx = (E.x << 1) | (i & 1u), cy = (E.y << 1) | ((i >> 1) & 1u), cz = (E.z << 1) | ((i >> 2) & 1u);
	u32 lo = cs.z, ln = cs.w;
	if (ln > 1) { // the child's filtered list, in the parent's (submission) order
		lo = loffs[t];
		u32 w = lo;
		for (u32 k = 0; k < E.llen; ++k) {
			u32 ix = lists[E.loff + k];
			if (P.editor(ix).geom(l + 1, cx, cy, cz) != kNotAffected)
				next_lists[w++] = ix;
		}
	}
	next[idx] = Entry{cx, cy, cz, kids[t], e, sf & ~kSfNewJob, cs.x, job, lo, ln};
}

// ---------------------------------------------------------------- colors: dense decode + pushes
// Decoded color of voxel v of the chunk stored at leaf word index idx (layout of
// DAGColorPool::write_leaf_chunk, DAGColorPool.hpp:66-96; lookup as Color_GetLeafColor).
__device__ inline Color chunk_voxel(const u32 *__restrict__ L, u32 idx, u32 v) {
	u32 M = L[idx + 1u], B = L[idx + 2u];
	u32 mo = idx + 4u, bo = mo + 2u * M, wo = bo + 2u * B;
	u32 m = v >> 14, vid = v & 0x3FFFu;
	if (m >= M)
		return Color::none();
	u32 fb = L[mo + 2u * m], ws = L[mo + 2u * m + 1u];
	u32 lb = m + 1u < M ? L[mo + 2u * (m + 1u)] : B;
	u32 lo = fb, cnt = lb - fb;
	while (cnt > 0) { // last header with voxoff <= vid
		u32 step = cnt >> 1;
		if ((L[bo + 2u * (lo + step) + 1u] >> 18) <= vid) {
			lo += step + 1u;
			cnt -= step + 1u;
		} else
			cnt = step;
	}
	u32 h = lo - 1u;
	u32 hc = L[bo + 2u * h], hp = L[bo + 2u * h + 1u];
	u32 bpw = (hp >> 16) & 3u;
	if (bpw == 0)
		return Color{hc, 1u << 16};
	u32 bit = ws + (hp & 0xFFFFu) + (vid - (hp >> 18)) * bpw;
	u32 o = bit & 31u, w = L[wo + (bit >> 5)] >> o;
	if (o + bpw > 32u)
		w |= L[wo + (bit >> 5) + 1u] << (32u - o);
	w &= (1u << bpw) - 1u;
	return Color{hc, w | (bpw << 8) | (1u << 16)};
}

__global__ void k_dense_init(const Entry *__restrict__ ent, u32 n, u32 V, u32 blocks_per_job,
                             const u32 *__restrict__ cleaves, Color *__restrict__ dense) {
	u32 e = blockIdx.x / blocks_per_job, b = blockIdx.x % blocks_per_job;
	if (e >= n)
		return;
	Entry E = ent[e];
	if (!(E.sf & kSfWriter))
		return;
	Color *out = dense + (u64)e * V;
	bool leaf = ctag(E.octree) == kTagLeaf;
	Color fill = get_fill(E.octree);
	for (u32 v = b * blockDim.x + threadIdx.x; v < V; v += blocks_per_job * blockDim.x)
		out[v] = leaf ? chunk_voxel(cleaves, cdata(E.octree), v) : fill;
}

__global__ void k_push(const PushEvent *__restrict__ ev, u32 nev, u32 V, Color *__restrict__ dense) {
	u32 k = blockIdx.x;
	if (k >= nev)
		return;
	PushEvent p = ev[k];
	Color c{p.colors, p.wb};
	Color *out = dense + (u64)p.job * V + p.offset;
	for (u32 v = threadIdx.x; v < p.count; v += blockDim.x)
		out[v] = c;
}

// ---------------------------------------------------------------- leaves
// edit_leaf (NodePool.hpp:319-343): one wavefront per leaf, lane i = voxel i (GetLeafCoord order).
// A leaf with an edit list folds the voxel through the edits in order.
__global__ void k_leaf(EditParams P, const Entry *__restrict__ ent, u32 n, const u32 *__restrict__ lists,
                       u32 *__restrict__ status, u32 *__restrict__ pend, u32 *__restrict__ content,
                       Color *__restrict__ dense, u32 V) {
	const u32 lane = threadIdx.x & 63u;
	const u32 e = (u32)(((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
	if (e >= n)
		return; // whole wavefront
	Entry E = ent[e];
	u32 w0 = 0, w1 = 0;
	if (E.ptr != kNull) {
		w0 = P.words[E.ptr];
		w1 = P.words[E.ptr + 1u];
	}
	bool v = ((lane < 32u ? (w0 >> lane) : (w1 >> (lane - 32u))) & 1u) != 0;
	u32 vx = (E.x << 2) | ((lane >> 2) & 2u) | (lane & 1u);
	u32 vy = (E.y << 2) | ((lane >> 3) & 2u) | ((lane >> 1) & 1u);
	u32 vz = (E.z << 2) | ((lane >> 4) & 2u) | ((lane >> 2) & 1u);
	const bool writer = (E.sf & kSfWriter) && !(E.sf & kSfFinal);
	Color *cp = nullptr;
	Color col = Color::none();
	if (writer) { // VBRChunkWriter::Edit (VBRColor.hpp:477-486) on the dense color of this voxel
		u32 msk = (1u << P.leaf_bits) - 1u;
		cp = dense + (u64)E.job * V + morton3(vx & msk, vy & msk, vz & msk);
		col = *cp;
	}
	bool nv = v;
	for (u32 k = 0; k < E.llen; ++k) {
		DevEditor ed = P.editor(E.llen == 1 ? E.loff : lists[E.loff + k]);
		if (!ed.stateful())
			nv = ed.edit_voxel(vx, vy, vz, nv);
		else if (writer)
			nv = ed.edit_voxel_vbr(vx, vy, vz, nv, col);
		else {
			Color dummy = Color::none();
			nv = ed.edit_voxel_vbr(vx, vy, vz, nv, dummy);
		}
	}
	if (writer && (E.sf & kSfActive))
		*cp = col;
	unsigned long long m = __ballot(nv);
	if (lane == 0) {
		u32 n0 = (u32)m, n1 = (u32)(m >> 32);
		bool changed = n0 != w0 || n1 != w1;
		pend[e] = 0;
		if (!changed)
			status[e] = E.ptr;
		else if (m == 0ull)
			status[e] = kNull;
		else {
			pend[e] = 1;
			content[2 * (u64)e] = n0;
			content[2 * (u64)e + 1] = n1;
		}
	}
}

// ---------------------------------------------------------------- bottom-up
// edit_node's join (NodePool.hpp:375-395): new mask from the children's occupancy, pack.
__global__ void k_build(EditParams P, const Entry *__restrict__ ent, u32 n, const u32 *__restrict__ kids,
                        u32 *__restrict__ status, u32 *__restrict__ pend, u32 *__restrict__ content,
                        u32 *__restrict__ csize) {
	u32 e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n)
		return;
	Entry E = ent[e];
	u32 om = 0, k = 1;
	if (E.ptr != kNull)
		om = P.words[E.ptr];
	bool changed = false;
	u32 nm = 0, c[9];
	for (u32 i = 0; i < 8; ++i) {
		u32 o = ((om >> i) & 1u) ? P.words[E.ptr + k++] : kNull;
		u32 nk = kids[(u64)e * 8u + i];
		changed |= nk != o;
		if (nk != kNull)
			nm |= 1u << i;
	}
	pend[e] = 0;
	if (!changed) {
		status[e] = E.ptr;
		return;
	}
	if (nm == 0) {
		status[e] = kNull;
		return;
	}
	c[0] = nm;
	k = 1;
	for (u32 i = 0; i < 8; ++i) {
		u32 nk = kids[(u64)e * 8u + i];
		if (nk != kNull)
			c[k++] = nk;
	}
	pend[e] = 1;
	csize[e] = k;
	for (u32 j = 0; j < k; ++j)
		content[(u64)e * 9u + j] = c[j];
}

__device__ inline bool same_node(const u32 *content, u32 stride, const u32 *csize, u32 a, u32 b) {
	u32 na = csize ? csize[a] : 2u, nb = csize ? csize[b] : 2u;
	if (na != nb)
		return false;
	for (u32 j = 0; j < na; ++j)
		if (content[(u64)a * stride + j] != content[(u64)b * stride + j])
			return false;
	return true;
}

// Within-level dedup: identical new nodes elect the lowest entry index (the first in DFS order,
// the one the reference appends; the others find it).
__global__ void k_dedup(EditParams P, u32 level, u32 n, const u32 *__restrict__ pend, const u32 *__restrict__ content,
                        const u32 *__restrict__ csize, u32 stride, u32 *__restrict__ table, u32 tmask,
                        u32 *__restrict__ slot, u32 *__restrict__ bucket) {
	u32 e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n || !pend[e])
		return;
	const u32 *c = content + (u64)e * stride;
	bool leaf = csize == nullptr;
	u32 nsz = leaf ? 2u : csize[e];
	u32 h = leaf ? murmur_leaf(c[0], c[1]) : murmur_span(c, nsz);
	bucket[e] = P.c.level_base[level] + (h & ((1u << P.c.bucket_bits[level]) - 1u));
	u32 s = (h ^ (h >> 15)) * 0x2c1b3c6du;
	s = (s ^ (s >> 12)) & tmask;
	for (;;) {
		u32 cur = atomicCAS(&table[s], kNull, e);
		if (cur == kNull || same_node(content, stride, csize, cur, e)) {
			if (cur != kNull)
				atomicMin(&table[s], e);
			slot[e] = s;
			return;
		}
		s = (s + 1u) & tmask;
	}
}

__global__ void k_find(EditParams P, u32 n, const u32 *__restrict__ pend, const u32 *__restrict__ content,
                       const u32 *__restrict__ csize, u32 stride, const u32 *__restrict__ table,
                       const u32 *__restrict__ slot, const u32 *__restrict__ bucket, u32 *__restrict__ res,
                       u32 *__restrict__ nf) {
	u32 e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n) {
		if (e == n)
			nf[e] = 0;
		return;
	}
	nf[e] = 0;
	if (!pend[e] || table[slot[e]] != e)
		return;
	bool leaf = csize == nullptr;
	u32 nsz = leaf ? 2u : csize[e];
	u32 b = bucket[e];
	u32 r = find_in_bucket(P.words, P.c, b, P.bucket_words[b], content + (u64)e * stride, nsz, leaf);
	res[e] = r;
	nf[e] = r == kNull ? 1u : 0u;
}

__global__ void k_compact_nf(u32 n, const u32 *__restrict__ nf, const u32 *__restrict__ off,
                             const u32 *__restrict__ bucket, u32 level_base, u32 *__restrict__ keys,
                             u32 *__restrict__ vals) {
	u32 e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n || !nf[e])
		return;
	keys[off[e]] = bucket[e] - level_base;
	vals[off[e]] = e;
}

// Per-bucket sequential appends in DFS order (the sorted segment of each bucket).
__global__ void k_append(EditParams P, u32 m, const u32 *__restrict__ keys, const u32 *__restrict__ vals,
                         u32 level_base, const u32 *__restrict__ content, const u32 *__restrict__ csize, u32 stride,
                         u32 *__restrict__ res, u32 *__restrict__ failed) {
	u32 j = blockIdx.x * blockDim.x + threadIdx.x;
	if (j >= m || (j > 0 && keys[j] == keys[j - 1]))
		return;
	u32 key = keys[j], b = key + level_base;
	u32 bw = P.bucket_words[b];
	bool any_fail = false;
	for (u32 k = j; k < m && keys[k] == key; ++k) {
		u32 e = vals[k];
		u32 nsz = csize ? csize[e] : 2u;
		u32 p = append_to_bucket(P.words, P.c, b, bw, content + (u64)e * stride, nsz);
		failed[e] = p == kNull ? 1u : 0u;
		any_fail |= p == kNull;
		res[e] = p;
	}
	P.bucket_words[b] = bw;
	if (any_fail)
		atomicOr(&P.meta[kMetaOverflow], 1u);
}

__global__ void k_resolve(u32 n, const Entry *__restrict__ ent, const u32 *__restrict__ pend,
                          const u32 *__restrict__ status, const u32 *__restrict__ table, const u32 *__restrict__ slot,
                          const u32 *__restrict__ res, const u32 *__restrict__ nf, const u32 *__restrict__ failed,
                          u32 *__restrict__ parent_kids, u32 *__restrict__ root_out) {
	u32 e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n)
		return;
	Entry E = ent[e];
	u32 r;
	if (!pend[e])
		r = status[e];
	else {
		u32 rep = table[slot[e]];
		// a rep that was not found and failed to append: each entry keeps its own old pointer
		r = (nf[rep] && failed[rep]) ? E.ptr : res[rep];
	}
	if (parent_kids)
		parent_kids[(u64)E.parent * 8u + (E.sf & 7u)] = r;
	else
		root_out[0] = r;
}

// ---------------------------------------------------------------- colors: encode + SetLeaf
// Canonical VBR encoding of a dense color sequence (VBRChunkWriter::append, VBRColor.hpp:384-411):
// a macro block every 2^14 voxels, a header at each macro start and where (colors, bpw) changes,
// weights concatenated LSB first.
constexpr u32 kEncThreads = 256;

struct EncScan {
	u32 heads, bits;
};

__device__ inline bool is_head(const Color *d, u32 v) {
	if (v == 0 || (v & 0x3FFFu) == 0)
		return true;
	Color a = d[v - 1], b = d[v];
	return a.colors != b.colors || ((a.wb >> 8) & 0xFFu) != ((b.wb >> 8) & 0xFFu);
}

__device__ inline EncScan block_excl_scan(EncScan x, EncScan *sh, EncScan &total) {
	u32 t = threadIdx.x;
	sh[t] = x;
	__syncthreads();
	for (u32 o = 1; o < kEncThreads; o <<= 1) {
		EncScan y = t >= o ? sh[t - o] : EncScan{0, 0};
		__syncthreads();
		sh[t].heads += y.heads;
		sh[t].bits += y.bits;
		__syncthreads();
	}
	EncScan incl = sh[t];
	total = sh[kEncThreads - 1];
	__syncthreads();
	return EncScan{incl.heads - x.heads, incl.bits - x.bits};
}

// Size pass: one block per color-leaf entry with a writer. info[e] = {data_size, append_size, inplace, B}
__global__ void __launch_bounds__(kEncThreads) k_encode_size(EditParams P, const Entry *__restrict__ ent, u32 n,
                                                              const Color *__restrict__ dense, u32 V,
                                                              u32 *__restrict__ info, u32 *__restrict__ app) {
	__shared__ EncScan sh[kEncThreads];
	u32 e = blockIdx.x;
	if (e >= n)
		return;
	Entry E = ent[e];
	if (!(E.sf & kSfWriter)) {
		if (threadIdx.x == 0)
			app[e] = 0;
		return;
	}
	const Color *d = dense + (u64)e * V;
	u32 per = (V + kEncThreads - 1) / kEncThreads, v0 = threadIdx.x * per, v1 = min(V, v0 + per);
	EncScan x{0, 0};
	for (u32 v = v0; v < v1; ++v) {
		x.heads += is_head(d, v) ? 1u : 0u;
		x.bits += (d[v].wb >> 8) & 0xFFu;
	}
	EncScan tot;
	block_excl_scan(x, sh, tot);
	if (threadIdx.x == 0) {
		u32 M = (V + 0x3FFFu) >> 14, W = (tot.bits + 31u) >> 5;
		u32 ds = 2u * M + 2u * tot.heads + W + 4u;
		u32 as = (ds & 1u) ? ds + 1u : ds;
		u32 inplace = 0;
		if (!P.c.keep_history && ctag(E.octree) == kTagLeaf) { // DAGColorPool::SetLeaf (:172-204)
			u32 block = P.cleaves[cdata(E.octree)];
			if (ds <= block)
				inplace = 1, as = 0;
			else
				as = max(block << 1, as);
		}
		info[4 * e] = ds;
		info[4 * e + 1] = as;
		info[4 * e + 2] = inplace;
		info[4 * e + 3] = tot.heads;
		app[e] = as;
	}
}

__global__ void __launch_bounds__(kEncThreads) k_encode_write(EditParams P, const Entry *__restrict__ ent, u32 n,
                                                               const Color *__restrict__ dense, u32 V,
                                                               const u32 *__restrict__ info,
                                                               const u32 *__restrict__ app_off, u32 *__restrict__ ores,
                                                               u32 *__restrict__ parent_okids, u64 leaf_cap) {
	__shared__ EncScan sh[kEncThreads];
	__shared__ u32 macro_ws[160];
	u32 e = blockIdx.x;
	if (e >= n)
		return;
	Entry E = ent[e];
	u32 result = E.octree;
	if (E.sf & kSfWriter) {
		u32 ds = info[4 * e], as = info[4 * e + 1], inplace = info[4 * e + 2], B = info[4 * e + 3];
		u32 idx = inplace ? cdata(E.octree) : P.meta[kMetaColorLeaves] + app_off[e];
		if (inplace || (u64)idx + as <= leaf_cap) {
			u32 *L = P.cleaves;
			const Color *d = dense + (u64)e * V;
			u32 M = (V + 0x3FFFu) >> 14;
			u32 per = (V + kEncThreads - 1) / kEncThreads, v0 = threadIdx.x * per, v1 = min(V, v0 + per);
			EncScan x{0, 0};
			for (u32 v = v0; v < v1; ++v) {
				x.heads += is_head(d, v) ? 1u : 0u;
				x.bits += (d[v].wb >> 8) & 0xFFu;
			}
			EncScan tot;
			EncScan pre = block_excl_scan(x, sh, tot);
			u32 W = (tot.bits + 31u) >> 5;
			u32 mo = idx + 4u, bo = mo + 2u * M, wo = bo + 2u * B;
			// macro blocks: the voxel at each macro start is a head
			EncScan run = pre;
			for (u32 v = v0; v < v1; ++v) {
				bool h = is_head(d, v);
				if ((v & 0x3FFFu) == 0) {
					macro_ws[v >> 14] = run.bits;
					L[mo + 2u * (v >> 14)] = run.heads;
					L[mo + 2u * (v >> 14) + 1u] = run.bits;
				}
				run.heads += h ? 1u : 0u;
				run.bits += (d[v].wb >> 8) & 0xFFu;
			}
			for (u32 w = threadIdx.x; w < W; w += kEncThreads)
				L[wo + w] = 0u;
			__syncthreads();
			run = pre;
			for (u32 v = v0; v < v1; ++v) {
				Color c = d[v];
				u32 bpw = (c.wb >> 8) & 0xFFu;
				if (is_head(d, v)) {
					L[bo + 2u * run.heads] = c.colors;
					L[bo + 2u * run.heads + 1u] = ((v & 0x3FFFu) << 18) | (bpw << 16) | (run.bits - macro_ws[v >> 14]);
					run.heads += 1u;
				}
				if (bpw) {
					u32 w = c.wb & ((1u << bpw) - 1u), o = run.bits & 31u;
					atomicOr(&L[wo + (run.bits >> 5)], w << o);
					if (o + bpw > 32u)
						atomicOr(&L[wo + (run.bits >> 5) + 1u], w >> (32u - o));
				}
				run.bits += bpw;
			}
			if (threadIdx.x == 0) {
				L[idx + 1u] = M;
				L[idx + 2u] = B;
				L[idx + 3u] = W;
				if (!inplace)
					L[idx] = as;
			}
			if (!inplace)
				for (u32 w = ds + threadIdx.x; w < as; w += kEncThreads)
					L[idx + w] = 0u;
			result = inplace ? E.octree : cmake(kTagLeaf, idx);
		} else if (threadIdx.x == 0)
			atomicOr(&P.meta[kMetaOverflow], 2u);
	}
	if (threadIdx.x == 0) {
		ores[e] = result;
		if (parent_okids)
			parent_okids[(u64)E.parent * 8u + (E.sf & 7u)] = result;
	}
}

// ---------------------------------------------------------------- colors: SetNode in post-order
// Pass A (bottom-up): SetNode's decision (DAGColorPool.hpp:147-162) with new children marked;
// cnt = color nodes appended in the subtree.
__global__ void k_setnode_a(EditParams P, const Entry *__restrict__ ent, u32 n, const u32 *__restrict__ okids,
                            const u32 *__restrict__ first_child, const u32 *__restrict__ nchild,
                            const u32 *__restrict__ child_cnt, u32 *__restrict__ cnt, u32 *__restrict__ ores,
                            u32 *__restrict__ parent_okids) {
	u32 e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n)
		return;
	Entry E = ent[e];
	u32 r = E.octree, c = 0;
	if (!(E.sf & kSfFinal) && (E.sf & kSfActive)) {
		u32 ch[8];
		bool all_null = true, same_color = true;
		for (u32 i = 0; i < 8; ++i) {
			ch[i] = okids[(u64)e * 8u + i];
			all_null = all_null && ctag(ch[i]) == kTagNull;
		}
		same_color = ctag(ch[0]) == kTagColor;
		for (u32 i = 1; i < 8 && same_color; ++i)
			same_color = ch[i] == ch[0];
		if (all_null)
			r = kColorNull;
		else if (same_color)
			r = ch[0];
		else {
			bool unchanged = ctag(E.octree) == kTagNode;
			for (u32 i = 0; i < 8 && unchanged; ++i)
				unchanged = P.cnodes[(u64)cdata(E.octree) * 8u + i] == ch[i];
			r = unchanged ? E.octree : kColorNewNode;
		}
		if (child_cnt)
			for (u32 k = 0; k < nchild[e]; ++k)
				c += child_cnt[first_child[e] + k];
		c += r == kColorNewNode ? 1u : 0u;
	}
	cnt[e] = c;
	ores[e] = r;
	if (parent_okids)
		parent_okids[(u64)E.parent * 8u + (E.sf & 7u)] = r;
}
// Pass B (top-down): post-order indices. base[e] is the first index of e's subtree.
__global__ void k_setnode_b(u32 n, const u32 *__restrict__ base, const u32 *__restrict__ first_child,
                            const u32 *__restrict__ nchild, const u32 *__restrict__ child_cnt,
                            u32 *__restrict__ child_base, u32 *__restrict__ own) {
	u32 e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n)
		return;
	u32 run = base[e];
	if (child_cnt)
		for (u32 k = 0; k < nchild[e]; ++k) {
			child_base[first_child[e] + k] = run;
			run += child_cnt[first_child[e] + k];
		}
	own[e] = run;
}
// Pass C (bottom-up): write the appended nodes with their final children.
__global__ void k_setnode_c(EditParams P, const Entry *__restrict__ ent, u32 n, const u32 *__restrict__ okids,
                            const u32 *__restrict__ own, u32 *__restrict__ ores, u32 *__restrict__ parent_okids,
                            u64 cap) {
	u32 e = blockIdx.x * blockDim.x + threadIdx.x;
	if (e >= n || ores[e] != kColorNewNode)
		return;
	Entry E = ent[e];
	u32 idx = own[e], r;
	if ((u64)idx < cap) {
		for (u32 i = 0; i < 8; ++i)
			P.cnodes[(u64)idx * 8u + i] = okids[(u64)e * 8u + i];
		r = cmake(kTagNode, idx);
	} else {
		r = E.octree;
		atomicOr(&P.meta[kMetaOverflow], 2u);
	}
	ores[e] = r;
	if (parent_okids)
		parent_okids[(u64)E.parent * 8u + (E.sf & 7u)] = r;
}
__global__ void k_add_meta(u32 *meta, u32 which, const u32 *a, const u32 *b, u32 ib) {
	if (threadIdx.x | blockIdx.x)
		return;
	meta[which] += a[0] + (b ? b[ib] : 0u);
}

// ---------------------------------------------------------------- host orchestration
struct Level {
	Entry *ent = nullptr;
	u32 n = 0;
	u32 *kids = nullptr, *okids = nullptr, *first_child = nullptr, *nchild = nullptr;
	u32 *cnt = nullptr, *base = nullptr, *own = nullptr, *ores = nullptr;
};

template <typename T> static int buf(Pool &p, const char *name, u32 l, size_t count, T **out) {
	char key[64];
	snprintf(key, sizeof(key), "%s%u", name, l);
	return p.scratch.get<T>(key, count ? count : 1, out);
}

static int read_u32s(Pool &p, const u32 *d, u32 *h, u32 n) {
	HD_TRY(hipMemcpyAsync(h, d, n * 4, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

static int excl_scan(Pool &p, const u32 *in, u32 *out, u64 n) {
	size_t tmp = 0;
	HD_TRY(rocprim::exclusive_scan(nullptr, tmp, in, out, 0u, n, rocprim::plus<u32>(), p.stream));
	void *t;
	HD_RET(p.scratch.get<char>("scan_tmp", tmp + 16, (char **)&t));
	HD_TRY(rocprim::exclusive_scan(t, tmp, in, out, 0u, n, rocprim::plus<u32>(), p.stream));
	return HD_OK;
}

static int sort_pairs(Pool &p, u32 *keys, u32 *vals, u32 *keys2, u32 *vals2, u32 m, u32 bits) {
	size_t tmp = 0;
	u32 eb = bits ? bits : 1u;
	HD_TRY(rocprim::radix_sort_pairs(nullptr, tmp, keys, keys2, vals, vals2, m, 0, eb, p.stream));
	void *t;
	HD_RET(p.scratch.get<char>("sort_tmp", tmp + 16, (char **)&t));
	HD_TRY(rocprim::radix_sort_pairs(t, tmp, keys, keys2, vals, vals2, m, 0, eb, p.stream));
	return HD_OK;
}

// find-or-insert of the pending entries of one level, then write results into the parent level.
static int upsert_level(Pool &p, const EditParams &P, u32 l, Level &lv, u32 *status, u32 *pend, u32 *content,
                        u32 *csize, u32 *parent_kids, u32 *root_out) {
	const u32 n = lv.n;
	const bool leaf = l == p.dc.L - 1;
	const u32 stride = leaf ? 2u : 9u;
	u32 T = 64;
	while (T < 2u * n)
		T <<= 1;
	u32 *table, *slot, *bucket, *res, *nf, *nfo, *failed, *keys, *vals, *keys2, *vals2;
	HD_RET(buf(p, "u_table", 0, T, &table));
	HD_RET(buf(p, "u_slot", 0, n, &slot));
	HD_RET(buf(p, "u_bucket", 0, n, &bucket));
	HD_RET(buf(p, "u_res", 0, n, &res));
	HD_RET(buf(p, "u_nf", 0, n + 1, &nf));
	HD_RET(buf(p, "u_nfo", 0, n + 1, &nfo));
	HD_RET(buf(p, "u_failed", 0, n, &failed));
	HD_TRY(hipMemsetAsync(table, 0xFF, (size_t)T * 4, p.stream));
	const u32 B = 256;
	k_dedup<<<div_up(n, B), B, 0, p.stream>>>(P, l, n, pend, content, leaf ? nullptr : csize, stride, table, T - 1,
	                                          slot, bucket);
	k_find<<<div_up(n + 1, B), B, 0, p.stream>>>(P, n, pend, content, leaf ? nullptr : csize, stride, table, slot,
	                                             bucket, res, nf);
	HD_RET(excl_scan(p, nf, nfo, (u64)n + 1));
	u32 m;
	HD_RET(read_u32s(p, nfo + n, &m, 1));
	p.last_upserts += m;
	if (m) {
		HD_RET(buf(p, "u_keys", 0, m, &keys));
		HD_RET(buf(p, "u_vals", 0, m, &vals));
		HD_RET(buf(p, "u_keys2", 0, m, &keys2));
		HD_RET(buf(p, "u_vals2", 0, m, &vals2));
		k_compact_nf<<<div_up(n, B), B, 0, p.stream>>>(n, nf, nfo, bucket, p.dc.level_base[l], keys, vals);
		HD_RET(sort_pairs(p, keys, vals, keys2, vals2, m, p.dc.bucket_bits[l]));
		k_append<<<div_up(m, B), B, 0, p.stream>>>(P, m, keys2, vals2, p.dc.level_base[l], content,
		                                           leaf ? nullptr : csize, stride, res, failed);
	}
	k_resolve<<<div_up(n, B), B, 0, p.stream>>>(n, lv.ent, pend, status, table, slot, res, nf, failed, parent_kids,
	                                             root_out);
	HD_TRY(hipGetLastError());
	return HD_OK;
}

static bool is_stateful(const hd_edit_desc &d) {
	bool dig = d.kind == HD_EDIT_SPHERE_DIG || d.kind == HD_EDIT_AABB_DIG;
	bool paint = d.kind == HD_EDIT_SPHERE_PAINT || d.kind == HD_EDIT_AABB_PAINT;
	return (d.colored || paint) && !dig;
}

// One level-synchronous pass over a group of edits with pairwise disjoint voxel AABBs (a single
// edit is the plain path: NodePoolBase::Edit).
static int edit_pass(Pool &p, u32 root, u32 croot, const hd_edit_desc *edits, u32 n_edits, u32 *root_out,
                     u32 *croot_out) {
	HD_RET(make_filled(p));
	const u32 L = p.dc.L, leafL = p.dc.color_leaf;
	EditParams P;
	P.c = p.dc;
	P.hm = p.d_hm;
	P.n_edits = n_edits;
	P.words = p.words;
	P.bucket_words = p.bucket_words;
	P.filled = p.d_filled;
	P.cnodes = p.cnodes;
	P.cleaves = p.cleaves;
	P.meta = p.meta;
	P.stateful = 0;
	for (u32 k = 0; k < n_edits; ++k) {
		const hd_edit_desc &d = edits[k];
		if (d.kind > HD_EDIT_COLUMN_FILL)
			return set_error("unknown edit kind %u", d.kind), HD_ERR_ARG;
		if (d.kind == HD_EDIT_COLUMN_FILL && !p.hm_mn)
			return set_error("column fill needs hd_pool_set_heightmap"), HD_ERR_ARG;
		P.stateful |= is_stateful(d) ? 1u : 0u;
	}
	P.octree_root = croot;
	P.leaf_bits = p.dc.voxel_level - leafL;
	if (P.stateful && P.leaf_bits > 7)
		return set_error("color leaves deeper than 8^7 voxels are not supported"), HD_ERR_UNSUPPORTED;
	const u32 V = P.stateful ? 1u << (3u * P.leaf_bits) : 0u;
	hd_edit_desc *d_edits;
	HD_RET(p.scratch.get<hd_edit_desc>("pass_edits", n_edits, &d_edits));
	HD_TRY(hipMemcpyAsync(d_edits, edits, n_edits * sizeof(hd_edit_desc), hipMemcpyHostToDevice, p.stream));
	P.edits = d_edits;

	std::vector<Level> lv(L);
	std::vector<u32 *> lists(L, nullptr);
	u32 *rootbuf;
	HD_RET(p.scratch.get<u32>("rootbuf", 8, &rootbuf));
	HD_RET(buf(p, "ent", 0, 1, &lv[0].ent));
	HD_RET(buf(p, "lists", 0, n_edits, &lists[0]));
	k_root<<<1, 64, 0, p.stream>>>(P, root, lv[0].ent, lists[0], rootbuf);
	u32 rb[3];
	HD_RET(read_u32s(p, rootbuf, rb, 3));
	if (rb[0] == 0) {
		*root_out = rb[1];
		*croot_out = P.stateful ? rb[2] : croot;
		return HD_OK;
	}
	lv[0].n = 1;
	u64 entries = 1;
	// ---- top-down (push events of level l's children go to a per-level buffer of 8 n_l entries)
	std::vector<PushEvent *> ev(L, nullptr);
	u32 *ev_count;
	HD_RET(p.scratch.get<u32>("ev_count", L, &ev_count));
	HD_TRY(hipMemsetAsync(ev_count, 0, 4 * L, p.stream));
	for (u32 l = 0; l + 1 < L; ++l) {
		u32 n = lv[l].n;
		if (!n)
			break;
		u64 n8 = (u64)n * 8u;
		u32 *flag, *off, *lcnt, *loffs;
		uint4 *cstate;
		HD_RET(buf(p, "kids", l, n8, &lv[l].kids));
		if (P.stateful && l < leafL) {
			HD_RET(buf(p, "okids", l, n8, &lv[l].okids));
			HD_RET(buf(p, "fchild", l, n, &lv[l].first_child));
			HD_RET(buf(p, "nchild", l, n, &lv[l].nchild));
		}
		const bool events = P.stateful && l + 1 > leafL;
		if (events)
			HD_RET(buf(p, "events", l, n8, &ev[l]));
		HD_RET(buf(p, "flag", 0, n8 + 1, &flag));
		HD_RET(buf(p, "off", 0, n8 + 1, &off));
		HD_RET(buf(p, "lcnt", 0, n8 + 1, &lcnt));
		HD_RET(buf(p, "loffs", 0, n8 + 1, &loffs));
		HD_RET(buf(p, "cstate", 0, n8, &cstate));
		k_expand<<<div_up(n8 + 1, 256), 256, 0, p.stream>>>(P, l, lv[l].ent, n, lists[l], lv[l].kids, lv[l].okids,
		                                                    flag, lcnt, cstate, ev[l], ev_count + l,
		                                                    events ? (u32)n8 : 0u);
		HD_TRY(hipGetLastError());
		HD_RET(excl_scan(p, flag, off, n8 + 1));
		u32 tot[2];
		if (n_edits > 1) {
			HD_RET(excl_scan(p, lcnt, loffs, n8 + 1));
			HD_TRY(hipMemcpyAsync(&tot[1], loffs + n8, 4, hipMemcpyDeviceToHost, p.stream));
		} else
			tot[1] = 0;
		HD_RET(read_u32s(p, off + n8, &tot[0], 1));
		u32 nn = tot[0];
		lv[l + 1].n = nn;
		entries += nn;
		HD_RET(buf(p, "ent", l + 1, nn, &lv[l + 1].ent));
		HD_RET(buf(p, "lists", l + 1, tot[1], &lists[l + 1]));
		if (nn || lv[l].first_child)
			k_emit<<<div_up(n8, 256), 256, 0, p.stream>>>(l, lv[l].ent, n, lists[l], P, lv[l].kids, flag, off, loffs,
			                                              cstate, lv[l + 1].ent, lists[l + 1], lv[l].first_child,
			                                              lv[l].nchild);
		HD_TRY(hipGetLastError());
	}
	p.last_entries += entries;
	// ---- colors: dense decode + pushes
	Color *dense = nullptr;
	if (P.stateful && lv[leafL].n) {
		std::vector<u32> nev(L, 0);
		HD_RET(read_u32s(p, ev_count, nev.data(), L));
		u64 dense_words = (u64)lv[leafL].n * V;
		if (dense_words * sizeof(Color) > (48ull << 30))
			return set_error("color scratch too large (%llu leaves)", (unsigned long long)lv[leafL].n),
			       HD_ERR_CAPACITY;
		HD_RET(p.scratch.get<Color>("dense", dense_words, &dense));
		u32 bpj = std::max(1u, std::min(64u, V / 1024u));
		k_dense_init<<<lv[leafL].n * bpj, 256, 0, p.stream>>>(lv[leafL].ent, lv[leafL].n, V, bpj, p.cleaves, dense);
		for (u32 l = 0; l < L; ++l)
			if (nev[l])
				k_push<<<nev[l], 256, 0, p.stream>>>(ev[l], nev[l], V, dense);
		HD_TRY(hipGetLastError());
		p.last_jobs += lv[leafL].n;
	}
	// ---- leaves + bottom-up DAG
	for (int l = (int)L - 1; l >= 0; --l) {
		u32 n = lv[l].n;
		if (!n)
			continue;
		bool leaf = l == (int)L - 1;
		u32 *status, *pend, *content, *csize = nullptr;
		HD_RET(buf(p, "status", 0, n, &status));
		HD_RET(buf(p, "pend", 0, n, &pend));
		HD_RET(buf(p, "content", 0, (size_t)n * (leaf ? 2 : 9), &content));
		if (leaf)
			k_leaf<<<div_up((u64)n * 64u, 256), 256, 0, p.stream>>>(P, lv[l].ent, n, lists[l], status, pend, content,
			                                                         dense, V);
		else {
			HD_RET(buf(p, "csize", 0, n, &csize));
			k_build<<<div_up(n, 256), 256, 0, p.stream>>>(P, lv[l].ent, n, lv[l].kids, status, pend, content, csize);
		}
		HD_TRY(hipGetLastError());
		HD_RET(upsert_level(p, P, (u32)l, lv[l], status, pend, content, csize, l > 0 ? lv[l - 1].kids : nullptr,
		                    rootbuf + 4));
	}
	// ---- colors: SetLeaf, SetNode
	u32 new_croot = croot;
	if (P.stateful) {
		if (lv[leafL].n) {
			u32 n = lv[leafL].n;
			u32 *info, *app, *appo;
			HD_RET(buf(p, "info", 0, 4 * (size_t)n, &info));
			HD_RET(buf(p, "app", 0, n + 1, &app));
			HD_RET(buf(p, "appo", 0, n + 1, &appo));
			HD_RET(buf(p, "ores", leafL, n, &lv[leafL].ores));
			HD_TRY(hipMemsetAsync(app + n, 0, 4, p.stream));
			k_encode_size<<<n, kEncThreads, 0, p.stream>>>(P, lv[leafL].ent, n, dense, V, info, app);
			HD_RET(excl_scan(p, app, appo, (u64)n + 1));
			k_encode_write<<<n, kEncThreads, 0, p.stream>>>(P, lv[leafL].ent, n, dense, V, info, appo, lv[leafL].ores,
			                                                leafL > 0 ? lv[leafL - 1].okids : nullptr, p.cleaf_cap);
			k_add_meta<<<1, 64, 0, p.stream>>>(p.meta, kMetaColorLeaves, appo + n, nullptr, 0);
			HD_TRY(hipGetLastError());
		}
		if (leafL > 0 && lv[0].n) {
			for (int l = (int)leafL - 1; l >= 0; --l) {
				u32 n = lv[l].n;
				if (!n)
					continue;
				HD_RET(buf(p, "cnt", l, n, &lv[l].cnt));
				HD_RET(buf(p, "ores", l, n, &lv[l].ores));
				k_setnode_a<<<div_up(n, 256), 256, 0, p.stream>>>(
				    P, lv[l].ent, n, lv[l].okids, lv[l].first_child, lv[l].nchild,
				    (l + 1 < (int)leafL && lv[l + 1].n) ? lv[l + 1].cnt : nullptr, lv[l].cnt, lv[l].ores,
				    l > 0 ? lv[l - 1].okids : nullptr);
			}
			HD_RET(buf(p, "cbase", 0, 1, &lv[0].base));
			HD_TRY(hipMemcpyAsync(lv[0].base, p.meta + kMetaColorNodes, 4, hipMemcpyDeviceToDevice, p.stream));
			for (u32 l = 0; l < leafL; ++l) {
				u32 n = lv[l].n;
				if (!n)
					break;
				bool has_child = l + 1 < leafL && lv[l + 1].n;
				if (has_child)
					HD_RET(buf(p, "cbase", l + 1, lv[l + 1].n, &lv[l + 1].base));
				HD_RET(buf(p, "own", l, n, &lv[l].own));
				k_setnode_b<<<div_up(n, 256), 256, 0, p.stream>>>(n, lv[l].base, lv[l].first_child, lv[l].nchild,
				                                                  has_child ? lv[l + 1].cnt : nullptr,
				                                                  has_child ? lv[l + 1].base : nullptr, lv[l].own);
			}
			for (int l = (int)leafL - 1; l >= 0; --l) {
				u32 n = lv[l].n;
				if (!n)
					continue;
				k_setnode_c<<<div_up(n, 256), 256, 0, p.stream>>>(P, lv[l].ent, n, lv[l].okids, lv[l].own, lv[l].ores,
				                                                  l > 0 ? lv[l - 1].okids : nullptr, p.cnode_cap);
			}
			k_add_meta<<<1, 64, 0, p.stream>>>(p.meta, kMetaColorNodes, lv[0].cnt, nullptr, 0);
			HD_TRY(hipGetLastError());
		}
		HD_RET(read_u32s(p, lv[0].ores, &new_croot, 1));
	}
	u32 r;
	HD_RET(read_u32s(p, rootbuf + 4, &r, 1));
	*root_out = r;
	*croot_out = new_croot;
	return HD_OK;
}

// Voxel AABB of an edit, clipped to the grid (inclusive bounds). Column fills cover the grid.
static void edit_aabb(const hd_edit_desc &d, u32 res, i64 lo[3], i64 hi[3]) {
	for (int i = 0; i < 3; ++i)
		lo[i] = 0, hi[i] = (i64)res - 1;
	if (d.kind <= HD_EDIT_SPHERE_PAINT) {
		i64 r = (i64)std::sqrt((double)d.r2);
		while ((r + 1) * (r + 1) <= (i64)d.r2)
			++r;
		while (r * r > (i64)d.r2)
			--r;
		for (int i = 0; i < 3; ++i)
			lo[i] = std::max<i64>(lo[i], (i64)d.p0[i] - r), hi[i] = std::min<i64>(hi[i], (i64)d.p0[i] + r);
	} else if (d.kind <= HD_EDIT_AABB_PAINT) {
		for (int i = 0; i < 3; ++i)
			lo[i] = std::max<i64>(lo[i], d.p0[i]), hi[i] = std::min<i64>(hi[i], (i64)d.p1[i] - 1);
	}
}

// Level scheduling of a batch: edit k goes to the first group after every group holding an earlier
// edit whose voxel AABB meets k's. Edits within a group have disjoint voxel sets, so they commute
// and one pass applies them; groups run in order, so every voxel sees its edits in submission order.
// A coarse 3-D grid of cells keeps the overlap search local.
void schedule_groups(const hd_edit_desc *edits, u32 n, u32 res, std::vector<std::vector<u32>> &groups) {
	const u32 cell_bits = res >= 64 ? std::max(4u, (u32)std::log2((double)res) - 5u) : 2u; // ~32^3 cells
	const u32 cells = std::max(1u, res >> cell_bits);
	std::vector<std::vector<u32>> grid((size_t)cells * cells * cells);
	std::vector<i64> box(6 * (size_t)n);
	std::vector<u32> gid(n, 0);
	groups.clear();
	for (u32 k = 0; k < n; ++k) {
		i64 *b = &box[6 * (size_t)k];
		edit_aabb(edits[k], res, b, b + 3);
		u32 g = 0;
		bool empty = b[0] > b[3] || b[1] > b[4] || b[2] > b[5];
		i64 c0[3], c1[3];
		if (!empty) {
			for (int i = 0; i < 3; ++i)
				c0[i] = b[i] >> cell_bits, c1[i] = b[i + 3] >> cell_bits;
			for (i64 z = c0[2]; z <= c1[2]; ++z)
				for (i64 y = c0[1]; y <= c1[1]; ++y)
					for (i64 x = c0[0]; x <= c1[0]; ++x)
						for (u32 j : grid[((size_t)z * cells + y) * cells + x]) {
							const i64 *o = &box[6 * (size_t)j];
							if (gid[j] + 1 > g && o[0] <= b[3] && b[0] <= o[3] && o[1] <= b[4] && b[1] <= o[4] &&
							    o[2] <= b[5] && b[2] <= o[5])
								g = gid[j] + 1;
						}
		}
		gid[k] = g;
		if (g >= groups.size())
			groups.resize(g + 1);
		groups[g].push_back(k);
		if (!empty)
			for (i64 z = c0[2]; z <= c1[2]; ++z)
				for (i64 y = c0[1]; y <= c1[1]; ++y)
					for (i64 x = c0[0]; x <= c1[0]; ++x)
						grid[((size_t)z * cells + y) * cells + x].push_back(k);
	}
}

int edit_batched(Pool &p, u32 root, u32 croot, const hd_edit_desc *edits, u32 n, u32 *root_out, u32 *croot_out) {
	std::vector<std::vector<u32>> groups;
	schedule_groups(edits, n, 1u << p.dc.voxel_level, groups);
	std::vector<hd_edit_desc> g_edits;
	for (auto &g : groups) {
		g_edits.clear();
		for (u32 k : g)
			g_edits.push_back(edits[k]);
		HD_RET(edit_pass(p, root, croot, g_edits.data(), (u32)g_edits.size(), &root, &croot));
	}
	p.last_groups = (u32)groups.size();
	*root_out = root;
	*croot_out = croot;
	return HD_OK;
}

} // namespace hd

using namespace hd;

extern "C" int hd_edit(hd_pool *h, uint32_t root, uint32_t color_root, const hd_edit_desc *edits, uint32_t n,
                       uint32_t mode, uint32_t *root_out, uint32_t *color_root_out) {
	if (!h || (!edits && n) || !root_out || !color_root_out)
		return set_error("hd_edit: null argument"), HD_ERR_ARG;
	Pool &p = h->p;
	auto t0 = std::chrono::steady_clock::now();
	p.last_entries = p.last_upserts = p.last_jobs = 0;
	u32 r = root, c = color_root;
	if (mode == HD_MODE_BATCHED && n > 1) {
		HD_RET(edit_batched(p, r, c, edits, n, &r, &c));
	} else {
		for (u32 i = 0; i < n; ++i)
			HD_RET(edit_pass(p, r, c, edits + i, 1, &r, &c));
	}
	HD_TRY(hipStreamSynchronize(p.stream));
	p.last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	*root_out = r;
	*color_root_out = c;
	return HD_OK;
}
