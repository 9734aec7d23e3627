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
#include <cstdlib>
#include <cstring>
#include <string>

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
	// the pass's edits may overlap: node edit lists are folded in order (list_chain), and a node an
	// edit only crosses takes per-voxel colors instead of a whole-node fill color
	u32 folded = 0;
	u32 octree_root;
	u32 leaf_bits; // voxel_level - color_leaf_level
	__device__ DevEditor editor(u32 k) const {
		// whole words (scalar loads for a uniform k; the color's byte fields would otherwise take
		// vector byte loads)
		union {
			hd_edit_desc d;
			u32 w[sizeof(hd_edit_desc) / 4];
		} t;
		const u32 *src = reinterpret_cast<const u32 *>(edits + k);
#pragma unroll
		for (u32 i = 0; i < sizeof(hd_edit_desc) / 4; ++i)
			t.w[i] = src[i];
		DevEditor e;
		e.d = t.d;
		e.hm = hm;
		e.voxel_level = c.voxel_level;
		return e;
	}
};

__device__ inline u32 color_child(const u32 *cnodes, u32 p, u32 i) { // DAGColorPool::GetChild (:139-143)
	u32 t = ctag(p);
	return t == kTagNode ? ldg(cnodes, (u64)cdata(p) * 8u + i) : (t == kTagColor ? p : kColorNull);
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
	// A folded pass: the fill color a crossed node would take (empty, or already the edit's color)
	// comes from the node's state before the pass, which earlier edits of the pass may have changed;
	// per-voxel colors give the occupied voxels the same colors without it.
	if (P.folded && col.has() && ed.geom(cl, cx, cy, cz) != kFill)
		col = Color::none();
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
	u32 cnt = 1, single = E.loff, kid_base = old;
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

// Block-wide exclusive scan (shuffle scans per wavefront + one exchange of wavefront totals).
__device__ inline EncScan block_excl_scan(EncScan x, EncScan *sh, EncScan &total) {
	const u32 lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	EncScan v = x;
#pragma unroll
	for (u32 o = 1; o < 64; o <<= 1) {
		u32 a = __shfl_up(v.heads, o), b = __shfl_up(v.bits, o);
		if (lane >= o)
			v.heads += a, v.bits += b;
	}
	if (lane == 63)
		sh[w] = v;
	__syncthreads();
	EncScan pre{0, 0}, tot{0, 0};
#pragma unroll
	for (u32 k = 0; k < kEncThreads / 64; ++k) {
		EncScan s = sh[k];
		if (k < w)
			pre.heads += s.heads, pre.bits += s.bits;
		tot.heads += s.heads;
		tot.bits += s.bits;
	}
	__syncthreads();
	total = tot;
	return EncScan{pre.heads + v.heads - x.heads, pre.bits + v.bits - x.bits};
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

// ================================================================ sync-free pass
// The same pass with no host round trip: level sizes stay in device memory and grids are fixed
// by host-side upper bounds. Top-down compaction is stable (block totals, then block prefix + LDS
// scan), so entry indices keep DFS order. Bottom-up, the build kernel finds each new node among
// the bucket's existing words and queues the misses on a per-bucket list; one walker per bucket
// then replays the misses in entry (DFS) order with an exact find-or-append over the words this
// level appended -- the reference's sequential upsert, with no cross-workgroup visibility needed
// inside a launch. Small levels run in single-workgroup kernels; roots stay in device memory so
// consecutive passes queue back to back.

constexpr u32 kAT = 256;       // threads per block of the sync-free kernels
constexpr u32 kTopPairs = 4096; // a level whose 8 n pairs fit this bound expands inside a_top
constexpr u32 kGather = 32;    // bucket lists up to this length are sorted in registers
constexpr u32 kWalkList = 64;  // groups per bucket listed by arrival for the walker
constexpr u32 kWalkShort = 16; // at most this many: one thread walks the bucket from registers

struct Pair {
	u32 a, b;
};
// Block-wide exclusive scan of kAT pairs: shuffle scans inside each wavefront, one LDS exchange
// of the wavefront totals (two barriers). sh needs kAT / 64 entries.
__device__ inline Pair block_excl_scan2(Pair x, Pair *sh, Pair &total) {
	const u32 lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	Pair v = x;
#pragma unroll
	for (u32 o = 1; o < 64; o <<= 1) {
		u32 a = __shfl_up(v.a, o), b = __shfl_up(v.b, o);
		if (lane >= o)
			v.a += a, v.b += b;
	}
	if (lane == 63)
		sh[w] = v;
	__syncthreads();
	Pair pre{0, 0}, tot{0, 0};
#pragma unroll
	for (u32 k = 0; k < kAT / 64; ++k) {
		Pair s = sh[k];
		if (k < w)
			pre.a += s.a, pre.b += s.b;
		tot.a += s.a;
		tot.b += s.b;
	}
	__syncthreads(); // sh reuse
	total = tot;
	return Pair{pre.a + v.a - x.a, pre.b + v.b - x.b};
}

// per-pass device counters: pm[kPmHdr + l] entries of level l, pm[kPmHdr + L + l] list words
enum PassMeta : u32 { kPmOverflow = 0, kPmEvents = 1, kPmLeafTotal = 2, kPmGen = 3, kPmRest = 4, kPmHdr = 5 };
// counters accumulated over the passes of one hd_edit call
enum AccMeta : u32 { kAccEntries = 0, kAccUpserts = 1, kAccJobs = 2, kAccOverflow = 3, kAccCount = 4 };

// Per-level buffers of the sync-free pass (a device-resident table, indexed by level).
struct ALevel {
	Entry *ent;
	u32 *lists;               // edit lists of multi-edit entries
	u32 *kids;                // 8 per entry: new pointer of each terminal child
	u32 *okids;               // 8 per entry: color state of each child (levels above the color leaves)
	u32 *first_child, *pmask; // index of the first proceeding child at the next level + their slots
	u32 *status, *pend, *res, *slot, *lpos, *bucket;
	u32 *content, *csize;
	u32 *ores, *ccnt, *cbase, *own; // color results, SetNode passes
	uint4 *pev; // 8 per entry (levels color leaf .. L - 2, color leaves rebuilt in LDS): a pushed
	            // child's {voxel offset in its job, colors, wb, 0}, valid where pmask bit 16 + slot is set
	u32 cap, list_cap;
};

constexpr u32 kListBatch = 8; // edit-list words loaded per round trip

struct PairOut {
	u32 flag, lcnt;
	uint4 cs;
	u32 kid; // the child's new pointer if terminal, else its old one
};

// classification of one (node, child) pair: the body of k_expand. The old node is read whole
// (9 words, the pool is padded) so its mask and child arrive in one memory round trip.
// bits: where the classification records which list edits touch the child (bit k for list entry k;
// at least (llen + 31) / 32 words, private to the thread), for emit_pair_e; or null.
// Child i of entry E (level l): its old pointer (the node read whole, 9 words, the pool is padded,
// so its mask and child arrive in one memory round trip), coordinates, color state and inherited
// flags.
struct PairIn {
	u32 old, octree, sf_in, cx, cy, cz;
};
__device__ inline PairIn pair_in(const EditParams &P, u32 l, const Entry &E, u32 i) {
	PairIn q;
	q.old = kNull;
	if (E.ptr != kNull) {
		u32 w[9];
#pragma unroll
		for (u32 k = 0; k < 9; ++k)
			w[k] = ldg(P.words, (u64)E.ptr + k);
		const u32 m = w[0];
		if ((m >> i) & 1u) {
			const u32 r = (u32)__builtin_popcount(m & ((1u << i) - 1u));
#pragma unroll
			for (u32 k = 0; k < 8; ++k)
				if (k == r)
					q.old = w[1 + k];
		}
	}
	q.cx = (E.x << 1) | (i & 1u), q.cy = (E.y << 1) | ((i >> 1) & 1u), q.cz = (E.z << 1) | ((i >> 2) & 1u);
	q.octree = E.octree;
	if (P.stateful && l + 1 <= P.c.color_leaf)
		q.octree = color_child(P.cnodes, E.octree, i);
	q.sf_in = i | (E.sf & (kSfFinal | kSfWriter));
	return q;
}
// A list entry's child, folded (list_chain). pre: the list's last edit that resets the child, in
// pre->single, whether some edit covers it whole, in pre->cnt, and pre2: the edits touching it from
// the chain's stop on (count, first, any stateful), as a wavefront found them (coop_node); null to
// scan the list here.
struct ListHits {
	u32 cnt, single, active;
};
// A list entry's child after folding its list (list_chain): the state before its first remaining
// edit (pointer, octree state, flags), that edit's list index k (llen: none remains, the child is
// terminal) and the last color a whole-node edit pushed on it.
struct ChainRes {
	Classified c;
	u32 ptr, k;
	bool pushed;
	Color pcol;
};
// From list index rs (the last edit that fills or clears the child whole, or 0): each edit that
// touches the child is applied in turn while it covers the child whole (classify_single: fill,
// clear, whole-node paint, or nothing); the first one that crosses it stops the chain. A child whose
// list holds more than one edit starts without the is_final bit: each edit of a folded list runs
// its own traversal.
__device__ inline ChainRes list_chain(const EditParams &P, const ALevel &A, const Entry &E, u32 rs, u32 cl, u32 cx,
                                      u32 cy, u32 cz, u32 old, u32 octree, u32 sf_in) {
	ChainRes r{Classified{kNotAffected, octree, sf_in & ~kSfFinal, false, Color::none()}, old, E.llen, false,
	           Color::none()};
	for (u32 k = rs; k < E.llen; ++k) {
		const DevEditor ed = P.editor(ldg(A.lists, (u64)E.loff + k));
		if (ed.geom(cl, cx, cy, cz) == kNotAffected)
			continue;
		const Classified s = classify_single(P, ed, cl, cx, cy, cz, r.ptr, r.c.octree, r.c.sf & ~kSfFinal);
		if (s.t == kProceed) {
			r.k = k;
			return r;
		}
		r.ptr = s.t == kClear ? kNull : (s.t == kFill ? P.filled[cl] : r.ptr);
		r.c.octree = s.octree;
		r.c.sf = s.sf & ~kSfFinal;
		if (s.push)
			r.pushed = true, r.pcol = s.col;
	}
	return r;
}
__device__ inline PairOut expand_pair_e(const EditParams &P, u32 l, const ALevel &A, u64 t, const Entry &E,
                                        PushEvent *__restrict__ ev, u32 *__restrict__ pm, u32 ev_cap,
                                        u32 *bits = nullptr, const ListHits *pre = nullptr,
                                        const ListHits *pre2 = nullptr) {
	const u32 i = (u32)(t & 7u);
	const PairIn q = pair_in(P, l, E, i);
	const u32 old = q.old, cl = l + 1, leafL = P.c.color_leaf, cx = q.cx, cy = q.cy, cz = q.cz;
	const u32 octree = q.octree, sf_in = q.sf_in;
	Classified c{kNotAffected, octree, sf_in, false, Color::none()};
	u32 cnt = 1, single = E.loff, kid_base = old;
	if (E.llen == 1)
		c = classify_single(P, P.editor(E.loff), cl, cx, cy, cz, old, octree, sf_in);
	else {
		// the folded list: the last edit that resets the child (its touching edits marked in bits);
		// without an edit that covers the child whole, the scan's count is the list's (no chain)
		u32 rs = 0, cnt0 = 0, single0 = 0;
		bool full = true, active0 = false;
		if (pre)
			rs = pre->single, full = pre->cnt != 0u; // coop_reset's results
		else {
			full = false;
			if (bits)
				for (u32 w = 0; w < (E.llen + 31u) >> 5; ++w)
					bits[w] = 0u;
			for (u32 k0 = 0; k0 < E.llen; k0 += kListBatch) { // list words of a batch in one round trip
				u32 ix[kListBatch];
#pragma unroll
				for (u32 j = 0; j < kListBatch; ++j)
					ix[j] = k0 + j < E.llen ? ldg(A.lists, (u64)E.loff + k0 + j) : 0u;
#pragma unroll
				for (u32 j = 0; j < kListBatch; ++j)
					if (k0 + j < E.llen) {
						const DevEditor ed = P.editor(ix[j]);
						const EditType g = ed.geom(cl, cx, cy, cz);
						if (g != kNotAffected) {
							if (bits)
								bits[(k0 + j) >> 5] |= 1u << ((k0 + j) & 31u);
							if (cnt0 == 0)
								single0 = ix[j];
							++cnt0;
							active0 |= ed.stateful();
						}
						if (g == kFill) {
							full = true;
							if (!ed.is_paint())
								rs = k0 + j;
						}
					}
			}
		}
		// without a whole-node edit every touching edit remains: the list starts at index 0 (the count
		// is the wavefront's when it scanned the list, pre2, else this thread's)
		const u32 ntouch = pre2 ? pre2->cnt : cnt0;
		const ChainRes ch = full ? list_chain(P, A, E, rs, cl, cx, cy, cz, old, octree, sf_in)
		                         : ChainRes{Classified{kNotAffected, octree, sf_in & ~kSfFinal, false, Color::none()},
		                                    old, ntouch ? 0u : E.llen, false, Color::none()};
		kid_base = ch.ptr;
		cnt = 0;
		bool active = false;
		if (!full) { // every touching edit remains
			if (pre2)
				cnt = pre2->cnt, single = pre2->single, active = pre2->active != 0u;
			else
				cnt = cnt0, single = single0, active = active0;
		} else if (ch.k < E.llen) { // the remaining edits: those that touch the child from list index ch.k on
			if (pre2) {
				cnt = pre2->cnt;
				single = pre2->single;
				active = pre2->active != 0u;
			} else
				for (u32 k0 = ch.k; k0 < E.llen; k0 += kListBatch) {
					u32 ix[kListBatch];
#pragma unroll
					for (u32 j = 0; j < kListBatch; ++j)
						ix[j] = k0 + j < E.llen ? ldg(A.lists, (u64)E.loff + k0 + j) : 0u;
#pragma unroll
					for (u32 j = 0; j < kListBatch; ++j)
						if (k0 + j < E.llen) {
							const DevEditor ed = P.editor(ix[j]);
							if (ed.geom(cl, cx, cy, cz) != kNotAffected) {
								if (cnt == 0)
									single = ix[j];
								++cnt;
								active |= ed.stateful();
							}
						}
				}
			if (bits) // the child's list: the touching edits from ch.k on
				for (u32 w = 0; w < (E.llen + 31u) >> 5; ++w)
					bits[w] &= w < (ch.k >> 5) ? 0u : (w == (ch.k >> 5) ? ~((1u << (ch.k & 31u)) - 1u) : ~0u);
		}
		if (cnt == 1) { // one edit remains: its own classification from the chain's state
			c = classify_single(P, P.editor(single), cl, cx, cy, cz, ch.ptr, ch.c.octree, ch.c.sf);
		} else if (cnt > 1) {
			c = ch.c;
			c.t = kProceed;
			if (active)
				c.sf |= kSfActive | (cl == leafL ? (kSfWriter | kSfNewJob) : 0u);
			single = ch.k; // a list child's list starts at this index of the parent's
		} else { // the chain covered the child: terminal
			c = ch.c;
			c.t = ch.ptr == old ? kNotAffected : (ch.ptr == kNull ? kClear : kFill);
		}
		if (ch.pushed && !c.push) // the chain's last whole-node color, under what remains
			c.push = true, c.col = ch.pcol;
	}
	if (c.push && A.pev) { // the job rebuilds its colors from its pairs' records
		const u32 d = cl - leafL, msk = (1u << d) - 1u, sh = 3u * (P.c.voxel_level - cl);
		stg(A.pev, t, make_uint4(morton3(cx & msk, cy & msk, cz & msk) << sh, c.col.colors, c.col.wb, 0u));
	} else if (c.push) { // one bump of the shared event counter per wavefront (the active pushing lanes)
		u32 d = cl - leafL, msk = (1u << d) - 1u, sh = 3u * (P.c.voxel_level - cl);
		const u64 m = __ballot(1);
		const u32 lane = __lane_id(), leader = (u32)__builtin_ctzll(m);
		u32 k0 = 0;
		if (lane == leader)
			k0 = atomicAdd(&pm[kPmEvents], (u32)__popcll(m));
		const u32 k = (u32)__shfl((int)k0, (int)leader) + (u32)__popcll(m & ((1ull << lane) - 1ull));
		if (k < ev_cap)
			ev[k] = PushEvent{E.job, morton3(cx & msk, cy & msk, cz & msk) << sh, 1u << sh, c.col.colors, c.col.wb, 0, 0, 0};
		else
			atomicOr(&pm[kPmOverflow], 1u);
	}
	const u32 kid = c.t == kClear ? kNull : (c.t == kFill ? P.filled[cl] : kid_base);
	stg(A.kids, t, kid); // the child's starting pointer for proceeding and untouched slots
	if (A.okids)
		stg(A.okids, t, c.octree);
	bool proceed = c.t == kProceed;
	// bit 0: proceeds, bit 1: the child's pointer changed (a terminal result, or the state a folded
	// list leaves a proceeding child in: the join compares proceeding results with kids, so it must
	// know the slot changed even when the result equals that state), bit 2: pushed a color
	u32 flag = (proceed ? 1u : 0u) | (kid != old ? 2u : 0u) | (c.push ? 4u : 0u);
	return PairOut{flag, proceed && cnt > 1 ? cnt : 0u, make_uint4(c.octree, c.sf, single, cnt), kid};
}
__device__ inline PairOut expand_pair(const EditParams &P, u32 l, const ALevel &A, u64 t, PushEvent *__restrict__ ev,
                                      u32 *__restrict__ pm, u32 ev_cap) {
	return expand_pair_e(P, l, A, t, A.ent[(u32)(t >> 3)], ev, pm, ev_cap);
}

// Writes the child entry of a proceeding pair (k_emit). m8: proceed flags of the pair's node in
// bits 0-7, its changed children in bits 8-15 and its pushed children in bits 16-23 (valid
// on slot 0), idx / lo: the child's entry index and list offset at level l + 1.
// bits: expand_pair_e's record of the list edits that touch the child, or null to test them again.
// by_wave: the list is written by coop_emit (only the bound is checked here).
__device__ inline Entry emit_pair_e(const EditParams &P, u32 l, const ALevel &A, const ALevel &N,
                                    u32 *__restrict__ pm, u64 t, const Entry &E, u32 kid, u32 m8, const PairOut &o,
                                    u32 idx, u32 lo, const u32 *bits = nullptr, bool by_wave = false) {
	u32 e = (u32)(t >> 3), i = (u32)(t & 7u);
	if (i == 0) {
		stg(A.first_child, e, idx);
		stg(A.pmask, e, m8);
	}
	Entry out{};
	if (!(o.flag & 1u))
		return out;
	u32 sf = o.cs.y;
	u32 job = (sf & kSfNewJob) ? idx : ((sf & kSfWriter) ? E.job : kNull);
	const u32 cx = (E.x << 1) | (i & 1u), cy = (E.y << 1) | ((i >> 1) & 1u), cz = (E.z << 1) | ((i >> 2) & 1u);
	u32 ln = o.cs.w;
	if (ln > 1) { // the child's filtered list, in the parent's (submission) order
		if ((u64)lo + ln > N.list_cap) {
			atomicOr(&pm[kPmOverflow], 1u);
			ln = 0;
		} else if (!by_wave) {
			u32 w = lo;
			for (u32 k0 = 0; k0 < E.llen; k0 += kListBatch) { // loads and tests first, then the stores
				u32 ix[kListBatch], keep = 0;
#pragma unroll
				for (u32 j = 0; j < kListBatch; ++j)
					ix[j] = k0 + j < E.llen ? ldg(A.lists, (u64)E.loff + k0 + j) : 0u;
				static_assert(32 % kListBatch == 0, "a batch's bits sit in one word");
				if (bits)
					keep = (bits[k0 >> 5] >> (k0 & 31u)) & ((1u << kListBatch) - 1u);
				else // the touching edits from the fold's stop (cs.z) on
#pragma unroll
					for (u32 j = 0; j < kListBatch; ++j)
						if (k0 + j < E.llen && k0 + j >= o.cs.z && P.editor(ix[j]).geom(l + 1, cx, cy, cz) != kNotAffected)
							keep |= 1u << j;
#pragma unroll
				for (u32 j = 0; j < kListBatch; ++j)
					if ((keep >> j) & 1u)
						stg(N.lists, (u64)w++, ix[j]);
			}
		}
	} else
		lo = o.cs.z;
	out = Entry{cx, cy, cz, kid, e, sf & ~kSfNewJob, o.cs.x, job, lo, ln};
	if (idx < N.cap)
		stg(N.ent, idx, out);
	else
		atomicOr(&pm[kPmOverflow], 1u);
	return out;
}
__device__ inline void emit_pair(const EditParams &P, u32 l, const ALevel &A, const ALevel &N, u32 *__restrict__ pm,
                                 u64 t, u32 m8, const PairOut &o, u32 idx, u32 lo, bool by_wave = false) {
	if (o.flag & 1u)
		emit_pair_e(P, l, A, N, pm, t, A.ent[(u32)(t >> 3)], A.kids[t], m8, o, idx, lo, nullptr, by_wave);
	else
		emit_pair_e(P, l, A, N, pm, t, Entry{}, kNull, m8, o, idx, lo);
}

// Edit lists longer than kCoopList are split by a whole wavefront: lane j takes list entries
// j, j + 64, ... and tests each against all 8 children, so a list of n edits costs n / 64 coalesced
// loads (kCoopBatch of them in flight) instead of n / kListBatch dependent ones per pair. The
// callers keep the wavefront converged (the loops below are uniform).
constexpr u32 kCoopList = 16, kCoopBatch = 4;
__device__ inline u32 child_coord(u32 p, u32 i, u32 axis) { return (p << 1) | ((i >> axis) & 1u); }
// Per child i of E: the list edits that touch it from index ks[i] on (cnt[i]), the first of them
// (single[i]), and whether one of them is stateful (bit i of active).
struct CoopHits {
	u32 cnt[8], single[8], active;
};
// Per child i of E: the list index of the last edit that fills or clears it whole (0 if none), and
// in bit i of full whether some edit covers it whole (a paint too).
__device__ inline void coop_reset(const EditParams &P, u32 l, const ALevel &A, const Entry &E, u32 lane,
                                  u32 (&rs)[8], u32 &full) {
	full = 0u;
#pragma unroll
	for (u32 i = 0; i < 8; ++i)
		rs[i] = 0u;
	const u32 cl = l + 1u;
	for (u32 k0 = 0; k0 < E.llen; k0 += 64u * kCoopBatch) {
		u32 ix[kCoopBatch];
#pragma unroll
		for (u32 j = 0; j < kCoopBatch; ++j) {
			const u32 k = k0 + 64u * j + lane;
			ix[j] = k < E.llen ? ldg(A.lists, (u64)E.loff + k) : 0u;
		}
#pragma unroll
		for (u32 j = 0; j < kCoopBatch; ++j) {
			if (k0 + 64u * j >= E.llen) // uniform
				break;
			const bool v = k0 + 64u * j + lane < E.llen;
			const DevEditor ed = P.editor(ix[j]);
			const bool paint = ed.is_paint();
#pragma unroll
			for (u32 i = 0; i < 8; ++i) {
				const bool f =
				    v && ed.geom(cl, child_coord(E.x, i, 0), child_coord(E.y, i, 1), child_coord(E.z, i, 2)) == kFill;
				const u64 m = __ballot(f && !paint);
				if (m)
					rs[i] = k0 + 64u * j + 63u - (u32)__builtin_clzll(m);
				if (__ballot(f))
					full |= 1u << i;
			}
		}
	}
}
__device__ inline void coop_count(const EditParams &P, u32 l, const ALevel &A, const Entry &E, u32 lane,
                                  const u32 (&ks)[8], CoopHits &o) {
#pragma unroll
	for (u32 i = 0; i < 8; ++i)
		o.cnt[i] = 0u, o.single[i] = 0u;
	o.active = 0u;
	const u32 cl = l + 1u;
	u32 kmin = ks[0];
#pragma unroll
	for (u32 i = 1; i < 8; ++i)
		kmin = min(kmin, ks[i]);
	for (u32 k0 = kmin & ~63u; k0 < E.llen; k0 += 64u * kCoopBatch) {
		u32 ix[kCoopBatch];
#pragma unroll
		for (u32 j = 0; j < kCoopBatch; ++j) {
			const u32 k = k0 + 64u * j + lane;
			ix[j] = k < E.llen ? ldg(A.lists, (u64)E.loff + k) : 0u;
		}
#pragma unroll
		for (u32 j = 0; j < kCoopBatch; ++j) {
			if (k0 + 64u * j >= E.llen) // uniform
				break;
			const u32 k = k0 + 64u * j + lane;
			const bool v = k < E.llen;
			const DevEditor ed = P.editor(ix[j]);
			const bool st = ed.stateful();
#pragma unroll
			for (u32 i = 0; i < 8; ++i) {
				const bool hit = v && k >= ks[i] &&
				                 ed.geom(cl, child_coord(E.x, i, 0), child_coord(E.y, i, 1), child_coord(E.z, i, 2)) !=
				                     kNotAffected;
				const u64 m = __ballot(hit);
				if (m && o.cnt[i] == 0u)
					o.single[i] = __builtin_amdgcn_readlane(ix[j], (int)__builtin_ctzll(m));
				o.cnt[i] += (u32)__popcll(m);
				if (__ballot(hit && st))
					o.active |= 1u << i;
			}
		}
	}
}
__device__ inline ListHits coop_pick(const CoopHits &c, u32 i) {
	ListHits h{0u, 0u, (c.active >> i) & 1u};
#pragma unroll
	for (u32 k = 0; k < 8; ++k)
		if (k == i)
			h.cnt = c.cnt[k], h.single = c.single[k];
	return h;
}
// The whole wave folds the list of its node e (lanes 8 e .. 8 e + 7 hold the node's pairs): the
// last reset of every child, the chain on each pair's lane (list_chain, as expand_pair_e repeats
// it), then the touching edits from each chain's stop. h1 / h2: expand_pair_e's pre / pre2.
__device__ inline void coop_node(const EditParams &P, u32 l, const ALevel &A, const Entry &E, u32 e, u32 lane,
                                 ListHits &h1, ListHits &h2, bool &by_wave) {
	u32 rs[8], full;
	coop_reset(P, l, A, E, lane, rs, full);
	const bool mine = (lane >> 3) == e;
	const u32 i = lane & 7u;
	const bool my_full = (full >> i) & 1u;
	u32 my_rs = 0, kk = E.llen;
#pragma unroll
	for (u32 k = 0; k < 8; ++k)
		my_rs = k == i ? rs[k] : my_rs;
	if (mine) {
		if (my_full) {
			const PairIn q = pair_in(P, l, E, i);
			kk = list_chain(P, A, E, my_rs, l + 1u, q.cx, q.cy, q.cz, q.old, q.octree, q.sf_in).k;
		} else
			kk = 0; // no edit covers the child whole: all its touching edits remain
	}
	u32 ks[8];
#pragma unroll
	for (u32 k = 0; k < 8; ++k)
		ks[k] = __builtin_amdgcn_readlane(kk, (int)(8u * e + k));
	CoopHits co;
	coop_count(P, l, A, E, lane, ks, co);
	if (mine) {
		h1 = ListHits{my_full ? 1u : 0u, my_rs, 0u};
		h2 = coop_pick(co, i);
		by_wave = true;
	}
}
// The children's lists of E in list (submission) order: child i's at N.lists[lo[i]...] when bit i
// of need is set (a proceeding child with more than one edit whose list fits the level's bound):
// the edits that touch it from list index ks[i] on.
__device__ inline void coop_emit(const EditParams &P, u32 l, const ALevel &A, const ALevel &N, const Entry &E,
                                 u32 lane, const u32 (&lo)[8], const u32 (&ks)[8], u32 need) {
	const u32 cl = l + 1u;
	const u64 below = (1ull << lane) - 1ull;
	u32 run[8] = {0, 0, 0, 0, 0, 0, 0, 0};
	for (u32 k0 = 0; k0 < E.llen; k0 += 64u * kCoopBatch) {
		u32 ix[kCoopBatch];
#pragma unroll
		for (u32 j = 0; j < kCoopBatch; ++j) {
			const u32 k = k0 + 64u * j + lane;
			ix[j] = k < E.llen ? ldg(A.lists, (u64)E.loff + k) : 0u;
		}
#pragma unroll
		for (u32 j = 0; j < kCoopBatch; ++j) {
			if (k0 + 64u * j >= E.llen)
				break;
			const bool v = k0 + 64u * j + lane < E.llen;
			const DevEditor ed = P.editor(ix[j]);
#pragma unroll
			for (u32 i = 0; i < 8; ++i) {
				if (!((need >> i) & 1u))
					continue;
				const bool hit = v && k0 + 64u * j + lane >= ks[i] &&
				                 ed.geom(cl, child_coord(E.x, i, 0), child_coord(E.y, i, 1), child_coord(E.z, i, 2)) !=
				                     kNotAffected;
				const u64 m = __ballot(hit);
				if (hit)
					stg(N.lists, (u64)lo[i] + run[i] + (u32)__popcll(m & below), ix[j]);
				run[i] += (u32)__popcll(m);
			}
		}
	}
}
// chunk of the pair range [0, 8n) owned by this block (whole nodes per block)
__device__ inline void pair_chunk(u32 n, u64 &begin, u64 &end) {
	u64 P8 = (u64)n * 8u, chunk = (P8 + gridDim.x - 1) / gridDim.x;
	chunk = (chunk + 7u) & ~7ull;
	begin = (u64)blockIdx.x * chunk;
	if (begin > P8)
		begin = P8;
	end = begin + chunk < P8 ? begin + chunk : P8;
}

// Expands the pairs [begin, end) tile by tile and emits the proceeding children, starting at
// child index base (a block-wide helper; tiles start on node boundaries).
__device__ inline Pair expand_emit_range(const EditParams &P, u32 l, const ALevel &A, const ALevel &N,
                                         u32 *__restrict__ pm, PushEvent *ev, u32 ev_cap, u64 begin, u64 end, Pair base,
                                         const u32 *__restrict__ flag, const u32 *__restrict__ lcnt,
                                         const uint4 *__restrict__ cstate, Pair *sh, u32 *sflag) {
	const u32 lane = threadIdx.x & 63u;
	for (u64 t0 = begin; t0 < end; t0 += kAT) {
		u64 t = t0 + threadIdx.x;
		bool in = t < end;
		PairOut o{0, 0, make_uint4(0, 0, 0, 0)};
		bool by_wave = false; // the node's list is longer than kCoopList: the wave writes its children's
		if (in) {
			if (flag)
				o = PairOut{flag[t], lcnt[t], cstate[t]};
			else
				o = expand_pair(P, l, A, t, ev, pm, ev_cap);
			by_wave = ldg(&A.ent[(u32)(t >> 3)].llen, 0) > kCoopList;
		}
		sflag[threadIdx.x] = o.flag;
		Pair tile;
		Pair off = block_excl_scan2(Pair{o.flag & 1u, o.lcnt}, sh, tile); // synchronizes sflag
		if (in) {
			u32 m8 = 0;
			if ((t & 7u) == 0)
				for (u32 k = 0; k < 8; ++k) {
					u32 f = sflag[threadIdx.x + k];
					m8 |= ((f & 1u) << k) | (((f >> 1) & 1u) << (8 + k)) | (((f >> 2) & 1u) << (16 + k));
				}
			emit_pair(P, l, A, N, pm, t, m8, o, base.a + off.a, base.b + off.b, by_wave);
		}
		{ // the long lists' children, by the wave whose lanes hold the node's pairs (whole nodes per wave)
			const u32 lo = base.b + off.b;
			const u64 need = __ballot(in && by_wave && o.lcnt > 0u && (u64)lo + o.lcnt <= N.list_cap);
			const u64 wnode = (t0 + (threadIdx.x & ~63u)) >> 3;
			for (u32 e = 0; e < 8u; ++e) {
				const u32 need8 = (u32)(need >> (8u * e)) & 0xFFu;
				if (!need8)
					continue;
				u32 los[8], ks[8];
#pragma unroll
				for (u32 i = 0; i < 8; ++i) {
					los[i] = __builtin_amdgcn_readlane(lo, (int)(8u * e + i));
					ks[i] = __builtin_amdgcn_readlane(o.cs.z, (int)(8u * e + i));
				}
				coop_emit(P, l, A, N, A.ent[(u32)(wnode + e)], lane, los, ks, need8);
			}
		}
		base.a += tile.a;
		base.b += tile.b;
		__syncthreads(); // sflag reuse
	}
	return base;
}

__device__ inline void set_level_count(const EditParams &P, u32 l, const ALevel &N, u32 *pm, Pair total) {
	const u32 L = P.c.L;
	pm[kPmHdr + l + 1] = min(total.a, N.cap);
	pm[kPmHdr + L + l + 1] = total.b;
	if (total.a > N.cap)
		pm[kPmOverflow] |= 1u;
}

// Wave 0 (all 64 lanes): the root's folded edit list (as list_chain folds a child's): the last
// edit that fills or clears the whole grid, the whole-grid edits after it applied in turn, then the
// edits from the first that crosses the root on, 64 per step in submission order; lane 0 then
// classifies the root.
__device__ inline void root_entry(const EditParams &P, const ALevel &A0, u32 *__restrict__ roots, u32 *__restrict__ pm,
                                  u32 lane) {
	const u32 root = roots[0], croot = roots[1], L = P.c.L, n = P.n_edits;
	ChainRes ch{Classified{kNotAffected, croot, 0u, false, Color::none()}, root, 0u, false, Color::none()};
	u32 cnt = 0, single = 0;
	bool active = false;
	if (n == 1)
		cnt = 1;
	else {
		u32 rs = 0;
		for (u32 k0 = 0; k0 < n; k0 += 64u) {
			const u32 k = k0 + lane;
			bool z = false;
			if (k < n) {
				const DevEditor e = P.editor(k);
				z = !e.is_paint() && e.geom(0, 0, 0, 0) == kFill;
			}
			const u64 m = __ballot(z);
			if (m)
				rs = k0 + 63u - (u32)__builtin_clzll(m);
		}
		// the chain on every lane (the same values everywhere)
		ch.k = n;
		for (u32 k = rs; k < n; ++k) {
			const DevEditor ed = P.editor(k);
			if (ed.geom(0, 0, 0, 0) == kNotAffected)
				continue;
			const Classified sc = classify_single(P, ed, 0, 0, 0, 0, ch.ptr, ch.c.octree, ch.c.sf);
			if (sc.t == kProceed) {
				ch.k = k;
				break;
			}
			ch.ptr = sc.t == kClear ? kNull : (sc.t == kFill ? P.filled[0] : ch.ptr);
			ch.c.octree = sc.octree;
			ch.c.sf = sc.sf & ~kSfFinal;
		}
		for (u32 k0 = ch.k & ~63u; k0 < n; k0 += 64u) {
			const u32 k = k0 + lane;
			bool hit = false, st = false;
			if (k < n && k >= ch.k) {
				const DevEditor e = P.editor(k);
				hit = e.geom(0, 0, 0, 0) != kNotAffected;
				st = e.stateful();
			}
			const u64 m = __ballot(hit);
			const u32 at = cnt + (u32)__popcll(m & ((1ull << lane) - 1ull));
			if (hit && at < A0.list_cap)
				A0.lists[at] = k;
			if (m && cnt == 0)
				single = k0 + (u32)__builtin_ctzll(m);
			cnt += (u32)__popcll(m);
			active |= __ballot(hit && st) != 0ull;
		}
	}
	if (lane != 0)
		return;
	Classified c = ch.c;
	if (cnt == 1)
		c = classify_single(P, P.editor(single), 0, 0, 0, 0, ch.ptr, ch.c.octree, ch.c.sf);
	else if (cnt > 1) {
		c.t = kProceed;
		if (active)
			c.sf |= kSfActive | (P.c.color_leaf == 0 ? (kSfWriter | kSfNewJob) : 0u);
		if (cnt > A0.list_cap)
			pm[kPmOverflow] |= 1u;
	} else
		c.t = ch.ptr == root ? kNotAffected : (ch.ptr == kNull ? kClear : kFill);
	if (c.t == kProceed) {
		A0.ent[0] = Entry{0, 0, 0, ch.ptr, kNull, c.sf & ~kSfNewJob, c.octree, (c.sf & kSfWriter) ? 0u : kNull,
		                  cnt == 1 ? single : 0u, cnt};
		pm[kPmHdr] = 1;
		pm[kPmHdr + L] = cnt > 1 ? cnt : 0u;
	} else {
		roots[0] = c.t == kClear ? kNull : (c.t == kFill ? P.filled[0] : ch.ptr);
		if (P.stateful)
			roots[1] = c.octree;
	}
}

// Single workgroup: clears the pass counters, classifies the root and expands levels [0, l_end)
// (every one small enough for one block), with the current level's entries and count in LDS.
constexpr u32 kTopEnt = kTopPairs / 8;
constexpr u32 kTopEdits = 1024; // a pass with at most this many edits reads them from LDS in a_top
// kLists: the pass has more than one edit (nodes may carry edit lists).
template <bool kLists>
__global__ void __launch_bounds__(kAT) a_top(EditParams P0, const ALevel *__restrict__ lv, u32 *__restrict__ roots,
                                             u32 *__restrict__ pm, u32 pm_words, u32 l_end, PushEvent *ev, u32 ev_cap,
                                             u32 *__restrict__ gen_ctr) {
	__shared__ Pair sh[kAT];
	__shared__ u32 sflag[kAT];
	__shared__ Entry sent[2][kTopEnt];
	__shared__ u32 scount;
	__shared__ hd_edit_desc sedit[kTopEdits];
	__shared__ u32 sfilled[HD_MAX_LEVELS];
	__shared__ u32 sbits[kLists ? kAT : 1][kTopEdits / 32]; // per thread: the list edits touching its child
	for (u32 i = threadIdx.x; i < pm_words; i += kAT)
		pm[i] = 0;
	// the pass's edits and the filled chain are read by every classification: keep them in LDS
	EditParams P = P0;
	if (P.n_edits <= kTopEdits) {
		const u32 *src = reinterpret_cast<const u32 *>(P.edits);
		u32 *dst = reinterpret_cast<u32 *>(sedit);
		for (u32 i = threadIdx.x; i < P.n_edits * (u32)(sizeof(hd_edit_desc) / 4); i += kAT)
			dst[i] = src[i];
		P.edits = sedit;
	}
	if (threadIdx.x < P.c.L)
		sfilled[threadIdx.x] = P.filled[threadIdx.x];
	P.filled = sfilled;
	__syncthreads();
	if (threadIdx.x < 64) {
		root_entry(P, lv[0], roots, pm, threadIdx.x);
		if (threadIdx.x == 0) {
			pm[kPmGen] = atomicAdd(gen_ctr, P.c.L) + 1u; // this pass's dedup generations: pm[kPmGen] + level
			scount = pm[kPmHdr];
			if (scount)
				sent[0][0] = lv[0].ent[0];
		}
	}
	__syncthreads();
	// levels of at most 8 entries (64 pairs) by wave 0 alone: ballots and a wave scan replace the
	// block scan and its barriers
	__shared__ u32 s_level;
	if (threadIdx.x < 64) {
		const u32 lane = threadIdx.x;
		u32 l = 0, n = scount;
		for (; l < l_end && n <= 8u; ++l) {
			const ALevel A = lv[l], N = lv[l + 1];
			const Entry *cur = sent[l & 1];
			Entry *nxt = sent[(l + 1) & 1];
			const bool in = lane < 8u * n;
			PairOut o{0, 0, make_uint4(0, 0, 0, 0), kNull};
			Entry E{};
			u32 *bits = nullptr;
			// long lists: the whole wave classifies each such node's children (lane 8 e + i is pair i of e)
			ListHits hits{0u, 0u, 0u}, hits2{0u, 0u, 0u};
			bool by_wave = false;
			if (kLists)
				for (u32 e = 0; e < n; ++e) {
					const Entry Ee = cur[e];
					if (Ee.llen > kCoopList)
						coop_node(P, l, A, Ee, e, lane, hits, hits2, by_wave);
				}
			if (in) {
				E = cur[lane >> 3];
				bits = kLists && !by_wave && E.llen <= kTopEdits ? sbits[lane] : nullptr;
				o = expand_pair_e(P, l, A, lane, E, ev, pm, ev_cap, bits, by_wave ? &hits : nullptr,
				                  by_wave ? &hits2 : nullptr);
			}
			const u64 bp = __ballot(o.flag & 1u), bt = __ballot((o.flag >> 1) & 1u), bu = __ballot((o.flag >> 2) & 1u);
			const u64 below = (1ull << lane) - 1ull;
			u32 lc = o.lcnt; // inclusive wave scan of the list lengths (a single-edit pass has none)
			if (kLists)
#pragma unroll
				for (u32 d = 1; d < 64; d <<= 1) {
					const u32 y = __shfl_up(lc, d);
					if (lane >= d)
						lc += y;
				}
			const Pair tot{(u32)__popcll(bp), kLists ? (u32)__shfl(lc, 63) : 0u};
			if (in) {
				const u32 m8 = (lane & 7u) == 0 ? (u32)((bp >> lane) & 0xFFu) | ((u32)((bt >> lane) & 0xFFu) << 8) |
				                                      ((u32)((bu >> lane) & 0xFFu) << 16)
				                                : 0u;
				const u32 idx = (u32)__popcll(bp & below);
				Entry ne = emit_pair_e(P, l, A, N, pm, lane, E, o.kid, m8, o, idx, lc - o.lcnt, bits, by_wave);
				if ((o.flag & 1u) && idx < kTopEnt)
					nxt[idx] = ne;
			}
			if (kLists) { // the long lists' children, by the whole wave
				const u32 lo = lc - o.lcnt;
				const u64 need = __ballot(by_wave && o.lcnt > 0u && (u64)lo + o.lcnt <= N.list_cap);
				for (u32 e = 0; e < n; ++e) {
					const u32 need8 = (u32)(need >> (8u * e)) & 0xFFu;
					if (!need8)
						continue;
					u32 los[8], ks[8];
#pragma unroll
					for (u32 i = 0; i < 8; ++i) {
						los[i] = __builtin_amdgcn_readlane(lo, (int)(8u * e + i));
						ks[i] = __builtin_amdgcn_readlane(o.cs.z, (int)(8u * e + i));
					}
					coop_emit(P, l, A, N, cur[e], lane, los, ks, need8);
				}
			}
			if (lane == 0)
				set_level_count(P, l, N, pm, tot);
			n = min(tot.a, N.cap);
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
		}
		if (lane == 0) {
			s_level = l;
			scount = n;
		}
	}
	__syncthreads();
	for (u32 l = s_level; l < l_end; ++l) {
		const ALevel A = lv[l], N = lv[l + 1];
		const u32 n = min(scount, kTopEnt);
		const Entry *cur = sent[l & 1];
		Entry *nxt = sent[(l + 1) & 1];
		Pair base{0, 0};
		for (u64 t0 = 0; t0 < (u64)n * 8u; t0 += kAT) {
			const u64 t = t0 + threadIdx.x;
			const bool in = t < (u64)n * 8u;
			PairOut o{0, 0, make_uint4(0, 0, 0, 0), kNull};
			Entry E{};
			u32 *bits = nullptr;
			// long lists: each wave classifies the children of the 8 nodes whose pairs are its lanes
			const u32 lane = threadIdx.x & 63u;
			const u64 wnode = (t0 + (threadIdx.x & ~63u)) >> 3; // the wave's first node
			ListHits hits{0u, 0u, 0u}, hits2{0u, 0u, 0u};
			bool by_wave = false;
			if (kLists)
				for (u32 e = 0; e < 8u && wnode + e < n; ++e) {
					const Entry Ee = cur[wnode + e];
					if (Ee.llen > kCoopList)
						coop_node(P, l, A, Ee, e, lane, hits, hits2, by_wave);
				}
			if (in) {
				E = cur[t >> 3];
				bits = kLists && !by_wave && E.llen <= kTopEdits ? sbits[threadIdx.x] : nullptr;
				o = expand_pair_e(P, l, A, t, E, ev, pm, ev_cap, bits, by_wave ? &hits : nullptr,
				                  by_wave ? &hits2 : nullptr);
			}
			sflag[threadIdx.x] = o.flag;
			Pair tile;
			Pair off = block_excl_scan2(Pair{o.flag & 1u, o.lcnt}, sh, tile); // synchronizes sflag
			if (in) {
				u32 m8 = 0;
				if ((t & 7u) == 0)
					for (u32 k = 0; k < 8; ++k) {
						u32 f = sflag[threadIdx.x + k];
						m8 |= ((f & 1u) << k) | (((f >> 1) & 1u) << (8 + k)) | (((f >> 2) & 1u) << (16 + k));
					}
				const u32 idx = base.a + off.a;
				Entry ne = emit_pair_e(P, l, A, N, pm, t, E, o.kid, m8, o, idx, base.b + off.b, bits, by_wave);
				if ((o.flag & 1u) && idx < kTopEnt)
					nxt[idx] = ne;
			}
			if (kLists) { // the long lists' children, by the wave that classified them
				const u32 lo = base.b + off.b;
				const u64 need = __ballot(by_wave && o.lcnt > 0u && (u64)lo + o.lcnt <= N.list_cap);
				for (u32 e = 0; e < 8u && wnode + e < n; ++e) {
					const u32 need8 = (u32)(need >> (8u * e)) & 0xFFu;
					if (!need8)
						continue;
					u32 los[8], ks[8];
#pragma unroll
					for (u32 i = 0; i < 8; ++i) {
						los[i] = __builtin_amdgcn_readlane(lo, (int)(8u * e + i));
						ks[i] = __builtin_amdgcn_readlane(o.cs.z, (int)(8u * e + i));
					}
					coop_emit(P, l, A, N, cur[wnode + e], lane, los, ks, need8);
				}
			}
			base.a += tile.a;
			base.b += tile.b;
			__syncthreads(); // sflag reuse
		}
		if (threadIdx.x == 0) {
			set_level_count(P, l, N, pm, base);
			scount = min(base.a, N.cap);
		}
		__syncthreads();
	}
}

// Large levels: a_count (classify, block totals) then a_emit (block prefix, stable emit).
__global__ void __launch_bounds__(kAT) a_count(EditParams P, u32 l, const ALevel *__restrict__ lv, u32 *__restrict__ pm,
                                               u32 *__restrict__ flag, u32 *__restrict__ lcnt,
                                               uint4 *__restrict__ cstate, PushEvent *ev, u32 ev_cap,
                                               Pair *__restrict__ bsum) {
	__shared__ Pair sh[kAT];
	const ALevel A = lv[l];
	u64 begin, end;
	pair_chunk(pm[kPmHdr + l], begin, end);
	Pair acc{0, 0};
	const u32 lane = threadIdx.x & 63u;
	for (u64 t0 = begin; t0 < end; t0 += kAT) { // uniform trip count: waves classify long lists together
		const u64 t = t0 + threadIdx.x;
		const bool in = t < end;
		// nodes with lists longer than kCoopList: the wave whose lanes hold their pairs counts their
		// children's edits (whole nodes per wave: chunks start on node boundaries)
		const u64 wnode = (t0 + (threadIdx.x & ~63u)) >> 3;
		const u64 longl = __ballot(in && (t & 7u) == 0 && ldg(&A.ent[(u32)(t >> 3)].llen, 0) > kCoopList);
		ListHits hits{0u, 0u, 0u}, hits2{0u, 0u, 0u};
		bool by_wave = false;
		for (u32 e = 0; e < 8u; ++e)
			if ((longl >> (8u * e)) & 1ull)
				coop_node(P, l, A, A.ent[(u32)(wnode + e)], e, lane, hits, hits2, by_wave);
		if (!in)
			continue;
		PairOut o = expand_pair_e(P, l, A, t, A.ent[(u32)(t >> 3)], ev, pm, ev_cap, nullptr, by_wave ? &hits : nullptr,
		                          by_wave ? &hits2 : nullptr);
		flag[t] = o.flag;
		lcnt[t] = o.lcnt;
		cstate[t] = o.cs;
		acc.a += o.flag & 1u;
		acc.b += o.lcnt;
	}
	Pair tot;
	block_excl_scan2(acc, sh, tot);
	if (threadIdx.x == 0)
		bsum[blockIdx.x] = tot;
}
__global__ void __launch_bounds__(kAT) a_emit(EditParams P, u32 l, const ALevel *__restrict__ lv, u32 *__restrict__ pm,
                                              const u32 *__restrict__ flag, const u32 *__restrict__ lcnt,
                                              const uint4 *__restrict__ cstate, const Pair *__restrict__ bsum) {
	__shared__ Pair sh[kAT];
	__shared__ u32 sflag[kAT + 8];
	if (threadIdx.x < 8)
		sflag[kAT + threadIdx.x] = 0;
	u64 begin, end;
	pair_chunk(pm[kPmHdr + l], begin, end);
	Pair pre{0, 0};
	for (u32 b = threadIdx.x; b < blockIdx.x; b += kAT)
		pre.a += bsum[b].a, pre.b += bsum[b].b;
	Pair base;
	block_excl_scan2(pre, sh, base);
	Pair tot = expand_emit_range(P, l, lv[l], lv[l + 1], pm, nullptr, 0, begin, end, base, flag, lcnt, cstate, sh,
	                             sflag);
	if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0)
		set_level_count(P, l, lv[l + 1], pm, tot);
}

// ---- colors of a sync-free pass: dense decode, pushes
// One workgroup per color-leaf job (grid-stride): a writer job's voxels get the old chunk's colors
// (or its fill color). A chunk's macro table and run headers are staged in LDS, and each thread
// decodes a run of kDenseRun consecutive voxels: one binary search for its first voxel, then it
// walks the headers forward (runs are in voxel order).
// The color-leaf entries are visited 64 at a time per block: block b takes entries c0 + j G
// (j < 64, G blocks, c0 = b + 64 G i), so the writer jobs of a run of neighbouring entries go to
// different blocks. Wave 0 reads the 64 flags in one round trip and handles the entries without a
// writer (`other`, in parallel); the block then runs its writer jobs one after another (the
// returned mask is block-uniform). A block per entry would spend a dependent load on every entry
// without a writer.
template <class Other>
__device__ inline u64 writer_chunk(const Entry *__restrict__ ent, u32 n, u32 c0, u64 *s_wm, Other other) {
	if (threadIdx.x < 64u) {
		const u32 e = c0 + threadIdx.x * gridDim.x;
		bool w = false;
		if (e < n) {
			w = (ldg(&ent[e].sf, 0) & kSfWriter) != 0u;
			if (!w)
				other(e);
		}
		const u64 m = __ballot(w);
		if (threadIdx.x == 0)
			*s_wm = m;
	}
	__syncthreads();
	const u64 m = *s_wm;
	__syncthreads(); // the next chunk rewrites s_wm
	return m;
}
constexpr u32 kJobChunk = 64;

constexpr u32 kDenseThreads = 256, kDenseRun = 16, kDenseHdrWords = 8192;
// The job's current colors in voxel order (DAGColorPool's leaf decode): a uniform fill, or the
// chunk's headers staged in LDS (hdr, kDenseHdrWords) and runs of kDenseRun voxels decoded per
// thread by seeking the run's header and walking forward. out may be global or LDS memory; every
// thread of the block calls this.
__device__ inline void decode_job(const Entry &E, const u32 *__restrict__ cleaves, u32 *hdr, Color *out, u32 V) {
	if (ctag(E.octree) != kTagLeaf) {
		const Color fill = get_fill(E.octree);
		for (u32 v = threadIdx.x; v < V; v += blockDim.x)
			out[v] = fill;
		return;
	}
	const u32 idx = cdata(E.octree), M = cleaves[idx + 1u], B = cleaves[idx + 2u];
	if (2u * (M + B) > kDenseHdrWords) { // a chunk too large to stage: decode each voxel in place
		for (u32 v = threadIdx.x; v < V; v += blockDim.x)
			out[v] = chunk_voxel(cleaves, idx, v);
		return;
	}
	const u32 mo = idx + 4u, wo = mo + 2u * M + 2u * B;
	__syncthreads(); // the previous job's readers are done with hdr
	for (u32 k = threadIdx.x; k < 2u * (M + B); k += blockDim.x)
		hdr[k] = cleaves[mo + k];
	__syncthreads();
	const u32 *mac = hdr, *hd = hdr + 2u * M;
	for (u32 v0 = threadIdx.x * kDenseRun; v0 < V; v0 += blockDim.x * kDenseRun) {
		u32 m = v0 >> 14, h = 0, lb = 0, ws = 0;
		auto seek = [&](u32 vid) { // last header of macro m with voxoff <= vid
			const u32 fb = mac[2u * m];
			lb = m + 1u < M ? mac[2u * (m + 1u)] : B;
			ws = mac[2u * m + 1u];
			u32 lo = fb, cnt = lb - fb;
			while (cnt > 0) {
				const u32 step = cnt >> 1;
				if ((hd[2u * (lo + step) + 1u] >> 18) <= vid)
					lo += step + 1u, cnt -= step + 1u;
				else
					cnt = step;
			}
			h = lo > fb ? lo - 1u : fb; // a macro's first header starts at voxel 0
		};
		if (m < M)
			seek(v0 & 0x3FFFu);
		for (u32 k = 0; k < kDenseRun && v0 + k < V; ++k) {
			const u32 v = v0 + k, vid = v & 0x3FFFu;
			if ((v >> 14) != m) { // a new macro block starts
				m = v >> 14;
				if (m < M)
					seek(vid);
			}
			Color c = Color::none();
			if (m < M) {
				while (h + 1u < lb && (hd[2u * (h + 1u) + 1u] >> 18) <= vid)
					++h;
				const u32 hc = hd[2u * h], hp = hd[2u * h + 1u], bpw = (hp >> 16) & 3u;
				if (bpw == 0)
					c = Color{hc, 1u << 16};
				else {
					const u32 bit = ws + (hp & 0xFFFFu) + (vid - (hp >> 18)) * bpw, o = bit & 31u;
					u32 w = cleaves[wo + (bit >> 5)] >> o;
					if (o + bpw > 32u)
						w |= cleaves[wo + (bit >> 5) + 1u] << (32u - o);
					w &= (1u << bpw) - 1u;
					c = Color{hc, w | (bpw << 8) | (1u << 16)};
				}
			}
			out[v] = c;
		}
	}
}
__global__ void __launch_bounds__(kDenseThreads) a_dense_init(const ALevel *__restrict__ lv, u32 leafL,
                                                              const u32 *__restrict__ pm, u32 log2v,
                                                              const u32 *__restrict__ cleaves,
                                                              Color *__restrict__ dense) {
	__shared__ u32 hdr[kDenseHdrWords]; // macro table (2M words) then headers (2B words)
	__shared__ u64 s_wm;
	const Entry *ent = lv[leafL].ent;
	const u32 jobs = pm[kPmHdr + leafL], V = 1u << log2v;
	for (u32 c0 = blockIdx.x; c0 < jobs; c0 += gridDim.x * kJobChunk)
	for (u64 m = writer_chunk(ent, jobs, c0, &s_wm, [](u32) {}); m; m &= m - 1ull) {
		const u32 job = c0 + (u32)__builtin_ctzll(m) * gridDim.x;
		const Entry E = ent[job];
		Color *out = dense + (u64)job * V;
		decode_job(E, cleaves, hdr, out, V);
	}
}
// Pushes cover terminal subtrees and leaf edits cover proceeding leaves: disjoint voxel ranges,
// so this kernel is ordered only after a_dense_init.
__global__ void a_push(const PushEvent *__restrict__ ev, const u32 *__restrict__ pm, u32 ev_cap, u32 V, u32 jobs,
                       Color *__restrict__ dense) {
	u32 nev = min(pm[kPmEvents], ev_cap);
	for (u32 k = blockIdx.x; k < nev; k += gridDim.x) {
		PushEvent p = ev[k];
		if (p.job >= jobs) // a job dropped by a level overflow (reported)
			continue;
		Color c{p.colors, p.wb};
		Color *out = dense + (u64)p.job * V + p.offset;
		for (u32 v = threadIdx.x; v < p.count; v += blockDim.x)
			out[v] = c;
	}
}

// ---- leaves + bottom-up DAG
// The first 128 words are fetched together with the word count (the pool is padded past its end).
__device__ inline u32 wave_find_leaf(const u32 *__restrict__ W, const u32 *__restrict__ bucket_words,
                                     const DevConfig &c, u32 bucket, u32 n0, u32 n1, u32 lane) {
	const u64 base = (u64)bucket << c.wbpb();
	const u32 wpb = c.words_per_bucket();
	u32 x0 = 0, x1 = 0;
	if (2u * lane + 2u <= wpb) {
		x0 = ldg(W, base + 2u * lane);
		x1 = ldg(W, base + 2u * lane + 1u);
	}
	const u32 bw = ldg(bucket_words, bucket);
	for (u32 i0 = 0; i0 < bw; i0 += 128u) {
		u32 j = i0 + 2u * lane;
		if (i0) {
			x0 = j + 2u <= bw ? ldg(W, base + j) : 0u;
			x1 = j + 2u <= bw ? ldg(W, base + j + 1u) : 0u;
		}
		bool m = j + 2u <= bw && x0 == n0 && x1 == n1;
		unsigned long long b = __ballot(m);
		if (b)
			return (u32)(base + i0 + 2u * (u32)__builtin_ctzll(b));
	}
	return kNull;
}

// Within-level dedup of the new nodes that are not in the pool yet, in a table shared by all
// levels: slot = (gen << 32) | entry, any other generation is empty, so the table is never
// cleared. Identical contents share a slot (a group) whose entry is the lowest index (the first
// in DFS order, the one the reference appends; the others find it). The entry that created the
// group queues it on its bucket's list; per-group results live beside the slot.
struct GroupTable {
	u64 *tab;
	u32 *gres, *gfail, *gnext;
	u32 mask;
};

// Result of a child entry once its level's upserts are done (k_resolve's rule).
__device__ inline u32 child_result(const ALevel &C, const GroupTable &G, u32 c) {
	// every load is issued unconditionally (indices clamped into their arrays) so the lookups of a
	// node's children overlap: two dependent round trips in all
	const u32 cc = min(c, C.cap - 1u);
	const u32 pend = ldg(C.pend, cc), st = ldg(C.status, cc), r = ldg(C.res, cc), s = ldg(C.slot, cc);
	const u32 pt = ldg(&C.ent[cc].ptr, 0);
	const u32 si = s & G.mask;
	const u32 gf = ldg(G.gfail, si), gr = ldg(G.gres, si);
	// not pending: status; found among the bucket's existing nodes: res; else the group's result
	// (the old pointer when the bucket was full)
	const u32 v = !pend ? st : (r != kNull ? r : (gf ? pt : gr));
	return c < C.cap ? v : kNull; // c >= cap is unreachable with sound bounds (the pass reports it)
}
// child_result and whether it is a node its group appended in this pass. Appends go past the
// bucket's used words, where no earlier node could point, so a node holding such a child is new:
// the find among the bucket's existing nodes would fail and is skipped.
__device__ inline uint2 child_result_fresh(const ALevel &C, const GroupTable &G, u32 c) {
	const u32 cc = min(c, C.cap - 1u);
	const u32 pend = ldg(C.pend, cc), st = ldg(C.status, cc), r = ldg(C.res, cc), s = ldg(C.slot, cc);
	const u32 pt = ldg(&C.ent[cc].ptr, 0);
	const u32 si = s & G.mask;
	const u32 gf = ldg(G.gfail, si), gr = ldg(G.gres, si);
	const bool fresh = pend && r == kNull && !gf;
	const u32 v = !pend ? st : (r != kNull ? r : (gf ? pt : gr));
	return c < C.cap ? make_uint2(v, fresh ? 1u : 0u) : make_uint2(kNull, 0u);
}

// edit_leaf (k_leaf) + the find among the bucket's existing leaves: one wavefront per leaf. A
// leaf inside a color job records what its edits assign: on the dense path the colors (lcol, by
// lane) and which voxels they are (lmask), which a_leafcolor moves into the job's dense colors (so
// this kernel does not wait for them to be decoded); with LDS jobs the assigning edit per voxel.
__device__ inline bool leaf_writes_colors(const EditParams &P, const ALevel *__restrict__ lv, const Entry &E) {
	return (E.sf & kSfWriter) && !(E.sf & kSfFinal) && (E.sf & kSfActive) && E.job < lv[P.c.color_leaf].cap;
}
// kMode: kLeafAll = occupancy, find and colors; kLeafNodes = occupancy and find only; kLeafColors =
// the colors of the leaves inside a color job only. A leaf's colors depend on its starting words and
// its edit list alone, so in a colored pass the color stream runs its own kLeafColors launch right
// after the top-down, beside the kLeafNodes launch of the node levels. kLeafColors records, per
// voxel, which of the pass's edits assigned its color last (lidx, kNoEdit for none: 2 bytes instead
// of the color and a mask bit); the color job reads the color from that edit.
enum : u32 { kLeafAll = 0, kLeafNodes = 1, kLeafColors = 2 };
constexpr u32 kNoEdit = 0xFFFFu; // a pass with LDS color jobs holds fewer edits
template <u32 kMode>
__global__ void __launch_bounds__(kAT) a_leaf(EditParams P, const ALevel *__restrict__ lv, const u32 *__restrict__ pm,
                                              Color *__restrict__ lcol, u64 *__restrict__ lmask,
                                              unsigned short *__restrict__ lidx) {
	const u32 lane = threadIdx.x & 63u, level = P.c.L - 1;
	const ALevel A = lv[level];
	const u32 n = pm[kPmHdr + level];
	const u32 waves = gridDim.x * (kAT / 64u);
	// the wave's leaf index is uniform: keep it (and what is read through it) in scalar registers
	const u32 e0 = __builtin_amdgcn_readfirstlane((u32)(((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
	for (u32 e = e0; e < n; e += waves) {
		Entry E = ldg(A.ent, e);
		E.x = __builtin_amdgcn_readfirstlane(E.x), E.y = __builtin_amdgcn_readfirstlane(E.y);
		E.z = __builtin_amdgcn_readfirstlane(E.z), E.ptr = __builtin_amdgcn_readfirstlane(E.ptr);
		E.sf = __builtin_amdgcn_readfirstlane(E.sf), E.job = __builtin_amdgcn_readfirstlane(E.job);
		E.loff = __builtin_amdgcn_readfirstlane(E.loff), E.llen = __builtin_amdgcn_readfirstlane(E.llen);
		const bool writer = kMode != kLeafNodes && leaf_writes_colors(P, lv, E);
		if (kMode == kLeafColors && !writer)
			continue;
		u32 w0 = 0, w1 = 0;
		if (E.ptr != kNull) {
			w0 = ldg(P.words, E.ptr);
			w1 = ldg(P.words, E.ptr + 1ull);
		}
		bool v = ((lane < 32u ? (w0 >> lane) : (w1 >> (lane - 32u))) & 1u) != 0;
		u32 vx = (E.x << 2) | ((lane >> 2) & 2u) | (lane & 1u);
		u32 vy = (E.y << 2) | ((lane >> 3) & 2u) | ((lane >> 1) & 1u);
		u32 vz = (E.z << 2) | ((lane >> 4) & 2u) | ((lane >> 2) & 1u);
		Color col = Color::none();
		bool asg = false;
		bool nv = v;
		u32 last = kNoEdit; // kLeafColors: the last edit that assigned this voxel's color
		for (u32 k = 0; k < E.llen; ++k) {
			const u32 idx = E.llen == 1 ? E.loff : ldg(A.lists, (u64)E.loff + k);
			DevEditor ed = P.editor(idx);
			if (!ed.stateful())
				nv = ed.edit_voxel(vx, vy, vz, nv);
			else if (kMode == kLeafColors) {
				bool a = false;
				nv = ed.edit_voxel_vbr_asg(vx, vy, vz, nv, col, a);
				last = a ? idx : last;
			} else
				nv = ed.edit_voxel_vbr_asg(vx, vy, vz, nv, col, asg);
		}
		if (kMode == kLeafColors) {
			lidx[(u64)e * 64u + lane] = (unsigned short)last;
			continue;
		}
		if (writer) {
			const u64 am = __ballot(asg);
			if (asg)
				lcol[(u64)e * 64u + lane] = col;
			if (lane == 0)
				lmask[e] = am;
		}
		unsigned long long m = __ballot(nv);
		u32 n0 = (u32)m, n1 = (u32)(m >> 32);
		bool pending = (n0 != w0 || n1 != w1) && m != 0ull; // wavefront-uniform
		if (!pending) {
			if (lane == 0) {
				stg(A.pend, e, 0u);
				stg(A.status, e, (n0 != w0 || n1 != w1) ? kNull : E.ptr);
			}
			continue;
		}
		u32 h = murmur_leaf(n0, n1);
		u32 b = P.c.level_base[level] + (h & ((1u << P.c.bucket_bits[level]) - 1u));
		u32 f = wave_find_leaf(P.words, P.bucket_words, P.c, b, n0, n1, lane);
		if (lane == 0) {
			stg(A.pend, e, 1u);
			stg(A.content, 2 * (u64)e, n0);
			stg(A.content, 2 * (u64)e + 1, n1);
			stg(A.bucket, e, b);
			stg(A.res, e, f);
		}
	}
}

// The colors a_leaf's edits assigned, into the jobs' dense colors (after a_dense_init and a_push:
// pushes cover terminal subtrees, leaves proceeding ones, so the voxel ranges are disjoint). One
// wavefront per leaf.
__global__ void __launch_bounds__(kAT) a_leafcolor(EditParams P, const ALevel *__restrict__ lv,
                                                   const u32 *__restrict__ pm, const Color *__restrict__ lcol,
                                                   const u64 *__restrict__ lmask, Color *__restrict__ dense, u32 V) {
	const u32 lane = threadIdx.x & 63u, level = P.c.L - 1;
	const ALevel A = lv[level];
	const u32 n = pm[kPmHdr + level];
	const u32 waves = gridDim.x * (kAT / 64u);
	const u32 e0 = __builtin_amdgcn_readfirstlane((u32)(((u64)blockIdx.x * blockDim.x + threadIdx.x) >> 6));
	for (u32 e = e0; e < n; e += waves) {
		Entry E = A.ent[e];
		E.sf = __builtin_amdgcn_readfirstlane(E.sf), E.job = __builtin_amdgcn_readfirstlane(E.job);
		if (!leaf_writes_colors(P, lv, E))
			continue;
		const u64 am = lmask[e];
		if ((am >> lane) & 1ull) {
			const u32 vx = (E.x << 2) | ((lane >> 2) & 2u) | (lane & 1u);
			const u32 vy = (E.y << 2) | ((lane >> 3) & 2u) | ((lane >> 1) & 1u);
			const u32 vz = (E.z << 2) | ((lane >> 4) & 2u) | ((lane >> 2) & 1u);
			const u32 msk = (1u << P.leaf_bits) - 1u;
			dense[(u64)E.job * V + morton3(vx & msk, vy & msk, vz & msk)] = lcol[(u64)e * 64u + lane];
		}
	}
}

// edit_node's join (k_build) for entry e of level l, with the children's results resolved from
// level l + 1, then the find among the bucket's existing nodes.
// The starting pointer of each proceeding or untouched slot is in kids (the old child, or the state
// a folded edit list leaves it in); pmask bits 8-15 flag the slots whose pointer changed.
struct Built {
	u32 fin;   // final pointer when known (unchanged / emptied / found)
	bool need; // new content that is not in the bucket yet
	u32 k, h, b, c[9];
};
// The words of an inner entry that its join reads: old pointer, proceed mask, first proceeding
// child, the 8 old / terminal child pointers, and its place in its parent (parent entry << 3 |
// slot); loaded ahead of the join where latency allows.
struct JoinIn {
	u32 ptr, pmw, fc, pidx;
	uint4 k0, k1;
};
__device__ inline JoinIn load_join(const ALevel &A, u32 e) {
	JoinIn q;
	q.ptr = ldg(&A.ent[e].ptr, 0);
	q.pidx = (ldg(&A.ent[e].parent, 0) << 3) | (ldg(&A.ent[e].sf, 0) & 7u);
	q.pmw = ldg(A.pmask, e);
	q.fc = ldg(A.first_child, e);
	q.k0 = ldg(reinterpret_cast<const uint4 *>(A.kids), 2 * (u64)e);
	q.k1 = ldg(reinterpret_cast<const uint4 *>(A.kids), 2 * (u64)e + 1);
	return q;
}
// The results of an entry's proceeding children, by slot: they are entries fc .. fc + cnt - 1 of
// the next level in slot order; all lookups are issued before any result is used.
template <typename ChildFn>
__device__ inline void children_by_slot(const JoinIn &q, ChildFn child, u32 (&cs)[8], bool &fresh) {
	const u32 pmk = q.pmw & 0xFFu, cnt = (u32)__builtin_popcount(pmk);
	u32 cr[8];
	fresh = false;
#pragma unroll
	for (u32 j = 0; j < 8; ++j) {
		const uint2 r = cnt ? child(q.fc + min(j, cnt - 1u)) : make_uint2(kNull, 0u);
		cr[j] = r.x;
		fresh |= r.y != 0u;
	}
#pragma unroll
	for (u32 i = 0; i < 8; ++i) {
		const u32 rank = (u32)__builtin_popcount(pmk & ((1u << i) - 1u));
		u32 c = cr[0];
#pragma unroll
		for (u32 j = 1; j <= i; ++j)
			c = rank == j ? cr[j] : c;
		cs[i] = c;
	}
}
// The join itself: slot i takes cs[i] when the child proceeded (pmask bit i), else its old or
// terminal pointer; pmask bits 8-15 flag the slots whose pointer changed.
__device__ inline Built build_node(const EditParams &P, u32 l, const JoinIn &q, const u32 (&cs)[8]) {
	Built o;
	const u32 pmw = q.pmw, pmk = pmw & 0xFFu;
	const uint4 k0 = q.k0, k1 = q.k1;
	const u32 kv[8] = {k0.x, k0.y, k0.z, k0.w, k1.x, k1.y, k1.z, k1.w};
	bool changed = ((pmw >> 8) & 0xFFu) != 0;
	u32 nm = 0, nk[8];
#pragma unroll
	for (u32 i = 0; i < 8; ++i) {
		const bool pr = (pmk >> i) & 1u;
		const u32 r = pr ? cs[i] : kv[i];
		changed |= pr && r != kv[i];
		nk[i] = r;
		nm |= (r != kNull ? 1u : 0u) << i;
	}
	o.need = false;
	o.k = 0;
	if (!changed || nm == 0) {
		o.fin = changed ? kNull : q.ptr;
		return o;
	}
	pack_inner(nm, nk, o.c);
	const u32 k = (u32)__builtin_popcount(nm) + 1u;
	o.k = k;
	o.h = murmur_regs(o.c, k);
	o.b = P.c.level_base[l] + (o.h & ((1u << P.c.bucket_bits[l]) - 1u));
	return o;
}

// edit_node's join (k_build) for entry e of level l, with the children's results resolved from
// level l + 1. An entry whose content is known final (unchanged / emptied) is written here and
// returns o.k == 0.
__device__ inline Built join_entry(const EditParams &P, u32 l, const ALevel &A, const ALevel &C, const GroupTable &G,
                                   u32 e) {
	const JoinIn q = load_join(A, e);
	u32 cs[8];
	bool fresh;
	children_by_slot(q, [&](u32 c) { return child_result_fresh(C, G, c); }, cs, fresh);
	Built o = build_node(P, l, q, cs);
	o.need = fresh; // (build_node leaves need false)
	if (!o.k) {
		stg(A.pend, e, 0u);
		stg(A.status, e, o.fin);
	}
	return o;
}
// The find of a joined entry among its bucket's existing nodes. Returns false when the
// workgroup's stage was full (the find runs again after a reset; the join is kept).
__device__ inline bool find_entry(const EditParams &P, u32 l, const ALevel &A, u32 e, const Built &o, u32 *stage,
                                  u32 *top, u32 cap) {
	u32 bw;
	const u32 f = o.need ? kNull // a fresh child: a new node
	                     : find_in_bucket_staged(P.words, P.bucket_words, P.c, o.b, o.c, o.k, headers_by_value(P.c, l),
	                                             bw, stage, top, cap);
	if (f == kFindRetry)
		return false;
#pragma unroll
	for (u32 j = 0; j < 9; ++j)
		if (j < o.k)
			stg(A.content, (u64)e * 9u + j, o.c[j]);
	stg(A.pend, e, 1u);
	stg(A.csize, e, o.k);
	stg(A.bucket, e, o.b);
	stg(A.res, e, f);
	return true;
}

// Joins entry e (pending, not in the pool) to its content's group; the creator queues the group
// (its arrival index lpos; the first kWalkList groups of a bucket are also listed by arrival).
__device__ inline void dedup_entry(const ALevel &A, const GroupTable &G, u32 gen, bool leaf, u32 e,
                                   u32 *__restrict__ bcnt, u32 *__restrict__ head, u32 *__restrict__ blist) {
	A.lpos[e] = kNull;
	if (!A.pend[e] || A.res[e] != kNull)
		return;
	const u32 stride = leaf ? 2u : 9u, n = leaf ? 2u : A.csize[e];
	const u32 *ce = A.content + (u64)e * stride;
	u32 h = leaf ? murmur_leaf(ce[0], ce[1]) : murmur_span(ce, n);
	u32 s = (h ^ (h >> 15)) * 0x2c1b3c6du;
	s = (s ^ (s >> 12)) & G.mask;
	const u64 mine = ((u64)gen << 32) | e;
	for (;;) {
		u64 cur = __hip_atomic_load(&G.tab[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
		if ((u32)(cur >> 32) != gen) {
			if (atomicCAS((unsigned long long *)&G.tab[s], (unsigned long long)cur, (unsigned long long)mine) == cur) {
				A.slot[e] = s;
				u32 b = A.bucket[e];
				const u32 lp = atomicAdd(&bcnt[b], 1u);
				A.lpos[e] = lp;
				if (lp < kWalkList)
					blist[(u64)b * kWalkList + lp] = s;
				G.gnext[s] = atomicExch(&head[b], s);
				return;
			}
			continue; // the slot changed under us: look again
		}
		u32 j = (u32)cur;
		bool same = leaf || A.csize[j] == n;
		for (u32 k = 0; k < n && same; ++k)
			same = A.content[(u64)j * stride + k] == ce[k];
		if (same) {
			atomicMin((unsigned long long *)&G.tab[s], (unsigned long long)mine);
			A.slot[e] = s;
			return;
		}
		s = (s + 1u) & G.mask;
	}
}

// The walker of a bucket appends the bucket's new groups in the order of their first entries (DFS
// order), as append_node (NodePool.hpp:134-157, 187-215) would: the groups of a level hold pairwise
// distinct contents, so no group can match a node appended by an earlier group of this walk (the
// find among the existing nodes, in the build, was the whole find).
// Short lists (kWalkList groups at most, read by arrival index) are ranked, placed and copied with
// static register indexing: every load of one phase is independent of the others, so a walk costs a
// few memory round trips whatever its length.
template <u32 kG>
__device__ inline u32 walk_short(const EditParams &P, bool leaf, const ALevel &A, const GroupTable &G, u32 b, u32 k,
                                  u32 bw, const u32 *__restrict__ blist) {
	const u32 stride = leaf ? 2u : 9u, wpp = P.c.words_per_page(), wpb = P.c.words_per_bucket();
	u32 sl[kG], key[kG], nz[kG], rk[kG], pos[kG];
#pragma unroll
	for (u32 x = 0; x < kG; ++x)
		sl[x] = x < k ? ldg(blist, (u64)b * kWalkList + x) : 0u;
#pragma unroll
	for (u32 x = 0; x < kG; ++x)
		key[x] = x < k ? (u32)ldg(G.tab, sl[x]) : 0xFFFFFFFFu; // the group's first entry
#pragma unroll
	for (u32 x = 0; x < kG; ++x)
		nz[x] = x < k ? (leaf ? 2u : ldg(A.csize, key[x])) : 0u;
#pragma unroll
	for (u32 x = 0; x < kG; ++x) { // rank by first entry (distinct keys)
		u32 r = 0;
#pragma unroll
		for (u32 y = 0; y < kG; ++y)
			r += key[y] < key[x] ? 1u : 0u;
		rk[x] = r;
		pos[x] = kNull;
	}
	u32 appended = 0;
	bool any_fail = false;
	const u64 bb = (u64)b << P.c.wbpb();
	for (u32 r = 0; r < kG; ++r) { // append_to_bucket in rank order: positions and page-tail zeros
		if (r >= k)
			break;
		u32 n = 0;
#pragma unroll
		for (u32 x = 0; x < kG; ++x)
			n = rk[x] == r ? nz[x] : n;
		u32 at = kNull;
		if (bw + n <= wpb) {
			u32 slot = bw >> P.c.wbpp, off = bw & (wpp - 1u);
			if (off + n > wpp) {
				for (u32 q = off; q < wpp; ++q)
					stg(P.words, bb + ((u64)slot << P.c.wbpp) + q, 0u);
				++slot;
				off = 0;
			}
			const u32 rel = (slot << P.c.wbpp) | off;
			at = (u32)(bb + rel);
			bw = rel + n;
			++appended;
		} else
			any_fail = true;
#pragma unroll
		for (u32 x = 0; x < kG; ++x)
			pos[x] = rk[x] == r ? at : pos[x];
	}
	// results, then the contents, four groups per round of loads
#pragma unroll
	for (u32 x = 0; x < kG; ++x)
		if (x < k) {
			stg(G.gres, sl[x], pos[x]);
			stg(G.gfail, sl[x], pos[x] == kNull ? 1u : 0u);
		}
#pragma unroll
	for (u32 x0 = 0; x0 < kG; x0 += 4) {
		if (x0 >= k)
			break;
		u32 w[4][9];
#pragma unroll
		for (u32 d = 0; d < 4; ++d)
#pragma unroll
			for (u32 j = 0; j < 9; ++j)
				w[d][j] = x0 + d < k && pos[x0 + d] != kNull && j < nz[x0 + d]
				              ? ldg(A.content, (u64)key[x0 + d] * stride + j)
				              : 0u;
#pragma unroll
		for (u32 d = 0; d < 4; ++d)
#pragma unroll
			for (u32 j = 0; j < 9; ++j)
				if (x0 + d < k && pos[x0 + d] != kNull && j < nz[x0 + d])
					stg(P.words, (u64)pos[x0 + d] + j, w[d][j]);
	}
	stg(P.bucket_words, b, bw);
	if (any_fail)
		atomicOr(&P.meta[kMetaOverflow], 1u);
	return appended;
}

// A bucket of kWalkShort < k <= kWalkList new groups, walked by a whole wavefront (all 64 lanes,
// b and k uniform): lane x holds the x-th group by arrival; its rank by first entry counts the
// smaller keys across the wave, the positions follow in rank order from one uniform pass (page
// padding makes them sequential; the padding words are cleared by the wave), and each lane writes
// its group's result and node words. Returns the nodes appended on lane 0 (0 on the others).
__device__ inline u32 walk_wave(const EditParams &P, bool leaf, const ALevel &A, const GroupTable &G, u32 b, u32 k,
                                const u32 *__restrict__ blist, u32 lane) {
	const u32 stride = leaf ? 2u : 9u, wpp = P.c.words_per_page(), wpb = P.c.words_per_bucket();
	const bool in = lane < k;
	const u32 sl = in ? ldg(blist, (u64)b * kWalkList + lane) : 0u;
	const u32 key = in ? (u32)ldg(G.tab, sl) : 0xFFFFFFFFu; // the group's first entry
	const u32 nz = in ? (leaf ? 2u : ldg(A.csize, key)) : 0u;
	u32 rk = 0;
	for (u32 y = 0; y < k; ++y) // distinct keys
		rk += (u32)__builtin_amdgcn_readlane((int)key, (int)y) < key ? 1u : 0u;
	u32 nz_r = 0; // lane r: the size of the group ranked r
	for (u32 y = 0; y < k; ++y)
		if ((u32)__builtin_amdgcn_readlane((int)rk, (int)y) == lane)
			nz_r = (u32)__builtin_amdgcn_readlane((int)nz, (int)y);
	u32 bw = ldg(P.bucket_words, b), at_r = kNull, appended = 0;
	bool any_fail = false;
	const u64 bb = (u64)b << P.c.wbpb();
	for (u32 r = 0; r < k; ++r) { // append_to_bucket in rank order
		const u32 n = (u32)__builtin_amdgcn_readlane((int)nz_r, (int)r);
		u32 at = kNull;
		if (bw + n <= wpb) {
			u32 slot = bw >> P.c.wbpp, off = bw & (wpp - 1u);
			if (off + n > wpp) {
				for (u32 q = off + lane; q < wpp; q += 64u)
					stg(P.words, bb + ((u64)slot << P.c.wbpp) + q, 0u);
				++slot;
				off = 0;
			}
			const u32 rel = (slot << P.c.wbpp) | off;
			at = (u32)(bb + rel);
			bw = rel + n;
			++appended;
		} else
			any_fail = true;
		at_r = lane == r ? at : at_r;
	}
	const u32 pos = (u32)__shfl((int)at_r, (int)rk);
	if (in) {
		stg(G.gres, sl, pos);
		stg(G.gfail, sl, pos == kNull ? 1u : 0u);
		if (pos != kNull)
			for (u32 j = 0; j < nz; ++j)
				stg(P.words, (u64)pos + j, ldg(A.content, (u64)key * stride + j));
	}
	if (lane == 0) {
		stg(P.bucket_words, b, bw);
		if (any_fail)
			atomicOr(&P.meta[kMetaOverflow], 1u);
	}
	return lane == 0 ? appended : 0u;
}

// Returns the nodes appended (0 unless e walks its bucket). A bucket of more than kWalkShort and at
// most kWalkList new groups is left to the caller's wavefront (wb, wk).
__device__ inline u32 walk_entry(const EditParams &P, u32 l, const ALevel &A, const GroupTable &G, u32 e,
                                  u32 *__restrict__ bcnt, u32 *__restrict__ head, const u32 *__restrict__ blist,
                                  u32 &wb, u32 &wk) {
	if (A.lpos[e] != 0)
		return 0;
	const bool leaf = l == P.c.L - 1;
	const u32 stride = leaf ? 2u : 9u;
	const u32 b = A.bucket[e], k = bcnt[b];
	const u32 bw0 = ldg(P.bucket_words, b);
	u32 appended = 0;
	if (k <= 4u)
		appended = walk_short<4>(P, leaf, A, G, b, k, bw0, blist);
	else if (k <= kWalkShort)
		appended = walk_short<kWalkShort>(P, leaf, A, G, b, k, bw0, blist);
	else if (k <= kWalkList) {
		wb = b, wk = k;
		return 0;
	} else { // long lists (distinct new nodes of one bucket): repeated minimum search over the chain
		u32 bw = bw0;
		bool any_fail = false;
		u32 last = 0;
		for (u32 r = 0; r < k; ++r) {
			u32 best = kNull, bs = kNull;
			for (u32 s = head[b]; s != kNull; s = G.gnext[s]) {
				u32 j = (u32)G.tab[s];
				if ((r == 0 || j > last) && j < best)
					best = j, bs = s;
			}
			if (bs == kNull)
				break;
			const u32 *node = A.content + (u64)best * stride;
			const u32 f = append_to_bucket(P.words, P.c, b, bw, node, leaf ? 2u : A.csize[best]);
			if (f == kNull)
				any_fail = true;
			else
				++appended;
			G.gres[bs] = f;
			G.gfail[bs] = f == kNull ? 1u : 0u;
			last = best;
		}
		stg(P.bucket_words, b, bw);
		if (any_fail)
			atomicOr(&P.meta[kMetaOverflow], 1u);
	}
	bcnt[b] = 0;
	head[b] = kNull;
	return appended;
}

// kCap: staging words per workgroup. Entries whose buckets do not fit the stage wait for the next
// round, so a larger stage takes fewer rounds; 10240 words (40 KiB) fit beside three color-job
// workgroups on a CU, for the batched passes whose color jobs run beside their node levels.
template <u32 kCap>
__global__ void __launch_bounds__(kAT) a_build(EditParams P, u32 l, const ALevel *__restrict__ lv,
                                               const u32 *__restrict__ pm, GroupTable G) {
	__shared__ __align__(16) u32 stage[kCap];
	__shared__ u32 top;
	const ALevel A = lv[l], C = lv[l + 1];
	const u32 n = pm[kPmHdr + l];
	for (u32 e0 = blockIdx.x * blockDim.x; e0 < n; e0 += gridDim.x * blockDim.x) { // uniform trip count
		const u32 e = e0 + threadIdx.x;
		Built o{};
		o.k = 0;
		if (e < n)
			o = join_entry(P, l, A, C, G, e);
		bool todo = o.k != 0;
		do { // an entry whose bucket found the stage full looks again after the reset
			if (threadIdx.x == 0)
				top = 0;
			__syncthreads();
			if (todo)
				todo = !find_entry(P, l, A, e, o, stage, &top, kCap);
		} while (__syncthreads_or(todo));
	}
}
__global__ void __launch_bounds__(kAT) a_dedup(EditParams P, u32 l, const ALevel *__restrict__ lv,
                                               const u32 *__restrict__ pm, GroupTable G, u32 *__restrict__ bcnt,
                                               u32 *__restrict__ head, u32 *__restrict__ blist) {
	const ALevel A = lv[l];
	const u32 n = pm[kPmHdr + l], gen = pm[kPmGen] + l;
	const bool leaf = l == P.c.L - 1;
	for (u32 e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
		dedup_entry(A, G, gen, leaf, e, bcnt, head, blist);
}
__global__ void __launch_bounds__(kAT) a_walk(EditParams P, u32 l, const ALevel *__restrict__ lv,
                                              const u32 *__restrict__ pm, GroupTable G, u32 *__restrict__ bcnt,
                                              u32 *__restrict__ head, const u32 *__restrict__ blist,
                                              u32 *__restrict__ acc) {
	const ALevel A = lv[l];
	const u32 n = pm[kPmHdr + l], lane = threadIdx.x & 63u;
	const bool leaf = l == P.c.L - 1;
	u32 app = 0;
	for (u32 e0 = blockIdx.x * blockDim.x; e0 < n; e0 += gridDim.x * blockDim.x) { // uniform trip count
		const u32 e = e0 + threadIdx.x;
		u32 wb = kNull, wk = 0;
		if (e < n)
			app += walk_entry(P, l, A, G, e, bcnt, head, blist, wb, wk);
		// the buckets of kWalkShort..kWalkList new groups its lanes found: one at a time, by the wave
		for (u64 m = __ballot(wb != kNull); m; m &= m - 1ull) {
			const int src = (int)__builtin_ctzll(m);
			const u32 b = (u32)__builtin_amdgcn_readlane((int)wb, src), k = (u32)__builtin_amdgcn_readlane((int)wk, src);
			app += walk_wave(P, leaf, A, G, b, k, blist, lane);
			if (lane == 0) {
				bcnt[b] = 0;
				head[b] = kNull;
			}
		}
	}
	// one counter update per wavefront: a bump per walker would serialise on the one address
#pragma unroll
	for (u32 d = 32; d > 0; d >>= 1)
		app += (u32)__shfl_xor((int)app, (int)d);
	if ((threadIdx.x & 63u) == 0 && app)
		atomicAdd(&acc[kAccUpserts], app);
}

// ---- colors: encode + SetLeaf (one block per color leaf job, grid-stride)
// Register path of the encoder for chunks of up to kEncPer * kEncThreads voxels: each thread
// loads its contiguous voxels with wide loads once (one memory round trip instead of one per voxel).
constexpr u32 kEncPer = 16;
struct EncRegs {
	u32 c[kEncPer], w[kEncPer]; // colors and weight words of voxels v0 .. v0 + n - 1
	u32 v0, n, heads;           // heads: bit k set if voxel v0 + k starts a header
	EncScan x;
};
__device__ inline void enc_load(const Color *__restrict__ d, u32 V, EncRegs &R) {
	const u32 per = (V + kEncThreads - 1) / kEncThreads;
	R.v0 = threadIdx.x * per;
	const u32 v1 = min(V, R.v0 + per);
	R.n = R.v0 < v1 ? v1 - R.v0 : 0u;
	Color prev{0, 0};
	if (R.n && R.v0)
		prev = d[R.v0 - 1];
	if (R.n == kEncPer) {
		const uint4 *q = reinterpret_cast<const uint4 *>(d + R.v0);
#pragma unroll
		for (u32 k = 0; k < kEncPer / 2; ++k) {
			uint4 t = q[k];
			R.c[2 * k] = t.x;
			R.w[2 * k] = t.y;
			R.c[2 * k + 1] = t.z;
			R.w[2 * k + 1] = t.w;
		}
	} else {
#pragma unroll
		for (u32 k = 0; k < kEncPer; ++k)
			if (k < R.n) {
				Color t = d[R.v0 + k];
				R.c[k] = t.colors;
				R.w[k] = t.wb;
			}
	}
	R.heads = 0;
	R.x = EncScan{0, 0};
	u32 pc = prev.colors, pb = (prev.wb >> 8) & 0xFFu;
#pragma unroll
	for (u32 k = 0; k < kEncPer; ++k)
		if (k < R.n) {
			const u32 v = R.v0 + k, b = (R.w[k] >> 8) & 0xFFu;
			if (v == 0 || (v & 0x3FFFu) == 0 || R.c[k] != pc || b != pb)
				R.heads |= 1u << k;
			R.x.bits += b;
			pc = R.c[k];
			pb = b;
		}
	R.x.heads = (u32)__builtin_popcount(R.heads);
}

// Encoded chunks of at most kEncTmp words are also written, in their final layout, to the job's
// slot of a staging buffer by a_encode_size (built in LDS, copied out coalesced): a_encode_write
// then copies them instead of reading the job's dense colors again (larger chunks are encoded
// again from the dense colors).
constexpr u32 kEncTmp = 4096;
__global__ void __launch_bounds__(kEncThreads) a_encode_size(EditParams P, const ALevel *__restrict__ lv,
                                                              u32 *__restrict__ pm, const Color *__restrict__ dense,
                                                              u32 V, u32 *__restrict__ info, u32 *__restrict__ app,
                                                              u32 *__restrict__ etmp, u32 *__restrict__ rest) {
	__shared__ EncScan sh[kEncThreads];
	__shared__ u32 stmp[kEncTmp];
	__shared__ u64 s_wm;
	const u32 leafL = P.c.color_leaf;
	const Entry *ent = lv[leafL].ent;
	const u32 n = pm[kPmHdr + leafL];
	for (u32 c0 = blockIdx.x; c0 < n; c0 += gridDim.x * kJobChunk)
	for (u64 m = writer_chunk(ent, n, c0, &s_wm, [&](u32 e) { app[e] = 0; }); m; m &= m - 1ull) {
		const u32 e = c0 + (u32)__builtin_ctzll(m) * gridDim.x;
		const Entry E = ent[e];
		const Color *d = dense + (u64)e * V;
		EncScan tot{0, 0};
		bool staged = false;
		if (V <= kEncPer * kEncThreads) {
			EncRegs R;
			enc_load(d, V, R);
			EncScan run = block_excl_scan(R.x, sh, tot);
			// one macro block (V <= 2^14): [size, M = 1, B, W, macro heads, macro bits, B headers, weights]
			const u32 B = tot.heads, W = (tot.bits + 31u) >> 5, ds = 6u + 2u * B + W;
			staged = ds <= kEncTmp; // block-uniform
			if (staged) {
				const u32 bo = 6u, wo = bo + 2u * B;
				for (u32 w = threadIdx.x; w < W; w += kEncThreads)
					stmp[wo + w] = 0u;
				if (threadIdx.x == 0) {
					stmp[0] = 0u;
					stmp[1] = 1u;
					stmp[2] = B;
					stmp[3] = W;
					stmp[4] = 0u;
					stmp[5] = 0u;
				}
				__syncthreads();
#pragma unroll
				for (u32 k = 0; k < kEncPer; ++k)
					if (k < R.n) {
						const u32 v = R.v0 + k, bpw = (R.w[k] >> 8) & 0xFFu;
						if ((R.heads >> k) & 1u) {
							stmp[bo + 2u * run.heads] = R.c[k];
							stmp[bo + 2u * run.heads + 1u] = (v << 18) | (bpw << 16) | run.bits;
							run.heads += 1u;
						}
						if (bpw) {
							const u32 w = R.w[k] & ((1u << bpw) - 1u), o = run.bits & 31u;
							atomicOr(&stmp[wo + (run.bits >> 5)], w << o);
							if (o + bpw > 32u)
								atomicOr(&stmp[wo + (run.bits >> 5) + 1u], w >> (32u - o));
						}
						run.bits += bpw;
					}
				__syncthreads();
				for (u32 w = threadIdx.x; w < ds; w += kEncThreads)
					etmp[(u64)e * kEncTmp + w] = stmp[w];
			}
		} else {
			EncScan x{0, 0};
			u32 per = (V + kEncThreads - 1) / kEncThreads, v0 = threadIdx.x * per, v1 = min(V, v0 + per);
			for (u32 v = v0; v < v1; ++v) {
				x.heads += is_head(d, v) ? 1u : 0u;
				x.bits += (d[v].wb >> 8) & 0xFFu;
			}
			block_excl_scan(x, sh, tot);
		}
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
			info[4 * e + 2] = inplace | (staged ? 2u : 0u);
			info[4 * e + 3] = tot.heads;
			app[e] = as;
			if (!staged)
				rest[atomicAdd(&pm[kPmRest], 1u)] = e;
		}
	}
}
// Colored passes whose color leaves hold at most kJobV voxels (so at most two node levels lie
// between a job and its 4^3 leaves). a_job_ranges, one thread per writer job: its contiguous entry
// ranges at those levels (from first_child / pmask) in jr[2 j], and the header of its current
// chunk {M, B, block size, weight words} (or zeros for a uniform color) in jr[2 j + 1].
constexpr u32 kJobV = 4096;
// Also fills the pass's edit color table (ecol[k]: the color a_leaf's kLeafColors records of edit k
// stand for; a column fill's, which depends on the voxel, is flagged in bit 24 of the weight word).
__global__ void a_job_ranges(const ALevel *__restrict__ lv, const u32 *__restrict__ pm, u32 leafL, u32 L,
                             const u32 *__restrict__ cleaves, uint4 *__restrict__ jr, const hd_edit_desc *__restrict__ edits,
                             u32 n_edits, uint2 *__restrict__ ecol) {
	const Entry *ent = lv[leafL].ent;
	const u32 n = pm[kPmHdr + leafL];
	for (u32 k = blockIdx.x * blockDim.x + threadIdx.x; k < n_edits; k += gridDim.x * blockDim.x) {
		const hd_edit_desc d = edits[k];
		const Color c = from_hd(d.color);
		ecol[k] = make_uint2(c.colors, c.wb | (d.kind == HD_EDIT_COLUMN_FILL ? 1u << 24 : 0u));
	}
	for (u32 j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
		if (!(ldg(&ent[j].sf, 0) & kSfWriter))
			continue;
		const u32 oct = ldg(&ent[j].octree, 0);
		uint4 hdr = make_uint4(0, 0, 0, 0);
		if (ctag(oct) == kTagLeaf)
			hdr = make_uint4(ldg(cleaves, (u64)cdata(oct) + 1u), ldg(cleaves, (u64)cdata(oct) + 2u),
			                 ldg(cleaves, (u64)cdata(oct)), ldg(cleaves, (u64)cdata(oct) + 3u));
		u32 lo = j, hi = j + 1u, r[4] = {0, 0, 0, 0};
#pragma unroll
		for (u32 k = 0; k < 2; ++k) {
			const u32 l = leafL + k;
			if (l + 1u < L && hi > lo) { // an empty range stays empty
				const u32 *fc = lv[l].first_child, *pmk = lv[l].pmask;
				const u32 nlo = ldg(fc, lo);
				const u32 nhi = ldg(fc, hi - 1u) + (u32)__builtin_popcount(ldg(pmk, hi - 1u) & 0xFFu);
				lo = nlo;
				hi = max(nhi, nlo);
			} else if (l + 1u < L)
				lo = hi = 0;
			r[2 * k] = lo;
			r[2 * k + 1] = hi;
		}
		jr[2ull * j] = make_uint4(r[0], r[1], r[2], r[3]);
		jr[2ull * j + 1u] = hdr;
	}
}

// LDS color jobs: one workgroup per writer job rebuilds the job's colors in LDS (job_rebuild) and
// encodes them (a_job_encode, a_job_rest). Every load of a phase is independent of the others in
// it: a rebuild costs about three memory round trips.
constexpr u32 kJobHdr = 3072; // staged header words (larger chunks are decoded voxel by voxel); also the
                              // largest encoded chunk staged here (larger ones go to a_job_rest)
// A voxel's VBRColor in LDS: colors, and weight | bits_per_weight << 8 | has_value << 15 in 16 bits
// (hd_edit accepts bits_per_weight <= 3), so that four jobs fit a CU's LDS.
__device__ inline u32 pack_wb16(u32 wb) { return (wb & 0xFFu) | (((wb >> 8) & 0x7Fu) << 8) | (((wb >> 16) & 1u) << 15); }
__device__ inline u32 unpack_wb16(u32 w) { return (w & 0xFFu) | (((w >> 8) & 0x7Fu) << 8) | ((w >> 15) << 16); }
static_assert(kJobHdr <= kEncTmp, "a staged chunk fits its kEncTmp-word slot");
static_assert(kJobHdr % kEncThreads == 0 && kEncThreads >= 72, "staged words split evenly; one push per thread");
struct JobLds {
	u32 colc[kJobV];
	unsigned short colw[kJobV];
	union {
		u32 hdr[kJobHdr];
		u32 stmp[kJobHdr];
	};
	union { // the rebuild's records (dead once the encode starts) / the encode's wave totals
		struct {
			u32 cov[kJobV / 32];
			uint4 fill[kJobV / 64 + 8]; // pushes: {voxel offset, colors, wb, voxel count}
			u32 nfill;
			u32 lbase[kJobV / 64];      // per leaf of the job: its first voxel, or kNull without assigned colors
		};
		EncScan sh[kEncThreads / 64];
	};
};
__device__ inline void job_set(JobLds &S, u32 v, Color c) {
	S.colc[v] = c.colors;
	S.colw[v] = (unsigned short)pack_wb16(c.wb);
}
// enc_load over the job's LDS colors.
// A thread's voxels of a job (v0 .. v0 + n - 1): which start a header, and the totals for the
// block scan. The colors stay in LDS; the writers read them again.
struct JobEnc {
	u32 v0, n, heads;
	EncScan x;
};
// kFull: V == kJobV, so every thread holds exactly kEncPer voxels.
template <bool kFull> __device__ inline void enc_scan_job(const JobLds &S, u32 V, JobEnc &R) {
	const u32 per = kFull ? kEncPer : (V + kEncThreads - 1) / kEncThreads;
	R.v0 = threadIdx.x * per;
	const u32 v1 = kFull ? R.v0 + kEncPer : min(V, R.v0 + per);
	R.n = R.v0 < v1 ? v1 - R.v0 : 0u;
	u32 pc = 0, pb = 0;
	if (R.n && R.v0)
		pc = S.colc[R.v0 - 1], pb = (S.colw[R.v0 - 1] >> 8) & 0x7Fu;
	R.heads = 0;
	R.x = EncScan{0, 0};
	auto step = [&](u32 k, u32 c, u32 w16) {
		const u32 v = R.v0 + k, b = (w16 >> 8) & 0x7Fu;
		if (v == 0 || (v & 0x3FFFu) == 0 || c != pc || b != pb)
			R.heads |= 1u << k;
		R.x.bits += b;
		pc = c;
		pb = b;
	};
	if (kFull || R.n == kEncPer) {
		const uint4 *qc = reinterpret_cast<const uint4 *>(S.colc + R.v0);
		const uint4 *qw = reinterpret_cast<const uint4 *>(S.colw + R.v0);
#pragma unroll
		for (u32 q = 0; q < kEncPer / 8; ++q) { // 8 voxels a step: their colors and weight words
			const uint4 c0 = qc[2 * q], c1 = qc[2 * q + 1], w = qw[q];
			const u32 cs[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w}, ws[4] = {w.x, w.y, w.z, w.w};
#pragma unroll
			for (u32 j = 0; j < 8; ++j)
				step(8 * q + j, cs[j], (ws[j >> 1] >> (16 * (j & 1))) & 0xFFFFu);
		}
	} else {
#pragma unroll
		for (u32 k = 0; k < kEncPer; ++k)
			if (k < R.n)
				step(k, S.colc[R.v0 + k], S.colw[R.v0 + k]);
	}
	R.x.heads = (u32)__builtin_popcount(R.heads);
}
// Rebuilds job e's colors in S (every thread of the block calls it): the colors pushed onto its
// terminal subtrees (pev records of the pairs whose pmask push bit is set) and the colors its leaves'
// edits assigned (lcol / lmask) are written first and marked covered (pushes cover terminal
// subtrees, leaves proceeding ones: disjoint); the job's current colors (DAGColorPool's leaf decode,
// headers staged in LDS) fill the rest. Returns with S complete and visible to the block; E and jh
// are the job's entry and the header of its current chunk.
__device__ inline void job_rebuild(const EditParams &P, const ALevel *__restrict__ lv, const uint4 *__restrict__ jrange,
                                   const unsigned short *__restrict__ lidx, const uint2 *__restrict__ ecol, u32 V, u32 e,
                                   JobLds &S, Entry &E, uint4 &jh) {
	const u32 leafL = P.c.color_leaf, L = P.c.L, vl = P.c.voxel_level;
	const u32 *cleaves = P.cleaves;
	const Entry *ent = lv[leafL].ent;
	const u32 below = L - 1u - leafL; // node levels from the job down to the leaves (0, 1 or 2)
	__syncthreads(); // the previous job's readers are done with S
	for (u32 w = threadIdx.x; w < V / 32u; w += kEncThreads)
		S.cov[w] = 0u;
	if (threadIdx.x == 0)
		S.nfill = 0;
	E = ent[e];
	const uint4 jr = jrange[2ull * e];
	jh = jrange[2ull * e + 1u];
	__syncthreads();
	const bool chunk = ctag(E.octree) == kTagLeaf;
	const u32 idx = cdata(E.octree), M = jh.x, B0 = jh.y, W0 = jh.w;
	// pushes (the job's own 8 pairs, then the pairs of its entries one level down: at most 72),
	// the leaves' records, the chunk's headers and its weight words when they fit beside them (at
	// most 384 words for 4096 voxels of <= 3 bits) and the leaves' colors: every load is issued
	// before the first LDS store, so they share one round trip
	const u32 t0 = threadIdx.x;
	const u32 np0 = below ? 8u : 0u, np1 = below > 1u ? 8u * (jr.y - jr.x) : 0u;
	u32 pmw = 0, pl = leafL;
	uint4 q = make_uint4(0, 0, 0, 0);
	if (t0 < np0 + np1) {
		pl = t0 < np0 ? leafL : leafL + 1u;
		const u32 pe = t0 < np0 ? e : jr.x + ((t0 - np0) >> 3);
		pmw = ldg(lv[pl].pmask, pe);
		q = ldg(lv[pl].pev, 8ull * pe + (t0 & 7u)); // read ahead: valid where the push bit is set
	}
	const ALevel &LA = lv[L - 1u];
	const u32 llo = below == 0u ? e : (below == 1u ? jr.x : jr.z);
	const u32 nl = below == 0u ? 1u : (below == 1u ? jr.y - jr.x : jr.w - jr.z);
	Entry LE{};
	if (t0 < nl) {
		const u32 le = llo + t0;
		LE.sf = ldg(&LA.ent[le].sf, 0);
		LE.job = ldg(&LA.ent[le].job, 0);
		LE.x = ldg(&LA.ent[le].x, 0);
		LE.y = ldg(&LA.ent[le].y, 0);
		LE.z = ldg(&LA.ent[le].z, 0);
	}
	const bool staged_hdr = chunk && 2u * (M + B0) <= kJobHdr;
	const bool staged_w = staged_hdr && W0 <= kJobHdr && 2u * (M + B0) + W0 <= kJobHdr;
	const u32 nk = staged_hdr ? 2u * (M + B0) + (staged_w ? W0 : 0u) : 0u;
	constexpr u32 kHdrPer = kJobHdr / kEncThreads;
	u32 hw[kHdrPer];
#pragma unroll
	for (u32 r = 0; r < kHdrPer; ++r) {
		const u32 k = t0 + r * kEncThreads;
		if (k < nk)
			hw[r] = cleaves[idx + 4u + k];
	}
	constexpr u32 kPer = kJobV / kEncThreads;
	// which edit assigned each leaf voxel's color last: a thread's r-th voxel lies in leaf
	// wave + 4 r (wave-uniform)
	const u32 wv = __builtin_amdgcn_readfirstlane(t0 >> 6);
	u32 ix[kPer];
#pragma unroll
	for (u32 r = 0; r < kPer; ++r)
		ix[r] = wv + r * (kEncThreads / 64u) < nl ? (u32)ldg(lidx, 64ull * llo + t0 + r * kEncThreads) : kNoEdit;
	if (t0 < np0 + np1 && ((pmw >> (16u + (t0 & 7u))) & 1u))
		S.fill[atomicAdd(&S.nfill, 1u)] = make_uint4(q.x, q.y, q.z, 1u << (3u * (vl - (pl + 1u))));
	if (t0 < nl) {
		const u32 m2 = (1u << P.leaf_bits >> 2) - 1u;
		// a leaf's lane is the Morton index of its voxel inside the 4^3 leaf
		S.lbase[t0] = leaf_writes_colors(P, lv, LE) ? morton3(LE.x & m2, LE.y & m2, LE.z & m2) << 6 : kNull;
	}
#pragma unroll
	for (u32 r = 0; r < kHdrPer; ++r) {
		const u32 k = t0 + r * kEncThreads;
		if (k < nk)
			S.hdr[k] = hw[r];
	}
	__syncthreads();
	// the assigning edits' colors from the pass's table (their loads overlap the pushes below)
	uint2 ac[kPer];
#pragma unroll
	for (u32 r = 0; r < kPer; ++r) {
		const u32 li = wv + r * (kEncThreads / 64u);
		ac[r] = make_uint2(0u, 0u);
		if (ix[r] != kNoEdit && li < nl && S.lbase[li] != kNull)
			ac[r] = ldg(ecol, ix[r]);
	}
	// the pushed and assigned colors, marked covered. A folded pass can push a color on a proceeding
	// child (the edit that reset it) and again on its children or assign leaf colors inside it: the
	// larger pushes go first, the leaves last.
	const u32 big = 1u << (3u * (vl - (leafL + 1u)));
	for (u32 pass = 0; pass < 2; ++pass) {
		for (u32 f = 0; f < S.nfill; ++f) {
			const uint4 q = S.fill[f];
			if ((q.w >= big) != (pass == 0))
				continue;
			for (u32 v = threadIdx.x; v < q.w; v += kEncThreads)
				job_set(S, q.x + v, Color{q.y, q.z});
			if (q.w >= 32u)
				for (u32 w = (q.x >> 5) + threadIdx.x; w < (q.x + q.w) >> 5; w += kEncThreads)
					S.cov[w] = 0xFFFFFFFFu;
			else if (threadIdx.x == 0)
				atomicOr(&S.cov[q.x >> 5], ((1u << q.w) - 1u) << (q.x & 31u));
		}
		if (P.folded)
			__syncthreads();
	}
#pragma unroll
	for (u32 r = 0; r < kPer; ++r) {
		const u32 li = wv + r * (kEncThreads / 64u), lane = t0 & 63u;
		if (ix[r] != kNoEdit && li < nl && S.lbase[li] != kNull) {
			const u32 at = S.lbase[li] | lane;
			Color c{ac[r].x, ac[r].y};
			if (ac[r].y >> 24) { // a column fill: the voxel's absolute position from its Morton index in the job
				const u32 bits = P.leaf_bits;
				const u32 vx = (E.x << bits) | morton_compact(at), vy = (E.y << bits) | morton_compact(at >> 1),
				          vz = (E.z << bits) | morton_compact(at >> 2);
				c = Color::rgb8(P.editor(ix[r]).column_rgb8(vx, vy, vz));
			}
			job_set(S, at, c);
			atomicOr(&S.cov[at >> 5], 1u << (at & 31u));
		}
	}
	__syncthreads();
	// the current colors of the voxels not covered above
	if (!chunk || M == 0u || B0 == 0u) {
		const Color fill = chunk ? Color::none() : get_fill(E.octree);
		for (u32 v = threadIdx.x; v < V; v += kEncThreads)
			if (!((S.cov[v >> 5] >> (v & 31u)) & 1u))
				job_set(S, v, fill);
	} else if (!staged_hdr) {
		for (u32 v = threadIdx.x; v < V; v += kEncThreads)
			if (!((S.cov[v >> 5] >> (v & 31u)) & 1u))
				job_set(S, v, chunk_voxel(cleaves, idx, v));
	} else {
		const u32 *mac = S.hdr, *hd = S.hdr + 2u * M, wo = idx + 4u + 2u * M + 2u * B0;
		for (u32 v0 = threadIdx.x * kDenseRun; v0 < V; v0 += kEncThreads * kDenseRun) {
			// V <= 4096 < 2^14: one macro block; the last header with voxoff <= v0, then forward
			const u32 ws = mac[1];
			u32 lo = mac[0], cnt = (M > 1u ? mac[2] : B0) - lo;
			const u32 lb = lo + cnt;
			const u32 fb = lo;
			while (cnt > 0) {
				const u32 step = cnt >> 1;
				if ((hd[2u * (lo + step) + 1u] >> 18) <= v0)
					lo += step + 1u, cnt -= step + 1u;
				else
					cnt = step;
			}
			u32 h = lo > fb ? lo - 1u : fb;
			for (u32 k = 0; k < kDenseRun && v0 + k < V; ++k) {
				const u32 v = v0 + k;
				while (h + 1u < lb && (hd[2u * (h + 1u) + 1u] >> 18) <= v)
					++h;
				if ((S.cov[v >> 5] >> (v & 31u)) & 1u)
					continue;
				Color c;
				const u32 hc = hd[2u * h], hp = hd[2u * h + 1u], bpw = (hp >> 16) & 3u;
				if (bpw == 0)
					c = Color{hc, 1u << 16};
				else {
					const u32 bit = ws + (hp & 0xFFFFu) + (v - (hp >> 18)) * bpw, o = bit & 31u;
					u32 w, w1 = 0u;
					if (staged_w) {
						w = S.hdr[2u * (M + B0) + (bit >> 5)];
						if (o + bpw > 32u)
							w1 = S.hdr[2u * (M + B0) + (bit >> 5) + 1u];
					} else {
						w = cleaves[wo + (bit >> 5)];
						if (o + bpw > 32u)
							w1 = cleaves[wo + (bit >> 5) + 1u];
					}
					w = (w >> o) | (o + bpw > 32u ? w1 << (32u - o) : 0u);
					w &= (1u << bpw) - 1u;
					c = Color{hc, w | (bpw << 8) | (1u << 16)};
				}
				job_set(S, v, c);
			}
		}
	}
	__syncthreads();
}
// Staging words per job of a sync-free pass with LDS jobs (on average: each workgroup fills its
// share of the area in turn); a chunk that no longer fits goes to a_job_rest like one larger than
// kJobHdr.
constexpr u32 kEtmpJobWords = 512;
// One workgroup per writer job rebuilds the job's colors in LDS (job_rebuild) and encodes them as
// a_encode_size does (SetLeaf's size decision). A chunk of at most kJobHdr words is staged in the
// workgroup's share of etmp (etmp_cap words); a larger one, or one that no longer fits, is listed for
// a_job_rest, which rebuilds and encodes it in place once the append offsets are known.
template <bool kFull>
__global__ void __launch_bounds__(kEncThreads, 4) a_job_encode(EditParams P, const ALevel *__restrict__ lv,
                                                             u32 *__restrict__ pm, const uint4 *__restrict__ jrange,
                                                             const unsigned short *__restrict__ lidx,
                                                             const uint2 *__restrict__ ecol, u32 V,
                                                             u32 *__restrict__ info, u32 *__restrict__ app,
                                                             u32 *__restrict__ etmp, u32 etmp_cap,
                                                             u32 *__restrict__ rest) {
	__shared__ JobLds S;
	__shared__ u64 s_wm;
	const u32 leafL = P.c.color_leaf;
	const Entry *ent = lv[leafL].ent;
	const u32 n = pm[kPmHdr + leafL];
	// each workgroup stages its jobs' chunks in its own share of etmp, one after another (no
	// atomics: a counter shared by the whole grid would serialize across the chip)
	const u32 per_block = etmp_cap / gridDim.x, blk_base = blockIdx.x * per_block;
	u32 used = 0;
	for (u32 c0 = blockIdx.x; c0 < n; c0 += gridDim.x * kJobChunk)
	for (u64 m = writer_chunk(ent, n, c0, &s_wm, [&](u32 e) { app[e] = 0; }); m; m &= m - 1ull) {
		const u32 e = c0 + (u32)__builtin_ctzll(m) * gridDim.x;
		__syncthreads(); // the previous job's readers are done with S
		Entry E;
		uint4 jh;
		job_rebuild(P, lv, jrange, lidx, ecol, V, e, S, E, jh);
		const bool chunk = ctag(E.octree) == kTagLeaf;
		const u32 block = jh.z;
		// encode: a_encode_size's register path over the LDS colors
		JobEnc R;
		enc_scan_job<kFull>(S, V, R);
		EncScan tot;
		EncScan run = block_excl_scan(R.x, S.sh, tot);
		const u32 B = tot.heads, W = (tot.bits + 31u) >> 5, ds = 6u + 2u * B + W;
		const bool staged = ds <= kJobHdr && used + ds <= per_block; // block-uniform
		const u32 off = blk_base + used;
		used += staged ? ds : 0u;
		if (staged) {
			const u32 bo = 6u, wo = bo + 2u * B;
			for (u32 w = threadIdx.x; w < W; w += kEncThreads)
				S.stmp[wo + w] = 0u;
			if (threadIdx.x == 0) {
				S.stmp[0] = 0u;
				S.stmp[1] = 1u;
				S.stmp[2] = B;
				S.stmp[3] = W;
				S.stmp[4] = 0u;
				S.stmp[5] = 0u;
			}
			__syncthreads();
#pragma unroll
			for (u32 k = 0; k < kEncPer; ++k)
				if (kFull || k < R.n) {
					const u32 v = R.v0 + k, wb = unpack_wb16(S.colw[v]), bpw = (wb >> 8) & 0xFFu;
					if ((R.heads >> k) & 1u) {
						S.stmp[bo + 2u * run.heads] = S.colc[v];
						S.stmp[bo + 2u * run.heads + 1u] = (v << 18) | (bpw << 16) | run.bits;
						run.heads += 1u;
					}
					if (bpw) {
						const u32 w = wb & ((1u << bpw) - 1u), o = run.bits & 31u;
						atomicOr(&S.stmp[wo + (run.bits >> 5)], w << o);
						if (o + bpw > 32u)
							atomicOr(&S.stmp[wo + (run.bits >> 5) + 1u], w >> (32u - o));
					}
					run.bits += bpw;
				}
			__syncthreads();
			for (u32 w = threadIdx.x; w < ds; w += kEncThreads)
				etmp[(u64)off + w] = S.stmp[w];
		}
		if (threadIdx.x == 0) {
			u32 as = (ds & 1u) ? ds + 1u : ds;
			u32 inplace = 0;
			if (!P.c.keep_history && chunk) { // DAGColorPool::SetLeaf (:172-204)
				if (ds <= block)
					inplace = 1, as = 0;
				else
					as = max(block << 1, as);
			}
			info[4 * e] = ds;
			info[4 * e + 1] = as;
			info[4 * e + 2] = inplace | (staged ? 2u : 0u) | 4u; // 4: info[4 e + 3] is the staging offset
			info[4 * e + 3] = staged ? off : B;
			app[e] = as;
			if (!staged)
				rest[atomicAdd(&pm[kPmRest], 1u)] = e;
		}
	}
}
// The jobs a_job_encode listed (their chunk was not staged): rebuilt in LDS again (the same inputs
// give the same colors and size) and encoded straight into the color leaves at the offset the
// append scan gave, or over the old chunk when it fits in place (the rebuild has read it whole
// before the first store). One workgroup per job.
template <bool kFull>
__global__ void __launch_bounds__(kEncThreads, 3) a_job_rest(EditParams P, const ALevel *__restrict__ lv,
                                                           const u32 *__restrict__ pm, const uint4 *__restrict__ jrange,
                                                           const unsigned short *__restrict__ lidx,
                                                           const uint2 *__restrict__ ecol, u32 V,
                                                           const u32 *__restrict__ info, const u32 *__restrict__ appo,
                                                           u64 leaf_cap, const u32 *__restrict__ rest) {
	__shared__ JobLds S;
	const u32 leafL = P.c.color_leaf;
	u32 *ores = lv[leafL].ores;
	const u32 leaf_base = P.meta[kMetaColorLeaves];
	const u32 nrest = pm[kPmRest];
	for (u32 i = blockIdx.x; i < nrest; i += gridDim.x) {
		const u32 e = rest[i];
		__syncthreads(); // the previous job's readers are done with S
		Entry E;
		uint4 jh;
		job_rebuild(P, lv, jrange, lidx, ecol, V, e, S, E, jh);
		JobEnc R;
		enc_scan_job<kFull>(S, V, R);
		EncScan tot;
		EncScan run = block_excl_scan(R.x, S.sh, tot);
		const u32 ds = info[4 * e], as = info[4 * e + 1], inplace = info[4 * e + 2] & 1u;
		const u32 idx = inplace ? cdata(E.octree) : leaf_base + appo[e];
		u32 result = E.octree;
		if (inplace || (u64)idx + as <= leaf_cap) {
			u32 *Lw = P.cleaves;
			const u32 B = tot.heads, W = (tot.bits + 31u) >> 5;
			const u32 mo = idx + 4u, bo = mo + 2u, wo = bo + 2u * B; // one macro block (V <= 2^14)
			for (u32 w = threadIdx.x; w < W; w += kEncThreads)
				Lw[wo + w] = 0u;
			if (threadIdx.x == 0) {
				Lw[mo] = 0u;
				Lw[mo + 1u] = 0u;
				Lw[idx + 1u] = 1u;
				Lw[idx + 2u] = B;
				Lw[idx + 3u] = W;
				if (!inplace)
					Lw[idx] = as;
			}
			__syncthreads();
#pragma unroll
			for (u32 k = 0; k < kEncPer; ++k)
				if (kFull || k < R.n) {
					const u32 v = R.v0 + k, wb = unpack_wb16(S.colw[v]), bpw = (wb >> 8) & 0xFFu;
					if ((R.heads >> k) & 1u) {
						Lw[bo + 2u * run.heads] = S.colc[v];
						Lw[bo + 2u * run.heads + 1u] = (v << 18) | (bpw << 16) | run.bits;
						run.heads += 1u;
					}
					if (bpw) {
						const u32 w = wb & ((1u << bpw) - 1u), o = run.bits & 31u;
						atomicOr(&Lw[wo + (run.bits >> 5)], w << o);
						if (o + bpw > 32u)
							atomicOr(&Lw[wo + (run.bits >> 5) + 1u], w >> (32u - o));
					}
					run.bits += bpw;
				}
			if (!inplace)
				for (u32 w = ds + threadIdx.x; w < as; w += kEncThreads)
					Lw[idx + w] = 0u;
			result = inplace ? E.octree : cmake(kTagLeaf, idx);
		} else if (threadIdx.x == 0)
			atomicOr(&P.meta[kMetaOverflow], 2u);
		if (threadIdx.x == 0)
			ores[e] = result;
	}
}

// Exclusive scan of the jobs' encoded sizes in entry (DFS) order, in tiles of kScanThreads x
// kScanPer jobs: a_scan_app scans each tile in one workgroup (1024 threads x 8 consecutive jobs: a
// running sum in registers, a wave scan, the waves' totals through LDS) and records its total;
// a_scan_fix adds the totals of the tiles before each job's and writes the grand total.
constexpr u32 kScanThreads = 1024, kScanPer = 8, kScanTile = kScanThreads * kScanPer;
__global__ void __launch_bounds__(kScanThreads) a_scan_app(u32 leafL, const u32 *__restrict__ pm,
                                                           const u32 *__restrict__ app, u32 *__restrict__ appo,
                                                           u32 *__restrict__ tsum) {
	__shared__ u32 swave[kScanThreads / 64];
	const u32 n = pm[kPmHdr + leafL], lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
	for (u32 t = blockIdx.x; (u64)t * kScanTile < n; t += gridDim.x) {
		const u32 first = t * kScanTile + threadIdx.x * kScanPer;
		u32 v[kScanPer], sum = 0;
#pragma unroll
		for (u32 k = 0; k < kScanPer; ++k) {
			v[k] = first + k < n ? app[first + k] : 0u;
			sum += v[k];
		}
		u32 x = sum; // inclusive wave scan of the per-thread sums
#pragma unroll
		for (u32 d = 1; d < 64; d <<= 1) {
			const u32 y = __shfl_up(x, d);
			if (lane >= d)
				x += y;
		}
		if (lane == 63)
			swave[w] = x;
		__syncthreads();
		u32 pre = 0, tot = 0;
#pragma unroll
		for (u32 k = 0; k < kScanThreads / 64; ++k) {
			const u32 s = swave[k];
			pre += k < w ? s : 0u;
			tot += s;
		}
		u32 o = pre + x - sum;
#pragma unroll
		for (u32 k = 0; k < kScanPer; ++k) {
			if (first + k < n)
				appo[first + k] = o;
			o += v[k];
		}
		if (threadIdx.x == 0)
			tsum[t] = tot;
		__syncthreads(); // swave reuse
	}
}
__global__ void a_scan_fix(u32 leafL, u32 *__restrict__ pm, u32 *__restrict__ appo, const u32 *__restrict__ tsum) {
	const u32 n = pm[kPmHdr + leafL], nt = (u32)(((u64)n + kScanTile - 1) / kScanTile);
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		u32 total = 0;
		for (u32 t = 0; t < nt; ++t)
			total += tsum[t];
		pm[kPmLeafTotal] = total;
	}
	for (u32 i = kScanTile + blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
		const u32 t = i / kScanTile;
		u32 pre = 0;
		for (u32 k = 0; k < t; ++k)
			pre += ldg(tsum, k);
		appo[i] += pre;
	}
}
__global__ void __launch_bounds__(kEncThreads) a_encode_write(EditParams P, const ALevel *__restrict__ lv,
                                                               const u32 *__restrict__ pm, const Color *__restrict__ dense,
                                                               u32 V, const u32 *__restrict__ info,
                                                               const u32 *__restrict__ appo, u64 leaf_cap,
                                                               const u32 *__restrict__ etmp) {
	// the staged chunks: a copy each (the writers whose chunk was too large are in a_encode_rest's list)
	const u32 leafL = P.c.color_leaf;
	const Entry *ent = lv[leafL].ent;
	u32 *ores = lv[leafL].ores;
	const u32 n = pm[kPmHdr + leafL];
	const u32 leaf_base = P.meta[kMetaColorLeaves];
	__shared__ u64 s_wm;
	for (u32 c0 = blockIdx.x; c0 < n; c0 += gridDim.x * kJobChunk)
	for (u64 m = writer_chunk(ent, n, c0, &s_wm, [&](u32 e) { ores[e] = ldg(&ent[e].octree, 0); }); m;
	     m &= m - 1ull) {
		const u32 e = c0 + (u32)__builtin_ctzll(m) * gridDim.x;
		const Entry E = ent[e];
		u32 result = E.octree;
		if (E.sf & kSfWriter) {
			u32 ds = info[4 * e], as = info[4 * e + 1], inplace = info[4 * e + 2] & 1u, B = info[4 * e + 3];
			const bool staged = (info[4 * e + 2] >> 1) & 1u;
			u32 idx = inplace ? cdata(E.octree) : leaf_base + appo[e];
			if (!staged)
				continue; // a_encode_rest encodes it (and sets its result)
			if (inplace || (u64)idx + as <= leaf_cap) { // the chunk a_encode_size built
				u32 *L = P.cleaves;
				const u32 *T = etmp + ((info[4 * e + 2] & 4u) ? (u64)info[4 * e + 3] : (u64)e * kEncTmp);
				for (u32 w = 1u + threadIdx.x; w < ds; w += kEncThreads)
					L[idx + w] = T[w];
				if (!inplace) {
					if (threadIdx.x == 0)
						L[idx] = as;
					for (u32 w = ds + threadIdx.x; w < as; w += kEncThreads)
						L[idx + w] = 0u;
				}
				result = inplace ? E.octree : cmake(kTagLeaf, idx);
			} else if (threadIdx.x == 0)
				atomicOr(&P.meta[kMetaOverflow], 2u);
		}
		if (threadIdx.x == 0)
			ores[e] = result;
	}
}
// The writer jobs whose encoded chunk did not fit a staging slot (listed by a_encode_size): encoded
// from the dense colors, one workgroup per job.
__global__ void __launch_bounds__(kEncThreads) a_encode_rest(EditParams P, const ALevel *__restrict__ lv,
                                                              const u32 *__restrict__ pm, const Color *__restrict__ dense,
                                                              u32 V, const u32 *__restrict__ info,
                                                              const u32 *__restrict__ appo, u64 leaf_cap,
                                                              const u32 *__restrict__ rest) {
	__shared__ EncScan sh[kEncThreads];
	__shared__ u32 macro_ws[160];
	const u32 leafL = P.c.color_leaf;
	const Entry *ent = lv[leafL].ent;
	u32 *ores = lv[leafL].ores;
	const u32 leaf_base = P.meta[kMetaColorLeaves];
	const u32 nrest = pm[kPmRest];
	for (u32 i = blockIdx.x; i < nrest; i += gridDim.x) {
		const u32 e = rest[i];
		const Entry E = ent[e];
		u32 result = E.octree;
		if (E.sf & kSfWriter) {
			u32 ds = info[4 * e], as = info[4 * e + 1], inplace = info[4 * e + 2] & 1u, B = info[4 * e + 3];
			u32 idx = inplace ? cdata(E.octree) : leaf_base + appo[e];
			if ((inplace || (u64)idx + as <= leaf_cap) && V <= kEncPer * kEncThreads) {
				u32 *L = P.cleaves;
				EncRegs R;
				enc_load(dense + (u64)e * V, V, R);
				EncScan tot;
				EncScan run = block_excl_scan(R.x, sh, tot);
				const u32 W = (tot.bits + 31u) >> 5;
				const u32 mo = idx + 4u, bo = mo + 2u, wo = bo + 2u * B; // one macro block (V <= 2^14)
				for (u32 w = threadIdx.x; w < W; w += kEncThreads)
					L[wo + w] = 0u;
				if (threadIdx.x == 0) {
					L[mo] = 0u;
					L[mo + 1u] = 0u;
					L[idx + 1u] = 1u;
					L[idx + 2u] = B;
					L[idx + 3u] = W;
					if (!inplace)
						L[idx] = as;
				}
				__syncthreads();
#pragma unroll
				for (u32 k = 0; k < kEncPer; ++k)
					if (k < R.n) {
						const u32 v = R.v0 + k, bpw = (R.w[k] >> 8) & 0xFFu;
						if ((R.heads >> k) & 1u) {
							L[bo + 2u * run.heads] = R.c[k];
							L[bo + 2u * run.heads + 1u] = (v << 18) | (bpw << 16) | run.bits;
							run.heads += 1u;
						}
						if (bpw) {
							u32 w = R.w[k] & ((1u << bpw) - 1u), o = run.bits & 31u;
							atomicOr(&L[wo + (run.bits >> 5)], w << o);
							if (o + bpw > 32u)
								atomicOr(&L[wo + (run.bits >> 5) + 1u], w >> (32u - o));
						}
						run.bits += bpw;
					}
				if (!inplace)
					for (u32 w = ds + threadIdx.x; w < as; w += kEncThreads)
						L[idx + w] = 0u;
				result = inplace ? E.octree : cmake(kTagLeaf, idx);
			} else if (inplace || (u64)idx + as <= leaf_cap) { // color leaves of more than 4096 voxels
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
		if (threadIdx.x == 0)
			ores[e] = result;
		__syncthreads(); // macro_ws / sh reuse
	}
}

// SetNode (DAGColorPool.hpp:139-171) above the color leaves, level by level: bottom-up, each
// entry's decision (its new octree state; kColorNewNode when it appends a node) and the color
// nodes its subtree appends; then top-down, post-order indices (a node after its children's
// subtrees: a new child's index is its base + count - 1) and every appended node written with its
// children's final pointers. The per-entry steps below take the child level's decisions through
// an accessor, so the same code serves levels in LDS, levels in global memory and grid launches.
//
// Decision of entry e of level l. child(ci, c) returns proceeding child ci's decision and adds its
// append count to c.
template <class Child>
__device__ inline void setnode_decide(const EditParams &P, const ALevel &A, const ALevel &C, u32 e, Child child,
                                      u32 &r_out, u32 &c_out) {
	const Entry E = ldg(A.ent, e);
	u32 r = E.octree, c = 0;
	if (!(E.sf & kSfFinal) && (E.sf & kSfActive)) {
		const u32 m = ldg(A.pmask, e) & 0xFFu, fc = ldg(A.first_child, e);
		u32 ch[8];
		bool all_null = true, same_color;
		for (u32 i = 0; i < 8; ++i) {
			if ((m >> i) & 1u) {
				const u32 ci = fc + (u32)__builtin_popcount(m & ((1u << i) - 1u));
				ch[i] = ci >= C.cap ? kColorNull : child(ci, c);
			} else
				ch[i] = ldg(A.okids, (u64)e * 8u + i);
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
		c += r == kColorNewNode ? 1u : 0u;
	}
	r_out = r;
	c_out = c;
}
// Top-down step of entry e (first post-order index run, decision r): place(ci, run) hands proceeding
// child ci its base, advances run past its subtree and returns its final pointer. Writes the
// entry's node if it is new and returns the entry's final pointer.
template <class Place>
__device__ inline u32 setnode_place(const EditParams &P, const ALevel &A, const ALevel &C, u32 e, u32 run, u32 r,
                                    u64 node_cap, Place place) {
	const u32 m = ldg(A.pmask, e) & 0xFFu, fc = ldg(A.first_child, e);
	u32 ch[8];
	for (u32 i = 0; i < 8; ++i) {
		if ((m >> i) & 1u) {
			const u32 ci = fc + (u32)__builtin_popcount(m & ((1u << i) - 1u));
			ch[i] = ci >= C.cap ? kColorNull : place(ci, run);
		} else
			ch[i] = ldg(A.okids, (u64)e * 8u + i);
	}
	if (r == kColorNewNode) {
		if ((u64)run < node_cap) {
			for (u32 i = 0; i < 8; ++i)
				P.cnodes[(u64)run * 8u + i] = ch[i];
			r = cmake(kTagNode, run);
		} else {
			r = ldg(&A.ent[e].octree, 0);
			atomicOr(&P.meta[kMetaOverflow], 2u);
		}
	}
	return r;
}
// A new child's final pointer from its base + count (its own index is the last of its subtree).
__device__ inline u32 setnode_child_final(const ALevel &C, u32 ci, u32 cr, u32 end, u64 node_cap) {
	if (cr != kColorNewNode)
		return cr;
	const u32 idx = end - 1u;
	return (u64)idx < node_cap ? cmake(kTagNode, idx) : ldg(&C.ent[ci].octree, 0);
}
// Child accessors for a child level kept in global memory (ores / ccnt / cbase); the color leaves'
// results are final (no count, no base).
struct SetNodeGlobal {
	const ALevel &C;
	bool leaf;
	u64 node_cap;
	__device__ u32 operator()(u32 ci, u32 &c) const {
		if (!leaf)
			c += ldg(C.ccnt, ci);
		return ldg(C.ores, ci);
	}
	__device__ u32 place(u32 ci, u32 &run) const {
		const u32 cr = ldg(C.ores, ci);
		if (leaf)
			return cr;
		C.cbase[ci] = run;
		run += ldg(C.ccnt, ci);
		return setnode_child_final(C, ci, cr, run, node_cap);
	}
};
// Grid launches for the large levels [lb, leafL): the decisions bottom-up (one launch per level,
// before the single workgroup), then the placement top-down (after it).
__global__ void a_setnode_up(EditParams P, const ALevel *__restrict__ lv, const u32 *__restrict__ pm, u32 l) {
	const ALevel A = lv[l], C = lv[l + 1];
	const u32 n = pm[kPmHdr + l];
	const SetNodeGlobal g{C, l + 1u == P.c.color_leaf, 0};
	for (u32 e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x) {
		u32 r, c;
		setnode_decide(P, A, C, e, g, r, c);
		A.ores[e] = r;
		A.ccnt[e] = c;
	}
}
__global__ void a_setnode_down(EditParams P, const ALevel *__restrict__ lv, const u32 *__restrict__ pm, u32 l,
                               u64 node_cap) {
	const ALevel A = lv[l], C = lv[l + 1];
	const u32 n = pm[kPmHdr + l];
	const SetNodeGlobal g{C, l + 1u == P.c.color_leaf, node_cap};
	for (u32 e = blockIdx.x * blockDim.x + threadIdx.x; e < n; e += gridDim.x * blockDim.x)
		setnode_place(P, A, C, e, ldg(A.cbase, e), ldg(A.ores, e), node_cap,
		              [&](u32 ci, u32 &run) { return g.place(ci, run); });
}
// The levels [0, lb) in one workgroup with their decisions in global memory (the fallback when
// they do not fit LDS); the levels from lb on are in global memory already.
__device__ inline void setnode_block(const EditParams &P, const ALevel *__restrict__ lv, const u32 *__restrict__ pm,
                                     u32 *__restrict__ roots, u64 node_cap, u32 lb) {
	const u32 leafL = P.c.color_leaf;
	for (int l = (int)lb - 1; l >= 0; --l) {
		const ALevel A = lv[l], C = lv[l + 1];
		const u32 n = pm[kPmHdr + l];
		const SetNodeGlobal g{C, (u32)l + 1u == leafL, node_cap};
		for (u32 e = threadIdx.x; e < n; e += blockDim.x) {
			u32 r, c;
			setnode_decide(P, A, C, e, g, r, c);
			A.ores[e] = r;
			A.ccnt[e] = c;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0 && pm[kPmHdr] > 0)
		lv[0].cbase[0] = P.meta[kMetaColorNodes];
	__syncthreads();
	for (u32 l = 0; l < lb; ++l) {
		const ALevel A = lv[l], C = lv[l + 1];
		const u32 n = pm[kPmHdr + l];
		const SetNodeGlobal g{C, l + 1u == leafL, node_cap};
		for (u32 e = threadIdx.x; e < n; e += blockDim.x) {
			const u32 r = setnode_place(P, A, C, e, A.cbase[e], A.ores[e], node_cap,
			                            [&](u32 ci, u32 &run) { return g.place(ci, run); });
			if (l == 0 && e == 0) {
				roots[1] = r;
				P.meta[kMetaColorNodes] += A.ccnt[0];
			}
		}
		__syncthreads();
	}
}

// ---- small levels in one workgroup, with the level bookkeeping in LDS
constexpr u32 kSmallLevel = 1024; // entries of a level handled inside a_bottom
constexpr u32 kSTab = 2048;       // LDS dedup / bucket-map slots (load factor <= 1/2)
static_assert(kSTab == 1u << 11 && kSmallLevel * 2u <= kSTab);
constexpr u32 kColorLds = 4096;   // SetNode entries (all levels above the color leaves) kept in LDS

constexpr u32 kStage = 5 * kSTab; // bucket staging words of the build (shares the dedup tables' LDS)
// the wave-0 levels' slot counts in the stage: hash slots, then bucket slots
constexpr u32 kDupSlots = 2048, kDupBucketSlots = 8192;
static_assert(kDupSlots + kDupBucketSlots <= kStage, "the slot counts fit the stage");

struct BottomLds {
	// per entry; [2] = this level / its child level, whose results the build resolves
	u32 res[2][kSmallLevel];  // final pointer, or kNull while a group result is pending
	union {
		struct {
			u32 slot[2][kSmallLevel]; // dedup slot of an entry with new content, else kNull
			u32 ptr[2][kSmallLevel];  // old pointer (kept when the bucket is full)
		};
		// the wave-0 levels: every entry's join inputs (3 uint4 each), loaded by the whole block
		__align__(16) u32 jin[4 * kSmallLevel];
	};
	u32 content[kSmallLevel][9];
	u32 csz[kSmallLevel], hash[kSmallLevel], bucket[kSmallLevel], bw[kSmallLevel], bmap[kSmallLevel];
	u32 glist[kSmallLevel], gnext[kSmallLevel]; // groups (their slots) and the per-bucket group lists
	u32 pslot[kSmallLevel];                     // parent entry << 3 | slot in the parent
	u32 gres[kSTab];                            // result per dedup slot (read by the parent's build)
	union {
		struct { // dedup and walk: cleared after the build
			u32 tab[kSTab];                                        // lowest entry per content
			u32 bkey[kSTab], bmin[kSTab], bhead[kSTab], bn[kSTab]; // bucket map: lowest entry, groups
		};
		__align__(16) u32 stage[kStage]; // build: the scanned buckets' words
	};
	unsigned char gfail[kSTab];
	unsigned char cfresh[kSmallLevel][8]; // by parent and slot: the child's result is its fresh append
	u32 gcount, stage_top;
};
// the table clear writes the stage as uint4: tab, bkey, bmin, bhead (kNull) then bn (0)
static_assert(offsetof(BottomLds, tab) == offsetof(BottomLds, stage) &&
              offsetof(BottomLds, bn) == offsetof(BottomLds, stage) + 4 * kSTab * sizeof(u32) &&
              offsetof(BottomLds, stage) % 16 == 0);
struct ColorLds {
	u32 r[kColorLds], cnt[kColorLds], base[kColorLds];
	u32 off[HD_MAX_LEVELS + 1];
	u32 fits;
};
union ALds {
	BottomLds b;
	ColorLds c;
};

__device__ inline bool lds_same(const BottomLds &S, u32 a, u32 b) {
	if (S.csz[a] != S.csz[b])
		return false;
	for (u32 k = 0; k < S.csz[a]; ++k)
		if (S.content[a][k] != S.content[b][k])
			return false;
	return true;
}
// Final pointer of entry c of a finished small level (its results are resolved after the walk).
// The load is unconditional (index clamped) so the lookups of a node's children overlap.
__device__ inline u32 lds_result(const BottomLds &S, u32 p, u32 c) {
	const u32 v = S.res[p][min(c, kSmallLevel - 1u)];
	return c < kSmallLevel ? v : kNull;
}

// Walk of a small level's bucket by its lowest new entry (the first group's lowest entry): the
// bucket's groups are replayed in the order of their lowest entries (find among this walk's
// appends, else append_node).
__device__ inline void small_walk_bucket(const EditParams &P, bool leaf, BottomLds &S, u32 e, u32 *__restrict__ acc) {
	const u32 m = S.bmap[e], b = S.bucket[e], bw0 = S.bw[e], cnt = S.bn[m];
	u32 bw = bw0, appended = 0;
	bool any_fail = false;
	auto upsert = [&](u32 g2) {
		const u32 s2 = S.glist[g2], j = S.tab[s2];
		const u32 *node = S.content[j];
		u32 nsz = S.csz[j];
		u32 f = kNull; // distinct groups: nothing appended by this walk can match (see walk_entry)
		unsigned char fail = 0;
		{
			f = append_to_bucket(P.words, P.c, b, bw, node, nsz);
			if (f == kNull)
				fail = 1, any_fail = true;
			else
				++appended;
		}
		S.gres[s2] = f;
		S.gfail[s2] = fail;
	};
	if (cnt == 1)
		upsert(S.bhead[m]);
	else if (cnt <= kGather) {
		u32 ord[kGather], key[kGather], n = 0;
		for (u32 x = S.bhead[m]; x != kNull && n < kGather; x = S.gnext[x]) {
			u32 r2 = S.tab[S.glist[x]], y = n++;
			while (y > 0 && key[y - 1] > r2) {
				key[y] = key[y - 1];
				ord[y] = ord[y - 1];
				--y;
			}
			key[y] = r2;
			ord[y] = x;
		}
		for (u32 x = 0; x < n; ++x)
			upsert(ord[x]);
	} else {
		u32 last = 0;
		for (u32 k = 0; k < cnt; ++k) {
			u32 best = kNull, bg = kNull;
			for (u32 x = S.bhead[m]; x != kNull; x = S.gnext[x]) {
				u32 r2 = S.tab[S.glist[x]];
				if ((k == 0 || r2 > last) && r2 < best)
					best = r2, bg = x;
			}
			upsert(bg);
			last = best;
		}
	}
	stg(P.bucket_words, b, bw);
	if (appended)
		atomicAdd(&acc[kAccUpserts], appended);
	if (any_fail)
		atomicOr(&P.meta[kMetaOverflow], 1u);
}

// SetNode for the levels [0, lb) with their decisions, counts and bases in LDS (level lb and
// below, when lb < color_leaf, come from global memory: the grid launches' levels).
__device__ inline void setnode_lds(const EditParams &P, const ALevel *__restrict__ lv, const u32 *__restrict__ pm,
                                   u32 *__restrict__ roots, u64 node_cap, ColorLds &S, u32 lb) {
	const u32 leafL = P.c.color_leaf;
	for (int l = (int)lb - 1; l >= 0; --l) { // decisions, bottom-up
		const ALevel A = lv[l], C = lv[l + 1];
		const u32 n = pm[kPmHdr + l], o = S.off[l], oc = S.off[l + 1];
		const SetNodeGlobal g{C, (u32)l + 1u == leafL, node_cap};
		const bool cg = (u32)l + 1u >= lb;
		for (u32 e = threadIdx.x; e < n; e += blockDim.x) {
			u32 r, c;
			setnode_decide(P, A, C, e, [&](u32 ci, u32 &cc) {
				if (cg)
					return g(ci, cc);
				cc += S.cnt[oc + ci];
				return S.r[oc + ci];
			}, r, c);
			S.r[o + e] = r;
			S.cnt[o + e] = c;
		}
		__syncthreads();
	}
	if (threadIdx.x == 0 && pm[kPmHdr] > 0)
		S.base[S.off[0]] = P.meta[kMetaColorNodes];
	__syncthreads();
	for (u32 l = 0; l < lb; ++l) { // top-down: indices + writes
		const ALevel A = lv[l], C = lv[l + 1];
		const u32 n = pm[kPmHdr + l], o = S.off[l], oc = S.off[l + 1];
		const SetNodeGlobal g{C, l + 1u == leafL, node_cap};
		const bool cg = l + 1u >= lb;
		for (u32 e = threadIdx.x; e < n; e += blockDim.x) {
			const u32 r = setnode_place(P, A, C, e, S.base[o + e], S.r[o + e], node_cap, [&](u32 ci, u32 &run) {
				if (cg)
					return g.place(ci, run);
				const u32 cr = S.r[oc + ci];
				S.base[oc + ci] = run;
				run += S.cnt[oc + ci];
				return setnode_child_final(C, ci, cr, run, node_cap);
			});
			if (l == 0 && e == 0) {
				roots[1] = r;
				P.meta[kMetaColorNodes] += S.cnt[o];
			}
		}
		__syncthreads();
	}
}

// The levels [0, l_top] when each has at most 64 entries, by wave 0 alone (no workgroup barriers,
// no LDS tables): lane e builds entry e as the workgroup path does; the lanes whose content is new
// then form groups of identical contents, handled one by one in ascending order of their first
// entries -- the order in which the workgroup walk appends a bucket's groups -- with the group's
// first lane appending it and every member taking the result (its old pointer when the bucket is
// full); each result goes to the parent's content row at its slot.
__device__ inline void tiny_levels(const EditParams &P, int l_top, int l_from, const ALevel *s_lv, const u32 *s_n,
                                   const GroupTable &G, BottomLds &S, u32 *__restrict__ acc, JoinIn pre, bool have_pre,
                                   bool preloaded) {
	const u32 lane = threadIdx.x, L = P.c.L;
	u32 jbase = 0; // preloaded: the level's first entry in S.jin
	for (int l = l_top; l >= 0; --l) {
		const u32 n = s_n[l];
		const ALevel A = s_lv[l];
		const bool leaf = (u32)l + 1 == L, act = lane < n;
		const u32 cur = l & 1;
		u32 fin = kNull, optr = kNull, b = 0, bw = 0, h = 0, k = 0, pidx = 0;
		u32 c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
		bool need = false, fresh_out = false;
		if (act) {
			if (leaf) {
				pidx = (ldg(&A.ent[lane].parent, 0) << 3) | (ldg(&A.ent[lane].sf, 0) & 7u);
				if (!ldg(A.pend, lane))
					fin = ldg(A.status, lane);
				else if ((fin = ldg(A.res, lane)) == kNull) {
					need = true;
					c[0] = ldg(A.content, 2 * (u64)lane);
					c[1] = ldg(A.content, 2 * (u64)lane + 1);
					b = ldg(A.bucket, lane);
					k = 2;
					h = murmur_leaf(c[0], c[1]);
					bw = ldg(P.bucket_words, b);
					optr = ldg(&A.ent[lane].ptr, 0);
				}
			} else {
				JoinIn q;
				if (preloaded) {
					const uint4 *J = reinterpret_cast<const uint4 *>(S.jin) + 3u * (jbase + lane);
					const uint4 a = J[0];
					q.ptr = a.x, q.pmw = a.y, q.fc = a.z, q.pidx = a.w;
					q.k0 = J[1];
					q.k1 = J[2];
				} else
					q = have_pre ? pre : load_join(A, lane);
				pidx = q.pidx;
				u32 cs[8];
				bool fresh = false;
				if (l < l_from) {
#pragma unroll
					for (u32 i = 0; i < 8; ++i) {
						cs[i] = S.content[lane][i];
						fresh |= ((q.pmw >> i) & 1u) && S.cfresh[lane][i];
					}
				} else
					children_by_slot(q, [&](u32 ch) { return child_result_fresh(s_lv[l + 1], G, ch); }, cs, fresh);
				Built o = build_node(P, (u32)l, q, cs);
				fin = o.fin;
				if (o.k) {
					if (fresh) { // a fresh child: a new node, no find
						fin = kNull;
						bw = ldg(P.bucket_words, o.b);
					} else {
						fin = find_in_bucket_staged(P.words, P.bucket_words, P.c, o.b, o.c, o.k,
						                            headers_by_value(P.c, (u32)l), bw, S.stage, &S.stage_top, kStage);
						if (fin == kFindRetry) // the stage is full: scan in global memory
							fin = find_in_global(P.words, P.c, (u64)o.b << P.c.wbpb(), bw, o.c, o.k);
					}
					if (fin == kNull) {
						need = true;
#pragma unroll
						for (u32 j = 0; j < 9; ++j)
							c[j] = o.c[j];
						k = o.k;
						h = o.h;
						b = o.b;
						optr = q.ptr;
					}
				}
			}
		}
		// the next level's entries: their loads overlap this level's groups
		have_pre = !preloaded && l > 0 && lane < s_n[l - 1];
		if (have_pre)
			pre = load_join(s_lv[l - 1], lane);
		if (!leaf)
			jbase += n;
		// groups of identical new contents: each lane's leader (the group's lowest lane); a lane whose
		// hash no other new content shares leads its own group, the others are grouped by content.
		// dup: another new content has this lane's hash slot; below: a lower lane's new content has this
		// lane's bucket slot. Slots are counted in the stage, free until the next level's finds: equal
		// values always share a slot, so a collision of distinct values only sends lanes to the exact
		// steps below (the content grouping, the ordered appends).
		const u64 news = __ballot(need);
		bool dup = false, below = false;
		if (news) {
			u32 *hc = S.stage, *bl = S.stage + kDupSlots;
			const u32 hs = h & (kDupSlots - 1u), bs = b & (kDupBucketSlots - 1u);
			if (need) {
				hc[hs] = 0u;
				bl[bs] = 64u;
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			if (need) {
				atomicAdd(&hc[hs], 1u);
				atomicMin(&bl[bs], lane);
			}
			__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
			__builtin_amdgcn_wave_barrier();
			__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
			if (need) {
				dup = hc[hs] > 1u;
				below = bl[bs] < lane;
			}
		}
		u64 pending = __ballot(dup), leaders = news & ~pending;
		u32 lead = lane;
		while (pending) {
			const u32 ld = (u32)__builtin_ctzll(pending);
			const u32 lh = __builtin_amdgcn_readlane(h, ld), lk = __builtin_amdgcn_readlane(k, ld);
			bool same = dup && h == lh && k == lk;
#pragma unroll
			for (u32 j = 0; j < 9; ++j)
				same = same && c[j] == __builtin_amdgcn_readlane(c[j], ld);
			if (same)
				lead = ld;
			leaders |= 1ull << ld;
			pending &= ~__ballot(same);
		}
		const bool leader = (leaders >> lane) & 1ull;
		// leaders in distinct buckets append at once; a bucket with several (rare: a level here has at
		// most 64 entries) takes its groups in lane order, as the reference appends them. A leader that
		// shares its bucket with a lower leader has below set (that leader's content is in a lower lane,
		// in the same slot); below may also be set by a slot collision, which only costs the ordered
		// appends.
		u32 f = kNull;
		if (__ballot(leader && below) == 0ull) {
			if (leader) {
				u32 nbw = bw;
				f = append_regs(P.words, P.c, b, nbw, c, k);
				stg(P.bucket_words, b, nbw);
			}
		} else
			for (u64 m = leaders; m; m &= m - 1ull) {
				const u32 ld = (u32)__builtin_ctzll(m);
				const u32 lb = __builtin_amdgcn_readlane(b, ld);
				u32 g = kNull, nbw = bw;
				if (lane == ld) {
					g = append_regs(P.words, P.c, lb, nbw, c, k);
					stg(P.bucket_words, lb, nbw);
				}
				nbw = __builtin_amdgcn_readlane(nbw, ld);
				f = lane == ld ? g : f;
				if (need && b == lb)
					bw = nbw; // later groups of this bucket append after it
			}
		const u32 appended = (u32)__popcll(__ballot(leader && f != kNull));
		const bool full = __ballot(leader && f == kNull) != 0ull;
		const u32 fl = (u32)__shfl((int)f, (int)lead); // the group's result
		if (need) {
			fin = fl == kNull ? optr : fl;
			fresh_out = fl != kNull;
		}
		if (lane == 0) {
			if (appended)
				atomicAdd(&acc[kAccUpserts], appended);
			if (full)
				atomicOr(&P.meta[kMetaOverflow], 1u);
		}
		if (act) {
			S.res[cur][lane] = fin;
			if (l > 0 && (pidx >> 3) < kSmallLevel) {
				S.content[pidx >> 3][pidx & 7u] = fin;
				S.cfresh[pidx >> 3][pidx & 7u] = fresh_out;
			}
		}
		S.stage_top = 0; // every lane writes the same value; the next level allocates from 0
		__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
		__builtin_amdgcn_wave_barrier();
		__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
	}
}

// Single workgroup: the levels [0, l_from] (each at most kSmallLevel entries) bottom-up -- build,
// dedup and walk with the bookkeeping in LDS -- then the root result and the pass statistics (a
// colored pass's SetNode runs on the color stream, a_setnode).
__global__ void __launch_bounds__(kAT) a_bottom(EditParams P, const ALevel *__restrict__ lv, const u32 *__restrict__ pm,
                                                int l_from, GroupTable G, u32 *__restrict__ roots,
                                                u32 *__restrict__ acc, u64 node_cap) {
	__shared__ ALds U;
	__shared__ ALevel s_lv[HD_MAX_LEVELS]; // the small levels' arrays and entry counts, read once
	__shared__ u32 s_n[HD_MAX_LEVELS];
	__shared__ Pair s_scan[kAT / 64];
	BottomLds &S = U.b;
	const u32 L = P.c.L;
	if (l_from >= 0) { // levels 0 .. l_from + 1 (the first small level's children may be large)
		const u32 nw = min((u32)l_from + 2u, L) * (u32)(sizeof(ALevel) / 4);
		for (u32 i = threadIdx.x; i < nw; i += blockDim.x)
			reinterpret_cast<u32 *>(s_lv)[i] = reinterpret_cast<const u32 *>(lv)[i];
		for (u32 i = threadIdx.x; i <= (u32)l_from; i += blockDim.x)
			s_n[i] = min(pm[kPmHdr + i], kSmallLevel);
	}
	if (threadIdx.x == 0)
		S.stage_top = 0;
	__syncthreads();
	JoinIn pre{};         // this thread's first entry of the level, loaded during the previous level
	bool have_pre = false;
	for (int l = l_from; l >= 0; --l) {
		bool tiny = true; // this level and all above have at most 64 entries: wave 0 finishes alone
		for (int i = 0; i <= l; ++i)
			tiny = tiny && s_n[i] <= 64u;
		if (tiny) {
			// every wave-0 level's join inputs come from the top-down only: the whole block loads
			// them at once, so wave 0 waits for memory once per level (its bucket counters)
			u32 items = 0;
			for (int i = 0; i <= l; ++i)
				items += (u32)i + 1u < L ? s_n[i] : 0u;
			const bool preloaded = 3u * items <= kSmallLevel; // 4 kSmallLevel words, 12 per entry
			if (preloaded)
				for (u32 it = threadIdx.x; it < items; it += blockDim.x) {
					int li = (u32)l + 1u < L ? l : l - 1; // levels l, l - 1, .. 0 (no leaf level)
					u32 base = 0;
					while (it >= base + s_n[li])
						base += s_n[li--];
					const JoinIn q = load_join(s_lv[li], it - base);
					uint4 *J = reinterpret_cast<uint4 *>(S.jin) + 3u * it;
					J[0] = make_uint4(q.ptr, q.pmw, q.fc, q.pidx);
					J[1] = q.k0;
					J[2] = q.k1;
				}
			__syncthreads();
			if (threadIdx.x < 64)
				tiny_levels(P, l, l_from, s_lv, s_n, G, S, acc, pre, have_pre, preloaded);
			__syncthreads();
			break;
		}
		const ALevel A = s_lv[l];
		const u32 n = s_n[l];
		const bool leaf = (u32)l + 1 == L;
		const bool hv = headers_by_value(P.c, (u32)l);
		const u32 cur = l & 1;
		// build (or take a_leaf's result) + the find among the bucket's existing nodes, in rounds of
		// blockDim entries; an entry whose bucket found the stage full is built again after a reset
		for (u32 e0 = 0; e0 < n; e0 += blockDim.x) {
			const u32 e = e0 + threadIdx.x;
			if (leaf) {
				if (e < n) {
					u32 fin = kNull, optr = kNull;
					bool need = false;
					const u32 par = ldg(&A.ent[e].parent, 0), sf = ldg(&A.ent[e].sf, 0);
					S.pslot[e] = (par << 3) | (sf & 7u);
					if (!ldg(A.pend, e))
						fin = ldg(A.status, e);
					else if ((fin = ldg(A.res, e)) == kNull) {
						need = true;
						u32 c0 = ldg(A.content, 2 * (u64)e), c1 = ldg(A.content, 2 * (u64)e + 1);
						u32 b = ldg(A.bucket, e);
						S.content[e][0] = c0;
						S.content[e][1] = c1;
						S.csz[e] = 2;
						S.hash[e] = murmur_leaf(c0, c1);
						S.bucket[e] = b;
						S.bw[e] = ldg(P.bucket_words, b);
						optr = ldg(&A.ent[e].ptr, 0);
					}
					S.res[cur][e] = fin;
					S.slot[cur][e] = need ? 0u : kNull;
					S.ptr[cur][e] = optr;
				}
				__syncthreads();
				continue;
			}
			// join, then the find among the bucket's existing nodes: the block copies the buckets
			// of its entries to the stage with coalesced loads (a bucket per thread would need a
			// cache line per lane), rounds of as many buckets as the stage holds
			bool todo = e < n;
			Built o{};
			o.k = 0;
			u32 bwv = 0, optr = kNull, pidx = 0;
			if (todo) {
				const JoinIn q = (have_pre && e0 == 0) ? pre : load_join(A, e);
				u32 cs[8];
				bool fresh = false;
				if (l < l_from) { // the children's results sit in this entry's content row, by slot
#pragma unroll
					for (u32 i = 0; i < 8; ++i) {
						cs[i] = S.content[e][i];
						fresh |= ((q.pmw >> i) & 1u) && S.cfresh[e][i];
					}
				} else
					children_by_slot(q, [&](u32 c) { return child_result_fresh(s_lv[l + 1], G, c); }, cs, fresh);
				o = build_node(P, (u32)l, q, cs);
				o.need = fresh;
				optr = q.ptr;
				pidx = q.pidx;
				if (o.k)
					bwv = ldg(P.bucket_words, o.b);
			}
			// a bucket larger than the whole stage is scanned in global memory; an entry with a
			// fresh child skips the find
			const bool huge = ((bwv + 3u) & ~3u) > kStage;
			for (;;) {
				const u32 len = todo && o.k && !huge && !o.need ? (bwv + 3u) & ~3u : 0u;
				Pair tot;
				const Pair off = block_excl_scan2(Pair{len, 0u}, s_scan, tot);
				const bool fits = off.a + len <= kStage;
				S.glist[threadIdx.x] = fits ? off.a : kStage; // non-decreasing in thread order
				S.gnext[threadIdx.x] = o.b;
				S.bmap[threadIdx.x] = fits ? len : 0u; // a unit past its owner's words is not copied
				__syncthreads();
				const u32 units = min(tot.a, kStage) >> 2;
				const u32 wbpb = P.c.wbpb();
#pragma unroll
				for (u32 half = 0; half < 2; ++half) {
					uint4 v[5];
					u32 at[5];
#pragma unroll
					for (u32 k = 0; k < 5; ++k) {
						const u32 u = threadIdx.x + (half * 5u + k) * kAT;
						at[k] = kNull;
						if (u < units) { // owner: the last thread whose stage offset is <= 4u
							u32 lo = 0, hi = kAT;
							while (hi - lo > 1) {
								const u32 mid = (lo + hi) >> 1;
								if (S.glist[mid] <= 4u * u)
									lo = mid;
								else
									hi = mid;
							}
							const u32 w = 4u * u - S.glist[lo];
							if (w < S.bmap[lo]) {
								v[k] = ldg(reinterpret_cast<const uint4 *>(P.words + ((u64)S.gnext[lo] << wbpb) + w), 0);
								at[k] = u;
							}
						}
					}
#pragma unroll
					for (u32 k = 0; k < 5; ++k)
						if (at[k] != kNull)
							reinterpret_cast<uint4 *>(S.stage)[at[k]] = v[k];
				}
				__syncthreads();
				if (todo && (!o.k || len == 0 || fits)) {
					u32 fin = o.fin;
					bool need = false;
					if (o.k) {
						const u64 base = (u64)o.b << wbpb;
						fin = o.need   ? kNull
						      : huge   ? find_in_global(P.words, P.c, base, bwv, o.c, o.k)
						      : len == 0 ? kNull
						      : hv       ? find_in_words_hv(S.stage + off.a, base, bwv, o.c, o.k)
						                 : find_in_words(S.stage + off.a, P.c, base, bwv, o.c, o.k);
						if (fin == kNull) {
							need = true;
#pragma unroll
							for (u32 j = 0; j < 9; ++j)
								if (j < o.k)
									S.content[e][j] = o.c[j];
							S.csz[e] = o.k;
							S.hash[e] = o.h;
							S.bucket[e] = o.b;
							S.bw[e] = bwv;
						}
					}
					S.pslot[e] = pidx;
					S.res[cur][e] = fin;
					S.slot[cur][e] = need ? 0u : kNull;
					S.ptr[cur][e] = need ? optr : kNull;
					todo = false;
				}
				if (!__syncthreads_or(todo))
					break;
			}
		}
		// the next level's first entry of this thread: its loads overlap this level's dedup and walk
		have_pre = l > 0 && threadIdx.x < s_n[l - 1];
		if (have_pre)
			pre = load_join(s_lv[l - 1], threadIdx.x);
		// the stage becomes the dedup tables (tab, bkey, bmin, bhead = kNull; bn = 0), sized for
		// the level: tsz slots, a power of two >= 2 n
		const u32 tlg = min(11u, max(6u, 32u - (u32)__builtin_clz(max(2u * n, 2u) - 1u))); // kSTab = 1 << 11
		const u32 tsz = 1u << tlg;
		for (u32 i = threadIdx.x; i < 5u * tsz; i += blockDim.x) {
			const u32 a = i >> tlg, k = i & (tsz - 1u);
			S.stage[a * kSTab + k] = a < 4u ? kNull : 0u;
		}
		if (threadIdx.x == 0) {
			S.gcount = 0;
			S.stage_top = 0;
		}
		__syncthreads();
		// dedup (identical contents join the slot of their lowest entry; a slot's creator lists the
		// group on its bucket) and the bucket's walker election (its lowest new entry)
		const u32 tmask = tsz - 1u;
		for (u32 e = threadIdx.x; e < n; e += blockDim.x) {
			if (S.slot[cur][e] == kNull)
				continue;
			const u32 h = S.hash[e], b = S.bucket[e];
			u32 m = (b * 0x9E3779B1u >> 7) & tmask;
			for (;;) {
				u32 k = atomicCAS(&S.bkey[m], kNull, b);
				if (k == kNull || k == b)
					break;
				m = (m + 1u) & tmask;
			}
			atomicMin(&S.bmin[m], e);
			S.bmap[e] = m;
			u32 sl = ((h ^ (h >> 15)) * 0x2c1b3c6du) & tmask;
			for (;;) {
				u32 c = atomicCAS(&S.tab[sl], kNull, e);
				if (c == kNull) {
					u32 g = atomicAdd(&S.gcount, 1u);
					S.glist[g] = sl;
					S.gnext[g] = atomicExch(&S.bhead[m], g);
					atomicAdd(&S.bn[m], 1u);
					break;
				}
				if (S.hash[c] == h && lds_same(S, c, e)) {
					atomicMin(&S.tab[sl], e);
					break;
				}
				sl = (sl + 1u) & tmask;
			}
			S.slot[cur][e] = sl;
		}
		__syncthreads();
		for (u32 e = threadIdx.x; e < n; e += blockDim.x)
			if (S.slot[cur][e] != kNull && S.bmin[S.bmap[e]] == e)
				small_walk_bucket(P, leaf, S, e, acc);
		__syncthreads();
		// resolve: every entry's final pointer (its group's result; the old pointer when the bucket
		// was full), written to its parent's content row at its slot for the parent's join
		for (u32 e = threadIdx.x; e < n; e += blockDim.x) {
			const u32 sl = S.slot[cur][e];
			const u32 r = sl == kNull ? S.res[cur][e] : (S.gfail[sl] ? S.ptr[cur][e] : S.gres[sl]);
			S.res[cur][e] = r;
			if (l > 0) {
				const u32 ps = S.pslot[e];
				if ((ps >> 3) < kSmallLevel) {
					S.content[ps >> 3][ps & 7u] = r;
					S.cfresh[ps >> 3][ps & 7u] = sl != kNull && !S.gfail[sl]; // a group's result is its append
				}
			}
		}
		__syncthreads();
	}
	if (threadIdx.x == 0) {
		if (pm[kPmHdr] > 0)
			roots[0] = l_from >= 0 ? lds_result(S, 0, 0) : child_result(lv[0], G, 0);
		u32 tot = 0;
		for (u32 l = 0; l < L; ++l)
			tot += pm[kPmHdr + l];
		acc[kAccEntries] += tot;
		if (P.stateful)
			acc[kAccJobs] += pm[kPmHdr + P.c.color_leaf];
		acc[kAccOverflow] |= pm[kPmOverflow];
	}
}

// SetNode of a colored pass (VBREditorWrapper's joins above the color leaves), one workgroup. It
// needs the top-down entries and the encoded color leaves only, so it runs on the color stream
// beside the node levels.
constexpr u32 kSetNodeThreads = 1024;
__global__ void __launch_bounds__(kSetNodeThreads) a_setnode(EditParams P, const ALevel *__restrict__ lv,
                                                 const u32 *__restrict__ pm, u32 *__restrict__ roots, u64 node_cap,
                                                 u32 lb) {
	__shared__ ColorLds Cs;
	const u32 leafL = P.c.color_leaf;
	if (threadIdx.x == 0) {
		u32 o = 0;
		for (u32 l = 0; l <= lb; ++l) {
			Cs.off[l] = o;
			o += l < lb ? pm[kPmHdr + l] : 0u;
		}
		Cs.fits = o <= kColorLds ? 1u : 0u;
	}
	__syncthreads();
	if (Cs.fits)
		setnode_lds(P, lv, pm, roots, node_cap, Cs, lb);
	else
		setnode_block(P, lv, pm, roots, node_cap, lb);
	if (threadIdx.x == 0) {
		if (leafL == 0 && pm[kPmHdr] > 0)
			roots[1] = lv[0].ores[0];
		P.meta[kMetaColorLeaves] += pm[kPmLeafTotal];
	}
}

__global__ void a_set_roots(u32 *roots, u32 root, u32 croot, u32 *acc) {
	if (threadIdx.x | blockIdx.x)
		return;
	roots[0] = root;
	roots[1] = croot;
	for (u32 i = 0; i < kAccCount; ++i)
		acc[i] = 0;
}
__global__ void a_set_roots_keep(u32 *roots, u32 root, u32 croot) {
	if (threadIdx.x | blockIdx.x)
		return;
	roots[0] = root;
	roots[1] = croot;
}
__global__ void a_fill_u32(u32 *p, u64 n, u32 v) {
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
		p[i] = v;
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

// Whether edits a and b may share a voxel: their voxel AABBs meet and, for a sphere, its ball of
// radius isqrt(r2) + 1 (which holds every voxel p with |p - c|^2 <= r2) meets the other shape's
// ball or box. Conservative: false only for disjoint voxel sets.
static bool may_share_voxels(const hd_edit_desc &a, const i64 *ba, const hd_edit_desc &b, const i64 *bb) {
	if (!(ba[0] <= bb[3] && bb[0] <= ba[3] && ba[1] <= bb[4] && bb[1] <= ba[4] && ba[2] <= bb[5] && bb[2] <= ba[5]))
		return false;
	auto isqrt = [](u64 v) {
		i64 r = (i64)std::sqrt((double)v);
		while ((r + 1) * (r + 1) <= (i64)v)
			++r;
		while (r * r > (i64)v)
			--r;
		return r;
	};
	const bool sa = a.kind <= HD_EDIT_SPHERE_PAINT, sb = b.kind <= HD_EDIT_SPHERE_PAINT;
	const bool xa = a.kind >= HD_EDIT_AABB_FILL && a.kind <= HD_EDIT_AABB_PAINT;
	const bool xb = b.kind >= HD_EDIT_AABB_FILL && b.kind <= HD_EDIT_AABB_PAINT;
	if (sa && sb) {
		i64 d2 = 0;
		for (int i = 0; i < 3; ++i) {
			const i64 d = (i64)a.p0[i] - (i64)b.p0[i];
			d2 += d * d;
		}
		const i64 R = isqrt(a.r2) + isqrt(b.r2) + 2;
		return d2 <= R * R;
	}
	if ((sa && xb) || (sb && xa)) {
		const hd_edit_desc &sp = sa ? a : b;
		const i64 *bx = sa ? bb : ba; // the box's voxel range [lo, hi]
		i64 d2 = 0;
		for (int i = 0; i < 3; ++i) {
			const i64 c = (i64)sp.p0[i], q = std::min(std::max(c, bx[i]), bx[i + 3]);
			d2 += (c - q) * (c - q);
		}
		const i64 R = isqrt(sp.r2) + 1;
		return d2 <= R * R;
	}
	return true; // boxes (their AABBs are exact) and column fills
}

// Level scheduling of a batch: edit k goes to the first group after every group holding an earlier
// edit that may share a voxel with k (may_share_voxels). Edits within a group have disjoint voxel
// sets, so they commute and one pass applies them; groups run in order, so every voxel sees its
// edits in submission order. A coarse 3-D grid of cells keeps the overlap search local.
void schedule_groups(const hd_edit_desc *edits, u32 n, u32 res, std::vector<std::vector<u32>> &groups) {
	const u32 cell_bits = res >= 64 ? std::max(4u, (u32)std::log2((double)res) - 5u) : 2u; // ~32^3 cells
	const u32 cells = std::max(1u, res >> cell_bits);
	std::vector<std::vector<u32>> grid((size_t)cells * cells * cells);
	std::vector<i64> box(6 * (size_t)n);
	std::vector<u32> gid(n, 0);
	groups.clear();
	// A weighted (two R5G6B5 endpoints + weight) color that an edit fills at or above the color
	// leaves is stored RGB8 by the color octree (DAGColorPool::FillNode), while a pass's per-voxel
	// fold keeps it exact. Which of the two happens depends on the node's state when the edit runs
	// (an empty node takes the fill), so such an edit must see exactly the sequential state: it
	// runs in a group of its own, after every earlier edit and before every later one. RGB8 colors
	// are stored exactly either way.
	u32 barrier = 0;
	for (u32 k = 0; k < n; ++k) {
		i64 *b = &box[6 * (size_t)k];
		edit_aabb(edits[k], res, b, b + 3);
		const bool weighted = is_stateful(edits[k]) && edits[k].color.has_value && edits[k].color.bits_per_weight;
		u32 g = barrier;
		bool empty = b[0] > b[3] || b[1] > b[4] || b[2] > b[5];
		i64 c0[3], c1[3];
		if (!empty) {
			for (int i = 0; i < 3; ++i)
				c0[i] = b[i] >> cell_bits, c1[i] = b[i + 3] >> cell_bits;
			for (i64 z = c0[2]; z <= c1[2]; ++z)
				for (i64 y = c0[1]; y <= c1[1]; ++y)
					for (i64 x = c0[0]; x <= c1[0]; ++x)
						for (u32 j : grid[((size_t)z * cells + y) * cells + x]) {
							if (gid[j] + 1 > g && may_share_voxels(edits[j], &box[6 * (size_t)j], edits[k], b))
								g = gid[j] + 1;
						}
		}
		if (weighted) {
			g = std::max<u32>(g, (u32)groups.size());
			barrier = g + 1;
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

// ---------------------------------------------------------------- sync-free pass: host side
constexpr u64 kAsyncBudget = 48ull << 30; // scratch bytes a sync-free pass may use (a sixth of HBM)
// The budget a large pass gets: kAsyncBudget, or less when the device holds less free memory (the
// scratch this pool already holds counts as free: its buffers are reused; buffers grow with a
// quarter of headroom, hence 3/4).
static u64 async_budget(const Pool &p) {
	size_t free_b = 0, total_b = 0;
	if (hipMemGetInfo(&free_b, &total_b) != hipSuccess)
		return kAsyncBudget;
	u64 held = 0;
	for (const auto &kv : p.scratch.bufs)
		held += kv.second.bytes;
	return std::min<u64>(kAsyncBudget, ((u64)free_b + held) / 4 * 3);
}

// Upper bounds on the proceeding nodes of a pass, per level. A node proceeds only where some
// edit's boundary crosses it: for a sphere the closed node cube [lb, lb + S]^3 meets the sphere
// surface, so it lies in the shell r +- S sqrt(3) (cubes are interior-disjoint: count <= shell
// volume / S^3) and within the sphere's box widened by one node on the low side; for a box it
// overlaps the box without lying inside it. Every proceeding node is crossed by the boundary of
// one of its edits, so per-edit sums S bound the entries of every level.
struct PassPlan {
	u32 first = 0, count = 0;
	bool stateful = false, async = false;
	bool folded = false; // the edits may overlap (folded lists)
	std::vector<u64> cap, list_cap;
	u64 ev_cap = 0, bytes = 0;
};

// Color leaves of at most kJobV voxels (and no deeper than the node leaves) are rebuilt in LDS by
// a_job_encode: pushes go to per-pair records instead of the event list.
static bool lds_jobs(const DevConfig &c) {
	return c.color_leaf + 1u <= c.L && 3u * (c.voxel_level - c.color_leaf) <= 12u;
}

static u64 edit_level_bound(const hd_edit_desc &d, u32 voxel_level, u32 l) {
	const u32 s = voxel_level - l;
	const i64 N = (i64)1 << l, S = (i64)1 << s; // nodes per axis, node size
	if (d.kind <= HD_EDIT_SPHERE_PAINT) {
		// sphere_node tests the closed cube [lb, lb + S]: a node proceeds only if that cube meets
		// the sphere surface, so lb <= c + r and lb + S >= c - r (r = floor(sqrt(r2)))
		i64 r = (i64)std::sqrt((double)d.r2);
		while ((r + 1) * (r + 1) <= (i64)d.r2)
			++r;
		while (r * r > (i64)d.r2)
			--r;
		u64 a = 1;
		for (int i = 0; i < 3; ++i) {
			i64 lo = (i64)d.p0[i] - r - S, hi = (i64)d.p0[i] + r; // lb in [lo, hi]
			i64 f = lo <= 0 ? 0 : (lo + S - 1) / S, t = hi < 0 ? -1 : std::min(hi / S, N - 1);
			a *= t >= f ? (u64)(t - f + 1) : 0u;
		}
		double rr = std::sqrt((double)d.r2), tt = (double)S * std::sqrt(3.0);
		double ro = rr + tt, ri = rr > tt ? rr - tt : 0.0;
		double shell = 4.0 / 3.0 * M_PI * (ro * ro * ro - ri * ri * ri) / ((double)S * S * S);
		u64 b = (u64)std::ceil(shell * (1.0 + 1e-9)) + 8u;
		return std::min(a, b);
	}
	if (d.kind <= HD_EDIT_AABB_PAINT) {
		// aabb_node: affected iff ub > p0 and lb < p1, filled iff lb >= p0 and ub <= p1 (per axis);
		// an empty box (p1 <= p0) still proceeds through the nodes that straddle it
		u64 a = 1, in = 1;
		for (int i = 0; i < 3; ++i) {
			i64 p0 = d.p0[i], p1 = d.p1[i];
			i64 f = p0 / S, t = std::min((p1 + S - 1) / S - 1, N - 1); // ub > p0, lb < p1
			a *= t >= f ? (u64)(t - f + 1) : 0u;
			i64 fi = (p0 + S - 1) / S, ti = std::min(p1 / S, N); // [fi, ti) lie inside
			in *= ti > fi ? (u64)(ti - fi) : 0u;
		}
		return a - std::min(a, in);
	}
	return (u64)N * N * N;
}

static void plan_pass(const Pool &p, const hd_edit_desc *edits, PassPlan &pp) {
	const u32 L = p.dc.L, leafL = p.dc.color_leaf, vl = p.dc.voxel_level;
	pp.stateful = false;
	bool column = false;
	for (u32 k = 0; k < pp.count; ++k) {
		pp.stateful |= is_stateful(edits[pp.first + k]);
		column |= edits[pp.first + k].kind == HD_EDIT_COLUMN_FILL;
	}
	pp.async = !column && pp.count > 0;
	if (!pp.async)
		return;
	std::vector<u64> S(L, 0);
	for (u32 k = 0; k < pp.count; ++k)
		for (u32 l = 1; l < L; ++l)
			S[l] += edit_level_bound(edits[pp.first + k], vl, l);
	pp.cap.assign(L, 0);
	pp.list_cap.assign(L, 0);
	pp.cap[0] = 1;
	pp.list_cap[0] = pp.count > 1 ? pp.count : 0;
	const u64 lim = 1ull << 31;
	// A folded pass: a node's list holds the edits from the first that crosses it on (the last one
	// that fills or clears it whole, and the whole-node edits after that, are folded into its
	// state), so besides crossing edits (S) only paints that cover it whole: at most the nodes that
	// lie whole inside the paint's box, per paint.
	std::vector<u64> inner(L, 0);
	if (pp.folded)
		for (u32 k = 0; k < pp.count; ++k) {
			const hd_edit_desc &d = edits[pp.first + k];
			if (d.kind != HD_EDIT_SPHERE_PAINT && d.kind != HD_EDIT_AABB_PAINT)
				continue;
			i64 lo[3], hi[3];
			edit_aabb(d, 1u << vl, lo, hi); // voxels [lo, hi] inclusive
			for (u32 l = 1; l < L; ++l) {
				const i64 Sz = (i64)1 << (vl - l);
				u64 a = 1;
				for (int i = 0; i < 3; ++i) {
					const i64 f = (lo[i] + Sz - 1) / Sz, t = (hi[i] + 1) / Sz; // nodes [f, t) inside
					a *= t > f ? (u64)(t - f) : 0u;
				}
				inner[l] += a;
			}
		}
	for (u32 l = 1; l < L; ++l) {
		pp.cap[l] = std::min(std::min(S[l], 8 * pp.cap[l - 1]), lim);
		// a multi-edit node is crossed by the boundary of all its edits but at most one, which
		// fills it (two fills would overlap): list words <= 2 S
		if (pp.folded)
			pp.list_cap[l] = std::min(S[l] + std::min(inner[l], (u64)pp.count * pp.cap[l]), lim);
		else
			pp.list_cap[l] = pp.count > 1 ? std::min(2 * S[l], lim) : 0;
	}
	pp.ev_cap = 0;
	u64 pev_bytes = 0;
	if (pp.stateful)
		for (u32 l = leafL; l + 1 < L; ++l)
			(lds_jobs(p.dc) ? pev_bytes : pp.ev_cap) += 8 * pp.cap[l];
	pp.ev_cap = std::min<u64>(pp.ev_cap, lim);
	pev_bytes *= sizeof(uint4);
	// bytes: entries + per-level arrays, expand temporaries, lists, events, dense colors
	u64 b = 0, maxpairs = 0;
	for (u32 l = 0; l < L; ++l) {
		u64 c = pp.cap[l];
		b += c * (sizeof(Entry) + 4 * 9 /*content*/ + 4 * 10) + 4 * pp.list_cap[l];
		if (l + 1 < L) {
			b += c * 8 * 4 * (pp.stateful && l < leafL ? 2 : 1) + c * 8;
			maxpairs = std::max(maxpairs, 8 * c);
		}
	}
	b += maxpairs * 24 + pp.ev_cap * sizeof(PushEvent) + pev_bytes;
	if (pp.stateful) // per job: dense colors and a staging slot, or (LDS jobs) bump-allocated staging
		b += (lds_jobs(p.dc) ? pp.cap[leafL] * (60 + 4 * (u64)kEtmpJobWords)
		                     : pp.cap[leafL] * ((u64)sizeof(Color) << (3 * (vl - leafL))) +
		                           pp.cap[leafL] * (60 + 4 * (u64)kEncTmp)) +
		     pp.cap[L - 1] * (lds_jobs(p.dc) ? 64 * sizeof(unsigned short) : 64 * sizeof(Color) + 8);
	pp.bytes = b;
	if (b > kAsyncBudget || (b > (1ull << 30) && b > async_budget(p)) || (pp.stateful && vl - leafL > 7))
		pp.async = false;
}

// Staging words per LDS job; HD_EDIT_STAGING_WORDS overrides it (tests: less staging sends more
// jobs through a_job_rest, with the same result).
static u64 staging_words() {
	const char *e = getenv("HD_EDIT_STAGING_WORDS");
	return e && *e ? strtoull(e, nullptr, 10) : kEtmpJobWords;
}

struct AsyncBufs {
	std::vector<ALevel> lv;
	ALevel *d_lv = nullptr;
	u32 *pm = nullptr, *roots = nullptr, *acc = nullptr, *bcnt = nullptr, *head = nullptr, *blist = nullptr;
	u32 *flag = nullptr, *lcnt = nullptr, *info = nullptr, *app = nullptr, *appo = nullptr, *etmp = nullptr, *rest = nullptr;
	u32 *tsum = nullptr; // per-tile totals of the jobs' append scan
	uint4 *cstate = nullptr, *jrange = nullptr;
	u32 etmp_cap = 0; // staging words of LDS jobs (bump-allocated)
	Pair *bsum = nullptr;
	PushEvent *ev = nullptr;
	Color *dense = nullptr, *lcol = nullptr;
	u64 *lmask = nullptr;
	unsigned short *lidx = nullptr; // LDS jobs: per leaf voxel, the pass edit that assigned its color
	uint2 *ecol = nullptr;          // LDS jobs: per pass edit, the color lidx stands for
	hd_edit_desc *edits = nullptr;
	GroupTable G{};
	u64 tab_slots = 0;
	u32 *gen_ctr = nullptr;    // device: generations handed out so far (a_top takes L per pass)
	u32 *gen_mirror = nullptr; // host mirror of *gen_ctr
};
constexpr u32 kMaxEmitBlocks = 1024;

// Buffers for the largest bounds of the call's sync-free passes (allocated before any pass is
// queued: a growing buffer would be freed under queued work).
static int async_setup(Pool &p, const std::vector<PassPlan> &plans, const std::vector<hd_edit_desc> &flat,
                       AsyncBufs &B, const std::string &pre = "a_", u32 *gen_mirror = nullptr) {
	B.gen_mirror = gen_mirror ? gen_mirror : &p.edit_gen;
	const u32 L = p.dc.L, leafL = p.dc.color_leaf;
	std::vector<u64> cap(L, 0), lcap(L, 0);
	u64 ev = 0, maxpairs = 8;
	bool stateful = false;
	for (auto &pp : plans) {
		if (!pp.async)
			continue;
		for (u32 l = 0; l < L; ++l) {
			cap[l] = std::max(cap[l], pp.cap[l]);
			lcap[l] = std::max(lcap[l], pp.list_cap[l]);
			if (l + 1 < L)
				maxpairs = std::max(maxpairs, 8 * pp.cap[l]);
		}
		ev = std::max(ev, pp.ev_cap);
		stateful |= pp.stateful;
	}
	B.lv.assign(L, ALevel{});
	for (u32 l = 0; l < L; ++l) {
		ALevel &A = B.lv[l];
		const u64 c = std::max<u64>(cap[l], 1);
		A.cap = (u32)cap[l];
		A.list_cap = (u32)lcap[l];
		HD_RET(buf(p, (pre + "ent").c_str(), l, c, &A.ent));
		HD_RET(buf(p, (pre + "lists").c_str(), l, std::max<u64>(lcap[l], 1), &A.lists));
		if (l + 1 < L) {
			HD_RET(buf(p, (pre + "kids").c_str(), l, 8 * c, &A.kids));
			HD_RET(buf(p, (pre + "fchild").c_str(), l, c, &A.first_child));
			HD_RET(buf(p, (pre + "pmask").c_str(), l, c, &A.pmask));
			if (stateful && l < leafL)
				HD_RET(buf(p, (pre + "okids").c_str(), l, 8 * c, &A.okids));
		}
		HD_RET(buf(p, (pre + "status").c_str(), l, c, &A.status));
		HD_RET(buf(p, (pre + "pend").c_str(), l, c, &A.pend));
		HD_RET(buf(p, (pre + "res").c_str(), l, c, &A.res));
		HD_RET(buf(p, (pre + "slot").c_str(), l, c, &A.slot));
		HD_RET(buf(p, (pre + "lpos").c_str(), l, c, &A.lpos));
		HD_RET(buf(p, (pre + "bucket").c_str(), l, c, &A.bucket));
		HD_RET(buf(p, (pre + "content").c_str(), l, c * (l + 1 < L ? 9 : 2), &A.content));
		if (l + 1 < L)
			HD_RET(buf(p, (pre + "csize").c_str(), l, c, &A.csize));
		if (stateful && lds_jobs(p.dc) && l >= leafL && l + 1 < L)
			HD_RET(buf(p, (pre + "pev").c_str(), l, 8 * c, &A.pev));
		if (stateful && l <= leafL) {
			HD_RET(buf(p, (pre + "ores").c_str(), l, c, &A.ores));
			HD_RET(buf(p, (pre + "ccnt").c_str(), l, c, &A.ccnt));
			HD_RET(buf(p, (pre + "cbase").c_str(), l, c, &A.cbase));
			HD_RET(buf(p, (pre + "own").c_str(), l, c, &A.own));
		}
	}
	HD_RET(p.scratch.get<ALevel>(pre + "levels", L, &B.d_lv));
	HD_TRY(hipMemcpyAsync(B.d_lv, B.lv.data(), L * sizeof(ALevel), hipMemcpyHostToDevice, p.stream));
	HD_RET(p.scratch.get<u32>(pre + "pm", kPmHdr + 2 * L, &B.pm));
	HD_RET(p.scratch.get<u32>(pre + "roots", 2, &B.roots));
	HD_RET(p.scratch.get<u32>(pre + "acc", kAccCount, &B.acc));
	HD_RET(p.scratch.get<u32>(pre + "flag", maxpairs, &B.flag));
	HD_RET(p.scratch.get<u32>(pre + "lcnt", maxpairs, &B.lcnt));
	HD_RET(p.scratch.get<uint4>(pre + "cstate", maxpairs, &B.cstate));
	HD_RET(p.scratch.get<Pair>(pre + "bsum", kMaxEmitBlocks, &B.bsum));
	HD_RET(p.scratch.get<PushEvent>(pre + "events", std::max<u64>(ev, 1), &B.ev));
	if (stateful) {
		const u64 jobs = std::max<u64>(cap[leafL], 1);
		const bool lj = lds_jobs(p.dc); // LDS jobs rebuild what they do not stage: no dense colors
		HD_RET(p.scratch.get<Color>(pre + "dense", lj ? 1 : jobs << (3 * (p.dc.voxel_level - leafL)), &B.dense));
		HD_RET(p.scratch.get<u32>(pre + "info", 4 * jobs, &B.info));
		HD_RET(p.scratch.get<u32>(pre + "app", jobs, &B.app));
		HD_RET(p.scratch.get<u32>(pre + "appo", jobs, &B.appo));
		HD_RET(p.scratch.get<u32>(pre + "tsum", (jobs + kScanTile - 1) / kScanTile, &B.tsum));
		B.etmp_cap = (u32)std::min<u64>(jobs * staging_words(), 1ull << 31);
		HD_RET(p.scratch.get<u32>(pre + "etmp", lj ? std::max<u64>(B.etmp_cap, 1) : jobs * kEncTmp, &B.etmp));
		HD_RET(p.scratch.get<u32>(pre + "rest", jobs, &B.rest));
		HD_RET(p.scratch.get<uint4>(pre + "jrange", 2 * jobs, &B.jrange));
		const u64 leaves = std::max<u64>(cap[L - 1], 1);
		HD_RET(p.scratch.get<Color>(pre + "lcol", lj ? 1 : leaves * 64, &B.lcol));
		HD_RET(p.scratch.get<u64>(pre + "lmask", lj ? 1 : leaves, &B.lmask));
		HD_RET(p.scratch.get<unsigned short>(pre + "lidx", lj ? leaves * 64 : 1, &B.lidx));
		u32 max_count = 1;
		for (auto &pp : plans)
			max_count = std::max(max_count, pp.count);
		HD_RET(p.scratch.get<uint2>(pre + "ecol", max_count, &B.ecol));
	}
	HD_RET(p.scratch.get<hd_edit_desc>(pre + "edits", std::max<size_t>(flat.size(), 1), &B.edits));
	HD_TRY(hipMemcpyAsync(B.edits, flat.data(), flat.size() * sizeof(hd_edit_desc), hipMemcpyHostToDevice, p.stream));
	// group table: 2x the largest level; generations make clearing unnecessary except on growth
	u64 maxcap = 1;
	for (u32 l = 0; l < L; ++l)
		maxcap = std::max(maxcap, cap[l]);
	u64 T = 1024;
	while (T < 2 * maxcap)
		T <<= 1;
	{
		DevBuf &tb = p.scratch.bufs[pre + "tab"];
		void *before = tb.p;
		HD_RET(tb.ensure(T * 8));
		B.G.tab = tb.as<u64>();
		u64 slots = tb.bytes / 8; // ensure() may round up: use every slot it gave
		while (T * 2 <= slots)
			T <<= 1;
		B.tab_slots = slots;
		HD_RET(p.scratch.get<u32>(pre + "gen", 1, &B.gen_ctr));
		if (tb.p != before) { // a new table: empty, generations restart
			a_fill_u32<<<std::min<u64>(div_up(2 * slots, 256), 4096), 256, 0, p.stream>>>((u32 *)tb.p, 2 * slots, 0u);
			a_fill_u32<<<1, 64, 0, p.stream>>>(B.gen_ctr, 1, 0u);
			HD_TRY(hipGetLastError());
			*B.gen_mirror = 0;
		}
	}
	B.G.mask = (u32)(T - 1);
	HD_RET(p.scratch.get<u32>(pre + "gres", T, &B.G.gres));
	HD_RET(p.scratch.get<u32>(pre + "gfail", T, &B.G.gfail));
	HD_RET(p.scratch.get<u32>(pre + "gnext", T, &B.G.gnext));
	// per-bucket group lists: empty between levels (every walker resets its bucket)
	const char *kb = "a_bcnt", *kh = "a_head";
	bool fresh = p.scratch.bufs.find(kb) == p.scratch.bufs.end();
	HD_RET(p.scratch.get<u32>(kb, p.total_buckets, &B.bcnt));
	HD_RET(p.scratch.get<u32>(kh, p.total_buckets, &B.head));
	HD_RET(p.scratch.get<u32>("a_blist", (u64)p.total_buckets * kWalkList, &B.blist));
	if (fresh) {
		u32 g = std::min<u64>(div_up(p.total_buckets, 256), 4096);
		a_fill_u32<<<g, 256, 0, p.stream>>>(B.bcnt, p.total_buckets, 0u);
		a_fill_u32<<<g, 256, 0, p.stream>>>(B.head, p.total_buckets, kNull);
		HD_TRY(hipGetLastError());
	}
	HD_TRY(hipStreamSynchronize(p.stream)); // the host staging arrays may go out of scope
	return HD_OK;
}

// The device hands out L generations per pass (a_top); the host mirrors the count so that before
// the 32-bit counter wraps the table is emptied and the counter restarted (queued, not captured).
static int gen_reserve(Pool &p, AsyncBufs &B, u32 passes) {
	const u64 need = (u64)passes * p.dc.L;
	if ((u64)*B.gen_mirror + need > 0xFFFFFFF0ull) {
		a_fill_u32<<<4096, 256, 0, p.stream>>>((u32 *)B.G.tab, 2 * B.tab_slots, 0u);
		a_fill_u32<<<1, 64, 0, p.stream>>>(B.gen_ctr, 1, 0u);
		HD_TRY(hipGetLastError());
		*B.gen_mirror = 0;
	}
	*B.gen_mirror += (u32)need;
	return HD_OK;
}

static EditParams pass_params(const Pool &p, const hd_edit_desc *d_edits, u32 n, bool stateful, bool folded = false) {
	EditParams P;
	P.folded = folded && n > 1 ? 1u : 0u;
	P.c = p.dc;
	P.hm = p.d_hm;
	P.edits = d_edits;
	P.n_edits = n;
	P.words = p.words;
	P.bucket_words = p.bucket_words;
	P.filled = p.d_filled;
	P.cnodes = p.cnodes;
	P.cleaves = p.cleaves;
	P.meta = p.meta;
	P.stateful = stateful ? 1u : 0u;
	P.octree_root = 0;
	P.leaf_bits = p.dc.voxel_level - p.dc.color_leaf;
	return P;
}

static u32 grid_for(u64 items, u32 per_block, u32 max_blocks) {
	return (u32)std::max<u64>(1, std::min<u64>(div_up(std::max<u64>(items, 1), per_block), max_blocks));
}

// SetNode: the levels above the color leaves whose bound exceeds kSetNodeWide entries run as grid
// launches (their decisions before the single workgroup, their placement after it); the levels
// above them stay in the workgroup.
constexpr u64 kSetNodeWide = 4096;
static void queue_setnode(const Pool &p, const PassPlan &pp, const EditParams &P, const AsyncBufs &B, hipStream_t s) {
	const u32 leafL = p.dc.color_leaf;
	u32 lb = leafL;
	while (lb > 1 && pp.cap[lb - 1] > kSetNodeWide)
		--lb;
	for (u32 l = leafL; l-- > lb;)
		a_setnode_up<<<grid_for(pp.cap[l], 256, 4096), 256, 0, s>>>(P, B.d_lv, B.pm, l);
	a_setnode<<<1, kSetNodeThreads, 0, s>>>(P, B.d_lv, B.pm, B.roots, p.cnode_cap, lb);
	for (u32 l = lb; l < leafL; ++l)
		a_setnode_down<<<grid_for(pp.cap[l], 256, 4096), 256, 0, s>>>(P, B.d_lv, B.pm, l, p.cnode_cap);
}

// Queues one sync-free pass (no host synchronization).
static int queue_pass(Pool &p, const PassPlan &pp, AsyncBufs &B, bool reserve = true) {
	const u32 L = p.dc.L, leafL = p.dc.color_leaf;
	if (pp.stateful && lds_jobs(p.dc) && pp.count >= kNoEdit) // leaves record edits by 16-bit index
		return set_error("hd_edit: a colored pass holds at most %u edits", kNoEdit - 1u), HD_ERR_ARG;
	const EditParams P = pass_params(p, B.edits + pp.first, pp.count, pp.stateful, pp.folded);
	const u32 V = pp.stateful ? 1u << (3u * P.leaf_bits) : 0u;
	// A pair of a multi-edit pass tests its parent's edit list: count those tests too (lists hold at
	// most the pass's edits, and shrink with depth), so list-heavy levels go to the wide kernels
	// (a pass of thousands of edits leaves only the root to the one workgroup: the wide kernels take
	// the level below's nodes in parallel, 4096 mixed edits 0.8 ms faster)
	const u64 lf = (std::min<u32>(pp.count, 4096) + 7) / 8;
	u32 top_end = 0;
	const u64 top_pairs = std::min<u64>(g_tuning.top_pairs, kTopPairs);
	while (top_end + 1 < L && 8 * pp.cap[top_end] * lf <= top_pairs)
		++top_end;
	if (reserve)
		HD_RET(gen_reserve(p, B, 1));
	const auto top = pp.count > 1 ? a_top<true> : a_top<false>;
	top<<<1, kAT, 0, p.stream>>>(P, B.d_lv, B.roots, B.pm, kPmHdr + 2 * L, top_end, B.ev, (u32)pp.ev_cap, B.gen_ctr);
	// one 256-pair tile per block up to the block cap: a block's tiles are a chain of dependent round
	// trips (entry, node words, color child), so a mid-size level spread over few blocks waits on
	// them in turn (256 sphere edits on the 4096^3 terrain: 44.1 -> 41.0 ms per step vs 2048 pairs)
	for (u32 l = top_end; l + 1 < L; ++l) {
		u32 G = grid_for(8 * pp.cap[l] * lf, kAT, kMaxEmitBlocks);
		a_count<<<G, kAT, 0, p.stream>>>(P, l, B.d_lv, B.pm, B.flag, B.lcnt, B.cstate, B.ev, (u32)pp.ev_cap, B.bsum);
		a_emit<<<G, kAT, 0, p.stream>>>(P, l, B.d_lv, B.pm, B.flag, B.lcnt, B.cstate, B.bsum);
	}
	const bool colors = pp.stateful && pp.cap[leafL] > 0;
	// A colored pass that proceeds only above the color leaves (its fills and clears land at or above
	// them) has no color leaf job, yet the octree above the leaves may change (a fill color at a
	// proceeding node, cleared children): SetNode alone, once the top-down has classified them.
	if (pp.stateful && !colors && leafL > 0)
		queue_setnode(p, pp, P, B, p.stream);
	const bool ljobs = lds_jobs(p.dc);
	if (colors && !ljobs) { // the jobs' dense colors on the color stream, beside the leaves
		HD_TRY(hipEventRecord(p.ev_top, p.stream));
		HD_TRY(hipStreamWaitEvent(p.cstream, p.ev_top, 0));
		a_dense_init<<<grid_for(pp.cap[leafL], 1, 8192), kDenseThreads, 0, p.cstream>>>(B.d_lv, leafL, B.pm,
		                                                                             3u * P.leaf_bits, p.cleaves, B.dense);
		if (pp.ev_cap)
			a_push<<<grid_for(pp.ev_cap, 1, 8192), 256, 0, p.cstream>>>(B.ev, B.pm, (u32)pp.ev_cap, V,
			                                                           B.lv[leafL].cap, B.dense);
	}
	const u32 leaf_grid = grid_for(pp.cap[L - 1], kAT / 64, 16384);
	// with LDS jobs, the leaves' colors start the color stream right after the top-down (always: such
	// a pass has the edit-index buffer lidx, and no lcol / lmask for kLeafAll)
	const bool split = colors && ljobs;
	if (split) {
		HD_TRY(hipEventRecord(p.ev_top, p.stream));
		HD_TRY(hipStreamWaitEvent(p.cstream, p.ev_top, 0));
		a_leaf<kLeafColors><<<leaf_grid, kAT, 0, p.cstream>>>(P, B.d_lv, B.pm, B.lcol, B.lmask, B.lidx);
		a_leaf<kLeafNodes><<<leaf_grid, kAT, 0, p.stream>>>(P, B.d_lv, B.pm, B.lcol, B.lmask, B.lidx);
	} else if (colors)
		a_leaf<kLeafAll><<<leaf_grid, kAT, 0, p.stream>>>(P, B.d_lv, B.pm, B.lcol, B.lmask, B.lidx);
	else
		a_leaf<kLeafNodes><<<leaf_grid, kAT, 0, p.stream>>>(P, B.d_lv, B.pm, B.lcol, B.lmask, B.lidx);
	int l_from = -1; // levels [0, l_from] are small enough for a_bottom
	const u64 small_level = std::min<u32>(g_tuning.small_level, kSmallLevel);
	while (l_from + 1 < (int)L && pp.cap[l_from + 1] <= small_level)
		++l_from;
	// the color leaves (job rebuild and encode) and SetNode on the color stream, beside the node
	// levels; they start once the node levels have passed level color_after (L: right after the
	// leaves)
	bool colors_queued = !colors;
	auto queue_colors = [&]() -> int {
		colors_queued = true;
		if (!split) {
			HD_TRY(hipEventRecord(p.ev_leaf, p.stream));
			HD_TRY(hipStreamWaitEvent(p.cstream, p.ev_leaf, 0));
		}
		u32 G = grid_for(pp.cap[leafL], 1, 8192);
		if (ljobs) {
			a_job_ranges<<<grid_for(std::max<u64>(pp.cap[leafL], pp.count), 256, 4096), 256, 0, p.cstream>>>(
			    B.d_lv, B.pm, leafL, L, p.cleaves, B.jrange, P.edits, pp.count, B.ecol);
			const auto job = V == kJobV ? a_job_encode<true> : a_job_encode<false>;
			job<<<G, kEncThreads, 0, p.cstream>>>(P, B.d_lv, B.pm, B.jrange, B.lidx, B.ecol, V, B.info, B.app, B.etmp,
			                                      B.etmp_cap, B.rest);
		}
		else {
			a_leafcolor<<<leaf_grid, kAT, 0, p.cstream>>>(P, B.d_lv, B.pm, B.lcol, B.lmask, B.dense, V);
			a_encode_size<<<G, kEncThreads, 0, p.cstream>>>(P, B.d_lv, B.pm, B.dense, V, B.info, B.app, B.etmp,
			                                                B.rest);
		}
		a_scan_app<<<grid_for(pp.cap[leafL], kScanTile, 1024), kScanThreads, 0, p.cstream>>>(leafL, B.pm, B.app,
		                                                                                    B.appo, B.tsum);
		a_scan_fix<<<grid_for(pp.cap[leafL], 256 * 8, 1024), 256, 0, p.cstream>>>(leafL, B.pm, B.appo, B.tsum);
		a_encode_write<<<G, kEncThreads, 0, p.cstream>>>(P, B.d_lv, B.pm, B.dense, V, B.info, B.appo, p.cleaf_cap,
		                                                 B.etmp);
		if (ljobs) {
			const auto rest = V == kJobV ? a_job_rest<true> : a_job_rest<false>;
			rest<<<std::min(G, 1024u), kEncThreads, 0, p.cstream>>>(P, B.d_lv, B.pm, B.jrange, B.lidx, B.ecol, V, B.info,
			                                                        B.appo, p.cleaf_cap, B.rest);
		} else
			a_encode_rest<<<std::min(G, 1024u), kEncThreads, 0, p.cstream>>>(P, B.d_lv, B.pm, B.dense, V, B.info,
			                                                                 B.appo, p.cleaf_cap, B.rest);
		queue_setnode(p, pp, P, B, p.cstream);
		HD_TRY(hipEventRecord(p.ev_color, p.cstream));
		return HD_OK;
	};
	// without the split, the color stream starts after the leaf level's walk (measured best)
	const int color_after = split ? (int)L : (int)L - 1;
	if (color_after >= (int)L)
		HD_RET(queue_colors());
	for (int l = (int)L - 1; l > l_from; --l) {
		u32 G = grid_for(pp.cap[l], kAT, 4096);
		if (l + 1 < (int)L)
			// a single-edit pass has few color jobs beside it: twice the stage (256 sphere edits on
			// the 4096^3 terrain: 41.1 -> 40.3 ms per step; the 4096-edit batch 1 % slower with it)
			(pp.count == 1 ? a_build<20480> : a_build<10240>)<<<G, kAT, 0, p.stream>>>(P, (u32)l, B.d_lv, B.pm, B.G);
		a_dedup<<<G, kAT, 0, p.stream>>>(P, (u32)l, B.d_lv, B.pm, B.G, B.bcnt, B.head, B.blist);
		a_walk<<<G, kAT, 0, p.stream>>>(P, (u32)l, B.d_lv, B.pm, B.G, B.bcnt, B.head, B.blist, B.acc);
		if (!colors_queued && l <= color_after)
			HD_RET(queue_colors());
	}
	if (!colors_queued)
		HD_RET(queue_colors());
	a_bottom<<<1, kAT, 0, p.stream>>>(P, B.d_lv, B.pm, l_from, B.G, B.roots, B.acc, p.cnode_cap);
	if (colors) // the next pass (and any host read) follows both streams
		HD_TRY(hipStreamWaitEvent(p.stream, p.ev_color, 0));
	HD_TRY(hipGetLastError());
	return HD_OK;
}

static bool force_sync_path() {
	const char *e = getenv("HD_EDIT_PATH");
	return e && strcmp(e, "sync") == 0;
}

// Applies the passes in order: sync-free passes queue back to back with the roots in device
// memory; a pass the bounds rule out (column fills, scratch over budget) runs the synchronized
// pipeline in between.
static int edit_sequence(Pool &p, u32 root, u32 croot, const std::vector<hd_edit_desc> &flat,
                         std::vector<PassPlan> &plans, u32 *root_out, u32 *croot_out) {
	HD_RET(make_filled(p));
	const bool sync_only = force_sync_path();
	bool any_async = false;
	for (auto &pp : plans) {
		plan_pass(p, flat.data(), pp);
		if (sync_only)
			pp.async = false;
		any_async |= pp.async;
	}
	if (!any_async) {
		for (auto &pp : plans)
			HD_RET(edit_pass(p, root, croot, flat.data() + pp.first, pp.count, &root, &croot));
		*root_out = root;
		*croot_out = croot;
		return HD_OK;
	}
	AsyncBufs B;
	HD_RET(async_setup(p, plans, flat, B));
	a_set_roots<<<1, 64, 0, p.stream>>>(B.roots, root, croot, B.acc);
	bool on_device = true;
	for (auto &pp : plans) {
		if (pp.async) {
			if (!on_device)
				a_set_roots_keep<<<1, 64, 0, p.stream>>>(B.roots, root, croot);
			on_device = true;
			HD_RET(queue_pass(p, pp, B));
		} else {
			if (on_device) {
				u32 rr[2];
				HD_RET(read_u32s(p, B.roots, rr, 2));
				root = rr[0], croot = rr[1];
			}
			on_device = false;
			HD_RET(edit_pass(p, root, croot, flat.data() + pp.first, pp.count, &root, &croot));
		}
	}
	u32 out[2 + kAccCount];
	HD_TRY(hipMemcpyAsync(out, B.roots, 8, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipMemcpyAsync(out + 2, B.acc, 4 * kAccCount, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	if (on_device)
		root = out[0], croot = out[1];
	p.last_entries += out[2 + kAccEntries];
	p.last_upserts += out[2 + kAccUpserts];
	p.last_jobs += out[2 + kAccJobs];
	if (out[2 + kAccOverflow])
		return set_error("hd_edit: a level outgrew its bound (internal error)"), HD_ERR_CAPACITY;
	*root_out = root;
	*croot_out = croot;
	return HD_OK;
}

// The passes of a batched call: dependency groups (schedule_groups), split at batch_pass_edits.
static void batched_plans(const Pool &p, const hd_edit_desc *edits, u32 n, std::vector<hd_edit_desc> &flat,
                          std::vector<PassPlan> &plans) {
	std::vector<std::vector<u32>> groups;
	schedule_groups(edits, n, 1u << p.dc.voxel_level, groups);
	flat.reserve(n);
	// a group's edits are pairwise disjoint, so any split of it into consecutive passes applies them
	// in an order they commute under
	const u32 cap = g_tuning.batch_pass_edits ? g_tuning.batch_pass_edits : ~0u;
	for (auto &g : groups)
		for (size_t b = 0; b < g.size(); b += cap) {
			PassPlan pp;
			pp.first = (u32)flat.size();
			pp.count = (u32)std::min<size_t>(cap, g.size() - b);
			for (size_t j = b; j < b + pp.count; ++j)
				flat.push_back(edits[g[j]]);
			plans.push_back(pp);
		}
}
static void plain_plans(const hd_edit_desc *edits, u32 n, std::vector<hd_edit_desc> &flat,
                        std::vector<PassPlan> &plans) {
	flat.assign(edits, edits + n);
	plans.assign(n, PassPlan{});
	for (u32 i = 0; i < n; ++i)
		plans[i].first = i, plans[i].count = 1;
}

// Folded passes: runs of consecutive edits, each run one pass whose node edit lists keep submission
// order (list_chain folds a node's list; leaves apply it voxel by voxel), so every voxel sees its
// edits in order and overlapping edits share a pass. The occupied voxels' colors are the last
// color written to each, as in sequential application. An edit whose colors would depend on the
// path the sequential edits take runs in a pass of its own, between the runs: a weighted color (a
// whole-node fill stores it RGB8, a voxel keeps it exact), a column fill, and an uncolored fill
// among colored edits (it refills voxels with the colors earlier edits left on them). Runs of up
// to 4096 edits: a larger pass shares its fixed costs (the top-down's first levels, the small
// levels, SetNode) among more edits and re-encodes a color leaf that several of its edits touch
// once (4096 mixed edits on the 4096^3 terrain: runs of 512 43 ms, 1024 39 ms, 2048 34 ms, all
// 4096 in one pass 28 ms).
constexpr u32 kFoldedPassEdits = 4096;
static bool folds(const hd_edit_desc &d, bool colored_batch) {
	if (d.kind == HD_EDIT_COLUMN_FILL)
		return false;
	if (is_stateful(d) && d.color.has_value && d.color.bits_per_weight)
		return false;
	const bool dig = d.kind == HD_EDIT_SPHERE_DIG || d.kind == HD_EDIT_AABB_DIG;
	return !(colored_batch && !is_stateful(d) && !dig);
}
static void folded_plans(const hd_edit_desc *edits, u32 n, std::vector<hd_edit_desc> &flat,
                         std::vector<PassPlan> &plans) {
	flat.assign(edits, edits + n);
	bool colored = false;
	for (u32 k = 0; k < n; ++k)
		colored |= is_stateful(edits[k]);
	// (a pass's leaves record its edits by 16-bit index)
	const u32 cap = std::min<u32>(g_tuning.batch_pass_edits ? g_tuning.batch_pass_edits : kFoldedPassEdits, kNoEdit - 1u);
	PassPlan cur;
	auto flush = [&]() {
		if (cur.count) {
			cur.folded = cur.count > 1;
			plans.push_back(cur);
		}
		cur = PassPlan{};
	};
	for (u32 k = 0; k < n; ++k) {
		if (!folds(edits[k], colored)) {
			flush();
			PassPlan solo;
			solo.first = k, solo.count = 1;
			plans.push_back(solo);
			continue;
		}
		if (cur.count == cap)
			flush();
		if (!cur.count)
			cur.first = k;
		++cur.count;
	}
	flush();
}

// The passes of a batched call: folded runs, each split in halves until its buffers fit the
// sync-free budget; dependency groups when a colored pass would rebuild its color leaves from
// dense per-voxel colors (nested pushes need the leaves rebuilt in LDS).
static void batch_plans(const Pool &p, const hd_edit_desc *edits, u32 n, std::vector<hd_edit_desc> &flat,
                        std::vector<PassPlan> &plans) {
	if (!lds_jobs(p.dc) || force_sync_path()) {
		batched_plans(p, edits, n, flat, plans);
		return;
	}
	std::vector<PassPlan> runs;
	folded_plans(edits, n, flat, runs);
	std::vector<PassPlan> todo(runs.rbegin(), runs.rend()); // a stack, first run on top
	while (!todo.empty()) {
		PassPlan pp = todo.back();
		todo.pop_back();
		plan_pass(p, flat.data(), pp);
		if (pp.folded && !pp.async) { // over the budget: its halves, in order
			PassPlan a, b;
			a.first = pp.first, a.count = pp.count / 2, a.folded = a.count > 1;
			b.first = pp.first + a.count, b.count = pp.count - a.count, b.folded = b.count > 1;
			todo.push_back(b);
			todo.push_back(a);
			continue;
		}
		plans.push_back(pp);
	}
}

int edit_batched(Pool &p, u32 root, u32 croot, const hd_edit_desc *edits, u32 n, u32 *root_out, u32 *croot_out) {
	std::vector<hd_edit_desc> flat;
	std::vector<PassPlan> plans;
	batch_plans(p, edits, n, flat, plans);
	p.last_groups = (u32)plans.size();
	return edit_sequence(p, root, croot, flat, plans, root_out, croot_out);
}

int edit_plain(Pool &p, u32 root, u32 croot, const hd_edit_desc *edits, u32 n, u32 *root_out, u32 *croot_out) {
	std::vector<hd_edit_desc> flat;
	std::vector<PassPlan> plans;
	plain_plans(edits, n, flat, plans);
	p.last_groups = n;
	return edit_sequence(p, root, croot, flat, plans, root_out, croot_out);
}

// hd_edit_reserve: the buffers the call's sync-free passes would allocate, allocated now.
static int edit_reserve(Pool &p, const hd_edit_desc *edits, u32 n, u32 mode) {
	std::vector<hd_edit_desc> flat;
	std::vector<PassPlan> plans;
	if (mode == HD_MODE_BATCHED && n > 1)
		batch_plans(p, edits, n, flat, plans);
	else
		plain_plans(edits, n, flat, plans);
	bool any_async = false;
	for (auto &pp : plans) {
		plan_pass(p, flat.data(), pp);
		any_async |= pp.async;
	}
	if (!any_async || force_sync_path())
		return HD_OK;
	AsyncBufs B;
	HD_RET(async_setup(p, plans, flat, B));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

// ---------------------------------------------------------------- frames captured as hipGraphs
// A frame = up to max_passes sync-free passes (groups of pairwise-disjoint edits, max_edits each,
// every edit within the radius budget) followed by one multi-view trace of the new roots. The whole
// sequence is captured once; each frame then writes its edits (padded with an empty box, which no
// node classification can touch), views and tile prefix into the frame's device buffers and
// replays the graph. Roots stay on the device between frames.
struct FrameState {
	Pool *p = nullptr;
	hd_frame_config cfg{};
	AsyncBufs B;
	std::vector<PassPlan> plans;
	std::vector<u64> budget; // per-level node bound of one edit
	std::string pre;
	u32 gen_mirror = 0;
	// staging for the asynchronous uploads: two slots, each reused only after its copies completed
	hd_edit_desc *h_edits[2] = {nullptr, nullptr};
	hd_view *h_views[2] = {nullptr, nullptr};
	u32 *h_tiles[2] = {nullptr, nullptr};
	hipEvent_t copied[2] = {nullptr, nullptr};
	u32 slot = 0;
	u32 *h_roots = nullptr;
	hd_view *d_views = nullptr;
	u32 *d_tiles = nullptr, *d_rgba = nullptr;
	u32 max_tiles = 0;
	u32 tile = 16; // the frame kernel's pixel tile when the graph was captured
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	~FrameState() {
		if (exec)
			(void)hipGraphExecDestroy(exec);
		if (graph)
			(void)hipGraphDestroy(graph);
		for (int k = 0; k < 2; ++k) {
			for (void *h : {(void *)h_edits[k], (void *)h_views[k], (void *)h_tiles[k]})
				if (h)
					(void)hipHostFree(h);
			if (copied[k])
				(void)hipEventDestroy(copied[k]);
		}
		if (h_roots)
			(void)hipHostFree(h_roots);
		if (p) { // the frame's own buffers (its graph holds their addresses)
			for (auto it = p->scratch.bufs.begin(); it != p->scratch.bufs.end();)
				if (it->first.compare(0, pre.size(), pre) == 0) {
					it->second.release();
					it = p->scratch.bufs.erase(it);
				} else
					++it;
		}
	}
};

// Largest edit_level_bound of a sphere of radius <= R or a box of half-extent <= R, over every
// position: a span of n voxels meets at most ceil((n - 1) / S) + 1 nodes of size S and holds at
// least floor(n / S) - 1 of them whole; the sphere's shell bound does not depend on the position.
static u64 budget_level_bound(u32 R, u32 voxel_level, u32 l) {
	const u32 s = voxel_level - l;
	const u64 S = 1ull << s, N = 1ull << l;
	auto meet = [&](u64 n) { return std::min<u64>(N, (n + S - 2) / S + 1); };
	auto whole = [&](u64 n) { return n / S >= 1 ? std::min<u64>(N, n / S - 1) : 0ull; };
	const u64 nb = 2ull * R + 1; // box: [c - R, c + R]
	u64 a = meet(nb), in = whole(nb);
	const u64 box = a * a * a - in * in * in;
	const u64 ns = 2ull * R + S + 1; // sphere: lb in [c - R - S, c + R]
	const u64 as = meet(ns);
	double rr = (double)R, tt = (double)S * std::sqrt(3.0);
	double ro = rr + tt, ri = rr > tt ? rr - tt : 0.0;
	double shell = 4.0 / 3.0 * M_PI * (ro * ro * ro - ri * ri * ri) / ((double)S * S * S);
	const u64 sph = std::min<u64>(as * as * as, (u64)std::ceil(shell * (1.0 + 1e-9)) + 8u);
	return std::max(box, sph);
}

static hd_edit_desc noop_edit() { // AABB [0,0)^3: every node has lb >= p1, so nothing is affected
	hd_edit_desc d{};
	d.kind = HD_EDIT_AABB_DIG;
	return d;
}

static int frame_create(Pool &p, const hd_frame_config &cfg, u32 root, u32 croot, FrameState &F) {
	static u32 frame_ids = 0;
	const u32 L = p.dc.L;
	if (!cfg.max_edits || !cfg.max_passes || !cfg.n_views || !cfg.max_width || !cfg.max_height)
		return set_error("hd_frame_create: empty budget"), HD_ERR_ARG;
	HD_RET(make_filled(p));
	F.p = &p;
	F.cfg = cfg;
	F.pre = "f" + std::to_string(++frame_ids) + "_";
	// per-edit budget: what a sphere of radius <= R or a box of half-extent <= R can reach at any
	// position (worst alignment): frame_set holds every edit to it
	F.budget.assign(L, 0);
	for (u32 l = 1; l < L; ++l)
		F.budget[l] = budget_level_bound(cfg.max_radius, p.dc.voxel_level, l);
	// one plan per pass, as plan_pass would bound max_edits budget edits
	std::vector<hd_edit_desc> flat((size_t)cfg.max_passes * cfg.max_edits, noop_edit());
	const bool colors = (cfg.flags & HD_FRAME_COLORS) != 0;
	for (u32 k = 0; k < cfg.max_passes; ++k) {
		PassPlan pp;
		pp.first = k * cfg.max_edits;
		pp.count = cfg.max_edits;
		pp.stateful = colors;
		pp.async = true;
		pp.cap.assign(L, 0);
		pp.list_cap.assign(L, 0);
		pp.cap[0] = 1;
		pp.list_cap[0] = cfg.max_edits > 1 ? cfg.max_edits : 0;
		const u64 lim = 1ull << 31;
		for (u32 l = 1; l < L; ++l) {
			u64 S = F.budget[l] * cfg.max_edits;
			pp.cap[l] = std::min(std::min(S, 8 * pp.cap[l - 1]), lim);
			pp.list_cap[l] = cfg.max_edits > 1 ? std::min(2 * S, lim) : 0;
		}
		pp.ev_cap = 0;
		if (colors && !lds_jobs(p.dc))
			for (u32 l = p.dc.color_leaf; l + 1 < L; ++l)
				pp.ev_cap += 8 * pp.cap[l];
		pp.ev_cap = std::min<u64>(pp.ev_cap, lim);
		F.plans.push_back(pp);
	}
	if (colors && p.dc.voxel_level - p.dc.color_leaf > 7)
		return set_error("hd_frame_create: color leaves deeper than 8^7 voxels"), HD_ERR_UNSUPPORTED;
	HD_RET(async_setup(p, F.plans, flat, F.B, F.pre, &F.gen_mirror));
	// views, tiles, output
	F.tile = g_tuning.trace_tile;
	const u32 tiles1 = div_up(cfg.max_width, F.tile) * div_up(cfg.max_height, F.tile);
	F.max_tiles = tiles1 * cfg.n_views;
	HD_RET(p.scratch.get<hd_view>(F.pre + "views", cfg.n_views, &F.d_views));
	HD_RET(p.scratch.get<u32>(F.pre + "tiles", cfg.n_views + 1, &F.d_tiles));
	HD_RET(p.scratch.get<u32>(F.pre + "rgba", (u64)cfg.n_views * cfg.max_width * cfg.max_height, &F.d_rgba));
	for (int k = 0; k < 2; ++k) {
		HD_TRY(hipHostMalloc((void **)&F.h_edits[k], flat.size() * sizeof(hd_edit_desc), hipHostMallocDefault));
		HD_TRY(hipHostMalloc((void **)&F.h_views[k], cfg.n_views * sizeof(hd_view), hipHostMallocDefault));
		HD_TRY(hipHostMalloc((void **)&F.h_tiles[k], (cfg.n_views + 1) * 4, hipHostMallocDefault));
		HD_TRY(hipEventCreateWithFlags(&F.copied[k], hipEventDisableTiming));
		HD_TRY(hipEventRecord(F.copied[k], p.stream));
	}
	HD_TRY(hipHostMalloc((void **)&F.h_roots, 8 + 4 * kAccCount, hipHostMallocDefault));
	std::memset(F.h_views[0], 0, cfg.n_views * sizeof(hd_view));
	std::memset(F.h_tiles[0], 0, (cfg.n_views + 1) * 4);
	HD_TRY(hipMemcpyAsync(F.d_views, F.h_views[0], cfg.n_views * sizeof(hd_view), hipMemcpyHostToDevice, p.stream));
	HD_TRY(hipMemcpyAsync(F.d_tiles, F.h_tiles[0], (cfg.n_views + 1) * 4, hipMemcpyHostToDevice, p.stream));
	a_set_roots<<<1, 64, 0, p.stream>>>(F.B.roots, root, croot, F.B.acc);
	HD_TRY(hipGetLastError());
	const bool gc = (cfg.flags & HD_FRAME_GC) != 0;
	if (gc) // the GC's buffers are sized by the pool's configuration: the graph keeps their addresses
		HD_RET(p.gc_plan.prepare(p, ""));
	HD_TRY(hipStreamSynchronize(p.stream));
	// capture: the passes, the GC of the root they leave on the device, then the trace of it
	HD_TRY(hipStreamBeginCapture(p.stream, hipStreamCaptureModeThreadLocal));
	int rc = HD_OK;
	for (u32 k = 0; k < cfg.max_passes && rc == HD_OK; ++k)
		rc = queue_pass(p, F.plans[k], F.B, false);
	if (rc == HD_OK && gc)
		rc = gc_queue(p, p.gc_plan, p.stream, F.B.roots, 1, true);
	if (rc == HD_OK)
		rc = trace_on_stream(p, p.stream, 0, 0, F.d_views, F.d_tiles, cfg.n_views, F.max_tiles, F.d_rgba, F.B.roots);
	hipError_t e = hipStreamEndCapture(p.stream, &F.graph);
	HD_RET(rc);
	HD_TRY(e);
	HD_TRY(hipGraphInstantiate(&F.exec, F.graph, nullptr, nullptr, 0));
	return HD_OK;
}

static int frame_set(FrameState &F, const hd_edit_desc *edits, u32 n, const hd_view *views, u32 n_views) {
	Pool &p = *F.p;
	const hd_frame_config &cfg = F.cfg;
	const u32 L = p.dc.L;
	if (n_views > cfg.n_views)
		return set_error("hd_frame_set: %u views exceed the frame's %u", n_views, cfg.n_views), HD_ERR_ARG;
	for (u32 k = 0; k < n; ++k) {
		const hd_edit_desc &d = edits[k];
		if (d.kind > HD_EDIT_AABB_PAINT)
			return set_error("hd_frame_set: edit %u: column fills and unknown kinds are not frame edits", k), HD_ERR_ARG;
		if (!edit_color_ok(d))
			return set_error("hd_frame_set: edit %u: bits_per_weight > 3", k), HD_ERR_ARG;
		if (is_stateful(d) && !(cfg.flags & HD_FRAME_COLORS))
			return set_error("hd_frame_set: edit %u is colored but the frame has no color passes", k), HD_ERR_ARG;
		for (u32 l = 1; l < L; ++l)
			if (edit_level_bound(d, p.dc.voxel_level, l) > F.budget[l])
				return set_error("hd_frame_set: edit %u exceeds the frame's radius budget %u", k, cfg.max_radius),
				       HD_ERR_CAPACITY;
	}
	std::vector<std::vector<u32>> groups;
	if (n)
		schedule_groups(edits, n, 1u << p.dc.voxel_level, groups);
	if (groups.size() > cfg.max_passes)
		return set_error("hd_frame_set: the edits need %zu passes, the frame has %u", groups.size(), cfg.max_passes),
		       HD_ERR_CAPACITY;
	const u32 k = F.slot;
	F.slot ^= 1u;
	HD_TRY(hipEventSynchronize(F.copied[k])); // this slot's previous uploads have been read
	hd_edit_desc *he = F.h_edits[k];
	hd_view *hv = F.h_views[k];
	u32 *ht = F.h_tiles[k];
	for (u64 i = 0; i < (u64)cfg.max_passes * cfg.max_edits; ++i)
		he[i] = noop_edit();
	for (size_t g = 0; g < groups.size(); ++g) {
		if (groups[g].size() > cfg.max_edits)
			return set_error("hd_frame_set: a pass needs %zu edits, the frame has %u", groups[g].size(), cfg.max_edits),
			       HD_ERR_CAPACITY;
		for (size_t j = 0; j < groups[g].size(); ++j)
			he[g * cfg.max_edits + j] = edits[groups[g][j]];
	}
	u64 off = 0;
	ht[0] = 0;
	for (u32 v = 0; v < cfg.n_views; ++v) {
		hd_view w{};
		if (v < n_views) {
			w = views[v];
			if (!w.width || !w.height || w.width > cfg.max_width || w.height > cfg.max_height)
				return set_error("hd_frame_set: view %u is empty or larger than %ux%u", v, cfg.max_width,
				                 cfg.max_height),
				       HD_ERR_ARG;
			w.out_offset = off;
			off += (u64)w.width * w.height;
		}
		hv[v] = w;
		ht[v + 1] = ht[v] + (v < n_views ? div_up(w.width, F.tile) * div_up(w.height, F.tile) : 0u);
	}
	HD_TRY(hipMemcpyAsync(F.B.edits, he, (u64)cfg.max_passes * cfg.max_edits * sizeof(hd_edit_desc),
	                      hipMemcpyHostToDevice, p.stream));
	HD_TRY(hipMemcpyAsync(F.d_views, hv, cfg.n_views * sizeof(hd_view), hipMemcpyHostToDevice, p.stream));
	HD_TRY(hipMemcpyAsync(F.d_tiles, ht, (cfg.n_views + 1) * 4, hipMemcpyHostToDevice, p.stream));
	HD_TRY(hipEventRecord(F.copied[k], p.stream));
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
	for (u32 i = 0; i < n; ++i) {
		if (edits[i].kind > HD_EDIT_COLUMN_FILL)
			return set_error("unknown edit kind %u", edits[i].kind), HD_ERR_ARG;
		if (edits[i].kind == HD_EDIT_COLUMN_FILL && !p.hm_mn)
			return set_error("column fill needs hd_pool_set_heightmap"), HD_ERR_ARG;
		if (!edit_color_ok(edits[i]))
			return set_error("edit %u: bits_per_weight %u > 3", i, edits[i].color.bits_per_weight), HD_ERR_ARG;
	}
	if (n == 0) {
		p.last_groups = 0;
	} else if (mode == HD_MODE_BATCHED && n > 1) {
		HD_RET(edit_batched(p, r, c, edits, n, &r, &c));
	} else {
		HD_RET(edit_plain(p, r, c, edits, n, &r, &c));
	}
	HD_TRY(hipStreamSynchronize(p.stream));
	p.last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
	*root_out = r;
	*color_root_out = c;
	return HD_OK;
}

extern "C" int hd_edit_reserve(hd_pool *h, const hd_edit_desc *edits, uint32_t n, uint32_t mode) {
	if (!h || (!edits && n))
		return set_error("hd_edit_reserve: null argument"), HD_ERR_ARG;
	for (u32 i = 0; i < n; ++i) {
		if (edits[i].kind > HD_EDIT_COLUMN_FILL)
			return set_error("unknown edit kind %u", edits[i].kind), HD_ERR_ARG;
		if (!edit_color_ok(edits[i]))
			return set_error("edit %u: bits_per_weight %u > 3", i, edits[i].color.bits_per_weight), HD_ERR_ARG;
	}
	return edit_reserve(h->p, edits, n, mode);
}

// ---- frames (hipGraph): the C-ABI
struct hd_frame {
	hd::FrameState F;
};

extern "C" int hd_frame_create(hd_pool *h, const hd_frame_config *cfg, uint32_t root, uint32_t color_root,
                               hd_frame **out) {
	if (!h || !cfg || !out)
		return set_error("hd_frame_create: null argument"), HD_ERR_ARG;
	hd_frame *f = new hd_frame;
	int rc = frame_create(h->p, *cfg, root, color_root, f->F);
	if (rc != HD_OK) {
		delete f;
		return rc;
	}
	*out = f;
	return HD_OK;
}

extern "C" int hd_frame_set(hd_frame *f, const hd_edit_desc *edits, uint32_t n, const hd_view *views,
                            uint32_t n_views) {
	if (!f || (!edits && n) || (!views && n_views))
		return set_error("hd_frame_set: null argument"), HD_ERR_ARG;
	return frame_set(f->F, edits, n, views, n_views);
}

extern "C" int hd_frame_set_roots(hd_frame *f, uint32_t root, uint32_t color_root) {
	if (!f)
		return set_error("hd_frame_set_roots: null argument"), HD_ERR_ARG;
	Pool &p = *f->F.p;
	a_set_roots_keep<<<1, 64, 0, p.stream>>>(f->F.B.roots, root, color_root);
	HD_TRY(hipGetLastError());
	return HD_OK;
}

extern "C" int hd_frame_launch(hd_frame *f) {
	if (!f)
		return set_error("hd_frame_launch: null argument"), HD_ERR_ARG;
	FrameState &F = f->F;
	Pool &p = *F.p;
	HD_RET(gen_reserve(p, F.B, F.cfg.max_passes)); // outside the graph: may queue a table reset
	HD_TRY(hipGraphLaunch(F.exec, p.stream));
	return HD_OK;
}

extern "C" int hd_frame_wait(hd_frame *f, uint32_t *root_out, uint32_t *color_root_out) {
	if (!f)
		return set_error("hd_frame_wait: null argument"), HD_ERR_ARG;
	FrameState &F = f->F;
	Pool &p = *F.p;
	HD_TRY(hipMemcpyAsync(F.h_roots, F.B.roots, 8, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipMemcpyAsync(F.h_roots + 2, F.B.acc, 4 * kAccCount, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	if (F.h_roots[2 + kAccOverflow])
		return set_error("hd_frame_wait: a level outgrew its bound (internal error)"), HD_ERR_CAPACITY;
	if (F.cfg.flags & HD_FRAME_GC) { // the captured GC found the filled chain again on the device
		HD_TRY(hipMemcpy(p.filled.data(), p.d_filled, p.dc.L * 4, hipMemcpyDeviceToHost));
	}
	if (root_out)
		*root_out = F.h_roots[0];
	if (color_root_out)
		*color_root_out = F.h_roots[1];
	return HD_OK;
}

extern "C" uint32_t *hd_frame_rgba(hd_frame *f) { return f ? f->F.d_rgba : nullptr; }

extern "C" int hd_frame_destroy(hd_frame *f) {
	if (f) {
		(void)hipStreamSynchronize(f->F.p->stream);
		delete f;
	}
	return HD_OK;
}
