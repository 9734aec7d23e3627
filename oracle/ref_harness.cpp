// ref_harness.cpp -- binds the VkHashDAG reference headers (compiled from the reference tree,
// never copied) to in-memory storage so the test suite can run the reference's own CPU edit,
// GC, VBR and traversal code. TEST INFRASTRUCTURE ONLY; built by oracle/Makefile into
// oracle/_ref/libhdref.so when the reference tree is present.
//
// What is the reference here: hashdag::NodePoolBase (NodePool.hpp), NodePoolThreadedEdit,
// NodePoolThreadedGC, NodePoolTraversal, VBREditorWrapper (VBREditor.hpp), VBRChunkWriter /
// VBRBitsetWriter (VBRColor.hpp), MurmurHasher32 (Hasher.hpp), libfork's busy_pool.
// What is ours: the page store (semantics of src/DAGNodePool.hpp:56-74), the color octree
// store (semantics of src/DAGColorPool.hpp:139-204, which needs a Vulkan device to construct)
// and the editor predicates (src/main.cpp:32-150 cannot be compiled without Vulkan/GLFW; the
// harness adapts hd_oracle.hpp's Editor to the reference's Editor concepts).
#include <hashdag/NodePool.hpp>
#include <hashdag/NodePoolThreadedEdit.hpp>
#include <hashdag/NodePoolThreadedGC.hpp>
#include <hashdag/NodePoolTraversal.hpp>
#include <hashdag/VBREditor.hpp>

#include <libfork/schedule/busy_pool.hpp>
#include <parallel_hashmap/phmap.h>

#include "hd_oracle.hpp"

#include <atomic>
#include <chrono>
#include <mutex>

namespace ref_phmap {
template <typename K, typename V> using map = phmap::flat_hash_map<K, V>;
template <typename K> using set = phmap::flat_hash_set<K>;
} // namespace ref_phmap

using u32 = uint32_t;
using u64 = uint64_t;

class RefPool final : public hashdag::NodePoolBase<RefPool, u32>,
                      public hashdag::NodePoolTraversal<RefPool, u32>,
                      public hashdag::NodePoolThreadedEdit<RefPool, u32>,
                      public hashdag::NodePoolThreadedGC<RefPool, u32, ref_phmap::map, ref_phmap::set> {
public:
	using WordSpanHasher = hashdag::MurmurHasher32;

	std::unique_ptr<u32[]> m_bucket_words;
	std::unique_ptr<std::unique_ptr<u32[]>[]> m_pages;
	std::array<std::mutex, 1024> m_mutexes{};

	std::mutex &GetBucketRefMutex(u32 b) { return m_mutexes[b % m_mutexes.size()]; }
	u32 &GetBucketRefWords(u32 b) { return m_bucket_words[b]; }
	const u32 *ReadPage(u32 p) const { return m_pages[p].get(); }
	void ZeroPage(u32 p, u32 off, u32 n) { std::fill(m_pages[p].get() + off, m_pages[p].get() + off + n, 0u); }
	void WritePage(u32 p, u32 off, std::span<const u32> s) {
		if (!m_pages[p]) {
			m_pages[p] = std::make_unique<u32[]>(GetConfig().GetWordsPerPage());
		}
		std::copy(s.begin(), s.end(), m_pages[p].get() + off);
	}
	void FreePage(u32 p) { m_pages[p] = nullptr; }

	explicit RefPool(const hashdag::Config<u32> &c) : hashdag::NodePoolBase<RefPool, u32>(c) {
		m_bucket_words = std::make_unique<u32[]>(GetConfig().GetTotalBuckets());
		m_pages = std::make_unique<std::unique_ptr<u32[]>[]>(GetConfig().GetTotalPages());
	}
	u32 word(u64 w) const {
		const u32 *p = m_pages[w >> GetConfig().word_bits_per_page].get();
		return p ? p[w & (GetConfig().GetWordsPerPage() - 1u)] : 0u;
	}
};

// Color octree store with DAGColorPool's semantics; appends are atomic so ThreadedEdit can
// run VBREditorWrapper concurrently (as SafePagedVector allows in the reference).
struct RefOctree {
	struct Pointer {
		static constexpr u32 kDataBits = 30u;
		enum class Tag { kNode = 0, kColor, kLeaf, kNull };
		u32 pointer;
		Pointer() : Pointer(Tag::kNull, 0u) {}
		Pointer(Tag t, u32 d) : pointer{(static_cast<u32>(t) << kDataBits) | d} {}
		Tag GetTag() const { return static_cast<Tag>(pointer >> kDataBits); }
		u32 GetData() const { return pointer & ((1u << kDataBits) - 1u); }
		bool operator==(const Pointer &r) const { return pointer == r.pointer; }
	};
	using Node = std::array<Pointer, 8>;
	u32 leaf_level = 0;
	bool keep_history = false;
	std::unique_ptr<Node[]> nodes;
	std::unique_ptr<u32[]> leaves;
	u64 node_cap = 0, leaf_cap = 0;
	std::atomic<u64> node_count{0}, leaf_count{0};

	Pointer GetChild(Pointer p, auto idx) const {
		return p.GetTag() == Pointer::Tag::kNode ? nodes[p.GetData()][idx]
		                                         : (p.GetTag() == Pointer::Tag::kColor ? p : Pointer{});
	}
	static hashdag::VBRColor GetFill(Pointer p) {
		return p.GetTag() == Pointer::Tag::kColor ? hashdag::VBRColor{hashdag::RGB8Color{p.GetData()}}
		                                          : hashdag::VBRColor{};
	}
	Pointer SetNode(Pointer p, std::span<const Pointer, 8> ch) {
		if (std::ranges::all_of(ch, [](Pointer c) { return c.GetTag() == Pointer::Tag::kNull; }))
			return {};
		if (Pointer c = ch[0];
		    c.GetTag() == Pointer::Tag::kColor && std::ranges::all_of(ch.subspan(1), [c](Pointer q) { return q == c; }))
			return c;
		if (p.GetTag() == Pointer::Tag::kNode && std::ranges::equal(nodes[p.GetData()], ch))
			return p;
		u64 idx = node_count.fetch_add(1);
		if (idx >= node_cap)
			return p;
		std::ranges::copy(ch, nodes[idx].begin());
		return Pointer{Pointer::Tag::kNode, (u32)idx};
	}
	static Pointer ClearNode(Pointer) { return {}; }
	static Pointer FillNode(Pointer, hashdag::VBRColor c) {
		return Pointer{Pointer::Tag::kColor, hashdag::RGB8Color{c.Get()}.GetData()};
	}
	hashdag::VBRChunk<u32, hashdag::VBRWriterContainer> GetLeaf(Pointer p) const {
		if (p.GetTag() != Pointer::Tag::kLeaf)
			return {};
		u64 i = p.GetData() + 1;
		u32 M = leaves[i], B = leaves[i + 1], W = leaves[i + 2];
		i += 3;
		std::vector<hashdag::VBRMacroBlock> mb(M);
		std::vector<hashdag::VBRBlockHeader> bh(B);
		for (u32 k = 0; k < M; ++k, i += 2)
			mb[k] = {leaves[i], leaves[i + 1]};
		for (u32 k = 0; k < B; ++k, i += 2) {
			bh[k].colors = leaves[i];
			bh[k].packed_14_2_16 = leaves[i + 1];
		}
		std::vector<u32> w(leaves.get() + i, leaves.get() + i + W);
		return {std::move(mb), std::move(bh), hashdag::VBRBitset<u32, hashdag::VBRWriterContainer>{std::move(w)}};
	}
	void write_chunk(u64 i, const hashdag::VBRChunk<u32, hashdag::VBRWriterContainer> &c) {
		leaves[i++] = (u32)c.GetMacroBlocks().size();
		leaves[i++] = (u32)c.GetBlockHeaders().size();
		leaves[i++] = (u32)c.GetWeightBits().GetWords().size();
		for (auto &m : c.GetMacroBlocks())
			leaves[i++] = m.first_block, leaves[i++] = m.weight_start;
		for (auto &h : c.GetBlockHeaders())
			leaves[i++] = h.colors, leaves[i++] = h.packed_14_2_16;
		for (u32 w : c.GetWeightBits().GetWords())
			leaves[i++] = w;
	}
	Pointer SetLeaf(Pointer p, hashdag::VBRChunk<u32, hashdag::VBRWriterContainer> &&c) {
		u64 data_size = c.GetMacroBlocks().size() * 2 + c.GetBlockHeaders().size() * 2 +
		                c.GetWeightBits().GetWords().size() + 4;
		u64 append_size = (data_size & 1) ? data_size + 1 : data_size;
		if (!keep_history && p.GetTag() == Pointer::Tag::kLeaf) {
			u64 idx = p.GetData(), block = leaves[idx];
			if (data_size <= block) {
				write_chunk(idx + 1, c);
				return p;
			}
			append_size = std::max(block << 1, append_size);
		}
		u64 idx = leaf_count.fetch_add(append_size);
		if (idx + append_size > leaf_cap)
			return p;
		std::fill(leaves.get() + idx, leaves.get() + idx + append_size, 0u);
		leaves[idx] = (u32)append_size;
		write_chunk(idx + 1, c);
		return Pointer{Pointer::Tag::kLeaf, (u32)idx};
	}
	u32 GetLeafLevel() const { return leaf_level; }
};
static_assert(hashdag::VBROctree<RefOctree, u32>);

namespace {
hashdag::VBRColor to_ref(const hdo::VBRColor &c) {
	if (!c.has)
		return {};
	if (c.bpw == 0)
		return hashdag::VBRColor{hashdag::RGB8Color{c.colors}};
	return hashdag::VBRColor{hashdag::R5G6B5Color{(uint16_t)(c.colors & 0xFFFFu)},
	                         hashdag::R5G6B5Color{(uint16_t)(c.colors >> 16)}, c.weight, c.bpw};
}
hdo::VBRColor from_ref(const hashdag::VBRColor &c) {
	hdo::VBRColor v;
	v.has = c.HasValue();
	v.colors = c.GetColors();
	v.weight = c.GetWeight();
	v.bpw = c.GetBitsPerWeight();
	return v;
}

// Adapters: the oracle's editor predicates presented through the reference's concepts.
struct StatelessEd {
	const hdo::Editor *e;
	const hdo::Config *cfg;
	hashdag::EditType EditNode(const hashdag::Config<u32> &, const hashdag::NodeCoord<u32> &c,
	                           hashdag::NodePointer<u32>) const {
		u32 pos[3] = {c.pos.x, c.pos.y, c.pos.z};
		return (hashdag::EditType)e->edit_node(*cfg, c.level, pos);
	}
	bool EditVoxel(const hashdag::Config<u32> &, const hashdag::NodeCoord<u32> &c, bool v) const {
		u32 pos[3] = {c.pos.x, c.pos.y, c.pos.z};
		return e->edit_voxel(pos, v);
	}
};
struct VBREd {
	const hdo::Editor *e;
	const hdo::Config *cfg;
	hashdag::EditType EditNode(const hashdag::Config<u32> &, const hashdag::NodeCoord<u32> &c,
	                           hashdag::NodePointer<u32> ptr, hashdag::VBRColor &fc) const {
		u32 pos[3] = {c.pos.x, c.pos.y, c.pos.z};
		hdo::VBRColor x = from_ref(fc);
		auto t = e->edit_node_vbr(*cfg, c.level, pos, ptr.HasValue(), x);
		fc = to_ref(x);
		return (hashdag::EditType)t;
	}
	bool EditVoxel(const hashdag::Config<u32> &, const hashdag::NodeCoord<u32> &c, bool v,
	               hashdag::VBRColor &color) const {
		u32 pos[3] = {c.pos.x, c.pos.y, c.pos.z};
		hdo::VBRColor x = from_ref(color);
		bool r = e->edit_voxel_vbr(pos, v, x);
		color = to_ref(x);
		return r;
	}
};
static_assert(hashdag::StatelessEditor<StatelessEd, u32>);
static_assert(hashdag::VBREditor<VBREd, u32>);
static_assert((int)hashdag::EditType::kNotAffected == (int)hdo::EditType::kNotAffected &&
              (int)hashdag::EditType::kProceed == (int)hdo::EditType::kProceed &&
              (int)hashdag::EditType::kFill == (int)hdo::EditType::kFill &&
              (int)hashdag::EditType::kClear == (int)hdo::EditType::kClear);

struct Ref {
	hdo::Config ocfg;
	RefPool pool;
	RefOctree octree;
	hdo::Heightmap hm;
	std::unique_ptr<lf::busy_pool> workers;
	Ref(const hashdag::Config<u32> &c, const hdo::Config &oc) : ocfg(oc), pool(c) {}
};

template <typename Ed>
u32 run_edit(Ref *r, u32 root, const Ed &ed, u32 threaded, u32 max_task_level, u32 *color_out) {
	if (!threaded)
		return r->pool.Edit(hashdag::NodePointer<u32>{root}, ed, [&](hashdag::NodePointer<u32> p, auto &&st) {
			if constexpr (requires { st.octree_node; })
				*color_out = st.octree_node.pointer;
			return *p;
		});
	return r->pool.ThreadedEdit(r->workers.get(), hashdag::NodePointer<u32>{root}, ed, max_task_level,
	                            [&](hashdag::NodePointer<u32> p, auto &&st) {
		                            if constexpr (requires { st.octree_node; })
			                            *color_out = st.octree_node.pointer;
		                            return *p;
	                            });
}

u32 edit_one(Ref *r, u32 root, const hd_edit_desc &d, u32 *color_root, u32 threaded, u32 max_task_level) {
	hdo::Editor e;
	e.d = d;
	e.hm = &r->hm;
	if (e.stateful()) {
		RefOctree::Pointer cr;
		cr.pointer = *color_root;
		hashdag::VBREditorWrapper<u32, VBREd, RefOctree> w{
		    .editor = VBREd{&e, &r->ocfg}, .p_octree = &r->octree, .octree_root = cr};
		return run_edit(r, root, w, threaded, max_task_level, color_root);
	}
	u32 dummy;
	hashdag::StatelessEditorWrapper<u32, StatelessEd> w{.editor = StatelessEd{&e, &r->ocfg}};
	return run_edit(r, root, w, threaded, max_task_level, &dummy);
}
} // namespace

extern "C" {

void *hdr_pool_create(u32 wbpp, u32 pbpb, u32 L, const u32 *bits, u32 color_leaf_level, u32 keep_history,
                      u32 workers, u64 node_cap, u64 leaf_cap) {
	hashdag::Config<u32> c{.word_bits_per_page = wbpp, .page_bits_per_bucket = pbpb,
	                       .bucket_bits_each_level = std::vector<u32>(bits, bits + L)};
	if (!hashdag::Config<u32>::Validate(c))
		return nullptr;
	hdo::Config oc;
	oc.wbpp = wbpp;
	oc.pbpb = pbpb;
	oc.bucket_bits.assign(bits, bits + L);
	auto *r = new Ref(c, oc);
	r->octree.leaf_level = color_leaf_level;
	r->octree.keep_history = keep_history != 0;
	r->octree.node_cap = node_cap;
	r->octree.leaf_cap = leaf_cap;
	r->octree.nodes = std::make_unique<RefOctree::Node[]>(node_cap);
	r->octree.leaves = std::make_unique<u32[]>(leaf_cap);
	r->workers = std::make_unique<lf::busy_pool>(workers ? workers : 1);
	return r;
}
void hdr_pool_destroy(void *p) { delete static_cast<Ref *>(p); }
void hdr_set_heightmap(void *p, const u32 *h, u32 res) { static_cast<Ref *>(p)->hm.set(h, res); }

// Applies n edits in order with NodePoolBase::Edit (threaded == 0) or
// NodePoolThreadedEdit::ThreadedEdit (threaded != 0, max_task_level as main.cpp:214-223).
u32 hdr_edit(void *p, u32 root, const hd_edit_desc *edits, u32 n, u32 *color_root, u32 threaded,
             u32 max_task_level) {
	auto *r = static_cast<Ref *>(p);
	for (u32 i = 0; i < n; ++i)
		root = edit_one(r, root, edits[i], color_root, threaded, max_task_level);
	return root;
}
// Timed variant: steady_clock around each edit call (main.cpp:152-157); returns seconds.
double hdr_edit_timed(void *p, u32 root, const hd_edit_desc *edits, u32 n, u32 *color_root, u32 threaded,
                      u32 max_task_level, u32 *root_out) {
	auto *r = static_cast<Ref *>(p);
	double total = 0;
	for (u32 i = 0; i < n; ++i) {
		auto t0 = std::chrono::steady_clock::now();
		root = edit_one(r, root, edits[i], color_root, threaded, max_task_level);
		total += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
	}
	*root_out = root;
	return total;
}
void hdr_gc(void *p, u32 *roots, u32 n) {
	auto *r = static_cast<Ref *>(p);
	std::vector<hashdag::NodePointer<u32>> rp;
	for (u32 i = 0; i < n; ++i)
		rp.emplace_back(roots[i]);
	rp = r->pool.ThreadedGC(r->workers.get(), rp);
	for (u32 i = 0; i < n; ++i)
		roots[i] = *rp[i];
}
void hdr_read_words(void *p, u64 first, u64 count, u32 *out) {
	auto *r = static_cast<Ref *>(p);
	for (u64 i = 0; i < count; ++i)
		out[i] = r->pool.word(first + i);
}
void hdr_bucket_words(void *p, u32 *out) {
	auto *r = static_cast<Ref *>(p);
	std::copy(r->pool.m_bucket_words.get(), r->pool.m_bucket_words.get() + r->pool.GetConfig().GetTotalBuckets(), out);
}
// Used words of every bucket packed in bucket order; returns the count (out may be null).
u64 hdr_live_words(void *p, u32 *out, u64 cap) {
	auto *r = static_cast<Ref *>(p);
	const auto &cfg = r->pool.GetConfig();
	const u64 nb = cfg.GetTotalBuckets(), wbpb = (u64)cfg.word_bits_per_page + cfg.page_bits_per_bucket;
	u64 at = 0;
	for (u64 b = 0; b < nb; ++b)
		for (u32 i = 0; i < r->pool.m_bucket_words[b]; ++i, ++at)
			if (out && at < cap)
				out[at] = r->pool.word((b << wbpb) + i);
	return at;
}
u32 hdr_filled(void *p, u32 *out) {
	auto *r = static_cast<Ref *>(p);
	for (size_t i = 0; i < r->pool.m_filled_node_pointers.size(); ++i)
		out[i] = *r->pool.m_filled_node_pointers[i];
	return (u32)r->pool.m_filled_node_pointers.size();
}
u64 hdr_color_node_count(void *p) { return static_cast<Ref *>(p)->octree.node_count.load(); }
u64 hdr_color_leaf_count(void *p) { return static_cast<Ref *>(p)->octree.leaf_count.load(); }
void hdr_color_nodes(void *p, u32 *out) {
	auto *r = static_cast<Ref *>(p);
	for (u64 i = 0; i < r->octree.node_count.load(); ++i)
		for (int k = 0; k < 8; ++k)
			out[8 * i + k] = r->octree.nodes[i][k].pointer;
}
void hdr_color_leaves(void *p, u32 *out) {
	auto *r = static_cast<Ref *>(p);
	std::copy(r->octree.leaves.get(), r->octree.leaves.get() + r->octree.leaf_count.load(), out);
}
// NodePoolTraversal::Traversal<float> (NodePoolTraversal.hpp:93-256): out = hit xyz + flag.
void hdr_traversal(void *p, u32 root, const float *o, const float *d, u64 n, float *out) {
	auto *r = static_cast<Ref *>(p);
	for (u64 i = 0; i < n; ++i) {
		auto h = r->pool.Traversal<float>(hashdag::NodePointer<u32>{root}, glm::vec3{o[3 * i], o[3 * i + 1], o[3 * i + 2]},
		                                  glm::vec3{d[3 * i], d[3 * i + 1], d[3 * i + 2]});
		out[4 * i + 3] = h ? 1.0f : 0.0f;
		if (h)
			out[4 * i] = h->x, out[4 * i + 1] = h->y, out[4 * i + 2] = h->z;
		else
			out[4 * i] = out[4 * i + 1] = out[4 * i + 2] = 0.0f;
	}
}
u32 hdr_hash_span(const u32 *w, u32 n) { return hashdag::MurmurHasher32{}(std::span<const u32>{w, n}); }
u32 hdr_hash_leaf(u32 w0, u32 w1) {
	u32 a[2] = {w0, w1};
	return hashdag::MurmurHasher32{}(std::span<const u32, 2>{a, 2});
}

// The reference VBRChunkWriter driven by the op stream of hdo_vbr_write (oracle_capi.cpp).
u64 hdr_vbr_write(const u32 *src, u32 src_words, const u32 *ops, u32 nops, u32 *out, u64 cap) {
	hashdag::VBRChunk<u32, hashdag::VBRWriterContainer> sc;
	if (src && src_words) {
		u32 M = src[0], B = src[1], W = src[2], i = 3;
		std::vector<hashdag::VBRMacroBlock> mb(M);
		std::vector<hashdag::VBRBlockHeader> bh(B);
		for (u32 k = 0; k < M; ++k, i += 2)
			mb[k] = {src[i], src[i + 1]};
		for (u32 k = 0; k < B; ++k, i += 2) {
			bh[k].colors = src[i];
			bh[k].packed_14_2_16 = src[i + 1];
		}
		std::vector<u32> w(src + i, src + i + W);
		sc = {std::move(mb), std::move(bh), hashdag::VBRBitset<u32, hashdag::VBRWriterContainer>{std::move(w)}};
	}
	hashdag::VBRChunkWriter<u32, hashdag::VBRWriterContainer> w{std::move(sc)};
	for (u32 k = 0; k < nops; ++k) {
		const u32 *op = ops + 4 * k;
		hdo::VBRColor c;
		c.colors = op[2];
		c.weight = op[3] & 0xFFu;
		c.bpw = (op[3] >> 8) & 0xFFu;
		c.has = (op[3] >> 16) & 1u;
		hashdag::VBRColor rc = to_ref(c);
		if (op[0] == 0)
			w.Copy(op[1], rc);
		else if (op[0] == 1)
			w.Push(rc, op[1]);
		else
			for (u32 j = 0; j < op[1]; ++j)
				w.Edit([&](hashdag::VBRColor &x) {
					if (c.has)
						x = rc;
					return true;
				}, hashdag::VBRColor{});
	}
	auto ch = w.Flush();
	std::vector<u32> o = {(u32)ch.GetMacroBlocks().size(), (u32)ch.GetBlockHeaders().size(),
	                      (u32)ch.GetWeightBits().GetWords().size()};
	for (auto &m : ch.GetMacroBlocks())
		o.push_back(m.first_block), o.push_back(m.weight_start);
	for (auto &h : ch.GetBlockHeaders())
		o.push_back(h.colors), o.push_back(h.packed_14_2_16);
	for (u32 x : ch.GetWeightBits().GetWords())
		o.push_back(x);
	for (u64 i = 0; i < o.size() && i < cap; ++i)
		out[i] = o[i];
	return o.size();
}

// Load a pool state (per-bucket counters + their used words packed in bucket order, the filled
// chain and the color arrays), e.g. one downloaded from the device, so the reference continues
// from exactly the same state.
void hdr_load_live(void *p, const u32 *bucket_words, const u32 *live, u64 nlive, const u32 *filled, u32 nfilled,
                   const u32 *cnodes, u64 ncnodes, const u32 *cleaves, u64 ncleaves) {
	auto *r = static_cast<Ref *>(p);
	auto &pool = r->pool;
	const auto &cfg = pool.GetConfig();
	u64 nb = cfg.GetTotalBuckets(), at = 0;
	for (u64 b = 0; b < nb; ++b) {
		u32 bw = bucket_words[b];
		pool.m_bucket_words[b] = bw;
		for (u32 off = 0; off < bw;) {
			u32 page = (u32)((b << cfg.page_bits_per_bucket) | (off >> cfg.word_bits_per_page));
			u32 po = off & (cfg.GetWordsPerPage() - 1u);
			u32 n = std::min(bw - off, cfg.GetWordsPerPage() - po);
			pool.WritePage(page, po, std::span<const u32>{live + at, n});
			at += n;
			off += n;
		}
	}
	(void)nlive;
	pool.m_filled_node_pointers.clear();
	for (u32 i = 0; i < nfilled; ++i)
		pool.m_filled_node_pointers.emplace_back(filled[i]);
	for (u64 i = 0; i < ncnodes; ++i)
		for (int k = 0; k < 8; ++k)
			r->octree.nodes[i][k].pointer = cnodes[8 * i + k];
	r->octree.node_count = ncnodes;
	std::copy(cleaves, cleaves + ncleaves, r->octree.leaves.get());
	r->octree.leaf_count = ncleaves;
}

} // extern "C"
