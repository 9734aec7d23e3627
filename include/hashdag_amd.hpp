// hashdag_amd.hpp -- C++20 host facade over hd_capi.h with the reference's names.
//
// It mirrors the include/hashdag plug-in surface for the device-supported path:
//   hashdag::Config / DefaultConfig (Config.hpp:15-75)        -> hashdag_amd::Config / DefaultConfig
//   hashdag::NodePointer (NodePointer.hpp:13-31)              -> hashdag_amd::NodePointer
//   hashdag::EditType (Editor.hpp:18)                         -> hashdag_amd::EditType
//   AABBEditor / SphereEditor<Mode> (src/main.cpp:32-150)     -> descriptor editors (ToDesc())
//   VBRColor / RGB8Color (VBRColor.hpp:19-44, Color.hpp:17-50) -> hashdag_amd::VBRColor
//   NodePoolBase::Edit, NodePoolThreadedEdit::ThreadedEdit,
//   NodePoolThreadedGC::ThreadedGC, NodePoolTraversal::Traversal (NodePool.hpp:404-417,
//   NodePoolThreadedEdit.hpp:104-126, NodePoolThreadedGC.hpp:394-403, NodePoolTraversal.hpp:93-256)
//                                                             -> hashdag_amd::DeviceNodePool
//   DAGNodePool::Create / SetRoot / GetRoot (src/DAGNodePool.hpp:41-99) and DAGColorPool::Config
//   (src/DAGColorPool.hpp:36-41)                              -> DeviceNodePool::Create, ColorConfig
//
//   NodeCoord (NodeCoord.hpp:14-94), StatelessEditor (Editor.hpp:36-40) -> hashdag_amd::NodeCoord,
//   HostStatelessEditor: any functor with the reference's EditNode / EditVoxel runs on the host
//   (hd_edit_host) and composes with device edits on the same pool
//   VBREditorWrapper / StatelessEditorWrapper (VBREditor.hpp:26-107, Editor.hpp:42-55)
//                                                             -> the same names, over descriptors
//
// Error behaviour follows the reference (built with -fno-exceptions): Create returns nullptr, a
// failed edit returns the root it was given, GC returns the roots unchanged, Traversal returns
// nullopt; GetLastError() explains the last failure. Device editors are plain descriptors of the
// closed set the device runs (the integer predicates of src/main.cpp); any other stateless functor
// editor runs on the host. The busy_pool and max_task_level arguments of the reference's threaded
// calls are accepted and unused: the device schedules its own parallelism.
#pragma once

#include "hd_capi.h"

#include <array>
#include <concepts>
#include <cstdint>
#include <memory>
#include <optional>
#include <span>
#include <variant>
#include <vector>

namespace hashdag_amd {

using Word = uint32_t;

struct Config {
	inline static constexpr Word kWordsPerLeaf = 2;
	inline static constexpr Word kMinWordBitsPerPage = 4;
	Word word_bits_per_page;
	Word page_bits_per_bucket;
	std::vector<Word> bucket_bits_each_level;

	inline static constexpr Word GetWordsPerLeaf() { return kWordsPerLeaf; }
	inline Word GetWordsPerPage() const { return Word(1u) << word_bits_per_page; }
	inline Word GetPagesPerBucket() const { return Word(1u) << page_bits_per_bucket; }
	inline Word GetWordsPerBucket() const { return Word(1u) << (word_bits_per_page + page_bits_per_bucket); }
	inline Word GetWordBitsPerBucket() const { return word_bits_per_page + page_bits_per_bucket; }
	inline Word GetBucketsAtLevel(Word level) const { return Word(1u) << bucket_bits_each_level[level]; }
	inline Word GetNodeLevels() const { return (Word)bucket_bits_each_level.size(); }
	inline Word GetLeafLevel() const { return GetNodeLevels(); }
	inline Word GetVoxelLevel() const { return GetNodeLevels() + 1u; }
	inline Word GetResolution() const { return Word(1u) << GetVoxelLevel(); }
	inline std::vector<Word> GetLevelBaseBucketIndices() const {
		std::vector<Word> base(bucket_bits_each_level.size());
		for (Word i = 1; i < base.size(); ++i)
			base[i] = GetBucketsAtLevel(i - 1) + base[i - 1];
		return base;
	}
	inline Word GetTotalBuckets() const {
		Word total = 0;
		for (Word b : bucket_bits_each_level)
			total += Word(1u) << b;
		return total;
	}
	inline Word GetTotalPages() const { return GetTotalBuckets() << page_bits_per_bucket; }
	inline Word GetTotalWords() const { return GetTotalBuckets() << (word_bits_per_page + page_bits_per_bucket); }
	// Config::Validate (Config.hpp:48-56), plus the device limit of HD_MAX_LEVELS node levels
	inline static bool Validate(const Config &c) {
		if (c.word_bits_per_page < kMinWordBitsPerPage || c.bucket_bits_each_level.empty() ||
		    c.bucket_bits_each_level.size() > HD_MAX_LEVELS)
			return false;
		uint64_t buckets = 0;
		for (Word b : c.bucket_bits_each_level)
			buckets += 1ull << b;
		return buckets * c.GetWordsPerBucket() - 1 <= uint64_t(Word(-2));
	}
};

struct DefaultConfig {
	uint32_t level_count = 17, top_level_count = 9;
	Word word_bits_per_page = 9;
	Word page_bits_per_bucket = 2;
	Word bucket_bits_per_top_level = 10;
	Word bucket_bits_per_bottom_level = 16;
	inline Config operator()() const {
		std::vector<Word> bits;
		for (uint32_t l = 0; l + 1 < level_count; ++l)
			bits.push_back(l < top_level_count ? bucket_bits_per_top_level : bucket_bits_per_bottom_level);
		return {word_bits_per_page, page_bits_per_bucket, std::move(bits)};
	}
};

class NodePointer {
	Word m_node;

public:
	inline constexpr NodePointer() : m_node(HD_NULL) {}
	inline constexpr NodePointer(Word node) : m_node{node} {}
	inline constexpr bool HasValue() const { return m_node != HD_NULL; }
	inline constexpr explicit operator bool() const { return HasValue(); }
	inline constexpr bool operator==(NodePointer r) const { return m_node == r.m_node; }
	inline constexpr bool operator!=(NodePointer r) const { return m_node != r.m_node; }
	inline constexpr Word Value() const { return m_node; }
	inline constexpr Word operator*() const { return Value(); }
	inline constexpr static NodePointer Null() { return NodePointer{}; }
};

enum class EditType { kNotAffected = 0, kProceed, kFill, kClear };
static_assert((int)EditType::kProceed == HD_EDIT_TYPE_PROCEED && (int)EditType::kFill == HD_EDIT_TYPE_FILL &&
              (int)EditType::kClear == HD_EDIT_TYPE_CLEAR);

// hashdag::NodeCoord (NodeCoord.hpp:14-94): a node at `level`, pos in nodes of that level. The member
// names match the reference's (pos.x / .y / .z), so its editors' integer code compiles unchanged.
struct NodeCoord {
	struct Vec3 {
		Word x = 0, y = 0, z = 0;
	};
	Word level = 0;
	Vec3 pos{};
	inline constexpr NodeCoord GetChildCoord(Word child_index) const {
		return {level + 1, {(pos.x << 1u) | (child_index & 1u), (pos.y << 1u) | ((child_index >> 1u) & 1u),
		                    (pos.z << 1u) | ((child_index >> 2u) & 1u)}};
	}
	inline constexpr Word GetChildIndex() const { return (pos.x & 1u) | ((pos.y & 1u) << 1u) | ((pos.z & 1u) << 2u); }
	inline constexpr Vec3 GetLowerBoundAtLevel(Word at_level) const {
		const Word b = at_level - level;
		return {pos.x << b, pos.y << b, pos.z << b};
	}
	inline constexpr Vec3 GetUpperBoundAtLevel(Word at_level) const {
		const Word b = at_level - level;
		return {(pos.x + 1u) << b, (pos.y + 1u) << b, (pos.z + 1u) << b};
	}
};

// hashdag::VBRColor: RGB8 when bits_per_weight == 0, else two RGB565 endpoints + a weight.
struct VBRColor {
	uint32_t colors = 0;
	uint8_t weight = 0, bits_per_weight = 0;
	bool has_value = false;
	inline static VBRColor RGB8(uint32_t rgb) { return {rgb & 0xFFFFFFu, 0, 0, true}; }
	inline hd_color ToHD() const { return {colors, weight, bits_per_weight, uint8_t(has_value ? 1 : 0), 0}; }
};

// AABBEditor (main.cpp:32-70): fill [aabb_min, aabb_max) with color.
struct AABBEditor {
	std::array<Word, 3> aabb_min{}, aabb_max{};
	VBRColor color{};
	inline hd_edit_desc ToDesc() const { return box_desc(HD_EDIT_AABB_FILL, aabb_min, aabb_max, color); }
	inline static hd_edit_desc box_desc(uint32_t kind, const std::array<Word, 3> &lo, const std::array<Word, 3> &hi,
	                                    const VBRColor &c) {
		hd_edit_desc d{};
		d.kind = kind;
		d.colored = c.has_value && kind != HD_EDIT_AABB_DIG ? 1u : 0u;
		for (int i = 0; i < 3; ++i)
			d.p0[i] = lo[i], d.p1[i] = hi[i];
		d.color = c.ToHD();
		return d;
	}
};
// box dig / paint: the AABB predicate with the sphere editors' dig and paint rules
struct AABBDigEditor {
	std::array<Word, 3> aabb_min{}, aabb_max{};
	inline hd_edit_desc ToDesc() const { return AABBEditor::box_desc(HD_EDIT_AABB_DIG, aabb_min, aabb_max, {}); }
};
struct AABBPaintEditor {
	std::array<Word, 3> aabb_min{}, aabb_max{};
	VBRColor color{};
	inline hd_edit_desc ToDesc() const { return AABBEditor::box_desc(HD_EDIT_AABB_PAINT, aabb_min, aabb_max, color); }
};

// SphereEditor<kFill | kDig | kPaint> (main.cpp:73-150): |p - center|^2 <= r2.
enum class SphereEditMode { kFill, kDig, kPaint };
template <SphereEditMode Mode> struct SphereEditor {
	std::array<Word, 3> center{};
	uint64_t r2 = 0;
	VBRColor color{};
	inline hd_edit_desc ToDesc() const {
		hd_edit_desc d{};
		d.kind = Mode == SphereEditMode::kFill ? HD_EDIT_SPHERE_FILL
		                                       : (Mode == SphereEditMode::kDig ? HD_EDIT_SPHERE_DIG : HD_EDIT_SPHERE_PAINT);
		d.colored = Mode != SphereEditMode::kDig && color.has_value ? 1u : 0u;
		for (int i = 0; i < 3; ++i)
			d.p0[i] = center[i];
		d.r2 = r2;
		d.color = color.ToHD();
		return d;
	}
};

// Heightfield column fill (the terrain builder; needs DeviceNodePool::SetHeightmap).
struct ColumnFillEditor {
	VBRColor deep_color = VBRColor::RGB8(0x505A64);
	uint64_t seed = 0;
	uint32_t skin = 8;
	inline hd_edit_desc ToDesc() const {
		hd_edit_desc d{};
		d.kind = HD_EDIT_COLUMN_FILL;
		d.colored = 1;
		d.r2 = seed;
		d.skin = skin;
		d.color = deep_color.ToHD();
		return d;
	}
};

template <typename T>
concept DeviceEditor = requires(const T &e) {
	{ e.ToDesc() } -> std::same_as<hd_edit_desc>;
};

// hashdag::StatelessEditor (Editor.hpp:36-40) over this facade's Config / NodeCoord / NodePointer:
// any such functor that is not a device descriptor runs on the host (hd_edit_host).
template <typename T>
concept HostStatelessEditor = !DeviceEditor<T> && requires(const T &e) {
	{ e.EditNode(std::declval<const Config &>(), NodeCoord{}, NodePointer{}) } -> std::convertible_to<EditType>;
	{ e.EditVoxel(std::declval<const Config &>(), NodeCoord{}, bool{}) } -> std::convertible_to<bool>;
};

// VBREditorWrapper::NodeState after an edit (VBREditor.hpp:28-32): octree_node is the new color
// octree root (a DAGColorPool pointer word). Stateless edits hand over std::monostate, as the
// reference's StatelessEditorWrapper does (Editor.hpp:45).
struct VBRNodeState {
	Word octree_node = HD_COLOR_NULL;
};

// hashdag::VBREditorWrapper / StatelessEditorWrapper (VBREditor.hpp:26-107, Editor.hpp:42-55): the
// reference's call sites wrap editors like this; here they select colored or stateless device
// passes. p_octree is the pool that holds the color octree (the device pool itself); octree_root is
// the color root the edit starts from.
template <typename Word_T, DeviceEditor Editor_T, typename Octree_T> struct VBREditorWrapper {
	Editor_T editor;
	Octree_T *p_octree = nullptr;
	Word_T octree_root = HD_COLOR_NULL;
	using NodeState = VBRNodeState;
	inline hd_edit_desc ToDesc() const {
		hd_edit_desc d = editor.ToDesc();
		d.colored = d.kind == HD_EDIT_SPHERE_DIG || d.kind == HD_EDIT_AABB_DIG ? 0u : 1u;
		return d;
	}
};
template <typename Word_T, typename Editor_T> struct StatelessEditorWrapper {
	Editor_T editor;
	using NodeState = std::monostate;
};
template <typename T>
concept VBRWrapped = requires { typename T::NodeState; } && std::same_as<typename T::NodeState, VBRNodeState>;

// DAGColorPool::Config (src/DAGColorPool.hpp:36-41) + capacities of the device pool.
struct ColorConfig {
	Word leaf_level;
	bool keep_history = false;
	uint64_t node_capacity = 1ull << 22, leaf_capacity = 1ull << 26;
};

inline const char *GetLastError() { return hd_last_error(); }

class DeviceNodePool {
	hd_pool *m_pool = nullptr;
	Config m_config;
	ColorConfig m_color_config;
	NodePointer m_root{};
	Word m_color_root = HD_COLOR_NULL;

	inline DeviceNodePool(hd_pool *p, Config c, ColorConfig cc)
	    : m_pool{p}, m_config{std::move(c)}, m_color_config{cc} {}

public:
	inline ~DeviceNodePool() {
		if (m_pool)
			hd_pool_destroy(m_pool);
	}
	DeviceNodePool(const DeviceNodePool &) = delete;
	DeviceNodePool &operator=(const DeviceNodePool &) = delete;

	// DAGNodePool::Create + DAGColorPool::Create: nullptr on an invalid config or a device failure
	inline static std::shared_ptr<DeviceNodePool> Create(const Config &config, const ColorConfig &color_config) {
		if (!Config::Validate(config))
			return nullptr;
		hd_config c{};
		c.word_bits_per_page = config.word_bits_per_page;
		c.page_bits_per_bucket = config.page_bits_per_bucket;
		c.node_levels = config.GetNodeLevels();
		for (Word l = 0; l < c.node_levels; ++l)
			c.bucket_bits[l] = config.bucket_bits_each_level[l];
		c.color_leaf_level = color_config.leaf_level;
		c.color_keep_history = color_config.keep_history ? 1u : 0u;
		c.color_node_capacity = color_config.node_capacity;
		c.color_leaf_capacity = color_config.leaf_capacity;
		hd_pool *p = nullptr;
		if (hd_pool_create(&c, &p) != HD_OK)
			return nullptr;
		return std::shared_ptr<DeviceNodePool>(new DeviceNodePool(p, config, color_config));
	}

	inline const Config &GetConfig() const { return m_config; }
	inline const ColorConfig &GetColorConfig() const { return m_color_config; }
	inline hd_pool *GetHandle() const { return m_pool; }
	inline NodePointer GetRoot() const { return m_root; }
	inline void SetRoot(NodePointer root) { m_root = root; }
	inline Word GetColorRoot() const { return m_color_root; }
	inline void SetColorRoot(Word color_root) { m_color_root = color_root; }

	inline bool SetHeightmap(std::span<const uint32_t> heights, uint32_t resolution) {
		if (heights.size() != (size_t)resolution * resolution)
			return false;
		return hd_pool_set_heightmap(m_pool, heights.data(), resolution) == HD_OK;
	}

	// NodePoolBase::Edit(root, editor): a new root. Colored editors run through the
	// VBREditorWrapper on the pool's color octree (its root is tracked by the pool).
	template <DeviceEditor Editor_T> inline NodePointer Edit(NodePointer root, const Editor_T &editor) {
		hd_edit_desc d = editor.ToDesc();
		return edit_descs(root, std::span<const hd_edit_desc>(&d, 1), HD_MODE_PLAIN);
	}
	// A generic stateless functor editor: NodePoolBase::Edit's recursion runs on the host over the
	// device pool's words (hd_edit_host); the result equals the reference's Edit word for word.
	template <HostStatelessEditor Editor_T> inline NodePointer Edit(NodePointer root, const Editor_T &editor) {
		struct Ctx {
			const Editor_T *ed;
			const Config *cfg;
		} ctx{&editor, &m_config};
		hd_host_editor he{};
		he.user = &ctx;
		he.edit_node = [](void *u, uint32_t level, uint32_t x, uint32_t y, uint32_t z, uint32_t ptr) -> uint32_t {
			auto *c = static_cast<Ctx *>(u);
			return (uint32_t)(EditType)c->ed->EditNode(*c->cfg, NodeCoord{level, {x, y, z}}, NodePointer{ptr});
		};
		he.edit_voxel = [](void *u, uint32_t x, uint32_t y, uint32_t z, uint32_t v) -> uint32_t {
			auto *c = static_cast<Ctx *>(u);
			return c->ed->EditVoxel(*c->cfg, NodeCoord{c->cfg->GetVoxelLevel(), {x, y, z}}, v != 0u) ? 1u : 0u;
		};
		Word r = *root;
		if (hd_edit_host(m_pool, *root, &he, &r) != HD_OK)
			return root;
		m_root = NodePointer{r};
		return m_root;
	}
	// The reference's wrapped editors (main.cpp:214-238): a VBREditorWrapper edit starts from its
	// octree_root; a StatelessEditorWrapper edit runs without colors (on the host for a functor).
	template <typename W, DeviceEditor E, typename O>
	inline NodePointer Edit(NodePointer root, const VBREditorWrapper<W, E, O> &wrapped) {
		m_color_root = (Word)wrapped.octree_root;
		hd_edit_desc d = wrapped.ToDesc();
		return edit_descs(root, std::span<const hd_edit_desc>(&d, 1), HD_MODE_PLAIN);
	}
	template <typename W, typename E> inline NodePointer Edit(NodePointer root, const StatelessEditorWrapper<W, E> &w) {
		if constexpr (DeviceEditor<E>) {
			hd_edit_desc d = w.editor.ToDesc();
			d.colored = 0;
			return edit_descs(root, std::span<const hd_edit_desc>(&d, 1), HD_MODE_PLAIN);
		} else
			return Edit(root, w.editor);
	}
	// NodePoolBase::Edit(root, editor, on_edit_done) (NodePool.hpp:405-414): returns
	// on_edit_done(new_root, state), state = VBRNodeState for colored edits, std::monostate else.
	template <typename Editor_T, typename OnDone>
	    requires std::invocable<OnDone, NodePointer, std::monostate>
	inline auto Edit(NodePointer root, const Editor_T &editor, OnDone &&on_edit_done) {
		NodePointer r = Edit(root, editor);
		if constexpr (VBRWrapped<Editor_T>)
			return on_edit_done(r, VBRNodeState{m_color_root});
		else if constexpr (DeviceEditor<Editor_T> && std::invocable<OnDone, NodePointer, VBRNodeState>) {
			if (editor.ToDesc().colored)
				return on_edit_done(r, VBRNodeState{m_color_root});
			return on_edit_done(r, std::monostate{});
		} else
			return on_edit_done(r, std::monostate{});
	}
	template <typename Editor_T, typename OnDone>
	    requires(!std::invocable<OnDone, NodePointer, std::monostate>) &&
	            std::invocable<OnDone, NodePointer, VBRNodeState>
	inline auto Edit(NodePointer root, const Editor_T &editor, OnDone &&on_edit_done) {
		NodePointer r = Edit(root, editor);
		return on_edit_done(r, VBRNodeState{m_color_root});
	}
	// NodePoolThreadedEdit::ThreadedEdit (NodePoolThreadedEdit.hpp:104-126): the device pipeline is
	// parallel and its results equal Edit's; the busy_pool and max_task_level are unused.
	template <typename Editor_T> inline NodePointer ThreadedEdit(NodePointer root, const Editor_T &editor) {
		return Edit(root, editor);
	}
	template <typename BusyPool, typename Editor_T>
	inline NodePointer ThreadedEdit(BusyPool *, NodePointer root, const Editor_T &editor, Word = Word(-1)) {
		return Edit(root, editor);
	}
	template <typename BusyPool, typename Editor_T, typename OnDone>
	inline auto ThreadedEdit(BusyPool *, NodePointer root, const Editor_T &editor, Word, OnDone &&on_edit_done) {
		return Edit(root, editor, std::forward<OnDone>(on_edit_done));
	}
	// A batch of edits in one call: HD_MODE_PLAIN applies them one by one (bit-exact with a
	// sequence of Edit calls); HD_MODE_BATCHED groups commuting edits into shared passes.
	inline NodePointer EditBatch(NodePointer root, std::span<const hd_edit_desc> edits,
	                             uint32_t mode = HD_MODE_BATCHED) {
		return edit_descs(root, edits, mode);
	}

	// NodePoolThreadedGC::ThreadedGC(pool, root / roots): compacts, remaps the roots in place.
	inline bool ThreadedGC(NodePointer &root) {
		Word r = *root;
		if (hd_gc(m_pool, &r, 1) != HD_OK)
			return false;
		root = NodePointer{r};
		return true;
	}
	// The reference's signatures (NodePoolThreadedGC.hpp:394-403): the remapped root / roots (the
	// given ones when the GC fails); the busy_pool is unused.
	template <typename BusyPool> inline NodePointer ThreadedGC(BusyPool *, NodePointer root) {
		NodePointer r = root;
		return ThreadedGC(r) ? r : root;
	}
	template <typename BusyPool>
	inline std::vector<NodePointer> ThreadedGC(BusyPool *, std::vector<NodePointer> roots) {
		std::vector<NodePointer> r = roots;
		return ThreadedGC(r) ? r : roots;
	}
	inline bool ThreadedGC(std::vector<NodePointer> &roots) {
		std::vector<Word> r(roots.size());
		for (size_t i = 0; i < r.size(); ++i)
			r[i] = *roots[i];
		if (hd_gc(m_pool, r.data(), (uint32_t)r.size()) != HD_OK)
			return false;
		for (size_t i = 0; i < r.size(); ++i)
			roots[i] = NodePointer{r[i]};
		return true;
	}

	// NodePoolTraversal::Traversal<float>(root, o, d): the hit position in [0,1]^3 or nullopt.
	inline std::optional<std::array<float, 3>> Traversal(NodePointer root, const std::array<float, 3> &o,
	                                                     const std::array<float, 3> &d) const {
		void *buf = nullptr;
		if (hd_device_alloc(sizeof(float) * 10, &buf) != HD_OK)
			return std::nullopt;
		float host[6] = {o[0], o[1], o[2], d[0], d[1], d[2]}, out[4] = {0, 0, 0, 0};
		float *dev = static_cast<float *>(buf);
		bool ok = hd_memcpy_h2d(dev, host, sizeof(host)) == HD_OK &&
		          hd_traversal(m_pool, *root, dev, dev + 3, 1, dev + 6) == HD_OK &&
		          hd_memcpy_d2h(out, dev + 6, sizeof(out)) == HD_OK;
		hd_device_free(buf);
		if (!ok || out[3] == 0.0f)
			return std::nullopt;
		return std::array<float, 3>{out[0], out[1], out[2]};
	}

	inline hd_stats GetStats() const {
		hd_stats s{};
		hd_pool_stats(m_pool, &s);
		return s;
	}

private:
	inline NodePointer edit_descs(NodePointer root, std::span<const hd_edit_desc> edits, uint32_t mode) {
		Word r = *root, c = m_color_root;
		if (hd_edit(m_pool, *root, m_color_root, edits.data(), (uint32_t)edits.size(), mode, &r, &c) != HD_OK)
			return root; // the reference's fallback: the root it was given
		m_color_root = c;
		m_root = NodePointer{r};
		return m_root;
	}
};

} // namespace hashdag_amd
