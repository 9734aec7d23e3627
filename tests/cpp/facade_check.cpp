// Drives the C++ facade (include/hashdag_amd.hpp) and checks it against direct C-ABI calls on a
// second pool: the same edits give word-identical pools, GC and traversal agree, failures follow
// the reference's conventions. Exit status 0 when every check passes.
#include "hashdag_amd.hpp"

#include <cstdio>
#include <vector>

using namespace hashdag_amd;

static int failures = 0;
#define CHECK(cond)                                                                                               \
	do {                                                                                                          \
		if (!(cond)) {                                                                                            \
			std::printf("check failed: %s (line %d): %s\n", #cond, __LINE__, GetLastError());                    \
			++failures;                                                                                           \
		}                                                                                                         \
	} while (0)

static std::vector<uint32_t> live_words(hd_pool *p) {
	uint64_t n = 0;
	hd_pool_read_live_words(p, nullptr, 0, &n);
	std::vector<uint32_t> w(n ? n : 1);
	hd_pool_read_live_words(p, w.data(), w.size(), &n);
	w.resize(n);
	return w;
}

// SphereEditor<kDig> (main.cpp:73-150) written as a generic functor against the facade's Config /
// NodeCoord / NodePointer: it runs on the host through DeviceNodePool::Edit (hd_edit_host).
struct SphereDigFunctor {
	std::array<Word, 3> center;
	uint64_t r2;
	inline EditType EditNode(const Config &config, const NodeCoord &coord, NodePointer) const {
		auto lb = coord.GetLowerBoundAtLevel(config.GetVoxelLevel()), ub = coord.GetUpperBoundAtLevel(config.GetVoxelLevel());
		const int64_t ld[3] = {(int64_t)lb.x - center[0], (int64_t)lb.y - center[1], (int64_t)lb.z - center[2]};
		const int64_t ud[3] = {(int64_t)ub.x - center[0], (int64_t)ub.y - center[1], (int64_t)ub.z - center[2]};
		uint64_t mx = 0, mn = 0;
		for (int i = 0; i < 3; ++i) {
			const uint64_t a = (uint64_t)(ld[i] * ld[i]), b = (uint64_t)(ud[i] * ud[i]);
			mx += a > b ? a : b;
			mn += (ld[i] > 0 ? a : 0) + (ud[i] < 0 ? b : 0);
		}
		if (mx <= r2)
			return EditType::kClear;
		return mn > r2 ? EditType::kNotAffected : EditType::kProceed;
	}
	inline bool EditVoxel(const Config &, const NodeCoord &coord, bool voxel) const {
		const int64_t d[3] = {(int64_t)coord.pos.x - center[0], (int64_t)coord.pos.y - center[1],
		                      (int64_t)coord.pos.z - center[2]};
		return voxel && !((uint64_t)(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) <= r2);
	}
};
static_assert(HostStatelessEditor<SphereDigFunctor>);

struct FakeBusyPool {}; // stands in for the caller's lf::busy_pool: accepted, unused

int main() {
	Config config = DefaultConfig{.level_count = 7, .top_level_count = 3, .word_bits_per_page = 5,
	                              .page_bits_per_bucket = 2, .bucket_bits_per_top_level = 4,
	                              .bucket_bits_per_bottom_level = 6}();
	CHECK(config.GetNodeLevels() == 6 && config.GetResolution() == 128);
	ColorConfig cc{.leaf_level = 3};
	auto pool = DeviceNodePool::Create(config, cc);
	CHECK(pool != nullptr);
	if (!pool)
		return 1;
	Config bad = config;
	bad.word_bits_per_page = 3; // fewer than 16 words per page
	CHECK(DeviceNodePool::Create(bad, cc) == nullptr);

	// the same edits through the raw C-ABI on a second pool
	hd_config c{};
	c.word_bits_per_page = config.word_bits_per_page;
	c.page_bits_per_bucket = config.page_bits_per_bucket;
	c.node_levels = config.GetNodeLevels();
	for (Word l = 0; l < c.node_levels; ++l)
		c.bucket_bits[l] = config.bucket_bits_each_level[l];
	c.color_leaf_level = cc.leaf_level;
	c.color_node_capacity = cc.node_capacity;
	c.color_leaf_capacity = cc.leaf_capacity;
	hd_pool *raw = nullptr;
	CHECK(hd_pool_create(&c, &raw) == HD_OK);

	std::vector<hd_edit_desc> descs = {
	    SphereEditor<SphereEditMode::kFill>{{64, 64, 64}, 40 * 40, VBRColor::RGB8(0x3366CC)}.ToDesc(),
	    AABBEditor{{10, 10, 10}, {50, 30, 90}, VBRColor::RGB8(0x22AA22)}.ToDesc(),
	    SphereEditor<SphereEditMode::kDig>{{80, 70, 60}, 18 * 18, {}}.ToDesc(),
	    AABBPaintEditor{{0, 0, 0}, {128, 128, 64}, VBRColor::RGB8(0xAA2222)}.ToDesc(),
	    AABBDigEditor{{60, 60, 0}, {70, 70, 128}}.ToDesc(),
	    SphereEditor<SphereEditMode::kPaint>{{40, 40, 40}, 25 * 25, VBRColor::RGB8(0x123456)}.ToDesc(),
	};
	NodePointer root = NodePointer::Null();
	root = pool->Edit(root, SphereEditor<SphereEditMode::kFill>{{64, 64, 64}, 40 * 40, VBRColor::RGB8(0x3366CC)});
	root = pool->ThreadedEdit(root, AABBEditor{{10, 10, 10}, {50, 30, 90}, VBRColor::RGB8(0x22AA22)});
	root = pool->Edit(root, SphereEditor<SphereEditMode::kDig>{{80, 70, 60}, 18 * 18, {}});
	root = pool->Edit(root, AABBPaintEditor{{0, 0, 0}, {128, 128, 64}, VBRColor::RGB8(0xAA2222)});
	root = pool->Edit(root, AABBDigEditor{{60, 60, 0}, {70, 70, 128}});
	root = pool->Edit(root, SphereEditor<SphereEditMode::kPaint>{{40, 40, 40}, 25 * 25, VBRColor::RGB8(0x123456)});
	uint32_t rr = HD_NULL, rc = HD_COLOR_NULL;
	for (auto &d : descs)
		CHECK(hd_edit(raw, rr, rc, &d, 1, HD_MODE_PLAIN, &rr, &rc) == HD_OK);
	CHECK(root && *root == rr && pool->GetColorRoot() == rc);
	CHECK(live_words(pool->GetHandle()) == live_words(raw));

	// a generic functor on the host equals the device's descriptor edit word for word, through the
	// reference's call shapes: ThreadedEdit(busy_pool*, root, wrapper, max_task_level, on_done)
	FakeBusyPool busy;
	struct Result {
		NodePointer node_ptr;
		std::optional<Word> color;
	};
	auto on_done = [](NodePointer r, auto &&state) -> Result {
		if constexpr (requires { state.octree_node; })
			return {r, state.octree_node};
		else
			return {r, std::nullopt};
	};
	Result hr = pool->ThreadedEdit(&busy, root, StatelessEditorWrapper<Word, SphereDigFunctor>{{{70, 60, 50}, 20 * 20}},
	                               cc.leaf_level, on_done);
	hd_edit_desc dig = SphereEditor<SphereEditMode::kDig>{{70, 60, 50}, 20 * 20, {}}.ToDesc();
	CHECK(hd_edit(raw, rr, rc, &dig, 1, HD_MODE_PLAIN, &rr, &rc) == HD_OK);
	CHECK(!hr.color.has_value() && *hr.node_ptr == rr);
	CHECK(live_words(pool->GetHandle()) == live_words(raw));
	Result vr = pool->ThreadedEdit(
	    &busy, hr.node_ptr,
	    VBREditorWrapper<Word, AABBEditor, DeviceNodePool>{
	        .editor = AABBEditor{{30, 30, 30}, {70, 50, 60}, VBRColor::RGB8(0x7788AA)},
	        .p_octree = pool.get(),
	        .octree_root = pool->GetColorRoot()},
	    cc.leaf_level, on_done);
	hd_edit_desc box = AABBEditor{{30, 30, 30}, {70, 50, 60}, VBRColor::RGB8(0x7788AA)}.ToDesc();
	CHECK(hd_edit(raw, rr, rc, &box, 1, HD_MODE_PLAIN, &rr, &rc) == HD_OK);
	CHECK(vr.color.has_value() && *vr.color == rc && *vr.node_ptr == rr);
	CHECK(live_words(pool->GetHandle()) == live_words(raw));
	root = vr.node_ptr;

	// a failing edit (column fill without a heightmap) keeps the root, as the reference's fallback
	NodePointer same = pool->Edit(root, ColumnFillEditor{});
	CHECK(same == root);

	// ThreadedGC(busy_pool*, root) returns the remapped root like hd_gc
	NodePointer g = pool->ThreadedGC(&busy, root);
	uint32_t gr = rr;
	CHECK(hd_gc(raw, &gr, 1) == HD_OK);
	CHECK(*g == gr);
	CHECK(live_words(pool->GetHandle()) == live_words(raw));

	// Traversal: a ray through the filled sphere hits, one that leaves the cube at once misses
	auto hit = pool->Traversal(g, {0.05f, 0.5f, 0.5f}, {1.0f, 0.0f, 0.0f});
	CHECK(hit.has_value() && (*hit)[0] > 0.15f && (*hit)[0] < 0.6f);
	auto miss = pool->Traversal(g, {0.5f, 0.5f, 0.999f}, {0.0f, 0.0f, 1.0f});
	CHECK(!miss.has_value());
	hd_pool_destroy(raw);
	if (failures)
		return 1;
	std::printf("facade ok\n");
	return 0;
}
