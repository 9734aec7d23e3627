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

	// a failing edit (column fill without a heightmap) keeps the root, as the reference's fallback
	NodePointer same = pool->Edit(root, ColumnFillEditor{});
	CHECK(same == root);

	// ThreadedGC remaps the root like hd_gc
	NodePointer g = root;
	CHECK(pool->ThreadedGC(g));
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
