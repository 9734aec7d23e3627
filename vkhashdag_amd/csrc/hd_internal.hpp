// hd_internal.hpp -- host-side state of a device pool and helpers shared by the HIP sources.
#pragma once

#include <cstring> // rocPRIM's texture_cache_iterator uses memset
#include <hip/hip_runtime.h>

#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "hd_device.hpp"

namespace hd {

void set_error(const char *fmt, ...);
// VBR colors carry at most 3 bits per weight (the chunk headers' 2-bit field, VBRColor.hpp:368-487).
inline bool edit_color_ok(const hd_edit_desc &d) { return !d.color.has_value || d.color.bits_per_weight <= 3; }

#define HD_TRY(expr)                                                                                              \
	do {                                                                                                          \
		hipError_t e_ = (expr);                                                                                   \
		if (e_ != hipSuccess) {                                                                                   \
			::hd::set_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(e_));                \
			return HD_ERR_DEVICE;                                                                                 \
		}                                                                                                         \
	} while (0)
#define HD_RET(expr)                                                                                              \
	do {                                                                                                          \
		int r_ = (expr);                                                                                          \
		if (r_ != HD_OK)                                                                                          \
			return r_;                                                                                            \
	} while (0)

// Growable device buffer.
struct DevBuf {
	void *p = nullptr;
	size_t bytes = 0;
	int ensure(size_t n) {
		if (n <= bytes && p)
			return HD_OK;
		if (p)
			HD_TRY(hipFree(p));
		p = nullptr;
		size_t want = n < 256 ? 256 : n + n / 4;
		HD_TRY(hipMalloc(&p, want));
		bytes = want;
		return HD_OK;
	}
	template <typename T> T *as() const { return static_cast<T *>(p); }
	void release() {
		if (p)
			(void)hipFree(p);
		p = nullptr;
		bytes = 0;
	}
};

// Named scratch buffers, grown on demand and kept for the pool's lifetime.
struct Scratch {
	std::map<std::string, DevBuf> bufs;
	template <typename T> int get(const std::string &name, size_t count, T **out) {
		HD_RET(bufs[name].ensure(count * sizeof(T)));
		*out = bufs[name].as<T>();
		return HD_OK;
	}
	void release() {
		for (auto &kv : bufs)
			kv.second.release();
		bufs.clear();
	}
};

// Device-side counters of a pool (one small array so a single copy reads them all).
enum Meta : u32 {
	kMetaColorNodes = 0,  // color octree node count (SafePagedVector count of DAGColorPool::m_nodes)
	kMetaColorLeaves = 1, // color leaf word count
	kMetaOverflow = 2,    // bit 0: a bucket was full (upsert fell back), bit 1: color capacity, bit 2: GC bucket
	kMetaCount = 8
};

// Zero words allocated past the pool so a node can be read whole (up to 9 words) without a
// bounds test, wherever in the pool it starts.
constexpr u64 kWordsPad = 16;

struct Pool;
// Device buffers of a node-pool GC (hd_gc.hip), sized by the pool's configuration: per level a
// bitmap of live node starts over the level's words and its popcount prefix (node ranks); new
// pointers by rank for two adjacent levels; one level's staged words, new-bucket keys, bucket
// lists, counts and offsets.
struct GcPlan {
	struct Level {
		u32 *bits = nullptr, *rank = nullptr;
		u64 words = 0, base = 0, nw = 0, buckets = 0;
	};
	std::vector<Level> lv;
	u32 *newp[2] = {nullptr, nullptr};
	u32 *stage = nullptr, *key = nullptr, *list = nullptr, *cnt = nullptr, *bstart = nullptr, *fill = nullptr,
	    *bsum = nullptr;
	int prepare(Pool &p, const char *prefix);
};

struct Pool {
	hd_config cfg{};
	DevConfig dc{};
	u64 total_words = 0, total_buckets = 0;
	u32 *words = nullptr;        // total_words u32: the node pool (pages of every bucket, flat)
	u32 *bucket_words = nullptr; // total_buckets u32: used words per bucket (incl. page padding)
	u32 *meta = nullptr;         // kMetaCount u32
	std::vector<u32> filled;     // filled-node chain (NodePool.hpp:240-262), host copy
	u32 *d_filled = nullptr;
	u32 *cnodes = nullptr, *cleaves = nullptr; // color octree nodes (8 u32 each) and VBR leaf words
	u64 cnode_cap = 0, cleaf_cap = 0;
	u32 *hm_mn = nullptr, *hm_mx = nullptr;
	DevHeightmap hm{};
	DevHeightmap *d_hm = nullptr; // device copy of hm (editors read it through a pointer)
	hipStream_t stream = nullptr;
	// color side stream of a sync-free pass: the leaf encode and SetNode run there, beside the node
	// levels (joined before the next pass)
	hipStream_t cstream = nullptr;
	hipEvent_t ev_top = nullptr, ev_leaf = nullptr, ev_color = nullptr;
	Scratch scratch;
	// statistics of the last edit (host)
	u64 last_entries = 0, last_upserts = 0, last_jobs = 0;
	u32 last_groups = 0;
	double last_ms = 0;
	u32 edit_gen = 0; // generation of the sync-free pass's group table (hd_edit.hip)
	GcPlan gc_plan;   // hd_gc's buffers
};

// Launch parameters (hd_tuning_set; process-wide): speed only, except that batch_pass_edits also
// renumbers the nodes of batched edits.
struct Tuning {
	u32 trace_tile = 16;      // frame kernel pixel tile edge: 8 or 16
	u32 top_pairs = 4096;     // single-workgroup top-down levels: at most this many (entry, slot) pairs
	u32 small_level = 1024;   // single-workgroup bottom-up levels: at most this many entries
	u32 batch_pass_edits = 0; // most edits per batched pass (a larger scheduling group is split); 0: no limit
};
extern Tuning g_tuning; // hd_pool.hip

int make_filled(Pool &p);               // hd_pool.hip
void launch_make_filled(Pool &p, hipStream_t s); // hd_pool.hip: the chain's lookup into d_filled, queued only
// A whole GC queued on stream s over device roots (remapped in place); the filled chain is a root
// when has_filled and is found again into d_filled. No host synchronisation (hd_gc.hip).
int gc_queue(Pool &p, const GcPlan &G, hipStream_t s, u32 *d_roots, u32 n_roots, bool has_filled);
// Multi-view trace of (root, color_root) queued on stream s; views and tile prefix on the device.
// d_roots (optional): read the roots from device memory instead (graph replays).
int trace_on_stream(Pool &p, hipStream_t s, u32 root, u32 color_root, const hd_view *d_views, const u32 *d_tiles,
                    u32 n_views, u32 n_tiles, u32 *rgba, const u32 *d_roots = nullptr); // hd_trace.hip
int read_meta(Pool &p, u32 *host_meta); // hd_pool.hip

inline u32 div_up(u64 a, u64 b) { return (u32)((a + b - 1) / b); }

} // namespace hd

struct hd_pool {
	hd::Pool p;
};
