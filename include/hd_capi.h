/*
 * hd_capi.h -- C-ABI of the MI355X HashDAG engine (edit -> GC -> traversal).
 *
 * Plain C: u32 handles and node pointers, int status codes, no exceptions and
 * no torch/HIP types in the signatures. Device buffers are passed as plain
 * pointers obtained from hd_device_alloc (or any hipMalloc'd memory).
 *
 * Reference interfaces replaced (paths are relative to the VkHashDAG tree):
 *   include/hashdag/Config.hpp:15-75            -> hd_config
 *   include/hashdag/NodePointer.hpp:13-31       -> uint32_t, HD_NULL
 *   include/hashdag/Editor.hpp:18-55            -> hd_edit_desc (closed set of editors)
 *   src/main.cpp:32-150                         -> HD_EDIT_* kinds
 *   include/hashdag/VBRColor.hpp:19-44          -> hd_color
 *   include/hashdag/NodePool.hpp:405-417        -> hd_edit (mode HD_MODE_PLAIN)
 *   include/hashdag/NodePoolThreadedEdit.hpp:104-126 -> hd_edit (same semantics)
 *   include/hashdag/NodePoolThreadedGC.hpp:394-403   -> hd_gc
 *   include/hashdag/NodePoolTraversal.hpp:93-256     -> hd_traversal
 *   shader/src/trace.frag:69-402 + src/rg/TracePass.cpp:106-139 -> hd_trace
 *   src/DAGNodePool.cpp:9-85, src/DAGColorPool.cpp:9-80 -> hd_pool_create / read / write
 */
#ifndef HD_CAPI_H
#define HD_CAPI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HD_NULL 0xFFFFFFFFu            /* NodePointer::Null() (NodePointer.hpp:18) */
#define HD_COLOR_NULL 0xC0000000u      /* DAGColorPool::Pointer{} = kNull tag, data 0 */
#define HD_MAX_LEVELS 24

/* Status codes. */
enum {
	HD_OK = 0,
	HD_ERR_ARG = -1,        /* invalid argument / config (Config::Validate fails) */
	HD_ERR_DEVICE = -2,     /* a HIP call failed */
	HD_ERR_CAPACITY = -3,   /* scratch or color-pool capacity exceeded */
	HD_ERR_UNSUPPORTED = -4 /* configuration outside what the device path handles */
};

/* hashdag::Config<uint32_t> (Config.hpp:15-57) + DAGColorPool::Config (DAGColorPool.hpp:36-41). */
typedef struct hd_config {
	uint32_t word_bits_per_page;   /* >= 4 (Config.hpp:18) */
	uint32_t page_bits_per_bucket;
	uint32_t node_levels;          /* number of entries in bucket_bits */
	uint32_t bucket_bits[HD_MAX_LEVELS];
	uint32_t color_leaf_level;     /* <= node_levels - 1 */
	uint32_t color_keep_history;   /* DAGColorPool::Config::keep_history */
	uint64_t color_node_capacity;  /* max color octree nodes (8 words each), <= 2^30 */
	uint64_t color_leaf_capacity;  /* max color leaf words, <= 2^30 */
} hd_config;

/* hashdag::VBRColor (VBRColor.hpp:19-44). bits_per_weight == 0: colors = RGB8 0x00BBGGRR. */
typedef struct hd_color {
	uint32_t colors;
	uint8_t weight;
	uint8_t bits_per_weight;
	uint8_t has_value;
	uint8_t pad_;
} hd_color;

/* Editor kinds: main.cpp:32-150 plus box dig/paint and the heightfield column fill. */
enum {
	HD_EDIT_SPHERE_FILL = 0,  /* SphereEditor<kFill>  (main.cpp:73-150) */
	HD_EDIT_SPHERE_DIG = 1,   /* SphereEditor<kDig> */
	HD_EDIT_SPHERE_PAINT = 2, /* SphereEditor<kPaint> */
	HD_EDIT_AABB_FILL = 3,    /* AABBEditor (main.cpp:32-70) */
	HD_EDIT_AABB_DIG = 4,     /* box remove: kClear inside, voxel && !in_range */
	HD_EDIT_AABB_PAINT = 5,   /* box paint, the AABB analogue of SphereEditor<kPaint> */
	HD_EDIT_COLUMN_FILL = 6   /* heightfield column fill: voxel |= z < height(x, y) */
};

/*
 * One edit. colored != 0 runs the editor through VBREditorWrapper (VBREditor.hpp:26-107)
 * against the color pool; colored == 0 runs it through StatelessEditorWrapper
 * (Editor.hpp:42-55). Dig editors are always stateless.
 *   sphere: p0 = centre, r2 = squared radius (|p - c|^2 <= r2 is inside, main.cpp:127-132)
 *   box:    p0 = min, p1 = max (exclusive, main.cpp:57-59)
 *   column: heightmap of the pool (hd_heightmap_set); color = deep color,
 *           r2 = seed of the per-block color noise, skin = skin depth in voxels.
 */
typedef struct hd_edit_desc {
	uint32_t kind;
	uint32_t colored;
	uint32_t p0[3];
	uint32_t p1[3];
	uint64_t r2;
	hd_color color;
	uint32_t skin;
	uint32_t reserved;
} hd_edit_desc;

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* HD_CAPI_H */
