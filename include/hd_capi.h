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

/* hashdag::VBRColor (VBRColor.hpp:19-44). bits_per_weight == 0: colors = RGB8 0x00BBGGRR; otherwise two
 * R5G6B5 endpoints (low, high 16 bits) and a weight. bits_per_weight is at most 3 (the chunk headers'
 * 2-bit field, VBRColor.hpp:368-487): the edit entry points reject larger values with HD_ERR_ARG. */
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


/* Pool statistics (the reference prints these to stdout / ImGui, main.cpp:275-278,379-385). */
typedef struct hd_stats {
	uint64_t total_words, total_buckets, used_words;
	uint64_t color_nodes, color_leaf_words;
	uint32_t overflow;    /* bit0: a bucket was full (upsert fell back, NodePool.hpp:195,210), bit1: color capacity */
	uint32_t node_levels;
	uint64_t level_used_words[HD_MAX_LEVELS];
	uint64_t last_entries, last_upserts, last_jobs; /* work of the last hd_edit */
	double last_ms;
	uint32_t last_groups; /* passes of the last batched hd_edit */
	uint32_t pad_;
} hd_stats;

typedef struct hd_pool hd_pool;

/* Last error message of the calling thread. */
const char *hd_last_error(void);

/* Create / destroy a device-resident pool: flat HBM node words + per-bucket word counters + color
 * octree nodes + color leaf words (replaces DAGNodePool::Create, src/DAGNodePool.cpp:9-47, and
 * DAGColorPool::Create, src/DAGColorPool.cpp:9-48). All work runs on the pool's own HIP stream. */
int hd_pool_create(const hd_config *cfg, hd_pool **out);
int hd_pool_destroy(hd_pool *pool);
int hd_pool_sync(hd_pool *pool);
void *hd_pool_stream(hd_pool *pool); /* the pool's hipStream_t */

/* Heightfield for HD_EDIT_COLUMN_FILL: resolution x resolution u32 heights, row-major, x fastest. */
int hd_pool_set_heightmap(hd_pool *pool, const uint32_t *heights, uint32_t res);

/* The benchmark terrain's heightfield generated on the device and installed like
 * hd_pool_set_heightmap (resolution = the pool's): fractal value noise, base_cells lattice cells
 * across the grid, each octave lacunarity x the frequency and gain x the amplitude, normalised to
 * [0, 1] and mapped to rint(zlo + t (zhi - zlo)) clamped to [0, resolution]. Identical heights to
 * vkhashdag_amd.workloads.value_noise_heightmap (same float64 operations and order); heights_out
 * (host, resolution^2 u32, may be null) receives them. No counterpart in the reference (synthetic
 * workload input). */
int hd_pool_value_noise_heightmap(hd_pool *pool, uint32_t octaves, uint64_t seed, uint32_t zlo, uint32_t zhi,
                                  uint32_t base_cells, double lacunarity, double gain, uint32_t *heights_out);

/* Host interop (the Flush format of src/DAGNodePool.cpp:56-85 without sparse binding):
 * word ranges of the node pool, per-bucket word counters, color arrays. */
int hd_pool_read_words(hd_pool *pool, uint64_t first, uint64_t count, uint32_t *out);
int hd_pool_write_words(hd_pool *pool, uint64_t first, uint64_t count, const uint32_t *in);
int hd_pool_read_bucket_words(hd_pool *pool, uint64_t first, uint64_t count, uint32_t *out);
int hd_pool_write_bucket_words(hd_pool *pool, uint64_t first, uint64_t count, const uint32_t *in);
int hd_pool_read_color_nodes(hd_pool *pool, uint64_t first, uint64_t count, uint32_t *out); /* 8 u32 per node */
int hd_pool_read_color_leaves(hd_pool *pool, uint64_t first, uint64_t count, uint32_t *out);
/* Used words of every bucket packed in bucket order (*n = their count; out may be null to query). */
int hd_pool_read_live_words(hd_pool *pool, uint32_t *out, uint64_t cap, uint64_t *n);
/* Sets all bucket counters and the used words of every bucket (inverse of the above). */
int hd_pool_write_live_words(hd_pool *pool, const uint32_t *bucket_words, const uint32_t *live, uint64_t n);
/* NodePoolBase::make_filled_node_pointers (NodePool.hpp:240-262); creates the chain on first use. */
int hd_pool_filled(hd_pool *pool, uint32_t *out, uint32_t cap, uint32_t *n);
int hd_pool_stats(hd_pool *pool, hd_stats *stats);
int hd_pool_clear_overflow(hd_pool *pool);

/* Edit modes. */
enum {
	HD_MODE_PLAIN = 0,  /* one level-synchronous pass per edit, edits applied in order: the result
	                       (pool words, bucket counters, color arrays, roots) equals NodePoolBase::Edit
	                       applied to each edit in turn (NodePool.hpp:405-417) */
	HD_MODE_BATCHED = 1 /* folded passes: runs of up to 4096 consecutive edits, overlapping ones
	                       included, each applied by one level-synchronous pass whose nodes carry the
	                       edits that touch them in submission order, folded from the last edit that
	                       fills or clears the node whole; a run whose pass buffers exceed the 48 GB
	                       sync-free budget is split in halves (4096 mixed edits on the 4096^3 terrain
	                       take one pass); an edit with a weighted color, a column fill, and an
	                       uncolored fill among colored edits run alone. The reachable DAG and the
	                       colors of occupied voxels equal sequential application */
};

/* Apply n edits to (root, color_root); writes the new roots. Replaces NodePoolBase::Edit and
 * NodePoolThreadedEdit::ThreadedEdit with the VBREditorWrapper / StatelessEditorWrapper
 * (main.cpp:214-234). Blocking: returns after the device work completed. */
int hd_edit(hd_pool *pool, uint32_t root, uint32_t color_root, const hd_edit_desc *edits, uint32_t n,
            uint32_t mode, uint32_t *root_out, uint32_t *color_root_out);

/* Plans hd_edit(pool, ..., edits, n, mode) and allocates the device buffers its passes need,
 * without editing: a later hd_edit on these edits, or on any call whose passes are no larger,
 * allocates nothing. No counterpart in the reference (its edits allocate host memory as they go);
 * for serving loops that must not stall on device allocation. Blocking. */
int hd_edit_reserve(hd_pool *pool, const hd_edit_desc *edits, uint32_t n, uint32_t mode);

/* Editor decisions: hashdag::EditType (Editor.hpp:18), same order. */
enum { HD_EDIT_TYPE_NOT_AFFECTED = 0, HD_EDIT_TYPE_PROCEED = 1, HD_EDIT_TYPE_FILL = 2, HD_EDIT_TYPE_CLEAR = 3 };

/* A generic stateless editor on the host: hashdag::StatelessEditor (Editor.hpp:36-40) as two
 * callbacks. edit_node(user, level, x, y, z, node_ptr) is EditNode(config, NodeCoord{level, pos},
 * node_ptr) and returns an HD_EDIT_TYPE_*; edit_voxel(user, x, y, z, voxel) is EditVoxel at the
 * voxel level and returns the voxel's new occupancy (0 / 1). */
typedef struct hd_host_editor {
	void *user;
	uint32_t (*edit_node)(void *user, uint32_t level, uint32_t x, uint32_t y, uint32_t z, uint32_t node_ptr);
	uint32_t (*edit_voxel)(void *user, uint32_t x, uint32_t y, uint32_t z, uint32_t voxel);
} hd_host_editor;

/* NodePoolBase::Edit(root, StatelessEditorWrapper{editor}) (NodePool.hpp:405-417) for an arbitrary
 * host editor: the reference's single-threaded recursion runs on the host over the device pool's
 * used words, then the words it appended and the changed bucket counters are uploaded (the dirty
 * ranges of DAGNodePool::Flush, src/DAGNodePool.cpp:56-85). Host and device edits interleave on one
 * pool word for word like a sequence of Edit calls. Colors are untouched (stateless). Blocking. */
int hd_edit_host(hd_pool *pool, uint32_t root, const hd_host_editor *editor, uint32_t *root_out);

/* NodePoolThreadedGC::ThreadedGC (NodePoolThreadedGC.hpp:394-403): keeps the DAGs below roots[0..n)
 * and the filled chain, compacts every bucket, remaps roots in place. */
int hd_gc(hd_pool *pool, uint32_t *roots, uint32_t n);

/* One camera (trace.frag push constants, trace.frag:15-23 / TracePass.cpp:10-18). World = [0,1]^3. */
typedef struct hd_view {
	float pos[3], look[3], side[3], up[3];
	float proj_factor;  /* TracePass.cpp:115-118; +inf disables the LOD cut-off (trace.frag:148) */
	uint32_t width, height;
	uint32_t flags;     /* HD_TRACE_* */
	uint64_t out_offset; /* first pixel of this view in the output buffers */
} hd_view;

enum {
	HD_TRACE_SHADOW = 1, /* one shadow ray per hit toward normalize(4,5,3) (no LOD) */
	HD_TRACE_RECORDS = 2, /* also write 48-byte per-pixel hit records */
	HD_TRACE_BEAM = 4     /* beam prepass (beam.frag): primary rays start at the 8x8-block coarse depth */
};

/* Render views on the device (DAG_RayMarch + Color_Fetch + the diffuse shading of trace.frag:374-402).
 * rgba: device u32 per pixel (RGBA8); records: device, 12 u32 per pixel when HD_TRACE_RECORDS:
 * vox[3], vox_size | hit << 31, hit_pos[3] (f32), norm[3] (f32), iterations, shadow hit. */
int hd_trace(hd_pool *pool, uint32_t root, uint32_t color_root, const hd_view *views, uint32_t n_views,
             uint32_t *rgba, uint32_t *records);

/* hd_trace without any device allocation, for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the reference.
 * The staged views, tile and beam prefixes and the beam prepass image live in one static device
 * array per GPU (4 MiB + 6 KiB, held until the process ends): at most 64 views and 2^20 beam
 * texels (HD_ERR_CAPACITY beyond). Every call on a GPU shares that array, so calls that overlap
 * on one GPU race. Same output as hd_trace; blocking. */
int hd_trace_noalloc(hd_pool *pool, uint32_t root, uint32_t color_root, const hd_view *views, uint32_t n_views,
                     uint32_t *rgba, uint32_t *records);

/* Asynchronous trace with views and the per-view tile prefix (hd_trace_tiles) already on the device;
 * no copies, no synchronisation. counters: device u64[4] (zeroed by the caller) selects the
 * instrumented variant counting [primary rays, shadow rays, march iterations, pool words read]. */
int hd_trace_async(hd_pool *pool, uint32_t root, uint32_t color_root, const hd_view *d_views,
                   const uint32_t *d_tile_start, uint32_t n_views, uint32_t n_tiles, uint32_t *rgba,
                   uint32_t *records, uint64_t *counters);
/* Beam prepass + trace, asynchronous (beam.frag:59-223, BeamPass.cpp:9-111, TracePass.cpp:35-55,
 * trace.frag:384-390): d_beam_start = hd_trace_beam_sizes prefix on the device, d_beam = n_beam
 * floats. Views with HD_TRACE_BEAM start each primary ray at 0.98 x the smallest of the 2x2 beam
 * texels around the pixel (+inf: the pixel misses) with that distance as the LOD bias. */
int hd_trace_beam_async(hd_pool *pool, uint32_t root, uint32_t color_root, const hd_view *d_views,
                        const uint32_t *d_tile_start, const uint32_t *d_beam_start, uint32_t n_views,
                        uint32_t n_tiles, uint32_t n_beam, float *d_beam, uint32_t *rgba);
/* Host helper: beam_start[0..n_views] prefix of beam texels (ceil(W/8) x ceil(H/8) per beam view). */
int hd_trace_beam_sizes(const hd_view *views, uint32_t n_views, uint32_t *beam_start);
/* Host helper: tile_start[0..n_views] prefix of pixel tiles per view; the tile edge is the current
 * hd_tuning.trace_tile (8 or 16), so compute the prefix after any hd_tuning_set and before the
 * hd_trace_async / hd_trace_beam_async call it is passed to. */
int hd_trace_tiles(const hd_view *views, uint32_t n_views, uint32_t *tile_start);
/* HIP-event timing of the work queued on the pool's stream between the two calls. */
int hd_timer_begin(hd_pool *pool);
int hd_timer_end(hd_pool *pool, float *ms);

/* NodePoolTraversal::Traversal<float> (NodePoolTraversal.hpp:93-256) for n rays given as device
 * arrays o[3n], d[3n]; out: device float[4n] = hit position in [0,1]^3 + hit flag. */
int hd_traversal(hd_pool *pool, uint32_t root, const float *o, const float *d, uint64_t n, float *out);

/* ---- frames captured as hipGraphs: up to max_passes passes of pairwise-disjoint edits (max_edits
 * each, every edit within max_radius voxels: sphere radius or box half-extent) followed by one
 * multi-view trace (n_views views up to max_width x max_height, pixels packed view after view) of
 * the roots the passes leave on the device. The sequence is captured once at creation; each frame
 * sets its edits and views (hd_frame_set) and replays the graph (hd_frame_launch). Edits are grouped
 * into passes of pairwise-disjoint edits, in order, so a frame's reachable DAG and the colors of its
 * occupied voxels equal applying the edits one by one (HD_MODE_BATCHED's bar), and one edit per
 * frame equals HD_MODE_PLAIN word for word. With HD_FRAME_GC the GC runs inside the graph
 * (the frame's words then equal hd_edit batched + hd_gc); without it, after an hd_gc hand the
 * remapped roots back with hd_frame_set_roots. */
enum {
	HD_FRAME_COLORS = 1, /* colored (VBR) edits allowed: the color passes are captured */
	HD_FRAME_GC = 2      /* a node-pool GC (NodePoolThreadedGC::ThreadedGC of the frame's root and the
	                        filled chain, NodePoolThreadedGC.hpp:394-403) is captured between the passes
	                        and the trace: each frame is edit -> GC -> render; the root the frame
	                        returns is the compacted one and older roots become invalid */
};

typedef struct hd_frame_config {
	uint32_t max_edits, max_passes, max_radius;
	uint32_t n_views, max_width, max_height;
	uint32_t flags; /* HD_FRAME_* */
	uint32_t pad_;
} hd_frame_config;

typedef struct hd_frame hd_frame;
int hd_frame_create(hd_pool *pool, const hd_frame_config *cfg, uint32_t root, uint32_t color_root,
                    hd_frame **out);
/* Edits of the next frame (grouped into passes; errors when they need more passes, more edits per
 * pass or a larger radius than the frame was built for) and its views (out_offset is assigned). */
int hd_frame_set(hd_frame *frame, const hd_edit_desc *edits, uint32_t n, const hd_view *views,
                 uint32_t n_views);
int hd_frame_set_roots(hd_frame *frame, uint32_t root, uint32_t color_root);
int hd_frame_launch(hd_frame *frame); /* asynchronous, on the pool's stream */
/* Waits for the launched frames; writes the roots after the last one. */
int hd_frame_wait(hd_frame *frame, uint32_t *root_out, uint32_t *color_root_out);
uint32_t *hd_frame_rgba(hd_frame *frame); /* device RGBA8 pixels of the last frame */
int hd_frame_destroy(hd_frame *frame);

/* ---- serving: a request loop over one device-resident DAG (replaces the reference app's loop of
 * one edit in flight with later edits dropped, src/main.cpp:214-296 and :287). Requests, sorted by
 * arrival on the server clock (ms from the call), are admitted per tick: edits into
 * max_edits_per_tick slots while the node pool has at least min_free_pages free pages (else a GC
 * runs; edits wait if it frees too little), renders up to max_pixels_per_tick pixels. A tick's
 * edits run as one batched hd_edit while its renders run as one multi-view trace on a second stream
 * against the DAG version of the tick's start; the next tick admits what arrived meanwhile.
 * Concurrent rendering needs append-only colors: the pool must keep color history.
 * HD_SERVE_PLAIN instead serves one request at a time in arrival order (the baseline loop). */
enum { HD_REQ_EDIT = 0, HD_REQ_RENDER = 1 };
enum { HD_SERVE_PLAIN = 1 };

typedef struct hd_request {
	uint32_t kind;     /* HD_REQ_* */
	uint32_t client;
	double arrival_ms; /* server clock */
	hd_edit_desc edit; /* HD_REQ_EDIT */
	hd_view view;      /* HD_REQ_RENDER (out_offset is assigned by the server) */
} hd_request;

typedef struct hd_request_result {
	double admit_ms, done_ms; /* server clock */
	uint64_t pixel_offset;    /* renders: first pixel of the image in rgba */
	uint32_t tick;
	uint32_t root, color_root; /* renders: the version rendered; edits: the roots after their tick */
	uint32_t pad_;
} hd_request_result;

typedef struct hd_serve_config {
	uint32_t max_edits_per_tick;
	uint32_t flags; /* HD_SERVE_* */
	uint64_t max_pixels_per_tick;
	uint64_t min_free_pages;
} hd_serve_config;

typedef struct hd_serve_stats {
	uint32_t ticks, edits, renders, gcs, gated_ticks, max_edits_in_tick, max_renders_in_tick, pad_;
	uint64_t min_free_pages; /* fewest free pages seen at an edit admission */
	double wall_ms;
} hd_serve_stats;

/* Serves n requests. rgba: device u32 buffer of rgba_pixels; when it holds every render's pixels each
 * image stays at its pixel_offset, else images go to a tick scratch buffer and are not kept. The
 * roots after the last tick are written to root_out / color_root_out. */
int hd_serve_run(hd_pool *pool, const hd_serve_config *cfg, uint32_t root, uint32_t color_root,
                 const hd_request *reqs, uint32_t n, uint32_t *rgba, uint64_t rgba_pixels,
                 hd_request_result *results, uint32_t *root_out, uint32_t *color_root_out,
                 hd_serve_stats *stats);
/* Free node-pool pages (pages of 2^word_bits_per_page words; a bucket holds 2^page_bits_per_bucket). */
int hd_pool_free_pages(hd_pool *pool, uint64_t *free_pages, uint64_t *total_pages);

/* Checkpoint dump / load: config, per-bucket counters + used words, color octree nodes and leaf
 * words with their counters, overflow flags and n_roots caller roots in one file. hd_pool_load
 * creates a new pool from it (writes the roots back, *n_roots <= cap); edits, GC and traversal
 * continue from the loaded state word for word. The heightfield (an editor input) is not saved. */
int hd_pool_save(hd_pool *pool, const char *path, const uint32_t *roots, uint32_t n_roots);
int hd_pool_load(const char *path, hd_pool **out, uint32_t *roots, uint32_t cap, uint32_t *n_roots);
/* The config a pool was created with (a loaded pool's comes from its file). */
int hd_pool_get_config(hd_pool *pool, hd_config *out);

/* Launch parameters. trace_tile, top_pairs and small_level change speed only (the same words,
 * roots and pixels); batch_pass_edits also renumbers the nodes a batched edit appends (the DAG up
 * to renumbering, the voxels and the colors stay the same -- the batched path's own bar).
 * Process-wide; a call made after hd_tuning_set uses the new values (a frame keeps the tile it was
 * created with). The autotuner (vkhashdag_amd/autotune.py) measures the candidates on the device
 * and stores the best set. No counterpart in the reference (its parameters are compile-time). */
typedef struct hd_tuning {
	uint32_t trace_tile;       /* pixel tile edge of the frame kernel: 8 or 16 (default 16) */
	uint32_t top_pairs;        /* the one-workgroup top-down kernel takes levels of at most this many
	                            * (entry, slot) pairs: [64, 4096] (default 4096) */
	uint32_t small_level;      /* the one-workgroup bottom-up kernel takes levels of at most this many
	                            * entries: [64, 1024] (default 1024) */
	uint32_t batch_pass_edits; /* most edits one batched pass applies; a larger group of pairwise
	                            * disjoint edits is split into consecutive passes; 0 = no limit */
} hd_tuning;
int hd_tuning_get(hd_tuning *out);
int hd_tuning_set(const hd_tuning *t);

/* Device memory helpers for hosts without a GPU runtime of their own. */
int hd_device_alloc(uint64_t bytes, void **out);
int hd_device_free(void *ptr);
int hd_memcpy_h2d(void *dst, const void *src, uint64_t bytes);
int hd_memcpy_d2h(void *dst, const void *src, uint64_t bytes);
int hd_device_count(int *n);
/* "name (arch, N CUs)" of device dev, NUL-terminated within cap bytes. */
int hd_device_name(int dev, char *buf, uint32_t cap);
/* Device-to-device copy by a plain 16-byte-vector kernel on the pool's stream (pool may be null:
 * the null stream); bytes and both pointers 16-byte aligned. Its HBM traffic is exactly bytes read
 * and bytes written, which calibrates hardware byte counters. Blocking. */
int hd_device_copy(void *dst, const void *src, uint64_t bytes, hd_pool *pool);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* HD_CAPI_H */
