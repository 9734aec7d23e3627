// hd_pool.hip -- device-resident node pool + color pool: creation, host interop, heightmap,
// filled-node chain. Replaces the page store of src/DAGNodePool.hpp:41-99 (+ .cpp:9-85) and the
// paged color storage of src/DAGColorPool.hpp:51-64 with flat HBM arrays.
#include "hd_internal.hpp"

#include <algorithm>

#include <cstdarg>
#include <mutex>
#include <stdexcept>


namespace hd {

static thread_local char g_err[512] = "";
void set_error(const char *fmt, ...) {
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(g_err, sizeof(g_err), fmt, ap);
	va_end(ap);
}

// NodePool::make_filled_node_pointers (NodePool.hpp:240-262), one thread: the all-solid leaf,
// then the all-solid inner node of every level, each found or appended in its bucket.
__global__ void k_make_filled(u32 *words, u32 *bucket_words, DevConfig c, u32 *out, u32 *meta) {
	if (threadIdx.x | blockIdx.x)
		return;
	u32 node[9];
	u32 prev = kNull;
	for (int l = (int)c.L - 1; l >= 0; --l) {
		bool leaf = l == (int)c.L - 1;
		u32 n = leaf ? 2u : 9u;
		if (leaf)
			node[0] = node[1] = ~0u;
		else {
			node[0] = 0xFFu;
			for (int i = 1; i < 9; ++i)
				node[i] = prev;
		}
		u32 b = bucket_of(c, (u32)l, node, n, leaf);
		u32 bw = bucket_words[b];
		u32 p = find_in_bucket(words, c, b, bw, node, n, leaf);
		if (p == kNull) {
			p = append_to_bucket(words, c, b, bw, node, n);
			if (p == kNull)
				meta[kMetaOverflow] |= 1u;
			else
				bucket_words[b] = bw;
		}
		out[l] = p;
		prev = p;
	}
}

void launch_make_filled(Pool &p, hipStream_t s) {
	k_make_filled<<<1, 64, 0, s>>>(p.words, p.bucket_words, p.dc, p.d_filled, p.meta);
}

int make_filled(Pool &p) {
	if (!p.filled.empty())
		return HD_OK;
	k_make_filled<<<1, 64, 0, p.stream>>>(p.words, p.bucket_words, p.dc, p.d_filled, p.meta);
	HD_TRY(hipGetLastError());
	p.filled.resize(p.dc.L);
	HD_TRY(hipMemcpyAsync(p.filled.data(), p.d_filled, p.dc.L * 4, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

int read_meta(Pool &p, u32 *host_meta) {
	HD_TRY(hipMemcpyAsync(host_meta, p.meta, kMetaCount * 4, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

// Heightfield min/max pyramid, one level per launch.
__global__ void k_pyramid(const u32 *src_mn, const u32 *src_mx, u32 *dst_mn, u32 *dst_mx, u32 side) {
	u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (u64)side * side)
		return;
	u32 x = (u32)(i % side), y = (u32)(i / side), ps = side * 2;
	u64 a = (u64)(2 * y) * ps + 2 * x, b = a + ps;
	dst_mn[i] = min(min(src_mn[a], src_mn[a + 1]), min(src_mn[b], src_mn[b + 1]));
	dst_mx[i] = max(max(src_mx[a], src_mx[a + 1]), max(src_mx[b], src_mx[b + 1]));
}

// Used words of every bucket, packed in bucket order (off = exclusive prefix of bucket_words).
__global__ void k_gather_live(const u32 *__restrict__ words, const u32 *__restrict__ bucket_words,
                              const u64 *__restrict__ off, u64 nb, u32 wbpb, u32 *__restrict__ out) {
	u64 b = blockIdx.x;
	if (b >= nb)
		return;
	u32 n = bucket_words[b];
	const u32 *src = words + (b << wbpb);
	u32 *dst = out + off[b];
	for (u32 i = threadIdx.x; i < n; i += blockDim.x)
		dst[i] = src[i];
}
__global__ void k_scatter_live(u32 *__restrict__ words, const u32 *__restrict__ bucket_words,
                               const u64 *__restrict__ off, u64 nb, u32 wbpb, const u32 *__restrict__ in) {
	u64 b = blockIdx.x;
	if (b >= nb)
		return;
	u32 n = bucket_words[b];
	u32 *dst = words + (b << wbpb);
	const u32 *src = in + off[b];
	for (u32 i = threadIdx.x; i < n; i += blockDim.x)
		dst[i] = src[i];
}

static int live_offsets(Pool &p, u64 **d_off, u64 *total) {
	std::vector<u32> bw(p.total_buckets);
	HD_TRY(hipMemcpyAsync(bw.data(), p.bucket_words, p.total_buckets * 4, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	std::vector<u64> off(p.total_buckets + 1, 0);
	for (u64 b = 0; b < p.total_buckets; ++b)
		off[b + 1] = off[b] + bw[b];
	*total = off[p.total_buckets];
	HD_RET(p.scratch.get<u64>("live_off", p.total_buckets + 1, d_off));
	HD_TRY(hipMemcpyAsync(*d_off, off.data(), off.size() * 8, hipMemcpyHostToDevice, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

} // namespace hd

using namespace hd;

extern "C" {

const char *hd_last_error(void) { return g_err; }

int hd_pool_create(const hd_config *cfg, hd_pool **out) {
	if (!cfg || !out || cfg->node_levels == 0 || cfg->node_levels > HD_MAX_LEVELS - 1 || cfg->word_bits_per_page < 4 ||
	    cfg->color_leaf_level > cfg->node_levels - 1 || cfg->node_levels + 1 > 22) {
		set_error("invalid hd_config");
		return HD_ERR_ARG;
	}
	u64 buckets = 0;
	for (u32 l = 0; l < cfg->node_levels; ++l) {
		if (cfg->bucket_bits[l] > 30) {
			set_error("bucket_bits too large");
			return HD_ERR_ARG;
		}
		buckets += 1ull << cfg->bucket_bits[l];
	}
	u64 words = buckets << (cfg->word_bits_per_page + cfg->page_bits_per_bucket);
	if (words - 1 > 0xFFFFFFFEull) { // Config::Validate (Config.hpp:48-56)
		set_error("pool exceeds 2^32-1 words");
		return HD_ERR_ARG;
	}
	if (cfg->color_node_capacity >= (1ull << 30) - 1 || cfg->color_leaf_capacity > (1ull << 30)) {
		set_error("color capacities exceed 30-bit color pointers (DAGColorPool.hpp:24)");
		return HD_ERR_ARG;
	}
	auto *h = new hd_pool();
	Pool &p = h->p;
	p.cfg = *cfg;
	p.dc.wbpp = cfg->word_bits_per_page;
	p.dc.pbpb = cfg->page_bits_per_bucket;
	p.dc.L = cfg->node_levels;
	p.dc.color_leaf = cfg->color_leaf_level;
	p.dc.voxel_level = cfg->node_levels + 1;
	p.dc.keep_history = cfg->color_keep_history;
	u32 base = 0;
	for (u32 l = 0; l < cfg->node_levels; ++l) {
		p.dc.bucket_bits[l] = cfg->bucket_bits[l];
		p.dc.level_base[l] = base;
		base += 1u << cfg->bucket_bits[l];
	}
	p.total_buckets = buckets;
	p.total_words = words;
	p.cnode_cap = cfg->color_node_capacity;
	p.cleaf_cap = cfg->color_leaf_capacity;
	auto fail = [&](hipError_t e) {
		set_error("hd_pool_create: %s", hipGetErrorString(e));
		hd_pool_destroy(h);
		return HD_ERR_DEVICE;
	};
	hipError_t e;
	if ((e = hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking)) != hipSuccess)
		return fail(e);
	if ((e = hipStreamCreateWithFlags(&p.cstream, hipStreamNonBlocking)) != hipSuccess)
		return fail(e);
	if ((e = hipEventCreateWithFlags(&p.ev_top, hipEventDisableTiming)) != hipSuccess)
		return fail(e);
	if ((e = hipEventCreateWithFlags(&p.ev_leaf, hipEventDisableTiming)) != hipSuccess)
		return fail(e);
	if ((e = hipEventCreateWithFlags(&p.ev_color, hipEventDisableTiming)) != hipSuccess)
		return fail(e);
	if ((e = hipMalloc(&p.words, (words + kWordsPad) * 4)) != hipSuccess)
		return fail(e);
	if ((e = hipMalloc(&p.bucket_words, buckets * 4)) != hipSuccess)
		return fail(e);
	if ((e = hipMalloc(&p.meta, kMetaCount * 4)) != hipSuccess)
		return fail(e);
	if ((e = hipMalloc(&p.d_filled, HD_MAX_LEVELS * 4)) != hipSuccess)
		return fail(e);
	if ((e = hipMalloc(&p.d_hm, sizeof(DevHeightmap))) != hipSuccess)
		return fail(e);
	if ((e = hipMemsetAsync(p.d_hm, 0, sizeof(DevHeightmap), p.stream)) != hipSuccess)
		return fail(e);
	if ((e = hipMalloc(&p.cnodes, (p.cnode_cap ? p.cnode_cap : 1) * 32)) != hipSuccess)
		return fail(e);
	if ((e = hipMalloc(&p.cleaves, (p.cleaf_cap ? p.cleaf_cap : 1) * 4)) != hipSuccess)
		return fail(e);
	if ((e = hipMemsetAsync(p.words, 0, (words + kWordsPad) * 4, p.stream)) != hipSuccess)
		return fail(e);
	if ((e = hipMemsetAsync(p.bucket_words, 0, buckets * 4, p.stream)) != hipSuccess)
		return fail(e);
	if ((e = hipMemsetAsync(p.meta, 0, kMetaCount * 4, p.stream)) != hipSuccess)
		return fail(e);
	if ((e = hipStreamSynchronize(p.stream)) != hipSuccess)
		return fail(e);
	*out = h;
	return HD_OK;
}

int hd_pool_destroy(hd_pool *h) {
	if (!h)
		return HD_OK;
	Pool &p = h->p;
	if (p.stream)
		(void)hipStreamSynchronize(p.stream);
	if (p.cstream)
		(void)hipStreamSynchronize(p.cstream);
	p.scratch.release();
	for (void *ptr : {(void *)p.words, (void *)p.bucket_words, (void *)p.meta, (void *)p.d_filled, (void *)p.cnodes,
	                  (void *)p.cleaves, (void *)p.hm_mn, (void *)p.hm_mx, (void *)p.d_hm})
		if (ptr)
			(void)hipFree(ptr);
	if (p.stream)
		(void)hipStreamDestroy(p.stream);
	if (p.cstream)
		(void)hipStreamDestroy(p.cstream);
	for (hipEvent_t ev : {p.ev_top, p.ev_leaf, p.ev_color})
		if (ev)
			(void)hipEventDestroy(ev);
	delete h;
	return HD_OK;
}

int hd_pool_sync(hd_pool *h) {
	HD_TRY(hipStreamSynchronize(h->p.stream));
	return HD_OK;
}

void *hd_pool_stream(hd_pool *h) { return h ? (void *)h->p.stream : nullptr; }

// Installs a res x res heightmap as the column-fill editor's input: the heights either come from
// the host (heights) or are already in the pool's base level (device; heights == null), then the
// min / max pyramid is built above them.
static int install_heightmap(Pool &p, const uint32_t *heights, uint32_t res, bool generated);

int hd_pool_set_heightmap(hd_pool *h, const uint32_t *heights, uint32_t res) {
	Pool &p = h->p;
	if (!heights || res != (1u << p.dc.voxel_level)) {
		set_error("heightmap must be resolution x resolution (%u)", 1u << p.dc.voxel_level);
		return HD_ERR_ARG;
	}
	return install_heightmap(p, heights, res, false);
}

// ---- seeded value-noise heightfield on the device: the recipe of the benchmark's terrain, the same
// float64 operations in the same order as the host generator (vkhashdag_amd/workloads.py
// value_noise_heightmap; no contraction), so the heights are identical.
namespace {
__device__ inline double lattice(u64 ix, u64 iy, u64 c) { // splitmix64 finaliser of a lattice point
	u64 k = (ix * 0x9E3779B97F4A7C15ull) ^ (iy * 0xC2B2AE3D27D4EB4Full) ^ c;
	k ^= k >> 30;
	k *= 0xBF58476D1CE4E5B9ull;
	k ^= k >> 27;
	k *= 0x94D049BB133111EBull;
	k ^= k >> 31;
	return (double)(k >> 11) * (1.0 / 9007199254740992.0);
}
struct NoiseArgs {
	u32 res, octaves;
	double base, lacunarity, gain, norm;
	u64 consts[32];
};
__global__ void k_noise(NoiseArgs a, double *__restrict__ acc, unsigned long long *__restrict__ minmax) {
	const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
	const u64 n = (u64)a.res * a.res;
	double v = 0.0;
	if (i < n) {
		const u32 x = (u32)(i % a.res), y = (u32)(i / a.res);
		const double xs = ((double)x + 0.5) / (double)a.res, ys = ((double)y + 0.5) / (double)a.res;
		double amp = 1.0, freq = a.base;
		for (u32 o = 0; o < a.octaves; ++o) {
			const double fx = xs * freq, fy = ys * freq;
			const double ix = floor(fx), iy = floor(fy);
			double tx = fx - ix, ty = fy - iy;
			tx = tx * tx * (3.0 - 2.0 * tx);
			ty = ty * ty * (3.0 - 2.0 * ty);
			const u64 x0 = (u64)ix, y0 = (u64)iy, c = a.consts[o];
			const double top = lattice(x0, y0, c) * (1.0 - tx) + lattice(x0 + 1, y0, c) * tx;
			const double bot = lattice(x0, y0 + 1, c) * (1.0 - tx) + lattice(x0 + 1, y0 + 1, c) * tx;
			v = v + amp * (top * (1.0 - ty) + bot * ty);
			amp *= a.gain;
			freq *= a.lacunarity;
		}
		v = v / a.norm;
		acc[i] = v;
	}
	// non-negative doubles order like their bit patterns: the wave's min / max, then one atomic each
	unsigned long long lo = i < n ? (unsigned long long)__double_as_longlong(v) : ~0ull;
	unsigned long long hi = i < n ? (unsigned long long)__double_as_longlong(v) : 0ull;
	for (u32 d = 32; d >= 1; d >>= 1) {
		lo = min(lo, (unsigned long long)__shfl_xor((long long)lo, d));
		hi = max(hi, (unsigned long long)__shfl_xor((long long)hi, d));
	}
	if ((threadIdx.x & 63u) == 0) {
		atomicMin(&minmax[0], lo);
		atomicMax(&minmax[1], hi);
	}
}
__global__ void k_noise_heights(u32 res, u32 zlo, u32 zhi, const double *__restrict__ acc,
                                const unsigned long long *__restrict__ minmax, u32 *__restrict__ out_mn,
                                u32 *__restrict__ out_mx) {
	const u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
	if (i >= (u64)res * res)
		return;
	const double lo = __longlong_as_double((long long)minmax[0]), hi = __longlong_as_double((long long)minmax[1]);
	const double span = hi - lo > 1e-12 ? hi - lo : 1e-12;
	const double t = (acc[i] - lo) / span;
	double z = rint((double)zlo + t * (double)((i64)zhi - (i64)zlo));
	z = z < 0.0 ? 0.0 : (z > (double)res ? (double)res : z);
	out_mn[i] = out_mx[i] = (u32)z;
}
} // namespace

extern "C" int hd_pool_value_noise_heightmap(hd_pool *h, uint32_t octaves, uint64_t seed, uint32_t zlo, uint32_t zhi,
                                             uint32_t base_cells, double lacunarity, double gain,
                                             uint32_t *heights_out) {
	if (!h || !octaves || octaves > 32 || !base_cells || !(lacunarity > 0.0) || !(gain > 0.0))
		return set_error("hd_pool_value_noise_heightmap: invalid argument"), HD_ERR_ARG;
	Pool &p = h->p;
	const u32 res = 1u << p.dc.voxel_level;
	const u64 n = (u64)res * res;
	NoiseArgs a{};
	a.res = res;
	a.octaves = octaves;
	a.base = (double)base_cells;
	a.lacunarity = lacunarity;
	a.gain = gain;
	double amp = 1.0, norm = 0.0;
	for (u32 o = 0; o < octaves; ++o) {
		a.consts[o] = seed * 0x165667B19E3779F9ull + (u64)o * 0x27D4EB2F165667C5ull;
		norm += amp;
		amp *= gain;
	}
	a.norm = norm;
	double *acc;
	unsigned long long *mm;
	HD_RET(p.scratch.get<double>("noise_acc", n, &acc));
	HD_RET(p.scratch.get<unsigned long long>("noise_minmax", 2, &mm));
	const unsigned long long init[2] = {~0ull, 0ull};
	HD_TRY(hipMemcpyAsync(mm, init, sizeof(init), hipMemcpyHostToDevice, p.stream));
	// the base level of the pyramid is written in place, then install_heightmap builds the rest
	u64 total = 0;
	for (u32 s = res;; s >>= 1) {
		total += (u64)s * s;
		if (s == 1)
			break;
	}
	if (p.hm_mn)
		HD_TRY(hipFree(p.hm_mn));
	if (p.hm_mx)
		HD_TRY(hipFree(p.hm_mx));
	p.hm_mn = p.hm_mx = nullptr;
	HD_TRY(hipMalloc(&p.hm_mn, total * 4));
	HD_TRY(hipMalloc(&p.hm_mx, total * 4));
	k_noise<<<div_up(n, 256), 256, 0, p.stream>>>(a, acc, mm);
	HD_TRY(hipGetLastError());
	k_noise_heights<<<div_up(n, 256), 256, 0, p.stream>>>(res, zlo, zhi, acc, mm, p.hm_mn, p.hm_mx);
	HD_TRY(hipGetLastError());
	HD_RET(install_heightmap(p, nullptr, res, true));
	if (heights_out) {
		HD_TRY(hipMemcpyAsync(heights_out, p.hm_mn, n * 4, hipMemcpyDeviceToHost, p.stream));
		HD_TRY(hipStreamSynchronize(p.stream));
	}
	return HD_OK;
}

static int install_heightmap(Pool &p, const uint32_t *heights, uint32_t res, bool generated) {
	u64 total = 0;
	DevHeightmap hm;
	hm.res = res;
	u32 k = 0;
	for (u32 s = res;; s >>= 1, ++k) {
		hm.off[k] = total;
		total += (u64)s * s;
		if (s == 1)
			break;
	}
	if (!generated) {
		if (p.hm_mn)
			HD_TRY(hipFree(p.hm_mn));
		if (p.hm_mx)
			HD_TRY(hipFree(p.hm_mx));
		p.hm_mn = p.hm_mx = nullptr;
		HD_TRY(hipMalloc(&p.hm_mn, total * 4));
		HD_TRY(hipMalloc(&p.hm_mx, total * 4));
		HD_TRY(hipMemcpyAsync(p.hm_mn, heights, (u64)res * res * 4, hipMemcpyHostToDevice, p.stream));
		HD_TRY(hipMemcpyAsync(p.hm_mx, heights, (u64)res * res * 4, hipMemcpyHostToDevice, p.stream));
	}
	k = 1;
	for (u32 s = res >> 1; s >= 1; s >>= 1, ++k) {
		u64 n = (u64)s * s;
		k_pyramid<<<div_up(n, 256), 256, 0, p.stream>>>(p.hm_mn + hm.off[k - 1], p.hm_mx + hm.off[k - 1],
		                                                 p.hm_mn + hm.off[k], p.hm_mx + hm.off[k], s);
		HD_TRY(hipGetLastError());
	}
	hm.mn = p.hm_mn;
	hm.mx = p.hm_mx;
	p.hm = hm;
	HD_TRY(hipMemcpyAsync(p.d_hm, &p.hm, sizeof(DevHeightmap), hipMemcpyHostToDevice, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

int hd_pool_read_words(hd_pool *h, uint64_t first, uint64_t count, uint32_t *out) {
	Pool &p = h->p;
	if (first + count > p.total_words)
		return set_error("read beyond pool"), HD_ERR_ARG;
	HD_TRY(hipMemcpyAsync(out, p.words + first, count * 4, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}
int hd_pool_write_words(hd_pool *h, uint64_t first, uint64_t count, const uint32_t *in) {
	Pool &p = h->p;
	if (first + count > p.total_words)
		return set_error("write beyond pool"), HD_ERR_ARG;
	HD_TRY(hipMemcpyAsync(p.words + first, in, count * 4, hipMemcpyHostToDevice, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}
int hd_pool_read_bucket_words(hd_pool *h, uint64_t first, uint64_t count, uint32_t *out) {
	Pool &p = h->p;
	if (first + count > p.total_buckets)
		return set_error("read beyond buckets"), HD_ERR_ARG;
	HD_TRY(hipMemcpyAsync(out, p.bucket_words + first, count * 4, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}
// A bucket counter above the bucket's capacity would make appends and finds run into the next
// bucket (or past the pool for the last one): every counter is checked before anything is copied.
static int check_bucket_counters(const Pool &p, const uint32_t *bw, uint64_t first, uint64_t count,
                                 const char *who) {
	const u64 wpb = 1ull << p.dc.wbpb();
	for (u64 i = 0; i < count; ++i)
		if (bw[i] > wpb)
			return set_error("%s: bucket %llu holds %u words (> %llu)", who, (unsigned long long)(first + i), bw[i],
			                 (unsigned long long)wpb),
			       HD_ERR_ARG;
	return HD_OK;
}
int hd_pool_write_bucket_words(hd_pool *h, uint64_t first, uint64_t count, const uint32_t *in) {
	Pool &p = h->p;
	if (first + count > p.total_buckets)
		return set_error("write beyond buckets"), HD_ERR_ARG;
	HD_RET(check_bucket_counters(p, in, first, count, "hd_pool_write_bucket_words"));
	HD_TRY(hipMemcpyAsync(p.bucket_words + first, in, count * 4, hipMemcpyHostToDevice, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}
int hd_pool_filled(hd_pool *h, uint32_t *out, uint32_t cap, uint32_t *n) {
	Pool &p = h->p;
	HD_RET(make_filled(p));
	*n = (u32)p.filled.size();
	for (u32 i = 0; i < p.filled.size() && i < cap; ++i)
		out[i] = p.filled[i];
	return HD_OK;
}
int hd_pool_stats(hd_pool *h, hd_stats *s) {
	Pool &p = h->p;
	u32 meta[kMetaCount];
	HD_RET(read_meta(p, meta));
	std::memset(s, 0, sizeof(*s));
	s->total_words = p.total_words;
	s->total_buckets = p.total_buckets;
	s->node_levels = p.dc.L;
	s->color_nodes = meta[kMetaColorNodes];
	s->color_leaf_words = meta[kMetaColorLeaves];
	s->overflow = meta[kMetaOverflow];
	std::vector<u32> bw(p.total_buckets);
	HD_TRY(hipMemcpy(bw.data(), p.bucket_words, p.total_buckets * 4, hipMemcpyDeviceToHost));
	for (u32 l = 0; l < p.dc.L; ++l) {
		u64 sum = 0;
		for (u32 b = p.dc.level_base[l]; b < p.dc.level_base[l] + (1u << p.dc.bucket_bits[l]); ++b)
			sum += bw[b];
		s->level_used_words[l] = sum;
		s->used_words += sum;
	}
	s->last_entries = p.last_entries;
	s->last_upserts = p.last_upserts;
	s->last_jobs = p.last_jobs;
	s->last_ms = p.last_ms;
	s->last_groups = p.last_groups;
	return HD_OK;
}
int hd_pool_read_color_nodes(hd_pool *h, uint64_t first, uint64_t count, uint32_t *out) {
	Pool &p = h->p;
	if (first + count > p.cnode_cap)
		return set_error("read beyond color nodes"), HD_ERR_ARG;
	HD_TRY(hipMemcpyAsync(out, p.cnodes + first * 8, count * 32, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}
int hd_pool_read_color_leaves(hd_pool *h, uint64_t first, uint64_t count, uint32_t *out) {
	Pool &p = h->p;
	if (first + count > p.cleaf_cap)
		return set_error("read beyond color leaves"), HD_ERR_ARG;
	HD_TRY(hipMemcpyAsync(out, p.cleaves + first, count * 4, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}
int hd_pool_clear_overflow(hd_pool *h) {
	HD_TRY(hipMemsetAsync(h->p.meta + kMetaOverflow, 0, 4, h->p.stream));
	HD_TRY(hipStreamSynchronize(h->p.stream));
	return HD_OK;
}

int hd_device_alloc(uint64_t bytes, void **out) {
	HD_TRY(hipMalloc(out, bytes ? bytes : 4));
	return HD_OK;
}
int hd_device_free(void *ptr) {
	HD_TRY(hipFree(ptr));
	return HD_OK;
}
int hd_memcpy_h2d(void *dst, const void *src, uint64_t bytes) {
	HD_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
	return HD_OK;
}
int hd_memcpy_d2h(void *dst, const void *src, uint64_t bytes) {
	HD_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
	return HD_OK;
}
int hd_device_count(int *n) {
	HD_TRY(hipGetDeviceCount(n));
	return HD_OK;
}

} // extern "C"

namespace hd {
// Plain 16-byte-vector copy: a kernel whose HBM traffic is known (bytes read = bytes written = n).
__global__ void k_copy16(uint4 *__restrict__ dst, const uint4 *__restrict__ src, u64 n16) {
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += (u64)gridDim.x * blockDim.x)
		dst[i] = src[i];
}
} // namespace hd

extern "C" int hd_device_copy(void *dst, const void *src, uint64_t bytes, hd_pool *pool) {
	if (!dst || !src || (bytes & 15u) || ((uintptr_t)dst & 15u) || ((uintptr_t)src & 15u))
		return set_error("hd_device_copy: null or unaligned argument (16-byte multiples)"), HD_ERR_ARG;
	hipStream_t s = pool ? pool->p.stream : nullptr;
	const u64 n16 = bytes >> 4;
	if (n16) {
		u32 g = (u32)std::min<u64>(div_up(n16, 256), 8192);
		k_copy16<<<g, 256, 0, s>>>((uint4 *)dst, (const uint4 *)src, n16);
		HD_TRY(hipGetLastError());
	}
	HD_TRY(hipStreamSynchronize(s));
	return HD_OK;
}

// Used words of every bucket packed in bucket order (the dirty-range Flush payload of
// src/DAGNodePool.cpp:56-85 for a full flush). *n receives the count; out may be null to query it.
extern "C" int hd_pool_read_live_words(hd_pool *h, uint32_t *out, uint64_t cap, uint64_t *n) {
	Pool &p = h->p;
	u64 *off, total;
	HD_RET(live_offsets(p, &off, &total));
	*n = total;
	if (!out)
		return HD_OK;
	if (cap < total)
		return set_error("hd_pool_read_live_words: buffer too small"), HD_ERR_ARG;
	u32 *tmp;
	HD_RET(p.scratch.get<u32>("live_tmp", total + 1, &tmp));
	k_gather_live<<<(u32)p.total_buckets, 256, 0, p.stream>>>(p.words, p.bucket_words, off, p.total_buckets, p.dc.wbpb(),
	                                                          tmp);
	HD_TRY(hipGetLastError());
	HD_TRY(hipMemcpyAsync(out, tmp, total * 4, hipMemcpyDeviceToHost, p.stream));
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}
// Inverse: set every bucket's counter and used words (host interop / checkpoint load).
extern "C" int hd_pool_write_live_words(hd_pool *h, const uint32_t *bucket_words, const uint32_t *live, uint64_t n) {
	if (!h || !bucket_words || (!live && n))
		return set_error("hd_pool_write_live_words: null argument"), HD_ERR_ARG;
	Pool &p = h->p;
	HD_RET(check_bucket_counters(p, bucket_words, 0, p.total_buckets, "hd_pool_write_live_words"));
	u64 sum = 0;
	for (u64 b = 0; b < p.total_buckets; ++b)
		sum += bucket_words[b];
	if (sum != n)
		return set_error("hd_pool_write_live_words: %llu words given, counters sum to %llu", (unsigned long long)n,
		                 (unsigned long long)sum),
		       HD_ERR_ARG;
	HD_TRY(hipMemcpyAsync(p.bucket_words, bucket_words, p.total_buckets * 4, hipMemcpyHostToDevice, p.stream));
	u64 *off, total;
	HD_RET(live_offsets(p, &off, &total));
	u32 *tmp;
	HD_RET(p.scratch.get<u32>("live_tmp", total + 1, &tmp));
	HD_TRY(hipMemcpyAsync(tmp, live, total * 4, hipMemcpyHostToDevice, p.stream));
	k_scatter_live<<<(u32)p.total_buckets, 256, 0, p.stream>>>(p.words, p.bucket_words, off, p.total_buckets,
	                                                           p.dc.wbpb(), tmp);
	HD_TRY(hipGetLastError());
	HD_TRY(hipStreamSynchronize(p.stream));
	return HD_OK;
}

// ---------------------------------------------------------------- checkpoint dump / load
// The pool's state as one file: the config, every bucket's word counter and used words (packed in
// bucket order: what DAGNodePool::Flush, src/DAGNodePool.cpp:56, uploads in a full flush), the color
// octree nodes and leaf words with their counters (DAGColorPool::Flush, src/DAGColorPool.cpp:50),
// the overflow flags and the caller's roots. The reference keeps no checkpoint of its own. Loading into a fresh pool restores a state from which edits, GC and traversal continue
// word for word. The heightfield is an editor input, not pool state, and is not saved.
namespace {
struct CkptHeader {
	char magic[8];
	uint32_t version, n_roots;
	hd_config cfg;
	uint64_t total_buckets, n_live, color_nodes, color_leaf_words;
	uint32_t overflow, flags; // flags bit 0: the filled-node chain had been created (it is a GC root)
};
constexpr uint32_t kCkptFilled = 1;
constexpr char kCkptMagic[8] = {'H', 'D', 'A', 'G', 'C', 'K', 'P', '1'};

struct File {
	FILE *f = nullptr;
	~File() {
		if (f)
			fclose(f);
	}
};
int put(FILE *f, const void *p, size_t bytes) {
	if (bytes && fwrite(p, 1, bytes, f) != bytes)
		return set_error("checkpoint: write failed"), HD_ERR_ARG;
	return HD_OK;
}
int get(FILE *f, void *p, size_t bytes) {
	if (bytes && fread(p, 1, bytes, f) != bytes)
		return set_error("checkpoint: truncated file"), HD_ERR_ARG;
	return HD_OK;
}
} // namespace

extern "C" int hd_pool_save(hd_pool *h, const char *path, const uint32_t *roots, uint32_t n_roots) {
	if (!h || !path || (!roots && n_roots))
		return set_error("hd_pool_save: null argument"), HD_ERR_ARG;
	Pool &p = h->p;
	u32 meta[kMetaCount];
	HD_RET(read_meta(p, meta));
	CkptHeader hd{};
	std::memcpy(hd.magic, kCkptMagic, 8);
	hd.version = 1;
	hd.n_roots = n_roots;
	hd.cfg = p.cfg;
	hd.total_buckets = p.total_buckets;
	hd.color_nodes = meta[kMetaColorNodes];
	hd.color_leaf_words = meta[kMetaColorLeaves];
	hd.overflow = meta[kMetaOverflow];
	hd.flags = p.filled.empty() ? 0u : kCkptFilled;
	HD_RET(hd_pool_read_live_words(h, nullptr, 0, &hd.n_live));
	std::vector<u32> bw(p.total_buckets), live(hd.n_live ? hd.n_live : 1), cn(8 * hd.color_nodes + 1),
	    cl(hd.color_leaf_words + 1);
	HD_TRY(hipMemcpy(bw.data(), p.bucket_words, p.total_buckets * 4, hipMemcpyDeviceToHost));
	u64 n = 0;
	HD_RET(hd_pool_read_live_words(h, live.data(), live.size(), &n));
	if (hd.color_nodes)
		HD_TRY(hipMemcpy(cn.data(), p.cnodes, hd.color_nodes * 32, hipMemcpyDeviceToHost));
	if (hd.color_leaf_words)
		HD_TRY(hipMemcpy(cl.data(), p.cleaves, hd.color_leaf_words * 4, hipMemcpyDeviceToHost));
	File F;
	F.f = fopen(path, "wb");
	if (!F.f)
		return set_error("hd_pool_save: cannot open %s", path), HD_ERR_ARG;
	HD_RET(put(F.f, &hd, sizeof(hd)));
	HD_RET(put(F.f, bw.data(), bw.size() * 4));
	HD_RET(put(F.f, live.data(), hd.n_live * 4));
	HD_RET(put(F.f, cn.data(), hd.color_nodes * 32));
	HD_RET(put(F.f, cl.data(), hd.color_leaf_words * 4));
	HD_RET(put(F.f, roots, (size_t)n_roots * 4));
	if (fflush(F.f) != 0)
		return set_error("hd_pool_save: write failed"), HD_ERR_ARG;
	return HD_OK;
}

static int pool_load(const char *path, hd_pool **out, uint32_t *roots, uint32_t cap, uint32_t *n_roots);
extern "C" int hd_pool_load(const char *path, hd_pool **out, uint32_t *roots, uint32_t cap, uint32_t *n_roots) {
	if (!path || !out || !n_roots || (!roots && cap))
		return set_error("hd_pool_load: null argument"), HD_ERR_ARG;
	try {
		return pool_load(path, out, roots, cap, n_roots);
	} catch (const std::exception &e) {
		return set_error("hd_pool_load: %s", e.what()), HD_ERR_ARG;
	}
}
static int pool_load(const char *path, hd_pool **out, uint32_t *roots, uint32_t cap, uint32_t *n_roots) {
	File F;
	F.f = fopen(path, "rb");
	if (!F.f)
		return set_error("hd_pool_load: cannot open %s", path), HD_ERR_ARG;
	CkptHeader hd{};
	HD_RET(get(F.f, &hd, sizeof(hd)));
	if (std::memcmp(hd.magic, kCkptMagic, 8) != 0 || hd.version != 1)
		return set_error("hd_pool_load: %s is not a version-1 checkpoint", path), HD_ERR_ARG;
	if (hd.n_roots > cap)
		return set_error("hd_pool_load: %u roots, room for %u", hd.n_roots, cap), HD_ERR_ARG;
	if (hd.color_nodes > hd.cfg.color_node_capacity || hd.color_leaf_words > hd.cfg.color_leaf_capacity)
		return set_error("hd_pool_load: color counts exceed the saved capacities"), HD_ERR_ARG;
	if (hd.flags & ~kCkptFilled)
		return set_error("hd_pool_load: unknown header flags 0x%x", hd.flags), HD_ERR_ARG;
	// counts against the config and the file size, before any device call or large allocation
	if (hd.cfg.node_levels == 0 || hd.cfg.node_levels > HD_MAX_LEVELS || hd.cfg.word_bits_per_page > 31 ||
	    hd.cfg.page_bits_per_bucket > 31)
		return set_error("hd_pool_load: invalid config in the header"), HD_ERR_ARG;
	u64 buckets = 0;
	for (u32 l = 0; l < hd.cfg.node_levels; ++l) {
		if (hd.cfg.bucket_bits[l] > 32)
			return set_error("hd_pool_load: invalid bucket bits in the header"), HD_ERR_ARG;
		buckets += 1ull << hd.cfg.bucket_bits[l];
	}
	const u64 wbpb = (u64)hd.cfg.word_bits_per_page + hd.cfg.page_bits_per_bucket;
	if (buckets != hd.total_buckets || wbpb > 32 || hd.n_live > (buckets << wbpb))
		return set_error("hd_pool_load: header counts do not fit the config"), HD_ERR_ARG;
	const u64 body = hd.total_buckets * 4 + hd.n_live * 4 + hd.color_nodes * 32 + hd.color_leaf_words * 4 +
	                 (u64)hd.n_roots * 4;
	const long here = ftell(F.f);
	if (here < 0 || fseek(F.f, 0, SEEK_END) != 0)
		return set_error("hd_pool_load: cannot size %s", path), HD_ERR_ARG;
	const long end = ftell(F.f);
	if (end < 0 || fseek(F.f, here, SEEK_SET) != 0)
		return set_error("hd_pool_load: cannot size %s", path), HD_ERR_ARG;
	if ((u64)(end - here) != body)
		return set_error("hd_pool_load: %s is truncated or oversized: %lld body bytes, the header describes %llu", path,
		                 (long long)(end - here), (unsigned long long)body),
		       HD_ERR_ARG;
	hd_pool *h = nullptr;
	HD_RET(hd_pool_create(&hd.cfg, &h));
	struct Guard {
		hd_pool *h;
		~Guard() {
			if (h)
				hd_pool_destroy(h);
		}
	} guard{h};
	Pool &p = h->p;
	if (p.total_buckets != hd.total_buckets)
		return set_error("hd_pool_load: bucket count mismatch"), HD_ERR_ARG;
	std::vector<u32> bw(hd.total_buckets), live(hd.n_live ? hd.n_live : 1), cn(8 * hd.color_nodes + 1),
	    cl(hd.color_leaf_words + 1);
	HD_RET(get(F.f, bw.data(), bw.size() * 4));
	HD_RET(get(F.f, live.data(), hd.n_live * 4));
	HD_RET(get(F.f, cn.data(), hd.color_nodes * 32));
	HD_RET(get(F.f, cl.data(), hd.color_leaf_words * 4));
	HD_RET(get(F.f, roots, (size_t)hd.n_roots * 4));
	const u64 wpb = 1ull << p.dc.wbpb();
	for (u64 b = 0; b < hd.total_buckets; ++b)
		if (bw[b] > wpb)
			return set_error("hd_pool_load: bucket %llu holds %u words (> %llu)", (unsigned long long)b, bw[b],
			                 (unsigned long long)wpb),
			       HD_ERR_ARG;
	HD_RET(hd_pool_write_live_words(h, bw.data(), live.data(), hd.n_live));
	if (hd.color_nodes)
		HD_TRY(hipMemcpy(p.cnodes, cn.data(), hd.color_nodes * 32, hipMemcpyHostToDevice));
	if (hd.color_leaf_words)
		HD_TRY(hipMemcpy(p.cleaves, cl.data(), hd.color_leaf_words * 4, hipMemcpyHostToDevice));
	u32 meta[kMetaCount] = {};
	meta[kMetaColorNodes] = (u32)hd.color_nodes;
	meta[kMetaColorLeaves] = (u32)hd.color_leaf_words;
	meta[kMetaOverflow] = hd.overflow;
	HD_TRY(hipMemcpy(p.meta, meta, sizeof(meta), hipMemcpyHostToDevice));
	// the saved pool holds the chain's nodes, so this finds them and appends nothing
	if (hd.flags & kCkptFilled)
		HD_RET(make_filled(p));
	*n_roots = hd.n_roots;
	*out = h;
	guard.h = nullptr;
	return HD_OK;
}

extern "C" int hd_pool_get_config(hd_pool *h, hd_config *out) {
	if (!h || !out)
		return set_error("hd_pool_get_config: null argument"), HD_ERR_ARG;
	*out = h->p.cfg;
	return HD_OK;
}

// ---------------------------------------------------------------- tuning
namespace hd {
Tuning g_tuning;
}
extern "C" int hd_tuning_get(hd_tuning *out) {
	if (!out)
		return set_error("hd_tuning_get: null argument"), HD_ERR_ARG;
	out->trace_tile = g_tuning.trace_tile;
	out->top_pairs = g_tuning.top_pairs;
	out->small_level = g_tuning.small_level;
	out->batch_pass_edits = g_tuning.batch_pass_edits;
	return HD_OK;
}
extern "C" int hd_tuning_set(const hd_tuning *t) {
	if (!t)
		return set_error("hd_tuning_set: null argument"), HD_ERR_ARG;
	if (t->trace_tile != 8 && t->trace_tile != 16)
		return set_error("hd_tuning_set: trace_tile must be 8 or 16 (got %u)", t->trace_tile), HD_ERR_ARG;
	if (t->top_pairs < 64 || t->top_pairs > 4096)
		return set_error("hd_tuning_set: top_pairs must be in [64, 4096] (got %u)", t->top_pairs), HD_ERR_ARG;
	if (t->small_level < 64 || t->small_level > 1024)
		return set_error("hd_tuning_set: small_level must be in [64, 1024] (got %u)", t->small_level), HD_ERR_ARG;
	g_tuning.trace_tile = t->trace_tile;
	g_tuning.top_pairs = t->top_pairs;
	g_tuning.small_level = t->small_level;
	g_tuning.batch_pass_edits = t->batch_pass_edits;
	return HD_OK;
}

extern "C" int hd_device_name(int dev, char *buf, uint32_t cap) {
	if (!buf || !cap)
		return set_error("hd_device_name: null argument"), HD_ERR_ARG;
	hipDeviceProp_t prop;
	HD_TRY(hipGetDeviceProperties(&prop, dev));
	std::snprintf(buf, cap, "%s (%s, %d CUs)", prop.name[0] ? prop.name : "AMD GPU", prop.gcnArchName,
	              prop.multiProcessorCount);
	return HD_OK;
}
