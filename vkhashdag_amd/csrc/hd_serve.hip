// hd_serve.hip -- request loop over one device-resident DAG: edit and render requests from many
// clients are admitted per tick as edit slots, pixel budget and free node-pool pages allow; a
// tick's edits run as one batched edit (submission order kept per voxel) while its renders run as
// one multi-view trace on a second stream, reading the DAG version of the tick's start.
//
// This replaces the reference app's loop of one edit in flight with later edits dropped
// (src/main.cpp:214-296, :287) and rendering on the old root while an edit runs (the append-only
// pool makes that safe, main.cpp:244-250). Concurrent rendering needs append-only colors too:
// DAGColorPool::SetLeaf rewrites a leaf in place unless it keeps history (DAGColorPool.hpp:172-204),
// so the server requires color_keep_history.
#include "hd_internal.hpp"

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

namespace hd {

// used pages of the pool: sum over buckets of ceil(bucket words / words per page)
__global__ void k_used_pages(const u32 *__restrict__ bucket_words, u64 n, u32 wbpp, unsigned long long *out) {
	unsigned long long acc = 0;
	for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (u64)gridDim.x * blockDim.x)
		acc += (bucket_words[i] + (1u << wbpp) - 1u) >> wbpp;
	for (int o = 32; o > 0; o >>= 1)
		acc += __shfl_down(acc, o);
	if ((threadIdx.x & 63u) == 0 && acc)
		atomicAdd(out, acc);
}
__global__ void k_zero_u64(unsigned long long *p) {
	if (threadIdx.x == 0 && blockIdx.x == 0)
		*p = 0;
}

struct Server {
	Pool *p = nullptr;
	hipStream_t rstream = nullptr;                   // render stream (default priority)
	hipEvent_t ev_start = nullptr, ev_render = nullptr;
	unsigned long long *d_pages = nullptr, *h_pages = nullptr;
	hd_view *d_views = nullptr;
	u32 *d_tiles = nullptr, cap_views = 0;
	~Server() {
		if (rstream)
			(void)hipStreamDestroy(rstream);
		if (ev_start)
			(void)hipEventDestroy(ev_start);
		if (ev_render)
			(void)hipEventDestroy(ev_render);
		if (d_pages)
			(void)hipFree(d_pages);
		if (h_pages)
			(void)hipHostFree(h_pages);
		if (d_views)
			(void)hipFree(d_views);
		if (d_tiles)
			(void)hipFree(d_tiles);
	}
	int init(Pool &pool) {
		p = &pool;
		HD_TRY(hipStreamCreateWithFlags(&rstream, hipStreamNonBlocking));
		HD_TRY(hipEventCreate(&ev_start));
		HD_TRY(hipEventCreate(&ev_render));
		HD_TRY(hipMalloc(&d_pages, 8));
		HD_TRY(hipHostMalloc((void **)&h_pages, 8, hipHostMallocDefault));
		return HD_OK;
	}
	int views(u32 n) {
		if (n <= cap_views)
			return HD_OK;
		if (d_views)
			HD_TRY(hipFree(d_views));
		if (d_tiles)
			HD_TRY(hipFree(d_tiles));
		cap_views = std::max(n, 64u);
		HD_TRY(hipMalloc(&d_views, cap_views * sizeof(hd_view)));
		HD_TRY(hipMalloc(&d_tiles, (cap_views + 1) * 4));
		return HD_OK;
	}
	// free pages of the node pool (synchronizes the pool stream)
	int free_pages(u64 *out) {
		Pool &q = *p;
		k_zero_u64<<<1, 64, 0, q.stream>>>(d_pages);
		u32 g = (u32)std::min<u64>(div_up(q.total_buckets, 256), 2048);
		k_used_pages<<<g, 256, 0, q.stream>>>(q.bucket_words, q.total_buckets, q.dc.wbpp, d_pages);
		HD_TRY(hipGetLastError());
		HD_TRY(hipMemcpyAsync(h_pages, d_pages, 8, hipMemcpyDeviceToHost, q.stream));
		HD_TRY(hipStreamSynchronize(q.stream));
		u64 total = q.total_buckets << q.dc.pbpb;
		*out = total - std::min<u64>(total, *h_pages);
		return HD_OK;
	}
};

using Clock = std::chrono::steady_clock;
static double ms_since(Clock::time_point t0) {
	return std::chrono::duration<double, std::milli>(Clock::now() - t0).count();
}

} // namespace hd

using namespace hd;

extern "C" int hd_serve_run(hd_pool *h, const hd_serve_config *cfg, uint32_t root, uint32_t color_root,
                            const hd_request *reqs, uint32_t n, uint32_t *rgba, uint64_t rgba_pixels,
                            hd_request_result *results, uint32_t *root_out, uint32_t *color_root_out,
                            hd_serve_stats *stats) {
	if (!h || !cfg || (!reqs && n) || !results || !root_out || !color_root_out)
		return set_error("hd_serve_run: null argument"), HD_ERR_ARG;
	Pool &p = h->p;
	const bool plain = (cfg->flags & HD_SERVE_PLAIN) != 0;
	if (!plain && !p.dc.keep_history)
		return set_error("hd_serve_run: concurrent rendering needs color_keep_history (in-place color leaves)"),
		       HD_ERR_ARG;
	if (cfg->max_edits_per_tick == 0 || cfg->max_pixels_per_tick == 0)
		return set_error("hd_serve_run: empty tick budget"), HD_ERR_ARG;
	u64 need_pixels = 0;
	for (u32 i = 0; i < n; ++i) {
		if (i && reqs[i].arrival_ms < reqs[i - 1].arrival_ms)
			return set_error("hd_serve_run: requests must be sorted by arrival"), HD_ERR_ARG;
		if (reqs[i].kind == HD_REQ_RENDER) {
			const hd_view &v = reqs[i].view;
			if (!v.width || !v.height)
				return set_error("hd_serve_run: empty view (request %u)", i), HD_ERR_ARG;
			if ((u64)v.width * v.height > cfg->max_pixels_per_tick)
				return set_error("hd_serve_run: view larger than the tick's pixel budget (request %u)", i),
				       HD_ERR_ARG;
			need_pixels += (u64)v.width * v.height;
		} else if (reqs[i].kind != HD_REQ_EDIT)
			return set_error("hd_serve_run: unknown request kind %u", reqs[i].kind), HD_ERR_ARG;
		else if (!edit_color_ok(reqs[i].edit))
			return set_error("hd_serve_run: edit request %u: bits_per_weight > 3", i), HD_ERR_ARG;
	}
	// output pixels: every render keeps its image when rgba holds them all, else a tick scratch
	const bool keep = rgba && rgba_pixels >= need_pixels;
	u32 *scratch = nullptr;
	if (!keep)
		HD_RET(p.scratch.get<u32>("serve_rgba", cfg->max_pixels_per_tick, &scratch));
	Server S;
	HD_RET(S.init(p));
	HD_RET(make_filled(p));

	hd_serve_stats st{};
	st.min_free_pages = ~0ull;
	std::vector<u32> pend_edit, pend_render; // arrived, not yet admitted (FIFO)
	std::vector<hd_edit_desc> tick_edits;
	std::vector<hd_view> tick_views;
	std::vector<u32> tick_ids, tiles;
	u32 next = 0, done = 0, tick = 0;
	u64 pix_out = 0;
	u32 r = root, c = color_root;
	const auto t0 = Clock::now();
	while (done < n) {
		double now = ms_since(t0);
		while (next < n && reqs[next].arrival_ms <= now) {
			(reqs[next].kind == HD_REQ_EDIT ? pend_edit : pend_render).push_back(next);
			++next;
		}
		if (pend_edit.empty() && pend_render.empty()) {
			double wait = reqs[next].arrival_ms - now; // next < n here
			if (wait > 0.05)
				std::this_thread::sleep_for(std::chrono::microseconds((long long)(wait * 1000.0 * 0.9)));
			continue;
		}
		if (plain) { // one request at a time, in arrival order (the baseline loop)
			u32 id = pend_edit.empty() ? pend_render.front()
			                           : (pend_render.empty() ? pend_edit.front()
			                                                  : std::min(pend_edit.front(), pend_render.front()));
			auto &q = reqs[id].kind == HD_REQ_EDIT ? pend_edit : pend_render;
			q.erase(q.begin());
			hd_request_result &o = results[id];
			o.tick = tick;
			o.admit_ms = ms_since(t0);
			if (reqs[id].kind == HD_REQ_EDIT) {
				HD_RET(hd_edit(h, r, c, &reqs[id].edit, 1, HD_MODE_PLAIN, &r, &c));
				o.root = r, o.color_root = c;
				st.edits++;
			} else {
				hd_view v = reqs[id].view;
				u64 off = keep ? pix_out : 0;
				v.out_offset = off;
				HD_RET(hd_trace(h, r, c, &v, 1, keep ? rgba : scratch, nullptr));
				o.root = r, o.color_root = c, o.pixel_offset = off;
				pix_out += keep ? (u64)v.width * v.height : 0;
				st.renders++;
			}
			o.done_ms = ms_since(t0);
			++done, ++tick;
			continue;
		}
		// admission: edit slots gated on free pages, renders up to the pixel budget
		tick_edits.clear();
		tick_views.clear();
		tick_ids.clear();
		u32 n_e = 0;
		if (!pend_edit.empty()) {
			u64 fp = 0;
			HD_RET(S.free_pages(&fp));
			if (fp < cfg->min_free_pages) { // compact, then re-check
				u32 roots[1] = {r};
				HD_RET(hd_gc(h, roots, 1));
				r = roots[0];
				st.gcs++;
				HD_RET(S.free_pages(&fp));
			}
			st.min_free_pages = std::min<u64>(st.min_free_pages, fp);
			if (fp >= cfg->min_free_pages)
				n_e = std::min<u32>((u32)pend_edit.size(), cfg->max_edits_per_tick);
			else
				st.gated_ticks++;
		}
		const u32 rt2 = r, ct2 = c; // the DAG version this tick's renders see (after a GC: remapped)
		u64 pixels = 0;
		u32 n_r = 0;
		while (n_r < pend_render.size()) {
			const hd_view &v = reqs[pend_render[n_r]].view;
			u64 px = (u64)v.width * v.height;
			if (pixels + px > cfg->max_pixels_per_tick)
				break;
			pixels += px;
			++n_r;
		}
		if (n_e == 0 && n_r == 0)
			return set_error("hd_serve_run: %zu edits wait for free pages that GC cannot provide (%llu wanted)",
			                 pend_edit.size(), (unsigned long long)cfg->min_free_pages),
			       HD_ERR_CAPACITY;
		const double admit = ms_since(t0);
		// renders first, on their own stream, against the tick's starting version
		if (n_r) {
			tiles.assign(n_r + 1, 0);
			u64 off = keep ? pix_out : 0;
			for (u32 k = 0; k < n_r; ++k) {
				u32 id = pend_render[k];
				hd_view v = reqs[id].view;
				v.out_offset = off;
				results[id].pixel_offset = off;
				off += (u64)v.width * v.height;
				tick_views.push_back(v);
			}
			HD_RET(hd_trace_tiles(tick_views.data(), n_r, tiles.data()));
			HD_RET(S.views(n_r));
			HD_TRY(hipMemcpyAsync(S.d_views, tick_views.data(), n_r * sizeof(hd_view), hipMemcpyHostToDevice,
			                      S.rstream));
			HD_TRY(hipMemcpyAsync(S.d_tiles, tiles.data(), (n_r + 1) * 4, hipMemcpyHostToDevice, S.rstream));
			HD_TRY(hipEventRecord(S.ev_start, S.rstream));
			HD_RET(trace_on_stream(p, S.rstream, rt2, ct2, S.d_views, S.d_tiles, n_r, tiles[n_r],
			                       keep ? rgba : scratch));
			HD_TRY(hipEventRecord(S.ev_render, S.rstream));
			if (keep)
				pix_out = off;
		}
		// edits: one batched edit (its own stream), overlapping the renders
		double edit_done = admit;
		if (n_e) {
			for (u32 k = 0; k < n_e; ++k)
				tick_edits.push_back(reqs[pend_edit[k]].edit);
			HD_RET(hd_edit(h, r, c, tick_edits.data(), n_e, HD_MODE_BATCHED, &r, &c));
			edit_done = ms_since(t0);
		}
		double render_done = admit;
		if (n_r) {
			HD_TRY(hipEventSynchronize(S.ev_render));
			float dev_ms = 0;
			HD_TRY(hipEventElapsedTime(&dev_ms, S.ev_start, S.ev_render));
			render_done = std::min(ms_since(t0), admit + (double)dev_ms);
		}
		for (u32 k = 0; k < n_e; ++k) {
			hd_request_result &o = results[pend_edit[k]];
			o.tick = tick, o.admit_ms = admit, o.done_ms = edit_done, o.root = r, o.color_root = c;
		}
		for (u32 k = 0; k < n_r; ++k) {
			hd_request_result &o = results[pend_render[k]];
			o.tick = tick, o.admit_ms = admit, o.done_ms = render_done, o.root = rt2, o.color_root = ct2;
		}
		pend_edit.erase(pend_edit.begin(), pend_edit.begin() + n_e);
		pend_render.erase(pend_render.begin(), pend_render.begin() + n_r);
		st.edits += n_e;
		st.renders += n_r;
		st.max_edits_in_tick = std::max(st.max_edits_in_tick, n_e);
		st.max_renders_in_tick = std::max(st.max_renders_in_tick, n_r);
		done += n_e + n_r;
		++tick;
	}
	HD_TRY(hipStreamSynchronize(p.stream));
	st.ticks = tick;
	if (st.min_free_pages == ~0ull)
		st.min_free_pages = 0;
	st.wall_ms = ms_since(t0);
	if (stats)
		*stats = st;
	*root_out = r;
	*color_root_out = c;
	return HD_OK;
}

extern "C" int hd_pool_free_pages(hd_pool *h, uint64_t *free_pages, uint64_t *total_pages) {
	if (!h || !free_pages)
		return set_error("hd_pool_free_pages: null argument"), HD_ERR_ARG;
	Server S;
	HD_RET(S.init(h->p));
	HD_RET(S.free_pages(free_pages));
	if (total_pages)
		*total_pages = h->p.total_buckets << h->p.dc.pbpb;
	return HD_OK;
}
