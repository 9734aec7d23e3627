// hd_gc.hip -- device GC (placeholder until the GC kernels land).
#include "hd_internal.hpp"
using namespace hd;
extern "C" int hd_gc(hd_pool *h, uint32_t *roots, uint32_t n) {
	(void)h, (void)roots, (void)n;
	return set_error("hd_gc not built yet"), HD_ERR_UNSUPPORTED;
}
