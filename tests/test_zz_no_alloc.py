"""hd_trace_noalloc: the frame kernel without any device allocation.

Bar: bit-exact. Through the C-ABI, two calls in a row (the second with fewer, smaller views)
give the same RGBA8 pixels and hit records as hd_trace on the same views, beam prepass included;
more views or beam texels than its static array holds are rejected before any device work."""
import numpy as np
import pytest

from test_oracle_pins import small_cfg
from vkhashdag_amd import workloads as W

pytestmark = pytest.mark.gpu


def _capi():
    from vkhashdag_amd import capi
    return capi


def _views(capi, specs):
    arr = (capi.hd_view * len(specs))()
    off = 0
    for i, (vb, w, h, flags) in enumerate(specs):
        arr[i] = capi.make_view(vb, w, h, flags, off)
        off += w * h
    return arr, off


def _frame(capi, pool, fn, specs):
    vs, npx = _views(capi, specs)
    rgba, rec = capi.DeviceBuffer(npx * 4), capi.DeviceBuffer(npx * 48)
    capi.check(fn(pool.h, pool.root, pool.color_root, vs, len(specs), rgba.ptr, rec.ptr))
    return rgba.download(np.uint32, npx), rec.download(np.uint32, npx * 12).reshape(npx, 12)


def _scene(capi):
    res = 128
    h = W.value_noise_heightmap(res, 4, 9, 20, 90)
    cfg = small_cfg(res, 8, 9, 2, 3)
    pool = capi.DevicePool(cfg, heights=h)
    pool.edit(W.edits([W.terrain_edit(res, seed=3, skin=6)] + list(W.terrain_sphere_edits(10, res, h, 3, 20, seed=8))))
    return pool, res


def test_noalloc_trace_equals_trace_on_two_calls():
    capi = _capi()
    pool, res = _scene(capi)
    F = capi.TRACE_SHADOW | capi.TRACE_RECORDS
    v = lambda seed, w, h, lod=True: W.orbit_view((64, 64, 40), 150, seed, w, h, res=res, lod=lod)  # noqa: E731
    first = [(v(0, 320, 200), 320, 200, F | capi.TRACE_BEAM), (v(1, 160, 96), 160, 96, F),
             (v(2, 64, 48, lod=False), 64, 48, F)]
    second = [(v(3, 96, 64), 96, 64, F | capi.TRACE_BEAM)]
    for specs in (first, second):
        print("views", [(w, h, fl) for _, w, h, fl in specs], flush=True)
        rgba_a, rec_a = _frame(capi, pool, capi.lib().hd_trace, specs)
        rgba_b, rec_b = _frame(capi, pool, capi.lib().hd_trace_noalloc, specs)
        assert (rec_a[:, 3] >> 31).mean() > 0.2
        assert np.array_equal(rec_a, rec_b)
        assert np.array_equal(rgba_a, rgba_b)


def test_noalloc_trace_rejects_what_its_array_cannot_hold():
    capi = _capi()
    pool, res = _scene(capi)
    vb = W.orbit_view((64, 64, 40), 150, 0, 16, 16, res=res)
    dummy = capi.DeviceBuffer(4)
    many, _ = _views(capi, [(vb, 16, 16, 0)] * 65)
    with pytest.raises(capi.HDError, match="at most 64"):
        capi.check(capi.lib().hd_trace_noalloc(pool.h, pool.root, pool.color_root, many, 65, dummy.ptr, None))
    big, _ = _views(capi, [(vb, 1920, 1080, capi.TRACE_BEAM)] * 33)  # 33 x 32400 beam texels > 2^20
    with pytest.raises(capi.HDError, match="beam texels"):
        capi.check(capi.lib().hd_trace_noalloc(pool.h, pool.root, pool.color_root, big, 33, dummy.ptr, None))
