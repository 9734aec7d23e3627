"""The C-ABI library loads and exports every function include/hd_capi.h declares (no GPU needed)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "hd_capi.h")
LIB = os.path.join(ROOT, "vkhashdag_amd", "libhashdag_amd.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(hd_[a-z_0-9]+)\s*\(", text, flags=re.M)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("hd_pool_create", "hd_edit", "hd_gc", "hd_trace", "hd_traversal"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_checkpoint_load_rejects_bad_files_before_touching_the_gpu(tmp_path):
    # hd_pool_load reads and checks the file header before it creates a pool: no GPU call is made
    lib = ctypes.CDLL(LIB)
    lib.hd_last_error.restype = ctypes.c_char_p
    h, n = ctypes.c_void_p(), ctypes.c_uint32(0)
    roots = (ctypes.c_uint32 * 4)()
    rc = lib.hd_pool_load(str(tmp_path / "missing").encode(), ctypes.byref(h), roots, 4, ctypes.byref(n))
    assert rc != 0 and b"cannot open" in lib.hd_last_error()
    bad = tmp_path / "bad"
    bad.write_bytes(b"NOTADAG!" + bytes(200))
    rc = lib.hd_pool_load(str(bad).encode(), ctypes.byref(h), roots, 4, ctypes.byref(n))
    assert rc != 0 and b"not a version-1 checkpoint" in lib.hd_last_error()
    bad.write_bytes(b"HDAG")
    rc = lib.hd_pool_load(str(bad).encode(), ctypes.byref(h), roots, 4, ctypes.byref(n))
    assert rc != 0 and b"truncated" in lib.hd_last_error()
    assert not h.value


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_tuning_set_get_and_validation():
    # process-wide launch parameters: no GPU call is made
    import sys
    sys.path.insert(0, ROOT)
    from vkhashdag_amd import capi
    old = capi.get_tuning()
    try:
        capi.set_tuning(trace_tile=8, top_pairs=512, small_level=128, batch_pass_edits=3)
        assert capi.get_tuning() == dict(trace_tile=8, top_pairs=512, small_level=128, batch_pass_edits=3)
        for bad in (dict(trace_tile=4), dict(top_pairs=8192), dict(small_level=2048), dict(small_level=10)):
            with pytest.raises(capi.HDError):
                capi.set_tuning(**bad)
        assert capi.get_tuning()["trace_tile"] == 8  # a rejected set changes nothing
    finally:
        capi.set_tuning(**old)


def _ckpt_header(n_live, total_buckets=None, flags=0, n_roots=0):
    """A checkpoint header (hd_pool.hip's CkptHeader) for a 64^3 pool of 2^6 buckets per level."""
    import sys
    sys.path.insert(0, ROOT)
    from vkhashdag_amd import capi

    class Hdr(ctypes.Structure):
        _fields_ = [("magic", ctypes.c_char * 8), ("version", ctypes.c_uint32), ("n_roots", ctypes.c_uint32),
                    ("cfg", capi.hd_config), ("total_buckets", ctypes.c_uint64), ("n_live", ctypes.c_uint64),
                    ("color_nodes", ctypes.c_uint64), ("color_leaf_words", ctypes.c_uint64),
                    ("overflow", ctypes.c_uint32), ("flags", ctypes.c_uint32)]
    h = Hdr()
    h.magic, h.version, h.n_roots = b"HDAGCKP1", 1, n_roots
    h.cfg.word_bits_per_page, h.cfg.page_bits_per_bucket, h.cfg.node_levels = 5, 2, 5
    for i in range(5):
        h.cfg.bucket_bits[i] = 6
    h.cfg.color_leaf_level, h.cfg.color_node_capacity, h.cfg.color_leaf_capacity = 2, 1 << 10, 1 << 12
    h.total_buckets = 5 * 64 if total_buckets is None else total_buckets
    h.n_live, h.flags = n_live, flags
    return bytes(h), h.total_buckets


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_checkpoint_header_counts_checked_before_the_gpu(tmp_path):
    # a corrupt n_live, a bucket count that does not fit the config, a body of the wrong size and
    # unknown flags are errors (not exceptions or huge allocations), found before any device call
    lib = ctypes.CDLL(LIB)
    lib.hd_last_error.restype = ctypes.c_char_p
    h, n = ctypes.c_void_p(), ctypes.c_uint32(0)
    roots = (ctypes.c_uint32 * 4)()
    f = tmp_path / "ck"
    hdr, nb = _ckpt_header(n_live=1 << 62)
    f.write_bytes(hdr + bytes(4 * nb))
    assert lib.hd_pool_load(str(f).encode(), ctypes.byref(h), roots, 4, ctypes.byref(n)) != 0
    assert b"header counts" in lib.hd_last_error()
    hdr, nb = _ckpt_header(n_live=0, total_buckets=7)
    f.write_bytes(hdr + bytes(4 * 7))
    assert lib.hd_pool_load(str(f).encode(), ctypes.byref(h), roots, 4, ctypes.byref(n)) != 0
    assert b"header counts" in lib.hd_last_error()
    hdr, nb = _ckpt_header(n_live=16)
    f.write_bytes(hdr + bytes(4 * nb + 4 * 15))  # one live word short
    assert lib.hd_pool_load(str(f).encode(), ctypes.byref(h), roots, 4, ctypes.byref(n)) != 0
    assert b"truncated" in lib.hd_last_error()
    hdr, nb = _ckpt_header(n_live=0, flags=6)
    f.write_bytes(hdr + bytes(4 * nb))
    assert lib.hd_pool_load(str(f).encode(), ctypes.byref(h), roots, 4, ctypes.byref(n)) != 0
    assert b"flags" in lib.hd_last_error()
    assert not h.value
