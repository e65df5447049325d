"""C ABI of libbyo.so without a GPU: the library loads, exports every symbol
include/byo.h declares, and reports a missing device as an error."""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "byo.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(byo_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import byo_amd
    L = C.CDLL(byo_amd.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # the Python mirror binds exactly the declared entry points
    assert sorted(byo_amd.SIGNATURES) == syms


def test_version_and_error_plumbing():
    import byo_amd
    L = byo_amd.lib()
    assert b"gfx950" in L.byo_version()
    # null arguments are rejected before any device work
    assert L.byo_num_vertices(None, None) == byo_amd.E_INVAL
    assert b"null" in L.byo_last_error()
    assert L.byo_graph_destroy(None) == byo_amd.OK


def test_no_device_is_an_error_not_a_fallback():
    import numpy as np
    import byo_amd
    if byo_amd.device_count() > 0:
        pytest.skip("a device is visible; covered by the gpu tests")
    off = np.array([0, 1, 2], np.uint64)
    ids = np.array([1, 0], np.uint32)
    with pytest.raises(byo_amd.ByoError) as e:
        byo_amd.CsrGraph.from_csr(off, ids)
    assert e.value.code == byo_amd.E_NODEV


def test_bfs_stats_layout_matches_header():
    import byo_amd
    # uint32,uint32,uint64,double, 64 u64, 64 u64, 64 u8, 64 float
    assert C.sizeof(byo_amd.BfsStats) == 4 + 4 + 8 + 8 + 512 + 512 + 64 + 256
