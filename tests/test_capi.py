"""CPU checks of the C-ABI library: it loads, and it exports every function
that include/sbwt_gpu_amd.h declares (no compute call: there is no GPU here)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sbwt_gpu_amd.h")
LIB = os.path.join(ROOT, "sbwt_gpu_amd", "libsbwt_gpu_amd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(sbwt_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_the_api():
    names = declared_functions()
    for must in ("sbwt_index_from_rows", "sbwt_index_from_files", "sbwt_index_build", "sbwt_search_kmers",
                 "sbwt_search_positions", "sbwt_search_reads", "sbwt_presearch", "sbwt_pack", "sbwt_last_error"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_binding_covers_the_header():
    from sbwt_gpu_amd import _capi

    assert set(declared_functions()) == set(_capi.SIGNATURES)
    _capi.load()


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_host_only_entry_points():
    """Pure size helpers and argument checks need no device."""
    from sbwt_gpu_amd import _capi

    L = _capi.load()
    assert L.sbwt_packed_words(0) == 2 and L.sbwt_packed_words(64) == 4
    assert L.sbwt_nmask_words(64) == 3
    assert L.sbwt_reads_work_bound(10, 1000, 100) == 10 + 10 + 1
    cfg = _capi.Config()
    assert L.sbwt_config_init(ctypes.byref(cfg)) == 0
    assert cfg.presearch_depth == 12 and cfg.layout == 16
    # invalid k is rejected before any device call
    h = ctypes.c_void_p()
    rows = (ctypes.POINTER(ctypes.c_uint64) * 4)()
    st = L.sbwt_index_from_rows(rows, None, 10, 40, ctypes.byref(cfg), ctypes.byref(h))
    assert st != 0
    assert L.sbwt_last_error()
