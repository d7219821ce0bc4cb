"""ctypes binding of libsbwt_gpu_amd.so (the C ABI declared in
include/sbwt_gpu_amd.h). Loading fails loudly when the library is missing:
there is no CPU fallback in this package."""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsbwt_gpu_amd.so")

SBWT_OK = 0
ERRORS = {1: "SBWT_ERR_ARG", 2: "SBWT_ERR_HIP", 3: "SBWT_ERR_IO", 4: "SBWT_ERR_FORMAT", 5: "SBWT_ERR_UNSUPPORTED",
          6: "SBWT_ERR_INTERNAL"}
READS_N_AS_A = 1
READS_INDEPENDENT = 2
COUNT = 4
READS_PERSISTENT = 8
READS_NO_CERT = 16
READS_SEGMENT_LANES = 32
READS_RUN_ITEMS = 64
SEARCH_COOPERATIVE = 128
TEXT_LINES = 0
TEXT_RECORDS = 1
OUTPUT_TEXT = 0
OUTPUT_NODES = 0x100
OUTPUT_BITMAP = 0x200

# every function the header declares, with (restype, argtypes)
vp, u64, u32, i32 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int
u64p = C.POINTER(C.c_uint64)
cp = C.c_char_p


class Config(C.Structure):
    _fields_ = [("presearch_depth", u32), ("layout", u32), ("block_threads", u32),
                ("segment_kmers", u32), ("variant", u32), ("reserved", u32 * 3)]


class IndexInfo(C.Structure):
    _fields_ = [("n", u64), ("C", u64 * 5), ("device_bytes", u64), ("k", u32),
                ("presearch_depth", u32), ("layout", u32), ("has_groups", u32)]


SIGNATURES = {
    "sbwt_config_init": (i32, [C.POINTER(Config)]),
    "sbwt_index_from_rows": (i32, [u64p * 4, vp, u64, u32, C.POINTER(Config), C.POINTER(vp)]),
    "sbwt_index_from_files": (i32, [cp * 4, cp, u32, C.POINTER(Config), C.POINTER(vp)]),
    "sbwt_index_build": (i32, [vp, vp, u64, u32, C.POINTER(Config), vp, C.POINTER(vp)]),
    "sbwt_build_set_sort_chunk": (i32, [u64]),
    "sbwt_index_free": (i32, [vp]),
    "sbwt_index_get_info": (i32, [vp, C.POINTER(IndexInfo)]),
    "sbwt_index_digest": (i32, [vp, u64p]),
    "sbwt_index_get_rows": (i32, [vp, u64p * 4, vp, u64]),
    "sbwt_index_write_files": (i32, [vp, cp * 4, cp]),
    "sbwt_presearch": (i32, [vp, u32, vp]),
    "sbwt_presearch_get": (i32, [vp, vp, vp]),
    "sbwt_index_set_layout": (i32, [vp, u32, vp]),
    "sbwt_counters": (i32, [vp, u64p, i32, vp]),
    "sbwt_search_kmers_flags": (i32, [vp, vp, u64, vp, u32, vp]),
    "sbwt_search_kmers": (i32, [vp, vp, u64, vp, vp]),
    "sbwt_search_positions": (i32, [vp, vp, vp, u64, vp]),
    "sbwt_packed_words": (u64, [u64]),
    "sbwt_nmask_words": (u64, [u64]),
    "sbwt_pack": (i32, [vp, u64, vp, vp, vp]),
    "sbwt_kmer_offsets": (i32, [vp, u64, u32, vp, vp]),
    "sbwt_positions": (i32, [vp, vp, u64, u32, vp, vp]),
    "sbwt_extract_kmers": (i32, [vp, vp, u64, vp, vp]),
    "sbwt_search_kmers_masked": (i32, [vp, vp, u64, vp, vp, u32, vp]),
    "sbwt_search_reads_masked": (i32, [vp, vp, vp, vp, vp, vp, vp, u64, u64, u32, vp, vp, u32, vp]),
    "sbwt_reads_work_bound": (u64, [u64, u64, u32]),
    "sbwt_reads_work": (i32, [vp, u64, u32, vp, vp, vp]),
    "sbwt_search_reads": (i32, [vp, vp, vp, vp, vp, vp, vp, u64, u64, u32, vp, u32, vp]),
    "sbwt_select": (i32, [vp, u32, vp, u64, vp, vp]),
    "sbwt_node_labels": (i32, [vp, vp, vp, vp]),
    "sbwt_enumerate_kmers": (i32, [vp, vp, u64p, vp]),
    "sbwt_mlp_create": (i32, [u32, u64, u64, C.POINTER(vp)]),
    "sbwt_mlp_free": (i32, [vp]),
    "sbwt_mlp_param_count": (u64, [vp]),
    "sbwt_mlp_set_params": (i32, [vp, vp, u64]),
    "sbwt_mlp_get_params": (i32, [vp, vp, u64]),
    "sbwt_mlp_grad": (i32, [vp, vp, vp, u64, vp, vp, vp]),
    "sbwt_mlp_train_step": (i32, [vp, vp, vp, u64, C.c_float, vp, vp]),
    "sbwt_mlp_logits": (i32, [vp, vp, u64, vp, i32, vp]),
    "sbwt_bloom_create": (i32, [u32, u64, C.c_double, C.POINTER(vp)]),
    "sbwt_bloom_free": (i32, [vp]),
    "sbwt_bloom_bytes": (u64, [vp]),
    "sbwt_bloom_hashes": (u32, [vp]),
    "sbwt_bloom_get_words": (i32, [vp, vp, u64]),
    "sbwt_bloom_insert": (i32, [vp, vp, u64, vp]),
    "sbwt_bloom_query": (i32, [vp, vp, u64, vp, vp]),
    "sbwt_bloom_query_reads": (i32, [vp, vp, vp, vp, u64, vp, vp]),
    "sbwt_prefilter": (i32, [vp, C.c_float, vp, vp, u64, vp, vp]),
    "sbwt_gather_roof": (i32, [vp, u64, u32, u32, u32, u64, u64, vp, vp]),
    "sbwt_parse_lines": (i32, [cp, u64, vp, vp, u64, u64p]),
    "sbwt_parse_records": (i32, [cp, u64, vp, vp, u64, u64p]),
    "sbwt_search_text": (i32, [vp, cp, u64, u32, u32, u64, i32, vp]),
    "sbwt_found_bitmap": (i32, [vp, u64, u32, vp, vp, vp]),
    "sbwt_index_save": (i32, [vp, cp]),
    "sbwt_index_load": (i32, [cp, vp, vp]),
    "sbwt_format_text_scratch": (u64, [u64]),
    "sbwt_format_text": (i32, [vp, vp, u64, u64, vp, u64, vp, vp, u64, vp]),
    "sbwt_format_results": (i32, [vp, vp, u64, vp, u64, u64p]),
    "sbwt_write_results": (i32, [vp, vp, u64, i32]),
    "sbwt_reads_plan_create": (i32, [u64, u64, u32, C.POINTER(vp)]),
    "sbwt_reads_plan_free": (i32, [vp]),
    "sbwt_reads_plan_search": (i32, [vp, vp, vp, u64, vp, u64, vp, vp, u32, vp]),
    "sbwt_reads_search_static": (i32, [vp, vp, u64, vp, u64, u32, vp, vp, u32, vp]),
    "sbwt_last_error": (cp, []),
}

_lib = None


class SbwtError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def load():
    """Loads the in-tree library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m sbwt_gpu_amd.build` "
            "(the package has no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(status: int) -> None:
    if status != SBWT_OK:
        msg = load().sbwt_last_error().decode(errors="replace")
        raise SbwtError(status, msg)
