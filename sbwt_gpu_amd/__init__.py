"""sbwt_gpu_amd — SBWT k-mer membership search on one AMD Instinct MI355X.

The compute path is libsbwt_gpu_amd.so (hand-written HIP kernels for gfx950
behind the C ABI of include/sbwt_gpu_amd.h); this package binds it and uses
torch tensors only as device buffers. Importing the package does not load
the library; the first call into it does, and fails loudly if it is missing.
"""
from ._capi import SbwtError, load  # noqa: F401
from .api import (NONE, Bloom, Config, Index, Mlp, ReadBatch, ReadsPlan, format_results,  # noqa: F401
                  found_bitmap, pack, parse_lines, parse_records, prefilter, read_binary_results, search_text,
                  to_u64)

__all__ = ["Config", "Index", "ReadBatch", "ReadsPlan", "Mlp", "Bloom", "prefilter", "pack", "parse_lines", "parse_records", "search_text", "read_binary_results",
           "found_bitmap", "format_results", "to_u64", "NONE", "SbwtError",
           "load"]
