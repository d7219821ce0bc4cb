"""ctypes wrapper of the CPU oracle (``oracle/liboracle.so``).

Test infrastructure only: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py``, never by the product package.
Every function restates a piece of the reference algorithm; the C file cites
the reference file:line each one follows.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (gcc + OpenMP)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    L = C.CDLL(_LIB_PATH)
    vp = C.c_void_p
    L.orc_sbwt_build.restype = vp
    L.orc_sbwt_build.argtypes = [u8p, u64p, C.c_uint64, C.c_uint32]
    L.orc_sbwt_free.argtypes = [vp]
    for f in ("orc_sbwt_n", "orc_sbwt_nwords", "orc_sbwt_kmers"):
        getattr(L, f).restype = C.c_uint64
        getattr(L, f).argtypes = [vp]
    L.orc_sbwt_rows.argtypes = [vp, u64p, u64p, u64p, u64p, u64p]
    L.orc_sbwt_C.argtypes = [vp, u64p]
    L.orc_sbwt_labels.argtypes = [vp, u64p, u8p]
    L.orc_bv_write.argtypes = [C.c_char_p, u64p, C.c_uint64]
    L.orc_bv_header.restype = C.c_uint64
    L.orc_bv_header.argtypes = [C.c_char_p]
    L.orc_bv_read.argtypes = [C.c_char_p, u64p, C.c_uint64]
    L.orc_poppy_build.restype = vp
    L.orc_poppy_build.argtypes = [u64p, C.c_uint64, C.c_uint32, C.c_int]
    L.orc_poppy_free.argtypes = [vp]
    for f in ("orc_poppy_n0", "orc_poppy_n12", "orc_poppy_ones"):
        getattr(L, f).restype = C.c_uint64
        getattr(L, f).argtypes = [vp]
    for f in ("orc_poppy_L0", "orc_poppy_L12"):
        getattr(L, f).restype = C.POINTER(C.c_uint64)
        getattr(L, f).argtypes = [vp]
    L.orc_rank_poppy.restype = C.c_uint64
    L.orc_rank_poppy.argtypes = [vp, u64p, C.c_uint64]
    L.orc_rank_naive.restype = C.c_uint64
    L.orc_rank_naive.argtypes = [u64p, C.c_uint64]
    L.orc_index_open.restype = vp
    L.orc_index_open.argtypes = [u64p, u64p, u64p, u64p, C.c_uint64, C.c_uint32, C.c_int]
    L.orc_index_free.argtypes = [vp]
    L.orc_index_C.argtypes = [vp, u64p]
    L.orc_index_rank.restype = C.c_uint64
    L.orc_index_rank.argtypes = [vp, C.c_int, C.c_uint64]
    L.orc_presearch.argtypes = [vp, C.c_uint32, u64p, u64p]
    L.orc_search.argtypes = [vp, u64p, C.c_uint64, C.c_uint32, C.c_uint32, vp, vp, u64p, C.c_int]
    L.orc_search_count.argtypes = [vp, u64p, C.c_uint64, C.c_uint32, C.c_uint32, vp, vp,
                                   C.c_uint64, C.c_uint64, u64p, u32p, u32p]
    L.orc_pack.argtypes = [u8p, C.c_uint64, u64p]
    L.orc_extract.restype = C.c_uint64
    L.orc_extract.argtypes = [u64p, C.c_uint64]
    L.orc_positions.restype = C.c_uint64
    L.orc_positions.argtypes = [u64p, C.c_uint64, C.c_uint32, u64p]
    L.orc_kmer_offsets.argtypes = [u64p, C.c_uint64, C.c_uint32, u64p]
    L.orc_read_kmers.restype = C.c_uint64
    L.orc_read_kmers.argtypes = [u8p, u64p, C.c_uint64, C.c_uint32, C.c_int, u64p, u8p]
    L.orc_parse_lines.restype = C.c_uint64
    L.orc_parse_lines.argtypes = [u8p, C.c_uint64, u8p, u64p]
    L.orc_format_results.restype = C.c_uint64
    L.orc_format_results.argtypes = [u64p, u64p, C.c_uint64, C.c_char_p]
    _lib = L
    return L


def _u8(a) -> np.ndarray:
    if isinstance(a, (bytes, bytearray)):
        return np.frombuffer(bytes(a), dtype=np.uint8).copy()
    return np.ascontiguousarray(a, dtype=np.uint8)


class SBWT:
    """Brute-force SBWT of the k-mers of ``seqs`` (list of str/bytes)."""

    def __init__(self, seqs, k: int):
        if isinstance(seqs, (str, bytes)):
            seqs = [seqs]
        enc = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
        starts = np.zeros(len(enc) + 1, dtype=np.uint64)
        starts[1:] = np.cumsum([len(s) for s in enc])
        data = _u8(b"".join(enc)) if enc else np.zeros(1, np.uint8)
        if data.size == 0:
            data = np.zeros(1, np.uint8)
        L = lib()
        h = L.orc_sbwt_build(data, starts, len(enc), k)
        if not h:
            raise ValueError("oracle SBWT build failed")
        try:
            self.k = k
            self.n = int(L.orc_sbwt_n(h))
            self.nwords = int(L.orc_sbwt_nwords(h))
            self.num_kmers = int(L.orc_sbwt_kmers(h))
            nw = max(self.nwords, 1)
            self.rows = [np.zeros(nw, np.uint64) for _ in range(4)]
            self.group = np.zeros(nw, np.uint64)
            L.orc_sbwt_rows(h, *self.rows, self.group)
            self.C = np.zeros(5, np.uint64)
            L.orc_sbwt_C(h, self.C)
            self.keys = np.zeros(max(self.n, 1), np.uint64)
            self.lens = np.zeros(max(self.n, 1), np.uint8)
            L.orc_sbwt_labels(h, self.keys, self.lens)
            self.keys, self.lens = self.keys[: self.n], self.lens[: self.n]
        finally:
            L.orc_sbwt_free(h)

    def label(self, j: int) -> str:
        """Node j's label as a string with '$' padding."""
        key, ln = int(self.keys[j]), int(self.lens[j])
        chars = ["ACGT"[(key >> (2 * p)) & 3] for p in range(self.k)]
        return "$" * (self.k - ln) + "".join(chars[self.k - ln:])


class Poppy:
    """Reference poppy / cum-poppy L0/L12 over one row."""

    def __init__(self, words: np.ndarray, n: int, sb: int = 1024, version: int = 1):
        L = lib()
        self.words = np.ascontiguousarray(words, dtype=np.uint64)
        self.h = L.orc_poppy_build(self.words, n, sb, version)
        if not self.h:
            raise ValueError("unsupported poppy parameters")
        n0, n12 = int(L.orc_poppy_n0(self.h)), int(L.orc_poppy_n12(self.h))
        self.L0 = np.ctypeslib.as_array(L.orc_poppy_L0(self.h), shape=(n0,)).copy()
        self.L12 = np.ctypeslib.as_array(L.orc_poppy_L12(self.h), shape=(n12,)).copy()
        self.ones = int(L.orc_poppy_ones(self.h))

    def rank(self, i: int) -> int:
        return int(lib().orc_rank_poppy(self.h, self.words, i))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_poppy_free(self.h)
            self.h = None


def rank_naive(words: np.ndarray, i: int) -> int:
    return int(lib().orc_rank_naive(np.ascontiguousarray(words, np.uint64), i))


class Index:
    """Reference CPU search path: 4 rows + poppy rank + C (+ presearch)."""

    def __init__(self, rows, n: int, sb: int = 1024, version: int = 1):
        L = lib()
        self.rows = [np.ascontiguousarray(r, dtype=np.uint64) for r in rows]
        self.n = n
        self.h = L.orc_index_open(*self.rows, n, sb, version)
        if not self.h:
            raise ValueError("oracle index open failed")
        self.C = np.zeros(5, np.uint64)
        L.orc_index_C(self.h, self.C)
        self.m = 0
        self.left = self.right = None

    def rank(self, c: int, i: int) -> int:
        return int(lib().orc_index_rank(self.h, c, i))

    def presearch(self, m: int):
        self.m = m
        if m == 0:
            self.left = self.right = None
            return None
        self.left = np.zeros(1 << (2 * m), np.uint64)
        self.right = np.zeros(1 << (2 * m), np.uint64)
        lib().orc_presearch(self.h, m, self.left, self.right)
        return self.left, self.right

    def _lr(self):
        if self.m == 0:
            return None, None
        return self.left.ctypes.data, self.right.ctypes.data

    def search(self, kmers: np.ndarray, k: int, threads: int = 0) -> np.ndarray:
        kmers = np.ascontiguousarray(kmers, dtype=np.uint64)
        out = np.zeros(max(kmers.size, 1), np.uint64)
        l, r = self._lr()
        lib().orc_search(self.h, kmers if kmers.size else out, kmers.size, k, self.m, l, r, out, threads)
        return out[: kmers.size]

    def search_count(self, kmers: np.ndarray, k: int, cols_per_chunk: int, chunk_bytes: int):
        kmers = np.ascontiguousarray(kmers, dtype=np.uint64)
        n = kmers.size
        out = np.zeros(max(n, 1), np.uint64)
        steps = np.zeros(max(n, 1), np.uint32)
        lines = np.zeros(max(n, 1), np.uint32)
        l, r = self._lr()
        lib().orc_search_count(self.h, kmers if n else out, n, k, self.m, l, r,
                               cols_per_chunk, chunk_bytes, out, steps, lines)
        return out[:n], steps[:n], lines[:n]

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_index_free(self.h)
            self.h = None


def bv_write(path: str, words: np.ndarray, n: int) -> None:
    if lib().orc_bv_write(path.encode(), np.ascontiguousarray(words, np.uint64), n):
        raise OSError(path)


def bv_read(path: str):
    n = int(lib().orc_bv_header(path.encode()))
    if n == 0:
        raise OSError(path)
    nw = ((n + 4095) // 4096) * 64
    w = np.zeros(nw, np.uint64)
    if lib().orc_bv_read(path.encode(), w, nw):
        raise OSError(path)
    return w, n


def pack(raw: bytes) -> np.ndarray:
    a = _u8(raw)
    assert a.size % 32 == 0
    out = np.zeros(a.size // 32, np.uint64)
    lib().orc_pack(a, a.size // 32, out)
    return out


def extract(packed: np.ndarray, pos: int) -> int:
    return int(lib().orc_extract(np.ascontiguousarray(packed, np.uint64), pos))


def positions(row_starts: np.ndarray, k: int) -> np.ndarray:
    rs = np.ascontiguousarray(row_starts, np.uint64)
    tot = int(sum(max(0, int(rs[i + 1] - rs[i]) - k + 1) for i in range(rs.size - 1)))
    out = np.zeros(max(tot, 1), np.uint64)
    m = lib().orc_positions(rs, rs.size - 1, k, out)
    return out[:m]


def kmer_offsets(row_starts: np.ndarray, k: int) -> np.ndarray:
    rs = np.ascontiguousarray(row_starts, np.uint64)
    off = np.zeros(rs.size, np.uint64)
    lib().orc_kmer_offsets(rs, rs.size - 1, k, off)
    return off


def read_kmers(chars, row_starts: np.ndarray, k: int, n_breaks: bool = True):
    ch = _u8(chars)
    if ch.size == 0:
        ch = np.zeros(1, np.uint8)
    rs = np.ascontiguousarray(row_starts, np.uint64)
    tot = int(kmer_offsets(rs, k)[-1])
    km = np.zeros(max(tot, 1), np.uint64)
    valid = np.zeros(max(tot, 1), np.uint8)
    m = lib().orc_read_kmers(ch, rs, rs.size - 1, k, int(n_breaks), km, valid)
    return km[:m], valid[:m].astype(bool)


def parse_lines(text: bytes):
    t = _u8(text)
    if t.size == 0:
        t = np.zeros(1, np.uint8)
    chars = np.zeros(max(len(text), 1), np.uint8)
    rs = np.zeros(len(text) + 2, np.uint64)
    rows = int(lib().orc_parse_lines(t, len(text), chars, rs))
    return chars[: int(rs[rows])], rs[: rows + 1]


def format_results(results: np.ndarray, offsets: np.ndarray) -> bytes:
    res = np.ascontiguousarray(results, np.uint64)
    if res.size == 0:
        res = np.zeros(1, np.uint64)
    off = np.ascontiguousarray(offsets, np.uint64)
    rows = off.size - 1
    n = int(lib().orc_format_results(res, off, rows, None))
    buf = C.create_string_buffer(n + 1)
    lib().orc_format_results(res, off, rows, buf)
    return buf.raw[:n]


def search_reads(index: Index, chars, row_starts, k: int, n_breaks: bool = True) -> np.ndarray:
    """Reference search over every k-mer of every row (rows concatenated).
    With n_breaks, windows holding a non-ACGT character answer ~0 unsearched."""
    km, valid = read_kmers(chars, row_starts, k, n_breaks)
    out = index.search(km, k)
    if n_breaks:
        out[~valid] = np.uint64(0xFFFFFFFFFFFFFFFF)
    return out
