"""Host API over the C ABI, with torch tensors as device buffers.

Mirrors the reference's operator interface for the search path:
``Index.from_files`` / ``Index.from_rows`` stand for epic::MatrixSBWT::create
(include/matrix_sbwt.hpp:32), ``Index.presearch`` for epic::PreSearch
(include/presearch.hpp:32), ``Index.search_positions`` for the search kernel
ABI (include/gpu/search_kernel.hpp:62-63), and ``Index.search_reads`` for the
query half of epic::KmerSearch::search (include/epicseq/kmer_search.hpp:97-153).
Unsigned 64-bit values travel in torch.int64 tensors (same bits); ~0
(absent) reads as -1, the reference's printed value
(include/epicseq/results_persister.hpp:88-92).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _capi
from ._capi import check

NONE = np.uint64(0xFFFFFFFFFFFFFFFF)


def _stream(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _ptr(t: torch.Tensor | None):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _as_u64_np(a) -> np.ndarray:
    a = np.asarray(a)
    if a.dtype == np.int64:
        return a.view(np.uint64)
    return np.ascontiguousarray(a, dtype=np.uint64)


@dataclass
class Config:
    presearch_depth: int = 12
    layout: int = 16
    segment_kmers: int = 256

    def to_c(self) -> _capi.Config:
        c = _capi.Config()
        check(_capi.load().sbwt_config_init(C.byref(c)))
        c.presearch_depth = self.presearch_depth
        c.layout = self.layout
        c.segment_kmers = self.segment_kmers
        return c


class Index:
    """An SBWT index resident in HBM (owned by the library)."""

    def __init__(self, handle: C.c_void_p, config: Config):
        self._h = handle
        self.config = config
        self.lib = _capi.load()

    # ---- construction -------------------------------------------------
    @classmethod
    def from_rows(cls, rows, n: int, k: int, group=None, config: Config | None = None) -> "Index":
        """rows: four host arrays of MSB-first u64 words (A, C, G, T)."""
        config = config or Config()
        L = _capi.load()
        arrs = [np.ascontiguousarray(_as_u64_np(r)) for r in rows]
        ptrs = (C.POINTER(C.c_uint64) * 4)(*[a.ctypes.data_as(C.POINTER(C.c_uint64)) for a in arrs])
        g = None
        if group is not None:
            garr = np.ascontiguousarray(_as_u64_np(group))
            g = C.c_void_p(garr.ctypes.data)
        h = C.c_void_p()
        check(L.sbwt_index_from_rows(ptrs, g, n, k, C.byref(config.to_c()), C.byref(h)))
        return cls(h, config)

    @classmethod
    def from_files(cls, paths, k: int, group_path: str | None = None, config: Config | None = None) -> "Index":
        config = config or Config()
        L = _capi.load()
        ps = (C.c_char_p * 4)(*[p.encode() for p in paths])
        h = C.c_void_p()
        check(L.sbwt_index_from_files(ps, group_path.encode() if group_path else None, k,
                                      C.byref(config.to_c()), C.byref(h)))
        return cls(h, config)

    @classmethod
    def build(cls, seqs, k: int, config: Config | None = None, starts=None) -> "Index":
        """GPU SBWT construction. seqs: list of str/bytes, or a device uint8
        tensor of concatenated characters with `starts` (device int64, nseq+1)."""
        config = config or Config()
        L = _capi.load()
        if isinstance(seqs, torch.Tensor):
            chars, st = seqs, starts
        else:
            chars, st = _upload_rows(seqs)
        h = C.c_void_p()
        check(L.sbwt_index_build(_ptr(chars), _ptr(st), st.numel() - 1, k, C.byref(config.to_c()),
                                 _stream(), C.byref(h)))
        torch.cuda.synchronize()
        return cls(h, config)

    @classmethod
    def load(cls, path: str, config: Config | None = None) -> "Index":
        """Resume from an index image (sbwt_index_load): the saved
        configuration, or `config`."""
        L = _capi.load()
        h = C.c_void_p()
        check(L.sbwt_index_load(path.encode(), C.byref(config.to_c()) if config else None, C.byref(h)))
        ix = cls(h, config or Config())
        if config is None:
            info = ix.info
            ix.config = Config(presearch_depth=info["presearch_depth"], layout=info["layout"])
        return ix

    def save(self, path: str) -> None:
        """Checkpoint to an index image (sbwt_index_save)."""
        check(self.lib.sbwt_index_save(self._h, path.encode()))

    def close(self) -> None:
        if self._h:
            self.lib.sbwt_index_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ---- inspection ------------------------------------------------------
    @property
    def info(self) -> dict:
        i = _capi.IndexInfo()
        check(self.lib.sbwt_index_get_info(self._h, C.byref(i)))
        return {"n": i.n, "C": list(i.C), "device_bytes": i.device_bytes, "k": i.k,
                "presearch_depth": i.presearch_depth, "layout": i.layout, "has_groups": bool(i.has_groups)}

    def digest(self) -> dict:
        """Digests of the rows, rank chunks and presearch table in device
        memory (sbwt_index_digest): equal digests at two points mean
        nothing changed the index in between."""
        out = (C.c_uint64 * 3)()
        check(self.lib.sbwt_index_digest(self._h, out))
        return {"rows": int(out[0]), "chunks": int(out[1]), "table": int(out[2])}

    @property
    def n(self) -> int:
        return self.info["n"]

    @property
    def k(self) -> int:
        return self.info["k"]

    def rows(self):
        """(A, C, G, T, group) host copies, MSB-first words, ceil(n/4096)*64 words."""
        n = self.n
        nw = ((n + 4095) // 4096) * 64
        out = [np.zeros(nw, np.uint64) for _ in range(5)]
        ptrs = (C.POINTER(C.c_uint64) * 4)(*[a.ctypes.data_as(C.POINTER(C.c_uint64)) for a in out[:4]])
        g = C.c_void_p(out[4].ctypes.data) if self.info["has_groups"] else None
        check(self.lib.sbwt_index_get_rows(self._h, ptrs, g, nw))
        return out

    def write_files(self, paths, group_path: str | None = None) -> None:
        ps = (C.c_char_p * 4)(*[p.encode() for p in paths])
        check(self.lib.sbwt_index_write_files(self._h, ps, group_path.encode() if group_path else None))

    # ---- presearch / layout ------------------------------------------------
    def presearch(self, m: int) -> None:
        check(self.lib.sbwt_presearch(self._h, m, _stream()))

    def presearch_table(self):
        m = self.info["presearch_depth"]
        left = np.zeros(max(1, 1 << (2 * m)), np.uint64)
        right = np.zeros_like(left)
        check(self.lib.sbwt_presearch_get(self._h, C.c_void_p(left.ctypes.data), C.c_void_p(right.ctypes.data)))
        return (left, right) if m else (left[:0], right[:0])

    def set_layout(self, layout: int) -> None:
        check(self.lib.sbwt_index_set_layout(self._h, layout, _stream()))

    # ---- search ---------------------------------------------------------------
    def search_kmers(self, kmers: torch.Tensor, out: torch.Tensor | None = None, count: bool = False,
                     maybe: torch.Tensor | None = None, cooperative: bool = False) -> torch.Tensor:
        """kmers: device int64 tensor of top-aligned 2-bit k-mers; `maybe`
        (uint8 per k-mer, e.g. from prefilter()) skips those marked 0;
        cooperative: the wave-cooperative kernel (layouts 32 and 64)."""
        if out is None:
            out = torch.empty_like(kmers)
        flags = (_capi.COUNT if count else 0) | (_capi.SEARCH_COOPERATIVE if cooperative else 0)
        if maybe is None:
            check(self.lib.sbwt_search_kmers_flags(self._h, _ptr(kmers), kmers.numel(), _ptr(out), flags, _stream()))
        else:
            check(self.lib.sbwt_search_kmers_masked(self._h, _ptr(kmers), kmers.numel(), _ptr(maybe), _ptr(out),
                                                    flags, _stream()))
        return out

    def counters(self, reset: bool = True) -> dict:
        """Access counters of searches run with count=True (see sbwt_counters)."""
        out = (C.c_uint64 * 4)()
        check(self.lib.sbwt_counters(self._h, out, int(reset), _stream()))
        return {"lines": out[0], "steps": out[1], "kmers": out[2], "full_searches": out[3]}

    def select(self, c: int, ranks: torch.Tensor) -> torch.Tensor:
        """Column of the ranks[q]-th 1-bit of row c (~0 past the last)."""
        out = torch.empty_like(ranks)
        check(self.lib.sbwt_select(self._h, c, _ptr(ranks), ranks.numel(), _ptr(out), _stream()))
        return out

    def node_labels(self):
        """(labels, full): top-aligned label of every node, full=False for padded nodes."""
        n = self.n
        labels = torch.empty(n, dtype=torch.int64, device="cuda")
        full = torch.empty(n, dtype=torch.uint8, device="cuda")
        check(self.lib.sbwt_node_labels(self._h, _ptr(labels), _ptr(full), _stream()))
        return labels, full

    def enumerate_kmers(self) -> torch.Tensor:
        """Every k-mer of the index (top-aligned), in colex order."""
        out = torch.empty(self.n, dtype=torch.int64, device="cuda")
        cnt = C.c_uint64()
        check(self.lib.sbwt_enumerate_kmers(self._h, _ptr(out), C.byref(cnt), _stream()))
        return out[: cnt.value]

    def search_positions(self, packed: torch.Tensor, positions: torch.Tensor) -> torch.Tensor:
        """In place over `positions`, like the reference kernel."""
        check(self.lib.sbwt_search_positions(self._h, _ptr(packed), _ptr(positions), positions.numel(), _stream()))
        return positions

    def search_reads(self, reads: "ReadBatch", out: torch.Tensor | None = None, n_as_a: bool = False,
                     independent: bool = False, count: bool = False, persistent: bool = False,
                     certificates: bool = True, maybe: torch.Tensor | None = None,
                     lanes: bool = False) -> torch.Tensor:
        """Answers for every k-mer of the batch; `maybe` (uint8, one per k-mer,
        e.g. from prefilter()) skips the k-mers it marks 0."""
        if reads.k != self.k:
            raise ValueError(f"reads prepared for k={reads.k}, index has k={self.k}")
        if out is None:
            out = torch.empty(max(reads.total_kmers, 1), dtype=torch.int64, device=reads.chars.device)
        flags = ((_capi.READS_N_AS_A if n_as_a else 0) | (_capi.READS_INDEPENDENT if independent else 0)
                 | (_capi.COUNT if count else 0) | (_capi.READS_PERSISTENT if persistent else 0)
                 | (0 if certificates else _capi.READS_NO_CERT) | (_capi.READS_SEGMENT_LANES if lanes else 0))
        args = (self._h, _ptr(reads.packed), _ptr(reads.nmask), _ptr(reads.row_starts), _ptr(reads.offsets),
                _ptr(reads.work), _ptr(reads.items), reads.rows, reads.work_bound, reads.seg)
        if maybe is None:
            check(self.lib.sbwt_search_reads(*args, _ptr(out), flags, _stream()))
        else:
            check(self.lib.sbwt_search_reads_masked(*args, _ptr(maybe), _ptr(out), flags, _stream()))
        return out[: reads.total_kmers]


def _upload_rows(seqs, device="cuda"):
    enc = [s.encode() if isinstance(s, str) else bytes(s) for s in seqs]
    lens = np.array([len(s) for s in enc], dtype=np.int64)
    starts = np.zeros(len(enc) + 1, dtype=np.int64)
    starts[1:] = np.cumsum(lens)
    buf = np.frombuffer(b"".join(enc), dtype=np.uint8) if starts[-1] else np.zeros(1, np.uint8)
    chars = torch.from_numpy(buf.copy()).to(device)
    return chars, torch.from_numpy(starts).to(device)


class ReadBatch:
    """A set of rows (reads) resident in HBM, prepared for the read search:
    characters, row starts, packed 2-bit string + non-ACGT mask, per-row
    k-mer offsets and the work prefix."""

    def __init__(self, chars: torch.Tensor, row_starts: torch.Tensor, k: int, seg: int = 256,
                 total_kmers: int | None = None, host_starts: np.ndarray | None = None):
        L = _capi.load()
        self.chars, self.row_starts, self.k, self.seg = chars, row_starts, k, seg
        self.rows = row_starts.numel() - 1
        n_chars = chars.numel()
        dev = chars.device
        self.packed = torch.empty(L.sbwt_packed_words(n_chars), dtype=torch.int64, device=dev)
        self.nmask = torch.empty(L.sbwt_nmask_words(n_chars), dtype=torch.int64, device=dev)
        self.offsets = torch.empty(self.rows + 1, dtype=torch.int64, device=dev)
        s = _stream()
        check(L.sbwt_pack(_ptr(chars), n_chars, _ptr(self.packed), _ptr(self.nmask), s))
        check(L.sbwt_kmer_offsets(_ptr(row_starts), self.rows, k, _ptr(self.offsets), s))
        if total_kmers is None:
            if host_starts is not None:
                lens = np.diff(host_starts.astype(np.int64))
                total_kmers = int(np.maximum(lens - k + 1, 0).sum())
            else:
                total_kmers = int(self.offsets[-1].item())
        self.total_kmers = total_kmers
        self.work_bound = L.sbwt_reads_work_bound(self.rows, total_kmers, seg)
        self.work = torch.empty(self.rows + 1, dtype=torch.int64, device=dev)
        self.items = torch.empty(max(self.work_bound, 1), dtype=torch.int32, device=dev)
        check(L.sbwt_reads_work(_ptr(self.offsets), self.rows, seg, _ptr(self.work), _ptr(self.items), s))

    @classmethod
    def from_host(cls, seqs, k: int, seg: int = 256, device="cuda") -> "ReadBatch":
        chars, starts = _upload_rows(seqs, device)
        hs = starts.cpu().numpy()
        return cls(chars, starts, k, seg, host_starts=hs)

    def kmers(self) -> torch.Tensor:
        """Top-aligned code of every k-mer window, in output order."""
        L = _capi.load()
        pos = self.positions()
        out = torch.empty_like(pos)
        check(L.sbwt_extract_kmers(_ptr(self.packed), _ptr(pos), pos.numel(), _ptr(out), _stream()))
        return out

    def positions(self) -> torch.Tensor:
        """Start offset of every k-mer (include/gpu/positions_kernel.hpp:18-36)."""
        L = _capi.load()
        out = torch.empty(max(self.total_kmers, 1), dtype=torch.int64, device=self.chars.device)
        check(L.sbwt_positions(_ptr(self.row_starts), _ptr(self.offsets), self.rows, self.k, _ptr(out), _stream()))
        return out[: self.total_kmers]


def pack(chars: torch.Tensor):
    """2-bit packing + non-ACGT mask of a device uint8 tensor
    (include/gpu/create_kmer_string.hpp:22-42)."""
    L = _capi.load()
    n = chars.numel()
    packed = torch.empty(L.sbwt_packed_words(n), dtype=torch.int64, device=chars.device)
    nmask = torch.empty(L.sbwt_nmask_words(n), dtype=torch.int64, device=chars.device)
    check(L.sbwt_pack(_ptr(chars), n, _ptr(packed), _ptr(nmask), _stream()))
    return packed, nmask


def to_u64(t: torch.Tensor) -> np.ndarray:
    """Device int64 tensor -> host uint64 array (same bits)."""
    return t.detach().cpu().numpy().view(np.uint64)


def parse_lines(text: bytes):
    """Query text -> (chars uint8, row_starts int64) on the host
    (KmerSearch::read_file, include/epicseq/kmer_search.hpp:319-342)."""
    L = _capi.load()
    n = len(text)
    chars = np.zeros(max(n, 1), np.uint8)
    cap = text.count(b"\n") + 2
    rs = np.zeros(cap, np.uint64)
    rows = C.c_uint64()
    check(L.sbwt_parse_lines(text, n, C.c_void_p(chars.ctypes.data), C.c_void_p(rs.ctypes.data), cap,
                             C.byref(rows)))
    r = rows.value
    return chars[: int(rs[r])], rs[: r + 1].view(np.int64)


def parse_records(text: bytes):
    """FASTA / FASTQ text -> (chars uint8, row_starts int64), one row per
    record (plain text: one row per line, as parse_lines)."""
    L = _capi.load()
    n = len(text)
    chars = np.zeros(max(n, 1), np.uint8)
    cap = text.count(b"\n") + 2
    rs = np.zeros(cap, np.uint64)
    rows = C.c_uint64()
    check(L.sbwt_parse_records(text, n, C.c_void_p(chars.ctypes.data), C.c_void_p(rs.ctypes.data), cap,
                               C.byref(rows)))
    r = rows.value
    return chars[: int(rs[r])], rs[: r + 1].view(np.int64)


_OUTPUTS = {"text": _capi.OUTPUT_TEXT, "nodes": _capi.OUTPUT_NODES, "bitmap": _capi.OUTPUT_BITMAP}


def search_text(index: "Index", text: bytes, records: bool = False, fd: int = -1, batch_chars: int = 0,
                n_as_a: bool = False, independent: bool = False, output: str = "text") -> dict:
    """The pipelined query path over text in host memory (sbwt_search_text):
    results to file descriptor fd (-1: none) as result text, or as binary
    blocks of answers ("nodes") or found bits ("bitmap"; read either back
    with read_binary_results); returns the counts."""
    L = _capi.load()
    if output not in _OUTPUTS:
        raise ValueError(f"output {output!r}: one of {sorted(_OUTPUTS)}")
    stats = (C.c_uint64 * 3)()
    fmt = (_capi.TEXT_RECORDS if records else _capi.TEXT_LINES) | _OUTPUTS[output]
    check(L.sbwt_search_text(index._h, text, len(text), fmt,
                             _flags(n_as_a=n_as_a, independent=independent), batch_chars, fd, stats))
    return {"rows": stats[0], "kmers": stats[1], "found": stats[2]}


def read_binary_results(data: bytes) -> dict:
    """Decodes the binary output of search_text: {"kind": "nodes" | "bitmap",
    "k", "row_kmers" (k-mers per row, uint64), "values"} where values are the
    answers (uint64, ~0 = absent) or the found flags (bool), in row order."""
    if len(data) < 16 or data[:8] != b"SBWTRES1":
        raise ValueError("not a binary result file (no SBWTRES1 header)")
    kind_code, k = np.frombuffer(data, "<u4", 2, 8)
    if kind_code not in (1, 2):
        raise ValueError(f"unknown result kind {kind_code}")
    pos, counts, values = 16, [], []
    while pos < len(data):
        if pos + 16 > len(data):
            raise ValueError("truncated block header")
        rows, kmers = (int(x) for x in np.frombuffer(data, "<u8", 2, pos))
        pos += 16
        words = kmers if kind_code == 1 else (kmers + 63) // 64
        if pos + 8 * (rows + words) > len(data):
            raise ValueError("truncated block")
        c = np.frombuffer(data, "<u8", rows, pos)
        if int(c.sum()) != kmers:
            raise ValueError("block row counts do not add up to its k-mers")
        counts.append(c)
        pos += 8 * rows
        v = np.frombuffer(data, "<u8", words, pos)
        pos += 8 * words
        if kind_code == 2:
            v = np.unpackbits(v.view(np.uint8), bitorder="little")[:kmers].astype(bool)
        values.append(v)
    empty = np.zeros(0, np.uint64) if kind_code == 1 else np.zeros(0, bool)
    return {"kind": "nodes" if kind_code == 1 else "bitmap", "k": int(k),
            "row_kmers": np.concatenate(counts) if counts else np.zeros(0, np.uint64),
            "values": np.concatenate(values) if values else empty}


def found_bitmap(results: torch.Tensor, bit_offset: int = 0, found: bool = False, stream=None):
    """The found/absent bitmap of device answers (sbwt_found_bitmap): an
    int64 tensor of ceil((bit_offset + n) / 64) words, bit bit_offset + q set
    iff answer q is found; with found=True also the number found (a device
    int64 scalar tensor)."""
    L = _capi.load()
    n = results.numel()
    bits = torch.empty((bit_offset + n + 63) // 64, dtype=torch.int64, device=results.device)
    cnt = torch.empty(1, dtype=torch.int64, device=results.device) if found else None
    check(L.sbwt_found_bitmap(_ptr(results) if n else None, n, bit_offset, _ptr(bits) if bits.numel() else None,
                              _ptr(cnt), _stream(stream)))
    return (bits, cnt) if found else bits


def format_results(results: np.ndarray, offsets: np.ndarray) -> bytes:
    """Reference result text (include/epicseq/results_persister.hpp:82-96)."""
    L = _capi.load()
    res = np.ascontiguousarray(_as_u64_np(results))
    off = np.ascontiguousarray(_as_u64_np(offsets))
    rows = off.size - 1
    size = C.c_uint64()
    rp = C.c_void_p(res.ctypes.data) if res.size else None
    check(L.sbwt_format_results(rp, C.c_void_p(off.ctypes.data), rows, None, 0, C.byref(size)))
    buf = C.create_string_buffer(size.value + 1)
    check(L.sbwt_format_results(rp, C.c_void_p(off.ctypes.data), rows, buf, size.value, C.byref(size)))
    return buf.raw[: size.value]


def _flags(n_as_a=False, independent=False, count=False, persistent=False, certificates=True, lanes=False,
           runs=False) -> int:
    """runs: work items are the rows' N-free runs (plans only; see SBWT_READS_RUN_ITEMS)."""
    return ((_capi.READS_N_AS_A if n_as_a else 0) | (_capi.READS_INDEPENDENT if independent else 0)
            | (_capi.COUNT if count else 0) | (_capi.READS_PERSISTENT if persistent else 0)
            | (0 if certificates else _capi.READS_NO_CERT) | (_capi.READS_SEGMENT_LANES if lanes else 0)
            | (_capi.READS_RUN_ITEMS if runs else 0))


class ReadsPlan:
    """Preallocated scratch for batches of up to max_rows rows and max_chars
    characters: a batch's whole query path runs without allocating or
    synchronizing, so sequences of batches can be captured in a HIP graph."""

    def __init__(self, max_rows: int, max_chars: int, seg: int = 256):
        self.lib = _capi.load()
        self._h = C.c_void_p()
        check(self.lib.sbwt_reads_plan_create(max_rows, max_chars, seg, C.byref(self._h)))
        self.max_rows, self.max_chars, self.seg = max_rows, max_chars, seg

    def search(self, index: Index, chars: torch.Tensor, row_starts: torch.Tensor, out: torch.Tensor,
               offsets: torch.Tensor | None = None, stream=None, **flags) -> torch.Tensor:
        """chars / row_starts may be slices of larger device buffers; row_starts
        holds offsets into `chars` (rows + 1 entries)."""
        rows = row_starts.numel() - 1
        check(self.lib.sbwt_reads_plan_search(self._h, index._h, _ptr(chars), chars.numel(), _ptr(row_starts), rows,
                                              _ptr(out), _ptr(offsets), _flags(**flags), _stream(stream)))
        return out

    def close(self) -> None:
        if self._h:
            self.lib.sbwt_reads_plan_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Mlp:
    """The prefilter's MLP (4k -> 128 -> 64 -> 1), bf16 weights, FP32
    accumulation, trained on the GPU (see include/sbwt_gpu_amd.h)."""

    def __init__(self, k: int = 31, seed: int = 1, max_batch: int = 65536):
        self.lib = _capi.load()
        self._h = C.c_void_p()
        check(self.lib.sbwt_mlp_create(k, seed, max_batch, C.byref(self._h)))
        self.k, self.max_batch = k, max_batch
        self.param_count = int(self.lib.sbwt_mlp_param_count(self._h))
        self._loss = torch.zeros(1, dtype=torch.float32, device="cuda")

    @property
    def params(self) -> np.ndarray:
        p = np.zeros(self.param_count, np.float32)
        check(self.lib.sbwt_mlp_get_params(self._h, C.c_void_p(p.ctypes.data), p.size))
        return p

    @params.setter
    def params(self, p: np.ndarray) -> None:
        p = np.ascontiguousarray(p, np.float32)
        check(self.lib.sbwt_mlp_set_params(self._h, C.c_void_p(p.ctypes.data), p.size))

    def grad(self, kmers: torch.Tensor, labels: torch.Tensor):
        g = torch.empty(self.param_count, dtype=torch.float32, device="cuda")
        check(self.lib.sbwt_mlp_grad(self._h, _ptr(kmers), _ptr(labels), kmers.numel(), _ptr(g), _ptr(self._loss),
                                     _stream()))
        return g, self._loss

    def train_step(self, kmers: torch.Tensor, labels: torch.Tensor, lr: float) -> torch.Tensor:
        check(self.lib.sbwt_mlp_train_step(self._h, _ptr(kmers), _ptr(labels), kmers.numel(), C.c_float(lr),
                                           _ptr(self._loss), _stream()))
        return self._loss

    def logits(self, kmers: torch.Tensor, reference: bool = False) -> torch.Tensor:
        out = torch.empty(kmers.numel(), dtype=torch.float32, device="cuda")
        check(self.lib.sbwt_mlp_logits(self._h, _ptr(kmers), kmers.numel(), _ptr(out), int(reference), _stream()))
        return out

    def close(self):
        if self._h:
            self.lib.sbwt_mlp_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Bloom:
    """Blocked Bloom filter over top-aligned k-mers in HBM (only the top 2k
    bits of a key are hashed)."""

    def __init__(self, k: int, expected: int, fpr: float = 0.01):
        self.lib = _capi.load()
        self._h = C.c_void_p()
        check(self.lib.sbwt_bloom_create(k, expected, fpr, C.byref(self._h)))
        self.k = k
        self.hashes = int(self.lib.sbwt_bloom_hashes(self._h))
        self.bytes = int(self.lib.sbwt_bloom_bytes(self._h))

    def insert(self, keys: torch.Tensor) -> None:
        check(self.lib.sbwt_bloom_insert(self._h, _ptr(keys), keys.numel(), _stream()))

    def words(self) -> np.ndarray:
        w = np.zeros(self.bytes // 8, np.uint64)
        check(self.lib.sbwt_bloom_get_words(self._h, C.c_void_p(w.ctypes.data), w.size))
        return w

    def query(self, keys: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        if out is None:
            out = torch.empty(keys.numel(), dtype=torch.uint8, device="cuda")
        check(self.lib.sbwt_bloom_query(self._h, _ptr(keys), keys.numel(), _ptr(out), _stream()))
        return out

    def query_reads(self, reads: "ReadBatch", out: torch.Tensor | None = None) -> torch.Tensor:
        """The filter's answer for every k-mer of the batch (in output
        order), read from the packed rows without extracting k-mers."""
        if reads.k != self.k:
            raise ValueError(f"reads prepared for k={reads.k}, filter built for k={self.k}")
        if out is None:
            out = torch.empty(max(reads.total_kmers, 1), dtype=torch.uint8, device=reads.chars.device)
        check(self.lib.sbwt_bloom_query_reads(self._h, _ptr(reads.packed), _ptr(reads.row_starts),
                                              _ptr(reads.offsets), reads.rows, _ptr(out), _stream()))
        return out[: reads.total_kmers]

    def close(self):
        if self._h:
            self.lib.sbwt_bloom_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def prefilter(mlp: Mlp, threshold: float, bloom: "Bloom | None", kmers: torch.Tensor,
              out: torch.Tensor | None = None) -> torch.Tensor:
    """maybe[q] = logit >= threshold (matrix cores), or else a Bloom hit (a
    second pass over the rejected k-mers)."""
    L = _capi.load()
    if out is None:
        out = torch.empty(kmers.numel(), dtype=torch.uint8, device="cuda")
    check(L.sbwt_prefilter(mlp._h, C.c_float(threshold), bloom._h if bloom is not None else None, _ptr(kmers),
                           kmers.numel(), _ptr(out), _stream()))
    return out
