"""The pipelined query path (sbwt_search_text): parsing, copies, search and
result text in overlapping batches must write exactly the text of the
one-shot path, which is checked against the CPU oracle elsewhere."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from oracle import records as R

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("needs a GPU", allow_module_level=True)

import sbwt_gpu_amd as sg  # noqa: E402

ACGT = np.array(list("ACGT"))


def _genome(rng, n):
    return "".join(ACGT[rng.integers(0, 4, n)])


def _reads(rng, g, count, lo=20, hi=400, n_rate=0.01):
    out = []
    for _ in range(count):
        L = int(rng.integers(lo, hi))
        p = int(rng.integers(0, len(g) - L))
        r = np.array(list(g[p:p + L]))
        r[rng.random(L) < 0.01] = ACGT[rng.integers(0, 4)]
        r[rng.random(L) < n_rate] = "N"
        out.append("".join(r))
    return out


@pytest.fixture(scope="module")
def index():
    rng = np.random.default_rng(41)
    g = _genome(rng, 200_000)
    ix = sg.Index.build([g], 31, sg.Config(presearch_depth=10))
    yield ix, g
    ix.close()


def _expected(ix, chars, starts, k=31, **flags):
    rows = len(starts) - 1
    if rows == 0:
        return b"", 0, 0
    batch = sg.ReadBatch(torch.from_numpy(np.ascontiguousarray(chars, np.uint8).copy() if len(chars) else
                                          np.zeros(1, np.uint8)).cuda(),
                         torch.from_numpy(np.asarray(starts, np.int64).copy()).cuda(), k,
                         host_starts=np.asarray(starts, np.int64))
    res = sg.to_u64(ix.search_reads(batch, **flags))
    lens = np.diff(np.asarray(starts, np.int64))
    off = np.zeros(rows + 1, np.uint64)
    off[1:] = np.cumsum(np.maximum(lens - k + 1, 0))
    return sg.format_results(res, off), int(off[-1]), int((res != sg.NONE).sum())


def _run(ix, text, tmp_path, **kw):
    path = tmp_path / "out.txt"
    fd = os.open(path, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
    try:
        stats = sg.search_text(ix, text, fd=fd, **kw)
    finally:
        os.close(fd)
    return path.read_bytes(), stats


@pytest.mark.parametrize("batch", [0, 700, 5_000, 64_000])
def test_lines_text_equals_one_shot(index, tmp_path, batch):
    ix, g = index
    rng = np.random.default_rng(batch + 1)
    lines = _reads(rng, g, 1500) + ["", ">header", "ACGT", g[1000:21000]] + _reads(rng, _genome(rng, 50_000), 500)
    rng.shuffle(lines)
    text = ("\n".join(lines) + "\n").encode()
    got, stats = _run(ix, text, tmp_path, batch_chars=batch)
    chars, starts = O.parse_lines(text)
    want, kmers, found = _expected(ix, chars, starts)
    assert got == want
    assert stats == {"rows": len(starts) - 1, "kmers": kmers, "found": found}


def test_large_batches_parse_on_threads(index, tmp_path):
    """Batches of 4 MiB and more are parsed on host threads and their found
    k-mers counted on the device: the text and the counts still equal the
    one-shot path's."""
    ix, g = index
    rng = np.random.default_rng(99)
    lines = _reads(rng, g, 30_000, 100, 300) + [">h", "", "ACGT\r"] + _reads(rng, _genome(rng, 60_000), 10_000, 100, 300)
    rng.shuffle(lines)
    text = ("\n".join(lines) + "\n").encode()
    assert len(text) > (6 << 20)
    chars, starts = O.parse_lines(text)
    want, kmers, found = _expected(ix, chars, starts)
    for batch in (0, 5 << 20):
        got, stats = _run(ix, text, tmp_path, batch_chars=batch)
        assert got == want
        assert stats == {"rows": len(starts) - 1, "kmers": kmers, "found": found}
        got, stats = _run(ix, text, tmp_path, batch_chars=batch, output="bitmap")
        assert stats["found"] == found


@pytest.mark.parametrize("fmt", ["fasta", "fastq"])
def test_records_text_equals_one_shot(index, tmp_path, fmt):
    ix, g = index
    rng = np.random.default_rng(7 if fmt == "fasta" else 8)
    parts = []
    for i, s in enumerate(_reads(rng, g, 800, 1, 600) + [g[50_000:70_000]]):
        w = int(rng.integers(10, 90))
        wrap = [s[j:j + w] for j in range(0, len(s), w)]
        if fmt == "fasta":
            parts += [f">r{i}"] + wrap
        else:
            q = "".join(rng.choice(list("@+I#"), len(s)))
            parts += [f"@r{i}"] + wrap + ["+"] + [q[j:j + w] for j in range(0, len(q), w)]
    text = ("\r\n".join(parts) + "\r\n").encode()
    chars, starts = R.parse_records(text)
    want, kmers, found = _expected(ix, chars, starts)
    for batch in (3_000, 0):
        got, stats = _run(ix, text, tmp_path, records=True, batch_chars=batch)
        assert got == want
        assert stats["rows"] == len(starts) - 1 and stats["kmers"] == kmers and stats["found"] == found


def test_flags_and_edges(index, tmp_path):
    ix, g = index
    rng = np.random.default_rng(3)
    text = ("\n".join(_reads(rng, g, 300, n_rate=0.05)) + "\n").encode()
    chars, starts = O.parse_lines(text)
    for flags in ({"n_as_a": True}, {"independent": True}):
        got, _ = _run(ix, text, tmp_path, batch_chars=2_000, **flags)
        assert got == _expected(ix, chars, starts, **flags)[0]
    assert _run(ix, b"", tmp_path)[0] == b""
    # rows shorter than k only: empty result lines, in order
    got, stats = _run(ix, b"ACGT\nAC\n\nA\n", tmp_path, batch_chars=3)
    assert got == b"\n\n\n" and stats == {"rows": 3, "kmers": 0, "found": 0}
    with pytest.raises(sg.SbwtError):
        sg.search_text(ix, b"@r\nACGT\n+\nII\n", records=True)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_format_text_on_device_equals_host(seed):
    """The GPU's result text is byte for byte the host's (which pins the
    reference's format against the oracle), for node ids, ~0, values past
    2^32 and negative int64 values, rows with no answers; a buffer that is too
    small is left untouched and the size tells what is needed."""
    import ctypes as C

    from sbwt_gpu_amd import _capi
    from sbwt_gpu_amd.api import _ptr

    L = _capi.load()
    rng = np.random.default_rng(seed)
    rows = int(rng.integers(0, 3000)) if seed else 0
    counts = rng.integers(0, 40, rows)
    counts[rng.random(rows) < 0.2] = 0
    off = np.zeros(rows + 1, np.uint64)
    off[1:] = np.cumsum(counts)
    count = int(off[-1])
    res = rng.integers(0, 200_000_000, count, dtype=np.uint64)
    res[rng.random(count) < 0.3] = np.uint64(0xFFFFFFFFFFFFFFFF)
    big = rng.random(count) < 0.02
    res[big] = rng.integers(1 << 32, 1 << 63, int(big.sum()), dtype=np.uint64)
    neg = rng.random(count) < 0.02
    res[neg] = rng.integers(1 << 63, (1 << 64) - 1, int(neg.sum()), dtype=np.uint64)
    if count > 3:
        res[:3] = np.array([0, (1 << 63), (1 << 63) - 1], np.uint64)
    want = sg.format_results(res, off)
    d_res = torch.from_numpy(res.view(np.int64).copy() if count else np.zeros(1, np.int64)).cuda()
    d_off = torch.from_numpy(off.view(np.int64).copy()).cuda()
    cap = count * 21 + rows
    d_text = torch.zeros(max(cap, 1), dtype=torch.uint8, device="cuda")
    d_size = torch.zeros(1, dtype=torch.int64, device="cuda")
    sb = int(L.sbwt_format_text_scratch(count))
    scratch = torch.empty(sb, dtype=torch.uint8, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _capi.check(L.sbwt_format_text(_ptr(d_res), _ptr(d_off), rows, count, _ptr(d_text), cap, _ptr(d_size),
                                   _ptr(scratch), sb, s))
    n = int(d_size.item())
    assert n == len(want)
    assert bytes(d_text[:n].cpu().numpy()) == want
    if n > 1:
        d_small = torch.full((n - 1,), 7, dtype=torch.uint8, device="cuda")
        _capi.check(L.sbwt_format_text(_ptr(d_res), _ptr(d_off), rows, count, _ptr(d_small), n - 1, _ptr(d_size),
                                       _ptr(scratch), sb, s))
        assert int(d_size.item()) == n and bool((d_small == 7).all().item())
    with pytest.raises(sg.SbwtError):
        _capi.check(L.sbwt_format_text(_ptr(d_res), _ptr(d_off), rows, count, _ptr(d_text), cap, _ptr(d_size),
                                       _ptr(scratch), sb - 1 if sb else 0, s))


@pytest.mark.parametrize("output", ["nodes", "bitmap"])
def test_binary_outputs_match_oracle(tmp_path, output):
    """Binary results decode to the oracle's answers (nodes) or to their
    found flags (bitmap), row by row, whatever the batch size; the blocks
    cut rows where the text batches were cut."""
    rng = np.random.default_rng(11)
    g = _genome(rng, 120_000)
    k = 31
    s = O.SBWT([g], k)
    oix = O.Index(s.rows, s.n)
    oix.presearch(9)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=9))
    lines = _reads(rng, g, 600, 1, 300) + ["", "ACGT", g[:31]] + _reads(rng, _genome(rng, 9_000), 200)
    rng.shuffle(lines)
    text = ("\n".join(lines) + "\n").encode()
    chars, starts = O.parse_lines(text)
    want = O.search_reads(oix, chars, starts, k)
    lens = np.diff(np.asarray(starts, np.int64))
    row_kmers = np.maximum(lens - k + 1, 0).astype(np.uint64)
    for batch in (0, 900, 7_000):
        got, stats = _run(ix, text, tmp_path, batch_chars=batch, output=output)
        dec = sg.read_binary_results(got)
        assert dec["kind"] == output and dec["k"] == k
        np.testing.assert_array_equal(dec["row_kmers"], row_kmers)
        if output == "nodes":
            np.testing.assert_array_equal(dec["values"], want)
        else:
            np.testing.assert_array_equal(dec["values"], want != sg.NONE)
        assert stats == {"rows": len(lens), "kmers": int(row_kmers.sum()), "found": int((want != sg.NONE).sum())}
    ix.close()


def test_binary_output_edges(index, tmp_path):
    ix, _ = index
    for output, code in (("nodes", 1), ("bitmap", 2)):
        got, stats = _run(ix, b"", tmp_path, output=output)
        assert got == b"SBWTRES1" + np.array([code, 31], "<u4").tobytes()
        dec = sg.read_binary_results(got)
        assert dec["row_kmers"].size == 0 and dec["values"].size == 0
        # rows shorter than k only (an empty line is no row, as in sbwt_parse_lines)
        got, stats = _run(ix, b"ACGT\nAC\n\nA\n", tmp_path, output=output, batch_chars=3)
        dec = sg.read_binary_results(got)
        np.testing.assert_array_equal(dec["row_kmers"], [0, 0, 0])
        assert dec["values"].size == 0 and stats == {"rows": 3, "kmers": 0, "found": 0}
    with pytest.raises(sg.SbwtError):
        _raw_format(ix, 0x300)
    with pytest.raises(sg.SbwtError):
        _raw_format(ix, 2)
    with pytest.raises(ValueError):
        sg.search_text(ix, b"ACGT\n", output="json")


def _raw_format(ix, fmt):
    from sbwt_gpu_amd import _capi

    L = _capi.load()
    _capi.check(L.sbwt_search_text(ix._h, b"ACGT\n", 5, fmt, 0, 0, -1, None))


@pytest.mark.parametrize("count", [0, 1, 63, 64, 65, 1000, 100_003])
def test_found_bitmap(count):
    """sbwt_found_bitmap against numpy's bit packing, at every bit offset
    class (0, inside the first word, the last bit of the first word), with
    the found count."""
    rng = np.random.default_rng(count)
    res = rng.integers(0, 1 << 40, count, dtype=np.uint64)
    res[rng.random(count) < 0.4] = sg.NONE
    if count > 2:
        res[0], res[-1] = sg.NONE, np.uint64(0)
    d_res = torch.from_numpy(res.view(np.int64).copy()).cuda()
    for off in (0, 1, 37, 63):
        bits, found = sg.found_bitmap(d_res, off, found=True)
        flags = np.concatenate([np.zeros(off, bool), res != sg.NONE])
        nwords = (off + count + 63) // 64
        padded = np.zeros(nwords * 64, bool)
        padded[:flags.size] = flags
        want = np.packbits(padded, bitorder="little").view(np.uint64)
        np.testing.assert_array_equal(sg.to_u64(bits), want)
        assert int(found.item()) == int((res != sg.NONE).sum())


@pytest.mark.parametrize("output", ["nodes", "bitmap"])
def test_binary_outputs_of_records(index, tmp_path, output):
    """FASTQ records (one row each) with binary outputs: row counts and
    values equal the one-shot search of the parsed records."""
    ix, g = index
    rng = np.random.default_rng(21)
    parts = []
    for i, r in enumerate(_reads(rng, g, 300, 20, 300)):
        w = int(rng.integers(20, 80))
        wrap = [r[j:j + w] for j in range(0, len(r), w)]
        parts += [f"@r{i}"] + wrap + ["+"] + ["I" * len(x) for x in wrap]
    text = ("\n".join(parts) + "\n").encode()
    chars, starts = R.parse_records(text)
    batch = sg.ReadBatch(torch.from_numpy(np.ascontiguousarray(chars, np.uint8).copy()).cuda(),
                         torch.from_numpy(np.asarray(starts, np.int64).copy()).cuda(), 31,
                         host_starts=np.asarray(starts, np.int64))
    want = sg.to_u64(ix.search_reads(batch))
    lens = np.diff(np.asarray(starts, np.int64))
    for b in (0, 2_000):
        got, stats = _run(ix, text, tmp_path, records=True, batch_chars=b, output=output)
        dec = sg.read_binary_results(got)
        np.testing.assert_array_equal(dec["row_kmers"], np.maximum(lens - 30, 0))
        np.testing.assert_array_equal(dec["values"], want if output == "nodes" else want != sg.NONE)
        assert stats["found"] == int((want != sg.NONE).sum())
