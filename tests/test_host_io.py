"""Host-side ingest and result text of the library (no GPU needed) against
the oracle's restatement of the reference's parser and printer."""
import os

import numpy as np
import pytest

from oracle import oracle as O

LIB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sbwt_gpu_amd",
                   "libsbwt_gpu_amd.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")


def _texts():
    rng = np.random.default_rng(0)
    out = [b"", b"\n", b">only header\n", b"ACGT", b"ACGT\n", b">h\nAC\r\nGT\n\n>x\nTTTT", b"\n\n\nA\n"]
    for _ in range(20):
        lines = []
        for _ in range(rng.integers(0, 50)):
            r = rng.random()
            if r < 0.1:
                lines.append(b">" + bytes(rng.integers(65, 90, 5).astype(np.uint8)))
            elif r < 0.15:
                lines.append(b"")
            else:
                lines.append(bytes(np.frombuffer(b"ACGTNacgt", np.uint8)[rng.integers(0, 9, rng.integers(1, 200))]))
        t = b"\n".join(lines)
        if rng.random() < 0.5:
            t += b"\n"
        out.append(t)
    return out


@pytest.mark.parametrize("text", _texts())
def test_parse_lines_matches_oracle(text):
    import sbwt_gpu_amd as sg

    c1, r1 = sg.parse_lines(text)
    c2, r2 = O.parse_lines(text)
    assert bytes(c1) == bytes(c2)
    assert list(map(int, r1)) == list(map(int, r2))


@pytest.mark.parametrize("seed", [0, 1])
def test_parse_lines_threads_match_oracle(seed):
    """Texts of 4 MiB and more are parsed on host threads, in pieces cut at
    line starts: the rows equal the oracle's, whatever lands on a cut
    ('>' lines, empty lines, CRLF, a last line without a newline), and a
    row capacity too small is still refused."""
    import ctypes as C

    import sbwt_gpu_amd as sg
    from sbwt_gpu_amd import _capi

    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b"ACGTNacgt", np.uint8)
    lines = []
    size = 0
    while size < (6 << 20):
        kind = rng.random()
        if kind < 0.02:
            ln = b">header " + bytes(alphabet[rng.integers(0, 9, 20)])
        elif kind < 0.04:
            ln = b""
        else:
            ln = bytes(alphabet[rng.integers(0, 9, int(rng.integers(1, 400)))])
        if rng.random() < 0.1:
            ln += b"\r"
        lines.append(ln)
        size += len(ln) + 1
    text = b"\n".join(lines) + (b"\n" if seed else b"")
    c1, r1 = sg.parse_lines(text)
    c2, r2 = O.parse_lines(text)
    assert bytes(c1) == bytes(c2)
    np.testing.assert_array_equal(np.asarray(r1, np.uint64), np.asarray(r2, np.uint64))
    L = _capi.load()
    rows = C.c_uint64()
    chars = np.zeros(len(text), np.uint8)
    rs = np.zeros(len(r2) - 1, np.uint64)  # one entry short
    with pytest.raises(sg.SbwtError, match="more than"):
        _capi.check(L.sbwt_parse_lines(text, len(text), C.c_void_p(chars.ctypes.data), C.c_void_p(rs.ctypes.data),
                                       rs.size, C.byref(rows)))


def test_format_results_matches_oracle():
    import sbwt_gpu_amd as sg

    rng = np.random.default_rng(1)
    for rows in (0, 1, 7, 3000):
        counts = rng.integers(0, 130, rows)
        off = np.zeros(rows + 1, np.uint64)
        off[1:] = np.cumsum(counts)
        res = rng.integers(0, 2**40, int(off[-1]), dtype=np.uint64)
        res[rng.random(res.size) < 0.3] = np.uint64(0xFFFFFFFFFFFFFFFF)
        assert sg.format_results(res, off) == O.format_results(res, off)


# ---- FASTA / FASTQ records -------------------------------------------------

_RECORD_CASES = [
    # (text, rows)
    (b">r1\nACGT\nTTGA\n>r2\nGG\n", [b"ACGTTTGA", b"GG"]),
    (b">r1\r\nAC\r\nGT\r\n>r2\r\n\r\n>r3\r\nA", [b"ACGT", b"", b"A"]),
    (b"\n\n>r1\nAC\n\nGT\n", [b"ACGT"]),
    (b"@q1\nACGT\n+\nIIII\n@q2\nGG\n+q2\n@+\n", [b"ACGT", b"GG"]),
    # quality lines that start with '@' and '+', sequence and quality over two lines
    (b"@q1\nAC\nGT\n+\n@I\n+I\n@q2\nT\n+\n@\n", [b"ACGT", b"T"]),
    (b"@q1\r\nACGT\r\n+\r\nIIII\r\n\r\n@q2\r\n\r\n+\r\n\r\n", [b"ACGT", b""]),
    (b"ACGT\n>not a header here\nTT\n", None),  # not FASTA/FASTQ: the line parser
    (b"", []),
]


@pytest.mark.parametrize("text,rows", _RECORD_CASES)
def test_parse_records_cases(text, rows):
    import sbwt_gpu_amd as sg
    from oracle import records as R

    c1, r1 = sg.parse_records(text)
    c2, r2 = R.parse_records(text)
    assert bytes(c1) == bytes(c2)
    assert list(map(int, r1)) == list(map(int, r2))
    if rows is not None:
        got = [bytes(c1[int(r1[i]):int(r1[i + 1])]) for i in range(len(r1) - 1)]
        assert got == rows
    else:
        c3, r3 = O.parse_lines(text)
        assert bytes(c1) == bytes(c3) and list(map(int, r1)) == list(map(int, r3))


@pytest.mark.parametrize("bad", [b"@q\nACGT\n", b"@q\nACGT\n+\nII\n", b"@q\nAC\n+\nII\nACGT\n"])
def test_parse_records_rejects_malformed_fastq(bad):
    import sbwt_gpu_amd as sg
    from oracle import records as R

    with pytest.raises(sg.SbwtError):
        sg.parse_records(bad)
    with pytest.raises(R.RecordError):
        R.parse_records(bad)


@pytest.mark.parametrize("fmt,seed", [(f, s) for f in ("fasta", "fastq") for s in range(6)])
def test_parse_records_random(fmt, seed):
    """Random records wrapped at random widths, LF or CRLF: rows are the
    records' sequences, as the restatement finds them."""
    import sbwt_gpu_amd as sg
    from oracle import records as R

    rng = np.random.default_rng(seed)
    nl = b"\r\n" if seed % 2 else b"\n"
    seqs, parts = [], []
    for i in range(int(rng.integers(1, 40))):
        s = bytes(np.frombuffer(b"ACGTNacgt", np.uint8)[rng.integers(0, 9, int(rng.integers(0, 300)))])
        seqs.append(s)
        w = int(rng.integers(1, 80))
        wrap = [s[j:j + w] for j in range(0, len(s), w)] or ([b""] if fmt == "fastq" else [])
        if fmt == "fasta":
            parts.append(b">rec%d" % i)
            parts.extend(wrap)
        else:
            qual = bytes(np.frombuffer(b"@+I#5", np.uint8)[rng.integers(0, 5, len(s))])
            parts.append(b"@rec%d" % i)
            parts.extend(wrap)
            parts.append(b"+")
            parts.extend([qual[j:j + w] for j in range(0, len(qual), w)] or [b""])
    text = nl.join(parts) + (nl if rng.random() < 0.5 else b"")
    c1, r1 = sg.parse_records(text)
    c2, r2 = R.parse_records(text)
    assert bytes(c1) == bytes(c2)
    assert list(map(int, r1)) == list(map(int, r2))
    got = [bytes(c1[int(r1[i]):int(r1[i + 1])]) for i in range(len(r1) - 1)]
    assert got == seqs


def test_read_binary_results_blocks():
    """The binary result decoder: blocks concatenate in row order; bitmaps
    are LSB-first per word and cut at each block's k-mer count; malformed
    files raise."""
    from sbwt_gpu_amd import read_binary_results

    head = b"SBWTRES1" + np.array([1, 31], "<u4").tobytes()
    b1 = np.array([2, 3, 1, 2, 10, 2**64 - 1, 7], "<u8").tobytes()
    b2 = np.array([1, 0, 0], "<u8").tobytes()
    b3 = np.array([1, 1, 1, 5], "<u8").tobytes()
    d = read_binary_results(head + b1 + b2 + b3)
    assert d["kind"] == "nodes" and d["k"] == 31
    np.testing.assert_array_equal(d["row_kmers"], [1, 2, 0, 1])
    np.testing.assert_array_equal(d["values"], np.array([10, 2**64 - 1, 7, 5], np.uint64))
    bhead = b"SBWTRES1" + np.array([2, 25], "<u4").tobytes()
    bits = np.array([2, 65, 64, 1], "<u8").tobytes() + np.array([1 | (1 << 63), 1], "<u8").tobytes()
    d = read_binary_results(bhead + bits)
    want = np.zeros(65, bool)
    want[[0, 63, 64]] = True
    np.testing.assert_array_equal(d["values"], want)
    for bad in (b"SBWTRES2" + head[8:], head + b1[:-8], head + np.array([1, 4, 3], "<u8").tobytes() + b"\0" * 32,
                b"SBWTRES1" + np.array([3, 31], "<u4").tobytes()):
        with pytest.raises(ValueError):
            read_binary_results(bad)
