"""FASTA / FASTQ records -> rows: a restatement for the tests (test
infrastructure only; the library's sbwt_parse_records has no reference
counterpart). Written with Python's own line splitting rather than the C
scanner's byte walk:

* lines end in "\\n" or "\\r\\n"; the first non-empty line picks the format;
* '>': FASTA, one row per '>' line holding the following lines joined;
* '@': FASTQ, the lines up to the '+' line are the sequence, then lines
  holding as many characters as the sequence are its quality (skipped);
* otherwise the reference's line parser (oracle.parse_lines).

Pinned in tests/test_host_io.py by hand-written records (multi-line,
CRLF, empty records, quality lines starting with '@' and '+').
"""
from __future__ import annotations

import numpy as np

from . import oracle as O


class RecordError(ValueError):
    pass


def _lines(text: bytes) -> list[bytes]:
    lines = text.split(b"\n")
    if text.endswith(b"\n"):
        lines = lines[:-1]
    return [ln[:-1] if ln.endswith(b"\r") else ln for ln in lines]


def parse_records(text: bytes):
    """(chars uint8, row_starts uint64) like oracle.parse_lines."""
    lines = _lines(text)
    first = next((ln for ln in lines if ln), b"")
    if first[:1] not in (b">", b"@"):
        return O.parse_lines(text)
    rows: list[bytes] = []
    if first[:1] == b">":
        cur = None
        for ln in lines:
            if ln.startswith(b">"):
                cur = []
                rows.append(cur)
            elif ln and cur is not None:
                cur.append(ln)
        rows = [b"".join(r) for r in rows]
    else:
        i = 0
        while i < len(lines):
            if not lines[i]:
                i += 1
                continue
            if not lines[i].startswith(b"@"):
                raise RecordError(f"record {len(rows) + 1} does not start with '@'")
            i += 1
            seq = []
            while i < len(lines) and not lines[i].startswith(b"+"):
                seq.append(lines[i])
                i += 1
            if i == len(lines):
                raise RecordError("no '+' line")
            i += 1
            s = b"".join(seq)
            q = 0
            while q < len(s) and i < len(lines):
                q += len(lines[i])
                i += 1
            if q != len(s):
                raise RecordError("quality length differs from the sequence length")
            rows.append(s)
    starts = np.zeros(len(rows) + 1, np.uint64)
    starts[1:] = np.cumsum([len(r) for r in rows]) if rows else []
    chars = np.frombuffer(b"".join(rows), np.uint8) if rows else np.zeros(0, np.uint8)
    return chars, starts
