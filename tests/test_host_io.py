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
