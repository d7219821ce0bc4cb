"""The oracle against the golden fixtures (tests/golden, written from the
pure-Python brute force by tests/golden/make_golden.py): SBWT rows, group
row and C, every 5-mer's answer of a 64 bp index, the presearch tables and
2,000 query answers of a 1 kbp index at k = 31, read answers and result text,
and a packing known-answer table."""
import numpy as np
import pytest

import golden_fixtures as gf
from oracle import oracle as O


@pytest.mark.parametrize("fx", [gf.sbwt_64bp_k5, gf.sbwt_1kbp_k31])
def test_oracle_builder_gives_the_fixture_rows(fx):
    f = fx()
    s = O.SBWT([f["seq"]], f["k"])
    assert s.n == f["n"] and list(map(int, s.C)) == list(map(int, f["C"]))
    for c in range(4):
        np.testing.assert_array_equal(s.rows[c], f["row_words"][c], err_msg=f"row {c}")
    np.testing.assert_array_equal(s.group, f["group_words"])


@pytest.mark.parametrize("m", [0, 2, 5])
@pytest.mark.parametrize("sb", [256, 1024])
def test_oracle_answers_every_kmer_of_the_64bp_index(m, sb):
    f = gf.sbwt_64bp_k5()
    oix = O.Index(f["row_words"], f["n"], sb=sb)
    oix.presearch(m)
    np.testing.assert_array_equal(oix.search(f["codes"], f["k"]), f["want"])


def test_oracle_presearch_tables_and_queries_of_the_1kbp_index():
    f = gf.sbwt_1kbp_k31()
    oix = O.Index(f["row_words"], f["n"], sb=1024)
    for m in (5, 8):
        left, right = oix.presearch(m)
        np.testing.assert_array_equal(left, f[f"pre{m}_left"].astype(np.uint64), err_msg=f"m={m} left")
        np.testing.assert_array_equal(right, f[f"pre{m}_right"].astype(np.uint64), err_msg=f"m={m} right")
        np.testing.assert_array_equal(oix.search(f["queries"], 31), f["answers"], err_msg=f"m={m}")
    assert (f["answers"][:1000] != gf.NONE).all() and (f["answers"][1000:] == gf.NONE).mean() > 0.99


def test_oracle_read_answers_and_text():
    f = gf.sbwt_1kbp_k31()
    r = gf.reads_1kbp_k31()
    oix = O.Index(f["row_words"], f["n"])
    oix.presearch(8)
    chars, rs = O.parse_lines("\n".join(r["rows"]).encode())
    got = O.search_reads(oix, chars, rs, r["k"])
    np.testing.assert_array_equal(got, r["want"])
    assert O.format_results(got, O.kmer_offsets(rs, r["k"])).decode() == r["text"]


def test_oracle_packing_known_answers():
    for e in gf.pack_kat():
        raw = e["raw"] + b"\0" * (-len(e["raw"]) % 32)
        got = O.pack(raw)
        nw = (len(e["raw"]) + 31) // 32
        np.testing.assert_array_equal(got[:nw], e["packed"][:nw], err_msg=repr(e["raw"][:20]))
