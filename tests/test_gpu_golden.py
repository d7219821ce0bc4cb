"""The HIP path against the golden fixtures (tests/golden, written from the
pure-Python brute force): every 5-mer's answer of a 64 bp index on every
layout and kernel, with the group row given and derived; the presearch
tables and 2,000 query answers of a 1 kbp index at k = 31; its read answers
in streaming and per-k-mer form and the result text printed on the GPU; the
packing known-answer table, N mask included."""
import os

import numpy as np
import pytest

import golden_fixtures as gf

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a GPU", allow_module_level=True)

import sbwt_gpu_amd as sg  # noqa: E402


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a).view(np.int64)).cuda()


@pytest.mark.parametrize("given_group", [True, False])
@pytest.mark.parametrize("layout", [16, 32, 64])
def test_every_kmer_of_the_64bp_index(layout, given_group):
    f = gf.sbwt_64bp_k5()
    for m in (0, 2, 5):
        ix = sg.Index.from_rows(f["row_words"], f["n"], f["k"], group=f["group_words"] if given_group else None,
                                config=sg.Config(presearch_depth=m, layout=layout))
        assert ix.info["C"] == list(map(int, f["C"]))
        np.testing.assert_array_equal(ix.rows()[4][: f["group_words"].size], f["group_words"])
        q = dev(f["codes"])
        np.testing.assert_array_equal(sg.to_u64(ix.search_kmers(q)), f["want"], err_msg=f"m={m}")
        if layout >= 32:
            np.testing.assert_array_equal(sg.to_u64(ix.search_kmers(q, cooperative=True)), f["want"])
        ix.close()


def test_presearch_tables_and_queries_of_the_1kbp_index():
    f = gf.sbwt_1kbp_k31()
    for m in (5, 8):
        ix = sg.Index.from_rows(f["row_words"], f["n"], 31, config=sg.Config(presearch_depth=m))
        left, right = ix.presearch_table()
        np.testing.assert_array_equal(left, f[f"pre{m}_left"].astype(np.uint64), err_msg=f"m={m}")
        np.testing.assert_array_equal(right, f[f"pre{m}_right"].astype(np.uint64), err_msg=f"m={m}")
        np.testing.assert_array_equal(sg.to_u64(ix.search_kmers(dev(f["queries"]))), f["answers"])
        ix.close()


def test_gpu_builder_gives_the_fixture_rows():
    for f in (gf.sbwt_64bp_k5(), gf.sbwt_1kbp_k31()):
        ix = sg.Index.build([f["seq"]], f["k"], sg.Config(presearch_depth=4))
        rows = ix.rows()
        assert ix.n == f["n"] and ix.info["C"] == list(map(int, f["C"]))
        for c in range(4):
            np.testing.assert_array_equal(rows[c][: f["row_words"][c].size], f["row_words"][c])
        np.testing.assert_array_equal(rows[4][: f["group_words"].size], f["group_words"])
        ix.close()


def test_read_answers_and_gpu_text(tmp_path):
    f = gf.sbwt_1kbp_k31()
    r = gf.reads_1kbp_k31()
    ix = sg.Index.from_rows(f["row_words"], f["n"], 31, config=sg.Config(presearch_depth=8))
    batch = sg.ReadBatch.from_host(r["rows"], 31)
    np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch)), r["want"])
    np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, independent=True)), r["want"])
    path = tmp_path / "out.txt"
    fd = os.open(path, os.O_WRONLY | os.O_CREAT | os.O_TRUNC)
    try:
        st = sg.search_text(ix, "\n".join(r["rows"]).encode() + b"\n", fd=fd)
    finally:
        os.close(fd)
    assert st["kmers"] == r["want"].size
    assert path.read_text() == r["text"]
    ix.close()


def test_packing_known_answers():
    for e in gf.pack_kat():
        chars = torch.from_numpy(np.frombuffer(e["raw"], np.uint8).copy()).cuda()
        packed, nmask = sg.pack(chars)
        n = len(e["raw"])
        pw, mw = (n + 31) // 32, (n + 63) // 64
        np.testing.assert_array_equal(sg.to_u64(packed)[:pw], e["packed"][:pw], err_msg=repr(e["raw"][:20]))
        got = sg.to_u64(nmask)[:mw].copy()
        if n % 64:  # bits past the last character are not part of the mask (the GPU marks them)
            got[-1] &= np.uint64(((1 << (n % 64)) - 1) << (64 - n % 64))
        np.testing.assert_array_equal(got, e["nmask"][:mw], err_msg=repr(e["raw"][:20]))
