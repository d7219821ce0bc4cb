"""Index images (sbwt_index_save / sbwt_index_load): a saved and reloaded
index holds the same rows, C, presearch table and configuration, and answers
like the oracle; damaged or foreign files are refused."""
import numpy as np
import pytest

from oracle import oracle as O
from test_gpu_parity import dev_u64, genome, oracle_index, sample_queries, topk

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("needs a GPU", allow_module_level=True)

import sbwt_gpu_amd as sg  # noqa: E402

HEADER = 128


@pytest.fixture(scope="module")
def small():
    rng = np.random.default_rng(5)
    g = genome(rng, 150_000)
    k = 31
    s = O.SBWT([g], k)
    q = topk(sample_queries(rng, g, k, 20_000))
    return s, g, q, k


def _same_index(a, b):
    ia, ib = a.info, b.info
    for key in ("n", "C", "k", "presearch_depth", "layout", "has_groups", "device_bytes"):
        assert ia[key] == ib[key], key
    for x, y in zip(a.rows(), b.rows()):
        np.testing.assert_array_equal(x, y)
    for x, y in zip(a.presearch_table(), b.presearch_table()):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("layout,m,groups", [(16, 9, True), (32, 0, True), (64, 12, False)])
def test_image_round_trip(tmp_path, small, layout, m, groups):
    s, g, q, k = small
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group if groups else None,
                            config=sg.Config(presearch_depth=m, layout=layout))
    path = str(tmp_path / "ix.img")
    ix.save(path)
    back = sg.Index.load(path)
    _same_index(ix, back)
    want = oracle_index(s, m).search(q, k)
    np.testing.assert_array_equal(sg.to_u64(back.search_kmers(dev_u64(q))), want)
    if groups:
        batch = sg.ReadBatch.from_host([g[1000:9000], g[50_000:50_200] + "N" + g[60_000:60_300]], k)
        np.testing.assert_array_equal(sg.to_u64(back.search_reads(batch)), sg.to_u64(ix.search_reads(batch)))
    # another configuration on load: the rows are the image's, the layout and table the caller's
    other = sg.Index.load(path, sg.Config(presearch_depth=7, layout=32 if layout != 32 else 16))
    assert other.info["presearch_depth"] == 7 and other.info["C"] == ix.info["C"]
    np.testing.assert_array_equal(sg.to_u64(other.search_kmers(dev_u64(q))), want)
    for i in (ix, back, other):
        i.close()


def test_image_of_a_built_index(tmp_path):
    """An index built on the GPU keeps its k-mer count and suffix groups."""
    rng = np.random.default_rng(9)
    g = genome(rng, 80_000)
    ix = sg.Index.build([g, g[:5000][::-1]], 25, sg.Config(presearch_depth=8))
    path = str(tmp_path / "b.img")
    ix.save(path)
    back = sg.Index.load(path)
    _same_index(ix, back)
    np.testing.assert_array_equal(sg.to_u64(back.enumerate_kmers()), sg.to_u64(ix.enumerate_kmers()))
    ix.close()
    back.close()


def test_image_several_staging_chunks(tmp_path):
    """Rows of 144 MB cross three 64 MiB staging buffers (both buffers
    reused). The rows hold one bit in every column but the first, so
    C[4] = n, as the loader checks."""
    n = 230_000_000
    rng = np.random.default_rng(1)
    which = rng.integers(0, 4, n, dtype=np.uint8)
    which[0] = 4
    nw = ((n + 4095) // 4096) * 64
    rows = []
    for c in range(4):
        bits = np.packbits(which == c)
        w = np.zeros(nw * 8, np.uint8)
        w[:bits.size] = bits
        rows.append(w.view(">u8").astype(np.uint64))
    del which
    ix = sg.Index.from_rows(rows, n, 31, config=sg.Config(presearch_depth=6))
    path = str(tmp_path / "big.img")
    ix.save(path)
    back = sg.Index.load(path)
    _same_index(ix, back)
    q = rng.integers(0, 1 << 62, 100_000, dtype=np.uint64)
    np.testing.assert_array_equal(sg.to_u64(back.search_kmers(dev_u64(q))), sg.to_u64(ix.search_kmers(dev_u64(q))))
    ix.close()
    back.close()


def test_damaged_images_are_refused(tmp_path, small):
    s, _, _, k = small
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=6))
    path = tmp_path / "ix.img"
    ix.save(str(path))
    good = path.read_bytes()

    def refused(data, match):
        bad = tmp_path / "bad.img"
        bad.write_bytes(data)
        with pytest.raises(sg.SbwtError, match=match):
            sg.Index.load(str(bad))

    # one flipped bit in the rows, two rows words exchanged
    flip = bytearray(good)
    flip[HEADER + 1000] ^= 0x10
    refused(bytes(flip), "checksum")
    swap = bytearray(good)
    a, b = HEADER + 8 * 10, HEADER + 8 * 11
    swap[a:a + 8], swap[b:b + 8] = good[b:b + 8], good[a:a + 8]
    if swap != bytearray(good):
        refused(bytes(swap), "checksum")
    refused(good[:-8], "holds")
    refused(good + b"\0" * 8, "holds")
    refused(b"SBWTIMG2" + good[8:], "not an index image")
    refused(good[:64], "not an index image")
    # a header whose C disagrees with the rows
    hdr = bytearray(good)
    c1 = int.from_bytes(hdr[40:48], "little")
    hdr[40:48] = (c1 + 1).to_bytes(8, "little")
    refused(bytes(hdr), "C\\[1\\]")
    # an n the rows' size does not fit
    hdr = bytearray(good)
    hdr[16:24] = (s.n + 10_000).to_bytes(8, "little")
    refused(bytes(hdr), "row size")
    with pytest.raises(sg.SbwtError, match="can not be opened"):
        sg.Index.load(str(tmp_path / "missing.img"))
    ix.close()


@pytest.mark.parametrize("k", [9, 15, 32])
def test_image_round_trip_k(tmp_path, k):
    """Images of indexes at the edges of k (short k-mers, the largest k)
    reload with their k and answer as before."""
    rng = np.random.default_rng(k)
    g = genome(rng, 40_000)
    s = O.SBWT([g], k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=min(k, 8)))
    path = str(tmp_path / "k.img")
    ix.save(path)
    back = sg.Index.load(path)
    assert back.k == k
    _same_index(ix, back)
    batch = sg.ReadBatch.from_host([g[100:400], g[5000:5100] + "ACGTTGCA" + g[7000:7100]], k)
    np.testing.assert_array_equal(sg.to_u64(back.search_reads(batch)), sg.to_u64(ix.search_reads(batch)))
    ix.close()
    back.close()



def test_index_digest_tracks_every_region(small):
    """sbwt_index_digest: unchanged by searches and by rebuilding a region
    from the same rows; changed exactly where the layout, the table depth or
    a row changes."""
    s, g, q, k = small
    m = 9
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=m, layout=16))
    d0 = ix.digest()
    assert d0["table"] != 0
    ix.search_kmers(dev_u64(q))
    torch.cuda.synchronize()
    assert ix.digest() == d0
    ix.presearch(m)
    ix.set_layout(16)
    assert ix.digest() == d0  # rebuilt from the same rows: the same words
    ix.set_layout(32)
    d1 = ix.digest()
    assert d1["rows"] == d0["rows"] and d1["table"] == d0["table"] and d1["chunks"] != d0["chunks"]
    ix.presearch(m - 1)
    d2 = ix.digest()
    assert d2["rows"] == d0["rows"] and d2["chunks"] == d1["chunks"] and d2["table"] != d0["table"]
    ix.presearch(0)
    assert ix.digest()["table"] == 0
    # one bit moved inside a row (same C): another digest of the rows
    rows = [r.copy() for r in s.rows]
    w = next(i for i in range(rows[0].size // 2) if int(rows[0][i]) & ~(int(rows[0][i]) - 1) > 1)
    x = int(rows[0][w])
    lo = (x & -x).bit_length() - 1  # lowest set bit, > 0: bit lo - 1 is clear
    rows[0][w] = np.uint64(x ^ (1 << lo) ^ (1 << (lo - 1)))
    other = sg.Index.from_rows(rows, s.n, k, config=sg.Config(presearch_depth=0, layout=16))
    base = sg.Index.from_rows(s.rows, s.n, k, config=sg.Config(presearch_depth=0, layout=16))
    assert other.digest()["rows"] != base.digest()["rows"]
