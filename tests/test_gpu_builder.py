"""The GPU SBWT builder on a pangenome-style input, against the oracle's
builder: 20 haplotypes of a 1 Mbp base genome at 0.1% SNP and 0.01% indels
(1-10 bp), the shape of the 20-haplotype large-index workload at 1/150 of
its size. The rows must equal orc_sbwt_build's bit for bit, and 1M queries
(half sampled from the haplotypes, half uniform random) must get the
oracle's answers. The same input is built again with the bucket sorts cut
to small chunks, which drives the builder through several partition levels
and the bucket seams; tiny inputs with one-item chunks reach the
all-keys-equal buckets."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("needs a GPU", allow_module_level=True)

import sbwt_gpu_amd as sg  # noqa: E402
from sbwt_gpu_amd import _capi, synth  # noqa: E402

NONE = np.uint64(0xFFFFFFFFFFFFFFFF)
CODE = np.full(256, 255, np.uint8)
for i, c in enumerate(b"ACGT"):
    CODE[c] = i


def top_kmers(chars: np.ndarray, pos: np.ndarray, k: int) -> np.ndarray:
    """Top-aligned 2-bit k-mers at character positions of an ACGT array."""
    codes = CODE[chars]
    out = np.zeros(pos.size, np.uint64)
    for t in range(k):
        out |= codes[pos + t].astype(np.uint64) << np.uint64(62 - 2 * t)
    return out


@pytest.fixture(scope="module")
def pangenome():
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev)
    gen.manual_seed(71)
    base = torch.randint(0, 4, (1_000_000,), dtype=torch.uint8, device=dev, generator=gen)
    chars, starts = synth.gpu_haplotypes(base, 20, seed=72)
    torch.cuda.synchronize()
    chars_h = chars.cpu().numpy()
    starts_h = starts.cpu().numpy()
    seqs = [chars_h[starts_h[i]:starts_h[i + 1]].tobytes() for i in range(20)]
    ref = O.SBWT(seqs, 31)
    return chars, starts, chars_h, starts_h, ref


def _build(chars, starts, k, m=12, chunk=0):
    L = _capi.load()
    _capi.check(L.sbwt_build_set_sort_chunk(chunk))
    try:
        return sg.Index.build(chars, k, sg.Config(presearch_depth=m), starts=starts)
    finally:
        _capi.check(L.sbwt_build_set_sort_chunk(0))


def _rows_equal(ix, ref):
    rows = ix.rows()
    assert ix.n == ref.n and ix.info["C"] == list(map(int, ref.C))
    for c in range(4):
        np.testing.assert_array_equal(rows[c][: ref.nwords], ref.rows[c], err_msg=f"row {c}")
    np.testing.assert_array_equal(rows[4][: ref.nwords], ref.group, err_msg="group row")


def test_pangenome_rows_and_answers_equal_the_oracle(pangenome):
    chars, starts, chars_h, starts_h, ref = pangenome
    k = 31
    ix = _build(chars, starts, k)
    assert ix.info["has_groups"] == 1
    _rows_equal(ix, ref)
    # 1M queries: 500K at uniform positions of uniform haplotypes, 500K random
    rng = np.random.default_rng(73)
    lens = np.diff(starts_h)
    hap = rng.integers(0, lens.size, 500_000)
    pos = starts_h[:-1][hap] + (rng.random(500_000) * (lens[hap] - k + 1)).astype(np.int64)
    present = top_kmers(chars_h, pos, k)
    rnd = rng.integers(0, 2**62, 500_000, dtype=np.uint64) << np.uint64(2)
    q = np.empty(1_000_000, np.uint64)
    q[0::2], q[1::2] = present, rnd
    got = sg.to_u64(ix.search_kmers(torch.from_numpy(q.view(np.int64)).cuda()))
    oix = O.Index(ref.rows, ref.n, sb=1024)
    oix.presearch(10)
    want = oix.search(q, k)
    np.testing.assert_array_equal(got, want)
    assert (got[0::2] != NONE).all(), "a sampled haplotype k-mer is absent"
    assert ix.info["n"] == ref.n and ref.num_kmers > 1_000_000


@pytest.mark.parametrize("chunk", [1 << 14, 3001])
def test_pangenome_small_sort_chunks_equal_the_oracle(pangenome, chunk):
    chars, starts, _, _, ref = pangenome
    ix = _build(chars, starts, 31, m=8, chunk=chunk)
    _rows_equal(ix, ref)


@pytest.mark.parametrize("seqs,k", [(["ACGTACGTTGCA"], 3), (["AAAAAAA"], 3), (["ACGTNACGTTTACG", "acgtacgt"], 3),
                                    (["TTTT", "AC"], 5), (["CGTAC" * 4], 5), (["ACGT" * 10, "", "GGGGCCCC"], 1),
                                    (["ACGTTGCA" * 9], 32), (["A" * 400, "C" * 300 + "A" * 40], 9)])
@pytest.mark.parametrize("chunk", [1, 2, 7])
def test_tiny_inputs_with_tiny_sort_chunks(seqs, k, chunk):
    ref = O.SBWT(seqs, k)
    enc = [s.encode() for s in seqs]
    st = np.zeros(len(enc) + 1, np.int64)
    st[1:] = np.cumsum([len(s) for s in enc])
    chars = torch.from_numpy(np.frombuffer(b"".join(enc) or b"A", np.uint8)[: max(int(st[-1]), 1)].copy()).cuda()
    ix = _build(chars, torch.from_numpy(st).cuda(), k, m=min(k, 3), chunk=chunk)
    _rows_equal(ix, ref)


def test_many_short_sequences_share_padded_nodes():
    """Thousands of sources whose padded nodes collide: the padded-node sort
    and its deduplication at small chunks."""
    rng = np.random.default_rng(74)
    common = "".join("ACGT"[c] for c in rng.integers(0, 4, 20))
    seqs = [common + "".join("ACGT"[c] for c in rng.integers(0, 4, int(rng.integers(0, 30))))
            for _ in range(3000)]
    ref = O.SBWT(seqs, 15)
    for chunk in (0, 500):
        enc = [s.encode() for s in seqs]
        st = np.zeros(len(enc) + 1, np.int64)
        st[1:] = np.cumsum([len(s) for s in enc])
        chars = torch.from_numpy(np.frombuffer(b"".join(enc), np.uint8).copy()).cuda()
        ix = _build(chars, torch.from_numpy(st).cuda(), 15, m=6, chunk=chunk)
        _rows_equal(ix, ref)
