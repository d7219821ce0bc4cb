"""The two workloads the oracle can run in full, at their stated sizes:

* the 1 Mbp CPU-runnable case: a seeded uniform 1,000,000 bp genome, k = 31,
  presearch depth 8, 1,000,000 query 31-mers (half at uniform genome
  positions, half uniform random), against the oracle with poppy SB = 1024;
* the 100 Mbp read batch: the seeded 100 Mbp genome (GC 40%, 10% segmental
  copies at 1%), k = 31, presearch depth 12, 1,000,000 reads of 150 bp at 1%
  substitutions (120,000,000 k-mers).

The GPU-built rows must equal the oracle builder's, and every answer must
equal the oracle's."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a GPU", allow_module_level=True)

import sbwt_gpu_amd as sg  # noqa: E402
from sbwt_gpu_amd import synth  # noqa: E402

NONE = np.uint64(0xFFFFFFFFFFFFFFFF)


def _rows_equal(ix, ref):
    rows = ix.rows()
    assert ix.n == ref.n and ix.info["C"] == list(map(int, ref.C))
    for c in range(4):
        np.testing.assert_array_equal(rows[c][: ref.nwords], ref.rows[c], err_msg=f"row {c}")
    np.testing.assert_array_equal(rows[4][: ref.nwords], ref.group, err_msg="group row")


def test_one_mbp_uniform_one_million_queries():
    k, m = 31, 8
    g = synth.genome_codes(1_000_000, seed=101, gc=0.5)
    text = synth.to_ascii(g)
    ref = O.SBWT([text.tobytes()], k)
    ix = sg.Index.build([text.tobytes()], k, sg.Config(presearch_depth=m))
    _rows_equal(ix, ref)
    rng = np.random.default_rng(102)
    pos = rng.integers(0, g.size - k + 1, 500_000)
    present = np.zeros(pos.size, np.uint64)
    for t in range(k):
        present |= g[pos + t].astype(np.uint64) << np.uint64(62 - 2 * t)
    rnd = rng.integers(0, 2**62, 500_000, dtype=np.uint64) << np.uint64(2)
    q = np.concatenate([present, rnd])
    rng.shuffle(q)
    oix = O.Index(ref.rows, ref.n, sb=1024)
    oix.presearch(m)
    want = oix.search(q, k)
    got = sg.to_u64(ix.search_kmers(torch.from_numpy(q.view(np.int64)).cuda()))
    np.testing.assert_array_equal(got, want)
    assert (want != NONE).sum() >= 500_000


def test_hundred_mbp_read_batch_one_million_reads():
    k, m = 31, 12
    g = synth.genome_codes(100_000_000, seed=3, gc=0.4, seg_frac=0.10, seg_len=(1_000, 50_000), seg_mut=0.01)
    text = synth.to_ascii(g)
    dchars = torch.from_numpy(text).cuda()
    ix = sg.Index.build(dchars, k, sg.Config(presearch_depth=m),
                        starts=torch.tensor([0, g.size], dtype=torch.int64, device="cuda"))
    del dchars
    ref = O.SBWT([text.tobytes()], k)
    _rows_equal(ix, ref)
    rchars, rstarts = synth.reads(g, 1_000_000, 150, seed=4, sub=0.01)
    batch = sg.ReadBatch(torch.from_numpy(rchars).cuda(), torch.from_numpy(rstarts).cuda(), k, host_starts=rstarts)
    assert batch.total_kmers == 120_000_000
    got = sg.to_u64(ix.search_reads(batch))
    oix = O.Index(ref.rows, ref.n, sb=1024)
    oix.presearch(m)
    want = O.search_reads(oix, rchars, rstarts.astype(np.uint64), k)
    np.testing.assert_array_equal(got, want)
    # the reference's per-k-mer semantics give the same answers
    np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, independent=True)), want)
