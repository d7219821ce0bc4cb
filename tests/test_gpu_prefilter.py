"""The learned negative prefilter on the GPU: MLP gradients and one Adam step
against the float64 CPU oracle, matrix-core logits against the vector kernel
and the oracle, the Bloom filter, the fused prefilter, zero false negatives
over an index, and the masked read search against the plain one.

The MLP is floating point: gradients are compared at a relative tolerance of
2e-4 of the largest entry (fp32 sums over the batch against float64), logits
at 2e-3 (1 + |z|) (fp32 accumulation plus rare one-ulp differences of the
bf16-rounded first layer). Everything downstream of the filter (maybe flags,
search answers) is compared exactly."""
import numpy as np
import pytest

from oracle import mlp as om
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("needs a GPU", allow_module_level=True)

import sbwt_gpu_amd as sg  # noqa: E402
from sbwt_gpu_amd.learned_filter import train_prefilter  # noqa: E402

NONE = np.uint64(0xFFFFFFFFFFFFFFFF)
ACGT = np.array(list("ACGT"))


def dev_u64(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.uint64).view(np.int64)).cuda()


def rand_kmers(rng, n, k, p=None):
    c = rng.choice(4, size=(n, k), p=p).astype(np.uint64)
    x = np.zeros(n, np.uint64)
    for t in range(k):
        x |= c[:, t] << np.uint64(62 - 2 * t)
    return x


@pytest.mark.parametrize("k,B", [(31, 4096), (31, 1000), (32, 77), (9, 300)])
def test_mlp_gradient_matches_oracle(k, B):
    rng = np.random.default_rng(k * 1000 + B)
    km = rand_kmers(rng, B, k)
    y = (rng.random(B) < 0.5).astype(np.float32)
    p = om.he_params(k, 11 + B)
    nn = sg.Mlp(k, seed=1, max_batch=4096)
    nn.params = p
    g, loss = nn.grad(dev_u64(km), torch.from_numpy(y).cuda())
    g = g.cpu().numpy().astype(np.float64)
    want_loss, want = om.loss_and_grad(p, km, y, k)
    assert abs(loss.item() - want_loss) <= 1e-5 * (1 + abs(want_loss))
    tol = 2e-4 * np.max(np.abs(want))
    err = np.abs(g - want)
    assert err.max() <= tol, (int(err.argmax()), float(err.max()), tol)
    nn.close()


def test_mlp_adam_step_matches_oracle():
    k, B, lr = 31, 2048, 1e-3
    rng = np.random.default_rng(3)
    km = rand_kmers(rng, B, k)
    y = (rng.random(B) < 0.3).astype(np.float32)
    p = om.he_params(k, 4)
    nn = sg.Mlp(k, max_batch=B)
    nn.params = p
    nn.train_step(dev_u64(km), torch.from_numpy(y).cuda(), lr)
    got = nn.params.astype(np.float64)
    _, g = om.loss_and_grad(p, km, y, k)
    want, _, _ = om.adam_step(p.astype(np.float64), 0.0, 0.0, g, 1, lr)
    # the first Adam step moves each parameter by about lr * sign(g); where
    # |g| is tiny its sign is at the mercy of fp32 summation, so those
    # entries are only held to the step size
    big = np.abs(g) > 1e-3 * np.max(np.abs(g))
    assert big.mean() > 0.5
    np.testing.assert_allclose(got[big], want[big], atol=2e-6, rtol=0)
    assert np.all(np.abs(got - p) <= lr * 1.001 + 1e-7)
    nn.close()


@pytest.mark.parametrize("k", [31, 32, 20, 9])
def test_mlp_logits_matrix_cores_match_oracle(k):
    """The MFMA kernel's lane layout, the vector reference kernel and the
    oracle's inference forward agree."""
    rng = np.random.default_rng(k)
    n = 10_007
    km = rand_kmers(rng, n, k)
    p = om.he_params(k, 21)
    nn = sg.Mlp(k, max_batch=16)
    nn.params = p
    d = dev_u64(km)
    z = nn.logits(d).cpu().numpy().astype(np.float64)
    zr = nn.logits(d, reference=True).cpu().numpy().astype(np.float64)
    zo = om.logits(p, km, k)
    assert np.all(np.isfinite(z))
    tol = 2e-3 * (1 + np.abs(zo))
    assert np.all(np.abs(zr - zo) <= tol), float(np.max(np.abs(zr - zo)))
    assert np.all(np.abs(z - zo) <= tol), float(np.max(np.abs(z - zo)))
    # nearly all agree far more tightly: the layouts are right, not just close
    assert np.median(np.abs(z - zo)) < 1e-5 * (1 + np.median(np.abs(zo)))
    nn.close()


def test_mlp_logits_empty_and_tiny():
    nn = sg.Mlp(31, max_batch=16)
    p = om.he_params(31, 5)
    nn.params = p
    assert nn.logits(torch.empty(0, dtype=torch.int64, device="cuda")).numel() == 0
    km = rand_kmers(np.random.default_rng(1), 3, 31)
    z = nn.logits(dev_u64(km)).cpu().numpy()
    np.testing.assert_allclose(z, om.logits(p, km, 31), atol=2e-3 * (1 + np.abs(z).max()))
    nn.close()


def test_mlp_rejects_bad_arguments():
    with pytest.raises(sg.SbwtError):
        sg.Mlp(33)
    nn = sg.Mlp(31, max_batch=8)
    km = dev_u64(np.zeros(9, np.uint64))
    with pytest.raises(sg.SbwtError):
        nn.grad(km, torch.zeros(9, device="cuda"))
    with pytest.raises(sg.SbwtError):
        nn.params = np.zeros(5, np.float32)
    nn.close()


def test_mlp_training_lowers_loss():
    """A learnable task (GC-rich positives against uniform negatives): the
    GPU's own forward/backward/Adam kernels lower the loss and separate the
    classes."""
    k, B = 31, 4096
    rng = np.random.default_rng(8)
    pos = rand_kmers(rng, 100_000, k, p=[0.1, 0.4, 0.4, 0.1])
    neg = rand_kmers(rng, 100_000, k)
    x = dev_u64(np.concatenate([pos, neg]))
    y = torch.cat([torch.ones(100_000), torch.zeros(100_000)]).cuda()
    nn = sg.Mlp(k, seed=3, max_batch=B)
    gen = torch.Generator(device="cuda")
    gen.manual_seed(1)
    losses = []
    for step in range(300):
        idx = torch.randint(0, x.numel(), (B,), generator=gen, device="cuda")
        losses.append(nn.train_step(x[idx].contiguous(), y[idx].contiguous(), 2e-3).item())
    assert np.mean(losses[-20:]) < 0.5 * np.mean(losses[:5])
    z = nn.logits(x)
    acc = ((z >= 0).float() == y).float().mean().item()
    assert acc > 0.85
    nn.close()


@pytest.mark.parametrize("k,p", [(32, 0.01), (31, 0.001), (20, 0.05)])
def test_bloom_matches_oracle_bit_for_bit(k, p):
    """Geometry, every word of the filter and every query answer equal the
    numpy restatement; no false negatives; the false-positive rate is the
    one the sizing promises."""
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 1 << 63, 300_000, dtype=np.uint64) << np.uint64(1)
    other = rng.integers(0, 1 << 63, 300_000, dtype=np.uint64) << np.uint64(1)
    bl = sg.Bloom(k, keys.size, p)
    blocks, hashes = om.bloom_geometry(keys.size, p)
    assert bl.bytes == blocks * 128 and bl.hashes == hashes
    bl.insert(dev_u64(keys))
    want = om.bloom_build(keys, blocks, hashes, k)
    np.testing.assert_array_equal(bl.words(), want)
    assert bool((bl.query(dev_u64(keys)) == 1).all().item())
    got = bl.query(dev_u64(other)).cpu().numpy().astype(bool)
    np.testing.assert_array_equal(got, om.bloom_query(want, other, hashes, k))
    assert got.mean() <= 1.3 * p
    bl.close()


def test_bloom_query_reads_equals_extracted_kmers():
    """The fused read query answers what the k-mer query answers on the
    extracted k-mers, and what the numpy restatement answers, including rows
    shorter than k, non-ACGT characters and a 20 kbp row."""
    rng = np.random.default_rng(17)
    k = 31
    g = _genome(rng, 60_000)
    bl = sg.Bloom(k, len(g), 0.01)
    gk = sg.ReadBatch.from_host([g], k)
    bl.insert(gk.kmers())
    reads = _reads(rng, g, 200, 150) + _reads(rng, _genome(rng, 30_000), 200, 150)
    reads += ["ACG", "", g[100:131], g[5_000:25_000], "ACGTNNACGT" * 10, g[200:400].lower()]
    batch = sg.ReadBatch.from_host(reads, k)
    km = batch.kmers()
    want = bl.query(km)
    got = bl.query_reads(batch)
    assert torch.equal(got, want)
    words = bl.words()
    np.testing.assert_array_equal(got.cpu().numpy().astype(bool),
                                  om.bloom_query(words, sg.to_u64(km), bl.hashes, k))
    # every k-mer of the indexed sequence is in the filter
    assert bool((bl.query_reads(gk) == 1).all().item())
    bl.close()


def test_bloom_hashes_only_the_k_mer():
    """32-character windows (extraction) and k-mers with zeroed low bits
    (enumeration) are the same key; a filter for another k is refused."""
    k = 31
    rng = np.random.default_rng(12)
    km = rand_kmers(rng, 20_000, k)
    noisy = km | rng.integers(0, 4, km.size, dtype=np.uint64)
    bl = sg.Bloom(k, km.size, 0.01)
    bl.insert(dev_u64(km))
    assert bool((bl.query(dev_u64(noisy)) == 1).all().item())
    nn = sg.Mlp(k, max_batch=16)
    other = sg.Bloom(30, 10, 0.01)
    with pytest.raises(sg.SbwtError):
        sg.prefilter(nn, 0.0, other, dev_u64(km))
    assert bool((sg.prefilter(nn, 1e30, bl, dev_u64(noisy)) == 1).all().item())
    nn.close()
    bl.close()
    other.close()


def test_prefilter_fused_equals_parts():
    k = 31
    rng = np.random.default_rng(6)
    km = rand_kmers(rng, 50_001, k)
    nn = sg.Mlp(k, max_batch=16)
    nn.params = om.he_params(k, 7)
    d = dev_u64(km)
    z = nn.logits(d)
    thr = float(z.median().item())
    bl = sg.Bloom(k, 1000, 0.01)
    bl.insert(d[:1000])
    m = sg.prefilter(nn, thr, bl, d)
    want = ((z >= thr) | (bl.query(d) == 1)).to(torch.uint8)
    assert torch.equal(m, want)
    assert bool((m[:1000] == 1).all().item())
    m0 = sg.prefilter(nn, thr, None, d)
    assert torch.equal(m0, (z >= thr).to(torch.uint8))
    nn.close()
    bl.close()


def _genome(rng, n):
    return "".join(ACGT[rng.integers(0, 4, n)])


def _reads(rng, g, count, L, err=0.01):
    out = []
    for _ in range(count):
        p = int(rng.integers(0, len(g) - L))
        r = np.array(list(g[p:p + L]))
        flip = rng.random(L) < err
        r[flip] = ACGT[rng.integers(0, 4, int(flip.sum()))]
        out.append("".join(r))
    return out


@pytest.mark.parametrize("variant", [{}, {"certificates": False}, {"independent": True}, {"lanes": True},
                                     {"persistent": True, "independent": True}])
def test_masked_search_equals_plain_where_maybe(variant):
    rng = np.random.default_rng(9)
    g = _genome(rng, 100_000)
    k = 31
    ix = sg.Index.build([g], k, sg.Config(presearch_depth=10))
    reads = _reads(rng, g, 300, 150) + _reads(rng, _genome(rng, 50_000), 300, 150) + ["ACG", g[:31]]
    batch = sg.ReadBatch.from_host(reads, k)
    plain = sg.to_u64(ix.search_reads(batch, **variant))
    # any nonzero byte means "maybe"; the kernels read flags in aligned
    # 16-byte blocks, so also pass flags at every offset from a 16-byte
    # boundary, and in runs that span blocks
    vals = np.array([1, 2, 128, 255], np.uint8)
    mask = np.where(rng.random(plain.size) < 0.6, vals[rng.integers(0, 4, plain.size)], 0).astype(np.uint8)
    runs = np.repeat(rng.random(plain.size // 20 + 1) < 0.3, 20)[: plain.size].astype(np.uint8)
    for m in (mask, runs):
        want = np.where(m != 0, plain, NONE)
        for shift in (0, 1, 7, 15):
            store = torch.zeros(m.size + 32, dtype=torch.uint8, device="cuda")
            mt = store[shift: shift + m.size]
            mt.copy_(torch.from_numpy(m))
            got = sg.to_u64(ix.search_reads(batch, maybe=mt, **variant))
            np.testing.assert_array_equal(got, want, err_msg=f"flag offset {shift}")
    ones = torch.ones(plain.size, dtype=torch.uint8, device="cuda")
    np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, maybe=ones, **variant)), plain)


def test_trained_prefilter_zero_false_negatives_and_exact_answers():
    """Train on the index's own k-mers and search-verified negatives; every
    indexed k-mer passes the filter, and the filtered read search gives the
    plain search's answers exactly, which match the CPU oracle."""
    rng = np.random.default_rng(10)
    g = _genome(rng, 200_000)
    k = 31
    ix = sg.Index.build([g], k, sg.Config(presearch_depth=10))
    pf = train_prefilter(ix, positives=150_000, negatives=150_000, batch=4096, epochs=2, lr=2e-3, seed=3)
    assert pf.stats["zero_false_negatives_on_index"]
    allk = ix.enumerate_kmers()
    assert bool((pf.maybe(allk) == 1).all().item())
    reads = _reads(rng, g, 100, 150) + _reads(rng, _genome(rng, 100_000), 900, 150)
    batch = sg.ReadBatch.from_host(reads, k)
    plain = sg.to_u64(ix.search_reads(batch))
    maybe = pf.maybe(batch.kmers())
    filt = sg.to_u64(ix.search_reads(batch, maybe=maybe))
    np.testing.assert_array_equal(filt, plain)
    s = O.SBWT([g], k)
    oix = O.Index(s.rows, s.n)
    oix.presearch(10)
    chars, rs = O.parse_lines("\n".join(reads).encode())
    np.testing.assert_array_equal(plain, O.search_reads(oix, chars, rs, k))
    # the filter rejects a useful share of the negatives
    m = maybe.cpu().numpy()
    assert m[plain == NONE].mean() < 0.9


def test_learned_filter_gains_on_composition():
    """What the trained network gains over the filter without it, where
    there is something to learn: an index of a GC-rich genome queried with
    reads of an AT-rich one. The network learns the composition, so its
    Bloom filter only holds the indexed k-mers it rejects. The learned
    filter (network + that Bloom filter) keeps zero false negatives, passes
    few of the AT-rich k-mers, and is several times smaller than a Bloom
    filter over every indexed k-mer at the same false-positive rate. On
    uniform genomes there is no such signal (the network stays at chance)."""
    rng = np.random.default_rng(33)
    k = 31

    def biased(n, gc):
        p = [(1 - gc) / 2, gc / 2, gc / 2, (1 - gc) / 2]
        return "".join(ACGT[rng.choice(4, size=n, p=p)])

    g = biased(300_000, 0.72)
    ix = sg.Index.build([g], k, sg.Config(presearch_depth=10))
    pf = train_prefilter(ix, positives=250_000, negatives=250_000, batch=4096, epochs=3, lr=2e-3, seed=5)
    assert pf.stats["zero_false_negatives_on_index"]
    allk = ix.enumerate_kmers()
    assert bool((pf.maybe(allk) == 1).all().item())
    at_reads = sg.ReadBatch.from_host(_reads(rng, biased(200_000, 0.28), 2_000, 150), k)
    km = at_reads.kmers()
    assert bool((ix.search_kmers(km) == -1).all().item())  # no AT-rich k-mer is indexed
    learned_pass = float(pf.maybe(km).float().mean().item())
    whole = sg.Bloom(k, allk.numel(), 0.01)
    whole.insert(allk)
    whole_pass = float(whole.query(km).float().mean().item())
    learned_bytes = pf.stats["bloom_bytes"] + pf.mlp.param_count * 2  # bf16 weights
    print(f"learned filter: {learned_bytes} B, passes {learned_pass:.4f} of AT-rich k-mers; "
          f"Bloom over every indexed k-mer: {whole.bytes} B, passes {whole_pass:.4f}")
    assert learned_pass < 0.1, learned_pass
    assert learned_bytes * 2 < whole.bytes, (learned_bytes, whole.bytes)
    assert whole_pass < 0.05
    whole.close()


def test_masked_kmer_search_equals_plain_where_maybe():
    rng = np.random.default_rng(21)
    g = _genome(rng, 50_000)
    k = 31
    ix = sg.Index.build([g], k, sg.Config(presearch_depth=8))
    pos = rng.integers(0, len(g) - k, 3000)
    q = np.concatenate([rand_kmers(rng, 3000, k),
                        np.array([int("".join("0123"["ACGT".index(ch)] for ch in g[p:p + k]), 4) << (64 - 2 * k)
                                  for p in pos], np.uint64)])
    d = dev_u64(q)
    plain = sg.to_u64(ix.search_kmers(d))
    mask = (rng.random(q.size) < 0.5).astype(np.uint8)
    got = sg.to_u64(ix.search_kmers(d, maybe=torch.from_numpy(mask).cuda()))
    np.testing.assert_array_equal(got, np.where(mask == 1, plain, NONE))
    assert (plain[3000:] != NONE).all()


@pytest.mark.parametrize("case", ["uniform", "composition"])
def test_calibrated_prefilter_keeps_answers(case):
    """Calibration picks the cheapest of: the network at a quantile
    threshold, the Bloom filter alone, or no filter, from held-out k-mers
    and measured kernel costs. Whatever it picks keeps zero false negatives
    on the index and the plain search's answers. On a uniform genome the
    network is at chance (AUC near 0.5), so the network is not used; on a
    GC-rich index it separates the composition (AUC well above 0.5)."""
    rng = np.random.default_rng(40)
    k = 31
    if case == "uniform":
        g = _genome(rng, 200_000)
    else:
        p = [0.14, 0.36, 0.36, 0.14]
        g = "".join(ACGT[rng.choice(4, size=200_000, p=p)])
    ix = sg.Index.build([g], k, sg.Config(presearch_depth=10))
    pf = train_prefilter(ix, positives=150_000, negatives=150_000, batch=4096, epochs=2, lr=2e-3, seed=6,
                         calibrate_for=0.9)
    cal = pf.stats["calibration"]
    assert cal is not None and cal["best"] in cal["options"]
    assert min(o["cost"] for o in cal["options"]) == cal["best"]["cost"]
    assert pf.stats["network_used"] == (pf.mlp is not None) == (cal["best"]["option"] == "network")
    if case == "uniform":
        assert abs(cal["auc"] - 0.5) < 0.05, cal["auc"]
        assert not pf.stats["network_used"]
    else:
        assert cal["auc"] > 0.7, cal["auc"]
    allk = ix.enumerate_kmers()
    assert bool((pf.maybe(allk) == 1).all().item())
    reads = _reads(rng, g, 100, 150) + _reads(rng, _genome(rng, 100_000), 400, 150)
    batch = sg.ReadBatch.from_host(reads, k)
    plain = sg.to_u64(ix.search_reads(batch))
    km = batch.kmers()
    np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, maybe=pf.maybe(km))), plain)
    np.testing.assert_array_equal(sg.to_u64(ix.search_kmers(km, maybe=pf.maybe(km))), sg.to_u64(ix.search_kmers(km)))
