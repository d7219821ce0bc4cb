"""Bit-exact parity of the HIP path (through the C ABI) with the CPU oracle.

Every answer is an integer node id or ~0, so every comparison is exact."""
import itertools
import os

import numpy as np
import pytest

import brute
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover - CPU container
    pytest.skip("needs a GPU", allow_module_level=True)

import sbwt_gpu_amd as sg  # noqa: E402

NONE = np.uint64(0xFFFFFFFFFFFFFFFF)
ACGT = np.array(list("ACGT"))


def genome(rng, n, p=(0.25, 0.25, 0.25, 0.25)):
    return "".join(ACGT[rng.choice(4, size=n, p=p)])


def dev_u64(a: np.ndarray) -> "torch.Tensor":
    return torch.from_numpy(np.ascontiguousarray(a, np.uint64).view(np.int64)).cuda()


def oracle_index(s, m, sb=1024):
    ix = O.Index(s.rows, s.n, sb=sb)
    ix.presearch(m)
    return ix


def topk(kmers_str):
    return np.array([brute.encode_top(x) for x in kmers_str], np.uint64)


def sample_queries(rng, g, k, count):
    pos = rng.integers(0, len(g) - k + 1, count // 2)
    sampled = [g[p:p + k] for p in pos]
    rnd = ["".join(ACGT[rng.integers(0, 4, k)]) for _ in range(count - count // 2)]
    q = sampled + rnd
    rng.shuffle(q)
    return q


# ---------------------------------------------------------------------------
# index upload, C array, presearch table


@pytest.mark.parametrize("layout", [16, 32, 64])
def test_index_info_and_presearch_table(layout):
    rng = np.random.default_rng(5)
    g = genome(rng, 3000)
    k = 9
    s = O.SBWT([g], k)
    for m in (1, 4, 7):
        ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=m, layout=layout))
        info = ix.info
        assert info["n"] == s.n and info["C"] == list(map(int, s.C)) and info["layout"] == layout
        left, right = ix.presearch_table()
        ol, orr = oracle_index(s, m).presearch(m)
        np.testing.assert_array_equal(left, ol)
        np.testing.assert_array_equal(right, orr)
        ix.close()


@pytest.mark.parametrize("seqs,k", [(["ACGTACGTTGCAAGGCT", "TTGACCA"], 4), (["GATTACA", "CATTAG", "ACGTGTCA"], 3)])
def test_every_kmer_tiny_against_brute_force(seqs, k):
    order, _, _ = brute.sbwt(seqs, k)
    K = brute.kmer_set(seqs, k)
    s = O.SBWT(seqs, k)
    allk = ["".join(p) for p in itertools.product("ACGT", repeat=k)]
    want = np.array([brute.search(order, K, x) for x in allk], np.uint64)
    q = dev_u64(topk(allk))
    for layout in (16, 32, 64):
        for m in range(0, k + 1):
            ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=m, layout=layout))
            got = sg.to_u64(ix.search_kmers(q))
            np.testing.assert_array_equal(got, want, err_msg=f"layout={layout} m={m}")
            ix.close()


# ---------------------------------------------------------------------------
# search_kmers / search_positions against the oracle on the CPU-runnable case


@pytest.mark.parametrize("layout,m", [(16, 8), (32, 8), (64, 8), (16, 0), (16, 12), (32, 5)])
def test_search_kmers_one_mbp_scale(layout, m):
    rng = np.random.default_rng(100 + layout + m)
    g = genome(rng, 300_000)
    k = 31
    s = O.SBWT([g], k)
    q = topk(sample_queries(rng, g, k, 40_000))
    want = oracle_index(s, m).search(q, k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=m, layout=layout))
    got = sg.to_u64(ix.search_kmers(dev_u64(q)))
    np.testing.assert_array_equal(got, want)
    assert (got != NONE).sum() >= 20_000


def test_deepest_presearch_table():
    """Depth 16 (4^16 entries, the first table past a 32-bit index) and the
    largest depth, 17 (137 GB), answer like a shallow table; depth 18 is
    refused."""
    rng = np.random.default_rng(16)
    g = genome(rng, 50_000)
    k = 20
    s = O.SBWT([g], k)
    q = topk(sample_queries(rng, g, k, 4000))
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=6))
    want = sg.to_u64(ix.search_kmers(dev_u64(q)))
    np.testing.assert_array_equal(want, oracle_index(s, 6).search(q, k))
    ix.presearch(16)
    assert ix.info["presearch_depth"] == 16
    np.testing.assert_array_equal(sg.to_u64(ix.search_kmers(dev_u64(q))), want)
    batch = sg.ReadBatch.from_host([g[:5_000]], k)
    reads_want = sg.to_u64(ix.search_reads(batch))
    with pytest.raises(sg.SbwtError, match="presearch depth"):
        ix.presearch(18)
    ix.presearch(0)  # frees the depth-16 table
    torch.cuda.empty_cache()
    if torch.cuda.mem_get_info()[0] >= (8 << 34) + (16 << 30):
        ix.presearch(17)
        assert ix.info["presearch_depth"] == 17
        np.testing.assert_array_equal(sg.to_u64(ix.search_kmers(dev_u64(q))), want)
        np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch)), reads_want)
        ix.presearch(0)
    ix.close()


def test_random_rows_stress_rank():
    """Rows of random bits (total ones n-1, so C[4] = n): not an SBWT, but the
    narrowing is still well defined, and random k-mers drive rank through
    every chunk offset and chunk boundary of every layout."""
    rng = np.random.default_rng(9)
    n = 200_003
    col = rng.integers(0, 4, n - 1)  # one 1-bit per column for n-1 columns
    perm = rng.permutation(n)[: n - 1]
    rows = [np.zeros(((n + 4095) // 4096) * 64, np.uint64) for _ in range(4)]
    for c in range(4):
        pos = np.sort(perm[col == c])
        bits = np.zeros(len(rows[c]) * 64, np.uint8)
        bits[pos] = 1
        rows[c] = np.packbits(bits).view(">u8").astype(np.uint64)
    k = 12
    q = rng.integers(0, 2**63, 100_000, dtype=np.uint64) * np.uint64(2)
    for m in (0, 3, 6):
        oix = O.Index(rows, n)
        oix.presearch(m)
        want = oix.search(q, k)
        for layout in (16, 32, 64):
            ix = sg.Index.from_rows(rows, n, k, config=sg.Config(presearch_depth=m, layout=layout))
            got = sg.to_u64(ix.search_kmers(dev_u64(q)))
            np.testing.assert_array_equal(got, want, err_msg=f"m={m} layout={layout}")
            ix.close()


@pytest.mark.parametrize("layout", [32, 64])
def test_cooperative_kernel_matches_oracle(layout):
    """SBWT_SEARCH_COOPERATIVE: two or four lanes per k-mer. Random rows drive
    every chunk offset; an SBWT with half present queries, presearch depths 0
    and 8, a mask, a count that is no multiple of the group, and the counted
    variant (same lines as the one-lane kernel) must all answer alike."""
    rng = np.random.default_rng(200 + layout)
    n = 100_003
    col = rng.integers(0, 4, n - 1)
    perm = rng.permutation(n)[: n - 1]
    rows = []
    for c in range(4):
        bits = np.zeros(((n + 4095) // 4096) * 4096, np.uint8)
        bits[np.sort(perm[col == c])] = 1
        rows.append(np.packbits(bits).view(">u8").astype(np.uint64))
    q = rng.integers(0, 2**63, 50_001, dtype=np.uint64) * np.uint64(2)
    for m in (0, 4):
        oix = O.Index(rows, n)
        oix.presearch(m)
        ix = sg.Index.from_rows(rows, n, 12, config=sg.Config(presearch_depth=m, layout=layout))
        np.testing.assert_array_equal(sg.to_u64(ix.search_kmers(dev_u64(q), cooperative=True)), oix.search(q, 12),
                                      err_msg=f"random rows m={m}")
        ix.close()
    g = genome(rng, 200_000)
    k = 31
    s = O.SBWT([g], k)
    q = topk(sample_queries(rng, g, k, 30_001))
    for m in (0, 8):
        want = oracle_index(s, m).search(q, k)
        ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=m, layout=layout))
        dq = dev_u64(q)
        got = sg.to_u64(ix.search_kmers(dq, cooperative=True))
        np.testing.assert_array_equal(got, want, err_msg=f"m={m}")
        assert (got != NONE).sum() >= 15_000
        for cnt in (1, 2, 3, 5):  # groups cut by the end of the batch
            np.testing.assert_array_equal(sg.to_u64(ix.search_kmers(dq[:cnt], cooperative=True)), want[:cnt])
        maybe = torch.from_numpy((rng.random(q.size) < 0.7).astype(np.uint8)).cuda()
        masked = sg.to_u64(ix.search_kmers(dq, maybe=maybe, cooperative=True))
        np.testing.assert_array_equal(masked, np.where(maybe.cpu().numpy() == 1, want, NONE))
        ix.counters(reset=True)
        ix.search_kmers(dq, count=True)
        lane = ix.counters(reset=True)
        got_c = sg.to_u64(ix.search_kmers(dq, count=True, cooperative=True))
        coop = ix.counters(reset=True)
        np.testing.assert_array_equal(got_c, want)
        assert coop == lane and lane["kmers"] == q.size
        ix.close()


def test_search_positions_reference_abi():
    rng = np.random.default_rng(3)
    g = genome(rng, 100_000)
    k = 31
    s = O.SBWT([g], k)
    reads = [g[p:p + 150] if i % 3 else "".join(ACGT[rng.integers(0, 4, 100)])
             for i, p in enumerate(rng.integers(0, len(g) - 150, 500))]
    chars, rs = O.parse_lines("\n".join(reads).encode())
    raw = bytes(chars) + b"\0" * (32 - len(chars) % 32 + 32)
    packed_ref = O.pack(raw[: (len(raw) // 32) * 32])
    pos = O.positions(rs, k)
    oix = oracle_index(s, 10)
    want = oix.search(np.array([O.extract(packed_ref, int(p)) for p in pos], np.uint64), k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=10))
    dchars = torch.from_numpy(np.frombuffer(bytes(chars), np.uint8).copy()).cuda()
    packed, _ = sg.pack(dchars)
    m = min(packed.numel(), packed_ref.size)
    np.testing.assert_array_equal(sg.to_u64(packed)[:m], packed_ref[:m])
    batch = sg.ReadBatch(dchars, dev_u64(rs), k, host_starts=rs)
    dpos = batch.positions()
    np.testing.assert_array_equal(sg.to_u64(dpos), pos)
    got = sg.to_u64(ix.search_positions(packed, dpos))
    np.testing.assert_array_equal(got, want)


# ---------------------------------------------------------------------------
# read search: streaming / independent, N policies, ragged rows


def _reads(rng, g, n_reads, err=0.01, nrate=0.0, lens=(150, 150)):
    out = []
    for _ in range(n_reads):
        L = int(rng.integers(lens[0], lens[1] + 1))
        if L >= len(g):
            L = len(g) - 1
        p = int(rng.integers(0, len(g) - L))
        r = np.array(list(g[p:p + L]))
        flip = rng.random(L) < err
        r[flip] = ACGT[rng.integers(0, 4, flip.sum())]
        nn = rng.random(L) < nrate
        r[nn] = "N"
        out.append("".join(r))
    return out


READ_VARIANTS = [{}, {"certificates": False}, {"persistent": True}, {"persistent": True, "certificates": False},
                 {"independent": True}, {"independent": True, "persistent": True}, {"lanes": True},
                 {"lanes": True, "certificates": False}]


@pytest.mark.parametrize("layout", [16, 32, 64])
@pytest.mark.parametrize("seg", [256, 7, 1])
def test_search_reads_matches_oracle(layout, seg):
    rng = np.random.default_rng(layout * 10 + seg)
    g = genome(rng, 200_000, (0.3, 0.2, 0.2, 0.3))
    k = 31
    s = O.SBWT([g], k)
    reads = _reads(rng, g, 400, err=0.01, nrate=0.02, lens=(20, 400))
    reads += ["", "ACGT", "N" * 40, g[:31], g[:32] + "N" + g[33:80]]
    chars, rs = O.parse_lines("\n".join(reads).encode())
    oix = oracle_index(s, 10)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=10, layout=layout))
    batch = sg.ReadBatch(torch.from_numpy(np.frombuffer(bytes(chars), np.uint8).copy()).cuda(), dev_u64(rs), k,
                         seg=seg, host_starts=rs)
    for n_as_a in (False, True):
        want = O.search_reads(oix, chars, rs, k, n_breaks=not n_as_a)
        for kw in READ_VARIANTS:
            got = sg.to_u64(ix.search_reads(batch, n_as_a=n_as_a, **kw))
            np.testing.assert_array_equal(got, want, err_msg=f"n_as_a={n_as_a} {kw}")


@pytest.mark.parametrize("layout", [16, 32])
def test_search_reads_errors_near_row_end(layout):
    """Substitutions in a row's last k - 1 characters, where no k-mer starts
    at the dead string's end: every read kernel answers as the oracle, and
    with certificates the counted searches fall (one search of [rowk-1, d]
    answers the row's remaining k-mers)."""
    rng = np.random.default_rng(77 + layout)
    g = genome(rng, 120_000)
    k = 31
    s = O.SBWT([g], k)
    reads = []
    for _ in range(600):
        L = int(rng.integers(k, 200))
        p = int(rng.integers(0, len(g) - L))
        r = np.array(list(g[p:p + L]))
        for e in rng.integers(max(0, L - k + 1), L, int(rng.integers(1, 3))):
            r[e] = "ACGT"[("ACGT".index(r[e]) + int(rng.integers(1, 4))) % 4]
        if rng.random() < 0.2:  # and one near the start
            e = int(rng.integers(0, min(L, 10)))
            r[e] = "ACGT"[("ACGT".index(r[e]) + 1) % 4]
        reads.append("".join(r))
    chars, rs = O.parse_lines("\n".join(reads).encode())
    want = O.search_reads(oracle_index(s, 10), chars, rs, k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=10, layout=layout))
    batch = sg.ReadBatch(torch.from_numpy(np.frombuffer(bytes(chars), np.uint8).copy()).cuda(), dev_u64(rs), k,
                         seg=256, host_starts=rs)
    searches = {}
    for kw in READ_VARIANTS:
        np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, **kw)), want, err_msg=str(kw))
        ix.counters(reset=True)
        ix.search_reads(batch, count=True, **kw)
        torch.cuda.synchronize()
        searches[tuple(sorted(kw.items()))] = ix.counters(reset=True)["full_searches"]
    assert searches[()] < searches[(("certificates", False),)]
    assert searches[(("persistent", True),)] < searches[(("certificates", False), ("persistent", True))]
    assert searches[(("lanes", True),)] < searches[(("certificates", False), ("lanes", True))]


@pytest.mark.parametrize("e", [3, 75, 135])
def test_search_reads_same_error_position(e):
    """Every read holds one substitution at the same position, in a genome
    with segmental copies at 1%. The lanes of a wave reach their errors
    together; lanes whose certificate search runs into a copy search their
    k-mers one by one, and lanes that run out of work split those lanes'
    k-mers between them. Every read kernel answers as the oracle."""
    rng = np.random.default_rng(300 + e)
    base = genome(rng, 60_000)
    parts = [base]
    for _ in range(40):
        p = int(rng.integers(0, 59_000))
        seg = np.array(list(base[p:p + 1000]))
        flip = rng.random(1000) < 0.01
        seg[flip] = ACGT[rng.integers(0, 4, flip.sum())]
        parts.append("".join(seg))
    g = "".join(parts)
    k = 31
    s = O.SBWT([g], k)
    reads = []
    for _ in range(3000):
        p = int(rng.integers(0, len(g) - 150))
        r = np.array(list(g[p:p + 150]))
        r[e] = "ACGT"[("ACGT".index(r[e]) + int(rng.integers(1, 4))) % 4]
        reads.append("".join(r))
    chars, rs = O.parse_lines("\n".join(reads).encode())
    want = O.search_reads(oracle_index(s, 10), chars, rs, k)
    assert (want == np.uint64(~np.uint64(0))).sum() >= 3000  # each read's k-mers over the error are absent
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=10))
    batch = sg.ReadBatch(torch.from_numpy(np.frombuffer(bytes(chars), np.uint8).copy()).cuda(), dev_u64(rs), k,
                         seg=256, host_starts=rs)
    for kw in READ_VARIANTS:
        np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, **kw)), want, err_msg=str(kw))


@pytest.mark.parametrize("negative", [0.1, 0.5, 0.9])
def test_search_reads_mixed_negative_reads(negative):
    """Batches mixing reads from the index's genome (with substitutions,
    some near the read start) and reads from another genome, in several
    proportions: the certificate policy of the uniform kernel adapts to the
    mix, and every read kernel answers as the oracle."""
    rng = np.random.default_rng(int(negative * 100))
    g = genome(rng, 150_000)
    other = genome(rng, 150_000)
    k = 31
    s = O.SBWT([g], k)
    reads = []
    for _ in range(1500):
        src = other if rng.random() < negative else g
        L = int(rng.integers(40, 250))
        p = int(rng.integers(0, len(src) - L))
        r = np.array(list(src[p:p + L]))
        flip = rng.random(L) < 0.01
        r[flip] = ACGT[rng.integers(0, 4, flip.sum())]
        if rng.random() < 0.3:
            e = int(rng.integers(0, 12))
            r[e] = "ACGT"[("ACGT".index(r[e]) + 1) % 4]
        reads.append("".join(r))
    chars, rs = O.parse_lines("\n".join(reads).encode())
    want = O.search_reads(oracle_index(s, 11), chars, rs, k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=11))
    for seg in (256, 40):
        batch = sg.ReadBatch(torch.from_numpy(np.frombuffer(bytes(chars), np.uint8).copy()).cuda(), dev_u64(rs), k,
                             seg=seg, host_starts=rs)
        for kw in READ_VARIANTS:
            np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, **kw)), want, err_msg=f"{seg} {kw}")


def test_search_reads_unaligned_output():
    """Answers written through an output pointer that is 8- but not
    16-byte aligned equal the aligned ones, and nothing outside the array is
    written: the per-lane kernel stores such answers slot by slot, the
    uniform-step kernel passes the array one slot lower and keeps its 16-byte
    stores for whole groups."""
    rng = np.random.default_rng(31)
    g = genome(rng, 60_000)
    k = 31
    s = O.SBWT([g], k)
    reads = _reads(rng, g, 300, err=0.02, nrate=0.01, lens=(31, 200))
    chars, rs = O.parse_lines("\n".join(reads).encode())
    oix = oracle_index(s, 9)
    want = O.search_reads(oix, chars, rs, k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=9))
    batch = sg.ReadBatch.from_host(reads, k, seg=17)
    for kw in READ_VARIANTS:
        buf = torch.full((batch.total_kmers + 3,), 7, dtype=torch.int64, device="cuda")
        got = ix.search_reads(batch, out=buf[1:batch.total_kmers + 1], **kw)
        np.testing.assert_array_equal(sg.to_u64(got), want, err_msg=str(kw))
        assert int(buf[0]) == 7 and int(buf[-1]) == 7 and int(buf[-2]) == 7


def test_uniform_kernel_refuses_batches_that_may_hold_a_4g_row():
    """The uniform kernel keeps row-relative positions in u32: a forced
    uniform search whose work bound admits a row of 2^32 k-mers is refused
    before any launch (the default routes such a batch to the per-lane
    kernel)."""
    from sbwt_gpu_amd import _capi
    from sbwt_gpu_amd.api import _ptr, _stream

    s = O.SBWT(["ACGTACGTTGCAGGCATTACG" * 3], 9)
    ix = sg.Index.from_rows(s.rows, s.n, 9, group=s.group, config=sg.Config(presearch_depth=4))
    L = _capi.load()
    buf = torch.zeros(16, dtype=torch.int64, device="cuda")
    items = torch.zeros(16, dtype=torch.int32, device="cuda")
    p = _ptr(buf)
    st = L.sbwt_search_reads(ix._h, p, p, p, p, p, _ptr(items), 1, 1 << 32, 256, p, _capi.READS_PERSISTENT,
                             _stream())
    assert st == 5, L.sbwt_last_error()
    assert b"2^32" in L.sbwt_last_error()


GROUP_CASES = [
    (["ACGTACGTTGCA"], 3), (["AAAAAAA"], 3), (["ACGTNACGTTTACG", "acgtacgt"], 3), (["TTTT", "AC"], 5),
    (["CGTAC" * 4], 5), (["ACGT" * 10, "", "GGGGCCCC"], 1), (["ACGTTGCA" * 9], 32),
]


def _group_case(name, rng):
    """Inputs whose group rows hold long groups: segmental copies, many
    short sequences ('$'-padded nodes with shared suffixes), near-repeats."""
    if name == "segmental":
        g = genome(rng, 40_000)
        for _ in range(30):
            a, b, ln = int(rng.integers(0, 39_000)), int(rng.integers(0, 39_000)), int(rng.integers(50, 900))
            ln = min(ln, 40_000 - max(a, b))
            g = g[:b] + g[a:a + ln] + g[b + ln:]
        return [g], 31
    if name == "short_reads":
        base = genome(rng, 2000)
        return [base[p:p + int(rng.integers(5, 60))] for p in rng.integers(0, 1900, 800)], 21
    if name == "near_repeats":
        unit = genome(rng, 40)
        return ["".join(unit[:i] + "ACGT"[int(rng.integers(0, 4))] + unit[i + 1:] for i in range(0, 40, 3)) * 3], 15
    raise KeyError(name)


@pytest.mark.parametrize("seqs,k", GROUP_CASES)
def test_group_row_derived_tiny(seqs, k):
    """Without a group row the index derives it on the device: equal to the
    oracle builder's group row, and the read search streams."""
    s = O.SBWT(seqs, k)
    ix = sg.Index.from_rows(s.rows, s.n, k, config=sg.Config(presearch_depth=min(k, 3)))
    assert ix.info["has_groups"]
    np.testing.assert_array_equal(ix.rows()[4][: s.nwords], s.group)


@pytest.mark.parametrize("case", ["segmental", "short_reads", "near_repeats"])
@pytest.mark.parametrize("layout", [16, 64])
def test_group_row_derived_matches_oracle_and_answers(case, layout):
    rng = np.random.default_rng({"segmental": 81, "short_reads": 82, "near_repeats": 83}[case])
    seqs, k = _group_case(case, rng)
    s = O.SBWT(seqs, k)
    ix = sg.Index.from_rows(s.rows, s.n, k, config=sg.Config(presearch_depth=8, layout=layout))
    assert ix.info["has_groups"]
    np.testing.assert_array_equal(ix.rows()[4][: s.nwords], s.group)
    src = seqs[0] if len(seqs[0]) > 500 else "".join(seqs)
    reads = _reads(rng, src, 300, err=0.02, nrate=0.01, lens=(k, 200))
    chars, rs = O.parse_lines("\n".join(reads).encode())
    want = O.search_reads(oracle_index(s, 8), chars, rs, k)
    batch = sg.ReadBatch.from_host(reads, k)
    for kw in READ_VARIANTS:
        np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, **kw)), want, err_msg=str(kw))


def test_four_reference_files_give_the_streaming_index(tmp_path):
    """The reference's input is four bit-vector files and nothing else: the
    loaded index has the group row (equal to the oracle's) and its read
    search runs as fast as the built index's, streaming."""
    rng = np.random.default_rng(77)
    g = genome(rng, 200_000, (0.3, 0.2, 0.2, 0.3))
    k = 31
    s = O.SBWT([g], k)
    paths = [str(tmp_path / f"{c}.bin") for c in "ACGT"]
    for c in range(4):
        O.bv_write(paths[c], s.rows[c], s.n)
    ix = sg.Index.from_files(paths, k, config=sg.Config(presearch_depth=10))
    assert ix.info["has_groups"]
    np.testing.assert_array_equal(ix.rows()[4][: s.nwords], s.group)
    reads = _reads(rng, g, 2000, err=0.01, lens=(150, 150))
    batch = sg.ReadBatch.from_host(reads, k)
    chars, rs = O.parse_lines("\n".join(reads).encode())
    want = O.search_reads(oracle_index(s, 10), chars, rs, k)
    np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch)), want)
    # streaming: far fewer full searches than k-mers
    ix.counters(reset=True)
    ix.search_reads(batch, count=True)
    ctr = ix.counters(reset=True)
    assert ctr["kmers"] == batch.total_kmers and ctr["full_searches"] < 0.2 * ctr["kmers"]


def test_group_row_that_does_not_belong_is_refused(tmp_path):
    rng = np.random.default_rng(78)
    g = genome(rng, 20_000)
    s = O.SBWT([g], 17)
    bad = s.group.copy()
    # a column with an outgoing edge that is not marked as a group start
    edges = s.rows[0] | s.rows[1] | s.rows[2] | s.rows[3]
    w = int(np.flatnonzero(edges[1:])[0]) + 1
    bit = 63 - int(np.flatnonzero(np.unpackbits(np.array([edges[w]]).view(np.uint8)[::-1]))[0])
    bad[w] &= ~(np.uint64(1) << np.uint64(bit))
    with pytest.raises(sg.SbwtError, match="does not belong"):
        sg.Index.from_rows(s.rows, s.n, 17, group=bad)
    # a group file that passes the edge check but is not this index's row:
    # one group start without edges (the last k-mer of a short sequence)
    # dropped
    seqs, k = _group_case("short_reads", rng)
    r = O.SBWT(seqs, k)
    edges = r.rows[0] | r.rows[1] | r.rows[2] | r.rows[3]
    bits = lambda a: np.unpackbits(a.astype(">u8").view(np.uint8))[: r.n].astype(bool)
    cand = np.flatnonzero(bits(r.group) & ~bits(edges))
    cand = cand[cand > 0]
    assert cand.size, "no group start without edges"
    j = int(cand[0])
    og = r.group.copy()
    og[j >> 6] &= ~(np.uint64(1) << np.uint64(63 - (j & 63)))
    paths = [str(tmp_path / f"{c}.bin") for c in "ACGT"]
    for c in range(4):
        O.bv_write(paths[c], r.rows[c], r.n)
    gfile = str(tmp_path / "grp.bin")
    O.bv_write(gfile, og, r.n)
    with pytest.raises(sg.SbwtError, match="does not belong"):
        sg.Index.from_files(paths, k, group_path=gfile)
    O.bv_write(gfile, r.group, r.n)
    assert sg.Index.from_files(paths, k, group_path=gfile).info["has_groups"]


@pytest.mark.parametrize("k,m", [(31, 12), (31, 0), (20, 8), (32, 10), (9, 4)])
def test_certificates_on_negative_heavy_reads(k, m):
    """Reads from an unrelated genome (almost every k-mer absent), reads with
    dense errors, short rows and rows of exactly k: every variant agrees with
    the oracle."""
    rng = np.random.default_rng(k * 100 + m)
    g = genome(rng, 150_000)
    other = genome(rng, 50_000)
    s = O.SBWT([g], k)
    reads = _reads(rng, other, 150, err=0.0, lens=(k - 1, 300))
    reads += _reads(rng, g, 150, err=0.05, lens=(k, 300))
    reads += _reads(rng, g, 100, err=0.01, nrate=0.03, lens=(k, 500))
    reads += [g[:k], g[5:5 + k + 1], other[:k]]
    order = rng.permutation(len(reads))
    reads = [reads[i] for i in order]
    chars, rs = O.parse_lines("\n".join(reads).encode())
    oix = oracle_index(s, m)
    want = O.search_reads(oix, chars, rs, k)
    want_a = O.search_reads(oix, chars, rs, k, n_breaks=False)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=m))
    for seg in (256, 13):
        batch = sg.ReadBatch.from_host(reads, k, seg=seg)
        for kw in READ_VARIANTS:
            np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, **kw)), want, err_msg=str(kw))
        np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, n_as_a=True)), want_a)


def test_certificates_cut_work():
    """On reads with errors the certificates answer most absent k-mers
    without a search: fewer searches and lines than streaming alone."""
    rng = np.random.default_rng(77)
    g = genome(rng, 400_000)
    k = 31
    s = O.SBWT([g], k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=10))
    batch = sg.ReadBatch.from_host(_reads(rng, g, 3000, err=0.02, lens=(150, 150)), k)
    ix.counters(reset=True)
    a = sg.to_u64(ix.search_reads(batch, count=True))
    with_cert = ix.counters(reset=True)
    b = sg.to_u64(ix.search_reads(batch, count=True, certificates=False))
    without = ix.counters(reset=True)
    np.testing.assert_array_equal(a, b)
    assert with_cert["kmers"] == without["kmers"] == batch.total_kmers
    assert with_cert["lines"] < 0.8 * without["lines"]
    ix.counters(reset=True)
    c = sg.to_u64(ix.search_reads(batch, count=True, persistent=True))
    pers = ix.counters(reset=True)
    np.testing.assert_array_equal(a, c)
    assert pers["kmers"] == batch.total_kmers and pers["lines"] < 0.8 * without["lines"]


def test_segmental_copies_groups():
    """Repeats create suffix groups larger than one node, which exercises the
    backward group-start scan of the streaming step."""
    rng = np.random.default_rng(17)
    base = genome(rng, 20_000)
    parts = [base]
    for _ in range(30):
        p = int(rng.integers(0, 19_000))
        seg = np.array(list(base[p:p + 800]))
        flip = rng.random(800) < 0.02
        seg[flip] = ACGT[rng.integers(0, 4, flip.sum())]
        parts.append("".join(seg))
    g = "".join(parts)
    k = 21
    s = O.SBWT([g], k)
    groups = np.unpackbits(s.group.view(np.uint8).reshape(-1, 8)[:, ::-1].reshape(-1))[: s.n]
    assert groups.sum() < s.n  # some groups have more than one node
    reads = _reads(rng, g, 300, err=0.005, lens=(50, 300))
    chars, rs = O.parse_lines("\n".join(reads).encode())
    oix = oracle_index(s, 8)
    want = O.search_reads(oix, chars, rs, k)
    for layout in (16, 32, 64):
        ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=8, layout=layout))
        batch = sg.ReadBatch.from_host(reads, k)
        for kw in READ_VARIANTS:
            got = sg.to_u64(ix.search_reads(batch, **kw))
            np.testing.assert_array_equal(got, want, err_msg=str(kw))


# ---------------------------------------------------------------------------
# GPU SBWT builder against the oracle's builder


@pytest.mark.parametrize("seqs,k", [(["ACGTACGTTGCA"], 3), (["AAAAAAA"], 3), (["ACGTNACGTTTACG", "acgtacgt"], 3),
                                    (["TTTT", "AC"], 5), (["CGTAC" * 4], 5), (["ACGT" * 10, "", "GGGGCCCC"], 1),
                                    (["ACGTTGCA" * 9], 32)])
def test_gpu_builder_tiny(seqs, k):
    ref = O.SBWT(seqs, k)
    ix = sg.Index.build(seqs, k, sg.Config(presearch_depth=min(k, 3)))
    rows = ix.rows()
    assert ix.n == ref.n
    for c in range(4):
        np.testing.assert_array_equal(rows[c][: ref.nwords], ref.rows[c])
    np.testing.assert_array_equal(rows[4][: ref.nwords], ref.group)


@pytest.mark.parametrize("k", [31, 32, 17])
def test_gpu_builder_multi_sequence(k):
    rng = np.random.default_rng(k)
    seqs = [genome(rng, int(rng.integers(0, 3000))) for _ in range(40)]
    seqs += [genome(rng, 50_000, (0.3, 0.2, 0.2, 0.3))]
    seqs[3] = seqs[3][:100] + "NNN" + seqs[3][100:]
    ref = O.SBWT(seqs, k)
    ix = sg.Index.build(seqs, k, sg.Config(presearch_depth=8))
    rows = ix.rows()
    assert ix.n == ref.n and ix.info["C"] == list(map(int, ref.C))
    for c in range(4):
        np.testing.assert_array_equal(rows[c][: ref.nwords], ref.rows[c])
    np.testing.assert_array_equal(rows[4][: ref.nwords], ref.group)


# ---------------------------------------------------------------------------
# files


def test_files_round_trip(tmp_path):
    rng = np.random.default_rng(4)
    g = genome(rng, 30_000)
    k = 25
    s = O.SBWT([g], k)
    paths = [str(tmp_path / f"{c}.bin") for c in "ACGT"]
    for c in range(4):
        O.bv_write(paths[c], s.rows[c], s.n)
    O.bv_write(str(tmp_path / "grp.bin"), s.group, s.n)
    ix = sg.Index.from_files(paths, k, group_path=str(tmp_path / "grp.bin"), config=sg.Config(presearch_depth=9))
    assert ix.info["C"] == list(map(int, s.C))
    out = [str(tmp_path / f"o{c}.bin") for c in "ACGT"]
    ix.write_files(out)
    for c in range(4):
        assert open(out[c], "rb").read() == open(paths[c], "rb").read()
    q = topk(sample_queries(rng, g, k, 5000))
    np.testing.assert_array_equal(sg.to_u64(ix.search_kmers(dev_u64(q))), oracle_index(s, 9).search(q, k))


def test_file_errors(tmp_path):
    rng = np.random.default_rng(4)
    s = O.SBWT([genome(rng, 2000)], 11)
    paths = [str(tmp_path / f"{c}.bin") for c in "ACGT"]
    for c in range(4):
        O.bv_write(paths[c], s.rows[c], s.n)
    with pytest.raises(sg.SbwtError, match="can not be opened"):
        sg.Index.from_files(paths[:3] + [str(tmp_path / "missing.bin")], 11)
    O.bv_write(paths[2], s.rows[2], s.n - 1)
    with pytest.raises(sg.SbwtError, match="differ in length"):
        sg.Index.from_files(paths, 11)
    # swapping two rows keeps n but breaks nothing detectable; dropping a
    # 1-bit breaks C[4] = n
    bad = s.rows[0].copy()
    nz = np.flatnonzero(bad)[0]
    bad[nz] = np.uint64(0)
    with pytest.raises(sg.SbwtError, match="not an SBWT"):
        sg.Index.from_rows([bad, s.rows[1], s.rows[2], s.rows[3]], s.n, 11)


# ---------------------------------------------------------------------------
# scale: size-independent properties on a GPU-built index


def test_scale_properties_streaming_equals_independent():
    rng = np.random.default_rng(2024)
    g = genome(rng, 5_000_000, (0.3, 0.2, 0.2, 0.3))
    k = 31
    ix = sg.Index.build([g], k, sg.Config(presearch_depth=12))
    reads = _reads(rng, g, 20_000, err=0.01, lens=(150, 150))
    batch = sg.ReadBatch.from_host(reads, k)
    a = sg.to_u64(ix.search_reads(batch))
    b = sg.to_u64(ix.search_reads(batch, independent=True))
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, persistent=True)), a)
    np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, certificates=False)), a)
    # error-free reads are found everywhere, at distinct-or-equal nodes
    clean = sg.ReadBatch.from_host([g[p:p + 150] for p in rng.integers(0, len(g) - 150, 2000)], k)
    c = sg.to_u64(ix.search_reads(clean))
    assert (c != NONE).all()
    # the answers agree with search_kmers on the same k-mers
    chars, rs = O.parse_lines("\n".join(reads[:2000]).encode())
    km, valid = O.read_kmers(chars, rs, k)
    d = sg.to_u64(ix.search_kmers(dev_u64(km)))
    np.testing.assert_array_equal(d, a[: d.size])


# ---------------------------------------------------------------------------
# select and enumeration


@pytest.mark.parametrize("layout", [16, 32, 64])
def test_select_against_numpy(layout):
    rng = np.random.default_rng(layout)
    g = genome(rng, 60_000)
    k = 17
    s = O.SBWT([g], k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=6, layout=layout))
    for c in range(4):
        bits = np.unpackbits(s.rows[c].view(np.uint8).reshape(-1, 8)[:, ::-1].reshape(-1))[: s.n]
        ones = np.flatnonzero(bits)
        ranks = np.arange(ones.size + 3, dtype=np.uint64)
        got = sg.to_u64(ix.select(c, dev_u64(ranks)))
        np.testing.assert_array_equal(got[: ones.size], ones.astype(np.uint64))
        assert (got[ones.size:] == NONE).all()


@pytest.mark.parametrize("seqs,k", [(["ACGTACGTTGCAAGGCT", "TTGACCA"], 4), (["CGTAC" * 4], 5), (["ACGT" * 10, "GGGGCCCC"], 1)])
def test_node_labels_tiny(seqs, k):
    s = O.SBWT(seqs, k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=min(k, 2)))
    labels, full = ix.node_labels()
    labels, full = sg.to_u64(labels), full.cpu().numpy().astype(bool)
    for j in range(s.n):
        lab = s.label(j)
        assert full[j] == ("$" not in lab), (j, lab)
        if full[j]:
            assert int(labels[j]) == brute.encode_top(lab)
        else:
            assert labels[j] == NONE


def test_enumerate_kmers_matches_oracle_builder():
    rng = np.random.default_rng(12)
    seqs = [genome(rng, 40_000), genome(rng, 3000)[:1000] + "N" + genome(rng, 2000)]
    k = 31
    s = O.SBWT(seqs, k)
    ix = sg.Index.build(seqs, k, sg.Config(presearch_depth=10))
    km = sg.to_u64(ix.enumerate_kmers())
    assert km.size == s.num_kmers
    full = s.lens == k
    want = np.array([brute.encode_top(s.label(j)) for j in np.flatnonzero(full)], np.uint64)
    np.testing.assert_array_equal(km, want)
    # every enumerated k-mer is found, at its own node
    np.testing.assert_array_equal(sg.to_u64(ix.search_kmers(dev_u64(km))), np.flatnonzero(full).astype(np.uint64))


@pytest.mark.parametrize("layout", [16, 32, 64])
@pytest.mark.parametrize("m", [0, 8, 12])
def test_counted_lines_equal_the_oracle_count(layout, m):
    """The roofline's algorithmic bytes: the lines and steps the counted
    k_search_kmers adds up equal orc_search_count's (pinned against brute
    force) on the same k-mers, layout and depth."""
    rng = np.random.default_rng(300 + layout + m)
    g = genome(rng, 400_000)
    k = 31
    s = O.SBWT([g], k)
    q = topk(sample_queries(rng, g, k, 20_000))
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=m, layout=layout))
    oix = oracle_index(s, m)
    cols = {16: 96, 32: 224, 64: 480}[layout]
    want, steps, lines = oix.search_count(q, k, cols, layout)
    ix.counters(reset=True)
    got = sg.to_u64(ix.search_kmers(dev_u64(q), count=True))
    ctr = ix.counters(reset=True)
    np.testing.assert_array_equal(got, want)
    assert ctr["lines"] == int(lines.sum()) and ctr["steps"] == int(steps.sum())
    assert ctr["kmers"] == q.size and ctr["full_searches"] == q.size


def _short_rows(rng, g_codes, n, lo, hi, err):
    """n rows of lo..hi characters from uniform positions of the genome (2-bit
    codes), with substitutions at rate err, concatenated; returns the ASCII
    characters and the row starts."""
    lens = rng.integers(lo, hi + 1, n)
    starts = rng.integers(0, g_codes.size - hi, n)
    rs = np.zeros(n + 1, np.uint64)
    rs[1:] = np.cumsum(lens)
    total = int(rs[-1])
    idx = np.repeat(starts - rs[:-1].astype(np.int64), lens) + np.arange(total)
    codes = g_codes[idx].copy()
    flip = rng.random(total) < err
    codes[flip] = rng.integers(0, 4, int(flip.sum()))
    return np.frombuffer(b"ACGT", np.uint8)[codes], rs


@pytest.mark.parametrize("seg", [256, 7])
def test_search_reads_many_short_rows(seg):
    """More work items than the kernel has lanes, over short rows packed back
    to back: lanes claim the item after the one they finished, whose first
    k-mer often starts in the packed word after the last one they held (the
    held words of the uniform kernel must be dropped on every change of
    row). Every read-search form against the oracle, and run items through a
    plan."""
    rng = np.random.default_rng(71 + seg)
    k = 31
    g_codes = rng.integers(0, 4, 1_000_000).astype(np.uint8)
    g = np.frombuffer(b"ACGT", np.uint8)[g_codes].tobytes().decode()
    s = O.SBWT([g], k)
    chars, rs = _short_rows(rng, g_codes, 450_000, k, 72, 0.01)
    oix = oracle_index(s, 10)
    want = O.search_reads(oix, chars, rs, k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=10))
    dchars = torch.from_numpy(chars.copy()).cuda()
    batch = sg.ReadBatch(dchars, dev_u64(rs), k, seg=seg, host_starts=rs)
    assert batch.work_bound > 400_000  # more items than 256 CUs x 6 blocks x 256 lanes
    for kw in ({}, {"persistent": True}, {"certificates": False, "persistent": True}, {"lanes": True}):
        got = sg.to_u64(ix.search_reads(batch, **kw))
        np.testing.assert_array_equal(got, want, err_msg=str(kw))
    plan = sg.ReadsPlan(rs.size - 1, chars.size, seg)
    buf = torch.empty(want.size + 2, dtype=torch.int64, device="cuda")
    for kw in ({}, {"runs": True, "persistent": True}):
        for at in (0, 1):  # answers 16-byte aligned, or 8 bytes past (a batch inside a larger array)
            buf.fill_(7)
            out = buf[at:at + want.size]
            plan.search(ix, dchars, dev_u64(rs), out, **kw)
            torch.cuda.synchronize()
            np.testing.assert_array_equal(sg.to_u64(out), want, err_msg=f"plan {kw} at {at}")
            assert int(buf[at + want.size]) == 7 and (at == 0 or int(buf[0]) == 7)
