"""Pins the CPU oracle against brute force, closed forms and hand-computed
known answers (the reference ships no tests, fixtures or golden vectors).

Each check is independent of the oracle's own formula: SBWT labels come from
Python string sorting, search answers from set membership, rank from prefix
popcounts, poppy fields from per-block counts, packing from hand-computed
words."""
import itertools
import os
import random

import numpy as np
import pytest

import brute
from oracle import oracle as O

NONE = brute.NONE


def rand_seq(rng, n, alphabet="ACGT"):
    return "".join(rng.choice(alphabet) for _ in range(n))


CASES = [
    (["ACGTACGTTGCA"], 3),
    (["AAAAAAA"], 3),
    (["ACACACACAC", "GGGTTT"], 4),
    (["ACGTNACGTTTACG", "acgtacgt"], 3),
    (["TTTT", "AC"], 5),  # no k-mer at all: only the root
    (["CGTAC" * 4], 5),  # cyclic: k-mers with predecessors
]


def _rand_cases():
    rng = random.Random(7)
    out = []
    for k in (2, 3, 4, 5, 6):
        for _ in range(4):
            nseq = rng.randint(1, 4)
            seqs = [rand_seq(rng, rng.randint(0, 60), "ACGTN" if rng.random() < 0.3 else "ACGT")
                    for _ in range(nseq)]
            out.append((seqs, k))
    return out


@pytest.mark.parametrize("seqs,k", CASES + _rand_cases())
def test_sbwt_builder_matches_brute_force(seqs, k):
    order, rows, group = brute.sbwt(seqs, k)
    s = O.SBWT(seqs, k)
    assert s.n == len(order)
    assert [s.label(j) for j in range(s.n)] == order
    for c in range(4):
        assert brute.words_to_bits(s.rows[c], s.n) == rows[c]
    assert brute.words_to_bits(s.group, s.n) == group
    # C array: C[0]=1 (root), C[c+1]=C[c]+ones_c, and C[4]=n for a valid SBWT
    ones = [sum(r) for r in rows]
    assert list(map(int, s.C)) == [1, 1 + ones[0], 1 + sum(ones[:2]), 1 + sum(ones[:3]), 1 + sum(ones)]
    assert int(s.C[4]) == s.n
    assert s.num_kmers == len(brute.kmer_set(seqs, k))


@pytest.mark.parametrize("seqs,k", [(["ACGTACGTTGCAAGGCT", "TTGACCA"], 4), (["ACGTTGCATGCAAACCGGTT" * 3], 5),
                                    (["GATTACA", "CATTAG", "ACGTGTCA"], 3)])
@pytest.mark.parametrize("sb", [256, 1024, 4096])
def test_search_every_kmer_against_brute_force(seqs, k, sb):
    order, _, _ = brute.sbwt(seqs, k)
    K = brute.kmer_set(seqs, k)
    s = O.SBWT(seqs, k)
    idx = O.Index(s.rows, s.n, sb=sb)
    allk = ["".join(p) for p in itertools.product("ACGT", repeat=k)]
    codes = np.array([brute.encode_top(x) for x in allk], dtype=np.uint64)
    expect = np.array([brute.search(order, K, x) for x in allk], dtype=np.uint64)
    for m in range(0, k + 1):
        idx.presearch(m)
        got = idx.search(codes, k)
        np.testing.assert_array_equal(got, expect, err_msg=f"m={m}")


@pytest.mark.parametrize("seqs,k,m", [(["ACGTACGTTGCAAGGCT", "TTGACCA"], 4, 2), (["GATTACA", "CATTAG"], 5, 3),
                                      (["ACGTTGCATGCAAACCGGTT"], 6, 4)])
def test_presearch_intervals_are_suffix_ranges(seqs, k, m):
    """Entry P is the range of nodes whose label ends with P; an empty range
    has l = r + 1 at the insertion point."""
    order, _, _ = brute.sbwt(seqs, k)
    s = O.SBWT(seqs, k)
    idx = O.Index(s.rows, s.n)
    left, right = idx.presearch(m)
    for i, p in enumerate(itertools.product("ACGT", repeat=m)):
        l, r = brute.interval(order, "".join(p))
        assert (int(left[i]), int(right[i])) == (l, r), "".join(p)


def _random_rows(rng, n, density):
    nw = ((n + 4095) // 4096) * 64
    w = rng.integers(0, 2**63, size=max(nw, 1), dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, size=max(nw, 1), dtype=np.uint64)
    if density < 0.5:
        w &= rng.integers(0, 2**63, size=max(nw, 1), dtype=np.uint64) * np.uint64(2)
    # zero the padding past n
    if n % 64:
        w[n >> 6] &= np.uint64(((1 << (n % 64)) - 1) << (64 - n % 64))
    w[(n + 63) // 64:] = 0
    return w


@pytest.mark.parametrize("n", [1, 63, 64, 65, 255, 256, 1024, 4096, 5000, 8192, 12345])
@pytest.mark.parametrize("sb,version", [(256, 1), (512, 1), (1024, 1), (2048, 1), (4096, 1),
                                        (256, 2), (512, 2), (1024, 2), (2048, 2)])
def test_poppy_rank_equals_prefix_popcount(n, sb, version):
    rng = np.random.default_rng(n * 31 + sb + version)
    w = _random_rows(rng, n, 0.5)
    p = O.Poppy(w, n, sb, version)
    bits = np.unpackbits(w.byteswap().view(np.uint8))[:n]  # MSB-first bit order
    prefix = np.concatenate([[0], np.cumsum(bits)])
    for i in range(n + 1):
        assert p.rank(i) == prefix[i], (i, sb, version)
        assert O.rank_naive(w, i) == prefix[i]
    assert p.ones == prefix[n]


@pytest.mark.parametrize("sb", [256, 512, 1024, 2048, 4096])
def test_poppy_fields(sb):
    """L12 word = rel << 32 | BB0 << 20 | BB1 << 10 | BB2 (33/22/11 at 4096),
    with rel relative to the hyperblock (2^32 bits, 2^31 at SB=4096)."""
    n = 3 * sb + 17
    rng = np.random.default_rng(sb)
    w = _random_rows(rng, n, 0.5)
    p = O.Poppy(w, n, sb, 1)
    bits = np.unpackbits(w.byteswap().view(np.uint8)).astype(np.int64)
    bb = sb // 4
    nsb = (n + sb - 1) // sb
    assert p.L12.size == nsb + 1 and p.L0.size == 1 and p.L0[0] == 0
    big = 1 if sb == 4096 else 0
    rel = 0
    for s_ in range(nsb):
        blk = bits[s_ * sb:(s_ + 1) * sb]
        cnt = [int(blk[b * bb:(b + 1) * bb].sum()) for b in range(4)]
        want = (rel << (32 + big)) | (cnt[0] << (20 + 2 * big)) | (cnt[1] << (10 + big)) | cnt[2]
        assert int(p.L12[s_]) == want
        rel += sum(cnt)
    assert int(p.L12[nsb]) == rel << (32 + big)
    # cum-poppy stores prefix sums instead
    if sb < 4096:
        q = O.Poppy(w, n, sb, 2)
        sh = (22, 11) if sb == 2048 else (20, 10)
        rel = 0
        for s_ in range(nsb):
            blk = bits[s_ * sb:(s_ + 1) * sb]
            cnt = [int(blk[b * bb:(b + 1) * bb].sum()) for b in range(4)]
            want = (rel << 32) | (cnt[0] << sh[0]) | ((cnt[0] + cnt[1]) << sh[1]) | (cnt[0] + cnt[1] + cnt[2])
            assert int(q.L12[s_]) == want
            rel += sum(cnt)


def test_poppy_hyperblock_boundary():
    """L0 holds one absolute count per 2^31-bit hyperblock at SB=4096; check
    the second hyperblock's base against a closed form on an all-ones row
    (rank(i) = i below n)."""
    n = (1 << 31) + 5000
    nw = ((n + 4095) // 4096) * 64
    w = np.full(nw, np.uint64(0xFFFFFFFFFFFFFFFF), dtype=np.uint64)
    w[n >> 6] = np.uint64(((1 << (n % 64)) - 1) << (64 - n % 64)) if n % 64 else w[n >> 6]
    w[(n + 63) // 64:] = 0
    p = O.Poppy(w, n, 4096, 1)
    assert p.L0.size == 2 and int(p.L0[1]) == 1 << 31
    for i in [0, 1, (1 << 31) - 1, 1 << 31, (1 << 31) + 1, (1 << 31) + 4095, (1 << 31) + 4096, n - 1, n]:
        assert p.rank(i) == i


def test_bitvector_file_format(tmp_path):
    """bit i = byte 8+(i>>3), bit 7-(i&7); 8-byte little-endian bit count."""
    bits = [1, 0, 1, 1, 0, 0, 0, 1, 1, 1] + [0] * 60 + [1]
    n = len(bits)
    w = brute.bits_to_words(bits)
    path = str(tmp_path / "v.bin")
    O.bv_write(path, w, n)
    raw = open(path, "rb").read()
    assert raw[:8] == n.to_bytes(8, "little")
    assert len(raw) == 8 + (n + 7) // 8
    assert raw[8] == 0b10110001 and raw[9] == 0b11000000
    assert raw[8 + 70 // 8] == 1 << (7 - 70 % 8)
    w2, n2 = O.bv_read(path)
    assert n2 == n and brute.words_to_bits(w2, n) == bits


def test_pack_known_answers():
    # create_kmer_string.hpp:20 -- C/c=1, G/g=2, T/t=3, anything else 0
    raw = b"ACGT" * 8
    assert O.pack(raw)[0] == 0x1B1B1B1B1B1B1B1B
    raw = b"acgtNNNN" + b"T" * 24
    assert O.pack(raw)[0] == 0x1B00FFFFFFFFFFFF
    raw = b"C" + b"\0" * 31 + b"G" * 32
    p = O.pack(raw)
    assert p[0] == 1 << 62 and p[1] == 0xAAAAAAAAAAAAAAAA


def test_extract_matches_direct_encoding():
    rng = random.Random(3)
    s = rand_seq(rng, 200)
    raw = (s + "\0" * 32).encode()
    raw += b"\0" * (-len(raw) % 32)
    packed = O.pack(raw)
    for pos in range(0, 200 - 31):
        for k in (5, 31, 32):
            if pos + k > 200:
                continue
            got = O.extract(packed, pos) >> (64 - 2 * k)
            assert got == brute.encode_top(s[pos:pos + k]) >> (64 - 2 * k)


def test_positions_closed_form():
    lens = [10, 31, 30, 0, 45, 31, 100]
    rs = np.concatenate([[0], np.cumsum(lens)]).astype(np.uint64)
    k = 31
    pos = O.positions(rs, k)
    assert pos.size == sum(max(0, L - k + 1) for L in lens)
    want = [int(rs[r]) + i for r, L in enumerate(lens) for i in range(max(0, L - k + 1))]
    assert list(map(int, pos)) == want
    off = O.kmer_offsets(rs, k)
    assert list(map(int, off)) == [0, 0, 1, 1, 1, 16, 17, 87]


def test_read_kmers_and_n_policy():
    chars = b"ACGTNACGTA"
    rs = np.array([0, 10], np.uint64)
    km, valid = O.read_kmers(chars, rs, 4, True)
    assert valid.tolist() == [True, False, False, False, False, True, True]
    assert int(km[0]) == brute.encode_top("ACGT")
    km2, valid2 = O.read_kmers(chars, rs, 4, False)
    assert valid2.all() and int(km2[1]) == brute.encode_top("CGTA")  # N codes as A


def test_parse_lines():
    text = b">r1\nACGT\n\nGG\r\n>x\nTTTA"
    chars, rs = O.parse_lines(text)
    assert bytes(chars) == b"ACGTGGTTTA"
    assert list(map(int, rs)) == [0, 4, 6, 10]


def test_format_results():
    res = np.array([0, NONE, 5, 123456789], np.uint64)
    off = np.array([0, 2, 2, 4], np.uint64)
    assert O.format_results(res, off) == b"0 -1 \n\n5 123456789 \n"


def test_reference_cpu_path_one_mbp():
    """The CPU-runnable workload at reduced size (200 kbp): sampled k-mers are
    found at their colex rank, random ones almost surely absent, and the
    answer does not depend on presearch depth or superblock size."""
    rng = np.random.default_rng(11)
    g = "".join(np.array(list("ACGT"))[rng.integers(0, 4, 200_000)])
    k = 31
    s = O.SBWT([g], k)
    assert int(s.C[4]) == s.n
    pos = rng.integers(0, len(g) - k + 1, 5000)
    pos_k = [g[p:p + k] for p in pos]
    rnd_k = ["".join(np.array(list("ACGT"))[rng.integers(0, 4, k)]) for _ in range(5000)]
    codes = np.array([brute.encode_top(x) for x in pos_k + rnd_k], np.uint64)
    results = {}
    for sb, m in [(1024, 8), (1024, 0), (256, 5), (4096, 12)]:
        idx = O.Index(s.rows, s.n, sb=sb)
        idx.presearch(m)
        results[(sb, m)] = idx.search(codes, k)
    base = results[(1024, 8)]
    for v in results.values():
        np.testing.assert_array_equal(v, base)
    found = base[:5000]
    assert (found != NONE).all()
    # the found node's label is the k-mer
    for j in range(0, 5000, 97):
        assert s.label(int(found[j])) == pos_k[j]
    assert (base[5000:] == NONE).mean() > 0.99


def _brute_lines(order, x, k, m, cpc, cbytes):
    """Rank lines and steps of the reference search of x, from brute-force
    suffix intervals (Python sort of the labels) instead of rank: a
    presearch gather is one line; each narrowing step reads the lines of
    columns l and r + 1 of the interval it narrows, and the search stops
    after the step that empties the interval."""
    n = len(order)
    rev = [s[::-1] for s in order]
    import bisect

    def iv(i):
        if i == 0:
            return 0, n - 1
        col = [r[:i] for r in rev]  # sorted: colex order agrees with every suffix length
        p = x[:i][::-1]
        return bisect.bisect_left(col, p), bisect.bisect_right(col, p) - 1

    def line(c):
        return (c // cpc) * cbytes // 128

    lines, steps = (1 if m else 0), 0
    for i in range(m, k):
        l, r = iv(i)
        lines += 1 if line(l) == line(r + 1) else 2
        steps += 1
        l2, r2 = iv(i + 1)
        if l2 > r2:
            break
    return lines, steps


@pytest.mark.parametrize("cpc,cbytes", [(96, 16), (224, 32), (480, 64), (5, 16)])
@pytest.mark.parametrize("m", [0, 2, 4])
def test_search_count_lines_against_brute_force(cpc, cbytes, m):
    """orc_search_count's per-k-mer lines and steps (the algorithmic bytes of
    the roofline) equal a count over brute-force intervals."""
    rng = random.Random(cpc * 10 + m)
    seqs = [rand_seq(rng, 2500), rand_seq(rng, 400)]
    k = 7
    order, _, _ = brute.sbwt(seqs, k)
    s = O.SBWT(seqs, k)
    assert s.n == len(order)
    idx = O.Index(s.rows, s.n)
    idx.presearch(m)
    K = sorted(brute.kmer_set(seqs, k))
    xs = [rng.choice(K) for _ in range(150)] + [rand_seq(rng, k) for _ in range(150)]
    kmers = np.array([brute.encode_top(x) for x in xs], np.uint64)
    ans, steps, lines = idx.search_count(kmers, k, cpc, cbytes)
    for q, x in enumerate(xs):
        bl, bs = _brute_lines(order, x, k, m, cpc, cbytes)
        assert (int(lines[q]), int(steps[q])) == (bl, bs), (x, m)
        assert int(ans[q]) == brute.search(order, set(K), x)
