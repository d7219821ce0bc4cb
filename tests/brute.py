"""Pure-Python brute force used to pin the oracle: string labels with '$'
padding, Python's own sort for colex order, set membership for search."""
from __future__ import annotations

import numpy as np

ALPHA = "ACGT"
NONE = 0xFFFFFFFFFFFFFFFF


def kmer_set(seqs, k):
    K = set()
    for s in seqs:
        s = s.upper()
        for i in range(len(s) - k + 1):
            w = s[i:i + k]
            if all(ch in ALPHA for ch in w):
                K.add(w)
    return K


def sbwt(seqs, k):
    """Returns (labels in colex order, rows[4] as 0/1 lists, group marks)."""
    K = kmer_set(seqs, k)
    suffixes = {y[1:] for y in K}
    R = set(K)
    R.add("$" * k)
    for x in K:
        if x[:-1] not in suffixes:
            for i in range(1, k):
                R.add("$" * (k - i) + x[:i])
    order = sorted(R, key=lambda s: s[::-1])  # '$' < 'A' < 'C' < 'G' < 'T' in ASCII
    n = len(order)
    rows = [[0] * n for _ in range(4)]
    group = [0] * n
    for j, lab in enumerate(order):
        if j == 0 or order[j - 1][1:] != lab[1:]:
            group[j] = 1
            for c, ch in enumerate(ALPHA):
                if lab[1:] + ch in R:
                    rows[c][j] = 1
    return order, rows, group


def bits_to_words(bits, nwords=None):
    n = len(bits)
    nw = nwords if nwords is not None else ((n + 4095) // 4096) * 64
    w = np.zeros(max(nw, 1), dtype=np.uint64)
    for i, b in enumerate(bits):
        if b:
            w[i >> 6] |= np.uint64(1) << np.uint64(63 - (i & 63))
    return w


def words_to_bits(words, n):
    return [int((int(words[i >> 6]) >> (63 - (i & 63))) & 1) for i in range(n)]


def encode_top(kmer: str) -> int:
    """top-aligned 2-bit code, char 0 in bits 63:62"""
    v = 0
    for t, ch in enumerate(kmer):
        v |= ALPHA.index(ch) << (62 - 2 * t)
    return v


def search(order, K, x):
    if x in K:
        return order.index(x)
    return NONE


def interval(order, p):
    """[l, r] of nodes whose label ends with p; for absent p, l = r + 1 = the
    number of nodes whose last |p| characters sort before p."""
    rp = p[::-1]
    m = len(p)
    l = sum(1 for s in order if s[::-1][:m] < rp)
    r1 = sum(1 for s in order if s[::-1][:m] <= rp)
    return l, r1 - 1
