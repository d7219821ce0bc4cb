"""Writes the golden fixtures of this directory from the pure-Python brute
force of tests/brute.py: node labels with '$' padding in Python's own sort
(colex order), set membership for search, suffix ranges for the presearch
table, and a table-driven packer. Nothing here calls the C oracle or the
library, so the fixtures pin both.

    python tests/golden/make_golden.py

The reference holds no fixtures of its own (no tests, golden vectors or
answer files); these follow its semantics: the SBWT rows and C array of
include/matrix_sbwt.hpp:150-157, the search of include/gpu/search_kernel.hpp:
61-104 (node id or ~0), the presearch table of include/presearch.hpp:24-30
(interval of every m-mer, char 0 in the top bits), the packing code table of
include/create_kmer_string.hpp:18-42 and the result text of
include/epicseq/results_persister.hpp:82-96.
"""
from __future__ import annotations

import bisect
import itertools
import json
import os
import random
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import brute  # noqa: E402

ACGT = "ACGT"


def rand_seq(rng, n):
    return "".join(rng.choice(ACGT) for _ in range(n))


def set_bits(bits):
    return [i for i, b in enumerate(bits) if b]


def c_array(rows):
    C = [1]
    for c in range(4):
        C.append(C[-1] + sum(rows[c]))
    return C


def presearch_table(order, m):
    """[l, r] of every m-mer in code order (char 0 most significant): the
    nodes whose last m characters are the m-mer; l = r + 1 = the insertion
    point when there are none (brute.interval, by bisection)."""
    pre = [s[::-1][:m] for s in order]  # sorted, since order is colex
    left, right = [], []
    for p in itertools.product(ACGT, repeat=m):
        rp = "".join(p)[::-1]
        left.append(bisect.bisect_left(pre, rp))
        right.append(bisect.bisect_right(pre, rp) - 1)
    return np.array(left, np.uint32), np.array(right, np.uint32)


def pack_table(raw: bytes):
    """2 bits per character, 32 per u64, character 0 in the top bits; C/c = 1,
    G/g = 2, T/t = 3, anything else 0. The N mask: one bit per character, 64
    per u64, MSB first, set for every byte that is not A, C, G or T in either
    case."""
    code = {ord("C"): 1, ord("c"): 1, ord("G"): 2, ord("g"): 2, ord("T"): 3, ord("t"): 3}
    acgt = set(b"ACGTacgt")
    packed = [0] * (len(raw) // 32 + 2)
    nmask = [0] * (len(raw) // 64 + 2)
    for i, b in enumerate(raw):
        packed[i // 32] |= code.get(b, 0) << (62 - 2 * (i % 32))
        if b not in acgt:
            nmask[i // 64] |= 1 << (63 - i % 64)
    return packed, nmask


def read_answers(order, K, rows_txt, k):
    """Per row, every window's answer: ~0 (as -1) when it holds a character
    other than A, C, G, T in either case, or is absent; else its node. And
    the result text: each answer as a signed decimal and a space, a newline
    per row."""
    answers, text = [], []
    for row in rows_txt:
        line = []
        for i in range(len(row) - k + 1):
            w = row[i:i + k].upper()
            ok = all(ch in ACGT for ch in w)  # lowercase acgt are bases (the packing table)
            v = order.index(w) if ok and w in K else -1
            line.append(v)
        answers.extend(line)
        text.append("".join(f"{v} " for v in line) + "\n")
    return answers, "".join(text)


def main():
    rng = random.Random(2024)
    # 1. a 64 bp sequence, k = 5: every 5-mer's answer (about 200 columns)
    seq = rand_seq(rng, 64)
    k = 5
    order, rows, group = brute.sbwt([seq], k)
    K = brute.kmer_set([seq], k)
    allk = ["".join(p) for p in itertools.product(ACGT, repeat=k)]
    ans = [order.index(x) if x in K else -1 for x in allk]
    with open(os.path.join(HERE, "sbwt_64bp_k5.json"), "w") as f:
        json.dump({"seq": seq, "k": k, "n": len(order), "C": c_array(rows),
                   "rows": [set_bits(r) for r in rows], "group": set_bits(group),
                   "answers_all_kmers_in_code_order": ans}, f)

    # 2. a 1 kbp sequence, k = 31: presearch tables at m = 5 and 8, answers of
    #    1000 sampled and 1000 random 31-mers, and reads with N, lowercase and
    #    short rows (answers and result text)
    seq = rand_seq(rng, 1000)
    k = 31
    order, rows, group = brute.sbwt([seq], k)
    K = brute.kmer_set([seq], k)
    l5, r5 = presearch_table(order, 5)
    l8, r8 = presearch_table(order, 8)
    qs = [seq[p:p + k] for p in (rng.randrange(0, len(seq) - k + 1) for _ in range(1000))]
    qs += [rand_seq(rng, k) for _ in range(1000)]
    codes = np.array([brute.encode_top(x) for x in qs], np.uint64)
    qans = np.array([order.index(x) if x in K else brute.NONE for x in qs], np.uint64)
    reads = []
    for i in range(40):
        p = rng.randrange(0, len(seq) - 150)
        r = list(seq[p:p + rng.randrange(20, 150)])
        for _ in range(rng.randrange(0, 3)):  # substitutions
            q = rng.randrange(len(r))
            r[q] = rng.choice(ACGT)
        if i % 5 == 1:
            r[rng.randrange(len(r))] = "N"
        if i % 7 == 2:
            r = [ch.lower() for ch in r]
        reads.append("".join(r))
    reads.append(seq[100:100 + k])          # exactly one k-mer
    reads.append(seq[200:200 + k - 1])      # shorter than k: no k-mers
    rans, rtext = read_answers(order, K, reads, k)
    np.savez_compressed(os.path.join(HERE, "sbwt_1kbp_k31.npz"),
                        seq=np.frombuffer(seq.encode(), np.uint8), k=k, n=len(order),
                        C=np.array(c_array(rows), np.uint64),
                        rows=np.concatenate([np.array(set_bits(r), np.uint32) for r in rows]),  # set columns
                        row_ones=np.array([sum(r) for r in rows], np.uint64),  # ... per row, in order
                        group=np.array(set_bits(group), np.uint32),
                        pre5_left=l5, pre5_right=r5, pre8_left=l8, pre8_right=r8,
                        queries=codes, answers=qans)
    with open(os.path.join(HERE, "reads_1kbp_k31.json"), "w") as f:
        json.dump({"k": k, "rows": reads, "answers": rans, "text": rtext}, f)

    # 3. packing: mixed case, N, other bytes, NUL, unaligned lengths
    kat = []
    for s in (b"ACGT" * 8, b"acgtNNNN" + b"T" * 24, b"C" + b"\0" * 31 + b"G" * 32,
              b"GATTACA" * 11 + b"nRYKM.-" + b"tgca" * 5, bytes(rng.choice(b"ACGTacgtN") for _ in range(333))):
        packed, nmask = pack_table(s)
        kat.append({"chars": s.decode("latin-1"), "packed": [f"{w:016x}" for w in packed],
                    "nmask": [f"{w:016x}" for w in nmask]})
    with open(os.path.join(HERE, "pack_kat.json"), "w") as f:
        json.dump(kat, f)


if __name__ == "__main__":
    main()
