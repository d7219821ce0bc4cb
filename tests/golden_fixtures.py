"""Loads the golden fixtures of tests/golden (written by
tests/golden/make_golden.py from the pure-Python brute force)."""
from __future__ import annotations

import json
import os

import numpy as np

import brute

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NONE = np.uint64(0xFFFFFFFFFFFFFFFF)


def load_json(name: str):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)


def words(cols, n: int) -> np.ndarray:
    """MSB-first row words (ceil(n/4096)*64 of them) with the given set columns."""
    bits = [0] * n
    for c in cols:
        bits[int(c)] = 1
    return brute.bits_to_words(bits)


def sbwt_64bp_k5() -> dict:
    f = load_json("sbwt_64bp_k5.json")
    n, k = f["n"], f["k"]
    f["row_words"] = [words(r, n) for r in f["rows"]]
    f["group_words"] = words(f["group"], n)
    f["codes"] = (np.arange(4 ** k, dtype=np.uint64) << np.uint64(64 - 2 * k))
    f["want"] = np.array([NONE if a < 0 else a for a in f["answers_all_kmers_in_code_order"]], np.uint64)
    return f


def sbwt_1kbp_k31() -> dict:
    z = np.load(os.path.join(GOLDEN, "sbwt_1kbp_k31.npz"), allow_pickle=False)
    f = {key: z[key] for key in z.files}
    n = int(f["n"])
    cols = np.split(f["rows"], np.cumsum(f["row_ones"].astype(np.int64))[:-1])
    f["seq"] = f["seq"].tobytes().decode()
    f["k"], f["n"] = int(f["k"]), n
    f["row_words"] = [words(c, n) for c in cols]
    f["group_words"] = words(f["group"], n)
    return f


def reads_1kbp_k31() -> dict:
    f = load_json("reads_1kbp_k31.json")
    f["want"] = np.array([NONE if a < 0 else a for a in f["answers"]], np.uint64)
    return f


def pack_kat() -> list:
    out = []
    for e in load_json("pack_kat.json"):
        out.append({"raw": e["chars"].encode("latin-1"),
                    "packed": np.array([int(w, 16) for w in e["packed"]], np.uint64),
                    "nmask": np.array([int(w, 16) for w in e["nmask"]], np.uint64)})
    return out
