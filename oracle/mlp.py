"""CPU oracle of the learned prefilter's MLP (test infrastructure only).

The network (no reference counterpart): one-hot 2-bit k-mer (character t,
code c -> input 4t + c) -> 128 ReLU -> 64 ReLU -> 1 logit; loss = mean binary
cross-entropy with logits. The flat fp32 parameter vector is
W1[4k][128], b1[128], W2[128][64], b2[64], w3[64], b3.

Rounding rules the GPU kernels follow and this oracle restates:
  * W1, W2, w3 are used as bf16 (round to nearest even of the fp32 master);
    biases stay fp32;
  * training forward keeps h1, h2 in full precision; inference rounds h1
    (after ReLU) to bf16, because it feeds the second layer's bf16 MFMA;
  * gradients are taken with respect to the rounded weights and applied to
    the fp32 masters (straight-through), Adam with beta 0.9 / 0.999, eps 1e-8.

Everything here is float64 numpy. Pinned in tests/test_oracle_mlp.py by
central finite differences of the loss and by torch autograd over an
independently written forward (dense one-hot matmul). The Bloom filter's
restatement (below) is pinned there by splitmix64's published first output,
128-bit products in Python integers, the minimizer recomputed from the
k-mer's text, and its Poisson sizing model against simulation.
"""
from __future__ import annotations

import numpy as np

H1, H2 = 128, 64


def shape(k: int) -> dict:
    i = 4 * k
    o = {"in": i, "W1": 0}
    o["b1"] = o["W1"] + i * H1
    o["W2"] = o["b1"] + H1
    o["b2"] = o["W2"] + H1 * H2
    o["w3"] = o["b2"] + H2
    o["b3"] = o["w3"] + H2
    o["total"] = o["b3"] + 1
    return o


def bf16_round(x: np.ndarray) -> np.ndarray:
    """fp32 -> bf16 -> fp32 with round-to-nearest-even (finite inputs)."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(params: np.ndarray, k: int, rounded: bool = True):
    """Weights and biases as float64 arrays. rounded: the fp32 masters with
    W1, W2, w3 rounded to bf16 (what the kernels compute with); otherwise the
    values exactly as given (float64), for finite-difference checks."""
    o = shape(k)
    if rounded:
        p = np.asarray(params, dtype=np.float32)
        r = bf16_round
    else:
        p = np.asarray(params, dtype=np.float64)
        r = (lambda a: a)
    W1 = r(p[o["W1"]:o["b1"]]).astype(np.float64).reshape(4 * k, H1)
    b1 = p[o["b1"]:o["W2"]].astype(np.float64)
    W2 = r(p[o["W2"]:o["b2"]]).astype(np.float64).reshape(H1, H2)
    b2 = p[o["b2"]:o["w3"]].astype(np.float64)
    w3 = r(p[o["w3"]:o["b3"]]).astype(np.float64)
    b3 = float(p[o["b3"]])
    return W1, b1, W2, b2, w3, b3


def chars(kmers: np.ndarray, k: int) -> np.ndarray:
    """[B][k] 2-bit codes of top-aligned k-mers (character 0 in bits 63..62)."""
    km = np.asarray(kmers, dtype=np.uint64)
    sh = np.array([62 - 2 * t for t in range(k)], dtype=np.uint64)
    return ((km[:, None] >> sh[None, :]) & np.uint64(3)).astype(np.int64)


def first_layer(W1, kmers, k):
    """sum_t W1[4t + c_t] for every k-mer: the one-hot product as a gather."""
    c = chars(kmers, k)
    idx = 4 * np.arange(k)[None, :] + c
    return W1[idx].sum(axis=1)


def forward(params, kmers, k, inference: bool = False, rounded: bool = True):
    W1, b1, W2, b2, w3, b3 = unpack(params, k, rounded)
    a1 = first_layer(W1, kmers, k) + b1
    h1 = np.maximum(a1, 0.0)
    if inference:
        h1 = bf16_round(h1.astype(np.float32)).astype(np.float64)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0.0)
    z = h2 @ w3 + b3
    return z, (a1, h1, a2, h2)


def logits(params, kmers, k):
    """Inference logits (h1 rounded to bf16, as the matrix-core kernel does)."""
    return forward(params, kmers, k, inference=True)[0]


def loss_and_grad(params, kmers, labels, k, rounded: bool = True):
    """Mean BCE-with-logits and its gradient (flat, same layout as params)
    with respect to the bf16-rounded weights and fp32 biases (rounded=False:
    with respect to the float64 parameters as given)."""
    o = shape(k)
    W1, b1, W2, b2, w3, b3 = unpack(params, k, rounded)
    y = np.asarray(labels, dtype=np.float64)
    B = y.size
    z, (a1, h1, a2, h2) = forward(params, kmers, k, rounded=rounded)
    loss = float(np.mean(np.logaddexp(0.0, z) - y * z))
    dz = (1.0 / (1.0 + np.exp(-z)) - y) / B
    g = np.zeros(o["total"], dtype=np.float64)
    g[o["b3"]] = dz.sum()
    g[o["w3"]:o["b3"]] = h2.T @ dz
    dh2 = np.outer(dz, w3) * (a2 > 0)
    g[o["b2"]:o["w3"]] = dh2.sum(axis=0)
    g[o["W2"]:o["b2"]] = (h1.T @ dh2).reshape(-1)
    dh1 = (dh2 @ W2.T) * (a1 > 0)
    g[o["b1"]:o["W2"]] = dh1.sum(axis=0)
    c = chars(kmers, k)
    gW1 = np.zeros((4 * k, H1))
    for t in range(k):
        np.add.at(gW1, 4 * t + c[:, t], dh1)
    g[o["W1"]:o["b1"]] = gW1.reshape(-1)
    return loss, g


def adam_step(params, m, v, grad, step: int, lr: float):
    """One Adam update (beta1 0.9, beta2 0.999, eps 1e-8) in float64;
    returns (params, m, v)."""
    p = np.asarray(params, dtype=np.float64)
    m = 0.9 * m + 0.1 * grad
    v = 0.999 * v + 0.001 * grad * grad
    b1c = 1.0 - 0.9 ** step
    b2c = 1.0 - 0.999 ** step
    return p - lr * (m / b1c) / (np.sqrt(v / b2c) + 1e-8), m, v


def he_params(k: int, seed: int) -> np.ndarray:
    """A random parameter vector of the right size (numpy PCG64), He scale."""
    o = shape(k)
    rng = np.random.default_rng(seed)
    p = np.zeros(o["total"], dtype=np.float32)
    p[o["W1"]:o["b1"]] = rng.normal(0, np.sqrt(2.0 / k), o["b1"] - o["W1"])
    p[o["b1"]:o["W2"]] = rng.normal(0, 0.1, H1)
    p[o["W2"]:o["b2"]] = rng.normal(0, np.sqrt(2.0 / H1), o["b2"] - o["W2"])
    p[o["b2"]:o["w3"]] = rng.normal(0, 0.1, H2)
    p[o["w3"]:o["b3"]] = rng.normal(0, np.sqrt(1.0 / H2), H2)
    p[o["b3"]] = 0.05
    return p


# ---- Bloom filter: blocks chosen by the key's minimizer ----------------------
# The GPU filter's scheme, restated (sbwt_gpu_amd/csrc/mlp.hip, bloom_locate):
# a key's bits lie in one 16-byte pair of 64-bit words inside a 128-byte block
# of 16 words. block = high 64 bits of mix64(minimizer ^ BLOCK_SEED) * blocks,
# the minimizer being the key's m-mer code c (m = min(k, 16)) with the
# smallest o(c) = x ^ (x >> 16), x = c * 0x9E3779 mod 2^32 (leftmost on
# ties); h = mix64(key ^ golden), pair = h >> 61; bit i (i < hashes) = bits
# 7i..7i+6 of mix64(h), over the pair's 128 bits (bits 0..63 in the pair's
# first word). Sized by bloom_size: the smallest bits/key (grid of 1/20 of
# the unblocked -ln p / ln^2 2) and hash count 1..9 whose Poisson-averaged
# false-positive rate over 128-bit blocks reaches p (Putze, Sanders, Singler
# 2007, register-blocked filters), times CLUMP_MARGIN for the keys that share
# a minimizer and so land in one block together.

M64 = (1 << 64) - 1
GOLDEN = 0x9E3779B97F4A7C15
BLOCK_SEED = 0x632BE59BD9B4E019
BLOCK_WORDS = 16
CLUMP_MARGIN = 1.2
MAX_HASHES = 9


def mix64(x: int) -> int:
    """splitmix64's finaliser (Steele, Lea, Flood 2014)."""
    x ^= x >> 30
    x = (x * 0xBF58476D1CE4E5B9) & M64
    x ^= x >> 27
    x = (x * 0x94D049BB133111EB) & M64
    x ^= x >> 31
    return x


def mix64_np(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64).copy()
    with np.errstate(over="ignore"):
        x ^= x >> np.uint64(30)
        x *= np.uint64(0xBF58476D1CE4E5B9)
        x ^= x >> np.uint64(27)
        x *= np.uint64(0x94D049BB133111EB)
        x ^= x >> np.uint64(31)
    return x


def mulhi64_np(a: np.ndarray, b: int) -> np.ndarray:
    """High 64 bits of the 128-bit products a * b (uint64)."""
    m32 = np.uint64(0xFFFFFFFF)
    a = np.asarray(a, dtype=np.uint64)
    b = np.uint64(b)
    a_lo, a_hi = a & m32, a >> np.uint64(32)
    b_lo, b_hi = b & m32, b >> np.uint64(32)
    p0, p1, p2, p3 = a_lo * b_lo, a_lo * b_hi, a_hi * b_lo, a_hi * b_hi
    mid = (p0 >> np.uint64(32)) + (p1 & m32) + (p2 & m32)
    return p3 + (p1 >> np.uint64(32)) + (p2 >> np.uint64(32)) + (mid >> np.uint64(32))


def minimizer_order(c: np.ndarray) -> np.ndarray:
    """o(c) = x ^ (x >> 16), x = c * 0x9E3779 mod 2^32."""
    x = (np.asarray(c, dtype=np.uint64) * np.uint64(0x9E3779)) & np.uint64(0xFFFFFFFF)
    return x ^ (x >> np.uint64(16))


def minimizer_np(keys: np.ndarray, k: int) -> np.ndarray:
    """The minimizer code of every top-aligned key (leftmost on ties)."""
    m = min(k, 16)
    cmask = np.uint64((1 << (2 * m)) - 1)
    keys = np.asarray(keys, dtype=np.uint64)
    best = (keys >> np.uint64(64 - 2 * m)) & cmask
    bo = minimizer_order(best)
    for i in range(1, k - m + 1):
        c = (keys >> np.uint64(64 - 2 * m - 2 * i)) & cmask
        o = minimizer_order(c)
        sel = o < bo
        best = np.where(sel, c, best)
        bo = np.where(sel, o, bo)
    return best


def bloom_fpr(bits_per_key: float, h: int) -> float:
    """Poisson-averaged false-positive rate of 128-bit register blocks."""
    import math

    lam = 128.0 / bits_per_key
    per = math.log1p(-1.0 / 128.0) * h
    pmf, s = math.exp(-lam), 0.0
    nmax = int(lam + 12.0 * math.sqrt(lam) + 24.0)
    for n in range(nmax + 1):
        s += pmf * (1.0 - math.exp(per * n)) ** h
        pmf *= lam / (n + 1)
    return s


def bloom_model_size(p: float):
    """(bits per key, hashes) at which the Poisson model reaches p."""
    import math

    base = -math.log(p) / (math.log(2) * math.log(2))
    best, hashes = 0.0, 1
    for h in range(1, MAX_HASHES + 1):
        for j in range(201):
            bpk = base * (1.0 + 0.05 * j)
            if bloom_fpr(bpk, h) <= p:
                if best == 0.0 or bpk < best:
                    best, hashes = bpk, h
                break
    if best == 0.0:
        best, hashes = base * 11.0, MAX_HASHES
    return best, hashes


def bloom_size(p: float):
    """(bits per key, hashes) of a filter for false-positive rate p."""
    bpk, h = bloom_model_size(p)
    return bpk * CLUMP_MARGIN, h


def bloom_geometry(expected: int, fpr: float):
    """(128-byte blocks, hashes) of a filter for `expected` keys."""
    import math

    bpk, h = bloom_size(fpr)
    return (int(math.ceil(expected * bpk / (64.0 * BLOCK_WORDS))) if expected else 0), h


def key_mask(k: int) -> int:
    return M64 if k == 32 else (M64 ^ ((1 << (64 - 2 * k)) - 1))


def bloom_locate(keys: np.ndarray, blocks: int, hashes: int, k: int):
    """(index of the key's pair's first word, mask of that word, mask of the
    next word) of every key (numpy uint64 arrays)."""
    key = np.asarray(keys, dtype=np.uint64) & np.uint64(key_mask(k))
    blk = mulhi64_np(mix64_np(minimizer_np(key, k) ^ np.uint64(BLOCK_SEED)), blocks)
    h = mix64_np(key ^ np.uint64(GOLDEN))
    h2 = mix64_np(h)
    lo = np.zeros_like(h)
    hi = np.zeros_like(h)
    for i in range(hashes):
        b = (h2 >> np.uint64(7 * i)) & np.uint64(127)
        bit = np.uint64(1) << (b & np.uint64(63))
        first = b < np.uint64(64)
        lo |= np.where(first, bit, np.uint64(0))
        hi |= np.where(first, np.uint64(0), bit)
    return blk * np.uint64(BLOCK_WORDS) + (h >> np.uint64(61)) * np.uint64(2), lo, hi


def bloom_build(keys: np.ndarray, blocks: int, hashes: int, k: int) -> np.ndarray:
    w, lo, hi = bloom_locate(keys, blocks, hashes, k)
    out = np.zeros(blocks * BLOCK_WORDS, dtype=np.uint64)
    np.bitwise_or.at(out, w.astype(np.int64), lo)
    np.bitwise_or.at(out, (w + np.uint64(1)).astype(np.int64), hi)
    return out


def bloom_query(bits: np.ndarray, keys: np.ndarray, hashes: int, k: int) -> np.ndarray:
    w, lo, hi = bloom_locate(keys, bits.size // BLOCK_WORDS, hashes, k)
    wi = w.astype(np.int64)
    return ((bits[wi] & lo) == lo) & ((bits[wi + 1] & hi) == hi)
