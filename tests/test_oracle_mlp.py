"""Pins of the prefilter MLP's CPU oracle: finite differences, torch autograd
over an independently written forward, torch's bf16 rounding and Adam, and
the published splitmix64 test vector."""
import numpy as np
import pytest
import torch

from oracle import mlp as om


def _data(k, B, seed):
    rng = np.random.default_rng(seed)
    km = rng.integers(0, 1 << 62, B, dtype=np.uint64) << np.uint64(2)
    if k < 32:
        km &= np.uint64(~((1 << (64 - 2 * k)) - 1) & ((1 << 64) - 1))
    y = (rng.random(B) < 0.5).astype(np.float64)
    return km, y


def test_shape_layout():
    o = om.shape(31)
    assert o["in"] == 124
    assert o["total"] == 124 * 128 + 128 + 128 * 64 + 64 + 64 + 1


def test_bf16_round_matches_torch():
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 3, 100000), rng.normal(0, 1e-3, 1000), [0.0, -0.0, 1.0, 65504.0]])
    x = x.astype(np.float32)
    # exact ties (low 16 bits = 0x8000) on both sides of even/odd
    ties = (np.arange(64, dtype=np.uint32) << 16 | 0x3F800000 | 0x8000).view(np.float32)
    x = np.concatenate([x, ties, -ties])
    want = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(om.bf16_round(x), want)


def test_gradient_finite_differences():
    k, B = 6, 24
    km, y = _data(k, B, 2)
    p = om.he_params(k, 3).astype(np.float64)
    loss, g = om.loss_and_grad(p, km, y, k, rounded=False)
    rng = np.random.default_rng(4)
    o = om.shape(k)
    idx = np.concatenate([rng.choice(o["total"], 60, replace=False), [o["b3"], o["w3"], o["b2"], o["b1"]]])
    eps = 1e-6
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        lp, _ = om.loss_and_grad(pp, km, y, k, rounded=False)
        lm, _ = om.loss_and_grad(pm, km, y, k, rounded=False)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])


@pytest.mark.parametrize("k", [31, 32, 9])
def test_gradient_matches_torch_autograd(k):
    B = 200
    km, y = _data(k, B, 5 + k)
    p = om.he_params(k, 6)
    loss, g = om.loss_and_grad(p, km, y, k)
    W1, b1, W2, b2, w3, b3 = [torch.tensor(np.asarray(a, dtype=np.float64), requires_grad=True)
                              for a in om.unpack(p, k)]
    # dense one-hot input, written independently of the oracle's gather
    X = torch.zeros(B, 4 * k, dtype=torch.float64)
    for t in range(k):
        c = torch.from_numpy(((km >> np.uint64(62 - 2 * t)) & np.uint64(3)).astype(np.int64))
        X[torch.arange(B), 4 * t + c] = 1.0
    h1 = torch.relu(X @ W1 + b1)
    h2 = torch.relu(h1 @ W2 + b2)
    z = h2 @ w3 + b3
    L = torch.nn.functional.binary_cross_entropy_with_logits(z, torch.from_numpy(y))
    L.backward()
    o = om.shape(k)
    assert abs(L.item() - loss) < 1e-12
    got = {"W1": g[o["W1"]:o["b1"]], "b1": g[o["b1"]:o["W2"]], "W2": g[o["W2"]:o["b2"]],
           "b2": g[o["b2"]:o["w3"]], "w3": g[o["w3"]:o["b3"]], "b3": g[o["b3"]:]}
    for name, t in zip(["W1", "b1", "W2", "b2", "w3", "b3"], [W1, b1, W2, b2, w3, b3]):
        np.testing.assert_allclose(got[name], t.grad.numpy().reshape(-1), rtol=1e-9, atol=1e-14, err_msg=name)


def test_inference_rounds_h1_only():
    k = 31
    km, _ = _data(k, 64, 7)
    p = om.he_params(k, 8)
    zi = om.logits(p, km, k)
    zt, (a1, h1, a2, h2) = om.forward(p, km, k)
    # rounding h1 to bf16 moves the logit by at most a few bf16 ulps of the layer
    assert np.max(np.abs(zi - zt)) < 0.05 * (1 + np.max(np.abs(zt)))
    assert not np.array_equal(zi, zt)


def test_adam_matches_torch():
    rng = np.random.default_rng(9)
    p0 = rng.normal(size=50)
    grads = [rng.normal(size=50) for _ in range(3)]
    t = torch.tensor(p0, requires_grad=True)
    opt = torch.optim.Adam([t], lr=1e-2, betas=(0.9, 0.999), eps=1e-8)
    p, m, v = p0.copy(), np.zeros(50), np.zeros(50)
    for s, gr in enumerate(grads, 1):
        t.grad = torch.from_numpy(gr.copy())
        opt.step()
        p, m, v = om.adam_step(p, m, v, gr, s, 1e-2)
    np.testing.assert_allclose(p, t.detach().numpy(), rtol=1e-12, atol=1e-14)


def test_splitmix64_vector():
    # splitmix64 seeded with 0: the first output is 0xE220A8397B1DCDAF
    assert om.mix64(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF


def test_mulhi64_and_mix64_vectorised():
    rng = np.random.default_rng(3)
    a = rng.integers(0, 1 << 63, 2000, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    b = int(rng.integers(1, 1 << 62)) * 3
    got = om.mulhi64_np(a, b)
    want = np.array([(int(x) * b) >> 64 for x in a], dtype=np.uint64)
    np.testing.assert_array_equal(got, want)
    np.testing.assert_array_equal(om.mix64_np(a[:50]), np.array([om.mix64(int(x)) for x in a[:50]], np.uint64))


@pytest.mark.parametrize("p", [0.01, 0.001, 0.05])
def test_bloom_sizing_is_minimal_on_its_grid(p):
    bpk, h = om.bloom_model_size(p)
    assert 1 <= h <= 9
    assert om.bloom_fpr(bpk, h) <= p
    import math
    base = -math.log(p) / math.log(2) ** 2
    j = round(bpk / base / 0.05 - 20)
    if j > 0:
        assert om.bloom_fpr(base * (1 + 0.05 * (j - 1)), h) > p
    assert om.bloom_size(p) == (bpk * om.CLUMP_MARGIN, h)


def _minimizer_from_text(key: int, k: int) -> int:
    """The minimizer recomputed from the k-mer's characters: every m-mer as a
    base-4 number, ordered by x ^ (x >> 16) with x = code * 0x9E3779 mod
    2^32, the leftmost of the smallest."""
    text = "".join("ACGT"[(key >> (62 - 2 * t)) & 3] for t in range(k))
    m = min(k, 16)
    best = None
    for i in range(k - m + 1):
        code = int("".join(str("ACGT".index(ch)) for ch in text[i:i + m]), 4)
        x = (code * 0x9E3779) % (1 << 32)
        o = x ^ (x >> 16)
        if best is None or o < best[0]:
            best = (o, code)
    return best[1]


@pytest.mark.parametrize("k", [31, 32, 16, 9])
def test_minimizer_matches_text_restatement(k):
    rng = np.random.default_rng(k)
    keys = rng.integers(0, 1 << 63, 400, dtype=np.uint64) << np.uint64(1)
    keys &= np.uint64(om.key_mask(k))
    # repeats inside a k-mer make ties: a periodic key
    rep = int("0123" * 8, 4) << (64 - 2 * 32)
    keys = np.concatenate([keys, np.array([rep & om.key_mask(k), 0], np.uint64)])
    got = om.minimizer_np(keys, k)
    want = np.array([_minimizer_from_text(int(x), k) for x in keys], np.uint64)
    np.testing.assert_array_equal(got, want)


def test_bloom_blocks_follow_minimizers():
    """Consecutive k-mers of a sequence share their block whenever they share
    their minimizer: about 1 - 2 / (k - m + 2) of neighbours for random text."""
    rng = np.random.default_rng(8)
    k = 31
    seq = rng.integers(0, 4, 20_000).astype(np.uint64)
    n = seq.size - k + 1
    keys = np.zeros(n, np.uint64)
    for t in range(k):
        keys |= seq[t:t + n] << np.uint64(62 - 2 * t)
    blocks = 1 << 20
    w, _, _ = om.bloom_locate(keys, blocks, 5, k)
    blk = w // np.uint64(om.BLOCK_WORDS)
    mins = om.minimizer_np(keys, k)
    same_min = mins[1:] == mins[:-1]
    assert (blk[1:][same_min] == blk[:-1][same_min]).all()
    frac = same_min.mean()
    assert abs(frac - (1 - 2 / (k - 16 + 2))) < 0.03, frac


def test_bloom_fpr_formula_matches_simulation():
    """The Poisson model the sizing rests on predicts the measured rate of
    the filter built by the restated hash scheme (independent random keys,
    which do not share minimizers)."""
    rng = np.random.default_rng(5)
    k = 31
    keys = rng.integers(0, 1 << 62, 200_000, dtype=np.uint64) << np.uint64(2)
    blocks, h = om.bloom_geometry(keys.size, 0.01)
    bits = om.bloom_build(keys, blocks, h, k)
    assert bits.size == blocks * om.BLOCK_WORDS
    assert om.bloom_query(bits, keys, h, k).all()
    other = (rng.integers(0, 1 << 62, 1_000_000, dtype=np.uint64) << np.uint64(2)) | np.uint64(1 << 10)
    measured = om.bloom_query(bits, other, h, k).mean()
    predicted = om.bloom_fpr(bits.size * 64 / keys.size, h)
    assert abs(measured - predicted) < 0.15 * predicted, (measured, predicted)
    assert measured <= 0.01
    # low bits below the k-mer are ignored
    assert om.bloom_query(bits, keys | np.uint64(3), h, k).all()


def test_bloom_rate_on_sequence_kmers():
    """The k-mers of a sequence share minimizers in runs and so load blocks
    unevenly; with the clump margin the rate still meets the target."""
    rng = np.random.default_rng(11)
    k = 31
    seq = rng.integers(0, 4, 400_000).astype(np.uint64)
    n = seq.size - k + 1
    keys = np.zeros(n, np.uint64)
    for t in range(k):
        keys |= seq[t:t + n] << np.uint64(62 - 2 * t)
    blocks, h = om.bloom_geometry(keys.size, 0.01)
    bits = om.bloom_build(keys, blocks, h, k)
    assert om.bloom_query(bits, keys, h, k).all()
    other = rng.integers(0, 1 << 62, 1_000_000, dtype=np.uint64) << np.uint64(2)
    assert om.bloom_query(bits, other, h, k).mean() <= 0.01
