"""Seeded synthetic genomes and reads for the benchmark workloads.

All draws come from numpy's PCG64 with fixed seeds, so every run builds the
same index and queries the same reads.
"""
from __future__ import annotations

import numpy as np

ACGT = np.frombuffer(b"ACGT", dtype=np.uint8)


def genome_codes(length: int, seed: int, gc: float = 0.5, seg_frac: float = 0.0,
                 seg_len=(1_000, 50_000), seg_mut: float = 0.01) -> np.ndarray:
    """2-bit codes (0..3 = A,C,G,T) of a genome with the given GC fraction;
    seg_frac of it is then overwritten by copies of random source segments
    (lengths uniform in seg_len) carrying i.i.d. substitutions at seg_mut."""
    rng = np.random.default_rng(seed)
    at = (1.0 - gc) / 2
    codes = np.empty(length, dtype=np.uint8)
    chunk = 1 << 24
    for s in range(0, length, chunk):
        e = min(length, s + chunk)
        codes[s:e] = rng.choice(4, size=e - s, p=[at, gc / 2, gc / 2, at])
    covered, target = 0, int(seg_frac * length)
    while covered < target:
        L = int(rng.integers(seg_len[0], seg_len[1] + 1))
        L = min(L, length // 2)
        src = int(rng.integers(0, length - L))
        dst = int(rng.integers(0, length - L))
        seg = codes[src:src + L].copy()
        flip = np.flatnonzero(rng.random(L) < seg_mut)
        seg[flip] = (seg[flip] + rng.integers(1, 4, flip.size)) % 4
        codes[dst:dst + L] = seg
        covered += L
    return codes


def to_ascii(codes: np.ndarray) -> np.ndarray:
    return ACGT[codes]


def reads(genome_codes_: np.ndarray, count: int, length: int, seed: int, sub: float = 0.01,
          n_rate: float = 0.0):
    """`count` reads of `length` from uniform positions, each base substituted
    (by one of the three other bases) with probability `sub`, and set to N
    with probability `n_rate`. Returns (ASCII uint8 array of the concatenated
    reads, row starts int64[count+1])."""
    rng = np.random.default_rng(seed)
    G = genome_codes_.size
    out = np.empty(count * length, dtype=np.uint8)
    step = max(1, (1 << 24) // length)
    ar = np.arange(length, dtype=np.int64)
    for s in range(0, count, step):
        e = min(count, s + step)
        pos = rng.integers(0, G - length + 1, e - s)
        c = genome_codes_[pos[:, None] + ar[None, :]]
        flip = rng.random(c.shape) < sub
        c[flip] = (c[flip] + rng.integers(1, 4, int(flip.sum()))) % 4
        a = ACGT[c]
        if n_rate > 0:
            a[rng.random(a.shape) < n_rate] = ord("N")
        out[s * length:e * length] = a.reshape(-1)
    starts = np.arange(count + 1, dtype=np.int64) * length
    return out, starts


def lognormal_read_sets(genome_codes_: np.ndarray, sets: int, seed: int, max_reads: int = 5000,
                        median: float = 300.0, sigma: float = 0.8, lo: int = 100, hi: int = 20_000,
                        sub: float = 0.01, n_rate: float = 0.05, max_chars: int | None = None):
    """Read sets of uneven size: reads per set ~ U{1..max_reads}, read length
    ~ lognormal(ln median, sigma) clipped to [lo, hi], substitutions at `sub`,
    N at `n_rate`. Returns (ASCII chars, row starts, set starts over rows)."""
    rng = np.random.default_rng(seed)
    per_set = rng.integers(1, max_reads + 1, sets)
    nrows = int(per_set.sum())
    lens = np.clip(np.exp(rng.normal(np.log(median), sigma, nrows)).astype(np.int64), lo, hi)
    if max_chars is not None and lens.sum() > max_chars:
        keep = int(np.searchsorted(np.cumsum(lens), max_chars))
        lens = lens[:keep]
        cs = np.cumsum(per_set)
        per_set = np.diff(np.concatenate([[0], np.minimum(cs, keep)]))
        per_set = per_set[per_set > 0]
        nrows = keep
    starts = np.zeros(nrows + 1, dtype=np.int64)
    starts[1:] = np.cumsum(lens)
    out = np.empty(int(starts[-1]), dtype=np.uint8)
    G = genome_codes_.size
    for r in range(0, nrows, 4096):
        e = min(nrows, r + 4096)
        for i in range(r, e):
            L = int(lens[i])
            p = int(rng.integers(0, G - L + 1))
            out[starts[i]:starts[i + 1]] = genome_codes_[p:p + L]
    codes = out
    flip = rng.random(codes.size) < sub
    codes[flip] = (codes[flip] + rng.integers(1, 4, int(flip.sum()))) % 4
    chars = ACGT[codes]
    if n_rate > 0:
        chars[rng.random(chars.size) < n_rate] = ord("N")
    set_starts = np.zeros(per_set.size + 1, dtype=np.int64)
    set_starts[1:] = np.cumsum(per_set)
    return chars, starts, set_starts


def gpu_read_sets(genome_codes_dev, sets: int, seed: int, max_reads: int = 5000, median: float = 300.0,
                  sigma: float = 0.8, lo: int = 100, hi: int = 20_000, sub: float = 0.01, n_rate: float = 0.05,
                  chunk_chars: int = 1 << 25):
    """Uneven read sets generated on the GPU (torch, fixed seed): reads per
    set ~ U{1..max_reads}, read length ~ lognormal(ln median, sigma) clipped
    to [lo, hi], each read from a uniform position of the genome (2-bit codes
    on the device), substitutions at `sub`, N at `n_rate`. Returns (ASCII
    chars uint8, row starts int64[rows+1], set starts int64[sets+1]) on the
    device."""
    import torch

    dev = genome_codes_dev.device
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    G = genome_codes_dev.numel()
    per_set = torch.randint(1, max_reads + 1, (sets,), generator=gen, device=dev)
    set_starts = torch.zeros(sets + 1, dtype=torch.int64, device=dev)
    set_starts[1:] = torch.cumsum(per_set, 0)
    nrows = int(set_starts[-1].item())
    lens = torch.exp(torch.randn(nrows, generator=gen, device=dev) * sigma + float(np.log(median)))
    lens = lens.clamp(lo, hi).to(torch.int64)
    starts = torch.zeros(nrows + 1, dtype=torch.int64, device=dev)
    starts[1:] = torch.cumsum(lens, 0)
    total = int(starts[-1].item())
    pos = (torch.rand(nrows, generator=gen, device=dev, dtype=torch.float64) * (G - lens + 1)).to(torch.int64)
    chars = torch.empty(total, dtype=torch.uint8, device=dev)
    lut = torch.tensor(list(b"ACGT"), dtype=torch.uint8, device=dev)
    starts_h = starts.cpu().numpy()
    r0 = 0
    while r0 < nrows:
        r1 = int(np.searchsorted(starts_h, starts_h[r0] + chunk_chars, side="right")) - 1
        r1 = max(r1, r0 + 1)
        r1 = min(r1, nrows)
        c0, c1 = int(starts_h[r0]), int(starts_h[r1])
        L = lens[r0:r1]
        base = torch.repeat_interleave(pos[r0:r1] - (starts[r0:r1] - c0), L)
        src = base + torch.arange(c1 - c0, device=dev)
        codes = genome_codes_dev[src]
        flip = torch.rand(c1 - c0, generator=gen, device=dev) < sub
        codes = torch.where(flip, (codes + torch.randint(1, 4, (c1 - c0,), generator=gen, device=dev,
                                                           dtype=torch.uint8)) % 4, codes)
        a = lut[codes.to(torch.int64)]
        if n_rate > 0:
            a = torch.where(torch.rand(c1 - c0, generator=gen, device=dev) < n_rate,
                            torch.tensor(ord("N"), dtype=torch.uint8, device=dev), a)
        chars[c0:c1] = a
        r0 = r1
    return chars, starts, set_starts


def gpu_haplotypes(base_codes_dev, count: int, seed: int, snp: float = 1e-3, indel: float = 1e-4,
                   max_indel: int = 10):
    """`count` haplotypes of a base genome (2-bit codes on the device): each
    base substituted (by one of the three others) with probability `snp`,
    and indels at rate `indel` per base, half insertions of 1..max_indel
    random bases and half deletions of 1..max_indel bases, at uniform
    positions (seeded). Returns (ASCII chars of the haplotypes concatenated,
    row starts int64[count+1]) on the device, one row per haplotype."""
    import torch

    dev = base_codes_dev.device
    G = base_codes_dev.numel()
    gen = torch.Generator(device=dev)
    gen.manual_seed(seed)
    rng = np.random.default_rng(seed)
    lut = torch.tensor(list(b"ACGT"), dtype=torch.uint8, device=dev)
    haps, lens = [], []
    for _ in range(count):
        flip = torch.rand(G, generator=gen, device=dev) < snp
        delta = torch.randint(1, 4, (G,), generator=gen, device=dev, dtype=torch.uint8)
        codes = torch.where(flip, (base_codes_dev + delta) % 4, base_codes_dev)
        n = int(rng.binomial(G, indel))
        pos = np.unique(rng.integers(1, G - max_indel, n))
        L = rng.integers(1, max_indel + 1, pos.size)
        ins = rng.random(pos.size) < 0.5
        pool = torch.randint(0, 4, (int(L[ins].sum()) + 1,), generator=gen, device=dev, dtype=torch.uint8)
        segs, prev, used = [], 0, 0
        for p, ln, is_ins in zip(pos.tolist(), L.tolist(), ins.tolist()):
            if p < prev:  # inside the previous deletion
                continue
            segs.append(codes[prev:p])
            if is_ins:
                segs.append(pool[used:used + ln])
                used += ln
                prev = p
            else:
                prev = p + ln
        segs.append(codes[prev:])
        h = lut[torch.cat(segs).to(torch.int64)]
        haps.append(h)
        lens.append(h.numel())
        del codes, flip, delta
    starts = torch.zeros(count + 1, dtype=torch.int64, device=dev)
    starts[1:] = torch.cumsum(torch.tensor(lens, dtype=torch.int64, device=dev), 0)
    return torch.cat(haps), starts
