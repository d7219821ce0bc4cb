"""Training and assembly of the learned negative prefilter.

Positives are k-mers the SBWT itself enumerates (a backward label walk);
negatives are random k-mers (uniform, or at the index's base composition)
that the plain search confirms absent.
After training, the model runs over every indexed k-mer; those it scores
below the threshold (its false negatives) go into a Bloom filter, so the
filter's "maybe" covers every indexed k-mer and the exact search after it
gives the same answers as without it.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import torch

from .api import Bloom, Index, Mlp, prefilter

NONE64 = -1  # ~0 as int64


@dataclass
class Prefilter:
    mlp: Mlp
    threshold: float
    bloom: Bloom | None
    stats: dict = field(default_factory=dict)

    def maybe(self, kmers: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        return prefilter(self.mlp, self.threshold, self.bloom, kmers, out)


def random_kmers(n: int, k: int, gen: torch.Generator, composition=None) -> torch.Tensor:
    """n random top-aligned k-mers (int64): uniform characters, or i.i.d.
    characters drawn from `composition` (four probabilities for A, C, G, T)."""
    if composition is None:
        hi = torch.randint(0, 1 << 31, (n,), generator=gen, device="cuda", dtype=torch.int64)
        lo = torch.randint(0, 1 << 31, (n,), generator=gen, device="cuda", dtype=torch.int64)
        x = (hi << 33) ^ (lo << 2)
        mask = ~((1 << (64 - 2 * k)) - 1) if k < 32 else -1
        return x & mask
    p = torch.tensor(composition, dtype=torch.float64, device="cuda")
    cdf = torch.cumsum(p / p.sum(), 0)[:3].to(torch.float32)
    x = torch.zeros(n, dtype=torch.int64, device="cuda")
    for t in range(k):
        u = torch.rand(n, generator=gen, device="cuda")
        c = (u[:, None] >= cdf[None, :]).sum(1).to(torch.int64)
        x |= c << (62 - 2 * t)
    return x


def index_composition(index: Index) -> list:
    """Share of A, C, G, T as the last character of the index's nodes
    (C[c+1] - C[c]): the base composition of its k-mers."""
    C = index.info["C"]
    return [(C[c + 1] - C[c]) / max(C[4] - C[0], 1) for c in range(4)]


CHUNK = 1 << 28  # k-mers per pass over the enumerated index


def train_prefilter(index: Index, *, positives: int = 50_000_000, negatives: int = 50_000_000, epochs: int = 1,
                    batch: int = 65536, lr: float = 2e-3, seed: int = 7, threshold: float = 0.0,
                    fpr: float = 0.01, train: bool = True, negatives_like: str = "uniform", log=None) -> Prefilter:
    """negatives_like: "uniform" (uniform random k-mers) or "index" (i.i.d.
    characters at the index's base composition); either way each negative
    is confirmed absent by the plain search."""
    k = index.k
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    t0 = time.time()
    allpos = index.enumerate_kmers()
    npos_all = allpos.numel()
    sel = torch.randint(0, max(npos_all, 1), (min(positives, npos_all),), generator=gen, device="cuda")
    pos = allpos[sel]
    del sel
    comp = index_composition(index) if negatives_like == "index" else None
    neg = random_kmers(int(negatives * 1.01) + 16, k, gen, comp)
    found = index.search_kmers(neg)
    neg = neg[found == NONE64][:negatives]
    del found
    x = torch.cat([pos, neg])
    y = torch.cat([torch.ones(pos.numel(), device="cuda"), torch.zeros(neg.numel(), device="cuda")])
    t_data = time.time() - t0
    mlp = Mlp(k, seed=seed, max_batch=batch)
    losses = []
    t1 = time.time()
    if train:
        n = x.numel()
        for ep in range(epochs):
            perm = torch.randperm(n, generator=gen, device="cuda")
            for b0 in range(0, n - batch + 1, batch):
                idx = perm[b0:b0 + batch]
                loss = mlp.train_step(x[idx].contiguous(), y[idx].contiguous(), lr)
                if (b0 // batch) % 200 == 0:
                    losses.append(float(loss.item()))
                    if log:
                        log(f"prefilter train epoch {ep} step {b0 // batch}: loss {losses[-1]:.4f}")
    torch.cuda.synchronize()
    t_train = time.time() - t1
    del x, y, pos
    # The model's false negatives over every indexed k-mer go into the Bloom
    # filter: `~(z >= threshold)` is the exact complement of what the kernel
    # keeps, so a NaN logit goes into the filter too. Two passes in chunks
    # (count, then insert) keep the memory at one chunk of logits.
    nfn = 0
    for a in range(0, npos_all, CHUNK):
        z = mlp.logits(allpos[a:a + CHUNK])
        nfn += int((~(z >= threshold)).sum().item())
    bloom = Bloom(k, max(nfn, 1), fpr) if nfn else None
    if bloom is not None:
        for a in range(0, npos_all, CHUNK):
            part = allpos[a:a + CHUNK]
            bloom.insert(part[~(mlp.logits(part) >= threshold)])
    missed = 0
    for a in range(0, npos_all, CHUNK):
        missed += int((prefilter(mlp, threshold, bloom, allpos[a:a + CHUNK]) != 1).sum().item())
    del allpos
    if missed:
        raise RuntimeError(f"the prefilter rejects {missed} indexed k-mers; refusing a filter with false negatives")
    zn = mlp.logits(neg[: min(neg.numel(), 10_000_000)])
    stats = {"indexed_kmers": npos_all, "train_positives": min(positives, npos_all), "train_negatives": neg.numel(),
             "negatives_like": negatives_like, "epochs": epochs if train else 0, "batch": batch, "lr": lr,
             "losses": losses, "data_s": round(t_data, 2), "train_s": round(t_train, 2),
             "model_false_negatives": nfn, "model_recall_on_index": round(1 - nfn / max(npos_all, 1), 4),
             "model_pass_rate_on_training_negatives": round(float((zn >= threshold).float().mean().item()), 4),
             "bloom_bytes": bloom.bytes if bloom else 0, "zero_false_negatives_on_index": True}
    return Prefilter(mlp, threshold, bloom, stats)
