"""Training and assembly of the learned negative prefilter.

Positives are k-mers the SBWT itself enumerates (select-based label walk);
negatives are uniform random k-mers that the plain search confirms absent.
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


def random_kmers(n: int, k: int, gen: torch.Generator) -> torch.Tensor:
    """n uniform random top-aligned k-mers (int64)."""
    hi = torch.randint(0, 1 << 31, (n,), generator=gen, device="cuda", dtype=torch.int64)
    lo = torch.randint(0, 1 << 31, (n,), generator=gen, device="cuda", dtype=torch.int64)
    x = (hi << 33) ^ (lo << 2)
    mask = ~((1 << (64 - 2 * k)) - 1) if k < 32 else -1
    return x & mask


def train_prefilter(index: Index, *, positives: int = 50_000_000, negatives: int = 50_000_000, epochs: int = 1,
                    batch: int = 65536, lr: float = 2e-3, seed: int = 7, threshold: float = 0.0,
                    fpr: float = 0.01, train: bool = True, log=None) -> Prefilter:
    k = index.k
    gen = torch.Generator(device="cuda")
    gen.manual_seed(seed)
    t0 = time.time()
    allpos = index.enumerate_kmers()
    npos_all = allpos.numel()
    sel = torch.randperm(npos_all, generator=gen, device="cuda")[: min(positives, npos_all)]
    pos = allpos[sel]
    neg = random_kmers(int(negatives * 1.01) + 16, k, gen)
    found = index.search_kmers(neg)
    neg = neg[found == NONE64][:negatives]
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
    # the model's false negatives over every indexed k-mer go into the Bloom filter
    z = mlp.logits(allpos)
    # the exact complement of what the kernel keeps (z >= threshold), so a
    # NaN logit goes into the Bloom filter too
    fn = allpos[~(z >= threshold)]
    bloom = Bloom(k, max(fn.numel(), 1), fpr) if fn.numel() else None
    if bloom is not None:
        bloom.insert(fn)
    m = prefilter(mlp, threshold, bloom, allpos)
    zero_fn = bool((m == 1).all().item())
    if not zero_fn:
        raise RuntimeError(f"the prefilter rejects {int((m != 1).sum().item())} indexed k-mers; "
                           "refusing a filter with false negatives")
    zn = mlp.logits(neg[: min(neg.numel(), 10_000_000)])
    stats = {"indexed_kmers": npos_all, "train_positives": pos.numel(), "train_negatives": neg.numel(),
             "epochs": epochs if train else 0, "batch": batch, "lr": lr, "losses": losses,
             "data_s": round(t_data, 2), "train_s": round(t_train, 2),
             "model_false_negatives": int(fn.numel()),
             "model_recall_on_index": round(1 - fn.numel() / max(npos_all, 1), 4),
             "model_pass_rate_on_random_negatives": round(float((zn >= threshold).float().mean().item()), 4),
             "bloom_bytes": bloom.bytes if bloom else 0, "zero_false_negatives_on_index": zero_fn}
    return Prefilter(mlp, threshold, bloom, stats)
