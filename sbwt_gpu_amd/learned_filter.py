"""Training and assembly of the learned negative prefilter.

Positives are k-mers the SBWT itself enumerates (a backward label walk);
negatives are random k-mers (uniform, or at the index's base composition)
that the plain search confirms absent.
After training, the model runs over every indexed k-mer; those it scores
below the threshold (its false negatives) go into a Bloom filter, so the
filter's "maybe" covers every indexed k-mer and the exact search after it
gives the same answers as without it. With calibration the threshold comes
from held-out k-mers and the kernels' measured costs, and the network (or
the whole filter) is left out where it does not pay.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import torch

from .api import Bloom, Index, Mlp, prefilter

NONE64 = -1  # ~0 as int64


@dataclass
class Prefilter:
    """maybe = network logit >= threshold, or a Bloom-filter hit. After
    calibration the network may be left out (mlp None: the Bloom filter holds
    every indexed k-mer) and so may the Bloom filter (both None: every k-mer
    is searched)."""
    mlp: Mlp | None
    threshold: float
    bloom: Bloom | None
    stats: dict = field(default_factory=dict)
    net: Mlp | None = None  # the trained network, used or not

    def maybe(self, kmers: torch.Tensor, out: torch.Tensor | None = None) -> torch.Tensor:
        if self.mlp is not None:
            return prefilter(self.mlp, self.threshold, self.bloom, kmers, out)
        if out is None:
            out = torch.empty(kmers.numel(), dtype=torch.uint8, device="cuda")
        if self.bloom is None:
            return out.fill_(1)
        return self.bloom.query(kmers, out=out)


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


def _auc(zp: torch.Tensor, zn: torch.Tensor) -> float:
    """P(logit of a positive > logit of a negative), ties half (Mann-Whitney)."""
    z = torch.cat([zp, zn]).double()
    order = torch.argsort(z)
    ranks = torch.empty_like(z)
    ranks[order] = torch.arange(1, z.numel() + 1, device=z.device, dtype=torch.float64)
    # average ranks of tied logits
    zs = z[order]
    uniq, inv, counts = torch.unique_consecutive(zs, return_inverse=True, return_counts=True)
    first = torch.cumsum(counts, 0) - counts + 1
    avg = (first + first + counts - 1).double() / 2
    ranks[order] = avg[inv]
    npos, nneg = zp.numel(), zn.numel()
    return float((ranks[:npos].sum() - npos * (npos + 1) / 2) / (npos * nneg))


def _ms_per_kmer(fn, n: int, reps: int = 3) -> float:
    fn()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(torch.cuda.current_stream())
    for _ in range(reps):
        fn()
    e1.record(torch.cuda.current_stream())
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps / n


def calibrate(index: Index, mlp: Mlp, pos: torch.Tensor, neg: torch.Tensor, fpr: float,
              negative_fraction: float) -> dict:
    """The cheapest filter for a query stream with `negative_fraction`
    absent k-mers, from held-out positives and negatives: the network at a
    threshold t (rejected k-mers then go to a Bloom filter over the indexed
    k-mers the network rejects), the Bloom filter alone (every indexed k-mer
    in it), or no filter. The expected cost per query k-mer of each option
    adds the kernels' measured costs per k-mer: the network (when used), the
    Bloom lookup (of every k-mer, or of those the network rejects) and the
    exact search of the k-mers that stay "maybe"."""
    z_pos, z_neg = mlp.logits(pos), mlp.logits(neg)
    f_neg, f_pos = negative_fraction, 1.0 - negative_fraction
    probe = Bloom(index.k, pos.numel(), fpr)
    probe.insert(pos)
    c_model = _ms_per_kmer(lambda: mlp.logits(neg), neg.numel())
    c_bloom = _ms_per_kmer(lambda: probe.query(neg), neg.numel())
    c_neg = _ms_per_kmer(lambda: index.search_kmers(neg), neg.numel())
    c_pos = _ms_per_kmer(lambda: index.search_kmers(pos), pos.numel())
    options = [{"option": "no filter", "cost": f_neg * c_neg + f_pos * c_pos, "threshold": None},
               {"option": "bloom only", "cost": c_bloom + f_neg * fpr * c_neg + f_pos * c_pos, "threshold": None}]
    qs = torch.tensor([0.5, 0.8, 0.9, 0.95, 0.98, 0.99, 0.995], device="cuda", dtype=torch.float32)
    for t in torch.quantile(z_neg.float()[: 1 << 24], qs).tolist():
        pass_neg = float((z_neg >= t).float().mean())
        pass_pos = float((z_pos >= t).float().mean())
        reject = f_neg * (1 - pass_neg) + f_pos * (1 - pass_pos)
        maybe_neg = pass_neg + (1 - pass_neg) * fpr
        cost = c_model + c_bloom * reject + f_neg * maybe_neg * c_neg + f_pos * c_pos
        options.append({"option": "network", "threshold": t, "cost": cost, "pass_neg": round(pass_neg, 4),
                        "pass_pos": round(pass_pos, 4)})
    best = min(options, key=lambda o: o["cost"])
    return {"best": best, "options": options, "auc": round(_auc(z_pos, z_neg), 4),
            "ms_per_kmer": {"network": c_model, "bloom": c_bloom, "search_negative": c_neg,
                            "search_positive": c_pos}, "negative_fraction": negative_fraction}


def train_prefilter(index: Index, *, positives: int = 50_000_000, negatives: int = 50_000_000, epochs: int = 1,
                    batch: int = 65536, lr: float = 2e-3, seed: int = 7, threshold: float = 0.0,
                    fpr: float = 0.01, train: bool = True, negatives_like: str = "uniform",
                    calibrate_for: float | None = None, log=None) -> Prefilter:
    """negatives_like: "uniform" (uniform random k-mers) or "index" (i.i.d.
    characters at the index's base composition); either way each negative
    is confirmed absent by the plain search. calibrate_for: the fraction of
    absent k-mers in the queries the filter will see; the threshold (and
    whether the network or the Bloom filter is used at all) is then chosen
    by `calibrate` on held-out k-mers instead of `threshold`."""
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
    cal = None
    use_model, use_bloom = True, True
    if calibrate_for is not None:
        nh = min(2_000_000, npos_all)
        hpos = allpos[torch.randint(0, max(npos_all, 1), (nh,), generator=gen, device="cuda")]
        hneg = random_kmers(int(nh * 1.01) + 16, k, gen, comp)
        hneg = hneg[index.search_kmers(hneg) == NONE64][:nh]
        cal = calibrate(index, mlp, hpos, hneg, fpr, calibrate_for)
        best = cal["best"]
        use_model = best["option"] == "network"
        use_bloom = best["option"] != "no filter"
        threshold = best["threshold"] if use_model else float("inf")
        del hpos, hneg
        if log:
            log(f"prefilter calibration: {best['option']} (threshold {best['threshold']}), network AUC {cal['auc']}")
    # The model's false negatives over every indexed k-mer go into the Bloom
    # filter: `~(z >= threshold)` is the exact complement of what the kernel
    # keeps, so a NaN logit goes into the filter too. Two passes in chunks
    # (count, then insert) keep the memory at one chunk of logits.
    nfn = 0
    if use_model:
        for a in range(0, npos_all, CHUNK):
            z = mlp.logits(allpos[a:a + CHUNK])
            nfn += int((~(z >= threshold)).sum().item())
    elif use_bloom:
        nfn = npos_all  # the network is left out: the Bloom filter holds every indexed k-mer
    bloom = Bloom(k, max(nfn, 1), fpr) if nfn else None
    if bloom is not None:
        for a in range(0, npos_all, CHUNK):
            part = allpos[a:a + CHUNK]
            bloom.insert(part[~(mlp.logits(part) >= threshold)] if use_model else part)
    pf = Prefilter(mlp if use_model else None, threshold, bloom, net=mlp)
    missed = 0
    for a in range(0, npos_all, CHUNK):
        missed += int((pf.maybe(allpos[a:a + CHUNK]) != 1).sum().item())
    del allpos
    if missed:
        raise RuntimeError(f"the prefilter rejects {missed} indexed k-mers; refusing a filter with false negatives")
    zn = mlp.logits(neg[: min(neg.numel(), 10_000_000)])
    t_net = threshold if use_model else 0.0  # the network's own pass rate, used or not
    stats = {"indexed_kmers": npos_all, "train_positives": min(positives, npos_all), "train_negatives": neg.numel(),
             "negatives_like": negatives_like, "epochs": epochs if train else 0, "batch": batch, "lr": lr,
             "losses": losses, "data_s": round(t_data, 2), "train_s": round(t_train, 2),
             "model_false_negatives": nfn if use_model else None,
             "model_recall_on_index": round(1 - nfn / max(npos_all, 1), 4) if use_model else None,
             "model_pass_rate_on_training_negatives": round(float((zn >= t_net).float().mean().item()), 4),
             "bloom_bytes": bloom.bytes if bloom else 0, "zero_false_negatives_on_index": True,
             "network_used": use_model, "bloom_used": bloom is not None,
             "threshold": threshold if use_model else None,
             "calibration": cal}
    pf.stats = stats
    return pf
