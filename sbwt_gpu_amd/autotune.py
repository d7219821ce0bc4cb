"""Autotuner for the read search.

Times the read search of a sample batch under every combination of rank
layout (16/32/64-byte chunks), presearch depth, segment length (k-mers per
work item), kernel (one lane per work item, or the persistent uniform-step
kernel) and work items (fixed segments of rows, or the rows' N-free runs),
checks that every combination gives the same answers, and leaves the index
in the fastest layout and depth. The reference fixes these by
flags (superblock size, presearch depth and threads per block,
include/parameters.hpp:39-58); here they are measured.
"""
from __future__ import annotations

import itertools
import time

import torch

from .api import Index, ReadsPlan


KERNELS = ("lane", "persistent", "lane-runs", "persistent-runs")


def kernel_flags(kernel: str) -> dict:
    """Search flags of a kernel name: one lane per work item ("lane") or the
    persistent uniform-step kernel ("persistent"); "-runs": the work items
    are the rows' N-free runs instead of fixed segments."""
    return {"persistent": kernel.startswith("persistent"), "lanes": kernel.startswith("lane"),
            "runs": kernel.endswith("-runs")}


def autotune(index: Index, chars: torch.Tensor, row_starts: torch.Tensor, *, layouts=(16, 32, 64),
             depths=(10, 11, 12, 13), segs=(64, 256, 1024), kernels=KERNELS, reps: int = 3,
             max_kmers: int | None = None, budget_s: float = 60.0, n_as_a: bool = False, streams: int = 1,
             log=None) -> dict:
    """chars / row_starts: a sample batch on the device (row_starts int64,
    offsets into chars). Returns {"best": {...}, "trials": [...]} and sets
    the index to the best layout and presearch depth.

    streams > 1 times every trial the way batches run on several plans and
    streams: reps searches of the sample on each of `streams` streams at
    once, whose launches overlap each other's ends; "ms" is then the time
    per search. With one stream a trial's time includes the launch's end,
    where the last lanes finish alone, and the ranking can differ (a
    batched pass on three streams ran 4% slower with the one-stream pick)."""
    rows = row_starts.numel() - 1
    n_chars = chars.numel()
    cap = max_kmers if max_kmers is not None else n_chars
    streams = max(1, streams)
    outs = [torch.empty(max(cap, 1), dtype=torch.int64, device=chars.device) for _ in range(streams)]
    out = outs[0]
    lanes = [torch.cuda.Stream() for _ in range(streams)] if streams > 1 else []
    total = int(torch.clamp(torch.diff(row_starts) - index.k + 1, min=0).sum())  # answers per search
    ref = None
    trials = []
    t_start = time.time()
    stream = torch.cuda.current_stream()
    k = index.k
    depths = [m for m in depths if m <= min(k, 17)]
    for layout, m in itertools.product(layouts, depths):
        if time.time() - t_start > budget_s:
            break
        if index.info["layout"] != layout:
            index.set_layout(layout)
        if index.info["presearch_depth"] != m:
            index.presearch(m)
        for seg, kern in itertools.product(segs, kernels):
            plans = [ReadsPlan(rows, n_chars, seg) for _ in range(streams)]
            flags = {**kernel_flags(kern), "n_as_a": n_as_a}
            plans[0].search(index, chars, row_starts, out, **flags)  # warm
            if streams == 1:
                times = []
                for _ in range(reps):
                    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                    e0.record(stream)
                    plans[0].search(index, chars, row_starts, out, **flags)
                    e1.record(stream)
                    e1.synchronize()
                    times.append(e0.elapsed_time(e1))
                ms = sorted(times)[len(times) // 2]
            else:
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                for st in lanes:
                    st.wait_stream(stream)
                for _ in range(reps):
                    for p, o, st in zip(plans, outs, lanes):
                        p.search(index, chars, row_starts, o, stream=st, **flags)
                for st in lanes:
                    stream.wait_stream(st)
                e1.record(stream)
                e1.synchronize()
                ms = e0.elapsed_time(e1) / (reps * streams)
            for o in outs if streams > 1 else [out]:
                if ref is None:
                    ref = o[:total].clone()
                elif not torch.equal(o[:total], ref):
                    raise RuntimeError(f"autotune: answers differ for layout={layout} m={m} seg={seg} kernel={kern}")
            trials.append({"layout": layout, "presearch_depth": m, "segment_kmers": seg, "kernel": kern, "ms": ms})
            if log:
                log(f"autotune layout=L{layout} m={m} seg={seg} {kern}: {ms:.3f} ms")
            for p in plans:
                p.close()
    best = min(trials, key=lambda t: t["ms"])
    if index.info["layout"] != best["layout"]:
        index.set_layout(best["layout"])
    if index.info["presearch_depth"] != best["presearch_depth"]:
        index.presearch(best["presearch_depth"])
    return {"best": best, "trials": trials}


KMER_KERNELS = ("lane", "cooperative")


def autotune_kmers(index: Index, kmers: torch.Tensor, *, layouts=(16, 32, 64), sample: int = 1 << 26,
                   reps: int = 3, log=None) -> dict:
    """The k-mer search (sbwt_search_kmers_flags) of the first `sample`
    top-aligned k-mers under every rank layout, one lane per k-mer or the
    cooperative kernel (layouts 32 and 64); checks that all give the same
    answers and leaves the index in the fastest layout. Which wins depends
    on the index: on a 3 Gbp uniform index two lanes per 32-byte chunk, on a
    pangenome whose rows fit the last-level cache one lane per 64-byte
    chunk. Returns {"best": {"layout", "kernel", "ms"}, "trials": [...]}."""
    q = kmers[: min(sample, kmers.numel())]
    out = torch.empty_like(q)
    stream = torch.cuda.current_stream()
    ref = None
    trials = []
    for layout in layouts:
        if index.info["layout"] != layout:
            index.set_layout(layout)
        for kern in (KMER_KERNELS if layout >= 32 else ("lane",)):
            coop = kern == "cooperative"
            index.search_kmers(q, out=out, cooperative=coop)  # warm
            times = []
            for _ in range(reps):
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(stream)
                index.search_kmers(q, out=out, cooperative=coop)
                e1.record(stream)
                e1.synchronize()
                times.append(e0.elapsed_time(e1))
            if ref is None:
                ref = out.clone()
            elif not torch.equal(out, ref):
                raise RuntimeError(f"autotune_kmers: answers differ for layout={layout} kernel={kern}")
            ms = sorted(times)[len(times) // 2]
            trials.append({"layout": layout, "kernel": kern, "ms": ms})
            if log:
                log(f"autotune_kmers layout=L{layout} {kern}: {ms:.3f} ms for {q.numel()} k-mers")
    best = min(trials, key=lambda t: t["ms"])
    if index.info["layout"] != best["layout"]:
        index.set_layout(best["layout"])
    return {"best": best, "trials": trials}
