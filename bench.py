#!/usr/bin/env python3
"""Benchmark: k-mer membership queries/s of the SBWT search on one MI355X.

Default workload (`--workload large`): two large indexes on one GPU, each
generated and built on the GPU from fixed seeds and queried with 1,000,000,000
independent 31-mers (50% sampled from the indexed sequences, 50% random)
through the reference's per-k-mer search semantics:
  * a 3 Gbp uniform-random genome (~3e9 distinct 31-mers, n ~ 3.0e9 columns);
  * 20 haplotypes of a 150 Mbp uniform base genome at 0.1% SNP and 0.01%
    indels of 1-10 bp (~0.25e9 columns).
Each index runs three paths side by side on the same resident queries:
plain, the learned prefilter (an MLP trained on the GPU plus a Bloom filter
over its false negatives, or what its calibration keeps of it) before the
masked search, and a Bloom filter over every indexed k-mer before the masked
search. The uniform index uses 32-byte rank chunks searched by two lanes per
k-mer (the cooperative kernel), the pangenome 64-byte chunks searched by one
lane per k-mer (the fastest of each index's layout sweep); the other layouts,
one lane per k-mer and two or four, are timed beside them. Every path's answers must
equal the plain ones bit for bit. The plain path is also timed with
presearch tables of depth 13 and 12 beside the configured 17. `value` is the
plain path on the 3 Gbp index. The 100 Mbp read batch (1M reads of 150 bp,
~120M consecutive 31-mers, `--workload reads`) rides in the same line.

Prints one JSON line. `roofline` prices the search kernel's algorithmic
bytes (128 B per rank line touched, counted by the kernel's counted variant
on the same queries) against the 8 TB/s HBM peak and a measured
random-line roof; `traffic` is FETCH_SIZE + WRITE_SIZE per launch from
rocprofv3 --pmc child runs, each scaled by what it reads for a 1 GiB copy,
with the L2's read requests by size (32, 64, 128 bytes) beside it as a
check that needs no scale; `cpu_baseline` times the CPU oracle (a
restatement of the reference's search: poppy rank, SB=1024) on a bounded
sample of the same k-mers and checks that the GPU answers there are
bit-identical.
"""
from __future__ import annotations

import argparse
import re
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "k-mer queries/s, 1 MI355X: plain vs filtered; % of HBM random-access roof"
HBM_PEAK_GBPS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--genome-bp", type=int, default=100_000_000)
    ap.add_argument("--reads", type=int, default=1_000_000)
    ap.add_argument("--read-len", type=int, default=150)
    ap.add_argument("--k", type=int, default=31)
    ap.add_argument("--presearch", type=int, default=None,
                    help="presearch depth m (default: 12, the read workloads' configured table; 17 for the "
                         "large and pangenome indexes, where every step is an HBM line and 4^17 x 8 B = 137 GB "
                         "fits beside the index)")
    ap.add_argument("--layout", type=int, default=16)
    ap.add_argument("--seg", type=int, default=256)
    ap.add_argument("--mode", choices=["default", "nocert", "persistent", "persistent-nocert", "independent",
                                       "persistent-independent", "lanes"],
                    default="default",
                    help="read search variant: streaming + certificates on the uniform-step kernel (default), "
                         "without certificates (per-lane kernel), the uniform-step kernel forced (with or "
                         "without certificates, or independent), every k-mer searched independently as the "
                         "reference does, or the per-lane kernel with certificates (lanes)")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-roof", action="store_true", help="skip the random-gather roof measurement")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 3 times (HBM counter calibration) and exit")
    ap.add_argument("--workload", choices=["large", "uniform", "pangenome", "reads", "batched", "prefilter", "text",
                                           "small"],
                    default="large",
                    help="large (default): the 3 Gbp uniform-random index and the 20-haplotype pangenome index, "
                         "1B independent 31-mers each, 50%% present, plain / learned-filtered / Bloom-filtered "
                         "side by side, plus the 100 Mbp read batch as an extra object; uniform or pangenome: "
                         "one of the two large indexes alone; reads: the 100 Mbp read search (1M x 150 bp); "
                         "batched: 10,000 uneven read sets with 5%% N, autotuned, in HIP-graph-captured "
                         "batches; prefilter: the learned negative prefilter (trained on the GPU) on a "
                         "90%%-negative read mix, filtered vs plain; text: query text to result text; small: the "
                         "1 Mbp / 1M-query case, GPU beside the CPU oracle on all cores and on one")
    ap.add_argument("--no-reads", action="store_true", help="large: leave out the 100 Mbp read batch")
    ap.add_argument("--plain-only", action="store_true",
                    help="large indexes: the plain path only (no prefilter training or filtered paths), e.g. for a "
                         "kernel trace whose search dispatches are all timed ones")
    ap.add_argument("--shallow-depths", default="13,12",
                    help="large indexes: presearch depths also timed on the plain path (comma list; empty: none)")
    ap.add_argument("--large-layout", choices=["default", "auto", "16", "32", "64"], default="default",
                    help="large / uniform / pangenome: rank layout of the index; default: 32-byte chunks with "
                         "the cooperative kernel for the uniform index, 64-byte chunks one lane per k-mer for the "
                         "pangenome (the fastest of the layout sweep on each); auto: what the k-mer autotuner "
                         "times fastest on 64M of the queries")
    ap.add_argument("--kmer-kernel", choices=["lane", "cooperative"], default="cooperative",
                    help="with a given --large-layout: one lane per k-mer, or a group of CB/16 lanes per k-mer "
                         "(layouts 32 and 64)")
    ap.add_argument("--layout-sweep", default="16,32,64",
                    help="large: also time these rank layouts, one lane and (32, 64) cooperative per k-mer")
    ap.add_argument("--negatives-like", choices=["uniform", "index"], default="index",
                    help="prefilter training negatives: uniform random k-mers, or random k-mers at the "
                         "index's base composition (each confirmed absent by the plain search)")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the two rocprofv3 --pmc child runs that measure HBM traffic per launch")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--image-child", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--sets", type=int, default=10_000)
    ap.add_argument("--batch-chars", type=int, default=256 << 20)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--streams", type=int, default=3,
                    help="batched: plans and streams the batches alternate over (graph pass: 123.7 ms with one, "
                         "112.7 with two, 105.9 with three, 106-108 with four to eight)")
    ap.add_argument("--run-items", action="store_true",
                    help="batched with --no-autotune: work items are the rows' N-free runs (else segments)")
    ap.add_argument("--train-positives", type=int, default=50_000_000)
    ap.add_argument("--train-negatives", type=int, default=50_000_000)
    ap.add_argument("--train-batch", type=int, default=65536)
    ap.add_argument("--train-epochs", type=int, default=1)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--threshold", type=float, default=0.0, help="prefilter logit threshold")
    ap.add_argument("--negative-fraction", type=float, default=0.9)
    ap.add_argument("--large-bp", type=int, default=3_000_000_000)
    ap.add_argument("--base-bp", type=int, default=150_000_000)
    ap.add_argument("--haplotypes", type=int, default=20)
    ap.add_argument("--queries", type=int, default=1_000_000_000)
    ap.add_argument("--checkpoint", action="store_true",
                    help="large: also time an index image save and load (to a temporary directory)")
    args = ap.parse_args()
    if args.presearch is None:
        args.presearch = 17 if args.workload in ("large", "uniform", "pangenome") else 12

    if args.calib_copy:
        return calib_copy()
    if args.image_child:
        return image_child(args.image_child)
    traffic = None
    if args.pmc_child:
        calib_copy()
        args.steps, args.warmup, args.no_cpu, args.no_roof = 1, 1, True, True
    elif (not args.no_traffic and args.workload in ("reads", "large", "uniform") and args.gpus == 1
          and int(os.environ.get("WORLD_SIZE", "1")) == 1):
        # before this process touches the GPU: the counter runs are children
        traffic = measure_traffic(args)
    args.traffic = traffic
    if args.workload != "reads":
        if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
            print("bench.py measures one GPU: the index is replicated per GPU and queries shard by read set "
                  "with no exchange, so a multi-GPU figure is not produced by this build.", file=sys.stderr)
            return 2
        return {"large": run_large, "uniform": run_uniform, "batched": run_batched, "prefilter": run_prefilter,
                "pangenome": run_pangenome, "text": run_text, "small": run_small}[args.workload](args)

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the index is replicated per GPU and queries shard by read set "
              "with no exchange, so a multi-GPU figure is not produced by this build.", file=sys.stderr)
        return 2

    print(json.dumps(run_reads(args)), flush=True)
    return 0


def run_reads(args) -> dict:
    """The 100 Mbp read batch (the body of the `reads` workload); returns its line."""
    import torch

    import sbwt_gpu_amd as sg
    from sbwt_gpu_amd import _capi, synth
    from sbwt_gpu_amd.api import _ptr, _stream

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    k = args.k

    # ---- index -----------------------------------------------------------
    t = time.time()
    g = synth.genome_codes(args.genome_bp, seed=3, gc=0.4, seg_frac=0.10, seg_len=(1_000, 50_000), seg_mut=0.01)
    log(f"genome {g.size} bp generated in {time.time() - t:.1f}s")
    gchars = torch.from_numpy(synth.to_ascii(g)).to(dev)
    gstarts = torch.tensor([0, g.size], dtype=torch.int64, device=dev)
    t = time.time()
    ix = sg.Index.build(gchars, k, sg.Config(presearch_depth=args.presearch, layout=args.layout), starts=gstarts)
    info = ix.info
    log(f"index built on GPU in {time.time() - t:.1f}s: n={info['n']} columns, "
        f"{info['device_bytes'] / 2**20:.0f} MiB, layout L{info['layout']}, m={info['presearch_depth']}")
    del gchars, gstarts

    # ---- reads -----------------------------------------------------------
    t = time.time()
    rchars_h, rstarts_h = synth.reads(g, args.reads, args.read_len, seed=4, sub=0.01)
    log(f"{args.reads} reads generated in {time.time() - t:.1f}s")
    t = time.time()
    rchars = torch.from_numpy(rchars_h).to(dev)
    rstarts = torch.from_numpy(rstarts_h).to(dev)
    torch.cuda.synchronize()
    h2d_s = time.time() - t
    rows = args.reads
    total = int(np.maximum(np.diff(rstarts_h) - k + 1, 0).sum())
    L = _capi.load()
    n_chars = rchars.numel()
    packed = torch.empty(L.sbwt_packed_words(n_chars), dtype=torch.int64, device=dev)
    nmask = torch.empty(L.sbwt_nmask_words(n_chars), dtype=torch.int64, device=dev)
    offsets = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    work = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    out = torch.empty(total, dtype=torch.int64, device=dev)
    bound = L.sbwt_reads_work_bound(rows, total, args.seg)
    items = torch.empty(bound, dtype=torch.int32, device=dev)
    flags = {"default": 0, "nocert": _capi.READS_NO_CERT, "persistent": _capi.READS_PERSISTENT,
             "persistent-nocert": _capi.READS_PERSISTENT | _capi.READS_NO_CERT,
             "independent": _capi.READS_INDEPENDENT,
             "persistent-independent": _capi.READS_PERSISTENT | _capi.READS_INDEPENDENT,
             "lanes": _capi.READS_SEGMENT_LANES}[args.mode]
    check = _capi.check
    s = _stream()
    stream = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2 * max(args.steps, 1))]

    def step(i=None, extra_flags=0):
        check(L.sbwt_pack(_ptr(rchars), n_chars, _ptr(packed), _ptr(nmask), s))
        check(L.sbwt_kmer_offsets(_ptr(rstarts), rows, k, _ptr(offsets), s))
        check(L.sbwt_reads_work(_ptr(offsets), rows, args.seg, _ptr(work), _ptr(items), s))
        if i is not None:
            ev[2 * i].record(stream)
        check(L.sbwt_search_reads(ix._h, _ptr(packed), _ptr(nmask), _ptr(rstarts), _ptr(offsets), _ptr(work),
                                  _ptr(items), rows, bound, args.seg, _ptr(out), flags | extra_flags, s))
        if i is not None:
            ev[2 * i + 1].record(stream)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log(f"warmup done ({args.warmup} steps), timing {args.steps} steps of {total} k-mers")
    t0 = time.perf_counter()
    for i in range(args.steps):
        step(i)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    value = total * args.steps / elapsed
    kern_ms = [ev[2 * i].elapsed_time(ev[2 * i + 1]) for i in range(args.steps)]
    kern_avg_s = float(np.mean(kern_ms)) / 1e3
    log(f"{value / 1e9:.3f} G k-mers/s end to end, search kernel avg {kern_avg_s * 1e3:.3f} ms "
        f"({total / kern_avg_s / 1e9:.3f} G k-mers/s)")
    res_gpu = out.cpu().numpy().view(np.uint64)

    # ---- algorithmic bytes: counted variant on the same batch (untimed) ----
    ix.counters(reset=True)
    step(None, _capi.COUNT)
    torch.cuda.synchronize()
    ctr = ix.counters(reset=True)
    same = np.array_equal(out.cpu().numpy().view(np.uint64), res_gpu)
    lines = int(ctr["lines"])
    alg_bytes = lines * 128
    achieved = alg_bytes / kern_avg_s / 1e9
    found = int((res_gpu != np.uint64(0xFFFFFFFFFFFFFFFF)).sum())
    log(f"counted run: {lines} lines ({lines / total:.2f}/k-mer), {ctr['steps']} steps, "
        f"{ctr['full_searches']} full searches, found {found}/{total}; counted == timed: {same}")

    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": traffic_bytes(args.traffic),
                "traffic_detail": args.traffic,
                "kernel": "k_reads_uniform" if args.mode in ("default", "persistent", "persistent-nocert", "persistent-independent") else "k_search_reads", "kernel_avg_ms": round(kern_avg_s * 1e3, 4),
                "lines_per_launch": lines, "bytes_per_kmer": round(alg_bytes / total, 2),
                "steps_per_launch": int(ctr["steps"]),
                "floor_bytes_per_kmer": round(128 * int(ctr["steps"]) / total, 2),
                "index_bytes": info["device_bytes"]}
    fetch = ((args.traffic or {}).get("counters") or {}).get("FETCH_SIZE")
    if fetch:
        roofline["hbm_fetch_bytes_per_kmer"] = round(fetch["bytes"] / total, 2)
    if (args.traffic or {}).get("tcc"):
        roofline["tcc_hit_rate"] = args.traffic["tcc"]["hit_rate"]
    rq = (args.traffic or {}).get("read_requests") or {}
    if "read_bytes" in rq:
        roofline["read_bytes_per_kmer_from_request_sizes"] = round(rq["read_bytes"] / total, 2)
    if not args.no_roof:
        roof = gather_roof()
        roofline["measured_random_roof"] = roof
        roofline["frac_of_measured_roof"] = round(achieved / roof["GBps_128B_lines"], 4)
        # the same gather over a buffer the size of the index (rows, chunks and
        # table), which the last-level cache partly holds, as the search's does
        ws = gather_roof(nbytes=int(info["device_bytes"]))
        roofline["working_set_random_roof"] = ws
        roofline["frac_of_working_set_roof"] = round(achieved / ws["GBps_128B_lines"], 4)

    # ---- consecutive batches over several plans and streams -----------------
    # Step i on plan i % S and stream i % S (each plan counts its own work
    # items): one batch's tail overlaps the next batch's search. Each stream
    # writes its own answers, checked against the single-stream ones.
    pipelined = None
    S = max(1, args.streams)
    if S > 1 and args.mode == "default":
        plans = [sg.ReadsPlan(rows, n_chars, args.seg) for _ in range(S)]
        lanes = [torch.cuda.Stream() for _ in range(S)]
        outs = [torch.empty_like(out) for _ in range(S)]

        def batches(n):
            for st in lanes:
                st.wait_stream(stream)
            for i in range(n):
                plans[i % S].search(ix, rchars, rstarts, outs[i % S], stream=lanes[i % S])
            for st in lanes:
                stream.wait_stream(st)

        batches(max(args.warmup, S))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        batches(args.steps)
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / args.steps
        same_p = all(np.array_equal(o.cpu().numpy().view(np.uint64), res_gpu) for o in outs)
        pipelined = {"streams": S, "kmers_per_s": round(total / el, 1), "ms_per_step": round(el * 1e3, 4),
                     "answers_equal_single_stream": same_p,
                     "lines_per_s_GBps": round(lines * 128 / el / 1e9, 1)}
        log(f"{S} plans and streams: {total / el / 1e9:.3f} G k-mers/s ({el * 1e3:.3f} ms per batch); "
            f"answers equal: {same_p}")
        for p_ in plans:
            p_.close()
        del outs

    # ---- CPU baseline + parity on a sample -----------------------------------
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ix, rchars_h, rstarts_h, res_gpu, k, args)

    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "k-mer queries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(elapsed / args.steps * 1e3, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded genome + reads generated in-run)",
        "config": {"workload": f"plain GPU read search: {args.genome_bp / 1e6:g} Mbp synthetic genome (GC 40%, "
                               f"10% segmental copies at 1%), k={k}, presearch {args.presearch}; "
                               f"{rows} reads x {args.read_len} bp at 1% substitutions",
                   "kmers_per_step": total, "index_columns": info["n"], "layout": f"L{info['layout']}",
                   "mode": args.mode, "segment_kmers": args.seg,
                   "search_kernel_kmers_per_s": round(total / kern_avg_s, 1),
                   "consecutive_batches_on_streams": pipelined,
                   "found_fraction": round(found / total, 4),
                   "h2d_reads_GBps": round(rchars_h.nbytes / h2d_s / 1e9, 2)},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    return line



def _time_path(run, args):
    """W untimed runs, then K timed runs between synchronisations; run(ev)
    records ev[0] (start), ev[1] (filter done) and ev[2] (search done) on
    the current stream. Returns (seconds per step, filter ms, search ms)."""
    import torch

    for _ in range(args.warmup):
        run(None)
    torch.cuda.synchronize()
    nsteps = max(args.steps, 1)
    evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(nsteps)]
    t0 = time.perf_counter()
    for i in range(nsteps):
        run(evs[i])
    torch.cuda.synchronize()
    el = (time.perf_counter() - t0) / nsteps
    filt = float(np.mean([e[0].elapsed_time(e[1]) for e in evs]))
    search = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
    return el, filt, search


def large_index(kind: str, args) -> dict:
    """One large index on one GPU, generated and built on the GPU from fixed
    seeds, queried with `queries` independent 31-mers, 50/50 present and
    absent, through the reference's per-k-mer search semantics (presearch
    gather + up to k - m narrowing steps, include/gpu/search_kernel.hpp:61-104):
      uniform:   a 3 Gbp uniform-random genome (~3e9 distinct 31-mers); even
                 queries at uniform genome positions, odd ones from a second
                 random sequence;
      pangenome: `haplotypes` copies of a 150 Mbp uniform base genome, each
                 with 0.1% SNPs and 0.01% indels of 1-10 bp; even queries at
                 uniform positions of uniformly chosen haplotypes, odd ones
                 uniform random k-mers.
    Three paths run side by side on the same resident queries: plain; the
    learned prefilter (MLP trained on the GPU + Bloom filter over its false
    negatives) before the masked search; and a Bloom filter over every
    indexed k-mer before the masked search. Every path's answers must equal
    the plain ones bit for bit, and every sampled (even) query must be found.
    Beside the configured depth the plain path is also timed with presearch
    tables of depth 13 and 12 (the reference's clamp is 5..13,
    include/parameters.hpp:255-259)."""
    import torch

    import sbwt_gpu_amd as sg
    from sbwt_gpu_amd import _capi, synth
    from sbwt_gpu_amd.api import _ptr, _stream
    from sbwt_gpu_amd.learned_filter import random_kmers, train_prefilter

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    k, Q = args.k, args.queries
    L = _capi.load()
    check = _capi.check
    s = _stream()
    gen = torch.Generator(device=dev)
    t = time.time()
    if kind == "uniform":
        G, R = args.large_bp, 64_000_000  # R: the absent-query source
        gen.manual_seed(8)
        chars = torch.randint(0, 4, (G + R,), dtype=torch.uint8, device=dev, generator=gen)
        # 2-bit code -> ASCII A, C, G, T
        chars.mul_(2).add_(65)
        chars.add_((chars >= 69).to(torch.uint8) * 2)
        chars.add_((chars == 73).to(torch.uint8) * 11)
        starts = torch.tensor([0, G], dtype=torch.int64, device=dev)
        desc = f"{G / 1e9:g} Gbp uniform-random genome"
    else:
        gen.manual_seed(31)
        base = torch.randint(0, 4, (args.base_bp,), dtype=torch.uint8, device=dev, generator=gen)
        chars, starts = synth.gpu_haplotypes(base, args.haplotypes, seed=32)
        del base
        G, R = chars.numel(), 0
        desc = (f"pangenome: {args.haplotypes} haplotypes of a {args.base_bp / 1e6:g} Mbp uniform base genome "
                f"(0.1% SNP, 0.01% indel 1-10 bp), {G / 1e9:.3f} Gbp")
    torch.cuda.synchronize()
    gen_s = time.time() - t
    log(f"{kind}: {G + R} bases generated on the GPU in {gen_s:.1f}s")
    # The queries come first, before the index build's allocations: in a
    # process that started right after another large one, a run of a few
    # thousand entries of a query array written after the build was seen
    # holding foreign data (the index itself was bit-identical to a good
    # run's). Two independent extractions of the sampled queries must agree;
    # a disagreement regenerates them.
    half = (Q + 1) // 2
    if kind == "uniform":
        qpos = torch.empty(Q, dtype=torch.int64, device=dev)
        qpos[0::2] = torch.randint(0, G - k + 1, (half,), device=dev, generator=gen)
        qpos[1::2] = G + torch.randint(0, R - k + 1, (Q // 2,), device=dev, generator=gen)
        spos = qpos[0::2].clone()  # the sampled queries' positions
    else:
        lens = torch.diff(starts)
        hap = torch.randint(0, lens.numel(), (half,), device=dev, generator=gen)
        off = (torch.rand(half, device=dev, generator=gen, dtype=torch.float64) * (lens - k + 1)[hap]).to(torch.int64)
        spos = starts[:-1][hap] + off
        del hap, off
        qpos = None
        rand_half = random_kmers(Q // 2, k, gen)
    integrity = {"query_regenerations": 0}
    for attempt in range(3):
        packed = torch.empty(L.sbwt_packed_words(G + R), dtype=torch.int64, device=dev)
        check(L.sbwt_pack(_ptr(chars), G + R, _ptr(packed), None, s))
        kmers = torch.empty(Q, dtype=torch.int64, device=dev)
        if qpos is not None:
            check(L.sbwt_extract_kmers(_ptr(packed), _ptr(qpos), Q, _ptr(kmers), s))
        else:
            kpos = torch.empty(half, dtype=torch.int64, device=dev)
            check(L.sbwt_extract_kmers(_ptr(packed), _ptr(spos), half, _ptr(kpos), s))
            kmers[0::2] = kpos
            kmers[1::2] = rand_half
            del kpos
        # the sampled queries once more, from a second packing of the characters
        packed2 = torch.empty(L.sbwt_packed_words(G + R), dtype=torch.int64, device=dev)
        check(L.sbwt_pack(_ptr(chars), G + R, _ptr(packed2), None, s))
        again = torch.empty_like(spos)
        check(L.sbwt_extract_kmers(_ptr(packed2), _ptr(spos), spos.numel(), _ptr(again), s))
        torch.cuda.synchronize()
        same = bool(torch.equal(again, kmers[0::2]))
        del packed, packed2, again
        if same:
            break
        integrity["query_regenerations"] += 1
        log(f"{kind}: the two extractions of the sampled queries differ; generating them again")
        del kmers
    integrity["queries_equal_second_extraction"] = same
    integrity["query_digest_before_search"] = int(kmers.sum())
    del qpos
    if kind != "uniform":
        del rand_half
    out = torch.empty_like(kmers)
    maybe = torch.empty(Q, dtype=torch.uint8, device=dev)
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    t = time.time()
    # the default per index: the fastest of the layout sweep over several runs
    # (the others stay within 2-8%; the autotuner's pick varies among those
    # within 2% from run to run)
    default = {"uniform": (32, True), "pangenome": (64, False)}[kind]
    build_layout = {"default": default[0], "auto": 32}.get(args.large_layout) or int(args.large_layout)
    in_sum0 = _byte_sum(chars[:G])
    ix = sg.Index.build(chars[:G], k, sg.Config(presearch_depth=args.presearch, layout=build_layout),
                        starts=starts)
    # integrity: the build leaves its input and the queries alone, and the
    # searches leave the index alone
    integrity["input_unchanged_by_build"] = _byte_sum(chars[:G]) == in_sum0
    integrity["queries_unchanged_by_build"] = int(kmers.sum()) == integrity["query_digest_before_search"]
    integrity["index_digest_after_build"] = ix.digest()
    info = ix.info
    build_s = time.time() - t
    log(f"{kind}: index built on GPU in {build_s:.1f}s: n={info['n']} columns, "
        f"{info['device_bytes'] / 2**30:.2f} GiB, layout L{info['layout']}, m={info['presearch_depth']}")
    del starts
    log(f"{kind}: {Q} queries ready")
    tune = None
    if args.large_layout == "auto":
        from sbwt_gpu_amd.autotune import autotune_kmers
        tune = autotune_kmers(ix, kmers, log=log)
        layout, coop = tune["best"]["layout"], tune["best"]["kernel"] == "cooperative"
    elif args.large_layout == "default":
        layout, coop = default
    else:
        layout, coop = build_layout, args.kmer_kernel == "cooperative" and build_layout >= 32
    kflags = _capi.SEARCH_COOPERATIVE if coop else 0
    kname = "k_search_kmers_coop" if coop else "k_search_kmers"
    info = ix.info
    log(f"{kind}: kmer kernel: {kname}<{layout},")  # (read by the counter runs' parent)
    stream = torch.cuda.current_stream()

    def plain(ev):
        if ev:
            ev[0].record(stream)
            ev[1].record(stream)
        check(L.sbwt_search_kmers_flags(ix._h, _ptr(kmers), Q, _ptr(out), kflags, s))
        if ev:
            ev[2].record(stream)

    el, _, kern_ms = _time_path(plain, args)
    if args.pmc_child:  # the counter runs need the timed launches only
        return {}
    ref = out.clone()
    missing = int((ref[0::2] == -1).sum().item())
    found = int((ref != -1).sum().item())
    log(f"{kind} plain: {Q / el / 1e9:.3f} G k-mers/s ({el * 1e3:.2f} ms/step, kernel {kern_ms:.2f} ms), "
        f"found {found}/{Q}, sampled k-mers missing: {missing}")
    integrity["index_unchanged_by_queries"] = ix.digest() == integrity["index_digest_after_build"]
    integrity["queries_unchanged_by_search"] = int(kmers.sum()) == integrity["query_digest_before_search"]
    if missing:
        integrity["missing"] = _recheck_missing(ix, chars, spos, kmers, ref, k, kflags, args.presearch, s)
    log(f"{kind} integrity: {integrity}")
    del chars, spos
    # algorithmic bytes: the counted variant on the same queries (untimed)
    ix.counters(reset=True)
    check(L.sbwt_search_kmers_flags(ix._h, _ptr(kmers), Q, _ptr(out), kflags | _capi.COUNT, s))
    torch.cuda.synchronize()
    ctr = ix.counters(reset=True)
    counted_same = bool(torch.equal(out, ref))
    lines = int(ctr["lines"])
    kern_s = kern_ms / 1e3
    achieved = lines * 128 / kern_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None, "kernel": kname,
                "kernel_avg_ms": round(kern_ms, 3), "lines_per_launch": lines,
                "bytes_per_kmer": round(lines * 128 / Q, 2), "steps_per_launch": int(ctr["steps"]),
                "floor_bytes_per_kmer": round(128 * int(ctr["steps"]) / Q, 2),
                "index_bytes": info["device_bytes"], "counted_equals_timed": counted_same}
    res = {"workload": desc, "bases": G, "index_columns": info["n"], "index_build_s": round(build_s, 1),
           "generate_s": round(gen_s, 1), "layout": f"L{info['layout']}", "kmer_kernel": kname,
           "presearch": info["presearch_depth"],
           "kmer_autotune": tune,
           "queries": Q, "found_fraction": round(found / Q, 4), "sampled_kmers_all_found": missing == 0,
           "integrity": integrity, "roofline": roofline}

    paths = {"plain": {"kmers_per_s": round(Q / el, 1), "ms_per_step": round(el * 1e3, 3),
                       "filter_kernel_ms": 0.0, "search_kernel_ms": round(kern_ms, 3), "maybe_fraction": None,
                       "answers_equal_plain": True}}
    res["paths"] = paths
    if not args.plain_only:
        _filtered_paths(kind, ix, kmers, out, ref, maybe, paths, res, args, kflags)
    torch.cuda.empty_cache()

    # ---- the CPU baseline on a sample, with GPU parity ----------------------
    if not args.no_cpu:
        res_h = ref[: 50_000_000].cpu().numpy().view(np.uint64)
        res["cpu_baseline"] = cpu_baseline_kmers(ix, kmers, res_h, k, args)

    # ---- shallower presearch tables (untimed rebuild, then timed steps) ------
    depths = {}
    for m in [int(x) for x in args.shallow_depths.split(",") if x]:
        if m >= args.presearch:
            continue
        t = time.time()
        ix.presearch(m)
        torch.cuda.synchronize()
        build_m = time.time() - t
        sub = argparse.Namespace(**{**vars(args), "steps": max(3, min(args.steps, 5)), "warmup": 1})
        e, _, km = _time_path(plain, sub)
        same = bool(torch.equal(out, ref))
        ix.counters(reset=True)
        check(L.sbwt_search_kmers_flags(ix._h, _ptr(kmers), Q, _ptr(out), kflags | _capi.COUNT, s))
        torch.cuda.synchronize()
        c = ix.counters(reset=True)
        depths[str(m)] = {"kmers_per_s": round(Q / e, 1), "kernel_ms": round(km, 3),
                          "lines_per_kmer": round(int(c["lines"]) / Q, 3),
                          "achieved_GBps": round(int(c["lines"]) * 128 / (km / 1e3) / 1e9, 1),
                          "table_bytes": 8 << (2 * m), "table_build_s": round(build_m, 2),
                          "answers_equal_plain": same}
        log(f"{kind} m={m}: {Q / e / 1e9:.3f} G k-mers/s, kernel {km:.2f} ms, "
            f"{int(c['lines']) / Q:.2f} lines/k-mer; answers equal: {same}")
    res["shallower_presearch"] = depths
    if depths:
        ix.presearch(args.presearch)
    # ---- other rank layouts, one lane or a group of lanes per k-mer ----------
    layouts = {}
    sweep = [int(x) for x in args.layout_sweep.split(",") if x]
    for lay in sweep:
        t = time.time()
        ix.set_layout(lay)
        torch.cuda.synchronize()
        relayout_s = time.time() - t
        for coop_ in ([False, True] if lay >= 32 else [False]):
            if lay == layout and coop_ == coop:
                continue  # the configured path, timed above
            flags = _capi.SEARCH_COOPERATIVE if coop_ else 0

            def run(ev, flags=flags):
                if ev:
                    ev[0].record(stream)
                    ev[1].record(stream)
                check(L.sbwt_search_kmers_flags(ix._h, _ptr(kmers), Q, _ptr(out), flags, s))
                if ev:
                    ev[2].record(stream)

            sub = argparse.Namespace(**{**vars(args), "steps": max(3, min(args.steps, 5)), "warmup": 1})
            e, _, km = _time_path(run, sub)
            same = bool(torch.equal(out, ref))
            ix.counters(reset=True)
            check(L.sbwt_search_kmers_flags(ix._h, _ptr(kmers), Q, _ptr(out), flags | _capi.COUNT, s))
            torch.cuda.synchronize()
            c = ix.counters(reset=True)
            name = f"L{lay}" + ("-cooperative" if coop_ else "")
            layouts[name] = {"kmers_per_s": round(Q / e, 1), "kernel_ms": round(km, 3),
                             "lines_per_kmer": round(int(c["lines"]) / Q, 3),
                             "achieved_GBps": round(int(c["lines"]) * 128 / (km / 1e3) / 1e9, 1),
                             "index_bytes": ix.info["device_bytes"], "relayout_s": round(relayout_s, 2),
                             "answers_equal_plain": same}
            log(f"{kind} {name}: {Q / e / 1e9:.3f} G k-mers/s, kernel {km:.2f} ms, "
                f"{int(c['lines']) / Q:.2f} lines/k-mer; answers equal: {same}")
    if sweep:
        ix.set_layout(layout)
        res["layouts"] = layouts
    if args.checkpoint and kind == "uniform":
        ix.presearch(args.presearch)
        res["checkpoint"] = image_round_trip(ix, kmers, out, ref.cpu().numpy().view(np.uint64))  # closes ix
    ix.close()
    del kmers, out, ref, maybe
    torch.cuda.empty_cache()
    return res


def _byte_sum(t) -> int:
    """Sum of the bytes of a device uint8 tensor (in pieces: no large temporaries)."""
    import torch

    step = 1 << 28
    return sum(int(t[i:i + step].sum(dtype=torch.int64)) for i in range(0, t.numel(), step))


def _recheck_missing(ix, chars, spos, kmers, ref, k, kflags, m, s, limit=4096):
    """Where sampled queries were answered absent: are the queries the input's
    k-mers (recomputed from the characters, not from the packed string), and
    are they found by a second search, after the presearch table is rebuilt,
    and after the rank chunks are rebuilt from the rows? Localises a wrong
    answer to the queries, the table, the chunks or the rows."""
    import torch

    from sbwt_gpu_amd import _capi
    from sbwt_gpu_amd.api import _ptr

    L, check = _capi.load(), _capi.check
    miss = torch.nonzero(ref[0::2] == -1).flatten()
    n_all = int(miss.numel())
    miss = miss[:limit]
    pos = spos[miss]
    # the k-mers from the characters: codes of pos .. pos+k-1, top-aligned
    win = chars[pos[:, None] + torch.arange(k, device=pos.device)[None, :]].to(torch.int64)
    code = (win == 67) * 1 + (win == 71) * 2 + (win == 84) * 3
    shifts = 62 - 2 * torch.arange(k, device=pos.device)
    again = (code << shifts[None, :]).sum(1)
    mask = -(1 << (64 - 2 * k)) if k < 32 else -1
    q = kmers[0::2][miss]
    bad = q[(q & mask) != again]
    res = {"count": n_all, "checked": int(miss.numel()),
           "queries_equal_input": int(((q & mask) == again).sum()),
           "first_positions": [int(x) for x in pos[:8].tolist()],
           "first_sampled_indices": [int(x) for x in miss[:16].tolist()],
           "first_wrong_values": [f"{int(x) & 0xFFFFFFFFFFFFFFFF:016x}" for x in bad[:8].tolist()],
           "wrong_zero": int((bad == 0).sum()), "wrong_all_ones": int((bad == -1).sum()),
           "wrong_below_2_32": int(((bad >= 0) & (bad < (1 << 32))).sum())}

    def search_again():
        out = torch.empty_like(q)
        check(L.sbwt_search_kmers_flags(ix._h, _ptr(q), q.numel(), _ptr(out), kflags, s))
        torch.cuda.synchronize()
        return int((out != -1).sum())

    res["found_on_second_search"] = search_again()
    before = ix.digest()
    ix.presearch(m)
    after = ix.digest()
    res["table_equal_after_rebuild"] = after["table"] == before["table"]
    res["found_after_table_rebuild"] = search_again()
    ix.set_layout(ix.info["layout"])
    res["chunks_equal_after_rebuild"] = ix.digest()["chunks"] == before["chunks"]
    res["found_after_chunk_rebuild"] = search_again()
    return res


def _filtered_paths(kind, ix, kmers, out, ref, maybe, paths, res, args, kflags=0):
    """The learned prefilter and the Bloom filter over every indexed k-mer,
    each before the masked search, timed on the same queries as the plain
    path; their answers are compared with the plain ones (ref) on the device."""
    import torch

    import sbwt_gpu_amd as sg
    from sbwt_gpu_amd import _capi
    from sbwt_gpu_amd.api import _ptr, _stream
    from sbwt_gpu_amd.learned_filter import train_prefilter

    L, check, s = _capi.load(), _capi.check, _stream()
    k, Q = args.k, kmers.numel()
    stream = torch.cuda.current_stream()
    t = time.time()
    pf = train_prefilter(ix, positives=args.train_positives, negatives=args.train_negatives,
                         epochs=args.train_epochs, batch=args.train_batch, lr=args.lr, seed=7,
                         threshold=args.threshold, fpr=0.01, negatives_like=args.negatives_like,
                         calibrate_for=0.5, log=log)
    train_s = time.time() - t
    log(f"{kind}: prefilter ready in {train_s:.1f}s: {pf.stats}")
    t = time.time()
    allpos = ix.enumerate_kmers()
    bloom_all = sg.Bloom(k, allpos.numel(), 0.01)
    bloom_all.insert(allpos)
    del allpos
    torch.cuda.empty_cache()
    bloom_s = time.time() - t

    def filtered(kind_):
        def run(ev):
            if ev:
                ev[0].record(stream)
            if kind_ == "learned":
                pf.maybe(kmers, out=maybe)
            else:
                check(L.sbwt_bloom_query(bloom_all._h, _ptr(kmers), Q, _ptr(maybe), s))
            if ev:
                ev[1].record(stream)
            check(L.sbwt_search_kmers_masked(ix._h, _ptr(kmers), Q, _ptr(maybe), _ptr(out), kflags, s))
            if ev:
                ev[2].record(stream)
        return run

    for name, run in (("filtered", filtered("learned")), ("bloom_only", filtered("bloom"))):
        e, fm, sm = _time_path(run, args)
        same = bool(torch.equal(out, ref))
        mf = round(float(maybe.float().mean().item()), 4)
        paths[name] = {"kmers_per_s": round(Q / e, 1), "ms_per_step": round(e * 1e3, 3),
                       "filter_kernel_ms": round(fm, 3), "search_kernel_ms": round(sm, 3), "maybe_fraction": mf,
                       "answers_equal_plain": same}
        log(f"{kind} {name}: {Q / e / 1e9:.3f} G k-mers/s ({e * 1e3:.2f} ms/step; filter {fm:.2f} ms, "
            f"search {sm:.2f} ms, maybe {mf}); answers equal plain: {same}")
    res["filtered_vs_plain"] = round(paths["filtered"]["kmers_per_s"] / paths["plain"]["kmers_per_s"], 3)
    res["bloom_only_vs_plain"] = round(paths["bloom_only"]["kmers_per_s"] / paths["plain"]["kmers_per_s"], 3)
    res["prefilter"] = {"train": pf.stats, "train_total_s": round(train_s, 1),
                        "learned_bloom_bytes": pf.bloom.bytes if pf.bloom else 0,
                        "model_bytes": int(L.sbwt_mlp_param_count(pf.net._h)) * 2 if pf.mlp is not None else 0,
                        "bloom_only_bytes": bloom_all.bytes, "bloom_only_build_s": round(bloom_s, 1)}
    del pf, bloom_all
    torch.cuda.empty_cache()


def small_case(args) -> dict:
    """The CPU-runnable case: a seeded uniform 1,000,000 bp genome, k = 31,
    presearch depth 8, 1,000,000 query 31-mers (half at uniform genome
    positions, half uniform random, shuffled), searched independently with
    the reference's semantics; the CPU oracle (poppy SB = 1024, the same
    depth) runs all of them on every host core and on one, and its answers
    are compared with the GPU's."""
    import torch

    import sbwt_gpu_amd as sg
    from oracle import oracle as O
    from sbwt_gpu_amd import _capi, synth
    from sbwt_gpu_amd.api import _ptr, _stream

    k, m, Q = 31, 8, 1_000_000
    g = synth.genome_codes(1_000_000, seed=101, gc=0.5)
    text = synth.to_ascii(g)
    ix = sg.Index.build([text.tobytes()], k, sg.Config(presearch_depth=m))
    rng = np.random.default_rng(102)
    pos = rng.integers(0, g.size - k + 1, Q // 2)
    present = np.zeros(pos.size, np.uint64)
    for t in range(k):
        present |= g[pos + t].astype(np.uint64) << np.uint64(62 - 2 * t)
    q = np.concatenate([present, rng.integers(0, 2**62, Q - Q // 2, dtype=np.uint64) << np.uint64(2)])
    rng.shuffle(q)
    kmers = torch.from_numpy(q.view(np.int64)).cuda()
    out = torch.empty_like(kmers)
    L, s = _capi.load(), _stream()
    stream = torch.cuda.current_stream()

    def run(ev):
        if ev:
            ev[0].record(stream)
            ev[1].record(stream)
        _capi.check(L.sbwt_search_kmers(ix._h, _ptr(kmers), Q, _ptr(out), s))
        if ev:
            ev[2].record(stream)

    sub = argparse.Namespace(steps=max(args.steps, 20), warmup=max(args.warmup, 3))
    el, _, kern_ms = _time_path(run, sub)
    got = out.cpu().numpy().view(np.uint64)
    rows = ix.rows()
    oix = O.Index(rows[:4], ix.n, sb=1024)
    oix.presearch(m)
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    oix.search(q[:10_000], k)  # warm
    t = time.perf_counter()
    want = oix.search(q, k)
    cpu_all = Q / (time.perf_counter() - t)
    t = time.perf_counter()
    oix.search(q, k, threads=1)
    cpu_one = Q / (time.perf_counter() - t)
    parity = bool(np.array_equal(got, want))
    ix.close()
    log(f"1 Mbp case: GPU {Q / el / 1e9:.3f} G k-mers/s (kernel {kern_ms:.3f} ms), CPU oracle "
        f"{cpu_all / 1e6:.2f} M/s on {threads} threads, {cpu_one / 1e6:.2f} M/s on 1; parity: {parity}")
    return {"workload": "1 Mbp uniform genome, k=31, presearch 8, 1M query 31-mers (50% sampled, 50% random), "
                        "independent searches",
            "gpu_kmers_per_s": round(Q / el, 1), "gpu_ms_per_step": round(el * 1e3, 4),
            "gpu_kernel_ms": round(kern_ms, 4),
            "cpu_all_cores_kmers_per_s": round(cpu_all, 1), "cpu_cores": threads,
            "cpu_one_core_kmers_per_s": round(cpu_one, 1), "cpu_kind": "port", "cpu_model": _cpu_model(),
            "found_fraction": round(float((got != np.uint64(0xFFFFFFFFFFFFFFFF)).mean()), 4),
            "gpu_equals_oracle": parity}


def run_small(args) -> int:
    res = small_case(args)
    line = {"metric": METRIC, "value": res["gpu_kmers_per_s"], "unit": "k-mer queries/s", "n_gpus": 1,
            "steps": max(args.steps, 20), "warmup": max(args.warmup, 3), "ms_per_step": res["gpu_ms_per_step"],
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic (seeded)", "config": res, "roofline": None,
            "cpu_baseline": {"value": res["cpu_all_cores_kmers_per_s"], "unit": "k-mer queries/s",
                             "cores": res["cpu_cores"], "kind": "port", "sample": "all 1M queries",
                             "single_thread_value": res["cpu_one_core_kmers_per_s"],
                             "gpu_parity_on_sample": res["gpu_equals_oracle"]}}
    print(json.dumps(line), flush=True)
    return 0


def _traffic_into(roofline: dict, traffic: dict | None, queries: int) -> None:
    """Adds the rocprofv3 counter runs (which search queries_per_launch
    k-mers) to a large-index roofline, scaled to its launch."""
    roofline["traffic_detail"] = traffic
    t = traffic_bytes(traffic)
    if t is not None:
        roofline["traffic"] = round(t * queries / traffic["queries_per_launch"])
    fetch = ((traffic or {}).get("counters") or {}).get("FETCH_SIZE")
    if fetch:
        roofline["hbm_fetch_bytes_per_kmer"] = round(fetch["bytes"] / traffic["queries_per_launch"], 2)
    if (traffic or {}).get("tcc"):
        roofline["tcc_hit_rate"] = traffic["tcc"]["hit_rate"]
    rq = (traffic or {}).get("read_requests") or {}
    if "read_bytes" in rq:
        roofline["read_bytes_per_kmer_from_request_sizes"] = round(rq["read_bytes"] / traffic["queries_per_launch"], 2)


def _large_line(res: dict, args) -> dict:
    p = res.get("paths", {}).get("plain") or {}
    return {"metric": METRIC, "value": p.get("kmers_per_s"), "unit": "k-mer queries/s", "n_gpus": 1,
            "steps": args.steps, "warmup": args.warmup, "ms_per_step": p.get("ms_per_step"),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
            "data": "synthetic (seeded, generated on the GPU)"}


def run_uniform(args) -> int:
    res = large_index("uniform", args)
    if args.pmc_child:
        return 0
    _traffic_into(res["roofline"], args.traffic, args.queries)
    if not args.no_roof:
        roof = gather_roof()
        res["roofline"]["measured_random_roof"] = roof
        res["roofline"]["frac_of_measured_roof"] = round(res["roofline"]["achieved"] / roof["GBps_128B_lines"], 4)
    line = _large_line(res, args)
    line["config"] = {k: v for k, v in res.items() if k not in ("roofline", "cpu_baseline")}
    line["roofline"], line["cpu_baseline"] = res["roofline"], res.get("cpu_baseline")
    print(json.dumps(line), flush=True)
    return 0


def run_pangenome(args) -> int:
    res = large_index("pangenome", args)
    if args.pmc_child:
        return 0
    line = _large_line(res, args)
    line["config"] = {k: v for k, v in res.items() if k not in ("roofline", "cpu_baseline")}
    line["roofline"], line["cpu_baseline"] = res["roofline"], res.get("cpu_baseline")
    print(json.dumps(line), flush=True)
    return 0


def run_large(args) -> int:
    """The large-index workload on one GPU: the 3 Gbp uniform-random index
    and the 20-haplotype pangenome index, 1B queries each (50/50), plain and
    filtered side by side (large_index). `value` is the plain path on the
    3 Gbp index; the pangenome and the 100 Mbp read batch (run_reads) ride in
    the same line as their own objects."""
    import torch

    uni = large_index("uniform", args)
    _traffic_into(uni["roofline"], args.traffic, args.queries)
    if not args.no_roof:
        roof = gather_roof()
        uni["roofline"]["measured_random_roof"] = roof
        uni["roofline"]["frac_of_measured_roof"] = round(uni["roofline"]["achieved"] / roof["GBps_128B_lines"], 4)
    torch.cuda.empty_cache()
    pan = large_index("pangenome", args)
    torch.cuda.empty_cache()
    reads = None
    if not args.no_reads:
        rargs = argparse.Namespace(**{**vars(args), "presearch": 12, "traffic": None, "no_roof": True})
        reads = run_reads(rargs)
    line = _large_line(uni, args)
    line["config"] = {
        "workload": "large index, one GPU: a 3 Gbp uniform-random genome index and a 20-haplotype pangenome "
                    f"index, {args.queries} independent 31-mers each, 50% present; plain, learned-filtered and "
                    "Bloom-filtered searches side by side; value = plain on the 3 Gbp index",
        "uniform_3Gbp": {k: v for k, v in uni.items() if k not in ("roofline", "cpu_baseline")},
        "pangenome": {k: v for k, v in pan.items() if k != "cpu_baseline"},
        "filtered_vs_plain": uni.get("filtered_vs_plain"),
    }
    line["roofline"], line["cpu_baseline"] = uni["roofline"], uni.get("cpu_baseline")
    line["pangenome_cpu_baseline"] = pan.get("cpu_baseline")
    line["reads_100Mbp_1Mx150bp"] = reads
    line["small_1Mbp_1M_queries"] = None if args.no_cpu else small_case(args)
    print(json.dumps(line), flush=True)
    return 0


def image_round_trip(ix, kmers, out, res) -> dict:
    """Checkpoint and resume: saves the index as an image, frees it, and
    resumes it in a new process (bench.py --image-child), which loads the
    image with the saved configuration and searches a 10M-k-mer sample of the
    queries: wall times, file size, and whether the answers are the same."""
    import shutil
    import subprocess
    import tempfile

    d = tempfile.mkdtemp(prefix="sbwt_image_")
    try:
        path = os.path.join(d, "index.img")
        t = time.perf_counter()
        ix.save(path)
        save_s = time.perf_counter() - t
        size = os.path.getsize(path)
        log(f"image saved: {size / 1e9:.2f} GB in {save_s:.2f}s")
        sample = min(kmers.numel(), 10_000_000)
        np.save(os.path.join(d, "kmers.npy"), kmers[:sample].cpu().numpy())
        np.save(os.path.join(d, "want.npy"), res[:sample])
        ix.close()
        r = subprocess.run([sys.executable, "-u", os.path.abspath(__file__), "--image-child", d],
                           capture_output=True, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError(f"image child failed:\n{r.stderr[-3000:]}")
        child = json.loads(r.stdout.strip().splitlines()[-1])
        log(f"image resumed in a new process: loaded in {child['load_s']:.2f}s, "
            f"sample answers equal: {child['same_answers']}")
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"image_bytes": size, "save_s": round(save_s, 2), **child, "sample_kmers": sample}


def image_child(d: str) -> int:
    import torch

    import sbwt_gpu_amd as sg

    torch.zeros(1, device="cuda")  # the GPU context, outside the load time
    t = time.perf_counter()
    ix = sg.Index.load(os.path.join(d, "index.img"))
    torch.cuda.synchronize()
    load_s = time.perf_counter() - t
    kmers = torch.from_numpy(np.load(os.path.join(d, "kmers.npy"))).cuda()
    got = ix.search_kmers(kmers).cpu().numpy().view(np.uint64)
    same = bool(np.array_equal(got, np.load(os.path.join(d, "want.npy"))))
    info = ix.info
    ix.close()
    print(json.dumps({"load_s": round(load_s, 2), "same_answers": same,
                      "loaded_presearch_depth": info["presearch_depth"]}), flush=True)
    return 0


def run_batched(args) -> int:
    """10,000 read sets of 1-5,000 reads (lognormal length, median 300 bp,
    100 bp-20 kbp, 1% substitutions, 5% N) against the 100 Mbp index, split
    into batches of consecutive read sets; every batch is one
    sbwt_reads_plan_search call (packing, offsets, work prefix, search) on
    preallocated scratch. The timed step replays one HIP graph that holds all
    batches; the eager loop over the same batches is timed beside it."""
    import torch

    import sbwt_gpu_amd as sg
    from sbwt_gpu_amd import autotune, synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    k = args.k
    t = time.time()
    g = synth.genome_codes(args.genome_bp, seed=3, gc=0.4, seg_frac=0.10, seg_len=(1_000, 50_000), seg_mut=0.01)
    gcodes = torch.from_numpy(g).to(dev)
    gchars = torch.from_numpy(synth.to_ascii(g)).to(dev)
    ix = sg.Index.build(gchars, k, sg.Config(presearch_depth=args.presearch, layout=args.layout),
                        starts=torch.tensor([0, g.size], dtype=torch.int64, device=dev))
    del gchars
    log(f"index built: n={ix.n} ({time.time() - t:.1f}s with the genome)")
    t = time.time()
    chars, starts, set_starts = synth.gpu_read_sets(gcodes, args.sets, seed=5)
    torch.cuda.synchronize()
    del gcodes
    rows = starts.numel() - 1
    starts_h = starts.cpu().numpy()
    lens = np.diff(starts_h)
    kcount = np.maximum(lens - k + 1, 0)
    total = int(kcount.sum())
    log(f"{args.sets} read sets, {rows} reads, {starts_h[-1]} bases, {total} k-mers generated in "
        f"{time.time() - t:.1f}s")
    # batches of whole read sets, at most batch_chars characters
    ss = set_starts.cpu().numpy()
    cuts = [0]
    for si in range(1, ss.size):
        if starts_h[ss[si]] - starts_h[cuts[-1]] > args.batch_chars and ss[si - 1] > cuts[-1]:
            cuts.append(int(ss[si - 1]))
    if cuts[-1] != rows:
        cuts.append(rows)
    kcum = np.zeros(rows + 1, dtype=np.int64)
    kcum[1:] = np.cumsum(kcount)
    batches = []
    for b0, b1 in zip(cuts[:-1], cuts[1:]):
        c0, c1 = int(starts_h[b0]), int(starts_h[b1])
        rel = (starts[b0:b1 + 1] - c0).contiguous()
        batches.append((chars[c0:c1], rel, int(kcum[b0]), int(kcum[b1])))
    max_rows = max(b[1].numel() - 1 for b in batches)
    max_chars = max(b[0].numel() for b in batches)
    out = torch.empty(total, dtype=torch.int64, device=dev)
    log(f"{len(batches)} batches (max {max_rows} rows, {max_chars} chars)")

    tune = None
    seg = args.seg
    kflags = autotune.kernel_flags(("persistent" if args.mode.startswith("persistent") else "lane") +
                                   ("-runs" if args.run_items else ""))
    if not args.no_autotune:
        sample_chars, sample_rel, _, _ = batches[0]
        t = time.time()
        # depths up to 13: deeper tables win the trials (repeated searches of
        # one batch, whose table entries stay cached) but lose over the pass
        tune = autotune.autotune(ix, sample_chars, sample_rel, segs=(64, 256, 1024), budget_s=90.0,
                                 streams=max(1, args.streams))
        seg = tune["best"]["segment_kmers"]
        kflags = autotune.kernel_flags(tune["best"]["kernel"])
        log(f"autotune ({time.time() - t:.1f}s, {len(tune['trials'])} trials): {tune['best']}")
    # batch i runs on plan i % S and stream i % S: the packing and run pass of
    # one batch overlap the search of another (each plan counts its own work
    # items, so searches on different plans may overlap)
    S = max(1, args.streams)
    plans = [sg.ReadsPlan(max_rows, max_chars, seg) for _ in range(S)]
    plan = plans[0]
    lanes = [torch.cuda.Stream() for _ in range(S)]

    def run_all(stream=None):
        base = stream or torch.cuda.current_stream()
        for st in lanes:
            st.wait_stream(base)
        for i, (bc, brel, o0, o1) in enumerate(batches):
            plans[i % S].search(ix, bc, brel, out[o0:o1] if o1 > o0 else out[:1], stream=lanes[i % S], **kflags)
        for st in lanes:
            base.wait_stream(st)

    run_all()
    torch.cuda.synchronize()
    eager_ref = out.clone()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        run_all()
    torch.cuda.synchronize()
    eager_s = (time.perf_counter() - t0) / args.steps
    # capture every batch in one graph
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=side):
        run_all(stream=side)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        out.fill_(0)
        graph.replay()
    torch.cuda.synchronize()
    replay_ok = bool(torch.equal(out, eager_ref))
    t0 = time.perf_counter()
    for _ in range(args.steps):
        graph.replay()
    torch.cuda.synchronize()
    graph_s = (time.perf_counter() - t0) / args.steps
    replay_ok = replay_ok and bool(torch.equal(out, eager_ref))
    value = total / graph_s
    res = out.cpu().numpy().view(np.uint64)
    found = int((res != np.uint64(0xFFFFFFFFFFFFFFFF)).sum())
    # algorithmic lines: the counted instantiation over the same batches (untimed)
    ix.counters(reset=True)
    for bc, brel, o0, o1 in batches:
        plan.search(ix, bc, brel, out[o0:o1] if o1 > o0 else out[:1], count=True, **kflags)
    torch.cuda.synchronize()
    lines = int(ix.counters(reset=True)["lines"])
    counted_same = bool(torch.equal(out, eager_ref))
    achieved = lines * 128 / graph_s / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBPS, 4), "traffic": None,
                "kernel": "k_reads_uniform" if kflags["persistent"] else "k_search_reads", "kernel_avg_ms": None,
                "run_items": kflags["runs"],
                "pass_ms": round(graph_s * 1e3, 3), "lines_per_pass": lines,
                "bytes_per_kmer": round(lines * 128 / total, 2), "counted_equals_timed": counted_same,
                "note": "over the whole graph pass (every batch's packing, offsets, work prefix and search): "
                        "a lower bound of the search kernels' rate"}
    log(f"counted pass: {lines} lines ({lines / total:.3f}/k-mer), {achieved:.0f} GB/s of 128-B lines over the "
        f"pass; counted == timed: {counted_same}")
    # windows without N
    log(f"graph: {value / 1e9:.3f} G k-mers/s ({graph_s * 1e3:.1f} ms/pass), eager {total / eager_s / 1e9:.3f} G/s; "
        f"replays equal eager: {replay_ok}; found {found}/{total}")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_rows(ix, chars, starts_h, res, k, args)
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "k-mer queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(graph_s * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32", "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": f"batched uneven read sets: {args.sets} sets of 1-5000 reads, lognormal length "
                               f"(median 300, sigma 0.8, 100-20000 bp), 1% substitutions, 5% N, against the "
                               f"{args.genome_bp / 1e6:g} Mbp index; HIP-graph batches",
                   "reads": rows, "bases": int(starts_h[-1]), "kmers_per_step": total, "batches": len(batches),
                   "eager_kmers_per_s": round(total / eager_s, 1), "graph_replays_equal_eager": replay_ok,
                   "streams": S, "autotune_best": tune["best"] if tune else None,
                   "autotune_best_per_kernel": _best_per_kernel(tune), "found_fraction": round(found / total, 4)},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def _best_per_kernel(tune):
    """The fastest autotune trial of each kernel (work items x kernel)."""
    if not tune:
        return None
    best = {}
    for t in tune["trials"]:
        if t["kernel"] not in best or t["ms"] < best[t["kernel"]]["ms"]:
            best[t["kernel"]] = t
    return best


def run_text(args) -> int:
    """The whole query path from query text in host memory to result text:
    the read workload's 1M reads x 150 bp as a 151 MB text, one read per line
    (the reference's query file format), against the 100 Mbp index; result
    text written to /dev/null. Timed two ways on the same text: one shot
    (parse all, copy all in, search, copy all back, print: the reference's
    order of stages) and sbwt_search_text (batches of 16 MiB, parsing,
    copies and printing overlapping the searches). Both write the same
    bytes (compared on the first 20,000 lines)."""
    import torch

    import sbwt_gpu_amd as sg
    from sbwt_gpu_amd import synth

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    k = args.k
    g = synth.genome_codes(args.genome_bp, seed=3, gc=0.4, seg_frac=0.10, seg_len=(1_000, 50_000), seg_mut=0.01)
    gchars = torch.from_numpy(synth.to_ascii(g)).to(dev)
    ix = sg.Index.build(gchars, k, sg.Config(presearch_depth=args.presearch, layout=args.layout),
                        starts=torch.tensor([0, g.size], dtype=torch.int64, device=dev))
    del gchars
    chars_h, _ = synth.reads(g, args.reads, args.read_len, seed=4, sub=0.01)
    rows_h = np.asarray(chars_h, np.uint8).reshape(args.reads, args.read_len)
    text = np.concatenate([rows_h, np.full((args.reads, 1), ord("\n"), np.uint8)], axis=1).tobytes()
    log(f"query text: {len(text) / 1e6:.1f} MB, {args.reads} lines")
    total = args.reads * max(args.read_len - k + 1, 0)
    null = os.open(os.devnull, os.O_WRONLY)
    import ctypes as C
    import tempfile

    lib = sg.load()

    def one_shot(txt, fd):
        chars, starts = sg.parse_lines(txt)
        batch = sg.ReadBatch(torch.from_numpy(chars).to(dev), torch.from_numpy(starts.copy()).to(dev), k,
                             host_starts=starts)
        res = np.ascontiguousarray(ix.search_reads(batch).cpu().numpy().view(np.uint64))
        off = np.zeros(len(starts), np.uint64)
        off[1:] = np.cumsum(np.maximum(np.diff(starts) - k + 1, 0))
        sg.api.check(lib.sbwt_write_results(C.c_void_p(res.ctypes.data), C.c_void_p(off.ctypes.data),
                                            len(starts) - 1, fd))

    def piped(txt, fd, batch=16 << 20):
        return sg.search_text(ix, txt, fd=fd, batch_chars=batch)

    # the same bytes either way (the first 20,000 lines, in batches of 200 kB)
    head = text[: 20_000 * (args.read_len + 1)]
    outs = []
    for fn in (lambda fd: one_shot(head, fd), lambda fd: piped(head, fd, 200_000)):
        with tempfile.TemporaryFile() as f:
            fn(f.fileno())
            f.seek(0)
            outs.append(f.read())
    same = outs[0] == outs[1] and len(outs[0]) > 0

    times = {}
    for name, fn in (("one_shot", one_shot), ("pipelined", piped)):
        fn(text, null)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(max(args.steps // 4, 2)):
            fn(text, null)
        torch.cuda.synchronize()
        times[name] = (time.perf_counter() - t0) / max(args.steps // 4, 2)
        log(f"{name}: {total / times[name] / 1e9:.3f} G k-mers/s text to text ({times[name] * 1e3:.1f} ms)")
    os.close(null)
    log(f"same result text: {same}")
    line = {
        "metric": METRIC, "value": round(total / times["pipelined"], 1), "unit": "k-mer queries/s", "n_gpus": 1,
        "steps": max(args.steps // 4, 2), "warmup": 1, "ms_per_step": round(times["pipelined"] * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32",
        "data": "synthetic (seeded genome + reads generated in-run)",
        "config": {"workload": f"query text to result text: {args.reads} reads x {args.read_len} bp as one line "
                               f"each ({len(text) / 1e6:.0f} MB), {args.genome_bp / 1e6:g} Mbp index, k={k}, "
                               f"presearch {args.presearch}; host memory in, result text to /dev/null",
                   "kmers_per_step": total, "one_shot_kmers_per_s": round(total / times["one_shot"], 1),
                   "one_shot_ms": round(times["one_shot"] * 1e3, 3),
                   "pipelined_vs_one_shot": round(times["one_shot"] / times["pipelined"], 3),
                   "same_result_text": same, "batch_chars": 16 << 20},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)
    return 0


MLP_FLOPS_PER_KMER = 2 * (124 * 128 + 128 * 64 + 64)  # one-hot 31x4 -> 128 -> 64 -> 1


def run_prefilter(args) -> int:
    """The learned negative prefilter on a 90%-negative read mix. Index: the
    100 Mbp genome of the read-search workload. Queries: 1,000,000 reads of
    150 bp at 1% substitutions, 90% drawn from a second independent seeded
    100 Mbp genome (same generator, other seed) and 10% from the indexed one,
    shuffled. The MLP is trained on the GPU on SBWT-enumerated positives and
    uniform random k-mers the plain search confirms absent; a 1%-FPR Bloom
    filter holds the model's false negatives over every indexed k-mer.

    Timed, each over the same batch resident in HBM (one step = the whole
    query path): plain streaming search (certificates), plain independent
    search (the reference's per-k-mer semantics), filtered (k-mer
    extraction + fused MLP/Bloom prefilter + masked search, in streaming and
    independent form), and the filter without its learned part (a Bloom
    filter over all indexed k-mers). Every filtered answer vector must equal
    the plain one bit for bit."""
    import torch

    import sbwt_gpu_amd as sg
    from sbwt_gpu_amd import _capi, synth
    from sbwt_gpu_amd.api import _ptr, _stream
    from sbwt_gpu_amd.learned_filter import train_prefilter

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    dev = torch.device("cuda:0")
    k = args.k
    L = _capi.load()
    check = _capi.check
    t = time.time()
    g1 = synth.genome_codes(args.genome_bp, seed=3, gc=0.4, seg_frac=0.10, seg_len=(1_000, 50_000), seg_mut=0.01)
    g2 = synth.genome_codes(args.genome_bp, seed=13, gc=0.4, seg_frac=0.10, seg_len=(1_000, 50_000), seg_mut=0.01)
    log(f"two genomes of {g1.size} bp generated in {time.time() - t:.1f}s")
    gchars = torch.from_numpy(synth.to_ascii(g1)).to(dev)
    t = time.time()
    ix = sg.Index.build(gchars, k, sg.Config(presearch_depth=args.presearch, layout=args.layout),
                        starts=torch.tensor([0, g1.size], dtype=torch.int64, device=dev))
    del gchars
    info = ix.info
    log(f"index built on GPU in {time.time() - t:.1f}s: n={info['n']}")
    n_neg = int(round(args.reads * args.negative_fraction))
    c_neg, _ = synth.reads(g2, n_neg, args.read_len, seed=14, sub=0.01)
    c_pos, _ = synth.reads(g1, args.reads - n_neg, args.read_len, seed=4, sub=0.01)
    perm = np.random.default_rng(15).permutation(args.reads)
    rchars_h = np.concatenate([c_neg, c_pos]).reshape(args.reads, args.read_len)[perm].reshape(-1)
    is_pos_read = (perm >= n_neg)
    rstarts_h = np.arange(args.reads + 1, dtype=np.int64) * args.read_len
    del c_neg, c_pos
    rchars = torch.from_numpy(rchars_h).to(dev)
    rstarts = torch.from_numpy(rstarts_h).to(dev)
    rows = args.reads
    total = int(np.maximum(np.diff(rstarts_h) - k + 1, 0).sum())

    # ---- train the prefilter (untimed) --------------------------------------
    t = time.time()
    pf = train_prefilter(ix, positives=args.train_positives, negatives=args.train_negatives,
                         epochs=args.train_epochs, batch=args.train_batch, lr=args.lr, seed=7,
                         threshold=args.threshold, fpr=0.01, negatives_like=args.negatives_like,
                         calibrate_for=args.negative_fraction, log=log)
    train_total_s = time.time() - t
    log(f"prefilter ready in {train_total_s:.1f}s: {pf.stats}")
    # the method without its learned part: a Bloom filter over every indexed k-mer
    allpos = ix.enumerate_kmers()
    bloom_all = sg.Bloom(k, allpos.numel(), 0.01)
    bloom_all.insert(allpos)
    bloom_all_zero_fn = bool((bloom_all.query(allpos) == 1).all().item())
    # logit distribution on indexed k-mers and on fresh verified negatives (threshold table)
    gen = torch.Generator(device=dev)
    gen.manual_seed(99)
    from sbwt_gpu_amd.learned_filter import random_kmers
    vneg = random_kmers(4_000_000, k, gen)
    vneg = vneg[ix.search_kmers(vneg) == -1]
    zpos = pf.net.logits(allpos)
    zneg = pf.net.logits(vneg)
    thr_table = []
    for thr in (-4.0, -2.0, -1.0, 0.0, 1.0, 2.0, 4.0):
        thr_table.append({"threshold": thr, "model_fn_fraction": round(float((zpos < thr).float().mean()), 4),
                          "random_negative_pass": round(float((zneg >= thr).float().mean()), 4)})
    del allpos, zpos

    # ---- query-path buffers ------------------------------------------------
    n_chars = rchars.numel()
    packed = torch.empty(L.sbwt_packed_words(n_chars), dtype=torch.int64, device=dev)
    nmask = torch.empty(L.sbwt_nmask_words(n_chars), dtype=torch.int64, device=dev)
    offsets = torch.empty(rows + 1, dtype=torch.int64, device=dev)
    # work decompositions for the segment lengths tried per path (per-k-mer
    # searches gain nothing from long segments, so they also try short ones)
    segs = sorted({args.seg, 8, 16, 32, 64})
    work_of = {}
    for sl in segs:
        b = L.sbwt_reads_work_bound(rows, total, sl)
        work_of[sl] = (torch.empty(rows + 1, dtype=torch.int64, device=dev),
                       torch.empty(b, dtype=torch.int32, device=dev), b)
    pos = torch.empty(total, dtype=torch.int64, device=dev)
    kmers = torch.empty(total, dtype=torch.int64, device=dev)
    maybe = torch.empty(total, dtype=torch.uint8, device=dev)
    out = torch.empty(total, dtype=torch.int64, device=dev)
    s = _stream()
    stream = torch.cuda.current_stream()
    nsteps = max(args.steps, 1)

    def prep(seg):
        work, items, _ = work_of[seg]
        check(L.sbwt_pack(_ptr(rchars), n_chars, _ptr(packed), _ptr(nmask), s))
        check(L.sbwt_kmer_offsets(_ptr(rstarts), rows, k, _ptr(offsets), s))
        check(L.sbwt_reads_work(_ptr(offsets), rows, seg, _ptr(work), _ptr(items), s))

    def search(flags, masked, seg):
        work, items, bound = work_of[seg]
        a = (ix._h, _ptr(packed), _ptr(nmask), _ptr(rstarts), _ptr(offsets), _ptr(work), _ptr(items), rows, bound,
             seg)
        if masked:
            check(L.sbwt_search_reads_masked(*a, _ptr(maybe), _ptr(out), flags, s))
        else:
            check(L.sbwt_search_reads(*a, _ptr(out), flags, s))

    def make_path(flags, filt, seg):
        """filt: None (plain), "learned" (MLP + Bloom over its false
        negatives) or "bloom" (Bloom over every indexed k-mer)."""
        def run(ev=None):
            prep(seg)
            if filt == "learned" and pf.mlp is not None:  # the matrix-core kernel takes extracted k-mers
                check(L.sbwt_positions(_ptr(rstarts), _ptr(offsets), rows, k, _ptr(pos), s))
                check(L.sbwt_extract_kmers(_ptr(packed), _ptr(pos), total, _ptr(kmers), s))
                if ev:
                    ev[0].record(stream)
                pf.maybe(kmers, out=maybe)
            elif filt == "learned":  # calibrated without the network: its Bloom filter, from the reads
                if ev:
                    ev[0].record(stream)
                if pf.bloom is not None:
                    check(L.sbwt_bloom_query_reads(pf.bloom._h, _ptr(packed), _ptr(rstarts), _ptr(offsets), rows,
                                                   _ptr(maybe), s))
                else:
                    maybe.fill_(1)
            elif filt == "bloom":  # straight from the packed reads
                if ev:
                    ev[0].record(stream)
                check(L.sbwt_bloom_query_reads(bloom_all._h, _ptr(packed), _ptr(rstarts), _ptr(offsets), rows,
                                               _ptr(maybe), s))
            if ev:
                ev[1].record(stream)
            search(flags, filt is not None, seg)
            if ev:
                ev[2].record(stream)
        return run

    def tuned(flags, filt):
        """Each path at its fastest segment length and kernel on this batch
        (untimed trials; every trial must give the same answers)."""
        kern = ((0, "uniform"), (_capi.READS_SEGMENT_LANES, "lanes")) if not flags & _capi.READS_INDEPENDENT else \
            ((0, "lanes"), (_capi.READS_PERSISTENT, "uniform"))
        best, ref = None, None
        for sl in segs:
            if sl < 32 and not flags & _capi.READS_INDEPENDENT:
                continue
            for kf, kname in kern:
                run = make_path(flags | kf, filt, sl)
                run()
                ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
                ev[0].record(stream)
                run()
                run()
                ev[1].record(stream)
                torch.cuda.synchronize()
                ms = ev[0].elapsed_time(ev[1]) / 2
                if ref is None:
                    ref = out.clone()
                elif not torch.equal(out, ref):
                    raise RuntimeError(f"answers differ at segment {sl}, kernel {kname}")
                if best is None or ms < best[0]:
                    best = (ms, sl, kname, run)
        return best

    specs = {
        "plain": (0, None),
        "plain_independent": (_capi.READS_INDEPENDENT, None),
        "filtered": (0, "learned"),
        "filtered_independent": (_capi.READS_INDEPENDENT, "learned"),
        "bloom_only": (0, "bloom"),
        "bloom_only_independent": (_capi.READS_INDEPENDENT, "bloom"),
    }
    paths, chosen = {}, {}
    for name, (fl, filt) in specs.items():
        ms, sl, kname, run = tuned(fl, filt)
        paths[name] = run
        chosen[name] = {"segment_kmers": sl, "kernel": kname}
    log(f"per-path segment length and kernel: {chosen}")
    results, answers = {}, {}
    for name, run in paths.items():
        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        evs = [[torch.cuda.Event(enable_timing=True) for _ in range(3)] for _ in range(nsteps)]
        t0 = time.perf_counter()
        for i in range(nsteps):
            run(evs[i])
        torch.cuda.synchronize()
        el = (time.perf_counter() - t0) / nsteps
        filt_ms = float(np.mean([e[0].elapsed_time(e[1]) for e in evs])) if name not in (
            "plain", "plain_independent") else 0.0
        search_ms = float(np.mean([e[1].elapsed_time(e[2]) for e in evs]))
        answers[name] = out.cpu().numpy().view(np.uint64).copy()
        mrate = float(maybe.float().mean().item()) if "plain" not in name else None
        results[name] = {"kmers_per_s": round(total / el, 1), "ms_per_step": round(el * 1e3, 3),
                         "filter_kernel_ms": round(filt_ms, 3), "search_kernel_ms": round(search_ms, 3),
                         "maybe_fraction": round(mrate, 4) if mrate is not None else None, **chosen[name]}
        log(f"{name}: {total / el / 1e9:.3f} G k-mers/s ({el * 1e3:.2f} ms/step; filter {filt_ms:.2f} ms, "
            f"search {search_ms:.2f} ms, maybe {mrate})")
    plain = answers["plain"]
    exact = {n: bool(np.array_equal(a, plain)) for n, a in answers.items()}
    log(f"answers equal the plain search: {exact}")
    none = np.uint64(0xFFFFFFFFFFFFFFFF)
    found = int((plain != none).sum())
    per_read = args.read_len - k + 1
    pos_kmer = np.repeat(is_pos_read, per_read)
    neg_frac_kmers = float(1.0 - (plain != none).mean())
    # the batch's k-mers (the paths extract them only when the network runs)
    check(L.sbwt_positions(_ptr(rstarts), _ptr(offsets), rows, k, _ptr(pos), s))
    check(L.sbwt_extract_kmers(_ptr(packed), _ptr(pos), total, _ptr(kmers), s))
    maybe_l = pf.maybe(kmers).cpu().numpy().astype(bool)
    m_on_absent = float(maybe_l[plain == none].mean())
    m_on_negreads = float(maybe_l[~pos_kmer].mean())

    # roofline of the filtered path's dominant kernel: the matrix-core pass
    # (k_prefilter) timed alone, without the Bloom pass over its rejects
    fr = results["filtered"]
    mlp_ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    # (the network's kernel is timed whether or not calibration kept it)
    check(L.sbwt_positions(_ptr(rstarts), _ptr(offsets), rows, k, _ptr(pos), s))
    check(L.sbwt_extract_kmers(_ptr(packed), _ptr(pos), total, _ptr(kmers), s))
    net_thr = pf.threshold if pf.mlp is not None else 0.0
    sg.prefilter(pf.net, net_thr, None, kmers, out=maybe)
    mlp_ev[0].record(stream)
    for _ in range(5):
        sg.prefilter(pf.net, net_thr, None, kmers, out=maybe)
    mlp_ev[1].record(stream)
    torch.cuda.synchronize()
    mlp_ms = mlp_ev[0].elapsed_time(mlp_ev[1]) / 5
    pf_s = mlp_ms / 1e3
    ix.counters(reset=True)
    pf.maybe(kmers, out=maybe)
    fk = _capi.READS_SEGMENT_LANES if chosen["filtered"]["kernel"] == "lanes" else 0
    prep(chosen["filtered"]["segment_kmers"])
    search(_capi.COUNT | fk, True, chosen["filtered"]["segment_kmers"])
    torch.cuda.synchronize()
    ctr = ix.counters(reset=True)
    search_gbps = int(ctr["lines"]) * 128 / (fr["search_kernel_ms"] / 1e3) / 1e9
    mlp_tflops = total * MLP_FLOPS_PER_KMER / pf_s / 1e12
    mlp_roof = {"bound": "mfma", "achieved": round(mlp_tflops, 2), "peak": 2500.0, "unit": "TFLOP/s",
                "frac": round(mlp_tflops / 2500.0, 4), "traffic": None, "kernel": "k_prefilter",
                "kernel_avg_ms": round(mlp_ms, 3), "flops_per_kmer": MLP_FLOPS_PER_KMER,
                "filter_with_bloom_pass_ms": fr["filter_kernel_ms"]}
    search_roof = {"bound": "hbm", "achieved": round(search_gbps, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                   "frac": round(search_gbps / HBM_PEAK_GBPS, 4), "traffic": None,
                   "kernel": ("k_search_reads" if fk else "k_reads_uniform") + " (masked)",
                   "kernel_avg_ms": fr["search_kernel_ms"],
                   "lines_per_launch": int(ctr["lines"])}
    roofline = dict(mlp_roof if pf.mlp is not None and mlp_ms >= fr["search_kernel_ms"] else search_roof)
    roofline["other_kernel"] = search_roof if roofline["kernel"] == "k_prefilter" else mlp_roof
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(ix, rchars_h, rstarts_h, plain, k, args)
    value = results["filtered"]["kmers_per_s"]
    line = {
        "metric": METRIC, "value": value, "unit": "k-mer queries/s", "n_gpus": 1, "steps": nsteps,
        "warmup": args.warmup, "ms_per_step": results["filtered"]["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u32 search, bf16/fp32 MLP",
        "data": "synthetic (seeded genomes + reads generated in-run)",
        "config": {"workload": f"learned prefilter: {args.genome_bp / 1e6:g} Mbp index (GC 40%, 10% segmental "
                               f"copies), k={k}, presearch {args.presearch}; {rows} reads x {args.read_len} bp at "
                               f"1% substitutions, {args.negative_fraction:.0%} from a second independent "
                               f"{args.genome_bp / 1e6:g} Mbp genome; MLP 124-128-64-1 bf16/fp32 trained on "
                               f"{pf.stats['train_positives']} enumerated positives + "
                               f"{pf.stats['train_negatives']} verified negatives; Bloom 1% FPR over its false "
                               f"negatives",
                   "kmers_per_step": total, "absent_fraction": round(neg_frac_kmers, 4), "found": found,
                   "paths": results, "answers_equal_plain": exact,
                   "filtered_vs_plain": round(results["filtered"]["kmers_per_s"] / results["plain"]["kmers_per_s"], 3),
                   "filtered_independent_vs_plain_independent": round(
                       results["filtered_independent"]["kmers_per_s"] / results["plain_independent"]["kmers_per_s"], 3),
                   "learned_vs_bloom_only": round(results["filtered"]["kmers_per_s"] /
                                                  results["bloom_only"]["kmers_per_s"], 3),
                   "maybe_on_absent_kmers": round(m_on_absent, 4),
                   "maybe_on_negative_read_kmers": round(m_on_negreads, 4),
                   "threshold": pf.stats.get("threshold"), "threshold_table": thr_table,
                   "network_used": pf.stats.get("network_used"),
                   "bloom_only_bytes": bloom_all.bytes, "bloom_only_zero_fn": bloom_all_zero_fn,
                   "train": pf.stats, "train_total_s": round(train_total_s, 1)},
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


def cpu_baseline_rows(ix, chars, starts_h, res_gpu, k, args):
    """The CPU oracle's read search on the first rows of the batch + parity."""
    from oracle import oracle as O

    rows = ix.rows()
    oix = O.Index(rows[:4], ix.n, sb=1024)
    oix.presearch(min(ix.info["presearch_depth"], 12))  # answers do not depend on the depth
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    nrows = min(len(starts_h) - 1, 200_000)
    c_end = int(starts_h[nrows])
    ch = chars[:c_end].cpu().numpy()
    rs = starts_h[: nrows + 1].astype(np.uint64)
    t = time.perf_counter()
    want = O.search_reads(oix, ch, rs, k)
    dt = time.perf_counter() - t
    parity = bool(np.array_equal(want, res_gpu[: want.size]))
    log(f"CPU oracle: {want.size / dt / 1e6:.2f} M k-mers/s on {threads} threads; GPU parity on {nrows} reads: "
        f"{parity}")
    return {"value": round(want.size / dt, 1), "unit": "k-mer queries/s", "cores": threads, "kind": "port",
            "sample": f"first {nrows} reads ({want.size} k-mers), independent search with N-breaking, poppy "
                      f"SB=1024", "gpu_parity_on_sample": parity, "cpu_model": _cpu_model()}


def cpu_baseline_kmers(ix, kmers, res_gpu, k, args):
    """The CPU oracle on a bounded sample of the same k-mers + GPU parity."""
    from oracle import oracle as O

    t = time.time()
    rows = ix.rows()
    oix = O.Index(rows[:4], ix.n, sb=1024)
    m_cpu = min(args.presearch, 12)  # a 4^16 or 4^17 table is not built on the host
    oix.presearch(m_cpu)
    log(f"oracle index + presearch built in {time.time() - t:.1f}s")
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    km_all = kmers[: 50_000_000].cpu().numpy().view(np.uint64)
    probe = km_all[:200_000]
    t = time.perf_counter()
    oix.search(probe, k)
    rate = probe.size / (time.perf_counter() - t)
    n = int(min(km_all.size, max(200_000, rate * args.cpu_seconds)))
    km = km_all[:n]
    t = time.perf_counter()
    want = oix.search(km, k)
    dt = time.perf_counter() - t
    parity = bool(np.array_equal(want, res_gpu[:n]))
    km1 = km[: max(1, n // max(threads, 1) // 4)]
    t = time.perf_counter()
    oix.search(km1, k, threads=1)
    rate1 = km1.size / (time.perf_counter() - t)
    log(f"CPU oracle: {n / dt / 1e6:.2f} M k-mers/s on {threads} threads, {rate1 / 1e6:.2f} M/s on 1; "
        f"GPU parity on the sample: {parity}")
    return {"value": round(n / dt, 1), "unit": "k-mer queries/s", "cores": threads, "kind": "port",
            "sample": f"first {n} of the same k-mers, poppy SB=1024, presearch {m_cpu}, OpenMP over k-mers",
            "single_thread_value": round(rate1, 1), "gpu_parity_on_sample": parity, "cpu_model": _cpu_model()}


def measure_traffic(args):
    """HBM bytes per launch of the search kernel from rocprofv3 --pmc child
    runs of this workload with --pmc-child (which also copies 1 GiB three
    times): FETCH_SIZE, then WRITE_SIZE (KiB per dispatch), each scaled by
    the factor that makes it read 1 GiB for that copy (HIP's device-to-device
    copy kernel, __amd_rocclr_copyBuffer); a third run reads TCC_HIT_sum and
    TCC_MISS_sum, the L2 hit rate of the search's requests; a fourth the
    L2's read requests to memory by size (TCC_EA0_RDREQ_32B/64B/128B), whose
    bytes need no scale."""
    import csv
    import shutil
    import subprocess
    import tempfile

    prof = shutil.which("rocprofv3")
    if not prof:
        return {"error": "rocprofv3 not found"}
    large = args.workload in ("large", "uniform")
    # large: the child names the k-mer kernel it times (its autotuner may pick
    # any layout and kernel) and its timed dispatch is that kernel's last
    kern = None if large else (
        "k_reads_uniform" if args.mode in ("default", "persistent", "persistent-nocert", "persistent-independent") else "k_search_reads")
    child = [sys.executable, "-u", os.path.abspath(__file__), "--pmc-child",
             "--workload", "uniform" if large else args.workload,
             "--mode", args.mode, "--layout", str(args.layout), "--presearch", str(args.presearch),
             "--large-layout", str(args.large_layout), "--kmer-kernel", args.kmer_kernel,
             "--seg", str(args.seg), "--genome-bp", str(args.genome_bp), "--reads", str(args.reads),
             "--read-len", str(args.read_len), "--k", str(args.k), "--large-bp", str(args.large_bp),
             "--queries", str(min(args.queries, 250_000_000))]
    out = {"kernel": kern, "counters": {}}
    env = dict(os.environ, TMPDIR=tempfile.gettempdir())

    def run_pass(ctrs):
        """Per-dispatch values of ctrs for the copies (in dispatch order) and
        the timed search kernel; raises RuntimeError with what went wrong."""
        tag = ctrs[0]
        d = tempfile.mkdtemp(prefix=f"sbwt_pmc_{tag}_")
        cmd = [prof, "--pmc", *ctrs, "-f", "csv", "-d", d, "-o", "run", "--", *child]
        log(f"traffic: {' '.join(cmd[:8])} ... (child run)")
        try:
            r = subprocess.run(cmd, capture_output=True, text=True, timeout=300, env=env, cwd=tempfile.gettempdir())
        except subprocess.TimeoutExpired:
            raise RuntimeError(f"{tag} run timed out")
        files = [os.path.join(dp, f) for dp, _, fs in os.walk(d) for f in fs if f.endswith("counter_collection.csv")]
        if not files:
            raise RuntimeError(f"{tag} run produced no counters (rc {r.returncode}): {r.stderr[-300:]}")
        name_ = kern
        if large:
            m = re.search(r"kmer kernel: (\S+)", r.stderr)
            if not m:
                raise RuntimeError(f"{tag}: the child did not name its k-mer kernel")
            name_ = m.group(1)
            out["kernel"] = name_
        copy, search = {c: [] for c in ctrs}, {c: [] for c in ctrs}
        with open(files[0]) as fh:
            rows_ = list(csv.DictReader(fh))
        if rows_ and "Dispatch_Id" in rows_[0]:
            rows_.sort(key=lambda r: int(r["Dispatch_Id"]))
        for row in rows_:
            name, c, v = row["Kernel_Name"], row.get("Counter_Name", tag), float(row["Counter_Value"])
            if c not in search:
                continue
            if "__amd_rocclr_copyBuffer" in name:
                copy[c].append(v)
            elif name_ in name and ", true>" not in name.split("(")[0][-8:]:
                search[c].append(v)
        shutil.rmtree(d, ignore_errors=True)
        if not all(search.values()):
            raise RuntimeError(f"{tag}: search kernel missing from the counters")
        # reads: the median dispatch (the build's own check also launches the
        # search kernel, on a few million sampled k-mers); large: the last
        # (the autotuner's trials run before it, on fewer k-mers)
        return copy, {c: float(v[-1] if large else np.median(v)) for c, v in search.items()}

    for ctrs in (("FETCH_SIZE",), ("WRITE_SIZE",), ("TCC_HIT_sum", "TCC_MISS_sum")):
        tag = ctrs[0]
        try:
            copy, med = run_pass(ctrs)
        except RuntimeError as e:
            return {"error": str(e)}
        if len(ctrs) == 2:  # the L2 hit rate: a ratio, no calibration
            hit, miss = med[ctrs[0]], med[ctrs[1]]
            out["tcc"] = {"hit_per_dispatch": round(hit), "miss_per_dispatch": round(miss),
                          "hit_rate": round(hit / max(hit + miss, 1.0), 4)}
            continue
        if not copy[tag]:
            return {"error": f"{tag}: copy kernel missing from the counters"}
        # the three 1 GiB copies are the child's first copy dispatches (it
        # makes them before anything else; later copies, such as tensor
        # clones, have other sizes)
        big = copy[tag][:3]
        copy_kib = sum(big) / len(big)
        if copy_kib < 100_000:
            return {"error": f"{tag}: the 1 GiB calibration copies were not found"}
        factor = (1 << 20) / copy_kib
        kib = med[tag]
        out["counters"][tag] = {"search_KiB_per_dispatch": round(kib, 1), "copy_KiB": round(copy_kib, 1),
                                "scale": round(factor, 4), "bytes": round(kib * 1024 * factor)}
    # The L2's read requests to memory by size. FETCH_SIZE counts a request as
    # 64 bytes unless TCC_BUBBLE marks it 128 (on gfx950 TCC_BUBBLE reads 0),
    # so 128-byte requests count half: the copy's x2 factor. Read bytes from
    # the request sizes need no factor; the copies check them (1 GiB each).
    sizes = ("TCC_EA0_RDREQ_sum", "TCC_EA0_RDREQ_32B_sum", "TCC_EA0_RDREQ_64B_sum", "TCC_EA0_RDREQ_128B_sum")

    def req_bytes(m):
        tot, r32, r64, r128 = (m[c] for c in sizes)
        return 128 * r128 + 64 * r64 + 32 * r32 + 64 * max(tot - r32 - r64 - r128, 0.0)

    try:
        copy, med = run_pass(sizes)
        big = [dict(zip(sizes, v)) for v in zip(*(copy[c][:3] for c in sizes))]
        copy_b = float(np.mean([req_bytes(m) for m in big])) if big else 0.0
        out["read_requests"] = {
            "per_dispatch": {c.replace("TCC_EA0_RDREQ", "").strip("_").replace("_sum", "").replace("sum", "all"):
                             round(med[c]) for c in sizes},
            "read_bytes": round(req_bytes(med)),
            "copy_read_bytes_over_1GiB": round(copy_b / (1 << 30), 4) if copy_b else None}
    except RuntimeError as e:
        out["read_requests"] = {"error": str(e)}
    out["queries_per_launch"] = min(args.queries, 250_000_000) if large else None
    return out


def traffic_bytes(t):
    """FETCH + WRITE bytes per launch (scaled), or None."""
    if not t or "counters" not in t or len(t["counters"]) != 2:
        return None
    return t["counters"]["FETCH_SIZE"]["bytes"] + t["counters"]["WRITE_SIZE"]["bytes"]


def gather_roof(gib: int = 4, nbytes: int | None = None):
    """Uniform-random 128-byte-line gathers over a `gib` GiB buffer (far
    beyond the 256 MiB last-level cache): the random-access roof. Two shapes:
    one 16-byte read per random line (the search's access), and whole lines
    read by 8 lanes together. `nbytes` (a multiple of 128) replaces the size:
    the roof for a working set of that size, which the caches partly hold."""
    import torch

    from sbwt_gpu_amd import _capi
    from sbwt_gpu_amd.api import _ptr, _stream

    L = _capi.load()
    nbytes = (gib << 30) if nbytes is None else max(128, nbytes // 128 * 128)
    buf = torch.empty(nbytes // 8, dtype=torch.int64, device="cuda")
    buf.fill_(0x5A5A)
    lanes, per = 256 * 1024 * 8, 64
    sink = torch.empty(lanes, dtype=torch.int64, device="cuda")
    s = _stream()
    st = torch.cuda.current_stream()
    out = {"buffer_GiB": round(nbytes / (1 << 30), 3)}
    for name, full in (("one_16B_read_per_line", 0), ("whole_line_8_lanes", 1)):
        _capi.check(L.sbwt_gather_roof(_ptr(buf), nbytes, 128, full, per, lanes, 7, _ptr(sink), s))
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        reps = 5
        e0.record(st)
        for r in range(reps):
            _capi.check(L.sbwt_gather_roof(_ptr(buf), nbytes, 128, full, per, lanes, 11 + r, _ptr(sink), s))
        e1.record(st)
        torch.cuda.synchronize()
        secs = e0.elapsed_time(e1) / 1e3
        nl = lanes * per * reps / (8 if full else 1)
        out[name] = {"lines_per_s": round(nl / secs, 1), "GBps_128B_lines": round(nl * 128 / secs / 1e9, 1)}
    out["GBps_128B_lines"] = max(out["one_16B_read_per_line"]["GBps_128B_lines"],
                                 out["whole_line_8_lanes"]["GBps_128B_lines"])
    del buf
    torch.cuda.empty_cache()
    log(f"random-line roof: {out}")
    return out


def calib_copy() -> int:
    """A plain copy of 1 GiB, three times: reads 1 GiB and writes 1 GiB per
    launch, for scaling the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    a = torch.ones(1 << 27, dtype=torch.int64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    log("copied 1 GiB three times")
    return 0


def cpu_baseline(ix, rchars_h, rstarts_h, res_gpu, k, args):
    """The CPU oracle (reference search restated: poppy SB=1024, presearch m)
    on a bounded sample of the same k-mers; also checks GPU parity there."""
    from oracle import oracle as O

    t = time.time()
    rows = ix.rows()
    oix = O.Index(rows[:4], ix.n, sb=1024)
    oix.presearch(args.presearch)
    log(f"oracle index + presearch built in {time.time() - t:.1f}s")
    per_read = args.read_len - k + 1
    threads = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1

    def sample(nreads):
        st = rstarts_h[: nreads + 1]
        km, _ = O.read_kmers(rchars_h[: st[-1]], st.astype(np.uint64), k)
        return km

    probe = sample(20_000)
    t = time.perf_counter()
    oix.search(probe, k)
    rate = probe.size / (time.perf_counter() - t)
    nreads = int(min(len(rstarts_h) - 1, max(20_000, rate * args.cpu_seconds / per_read)))
    km = sample(nreads)
    t = time.perf_counter()
    want = oix.search(km, k)
    dt = time.perf_counter() - t
    parity = bool(np.array_equal(want, res_gpu[: km.size]))
    km1 = km[: max(1, km.size // max(threads, 1) // 4)]
    t = time.perf_counter()
    oix.search(km1, k, threads=1)
    rate1 = km1.size / (time.perf_counter() - t)
    log(f"CPU oracle: {km.size / dt / 1e6:.2f} M k-mers/s on {threads} threads, {rate1 / 1e6:.2f} M/s on 1; "
        f"GPU parity on the sample: {parity}")
    return {"value": round(km.size / dt, 1), "unit": "k-mer queries/s", "cores": threads, "kind": "port",
            "sample": f"first {nreads} reads ({km.size} k-mers) of the same batch, independent search, "
                      f"poppy SB=1024, presearch {args.presearch}, OpenMP over k-mers",
            "single_thread_value": round(rate1, 1), "gpu_parity_on_sample": parity,
            "cpu_model": _cpu_model()}


def _cpu_model():
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


if __name__ == "__main__":
    sys.exit(main())
