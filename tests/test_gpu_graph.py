"""Batched read search on preallocated scratch (sbwt_reads_plan_search) and
its capture in a HIP graph: every replay, with eager work and
synchronizations in between, must equal the eager run and the oracle."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a GPU", allow_module_level=True)

import sbwt_gpu_amd as sg  # noqa: E402
from sbwt_gpu_amd import autotune, synth  # noqa: E402

ACGT = np.array(list("ACGT"))


def _setup(seed=1, k=31):
    rng = np.random.default_rng(seed)
    g = synth.genome_codes(150_000, seed, gc=0.45)
    gs = synth.to_ascii(g).tobytes().decode()
    s = O.SBWT([gs], k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=10))
    chars, starts, set_starts = synth.gpu_read_sets(torch.from_numpy(g).cuda(), 12, seed, max_reads=60,
                                                    median=200, lo=20, hi=2000, n_rate=0.03)
    oix = O.Index(s.rows, s.n)
    oix.presearch(10)
    return rng, s, ix, oix, chars, starts, set_starts


def _batches(chars, starts, set_starts, per=3):
    ss = set_starts.cpu().numpy()
    sh = starts.cpu().numpy()
    out = []
    for i in range(0, ss.size - 1, per):
        b0, b1 = int(ss[i]), int(ss[min(i + per, ss.size - 1)])
        c0, c1 = int(sh[b0]), int(sh[b1])
        out.append((chars[c0:c1], (starts[b0:b1 + 1] - c0).contiguous(), b0, b1))
    return out


@pytest.mark.parametrize("persistent", [False, True])
def test_plan_batches_match_oracle_and_graph_replays(persistent):
    rng, s, ix, oix, chars, starts, set_starts = _setup()
    k = 31
    sh = starts.cpu().numpy()
    kc = np.maximum(np.diff(sh) - k + 1, 0)
    kcum = np.concatenate([[0], np.cumsum(kc)])
    total = int(kcum[-1])
    want = O.search_reads(oix, chars.cpu().numpy(), sh.astype(np.uint64), k)
    batches = _batches(chars, starts, set_starts)
    plan = sg.ReadsPlan(max(b[1].numel() - 1 for b in batches), max(b[0].numel() for b in batches), seg=64)
    out = torch.empty(total, dtype=torch.int64, device="cuda")

    def run_all(stream=None):
        for bc, brel, b0, b1 in batches:
            plan.search(ix, bc, brel, out[int(kcum[b0]):int(kcum[b1])], stream=stream, persistent=persistent)

    run_all()
    torch.cuda.synchronize()
    eager = out.clone()
    np.testing.assert_array_equal(sg.to_u64(eager), want)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        run_all(stream=side)
    torch.cuda.synchronize()
    for rep in range(4):
        out.fill_(-7)
        torch.cuda.synchronize()
        _ = (torch.arange(1000, device="cuda") * rep).sum().item()  # unrelated eager work
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), f"replay {rep}"


@pytest.mark.parametrize("kw", [{"persistent": True}, {"persistent": True, "runs": True}, {"lanes": True}])
def test_two_plans_on_two_streams_overlap_safely(kw):
    """Batches alternate over two plans on two streams, so the searches of
    one index overlap (each plan counts its own work items): eagerly and in
    one captured graph, every replay must equal the oracle."""
    rng, s, ix, oix, chars, starts, set_starts = _setup(seed=8)
    k = 31
    sh = starts.cpu().numpy()
    kc = np.maximum(np.diff(sh) - k + 1, 0)
    kcum = np.concatenate([[0], np.cumsum(kc)])
    want = O.search_reads(oix, chars.cpu().numpy(), sh.astype(np.uint64), k)
    batches = _batches(chars, starts, set_starts, per=1)
    mr, mc = max(b[1].numel() - 1 for b in batches), max(b[0].numel() for b in batches)
    plans = [sg.ReadsPlan(mr, mc, seg=64) for _ in range(2)]
    lanes = [torch.cuda.Stream() for _ in range(2)]
    out = torch.full((int(kcum[-1]),), 7, dtype=torch.int64, device="cuda")

    def run_all(base):
        for st in lanes:
            st.wait_stream(base)
        for i, (bc, brel, b0, b1) in enumerate(batches):
            plans[i % 2].search(ix, bc, brel, out[int(kcum[b0]):int(kcum[b1])], stream=lanes[i % 2], **kw)
        for st in lanes:
            base.wait_stream(st)

    run_all(torch.cuda.current_stream())
    torch.cuda.synchronize()
    np.testing.assert_array_equal(sg.to_u64(out), want)
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=side):
        run_all(side)
    torch.cuda.synchronize()
    for rep in range(3):
        out.fill_(-7)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        np.testing.assert_array_equal(sg.to_u64(out), want, err_msg=f"replay {rep}")


def test_plan_rejects_oversized_batch():
    rng, s, ix, oix, chars, starts, set_starts = _setup(seed=2)
    plan = sg.ReadsPlan(3, 100)
    out = torch.empty(chars.numel(), dtype=torch.int64, device="cuda")
    with pytest.raises(sg.SbwtError, match="exceeds the plan"):
        plan.search(ix, chars, starts, out)


@pytest.mark.parametrize("streams", [1, 3])
def test_autotune_keeps_answers_and_applies_best(streams):
    """The trials (one stream, or three concurrent streams with a plan and an
    output each) all give the same answers, the best is applied, and it
    answers as the oracle does."""
    rng, s, ix, oix, chars, starts, set_starts = _setup(seed=3)
    res = autotune.autotune(ix, chars, starts, layouts=(16, 32), depths=(8, 10), segs=(64, 256), reps=1,
                            streams=streams)
    best = res["best"]
    assert len(res["trials"]) == 2 * 2 * 2 * len(autotune.KERNELS)
    assert ix.info["layout"] == best["layout"] and ix.info["presearch_depth"] == best["presearch_depth"]
    want = O.search_reads(oix, chars.cpu().numpy(), starts.cpu().numpy().astype(np.uint64), 31)
    plan = sg.ReadsPlan(starts.numel() - 1, chars.numel(), seg=best["segment_kmers"])
    out = torch.empty(want.size, dtype=torch.int64, device="cuda")
    plan.search(ix, chars, starts, out, **autotune.kernel_flags(best["kernel"]))
    np.testing.assert_array_equal(sg.to_u64(out), want)


def test_autotune_kmers_keeps_answers_and_applies_best():
    rng, s, ix, oix, chars, starts, set_starts = _setup(seed=4)
    g = synth.genome_codes(150_000, 4, gc=0.45)
    pos = rng.integers(0, g.size - 31 + 1, 20_000)
    q = np.zeros(pos.size, np.uint64)
    for t in range(31):
        q |= g[pos + t].astype(np.uint64) << np.uint64(62 - 2 * t)
    q = np.concatenate([q, rng.integers(0, 2**62, 20_000, dtype=np.uint64) << np.uint64(2)])
    want = oix.search(q, 31)
    dq = torch.from_numpy(q.view(np.int64)).cuda()
    res = autotune.autotune_kmers(ix, dq, sample=30_000, reps=1)
    best = res["best"]
    assert len(res["trials"]) == 5 and ix.info["layout"] == best["layout"]
    got = sg.to_u64(ix.search_kmers(dq, cooperative=best["kernel"] == "cooperative"))
    np.testing.assert_array_equal(got, want)
    assert (want != sg.NONE).sum() >= 20_000


# ---------------------------------------------------------------------------
# run items (SBWT_READS_RUN_ITEMS): work items are the rows' N-free runs


def _ragged_rows(rng, g, k):
    """Rows that reach every case of the run pass: shorter than k, exactly k,
    all N, no N at all (longer than 64 windows and than a segment), N at
    window-chunk boundaries (window 63, 64, 127), N in the first and last
    character, dense N, and lowercase / other non-ACGT bytes."""
    def piece(n):
        p = int(rng.integers(0, len(g) - n))
        return g[p:p + n]
    rows = [piece(k - 1), piece(k), "N" * 80, piece(2000), piece(700), "N" + piece(300), piece(300) + "N",
            piece(200).lower(), piece(90) + "RYKM" + piece(90), ""]
    for pos in (63, 64, 127, 128, 63 + k - 1, 64 + k - 1):
        r = list(piece(400))
        r[pos] = "N"
        rows.append("".join(r))
    for rate in (0.05, 0.2, 0.5):
        r = np.array(list(piece(1500)))
        r[rng.random(r.size) < rate] = "N"
        rows.append("".join(r))
    r = list(piece(500))  # a sequencing error inside a run, so certificates fire
    r[250] = "ACGT"[("ACGT".index(r[250]) + 1) % 4]
    rows.append("".join(r))
    return rows


@pytest.mark.parametrize("seg", [17, 64, 256])
@pytest.mark.parametrize("kernel", ["persistent", "lane"])
def test_run_items_match_oracle(seg, kernel):
    rng, s, ix, oix, chars, starts, set_starts = _setup(seed=5)
    k = 31
    gtxt = synth.to_ascii(synth.genome_codes(150_000, 5, gc=0.45)).tobytes().decode()
    rows = _ragged_rows(rng, gtxt, k)
    ch, rs = O.parse_lines("\n".join(rows).encode())
    # parse_lines skips empty lines; run the batch as the parsed rows
    want = O.search_reads(oix, ch, rs, k)
    dch = torch.from_numpy(np.frombuffer(bytes(ch), np.uint8).copy()).cuda()
    drs = torch.from_numpy(rs.astype(np.int64)).cuda()
    plan = sg.ReadsPlan(drs.numel() - 1, dch.numel(), seg=seg)
    out = torch.full((max(want.size, 1),), 7, dtype=torch.int64, device="cuda")
    flags = {"persistent": kernel == "persistent", "lanes": kernel == "lane"}
    plan.search(ix, dch, drs, out, runs=True, **flags)
    np.testing.assert_array_equal(sg.to_u64(out)[: want.size], want)
    # the N-heavy read sets of the batched workload, batch by batch
    sh = starts.cpu().numpy()
    kc = np.maximum(np.diff(sh) - k + 1, 0)
    kcum = np.concatenate([[0], np.cumsum(kc)])
    want2 = O.search_reads(oix, chars.cpu().numpy(), sh.astype(np.uint64), k)
    batches = _batches(chars, starts, set_starts)
    plan2 = sg.ReadsPlan(max(b[1].numel() - 1 for b in batches), max(b[0].numel() for b in batches), seg=seg)
    out2 = torch.full((int(kcum[-1]),), 7, dtype=torch.int64, device="cuda")
    for bc, brel, b0, b1 in batches:
        plan2.search(ix, bc, brel, out2[int(kcum[b0]):int(kcum[b1])], runs=True, **flags)
    np.testing.assert_array_equal(sg.to_u64(out2), want2)
    # counted: the same answers
    ix.counters(reset=True)
    plan.search(ix, dch, drs, out, runs=True, count=True, **flags)
    assert ix.counters(reset=True)["kmers"] <= want.size
    np.testing.assert_array_equal(sg.to_u64(out)[: want.size], want)


def test_run_items_graph_replays():
    rng, s, ix, oix, chars, starts, set_starts = _setup(seed=6)
    k = 31
    sh = starts.cpu().numpy()
    kc = np.maximum(np.diff(sh) - k + 1, 0)
    kcum = np.concatenate([[0], np.cumsum(kc)])
    want = O.search_reads(oix, chars.cpu().numpy(), sh.astype(np.uint64), k)
    batches = _batches(chars, starts, set_starts)
    plan = sg.ReadsPlan(max(b[1].numel() - 1 for b in batches), max(b[0].numel() for b in batches), seg=64)
    out = torch.empty(int(kcum[-1]), dtype=torch.int64, device="cuda")

    def run_all(stream=None):
        for bc, brel, b0, b1 in batches:
            plan.search(ix, bc, brel, out[int(kcum[b0]):int(kcum[b1])], stream=stream, runs=True)

    run_all()
    torch.cuda.synchronize()
    np.testing.assert_array_equal(sg.to_u64(out), want)
    eager = out.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    gr = torch.cuda.CUDAGraph()
    with torch.cuda.graph(gr, stream=side):
        run_all(stream=side)
    torch.cuda.synchronize()
    for rep in range(4):
        out.fill_(-7)
        torch.cuda.synchronize()
        _ = (torch.arange(1000, device="cuda") * rep).sum().item()
        gr.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, eager), f"replay {rep}"


def test_run_items_ignored_where_unsupported():
    """k < 16 or N coded as A: the plan searches segments as before."""
    rng = np.random.default_rng(7)
    g = "".join(ACGT[rng.integers(0, 4, 20_000)])
    s = O.SBWT([g], 11)
    ix = sg.Index.from_rows(s.rows, s.n, 11, group=s.group, config=sg.Config(presearch_depth=6))
    rows = [g[p:p + 120] if i % 2 else g[p:p + 60] + "N" + g[p + 61:p + 120]
            for i, p in enumerate(rng.integers(0, 19_000, 50))]
    ch, rs = O.parse_lines("\n".join(rows).encode())
    oix = O.Index(s.rows, s.n)
    oix.presearch(6)
    want = O.search_reads(oix, ch, rs, 11)
    dch = torch.from_numpy(np.frombuffer(bytes(ch), np.uint8).copy()).cuda()
    drs = torch.from_numpy(rs.astype(np.int64)).cuda()
    plan = sg.ReadsPlan(drs.numel() - 1, dch.numel(), seg=32)
    out = torch.empty(want.size, dtype=torch.int64, device="cuda")
    plan.search(ix, dch, drs, out, runs=True)
    np.testing.assert_array_equal(sg.to_u64(out), want)
    plan.search(ix, dch, drs, out, runs=True, n_as_a=True)
    np.testing.assert_array_equal(sg.to_u64(out), O.search_reads(oix, ch, rs, 11, n_breaks=False))
