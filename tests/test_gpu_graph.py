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


def test_plan_rejects_oversized_batch():
    rng, s, ix, oix, chars, starts, set_starts = _setup(seed=2)
    plan = sg.ReadsPlan(3, 100)
    out = torch.empty(chars.numel(), dtype=torch.int64, device="cuda")
    with pytest.raises(sg.SbwtError, match="exceeds the plan"):
        plan.search(ix, chars, starts, out)


def test_autotune_keeps_answers_and_applies_best():
    rng, s, ix, oix, chars, starts, set_starts = _setup(seed=3)
    res = autotune.autotune(ix, chars, starts, layouts=(16, 32), depths=(8, 10), segs=(64, 256), reps=1)
    best = res["best"]
    assert len(res["trials"]) == 2 * 2 * 2 * 2
    assert ix.info["layout"] == best["layout"] and ix.info["presearch_depth"] == best["presearch_depth"]
    batch = sg.ReadBatch(chars, starts, 31, seg=best["segment_kmers"])
    want = O.search_reads(oix, chars.cpu().numpy(), starts.cpu().numpy().astype(np.uint64), 31)
    np.testing.assert_array_equal(sg.to_u64(ix.search_reads(batch, persistent=best["kernel"] == "persistent",
                                                            lanes=best["kernel"] == "lane")), want)
