"""The allocation-free read search (sbwt_reads_search_static) against the
read-plan search it is a variant of (sbwt_reads_plan_search), through the
C ABI: two calls in a row, the second with a smaller batch, must each give
the plan's answers and k-mer offsets bit for bit; batches beyond the static
scratch are refused before anything runs."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("needs a GPU", allow_module_level=True)

import sbwt_gpu_amd as sg  # noqa: E402
from sbwt_gpu_amd import _capi  # noqa: E402
from sbwt_gpu_amd.api import _flags, _ptr, _stream  # noqa: E402

ACGT = np.array(list("ACGT"))


def _batch(rng, g, n, k):
    rows = []
    for i in range(n):
        ln = int(rng.integers(k - 3, 400))
        p = int(rng.integers(0, len(g) - ln))
        r = np.array(list(g[p:p + ln]))
        r[rng.random(r.size) < (0.05 if i % 3 else 0.0)] = "N"
        if i % 7 == 0 and r.size > 40:
            r[20] = "ACGT"[("ACGT".index(r[20]) + 1) % 4] if r[20] in "ACGT" else "A"
        rows.append("".join(r))
    ch, rs = O.parse_lines("\n".join(rows).encode())
    dch = torch.from_numpy(np.frombuffer(bytes(ch), np.uint8).copy()).cuda()
    drs = torch.from_numpy(rs.astype(np.int64)).cuda()
    total = int(np.maximum(np.diff(rs.astype(np.int64)) - k + 1, 0).sum())
    return dch, drs, total


def _static(L, ix, dch, drs, seg, out, off, flags):
    return L.sbwt_reads_search_static(ix._h, _ptr(dch), dch.numel(), _ptr(drs), drs.numel() - 1, seg, _ptr(out),
                                      _ptr(off), flags, _stream())


@pytest.mark.parametrize("kw", [{}, {"runs": True}, {"lanes": True}, {"persistent": True, "runs": True},
                                {"independent": True}])
def test_static_search_equals_plan_search_twice(kw):
    rng = np.random.default_rng(91)
    g = "".join(ACGT[rng.integers(0, 4, 120_000)])
    k = 31
    s = O.SBWT([g], k)
    ix = sg.Index.from_rows(s.rows, s.n, k, group=s.group, config=sg.Config(presearch_depth=10))
    L = _capi.load()
    flags = _flags(**kw)
    for n, seg in ((3000, 64), (500, 256)):  # the second call smaller than the first
        dch, drs, total = _batch(rng, g, n, k)
        plan = sg.ReadsPlan(drs.numel() - 1, dch.numel(), seg=seg)
        want = torch.full((total + 1,), 5, dtype=torch.int64, device="cuda")
        woff = torch.empty(drs.numel(), dtype=torch.int64, device="cuda")
        plan.search(ix, dch, drs, want[:total], offsets=woff, **kw)
        got = torch.full((total + 1,), 9, dtype=torch.int64, device="cuda")
        goff = torch.full((drs.numel(),), -3, dtype=torch.int64, device="cuda")
        _capi.check(_static(L, ix, dch, drs, seg, got, goff, flags))
        torch.cuda.synchronize()
        assert torch.equal(got[:total], want[:total]), (n, seg, kw)
        assert int(got[total]) == 9  # nothing written past the answers
        assert torch.equal(goff, woff)
        # and the oracle agrees
        ch = dch.cpu().numpy()
        rs = drs.cpu().numpy().astype(np.uint64)
        oix = O.Index(s.rows, s.n)
        oix.presearch(10)
        np.testing.assert_array_equal(sg.to_u64(got[:total]), O.search_reads(oix, ch, rs, k))
        plan.close()


def test_static_search_refuses_oversized_batches():
    rng = np.random.default_rng(92)
    g = "".join(ACGT[rng.integers(0, 4, 5000)])
    s = O.SBWT([g], 15)
    ix = sg.Index.from_rows(s.rows, s.n, 15, group=s.group, config=sg.Config(presearch_depth=6))
    L = _capi.load()
    buf = torch.zeros(16, dtype=torch.int64, device="cuda")
    p = _ptr(buf)
    too_many_chars = L.sbwt_reads_search_static(ix._h, p, 16_777_217, p, 1, 64, p, None, 0, _stream())
    assert too_many_chars == 1 and b"exceeds the static scratch" in L.sbwt_last_error()
    too_many_rows = L.sbwt_reads_search_static(ix._h, p, 100, p, 262_145, 64, p, None, 0, _stream())
    assert too_many_rows == 1 and b"exceeds the static scratch" in L.sbwt_last_error()
    bad_seg = L.sbwt_reads_search_static(ix._h, p, 100, p, 1, 8, p, None, 0, _stream())
    assert bad_seg == 1 and b"seg" in L.sbwt_last_error()
