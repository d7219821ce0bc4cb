// Run items of the read search (no reference counterpart: the reference
// searches every k-mer, N or not, include/gpu/search_kernel.hpp:61-104,
// after mapping N to A, include/gpu/create_kmer_string.hpp:20).
//
// A row's k-mer windows that hold no non-ACGT character come in maximal
// runs: a stretch of L characters between two non-ACGT characters (or the
// row's ends) holds L - k + 1 of them when L >= k. A run needs k characters
// and the next one starts after a non-ACGT character, so a batch of C
// characters and R rows has at most R + C / (k + 1) runs.
//
// An item {row, a, b, c} covers the row's windows [a, c): [a, b) hold a
// non-ACGT character and answer ~0, [b, c) is an N-free run of k-mers (a
// piece of at most seg of them). The gap before a row's first run and the
// gaps between runs lead the items that follow them; a row's trailing gap,
// and a row with no valid window, get an item with an empty run. So the
// items tile every row, and the search kernels never test a character for
// N: a lane skips its item's gap and then walks the run.
//
// The ~0 answers come from a streaming fill of the whole output by the same
// kernel, before the search (16-byte stores, which overlap the row walks).
// Written by the search kernels lane by lane they took the uniform-step
// kernel from 2.35 to 4.58 ms per batch of 5%-N reads.
//
// One thread per row finds the runs from the row's non-ACGT mask, four words
// at a time (loaded together, then their set bits consumed in registers). A wave's 64 rows take their item
// slots with one atomic (about 10^4 per batch of 650K rows), so items of 64
// consecutive rows are contiguous and in row order. Earlier versions: slots
// taken per 64-window chunk (2 million atomics on one counter per batch,
// 32 ms for a batch the search takes 5 ms for); two passes with a scan of
// the rows' counts between them (0.58 ms, with one load per non-ACGT
// character).
#include "index.hpp"

namespace sbwt {

namespace {

// out[0 .. n) = ~0 in 16-byte stores (the first answer alone when out is not
// 16-byte aligned): thread tid of `threads` takes every threads-th pair.
__device__ __forceinline__ void fill_none_part(u64* __restrict__ out, u64 n, u64 tid, u64 threads) {
  if (n == 0) return;
  const u64 head = (reinterpret_cast<uintptr_t>(out) & 15) ? 1 : 0;
  if (tid == 0 && head) out[0] = kNone;
  uint4* o = reinterpret_cast<uint4*>(out + head);
  const u64 pairs = (n - head) / 2;
  const uint4 v = make_uint4(~0u, ~0u, ~0u, ~0u);
  for (u64 i = tid; i < pairs; i += threads) o[i] = v;
  if (tid == 0 && ((n - head) & 1)) out[n - 1] = kNone;
}

// The items of a row, or only their number (runs == nullptr), from slot at.
__device__ __forceinline__ u64 row_items(const u64* __restrict__ nmask, u64 cb, u32 rowk, u32 k, u32 seg, u32 row,
                                         uint4* __restrict__ runs, u64 at, u64 cap) {
  if (!rowk) return 0;
  const u64 end = cb + rowk + k - 1;  // the row's characters
  u64 n = 0;
  u32 gap = 0;  // first window not yet covered by an item
  u64 c = cb;   // first character of the current N-free stretch
  // the N-free stretch [c, q) ends at q (a non-ACGT character or the row's end)
  auto stretch_end = [&](u64 q) {
    if (q - c >= k) {  // windows [c, q - k] are valid
      const u32 b = static_cast<u32>(c - cb), e = static_cast<u32>(q - cb) - k + 1;
      for (u32 x = b; x < e; x += seg) {
        const u32 xe = e - x < seg ? e : x + seg;
        if (runs && at + n < cap) runs[at + n] = make_uint4(row, x == b ? gap : x, x, xe);
        n++;
      }
      gap = e;
    }
    c = q + 1;
  };
  const u64 w0 = cb >> 6, w1 = (end - 1) >> 6;
  for (u64 wb = w0; wb <= w1; wb += 4) {  // four mask words loaded before any is used
    u64 xs[4];
#pragma unroll
    for (u32 t = 0; t < 4; t++) xs[t] = wb + t <= w1 ? nmask[wb + t] : 0ull;
#pragma unroll
    for (u32 t = 0; t < 4; t++) {
      const u64 w = wb + t;
      u64 x = xs[t];
      if (w == w0) x &= ~0ull >> (cb & 63);                    // characters before the row
      if (w == w1 && (end & 63)) x &= ~(~0ull >> (end & 63));  // characters after it
      while (x) {
        const u32 b = static_cast<u32>(__clzll(x));
        stretch_end((w << 6) + b);
        x &= ~(1ull << (63 - b));
      }
    }
  }
  stretch_end(end);
  if (gap < rowk) {  // the trailing gap (or a row without valid windows)
    if (runs && at + n < cap) runs[at + n] = make_uint4(row, gap, rowk, rowk);
    n++;
  }
  return n;
}

// One thread per row: fill a slice of the output with ~0, count the row's
// items, take slots for the wave's 64 rows with one atomic, then walk the
// row again (its mask words are cached) and write the items.
__global__ __launch_bounds__(256) void k_run_items(const u64* __restrict__ nmask, const u64* __restrict__ row_starts,
                                                   const u64* __restrict__ offsets, u64 rows, u32 k, u32 seg,
                                                   uint4* __restrict__ runs, u64 cap,
                                                   unsigned long long* __restrict__ count, u64* __restrict__ out) {
  const u64 row = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  const u32 lane = threadIdx.x & 63u;
  fill_none_part(out, offsets[rows], row, gridDim.x * static_cast<u64>(blockDim.x));
  u32 rowk = 0;
  u64 cb = 0;
  if (row < rows) {
    rowk = static_cast<u32>(offsets[row + 1] - offsets[row]);
    cb = row_starts[row];
  }
  const u64 n = row < rows ? row_items(nmask, cb, rowk, k, seg, static_cast<u32>(row), nullptr, 0, 0) : 0;
  u64 incl = n;
#pragma unroll
  for (u32 d = 1; d < 64; d <<= 1) {
    const u64 y = __shfl_up(incl, d, 64);
    if (lane >= d) incl += y;
  }
  const u64 total = __shfl(incl, 63, 64);
  unsigned long long at = 0;
  if (lane == 0 && total) at = atomicAdd(count, static_cast<unsigned long long>(total));
  at = __shfl(at, 0, 64);
  if (n) (void)row_items(nmask, cb, rowk, k, seg, static_cast<u32>(row), runs, at + incl - n, cap);
}

__global__ void k_zero_count(unsigned long long* c) { *c = 0ull; }

}  // namespace

u64 run_items_bound(u64 rows, u64 chars, u32 seg) {
  // runs: rows + chars / (k + 1) for k >= kRunItemsMinK; pieces: one more per
  // seg k-mers; one trailing-gap item per row
  return 2 * rows + chars / (kRunItemsMinK + 1) + chars / (seg ? seg : 256) + 64;
}

int run_items(const u64* nmask, const u64* row_starts, const u64* offsets, u64 rows, u32 k, u32 seg, uint4* runs,
              u64 cap, unsigned long long* count, u64* out, hipStream_t s) {
  if (k < kRunItemsMinK) return fail(SBWT_ERR_UNSUPPORTED, "run items need k >= %u (k = %u)", kRunItemsMinK, k);
  if (rows >= 0xFFFFFFFFull) return fail(SBWT_ERR_UNSUPPORTED, "run items hold u32 row numbers");
  k_zero_count<<<1, 1, 0, s>>>(count);
  k_run_items<<<static_cast<u32>(rows ? (rows + 255) / 256 : 1), 256, 0, s>>>(nmask, row_starts, offsets, rows, k,
                                                                             seg, runs, cap, count, out);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

}  // namespace sbwt
