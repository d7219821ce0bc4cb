// Run items of the read search (no reference counterpart: the reference
// searches every k-mer, N or not, include/gpu/search_kernel.hpp:61-104,
// after mapping N to A, include/gpu/create_kmer_string.hpp:20).
//
// A row's k-mer windows that hold no non-ACGT character come in maximal
// runs; between two runs lies at least one non-ACGT character, and a run
// needs k characters, so a batch of C characters and R rows has at most
// R + C / (k + 1) runs. This pass finds the runs with one wave per row: each
// lane tests one window of a 64-window chunk, a ballot gives the chunk's
// valid windows, and run ends, starts and the run carried across chunks
// follow from bit operations on that mask (a carried run that ends at a
// chunk's last window is emitted by lane 0 of the next chunk). Two passes,
// a count per row and then the items at the scanned row offsets, place the
// items in row order without atomics (one global counter for all waves'
// items stalled them: 32 ms for a batch the search takes 5 ms for). The pass answers every window
// that holds a non-ACGT character itself (~0, 64 consecutive answers per
// store instruction), and emits one work item {row, first k-mer, k-mers}
// per run, cut into pieces of at most `seg` k-mers. The search kernels then
// walk only N-free runs: no non-ACGT checks, and every lane of a wave
// starts its item the same way (a search, then streaming steps).
#include "index.hpp"

namespace sbwt {

namespace {

constexpr u32 kNoRun = 0xFFFFFFFFu;

__device__ __forceinline__ u64 mask64_at(const u64* __restrict__ w, u64 pos) {
  return (w[pos >> 6] << (pos & 63)) | ((w[(pos >> 6) + 1] >> (63 - (pos & 63))) >> 1);
}

// Pass 1 (WRITE = false): the number of items of every row, into
// row_items[row]. Pass 2 (WRITE = true): the items, at row_base[row] (the
// exclusive scan of pass 1) in row order, and ~0 for every N window. No
// atomics: a wave walks its rows and keeps the running count itself.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_run_items(const u64* __restrict__ nmask, const u64* __restrict__ row_starts,
                                                   const u64* __restrict__ offsets, u64 rows, u32 k, u32 seg,
                                                   u64* __restrict__ out, uint4* __restrict__ runs, u64 cap,
                                                   u64* __restrict__ row_items, const u64* __restrict__ row_base) {
  const u32 lane = threadIdx.x & 63u;
  const u64 wave = (blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x) >> 6;
  const u64 nwaves = (gridDim.x * static_cast<u64>(blockDim.x)) >> 6;
  for (u64 row = wave; row < rows; row += nwaves) {
    const u64 o0 = offsets[row];
    const u32 rowk = static_cast<u32>(offsets[row + 1] - o0);
    const u64 cb = row_starts[row];
    u64 at = WRITE ? row_base[row] : 0;  // the row's next item slot (pass 1: its item count)
    u32 open = kNoRun;  // first k-mer of the run that reaches the previous chunk's last window
    for (u32 base = 0; base < rowk; base += 64) {
      const u32 j = base + lane;
      bool valid = false;
      if (j < rowk) {
        valid = (mask64_at(nmask, cb + j) >> (64 - k)) == 0;
        if (WRITE && !valid) out[o0 + j] = kNone;
      }
      const u64 v = __ballot(valid);
      const bool last = base + 64 >= rowk;
      // a valid window followed by an invalid one ends a run; the chunk's last
      // window ends one only at the row's end (windows past it are invalid)
      u64 ends = v & ~(v >> 1);
      if (!last) ends &= ~(1ull << 63);
      u32 first = 0, len = 0, pieces = 0;
      if ((ends >> lane) & 1ull) {
        const u64 below = lane ? (~v & ((1ull << lane) - 1ull)) : 0ull;  // invalid windows before this one
        first = below ? base + 64u - static_cast<u32>(__clzll(below)) : (open != kNoRun ? open : base);
        len = j + 1 - first;
      } else if (lane == 0 && open != kNoRun && !(v & 1ull)) {
        // the run carried from the previous chunk ended at its last window
        first = open;
        len = base - open;
      }
      pieces = (len + seg - 1) / seg;
      // exclusive prefix of the pieces over the wave
      u32 incl = pieces;
#pragma unroll
      for (u32 d = 1; d < 64; d <<= 1) {
        const u32 y = __shfl_up(incl, d, 64);
        if (lane >= d) incl += y;
      }
      if (WRITE) {
        for (u32 q = 0; q < pieces; q++) {
          const u64 slot = at + (incl - pieces) + q;
          const u32 from = first + q * seg;
          const u32 n = len - q * seg < seg ? len - q * seg : seg;
          if (slot < cap) runs[slot] = make_uint4(static_cast<u32>(row), from, n, 0u);
        }
      }
      at += __shfl(incl, 63, 64);
      // the run holding the chunk's last window goes on into the next chunk
      if (!last && (v >> 63)) {
        const u64 inv = ~v;
        open = inv ? base + 64u - static_cast<u32>(__clzll(inv)) : (open != kNoRun ? open : base);
      } else {
        open = kNoRun;
      }
    }
    if (!WRITE && lane == 0) row_items[row] = at;
  }
}

}  // namespace

u64 run_items_bound(u64 rows, u64 chars, u32 seg) {
  // runs: rows + chars / (k + 1) for k >= kRunItemsMinK; pieces: one more per seg k-mers
  return rows + chars / (kRunItemsMinK + 1) + chars / (seg ? seg : 256) + 64;
}

static u32 run_grid(u64 rows) {
  const u64 want = (rows + 3) / 4;  // four waves (rows) per block
  return static_cast<u32>(want < 65536 ? (want ? want : 1) : 65536);
}

int run_item_counts(const u64* nmask, const u64* row_starts, const u64* offsets, u64 rows, u32 k, u32 seg,
                    u64* row_items, hipStream_t s) {
  if (k < kRunItemsMinK) return fail(SBWT_ERR_UNSUPPORTED, "run items need k >= %u (k = %u)", kRunItemsMinK, k);
  if (rows >= 0xFFFFFFFFull) return fail(SBWT_ERR_UNSUPPORTED, "run items hold u32 row numbers");
  k_run_items<false><<<run_grid(rows), 256, 0, s>>>(nmask, row_starts, offsets, rows, k, seg, nullptr, nullptr, 0,
                                                    row_items, nullptr);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

int run_item_write(const u64* nmask, const u64* row_starts, const u64* offsets, u64 rows, u32 k, u32 seg, u64* out,
                   uint4* runs, u64 cap, const u64* row_base, hipStream_t s) {
  k_run_items<true><<<run_grid(rows), 256, 0, s>>>(nmask, row_starts, offsets, rows, k, seg, out, runs, cap,
                                                   nullptr, row_base);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

}  // namespace sbwt
