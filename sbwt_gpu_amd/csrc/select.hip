// Select and k-mer enumeration over the rank-chunk layout (no reference
// counterpart: the reference has rank only).
//
// select_c(t) = the column of the t-th (0-based) 1-bit of row c. A sampled
// index holds, for every 4096th 1-bit, the chunk that contains it; the chunk
// of any t is then found by binary search over the chunks between two
// samples (each chunk's word 0 is its prefix count), and the bit inside the
// chunk by popcounts.
//
// Enumeration takes every node's label (labels.hip: a backward walk over
// the predecessor array), keeps the nodes with k real characters and
// compacts them in node order, in calls of at most 2^28 items.
#include <rocprim/device/device_select.hpp>

#include "index.hpp"

namespace sbwt {

constexpr u32 kSelectSample = 4096;

template <int CB>
__global__ void k_select_samples(DevIndex ix, u64 nchunks, u32* __restrict__ samples, u64 per_row) {
  const u64 ch = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  const u32 c = blockIdx.y;
  if (ch >= nchunks) return;
  Chunk<CB> q;
  load_chunk<CB>(row_of(ix, c), static_cast<u32>(ch), q);
  u32 ones = 0;
#pragma unroll
  for (u32 t = 0; t < Layout<CB>::PW; t++) ones += __popc(q.w[1 + t]);
  const u32 p = q.w[0];
  // the one multiple of kSelectSample in [p, p + ones), if any (ones < kSelectSample)
  const u32 m = (p + kSelectSample - 1) / kSelectSample * kSelectSample;
  if (ones && m < p + ones && m / kSelectSample < per_row) samples[c * per_row + m / kSelectSample] = static_cast<u32>(ch);
}

// Column of the t-th 1-bit of row c (t < ones of the row).
template <int CB>
__device__ __forceinline__ u32 select_dev(const DevIndex& ix, const u32* __restrict__ samples, u64 per_row,
                                          u64 nchunks, u32 c, u32 t) {
  constexpr u32 COLS = Layout<CB>::COLS;
  const uint4* row = row_of(ix, c);
  const u64 si = t / kSelectSample;
  const u32 ones = (c == 3 ? ix.C[4] : c_of(ix, c + 1)) - c_of(ix, c);
  u32 lo = samples[c * per_row + si];
  // samples exist for 0, S, 2S, ... below `ones`
  u32 hi = ((si + 1) * kSelectSample < ones) ? samples[c * per_row + si + 1] : static_cast<u32>(nchunks - 1);
  // last chunk lo..hi with prefix count <= t
  while (lo < hi) {
    const u32 mid = (lo + hi + 1) >> 1;
    const u32 p = reinterpret_cast<const u32*>(row + static_cast<u64>(mid) * Layout<CB>::Q)[0];
    if (p <= t) lo = mid; else hi = mid - 1;
  }
  Chunk<CB> q;
  load_chunk<CB>(row, lo, q);
  u32 need = t - q.w[0];
  u32 col = lo * COLS;
#pragma unroll
  for (u32 w = 0; w < Layout<CB>::PW; w++) {
    const u32 x = q.w[1 + w];
    const u32 pc = __popc(x);
    if (need < pc) {
      // the need-th set bit of x (LSB-first)
      u32 y = x;
      for (u32 s = 0; s < need; s++) y &= y - 1;
      return col + w * 32 + (__ffs(y) - 1);
    }
    need -= pc;
  }
  return col + COLS;  // not reached for t < ones
}

template <int CB>
__global__ void k_select(DevIndex ix, const u32* __restrict__ samples, u64 per_row, u64 nchunks, u32 c,
                         const u64* __restrict__ in, u64 count, u64* __restrict__ out) {
  const u64 q = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (q >= count) return;
  const u64 t = in[q];
  const u32 ones = (c == 3 ? ix.C[4] : c_of(ix, c + 1)) - c_of(ix, c);
  out[q] = t < ones ? static_cast<u64>(select_dev<CB>(ix, samples, per_row, nchunks, c, static_cast<u32>(t))) : kNone;
}

static int ensure_samples(sbwt_index* ix, hipStream_t s) {
  if (ix->d_select && ix->select_layout == ix->layout) return SBWT_OK;
  if (ix->d_select) {
    SBWT_HIP(hipFree(ix->d_select));
    ix->d_select = nullptr;
  }
  const u32 cols = cols_for_layout(ix->layout);
  const u64 nchunks = chunks_for(ix->n, cols);
  const u64 per_row = ix->n / kSelectSample + 2;
  SBWT_HIP(hipMalloc(&ix->d_select, 4 * per_row * sizeof(u32)));
  SBWT_HIP(hipMemsetAsync(ix->d_select, 0, 4 * per_row * sizeof(u32), s));
  dim3 grid(static_cast<u32>((nchunks + 255) / 256), 4);
  DevIndex v = dev_view(ix);
  switch (ix->layout) {
    case 16: k_select_samples<16><<<grid, 256, 0, s>>>(v, nchunks, ix->d_select, per_row); break;
    case 32: k_select_samples<32><<<grid, 256, 0, s>>>(v, nchunks, ix->d_select, per_row); break;
    default: k_select_samples<64><<<grid, 256, 0, s>>>(v, nchunks, ix->d_select, per_row); break;
  }
  SBWT_HIP(hipGetLastError());
  ix->select_layout = ix->layout;
  ix->select_per_row = per_row;
  return SBWT_OK;
}

}  // namespace sbwt

using namespace sbwt;

extern "C" {

int sbwt_select(sbwt_index* ix, uint32_t c, const uint64_t* d_ranks, uint64_t count, uint64_t* d_out, void* stream) {
  clear_error();
  SBWT_REQUIRE(ix && c < 4, "index must not be NULL and c must be 0..3");
  if (count == 0) return SBWT_OK;
  SBWT_REQUIRE(d_ranks && d_out, "device buffers must not be NULL");
  hipStream_t s = as_stream(stream);
  SBWT_TRY(ensure_samples(ix, s));
  const u64 nchunks = chunks_for(ix->n, cols_for_layout(ix->layout));
  DevIndex v = dev_view(ix);
  const u64 grid = (count + 255) / 256;
  switch (ix->layout) {
    case 16: k_select<16><<<grid, 256, 0, s>>>(v, ix->d_select, ix->select_per_row, nchunks, c, d_ranks, count, d_out); break;
    case 32: k_select<32><<<grid, 256, 0, s>>>(v, ix->d_select, ix->select_per_row, nchunks, c, d_ranks, count, d_out); break;
    default: k_select<64><<<grid, 256, 0, s>>>(v, ix->d_select, ix->select_per_row, nchunks, c, d_ranks, count, d_out); break;
  }
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

int sbwt_node_labels(sbwt_index* ix, uint64_t* d_labels, uint8_t* d_full, void* stream) {
  clear_error();
  SBWT_REQUIRE(ix && d_labels && d_full, "arguments must not be NULL");
  return node_labels(ix, d_labels, d_full, as_stream(stream));
}

int sbwt_enumerate_kmers(sbwt_index* ix, uint64_t* d_kmers, uint64_t* count_out, void* stream) {
  clear_error();
  SBWT_REQUIRE(ix && d_kmers && count_out, "arguments must not be NULL");
  hipStream_t s = as_stream(stream);
  constexpr u64 kPiece = 1ull << 28;
  u64* labels = nullptr;
  u8* full = nullptr;
  void* tmp = nullptr;
  u64* d_cnt = nullptr;
  int st = SBWT_OK;
  u64 done = 0;
  do {
    if (hipMalloc(&labels, ix->n * 8) != hipSuccess || hipMalloc(&full, ix->n) != hipSuccess ||
        hipMalloc(&d_cnt, 8) != hipSuccess) {
      st = fail(SBWT_ERR_HIP, "out of device memory for %llu labels", (unsigned long long)ix->n);
      break;
    }
    if ((st = node_labels(ix, labels, full, s)) != SBWT_OK) break;
    size_t tb = 0;
    const u64 first = ix->n < kPiece ? ix->n : kPiece;
    if (rocprim::select(nullptr, tb, labels, full, d_kmers, d_cnt, first, s) != hipSuccess ||
        hipMalloc(&tmp, tb ? tb : 1) != hipSuccess) {
      st = fail(SBWT_ERR_HIP, "k-mer compaction setup failed");
      break;
    }
    for (u64 off = 0; off < ix->n && st == SBWT_OK; off += kPiece) {
      const u64 len = ix->n - off < kPiece ? ix->n - off : kPiece;
      size_t need = 0;
      u64 got = 0;
      if (rocprim::select(nullptr, need, labels + off, full + off, d_kmers + done, d_cnt, len, s) != hipSuccess ||
          need > tb ||
          rocprim::select(tmp, need, labels + off, full + off, d_kmers + done, d_cnt, len, s) != hipSuccess ||
          hipMemcpyAsync(&got, d_cnt, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        st = fail(SBWT_ERR_HIP, "k-mer compaction failed");
      done += got;
    }
  } while (false);
  if (st == SBWT_OK) *count_out = done;
  (void)hipStreamSynchronize(s);
  if (labels) (void)hipFree(labels);
  if (full) (void)hipFree(full);
  if (tmp) (void)hipFree(tmp);
  if (d_cnt) (void)hipFree(d_cnt);
  return st;
}

}  // extern "C"
