// Result text on the device: the bytes of sbwt_format_results (the
// reference's ResultsPersister::save_results_to_file,
// include/epicseq/results_persister.hpp:82-96: each answer as a signed
// decimal, ~0 printing as -1, then a space; a newline after each row),
// written by the GPU so that the host only moves bytes.
//
// Two passes: a prefix sum of the answers' printed lengths gives every
// answer's place (row r's text starts r newlines later), then one wave per
// row writes its answers' digits and the row's newline.
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "common.hpp"

namespace sbwt {

namespace {

__constant__ u64 kPow10[20] = {1ull,
                               10ull,
                               100ull,
                               1000ull,
                               10000ull,
                               100000ull,
                               1000000ull,
                               10000000ull,
                               100000000ull,
                               1000000000ull,
                               10000000000ull,
                               100000000000ull,
                               1000000000000ull,
                               10000000000000ull,
                               100000000000000ull,
                               1000000000000000ull,
                               10000000000000000ull,
                               100000000000000000ull,
                               1000000000000000000ull,
                               10000000000000000000ull};

// decimal digits of x (x >= 0)
__device__ __forceinline__ u32 digits(u64 x) {
  u32 n = 1;
#pragma unroll
  for (int i = 1; i < 20; i++) n += x >= kPow10[i] ? 1u : 0u;
  return n;
}

// the answer as printed: its magnitude and sign (int64 of the bits)
__device__ __forceinline__ u64 magnitude(u64 v, bool& neg) {
  neg = v >> 63;
  return neg ? ~v + 1 : v;
}

struct PrintedLength {
  const u64* res;
  u64 count;
  __device__ u64 operator()(u64 q) const {
    if (q >= count) return 0;
    bool neg;
    const u64 m = magnitude(res[q], neg);
    return digits(m) + (neg ? 1u : 0u) + 1u;  // digits, sign, space
  }
};

__device__ __forceinline__ void print(char* __restrict__ at, u64 v) {
  bool neg;
  u64 m = magnitude(v, neg);
  const u32 nd = digits(m);
  if (neg) *at++ = '-';
  at[nd] = ' ';
  if (m >> 32) {
    for (u32 i = nd; i-- > 0; m /= 10) at[i] = static_cast<char>('0' + m % 10);
  } else {
    u32 x = static_cast<u32>(m);  // 32-bit division: the common case (node ids)
    for (u32 i = nd; i-- > 0; x /= 10u) at[i] = static_cast<char>('0' + x % 10u);
  }
}

__global__ __launch_bounds__(64) void k_text_write(const u64* __restrict__ res, const u64* __restrict__ off,
                                                   u64 rows, const u64* __restrict__ pos, u64 count, u64 cap,
                                                   char* __restrict__ text, u64* __restrict__ size) {
  const u64 total = pos[count] + rows;
  if (blockIdx.x == 0 && threadIdx.x == 0) *size = total;
  if (total > cap) return;  // too small: nothing written, *size tells how much is needed
  for (u64 r = blockIdx.x; r < rows; r += gridDim.x) {
    const u64 q0 = off[r], q1 = off[r + 1];
    for (u64 q = q0 + threadIdx.x; q < q1; q += 64) print(text + pos[q] + r, res[q]);
    if (threadIdx.x == 0) text[pos[q1] + r] = '\n';
  }
}

size_t scan_temp_bytes(u64 count) {
  auto it = rocprim::make_transform_iterator(rocprim::make_counting_iterator<u64>(0), PrintedLength{nullptr, 0});
  size_t b = 0;
  if (rocprim::exclusive_scan(nullptr, b, it, static_cast<u64*>(nullptr), u64(0), count + 1,
                              rocprim::plus<u64>()) != hipSuccess)
    return 0;
  return b;
}

inline u64 align256(u64 x) { return (x + 255) & ~255ull; }

}  // namespace

}  // namespace sbwt

using namespace sbwt;

extern "C" {

uint64_t sbwt_format_text_scratch(uint64_t count) {
  return align256((count + 1) * sizeof(u64)) + align256(scan_temp_bytes(count)) + 256;
}

int sbwt_format_text(const uint64_t* d_results, const uint64_t* d_offsets, uint64_t rows, uint64_t count,
                     char* d_text, uint64_t cap, uint64_t* d_size, void* d_scratch, uint64_t scratch_bytes,
                     void* stream) {
  clear_error();
  SBWT_REQUIRE(d_offsets && d_size && (d_results || count == 0) && (d_text || cap == 0),
               "device buffers must not be NULL");
  SBWT_REQUIRE(d_scratch && scratch_bytes >= sbwt_format_text_scratch(count),
               "scratch of %llu bytes: sbwt_format_text_scratch(%llu) = %llu needed",
               (unsigned long long)scratch_bytes, (unsigned long long)count,
               (unsigned long long)sbwt_format_text_scratch(count));
  hipStream_t s = as_stream(stream);
  u64* pos = reinterpret_cast<u64*>((reinterpret_cast<uintptr_t>(d_scratch) + 255) & ~uintptr_t(255));
  void* tmp = reinterpret_cast<char*>(pos) + align256((count + 1) * sizeof(u64));
  size_t tb = scan_temp_bytes(count);
  auto it = rocprim::make_transform_iterator(rocprim::make_counting_iterator<u64>(0), PrintedLength{d_results, count});
  SBWT_HIP(rocprim::exclusive_scan(tmp, tb, it, pos, u64(0), count + 1, rocprim::plus<u64>(), s));
  const u64 grid = rows == 0 ? 1 : (rows < (1u << 20) ? rows : (1u << 20));
  k_text_write<<<grid, 64, 0, s>>>(d_results, d_offsets, rows, pos, count, cap, d_text, d_size);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

}  // extern "C"
