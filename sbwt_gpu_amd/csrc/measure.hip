// Random-access roof of the device: uniform-random line gathers over a large
// buffer. The search reads one rank chunk (16-64 bytes) of one random
// 128-byte line per rank, so its ceiling is the rate at which random lines
// can be fetched, not the streaming bandwidth. No reference counterpart.
#include "common.hpp"

namespace sbwt {

__device__ __forceinline__ u64 xorshift(u64 x) {
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return x;
}

// Each lane (or each group of `group` lanes sharing one line) reads `per`
// random lines; part (lane % group) of the line is its 16-byte load. Eight
// independent loads are kept in flight per lane.
__global__ __launch_bounds__(256) void k_gather(const uint4* __restrict__ buf, u64 nlines, u32 line_q, u32 group,
                                                u32 per, u64 seed, u64* __restrict__ sink) {
  const u64 tid = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  const u64 gid = tid / group;
  const u32 part = static_cast<u32>(tid % group);
  u64 x = seed ^ ((gid + 1) * 0x9E3779B97F4A7C15ull);
  x = xorshift(xorshift(x));
  uint4 acc = make_uint4(0, 0, 0, 0);
  for (u32 i = 0; i < per; i += 8) {
    uint4 v[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      x = xorshift(x);
      const u64 line = static_cast<u64>((static_cast<unsigned __int128>(x) * nlines) >> 64);
      v[j] = buf[line * line_q + part];
    }
#pragma unroll
    for (int j = 0; j < 8; j++) {
      acc.x ^= v[j].x;
      acc.y ^= v[j].y;
      acc.z ^= v[j].z;
      acc.w ^= v[j].w;
    }
  }
  sink[tid] = (static_cast<u64>(acc.x ^ acc.z) << 32) | (acc.y ^ acc.w);
}

}  // namespace sbwt

using namespace sbwt;

extern "C" int sbwt_gather_roof(const void* d_buf, uint64_t bytes, uint32_t line_bytes, uint32_t full_line,
                                uint32_t per_lane, uint64_t lanes, uint64_t seed, uint64_t* d_sink, void* stream) {
  clear_error();
  SBWT_REQUIRE(d_buf && d_sink, "device buffers must not be NULL");
  SBWT_REQUIRE(line_bytes == 16 || line_bytes == 32 || line_bytes == 64 || line_bytes == 128,
               "line_bytes must be 16, 32, 64 or 128");
  SBWT_REQUIRE(per_lane % 8 == 0 && per_lane > 0, "per_lane must be a positive multiple of 8");
  SBWT_REQUIRE((reinterpret_cast<uintptr_t>(d_buf) & 127) == 0, "buffer must be 128-byte aligned");
  const u64 nlines = bytes / line_bytes;
  SBWT_REQUIRE(nlines > 0, "buffer smaller than one line");
  const u32 line_q = line_bytes / 16;
  const u32 group = full_line ? line_q : 1;
  SBWT_REQUIRE(lanes % 256 == 0 && lanes % group == 0, "lanes must be a multiple of 256");
  k_gather<<<lanes / 256, 256, 0, as_stream(stream)>>>(static_cast<const uint4*>(d_buf), nlines, line_q, group,
                                                      per_lane, seed, d_sink);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}
