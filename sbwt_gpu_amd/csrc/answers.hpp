// Answer stores of the read-search kernels.
//
// A lane answers the k-mers of its own rows in order, 8 bytes at a time. With
// the index reads streaming through the L2, a partially written output line
// is evicted before the lane comes back to it, and is written back piecemeal
// (measured: 3.3 GB of HBM writes per 0.96 GB of answers). So each lane
// collects its answers per aligned group of four output slots (32 bytes) in
// LDS, slot-major (slot s of thread t at obuf[s * kAnswerStride + t], so a
// wave's accesses never conflict), and a full group goes out as two 16-byte
// stores; partial groups (row and item edges) go out slot by slot.
#pragma once
#include "common.hpp"

namespace sbwt {

constexpr u32 kAnswerStride = 256;  // threads per block of the kernels that use it

struct AnswerBuffer {
  u64* obuf;  // this thread's slot 0 in LDS
  u64 group = ~0ull;
  u32 mask = 0;

  __device__ __forceinline__ void flush(u64* __restrict__ out, bool vec) {
    if (!mask) return;
    u64* o = out + group * 4;
    const u64 b0 = obuf[0], b1 = obuf[kAnswerStride], b2 = obuf[2 * kAnswerStride],
              b3 = obuf[3 * kAnswerStride];
    if (mask == 15 && vec) {
      reinterpret_cast<uint4*>(o)[0] = make_uint4(static_cast<u32>(b0), static_cast<u32>(b0 >> 32),
                                                  static_cast<u32>(b1), static_cast<u32>(b1 >> 32));
      reinterpret_cast<uint4*>(o)[1] = make_uint4(static_cast<u32>(b2), static_cast<u32>(b2 >> 32),
                                                  static_cast<u32>(b3), static_cast<u32>(b3 >> 32));
    } else {
      if (mask & 1) o[0] = b0;
      if (mask & 2) o[1] = b1;
      if (mask & 4) o[2] = b2;
      if (mask & 8) o[3] = b3;
    }
    mask = 0;
  }

  // out[q] = v; `last` flushes (the lane's item ends at q). vec: out is
  // 16-byte aligned, so a group of four is two aligned 16-byte stores.
  __device__ __forceinline__ void put(u64* __restrict__ out, bool vec, u64 q, u64 v, bool last) {
    const u64 g = q >> 2;
    const u32 sl = static_cast<u32>(q) & 3u;
    if (g != group) {
      flush(out, vec);
      group = g;
    }
    obuf[sl * kAnswerStride] = v;
    mask |= 1u << sl;
    if (mask == 15 || last) flush(out, vec);
  }
};

}  // namespace sbwt
