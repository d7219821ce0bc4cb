// Answer stores and prefilter flags of the read-search kernels.
//
// A lane answers the k-mers of its own rows in order, 8 bytes at a time. With
// the index reads streaming through the L2, a partially written output line
// is evicted before the lane comes back to it, and is written back piecemeal
// (measured: 3.3 GB of HBM writes per 0.96 GB of answers). So each lane
// collects its answers per aligned group of S output slots (S * 8 bytes) in
// LDS, slot-major (slot s of thread t at obuf[s * kAnswerStride + t], so a
// wave's accesses never conflict), and a full group goes out as S / 2
// 16-byte stores; partial groups (row and item edges) go out slot by slot.
// Runs of one value (k-mers answered ~0 without a search) skip the buffer for
// their whole groups.
//
// A run's part inside its first group and its part after the last whole
// group take one predicated LDS write per slot, not a loop iteration per
// answer: on reads with 5% N (runs of a few dozen k-mers) that took the
// uniform-step kernel from 7.18 to 6.13 ms per batch and the per-lane kernel
// from 5.86 to 5.65 ms.
//
// S = 4 (32-byte groups) by default. The uniform-step kernel on 16-byte
// chunks uses S = 8 (64 bytes, 16 KB of LDS per 256-thread block): the
// answers' writes still cost that kernel about 0.2 ms per 1M-read batch
// (3.18 ms with the same stores sent to an L2-resident region), and 64-byte
// groups take back some of it (3.39 -> 3.36 ms; 2.59 -> 2.50 ms on a 1 Mbp
// index). Non-temporal stores were slower (3.45 ms at S = 4, 3.53 at S = 8).
#pragma once
#include "common.hpp"

namespace sbwt {

constexpr u32 kAnswerStride = 256;  // threads per block of the kernels that use it

template <u32 S>
struct AnswerBufferT {
  static_assert(S >= 2 && S <= 16 && (S & (S - 1)) == 0, "a power of two, whole 16-byte stores");
  u64* obuf;  // this thread's slot 0 in LDS
  u32 mask = 0;

  // Writes the buffered answers of the group holding out[q].
  __device__ __forceinline__ void flush(u64* __restrict__ out, bool vec, u64 q) {
    if (!mask) return;
    u64* o = out + (q & ~static_cast<u64>(S - 1));
    if (mask == (1u << S) - 1 && vec) {
#pragma unroll
      for (u32 h = 0; h < S / 2; h++) {
        const u64 x = obuf[(2 * h) * kAnswerStride], y = obuf[(2 * h + 1) * kAnswerStride];
        reinterpret_cast<uint4*>(o)[h] = make_uint4(static_cast<u32>(x), static_cast<u32>(x >> 32),
                                                    static_cast<u32>(y), static_cast<u32>(y >> 32));
      }
    } else {
#pragma unroll
      for (u32 sl = 0; sl < S; sl++)
        if (mask & (1u << sl)) o[sl] = obuf[sl * kAnswerStride];
    }
    mask = 0;
  }

  // out[q] = v; `last`: the lane's item ends at q. A lane answers its items'
  // k-mers in order, so the buffer only ever holds the group of q - 1 (and is
  // empty at a group start): a group goes out when its last slot is written
  // or its item ends.
  __device__ __forceinline__ void put(u64* __restrict__ out, bool vec, u64 q, u64 v, bool last) {
    const u32 sl = static_cast<u32>(q) & (S - 1);
    obuf[sl * kAnswerStride] = v;
    mask |= 1u << sl;
    if (sl == S - 1 || last) flush(out, vec, q);
  }

  // out[q .. q+n-1] = v (n >= 1); `last`: the lane's item ends at q+n-1.
  // The part inside q's group and the part after the last whole group go
  // through the buffer (one predicated LDS write per slot); whole aligned
  // groups go straight from registers to memory.
  __device__ __forceinline__ void put_run(u64* __restrict__ out, bool vec, u64 q, u64 n, u64 v, bool last) {
    const u32 sl = static_cast<u32>(q) & (S - 1);
    if (sl != 0 || n < S) {
      const u32 c = n < S - sl ? static_cast<u32>(n) : S - sl;
#pragma unroll
      for (u32 t = 0; t < S; t++)
        if (t >= sl && t < sl + c) obuf[t * kAnswerStride] = v;
      mask |= ((1u << c) - 1u) << sl;
      if (sl + c == S || (last && n == c)) flush(out, vec, q);
      q += c;
      n -= c;
    }
    const uint4 vv = make_uint4(static_cast<u32>(v), static_cast<u32>(v >> 32), static_cast<u32>(v),
                                static_cast<u32>(v >> 32));
    for (; n >= S; q += S, n -= S) {
      u64* o = out + q;
      if (vec) {
#pragma unroll
        for (u32 h = 0; h < S / 2; h++) reinterpret_cast<uint4*>(o)[h] = vv;
      } else {
#pragma unroll
        for (u32 h = 0; h < S; h++) o[h] = v;
      }
    }
    if (n) {
#pragma unroll
      for (u32 t = 0; t < S; t++)
        if (t < n) obuf[t * kAnswerStride] = v;
      mask |= (1u << n) - 1u;
      if (last) flush(out, vec, q);
    }
  }
};

constexpr u32 kAnswerSlots = 4;
using AnswerBuffer = AnswerBufferT<kAnswerSlots>;

// Prefilter flags (one byte per k-mer, nonzero = search it) read by a lane
// walking its own k-mers in order. Thousands of lanes per CU each walk their
// own region, so a flag read one byte at a time misses every cache and
// fetches a whole 128-byte line from HBM per k-mer; instead a lane reads the
// aligned 16-byte block holding the flag and keeps it as 16 bits.
struct MaybeFlags {
  u32 bits = 0;  // bit i: byte i of the cached block is nonzero
  u32 tag = 0;   // the cached block's address >> 4 (low 32 bits)

  __device__ __forceinline__ static u32 nonzero4(u32 x) {
    x |= x >> 4;
    x |= x >> 2;
    x |= x >> 1;  // bit 8i: byte i is nonzero
    return ((x & 0x01010101u) * 0x00204081u) >> 21 & 15u;
  }

  // Before a walk that starts at flag idx and only moves forward: marks the
  // block before idx's as cached, which the walk never asks for.
  __device__ __forceinline__ void reset(const u8* maybe, u64 idx) {
    tag = static_cast<u32>(reinterpret_cast<uintptr_t>(maybe + idx) >> 4) - 1u;
  }

  // The number of zero flags from idx on, up to the end of idx's 16-byte
  // block (0: maybe[idx] != 0), for a flag array of `total` bytes.
  __device__ __forceinline__ u32 zeros(const u8* __restrict__ maybe, u64 total, u64 idx) {
    const uintptr_t a = reinterpret_cast<uintptr_t>(maybe + idx);
    const u32 t = static_cast<u32>(a >> 4), o = static_cast<u32>(a & 15);
    if (t != tag) {
      const uintptr_t blk = a - o;
      // a block reaching outside the array (its first or last 15 bytes) is not cached
      if (blk < reinterpret_cast<uintptr_t>(maybe) || idx - o + 16 > total) return maybe[idx] ? 0u : 1u;
      const uint4 v = *reinterpret_cast<const uint4*>(maybe + (idx - o));  // (a global load, not flat)
      bits = nonzero4(v.x) | nonzero4(v.y) << 4 | nonzero4(v.z) << 8 | nonzero4(v.w) << 12;
      tag = t;
    }
    return static_cast<u32>(__builtin_ctz((bits >> o) | (0x10000u >> o)));
  }
};

}  // namespace sbwt
