// Device rank layout of libsbwt_gpu_amd.
//
// Each of the five rows (A, C, G, T and the suffix-group starts) is an array
// of fixed-size rank chunks. A chunk of CB bytes is CB/4 little-endian u32
// words: word 0 is the number of 1-bits of the row before the chunk's first
// column (absolute; the index supports n < 2^32 columns), words 1..CB/4-1
// hold the chunk's columns, column (chunk*COLS + t) in bit t%32 of word
// 1 + t/32 (LSB-first). rank(i) is therefore one chunk read: word 0 plus a
// masked popcount inside the chunk, and a 16-, 32- or 64-byte chunk never
// straddles a 128-byte cache line. This replaces the reference's separate
// L0 / L12 / data arrays (include/rank_data_structures.hpp:142-198, read by
// include/gpu/rank_kernel.hpp:253-296,409-415), which cost two or three
// random lines per rank.
#pragma once
#include "common.hpp"

namespace sbwt {

template <int CB>
struct Layout {
  static_assert(CB == 16 || CB == 32 || CB == 64, "chunk bytes");
  static constexpr u32 WORDS = CB / 4;      // u32 words per chunk
  static constexpr u32 PW = WORDS - 1;      // payload words
  static constexpr u32 COLS = PW * 32;      // columns per chunk
  static constexpr u32 Q = CB / 16;         // 16-byte loads per chunk
};

// Chunks per row for n columns: one more than n / COLS so that rank(n) has
// a chunk to read.
inline u64 chunks_for(u64 n, u32 cols) { return n / cols + 1; }
inline u32 cols_for_layout(u32 cb) { return (cb / 4 - 1) * 32; }

template <int CB>
struct Chunk {
  u32 w[Layout<CB>::WORDS];
};

template <int CB>
__device__ __forceinline__ void load_chunk(const uint4* __restrict__ row, u32 ch, Chunk<CB>& c) {
  const uint4* p = row + static_cast<u64>(ch) * Layout<CB>::Q;
#pragma unroll
  for (u32 t = 0; t < Layout<CB>::Q; t++) {
    const uint4 v = p[t];
    c.w[4 * t + 0] = v.x;
    c.w[4 * t + 1] = v.y;
    c.w[4 * t + 2] = v.z;
    c.w[4 * t + 3] = v.w;
  }
}

// Number of 1-bits before offset o (0 <= o < COLS) plus the chunk's count:
// the whole payload words below o's word, then o's word under a mask of its
// o % 32 low bits. The word comparisons are shared by the whole-word counts
// and the selection of o's word, so a rank in a 16-byte chunk is about 11
// vector instructions (a mask per word took about 28).
template <int CB>
__device__ __forceinline__ u32 rank_chunk(const Chunk<CB>& c, u32 o) {
  u32 r = c.w[0];
  u32 part = c.w[1];
#pragma unroll
  for (u32 t = 1; t < Layout<CB>::PW; t++) {
    const bool past = o >= 32 * t;  // word t - 1 lies wholly below o
    r += __popc(past ? c.w[t] : 0u);
    part = past ? c.w[1 + t] : part;
  }
  return r + __popc(part & ((1u << (o & 31)) - 1u));
}

template <int CB>
__device__ __forceinline__ u32 bit_chunk(const Chunk<CB>& c, u32 o) {
  u32 word = c.w[1];
#pragma unroll
  for (u32 t = 1; t < Layout<CB>::PW; t++)
    if (o >= 32 * t) word = c.w[1 + t];
  return (word >> (o & 31)) & 1u;
}

// Offset of the last 1-bit at or before o inside the chunk, or -1.
template <int CB>
__device__ __forceinline__ int last_one_upto(const Chunk<CB>& c, u32 o) {
  int res = -1;
#pragma unroll
  for (u32 t = 0; t < Layout<CB>::PW; t++) {
    const u32 lo = 32 * t;
    const u32 m = (o >= lo + 31) ? ~0u : ((o >= lo) ? ((2u << (o - lo)) - 1u) : 0u);
    const u32 x = c.w[1 + t] & m;
    if (x) res = static_cast<int>(lo + 31 - __clz(x));
  }
  return res;
}

// Read-only view of an index passed to kernels by value.
struct DevIndex {
  const uint4* lines;  // rows A,C,G,T,group; row c starts at lines + c * row_q
  u64 row_q;           // 16-byte units per row
  const uint2* pre;    // presearch table, 4^m intervals {l, r}
  u32 C[5];
  u32 n, k, m;
  u32 has_groups;
};

__device__ __forceinline__ u32 c_of(const DevIndex& ix, u32 c) {
  // explicit selects keep C in scalar registers (a dynamic index into a
  // kernel-argument array would spill it to scratch)
  return c == 0 ? ix.C[0] : (c == 1 ? ix.C[1] : (c == 2 ? ix.C[2] : ix.C[3]));
}

__device__ __forceinline__ const uint4* row_of(const DevIndex& ix, u32 c) {
  return ix.lines + static_cast<u64>(c) * ix.row_q;
}

template <int CB>
__device__ __forceinline__ u32 rank_dev(const DevIndex& ix, u32 c, u32 i) {
  constexpr u32 COLS = Layout<CB>::COLS;
  const u32 ch = i / COLS;
  Chunk<CB> q;
  load_chunk<CB>(row_of(ix, c), ch, q);
  return rank_chunk<CB>(q, i - ch * COLS);
}

// One narrowing step of the reference (search_kernel.hpp:87-88):
// l = C[c] + rank_c(l), r = C[c] + rank_c(r+1) - 1. Both chunk loads are
// issued before either is used.
template <int CB>
__device__ __forceinline__ void narrow(const DevIndex& ix, u32 c, u32& l, u32& r) {
  constexpr u32 COLS = Layout<CB>::COLS;
  const uint4* row = row_of(ix, c);
  const u32 a = l, b = r + 1u;
  const u32 cha = a / COLS, chb = b / COLS;
  Chunk<CB> qa, qb;
  load_chunk<CB>(row, cha, qa);
  // Both loads even when l and r + 1 share a chunk (as they mostly do on a
  // narrow interval): skipping the second one where cha == chb sped up 64-byte
  // chunks on the 3 Gbp index (1B k-mers: 193 -> 163 ms) but slowed them on
  // the pangenome index, whose chunks stay in the last-level cache (150.6 ->
  // 155 ms), and slowed 16-byte chunks (170.5 -> 172.7 ms).
  load_chunk<CB>(row, chb, qb);
  const u32 base = c_of(ix, c);
  l = base + rank_chunk<CB>(qa, a - cha * COLS);
  r = base + rank_chunk<CB>(qb, b - chb * COLS) - 1u;
}

}  // namespace sbwt
