// Search kernels of libsbwt_gpu_amd.
//
// The reference runs one thread per k-mer: a presearch gather, then up to
// k-m narrowing steps of two ranks each, early exit on an empty interval
// (include/gpu/search_kernel.hpp:61-104). Here every rank is one chunk read
// (layout.hpp) and the two chunk reads of a step are issued together. The
// read search adds streaming: once the k-mer x at node j is found, the next
// k-mer x[1..k-1]c is found from j in one step, because the suffix-group
// start g of j (the first node sharing x's last k-1 characters) carries the
// only c-edge into it: the answer is C[c] + rank_c(g) if g has a c-edge,
// else absent. This yields the same node ids as the full search.
#include "answers.hpp"
#include "index.hpp"

namespace sbwt {

// Per-lane access tally of the counted variants (compiled out otherwise).
struct Tally {
  u32 lines = 0, steps = 0, full = 0;
};

template <int CB>
__device__ __forceinline__ u32 line_of(u32 col) {
  return static_cast<u32>((static_cast<u64>(col / Layout<CB>::COLS) * CB) >> 7);
}

// Full search of a top-aligned k-mer (char 0 in bits 63:62).
template <int CB, bool COUNT>
__device__ __forceinline__ u64 full_search(const DevIndex& ix, u64 kmer, Tally& t) {
  u32 l, r, i;
  if (ix.m) {
    const uint2 e = ix.pre[kmer >> (64 - 2 * ix.m)];
    l = e.x;
    r = e.y;
    i = ix.m;
    if (COUNT) t.lines++;
  } else {
    l = 0;
    r = ix.C[4] - 1u;
    i = 0;
  }
  if (COUNT) t.full++;
  for (; i < ix.k; i++) {
    const u32 c = static_cast<u32>(kmer >> (62 - 2 * i)) & 3u;
    if (COUNT) {
      t.lines += line_of<CB>(l) == line_of<CB>(r + 1u) ? 1u : 2u;
      t.steps++;
    }
    narrow<CB>(ix, c, l, r);
    if (l > r) break;
  }
  return l > r ? kNone : static_cast<u64>(l);
}

// Searches the first len (<= 32) characters of a top-aligned window. Returns
// the interval's left end if it is non-empty after len characters, else ~0
// with `dead` = the number of characters after which it was empty (the
// prefix of that length is a dead string: no node label ends with it).
template <int CB, bool COUNT>
__device__ __forceinline__ u64 search_len(const DevIndex& ix, u64 win, u32 len, Tally& t, u32& dead) {
  u32 l, r, i;
  if (COUNT) t.full++;
  if (ix.m && len >= ix.m) {
    const uint2 e = ix.pre[win >> (64 - 2 * ix.m)];
    l = e.x;
    r = e.y;
    i = ix.m;
    if (COUNT) t.lines++;
    if (l > r) {
      dead = i;
      return kNone;
    }
  } else {
    l = 0;
    r = ix.C[4] - 1u;
    i = 0;
  }
  for (; i < len; i++) {
    const u32 c = static_cast<u32>(win >> (62 - 2 * i)) & 3u;
    if (COUNT) {
      t.lines += line_of<CB>(l) == line_of<CB>(r + 1u) ? 1u : 2u;
      t.steps++;
    }
    narrow<CB>(ix, c, l, r);
    if (l > r) {
      dead = i + 1;
      return kNone;
    }
  }
  dead = 0;
  return static_cast<u64>(l);
}

// Next k-mer after a found one at node j, appending character c.
template <int CB, bool COUNT>
__device__ __forceinline__ u64 stream_step(const DevIndex& ix, u32 j, u32 c, Tally& t) {
  constexpr u32 COLS = Layout<CB>::COLS;
  const uint4* row = row_of(ix, c);
  u32 ch = j / COLS;
  u32 o = j - ch * COLS;
  Chunk<CB> q;
  load_chunk<CB>(row, ch, q);
  if (COUNT) {
    t.lines++;
    t.steps++;
  }
  // a set bit means j starts its group and has the c-edge
  if (bit_chunk<CB>(q, o)) return static_cast<u64>(c_of(ix, c) + rank_chunk<CB>(q, o));
  // otherwise find the group start g <= j; column 0 always starts a group
  const uint4* grow = row_of(ix, kGroupRow);
  Chunk<CB> gq;
  load_chunk<CB>(grow, ch, gq);
  if (COUNT) t.lines++;
  int p = last_one_upto<CB>(gq, o);
  while (p < 0 && ch > 0) {
    ch--;
    load_chunk<CB>(grow, ch, gq);
    if (COUNT && line_of<CB>(ch * COLS) != line_of<CB>((ch + 1) * COLS)) t.lines++;
    p = last_one_upto<CB>(gq, COLS - 1);
  }
  const u32 g = p < 0 ? 0u : ch * COLS + static_cast<u32>(p);
  if (g == j) return kNone;
  const u32 gch = g / COLS, go = g - gch * COLS;
  load_chunk<CB>(row, gch, q);
  if (COUNT && line_of<CB>(g) != line_of<CB>(j)) t.lines++;
  if (!bit_chunk<CB>(q, go)) return kNone;
  return static_cast<u64>(c_of(ix, c) + rank_chunk<CB>(q, go));
}

// Adds a lane's tally to the index counters (one atomic per counter per wave).
template <bool COUNT>
__device__ __forceinline__ void flush(unsigned long long* ctr, const Tally& t, u32 kmers) {
  if (!COUNT) return;
  u64 v[4] = {t.lines, t.steps, kmers, t.full};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    u64 x = v[i];
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(ctr + i, static_cast<unsigned long long>(x));
  }
}

// The reference's two-word extraction (search_kernel.hpp:66-68): the 64
// bits starting at bit 2*pos of the packed string.
__device__ __forceinline__ u64 extract2(const u64* __restrict__ w, u64 pos) {
  const u64 b = pos << 1;
  return (w[b >> 6] << (b & 63)) | ((w[(b >> 6) + 1] >> (63 - (b & 63))) >> 1);
}

// The 64 bits of a one-bit-per-character mask starting at character pos.
__device__ __forceinline__ u64 extract1(const u64* __restrict__ w, u64 pos) {
  return (w[pos >> 6] << (pos & 63)) | ((w[(pos >> 6) + 1] >> (63 - (pos & 63))) >> 1);
}

template <int CB, bool COUNT>
__global__ __launch_bounds__(256) void k_search_kmers(DevIndex ix, const u64* __restrict__ kmers, u64 count,
                                                      const u8* __restrict__ maybe, u64* __restrict__ out,
                                                      unsigned long long* ctr) {
  const u64 q = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  Tally t;
  if (q < count) out[q] = maybe && !maybe[q] ? kNone : full_search<CB, COUNT>(ix, kmers[q], t);
  flush<COUNT>(ctr, t, q < count ? 1u : 0u);
}

// Wave-cooperative variant of k_search_kmers for chunks of 32 or 64 bytes:
// a group of G = CB / 16 adjacent lanes searches one k-mer. Every rank
// reads one 16-byte part of its chunk per lane (part p = lane mod G holds
// the chunk's words 4p .. 4p+3; word 0 is the count before the chunk), each
// lane counts the bits of its part below the rank's offset, and G - 1
// shuffles sum the parts; every lane of the group then holds the same
// interval and takes the same branch. So a rank is one 16-byte load per lane
// instead of G loads by one lane, and the G loads of a group fall in one
// chunk (the reference's basic-block popcount split over lanes with
// shuffles, include/gpu/rank_kernel.hpp:298-396).
template <int CB>
__device__ __forceinline__ u32 part_rank(const uint4& v, u32 part, u32 o) {
  const u32 w[4] = {v.x, v.y, v.z, v.w};
  u32 r = 0;
#pragma unroll
  for (u32 i = 0; i < 4; i++) {
    const u32 cw = 4 * part + i;  // chunk word
    if (cw == 0) {
      r += w[i];  // the count before the chunk
    } else {
      const u32 lo = 32 * (cw - 1);  // first column of payload word cw - 1
      r += __popc(o >= lo + 32 ? w[i] : (o > lo ? w[i] & ((1u << (o - lo)) - 1u) : 0u));
    }
  }
  return r;
}

template <int CB, bool COUNT>
__global__ __launch_bounds__(256) void k_search_kmers_coop(DevIndex ix, const u64* __restrict__ kmers, u64 count,
                                                           const u8* __restrict__ maybe, u64* __restrict__ out,
                                                           unsigned long long* ctr) {
  constexpr u32 G = CB / 16;
  constexpr u32 COLS = Layout<CB>::COLS;
  static_assert(G == 2 || G == 4, "two or four lanes per k-mer");
  const u64 tid = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  const u64 q = tid / G;
  const u32 part = static_cast<u32>(tid % G);
  Tally t;
  const bool live = q < count && !(maybe && !maybe[q]);
  u64 res = kNone;
  if (live) {
    const u64 kmer = kmers[q];
    u32 l, r, i;
    if (ix.m) {
      const uint2 e = ix.pre[kmer >> (64 - 2 * ix.m)];
      l = e.x;
      r = e.y;
      i = ix.m;
      if (COUNT && part == 0) t.lines++;
    } else {
      l = 0;
      r = ix.C[4] - 1u;
      i = 0;
    }
    if (COUNT && part == 0) t.full++;
    for (; i < ix.k; i++) {
      const u32 c = static_cast<u32>(kmer >> (62 - 2 * i)) & 3u;
      const u32 a = l, b = r + 1u;
      const u32 cha = a / COLS, chb = b / COLS;
      if (COUNT && part == 0) {
        t.lines += line_of<CB>(a) == line_of<CB>(b) ? 1u : 2u;
        t.steps++;
      }
      const uint4* row = row_of(ix, c);
      const uint4 va = row[static_cast<u64>(cha) * Layout<CB>::Q + part];
      uint4 vb = va;  // a single-column interval's l and r + 1 mostly share a chunk
      if (chb != cha) vb = row[static_cast<u64>(chb) * Layout<CB>::Q + part];
      u32 ra = part_rank<CB>(va, part, a - cha * COLS), rb = part_rank<CB>(vb, part, b - chb * COLS);
#pragma unroll
      for (u32 d = 1; d < G; d <<= 1) {
        ra += __shfl_xor(ra, static_cast<int>(d), static_cast<int>(G));
        rb += __shfl_xor(rb, static_cast<int>(d), static_cast<int>(G));
      }
      const u32 base = c_of(ix, c);
      l = base + ra;
      r = base + rb - 1u;
      if (l > r) break;
    }
    res = l > r ? kNone : static_cast<u64>(l);
  }
  if (q < count && part == 0) out[q] = res;
  flush<COUNT>(ctr, t, (q < count && part == 0) ? 1u : 0u);
}

template <int CB>
__global__ __launch_bounds__(256) void k_search_positions(DevIndex ix, const u64* __restrict__ packed,
                                                          u64* __restrict__ pos_res, u64 count) {
  const u64 q = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (q >= count) return;
  Tally t;
  pos_res[q] = full_search<CB, false>(ix, extract2(packed, pos_res[q]), t);
}

// The 64 bits of a packed string (bits_per_char = 2) or of a one-bit mask
// (1) starting at a character that advances by one per call, holding the
// two words it spans in registers: one load per 32 (or 64) characters.
template <u32 BITS>
struct Rolling {
  const u64* w;
  u64 a = 0, b = 0, idx = ~0ull - 1;
  __device__ __forceinline__ u64 at(u64 pos) {
    const u64 bit = pos * BITS;
    const u64 word = bit >> 6;
    if (word != idx) {
      a = (word == idx + 1) ? b : w[word];
      b = w[word + 1];
      idx = word;
    }
    const u32 sh = static_cast<u32>(bit & 63);
    return (a << sh) | ((b >> (63 - sh)) >> 1);
  }
};

// One lane per work item: up to `seg` consecutive k-mers of one row, in
// order. With STREAM, a k-mer after a found one takes one streaming step.
// With CERT, a k-mer found absent with a dead string ending at read
// position d (d = the new character for a failed streaming step, else where
// the search died) triggers two short searches that can answer the next
// absent k-mers without searching them: the k-mer starting at d (dead after
// L characters: every k-mer containing [d, d+L-1] is absent) and the string
// [lo-1, d] that the k-mers in between share; when no k-mer starts at d
// (the row's last k - 1 characters), a string [q, d] that the next k-mers up
// to q contain, long enough to be absent by chance. A dead string w proves every
// k-mer containing w absent, because every substring of an indexed k-mer
// ends some node label (follow the k-mer's incoming edges back to the root).
// With `maybe` (one byte per k-mer, in output order), k-mers marked 0 by a
// prefilter answer ~0 without a search.
template <int CB, bool STREAM, bool NBREAK, bool CERT, bool COUNT>
__global__ __launch_bounds__(256) void k_search_reads(DevIndex ix, const u64* __restrict__ packed,
                                                      const u64* __restrict__ nmask,
                                                      const u64* __restrict__ row_starts,
                                                      const u64* __restrict__ offsets,
                                                      const u64* __restrict__ work, u64 rows, u64 bound, u32 seg,
                                                      const u8* __restrict__ maybe, u64* __restrict__ out,
                                                      u32 vec_out, unsigned long long* ctr,
                                                      const uint4* __restrict__ runs,
                                                      const unsigned long long* __restrict__ run_count) {
  __shared__ u64 obuf[kAnswerSlots * kAnswerStride];
  AnswerBuffer ab{obuf + threadIdx.x};
  const u64 gid = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  Tally t;
  u32 done = 0;
  // the item: k-mers [j0, j0 + cnt) of row r, from a run item or from the
  // work prefix (work[r] <= gid < work[r+1], segments of seg k-mers)
  u64 r = 0, j0 = 0, cnt = 0, jv = 0;  // jv: first k-mer after the run item's leading gap
  bool have = false;
  if (runs) {
    have = gid < bound && gid < *run_count;
    if (have) {
      const uint4 it = runs[gid];
      r = it.x;
      j0 = it.y;
      jv = it.z;
      cnt = it.w - it.y;
    }
  } else if (gid < bound && rows != 0 && gid < work[rows]) {
    have = true;
    u64 lo = 0, hi = rows;
    while (hi - lo > 1) {
      const u64 mid = (lo + hi) >> 1;
      if (work[mid] <= gid) lo = mid; else hi = mid;
    }
    r = lo;
    const u64 rk = offsets[r + 1] - offsets[r];
    j0 = (gid - work[r]) * seg;
    cnt = (rk - j0) < seg ? (rk - j0) : seg;
  }
  if (have) {
    const u64 rowk = offsets[r + 1] - offsets[r];
    const u64 rs = row_starts[r];
    const u64 obase = offsets[r];
    const u64 nk = maybe ? offsets[rows] : 0;
    MaybeFlags mf;
    if (maybe) mf.reset(maybe, obase + j0);
    const u32 k = ix.k;
    bool prev = false;
    u32 pnode = 0;
    u64 clo = 1, chi = 0;  // certified-absent k-mers of the row
    u64 kpos = ~0ull;      // a k-mer already found by a certificate search
    u32 knode = 0;
    Rolling<2> code{packed};
    Rolling<1> bad{nmask};
    const u64 jend = j0 + cnt;
    // a run item's gap [j0, jv): answered ~0 before the search
    const u64 jfirst = jv > j0 ? (jv < jend ? jv : jend) : j0;
    for (u64 j = jfirst; j < jend; j++) {
      // a run of k-mers answered ~0 without a search: certified absent,
      // holding a non-ACGT character (every k-mer from j up to the first one
      // holds it), or filtered out
      u64 run = 0;
      if (CERT && j >= clo && j <= chi) {
        run = chi + 1 - j;
      } else {
        const u64 nb = NBREAK ? bad.at(rs + j) : 0;
        const u32 p = nb ? static_cast<u32>(__clzll(nb)) : 64u;
        if (p < k) run = p + 1;
        else if (maybe) run = mf.zeros(maybe, nk, obase + j);
      }
      if (run) {
        if (run > jend - j) run = jend - j;
        ab.put_run(out, vec_out != 0, obase + j, run, kNone, j + run == jend);
        prev = false;
        j += run - 1;
        continue;
      }
      u64 res;
      u64 d = 0;  // end of a dead string found at this k-mer (0: none)
      const u64 win = code.at(rs + j);
      if (CERT && j == kpos) {
        res = knode;
      } else if (STREAM && prev) {
        res = stream_step<CB, COUNT>(ix, pnode, static_cast<u32>(win >> (64 - 2 * k)) & 3u, t);
        if (res == kNone) d = j + k - 1;
      } else {
        u32 dead;
        res = search_len<CB, COUNT>(ix, win, k, t, dead);
        if (res == kNone) d = j + dead - 1;
      }
      prev = res != kNone;
      pnode = static_cast<u32>(res);
      ab.put(out, vec_out != 0, obase + j, res, j + 1 == jend);
      // a string long enough to be absent by chance: log4(n) + 2 characters
      const u32 dlen = min(k, (33u - static_cast<u32>(__clz(ix.n))) / 2u + 2u);
      if (CERT && d > j && d < rowk && !(clo <= chi && chi > j)) {
        // the search from d stops alive after dlen + 2 characters (d past
        // the error: nothing to certify, and the rest of the k-mer's search
        // would cost a whole search for one known k-mer)
        u32 lp;
        const u32 cap = min(k, dlen + 2);
        const u64 rp = search_len<CB, COUNT>(ix, extract2(packed, rs + d), cap, t, lp);
        if (rp != kNone) {
          if (cap == k) {
            kpos = d;
            knode = static_cast<u32>(rp);
          }
        } else {
          const u64 lo2 = d + lp > k ? d + lp - k : 0;
          clo = lo2 > j + 1 ? lo2 : j + 1;
          chi = d;
          if (lo2 > j + 1 && d - lo2 + 2 <= k) {
            // the k-mers j+1 .. lo2-1 all contain [lo2-1, d]
            const u64 q = lo2 - 1;
            u32 dl;
            if (search_len<CB, COUNT>(ix, extract2(packed, rs + q), static_cast<u32>(d - q + 1), t, dl) == kNone)
              clo = j + 1;
          }
        }
      } else if (CERT && d >= rowk && !(clo <= chi && chi > j)) {
        // no k-mer starts at d; the k-mers j+1 .. q all contain [s, d] for
        // any s >= q: a string of dlen characters, or of dlen + 2 when the
        // row's last k-mer starts further from d
        const u64 dl = dlen;
        const u64 q = d + 1 >= dl ? (rowk - 1 < d + 1 - dl ? rowk - 1 : d + 1 - dl) : 0;
        const u64 ls = min(k, dlen + 2);
        const u64 s0 = d + 1 >= ls ? d + 1 - ls : 0;
        const u64 sq = s0 > q ? s0 : q;
        u32 dd;
        if (d + 1 >= dl && q > j &&
            search_len<CB, COUNT>(ix, extract2(packed, rs + sq), static_cast<u32>(d - sq + 1), t, dd) == kNone) {
          clo = j + 1;
          chi = q;
        }
      }
    }
    done = static_cast<u32>(cnt);
  }
  flush<COUNT>(ctr, t, done);
}

// A dispatch holds fewer than 2^32 work-items, so a batch is searched in
// pieces of at most kKmerPiece work-items: kKmerPiece / G k-mers for G lanes
// per k-mer (a billion k-mers is one dispatch in every kernel).
constexpr u64 kKmerPiece = (1ull << 32) - (1ull << 16);

template <int CB>
static int launch_kmers_coop(const sbwt_index* ix, const u64* kmers, u64 count, const u8* maybe, u64* out,
                             bool count_lines, hipStream_t s) {
  const u32 bt = 256;
  constexpr u64 piece = kKmerPiece / (CB / 16);
  for (u64 at = 0; at < count; at += piece) {
    const u64 n = count - at < piece ? count - at : piece;
    const u64 grid = (n * (CB / 16) + bt - 1) / bt;
    const u8* mb = maybe ? maybe + at : nullptr;
    if (count_lines)
      k_search_kmers_coop<CB, true><<<grid, bt, 0, s>>>(dev_view(ix), kmers + at, n, mb, out + at, ix->d_counters);
    else
      k_search_kmers_coop<CB, false><<<grid, bt, 0, s>>>(dev_view(ix), kmers + at, n, mb, out + at, ix->d_counters);
    SBWT_HIP(hipGetLastError());
  }
  return SBWT_OK;
}

template <int CB>
static int launch_kmers(const sbwt_index* ix, const u64* kmers, u64 count, const u8* maybe, u64* out,
                        bool count_lines, hipStream_t s) {
  const u32 bt = 256;
  for (u64 at = 0; at < count; at += kKmerPiece) {
    const u64 n = count - at < kKmerPiece ? count - at : kKmerPiece;
    const u64 grid = (n + bt - 1) / bt;
    const u8* mb = maybe ? maybe + at : nullptr;
    if (count_lines)
      k_search_kmers<CB, true><<<grid, bt, 0, s>>>(dev_view(ix), kmers + at, n, mb, out + at, ix->d_counters);
    else
      k_search_kmers<CB, false><<<grid, bt, 0, s>>>(dev_view(ix), kmers + at, n, mb, out + at, ix->d_counters);
    SBWT_HIP(hipGetLastError());
  }
  return SBWT_OK;
}

template <int CB>
static int launch_positions(const sbwt_index* ix, const u64* packed, u64* pos, u64 count, hipStream_t s) {
  const u32 bt = 256;
  for (u64 at = 0; at < count; at += kKmerPiece) {
    const u64 n = count - at < kKmerPiece ? count - at : kKmerPiece;
    k_search_positions<CB><<<(n + bt - 1) / bt, bt, 0, s>>>(dev_view(ix), packed, pos + at, n);
    SBWT_HIP(hipGetLastError());
  }
  return SBWT_OK;
}

struct ReadArgs {
  const u64 *packed, *nmask, *rs, *off, *work;
  u64 rows, bound;
  u32 seg;
  const u8* maybe;
  u64* out;
  RunList runs;
};

template <int CB, bool STREAM, bool NBREAK, bool CERT, bool COUNT>
static int launch_reads5(const sbwt_index* ix, const ReadArgs& a, hipStream_t s) {
  const u32 bt = kAnswerStride;  // the answer buffer's LDS layout assumes this block size
  k_search_reads<CB, STREAM, NBREAK, CERT, COUNT><<<(a.bound + bt - 1) / bt, bt, 0, s>>>(
      dev_view(ix), a.packed, a.nmask, a.rs, a.off, a.work, a.rows, a.bound, a.seg, a.maybe, a.out,
      (reinterpret_cast<uintptr_t>(a.out) & 15) == 0 ? 1u : 0u, ix->d_counters, a.runs.runs, a.runs.count);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

template <int CB, bool STREAM, bool NBREAK, bool CERT>
static int launch_reads4(const sbwt_index* ix, const ReadArgs& a, bool count, hipStream_t s) {
  return count ? launch_reads5<CB, STREAM, NBREAK, CERT, true>(ix, a, s)
               : launch_reads5<CB, STREAM, NBREAK, CERT, false>(ix, a, s);
}

template <int CB, bool STREAM, bool NBREAK>
static int launch_reads3(const sbwt_index* ix, const ReadArgs& a, bool cert, bool count, hipStream_t s) {
  return cert ? launch_reads4<CB, STREAM, NBREAK, true>(ix, a, count, s)
              : launch_reads4<CB, STREAM, NBREAK, false>(ix, a, count, s);
}

template <int CB>
static int launch_reads(const sbwt_index* ix, const ReadArgs& a, bool stream, bool nbreak, bool cert, bool count,
                        hipStream_t s) {
  if (stream && nbreak) return launch_reads3<CB, true, true>(ix, a, cert, count, s);
  if (stream) return launch_reads3<CB, true, false>(ix, a, cert, count, s);
  if (nbreak) return launch_reads3<CB, false, true>(ix, a, cert, count, s);
  return launch_reads3<CB, false, false>(ix, a, cert, count, s);
}

int search_reads_runs(const sbwt_index* ix, const u64* packed, const u64* row_starts, const u64* offsets, u64 rows,
                      u32 seg, const u8* maybe, u64* out, u32 flags, const RunList& runs,
                      unsigned long long* item_ctr, hipStream_t s) {
  if (rows == 0 || runs.cap == 0) return SBWT_OK;
  // run items hold no non-ACGT window: nothing to break on
  const bool independent = (flags & SBWT_READS_INDEPENDENT) != 0;
  const bool stream_mode = !independent && ix->has_groups;
  const bool cert = !independent && !(flags & SBWT_READS_NO_CERT);
  const bool cnt = (flags & SBWT_COUNT) != 0;
  const bool uniform = (flags & SBWT_READS_PERSISTENT) ||
                       (!(flags & SBWT_READS_SEGMENT_LANES) && stream_mode && cert);
  if (uniform)
    return search_reads_persistent(ix, packed, nullptr, row_starts, offsets, nullptr, nullptr, rows, runs.cap, seg,
                                   maybe, out, false, stream_mode, cert, cnt, runs, item_ctr, s);
  const ReadArgs a{packed, nullptr, row_starts, offsets, nullptr, rows, runs.cap, seg, maybe, out, runs};
  switch (ix->layout) {
    case 16: return launch_reads<16>(ix, a, stream_mode, false, cert, cnt, s);
    case 32: return launch_reads<32>(ix, a, stream_mode, false, cert, cnt, s);
    default: return launch_reads<64>(ix, a, stream_mode, false, cert, cnt, s);
  }
}

}  // namespace sbwt

using namespace sbwt;

extern "C" {

static int search_kmers_impl(const sbwt_index* ix, const uint64_t* d_kmers, uint64_t count, const uint8_t* d_maybe,
                             uint64_t* d_out, uint32_t flags, void* stream) {
  SBWT_REQUIRE(ix, "index is NULL");
  if (count == 0) return SBWT_OK;
  SBWT_REQUIRE(d_kmers && d_out, "device buffers must not be NULL");
  const bool cnt = (flags & SBWT_COUNT) != 0;
  const bool coop = (flags & SBWT_SEARCH_COOPERATIVE) != 0;
  hipStream_t s = as_stream(stream);
  switch (ix->layout) {
    case 16: return launch_kmers<16>(ix, d_kmers, count, d_maybe, d_out, cnt, s);
    case 32: return coop ? launch_kmers_coop<32>(ix, d_kmers, count, d_maybe, d_out, cnt, s)
                         : launch_kmers<32>(ix, d_kmers, count, d_maybe, d_out, cnt, s);
    default: return coop ? launch_kmers_coop<64>(ix, d_kmers, count, d_maybe, d_out, cnt, s)
                         : launch_kmers<64>(ix, d_kmers, count, d_maybe, d_out, cnt, s);
  }
}

int sbwt_search_kmers_flags(const sbwt_index* ix, const uint64_t* d_kmers, uint64_t count, uint64_t* d_out,
                            uint32_t flags, void* stream) {
  clear_error();
  return search_kmers_impl(ix, d_kmers, count, nullptr, d_out, flags, stream);
}

int sbwt_search_kmers_masked(const sbwt_index* ix, const uint64_t* d_kmers, uint64_t count, const uint8_t* d_maybe,
                             uint64_t* d_out, uint32_t flags, void* stream) {
  clear_error();
  SBWT_REQUIRE(d_maybe || count == 0, "d_maybe must not be NULL");
  return search_kmers_impl(ix, d_kmers, count, d_maybe, d_out, flags, stream);
}

int sbwt_search_kmers(const sbwt_index* ix, const uint64_t* d_kmers, uint64_t count, uint64_t* d_out,
                      void* stream) {
  return sbwt_search_kmers_flags(ix, d_kmers, count, d_out, 0, stream);
}

int sbwt_search_positions(const sbwt_index* ix, const uint64_t* d_packed, uint64_t* d_positions, uint64_t count,
                          void* stream) {
  clear_error();
  SBWT_REQUIRE(ix, "index is NULL");
  if (count == 0) return SBWT_OK;
  SBWT_REQUIRE(d_packed && d_positions, "device buffers must not be NULL");
  hipStream_t s = as_stream(stream);
  switch (ix->layout) {
    case 16: return launch_positions<16>(ix, d_packed, d_positions, count, s);
    case 32: return launch_positions<32>(ix, d_packed, d_positions, count, s);
    default: return launch_positions<64>(ix, d_packed, d_positions, count, s);
  }
}

static int search_reads_impl(const sbwt_index* ix, const uint64_t* d_packed, const uint64_t* d_nmask,
                             const uint64_t* d_row_starts, const uint64_t* d_offsets, const uint64_t* d_work,
                             const uint32_t* d_items, uint64_t rows, uint64_t work_bound, uint32_t seg,
                             const uint8_t* d_maybe, uint64_t* d_out, uint32_t flags, void* stream,
                             unsigned long long* item_ctr = nullptr) {
  SBWT_REQUIRE(ix, "index is NULL");
  if (rows == 0 || work_bound == 0) return SBWT_OK;
  SBWT_REQUIRE(d_packed && d_row_starts && d_offsets && d_work && d_out, "device buffers must not be NULL");
  const bool nbreak = !(flags & SBWT_READS_N_AS_A);
  const bool independent = (flags & SBWT_READS_INDEPENDENT) != 0;
  SBWT_REQUIRE(!nbreak || d_nmask, "d_nmask is required unless SBWT_READS_N_AS_A is set");
  if (seg == 0) seg = 256;
  const bool cnt = (flags & SBWT_COUNT) != 0;
  // independent: every k-mer searched from the presearch table, nothing
  // streamed or certified (the reference's way)
  const bool stream_mode = !independent && ix->has_groups;
  const bool cert = !independent && !(flags & SBWT_READS_NO_CERT);
  // the uniform-step kernel wins where certificates mix short searches into
  // the streams; the per-lane kernel's tight loop wins for independent
  // searches and streams without certificates
  bool uniform = (flags & SBWT_READS_PERSISTENT) ||
                 (!(flags & SBWT_READS_SEGMENT_LANES) && stream_mode && cert);
  // The uniform kernel keeps row-relative positions in u32. A batch whose
  // work bound admits 2^32 - 64 k-mers or more could hold such a row, so it
  // goes to the per-lane kernel (u64 positions), or is refused when the
  // uniform kernel is forced.
  if (uniform && work_bound > 0xFFFFFFC0ull / seg) {
    if (flags & SBWT_READS_PERSISTENT)
      return fail(SBWT_ERR_UNSUPPORTED,
                  "SBWT_READS_PERSISTENT: a batch of work bound %llu x %u k-mers may hold a row of 2^32 k-mers "
                  "or more, which the uniform kernel does not support",
                  (unsigned long long)work_bound, seg);
    uniform = false;
  }
  if (uniform)
    return search_reads_persistent(ix, d_packed, d_nmask, d_row_starts, d_offsets, d_work, d_items, rows,
                                   work_bound, seg, d_maybe, d_out, nbreak, stream_mode, cert, cnt, RunList{},
                                   item_ctr, as_stream(stream));
  const ReadArgs a{d_packed, d_nmask, d_row_starts, d_offsets, d_work, rows, work_bound, seg, d_maybe, d_out,
                   RunList{}};
  hipStream_t s = as_stream(stream);
  switch (ix->layout) {
    case 16: return launch_reads<16>(ix, a, stream_mode, nbreak, cert, cnt, s);
    case 32: return launch_reads<32>(ix, a, stream_mode, nbreak, cert, cnt, s);
    default: return launch_reads<64>(ix, a, stream_mode, nbreak, cert, cnt, s);
  }
}

int sbwt_search_reads(const sbwt_index* ix, const uint64_t* d_packed, const uint64_t* d_nmask,
                      const uint64_t* d_row_starts, const uint64_t* d_offsets, const uint64_t* d_work,
                      const uint32_t* d_items, uint64_t rows, uint64_t work_bound, uint32_t seg, uint64_t* d_out,
                      uint32_t flags, void* stream) {
  clear_error();
  return search_reads_impl(ix, d_packed, d_nmask, d_row_starts, d_offsets, d_work, d_items, rows, work_bound, seg,
                           nullptr, d_out, flags, stream);
}

int sbwt_search_reads_masked(const sbwt_index* ix, const uint64_t* d_packed, const uint64_t* d_nmask,
                             const uint64_t* d_row_starts, const uint64_t* d_offsets, const uint64_t* d_work,
                             const uint32_t* d_items, uint64_t rows, uint64_t work_bound, uint32_t seg,
                             const uint8_t* d_maybe, uint64_t* d_out, uint32_t flags, void* stream) {
  clear_error();
  SBWT_REQUIRE(d_maybe, "d_maybe must not be NULL");
  return search_reads_impl(ix, d_packed, d_nmask, d_row_starts, d_offsets, d_work, d_items, rows, work_bound, seg,
                           d_maybe, d_out, flags, stream);
}

}  // extern "C"

namespace sbwt {

int search_reads_counted_on(const sbwt_index* ix, const u64* packed, const u64* nmask, const u64* row_starts,
                            const u64* offsets, const u64* work, const u32* items, u64 rows, u64 work_bound,
                            u32 seg, const u8* maybe, u64* out, u32 flags, unsigned long long* item_ctr,
                            hipStream_t s) {
  return search_reads_impl(ix, packed, nmask, row_starts, offsets, work, items, rows, work_bound, seg, maybe, out,
                           flags, s, item_ctr);
}

}  // namespace sbwt
