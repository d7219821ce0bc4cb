// Persistent read search of libsbwt_gpu_amd (the default read search).
//
// Same answers as searching every k-mer independently (the reference's
// search_kernel.hpp:61-104 applied to every position of every row), computed
// with less work:
//
// * Streaming: after k-mer x at node j is found, x[1..k-1]c is answered from
//   j in one step through the suffix-group start (see search.hip).
// * Dead-substring certificates: when a search of a string w ends with an
//   empty interval, no node label ends with w, so no k-mer of the index
//   contains w (every substring of an indexed k-mer ends some node label:
//   walk the k-mer's incoming edges backwards), and every k-mer of the row
//   that contains w is absent without a search. After a k-mer is found
//   absent with a dead string ending at read position d (typically a
//   sequencing error), the kernel searches the k-mer starting at d (dead
//   after L chars, certifying the k-mers that contain [d, d+L-1]) and the
//   string ending at d that the remaining k-mers share, so a run of up to k
//   absent k-mers costs two short searches instead of k.
// * Lane refill: each lane is a small state machine that issues exactly one
//   rank-chunk (or presearch) load per loop iteration and fetches a new work
//   item (a segment of a row) from a global counter as soon as it runs out,
//   so no lane idles while another finishes a long search.
#include "index.hpp"

namespace sbwt {

namespace {

enum : u32 { OP_NONE = 0, OP_PRE, OP_STEP, OP_STREAM, OP_GSCAN, OP_GEDGE, OP_DONE };
enum : u32 { P_KMER = 0, P_CERTP, P_CERTL };

__device__ __forceinline__ u64 window2(const u64* __restrict__ w, u64 pos) {
  const u64 b = pos << 1;
  return (w[b >> 6] << (b & 63)) | ((w[(b >> 6) + 1] >> (63 - (b & 63))) >> 1);
}

template <int CB>
__device__ __forceinline__ u32 chunk_line(u32 col) {
  return static_cast<u32>((static_cast<u64>(col / Layout<CB>::COLS) * CB) >> 7);
}

struct ReadsDev {
  const u64* packed;
  const u64* nmask;
  const u64* row_starts;
  const u64* offsets;
  const u64* work;
  const u32* items;
  u64 rows;
  u32 seg;
};

}  // namespace

__global__ void k_zero_counter(unsigned long long* c) { *c = 0ull; }

template <int CB, bool NBREAK, bool STREAM, bool CERT, bool COUNT>
__global__ __launch_bounds__(256) void k_reads_v2(DevIndex ix, ReadsDev rd, unsigned long long* __restrict__ items,
                                                  u64* __restrict__ out, unsigned long long* ctr) {
  constexpr u32 COLS = Layout<CB>::COLS;
  const u32 k = ix.k, m = ix.m;
  const u64 total_items = rd.rows ? rd.work[rd.rows] : 0;

  u32 op = OP_NONE;
  // work item
  u64 cbase = 0, obase = 0;
  u32 j = 0, jend = 0, rowk = 0;
  bool prev = false;
  u32 pnode = 0;
  u32 clo = 1, chi = 0;  // certified k-mers [clo, chi] of the row
  u32 kpos = ~0u, knode = 0;
  // rolling code window of k-mer j: packed words a, a+1
  // (kNoWord + 1 is no word index either, so a fresh lane always loads both)
  constexpr u64 kNoWord = ~0ull - 1;
  u64 wa = 0, wb = 0, widx = kNoWord;
  // search state
  u64 win = 0;
  u32 s = 0, len = 0, i = 0, l = 0, r = 0, purpose = 0;
  // certificate / stream state
  u32 cp = 0, sc = 0, gch = 0, glim = 0, gpos = 0;
  u32 t_lines = 0, t_steps = 0, t_kmers = 0, t_search = 0;

  auto code_window = [&](u32 jj) -> u64 {
    const u64 b = (cbase + jj) << 1;
    const u64 a = b >> 6;
    if (a != widx) {
      if (a == widx + 1) {
        wa = wb;
      } else {
        wa = rd.packed[a];
      }
      wb = rd.packed[a + 1];
      widx = a;
    }
    return (wa << (b & 63)) | ((wb >> (63 - (b & 63))) >> 1);
  };
  auto n_window = [&](u32 jj) -> u64 {
    const u64 b = cbase + jj;
    return (rd.nmask[b >> 6] << (b & 63)) | ((rd.nmask[(b >> 6) + 1] >> (63 - (b & 63))) >> 1);
  };
  auto start_search = [&](u32 pur, u32 start, u32 length, u64 w) {
    purpose = pur;
    s = start;
    len = length;
    win = w;
    if (COUNT) t_search++;
    if (m && length >= m) {
      op = OP_PRE;
    } else {
      l = 0;
      r = ix.C[4] - 1u;
      i = 0;
      op = OP_STEP;
    }
  };
  // a k-mer before j was found absent with a dead string ending at char d
  auto start_cert = [&](u32 d) {
    if (!CERT) return;
    if (clo <= chi && chi >= j) return;  // a pending range still covers what follows
    if (d < j || d >= rowk) return;      // nothing ahead to certify, or no k-mer starts at d
    cp = d;
    start_search(P_CERTP, d, k, window2(rd.packed, cbase + d));
  };
  auto fetch = [&]() -> bool {
    const u64 it = atomicAdd(items, 1ull);
    if (it >= total_items) return false;
    u64 lo = 0;
    if (rd.items) {
      lo = rd.items[it];
    } else {
      u64 hi = rd.rows;
      while (hi - lo > 1) {
        const u64 mid = (lo + hi) >> 1;
        if (rd.work[mid] <= it) lo = mid; else hi = mid;
      }
    }
    const u64 o0 = rd.offsets[lo];
    rowk = static_cast<u32>(rd.offsets[lo + 1] - o0);
    j = static_cast<u32>(it - rd.work[lo]) * rd.seg;
    jend = rowk - j < rd.seg ? rowk : j + rd.seg;
    cbase = rd.row_starts[lo];
    obase = o0;
    prev = false;
    clo = 1;
    chi = 0;
    kpos = ~0u;
    widx = kNoWord;
    return true;
  };
  auto answer = [&](u64 v) {
    out[obase + j] = v;
    j++;
    if (COUNT) t_kmers++;
  };

  while (true) {
    // ---- advance to this lane's next memory operation ----
    while (op == OP_NONE) {
      if (j >= jend) {
        if (!fetch()) {
          op = OP_DONE;
          break;
        }
        continue;
      }
      if (j >= clo && j <= chi) {
        prev = false;
        answer(kNone);
        continue;
      }
      if (NBREAK && (n_window(j) >> (64 - k)) != 0) {
        prev = false;
        answer(kNone);
        continue;
      }
      if (j == kpos) {
        prev = true;
        pnode = knode;
        answer(knode);
        continue;
      }
      const u64 w = code_window(j);
      if (STREAM && prev) {
        sc = static_cast<u32>(w >> (64 - 2 * k)) & 3u;
        op = OP_STREAM;
      } else {
        start_search(P_KMER, j, k, w);
      }
    }
    if (!__any(op != OP_DONE)) break;

    // ---- memory phase: one load (or two chunks of one step) per lane ----
    const uint4* pa = nullptr;
    const uint4* pb = nullptr;
    const uint2* pp = nullptr;
    u32 c = 0, ca = 0, cb = 0;
    if (op == OP_PRE) {
      pp = ix.pre + (win >> (64 - 2 * m));
    } else if (op == OP_STEP) {
      c = static_cast<u32>(win >> (62 - 2 * i)) & 3u;
      ca = l / COLS;
      cb = (r + 1u) / COLS;
      pa = row_of(ix, c) + static_cast<u64>(ca) * Layout<CB>::Q;
      pb = row_of(ix, c) + static_cast<u64>(cb) * Layout<CB>::Q;
    } else if (op == OP_STREAM) {
      c = sc;
      ca = pnode / COLS;
      pa = row_of(ix, c) + static_cast<u64>(ca) * Layout<CB>::Q;
    } else if (op == OP_GSCAN) {
      pa = row_of(ix, kGroupRow) + static_cast<u64>(gch) * Layout<CB>::Q;
    } else if (op == OP_GEDGE) {
      c = sc;
      ca = gpos / COLS;
      pa = row_of(ix, c) + static_cast<u64>(ca) * Layout<CB>::Q;
    }
    Chunk<CB> qa, qb;
    uint2 e = make_uint2(0, 0);
    if (pa) {
#pragma unroll
      for (u32 t = 0; t < Layout<CB>::Q; t++) {
        const uint4 v = pa[t];
        qa.w[4 * t] = v.x;
        qa.w[4 * t + 1] = v.y;
        qa.w[4 * t + 2] = v.z;
        qa.w[4 * t + 3] = v.w;
      }
    }
    if (pb) {
#pragma unroll
      for (u32 t = 0; t < Layout<CB>::Q; t++) {
        const uint4 v = pb[t];
        qb.w[4 * t] = v.x;
        qb.w[4 * t + 1] = v.y;
        qb.w[4 * t + 2] = v.z;
        qb.w[4 * t + 3] = v.w;
      }
    }
    if (pp) e = *pp;

    // ---- compute phase ----
    bool sdone = false, salive = false;  // a search finished; alive = survived len chars
    u32 sdead = 0;                        // dead after this many chars
    if (op == OP_PRE) {
      if (COUNT) t_lines++;
      l = e.x;
      r = e.y;
      i = m;
      if (l > r) {
        sdone = true;
        sdead = m;
      } else if (i >= len) {
        sdone = salive = true;
      } else {
        op = OP_STEP;
      }
    } else if (op == OP_STEP) {
      if (COUNT) {
        t_lines += chunk_line<CB>(l) == chunk_line<CB>(r + 1u) ? 1u : 2u;
        t_steps++;
      }
      const u32 base = c_of(ix, c);
      l = base + rank_chunk<CB>(qa, l - ca * COLS);
      r = base + rank_chunk<CB>(qb, (r + 1u) - cb * COLS) - 1u;
      i++;
      if (l > r) {
        sdone = true;
        sdead = i;
      } else if (i >= len) {
        sdone = salive = true;
      }
    } else if (op == OP_STREAM) {
      if (COUNT) {
        t_lines++;
        t_steps++;
      }
      const u32 o = pnode - ca * COLS;
      if (bit_chunk<CB>(qa, o)) {
        pnode = c_of(ix, c) + rank_chunk<CB>(qa, o);
        answer(pnode);
        op = OP_NONE;
      } else {
        gch = ca;
        glim = o;
        op = OP_GSCAN;
      }
    } else if (op == OP_GSCAN) {
      if (COUNT) t_lines++;
      const int p = last_one_upto<CB>(qa, glim);
      if (p >= 0 || gch == 0) {
        const u32 g = p >= 0 ? gch * COLS + static_cast<u32>(p) : 0u;
        if (g == pnode) {
          prev = false;
          answer(kNone);
          op = OP_NONE;
          start_cert(j - 1 + k - 1);
        } else {
          gpos = g;
          op = OP_GEDGE;
        }
      } else {
        gch--;
        glim = COLS - 1;
      }
    } else if (op == OP_GEDGE) {
      if (COUNT && chunk_line<CB>(gpos) != chunk_line<CB>(pnode)) t_lines++;
      const u32 go = gpos - ca * COLS;
      op = OP_NONE;
      if (bit_chunk<CB>(qa, go)) {
        pnode = c_of(ix, c) + rank_chunk<CB>(qa, go);
        answer(pnode);
      } else {
        prev = false;
        answer(kNone);
        start_cert(j - 1 + k - 1);
      }
    }

    if (sdone) {
      op = OP_NONE;
      if (purpose == P_KMER) {
        if (salive) {
          prev = true;
          pnode = l;
          answer(l);
        } else {
          prev = false;
          answer(kNone);
          start_cert(s + sdead - 1);
        }
      } else if (purpose == P_CERTP) {
        if (salive) {
          kpos = cp;  // the k-mer at cp is present; answered when j reaches it
          knode = l;
        } else {
          const u32 lo = cp + sdead > k ? cp + sdead - k : 0u;
          clo = lo > j ? lo : j;
          chi = cp;
          if (lo > j && cp - lo + 2 <= k) {
            // the k-mers j..lo-1 all contain [lo-1, cp]
            const u32 q = lo - 1;
            start_search(P_CERTL, q, cp - q + 1, window2(rd.packed, cbase + q));
          }
        }
      } else {  // P_CERTL
        if (!salive) clo = j;
      }
    }
  }

  if (COUNT) {
    u64 v[4] = {t_lines, t_steps, t_kmers, t_search};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      u64 x = v[q];
      for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
      if ((threadIdx.x & 63) == 0) atomicAdd(ctr + q, static_cast<unsigned long long>(x));
    }
  }
}

template <int CB, bool NBREAK, bool STREAM, bool CERT, bool COUNT>
static int launch_v2_5(const sbwt_index* ix, const ReadsDev& rd, u64 bound, u64* out, hipStream_t s) {
  const u32 bt = 256;
  u64 grid = (bound + bt - 1) / bt;
  const u64 cap = static_cast<u64>(ix->num_cus ? ix->num_cus : 256) * 8;
  if (grid > cap) grid = cap;
  if (grid == 0) grid = 1;
  k_zero_counter<<<1, 1, 0, s>>>(ix->d_item_ctr);
  k_reads_v2<CB, NBREAK, STREAM, CERT, COUNT><<<grid, bt, 0, s>>>(dev_view(ix), rd, ix->d_item_ctr, out,
                                                                 ix->d_counters);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

template <int CB, bool NBREAK, bool STREAM, bool CERT>
static int launch_v2_4(const sbwt_index* ix, const ReadsDev& rd, u64 bound, u64* out, bool count, hipStream_t s) {
  return count ? launch_v2_5<CB, NBREAK, STREAM, CERT, true>(ix, rd, bound, out, s)
               : launch_v2_5<CB, NBREAK, STREAM, CERT, false>(ix, rd, bound, out, s);
}

template <int CB>
static int launch_v2(const sbwt_index* ix, const ReadsDev& rd, u64 bound, u64* out, bool nbreak, bool stream,
                     bool cert, bool count, hipStream_t s) {
  if (nbreak) {
    if (stream) return cert ? launch_v2_4<CB, true, true, true>(ix, rd, bound, out, count, s)
                            : launch_v2_4<CB, true, true, false>(ix, rd, bound, out, count, s);
    return cert ? launch_v2_4<CB, true, false, true>(ix, rd, bound, out, count, s)
                : launch_v2_4<CB, true, false, false>(ix, rd, bound, out, count, s);
  }
  if (stream) return cert ? launch_v2_4<CB, false, true, true>(ix, rd, bound, out, count, s)
                          : launch_v2_4<CB, false, true, false>(ix, rd, bound, out, count, s);
  return cert ? launch_v2_4<CB, false, false, true>(ix, rd, bound, out, count, s)
              : launch_v2_4<CB, false, false, false>(ix, rd, bound, out, count, s);
}

int search_reads_v2(const sbwt_index* ix, const u64* packed, const u64* nmask, const u64* rs, const u64* off,
                    const u64* work, const u32* items, u64 rows, u64 bound, u32 seg, u64* out, bool nbreak,
                    bool stream, bool cert, bool count, hipStream_t s) {
  const ReadsDev rd{packed, nmask, rs, off, work, items, rows, seg};
  switch (ix->layout) {
    case 16: return launch_v2<16>(ix, rd, bound, out, nbreak, stream, cert, count, s);
    case 32: return launch_v2<32>(ix, rd, bound, out, nbreak, stream, cert, count, s);
    default: return launch_v2<64>(ix, rd, bound, out, nbreak, stream, cert, count, s);
  }
}

}  // namespace sbwt
