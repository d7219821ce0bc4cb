// Persistent uniform-step read search of libsbwt_gpu_amd.
//
// Same answers as searching every k-mer independently (the reference's
// search_kernel.hpp:61-104 applied to every position of every row), with
// less work:
//
// * Streaming: after k-mer x at node j is found, x[1..k-1]c is the node
//   C[c] + rank_c(g), g the start of j's suffix group, if g has the c-edge.
// * Dead-substring certificates: when a search of a string w ends with an
//   empty interval, no node label ends with w, so no k-mer of the index
//   contains w (every substring of an indexed k-mer ends some node label:
//   walk the k-mer's incoming edges backwards), and every k-mer of the row
//   that contains w is absent without a search. After a k-mer is found
//   absent with a dead string ending at read position d (typically a
//   sequencing error), the kernel searches the k-mer starting at d (dead
//   after L chars, certifying the k-mers that contain [d, d+L-1]) and the
//   string ending at d that the remaining k-mers share, so a run of up to k
//   absent k-mers costs two short searches instead of k. When d lies in the
//   row's last k - 1 characters (no k-mer starts there), the next k-mers up to
//   a start q all contain [q, d], and one search of that string answers them.
//
// Execution: every lane is a small state machine, and every loop iteration
// every busy lane performs exactly one step of the same shape: two rank
// reads in one row, rank_c(a) and rank_c(b). A search step narrows [l, r]
// with (a, b) = (l, r + 1); a streaming step is the narrowing of [j, j]
// (the c-edge exists iff rank_c(j + 1) > rank_c(j)); the edge check at a
// group start g is the narrowing of [g, g]. A presearch lookup and the rare
// backwards scan for a group start ride on the same load. So the lanes of a
// wave execute the same instructions whatever they are doing, instead of
// one lane running a k-step search while the others wait. Lanes take new
// work items (segments of rows) from their wave's pool, which claims 64
// items at a time from a global counter; once the counter runs out, lanes
// left without work take half of what other lanes of their wave still hold
// (donate()).
#include "answers.hpp"
#include "index.hpp"

namespace sbwt {

namespace {

enum : u32 { M_IDLE = 0, M_SEARCH, M_STREAM, M_EDGE, M_GSCAN };
// P_CERTS: a P_CERTP search from the dead end of a k-mer search that died
// early (see start_cert)
enum : u32 { P_KMER = 0, P_CERTP, P_CERTL, P_CERTS };
// a purpose flag: the search's window is still to be loaded
constexpr u32 kWinPending = 4;
constexpr u32 kNoPos = 0xFFFFFFFFu;
// Work items a wave claims per atomic on the shared counter (at least 64,
// so one claim serves every lane of a refill). One atomic per lane or per
// refill serialises on the counter's L2 line and stalls the waves.
constexpr u64 kItemBatch = 64;
static_assert(kItemBatch >= 64, "one claim serves every lane of a wave that asks at once");

struct ReadsDev {
  const u64* packed;
  const u64* nmask;
  const u64* row_starts;
  const u64* offsets;
  const u64* work;
  const u32* items;
  const u8* maybe;
  u64 rows;
  u32 seg;
  u32 nbreak, stream, cert;
  u32 vec_out;  // out is 16-byte aligned (full answer groups go out as two 16-byte stores)
  // 1: out is the caller's answer array minus one slot, so that an array
  // 8 bytes off a 16-byte boundary still takes 16-byte stores; answer q goes
  // to out[q + 1]. A group holding slot 0 is never full (the lane never
  // writes slot 0), so it goes out slot by slot and nothing before the
  // caller's array is written.
  u32 oshift;
  const uint4* runs;  // run items {row, first k-mer, k-mers}, or null: segments of the work prefix
  const unsigned long long* run_count;
  u64 run_cap;  // items the run buffer holds
};

// The kernel's variants. FX_ANY reads every option from ReadsDev at run
// time. The two read-search defaults fix them at compile time, which drops
// their branches and frees the scalar registers that would hold them:
// FX_SEGMENTS streams with certificates over segment items given by an item
// array, breaks k-mers at non-ACGT characters, has no prefilter flags and
// writes 16-byte-aligned answers; FX_RUNS does the same over run items
// (which hold no non-ACGT window, so nothing breaks).
enum : int { FX_ANY = 0, FX_SEGMENTS = 1, FX_RUNS = 2 };

template <int FX>
struct Fx {
  __device__ __forceinline__ static bool nbreak(const ReadsDev& rd) {
    return FX == FX_SEGMENTS ? true : (FX == FX_RUNS ? false : rd.nbreak != 0);
  }
  __device__ __forceinline__ static bool stream(const ReadsDev& rd) { return FX != FX_ANY || rd.stream != 0; }
  __device__ __forceinline__ static bool cert(const ReadsDev& rd) { return FX != FX_ANY || rd.cert != 0; }
  __device__ __forceinline__ static bool vec(const ReadsDev& rd) { return FX != FX_ANY || rd.vec_out != 0; }
  __device__ __forceinline__ static const u8* maybe(const ReadsDev& rd) { return FX == FX_ANY ? rd.maybe : nullptr; }
  __device__ __forceinline__ static bool runs(const ReadsDev& rd) {
    return FX == FX_RUNS ? true : (FX == FX_SEGMENTS ? false : rd.runs != nullptr);
  }
};

__device__ __forceinline__ u64 window2(const u64* __restrict__ w, u64 pos) {
  const u64 b = pos << 1;
  return (w[b >> 6] << (b & 63)) | ((w[(b >> 6) + 1] >> (63 - (b & 63))) >> 1);
}

template <int CB>
__device__ __forceinline__ u32 chunk_line(u32 col) {
  return static_cast<u32>((static_cast<u64>(col / Layout<CB>::COLS) * CB) >> 7);
}

}  // namespace

__global__ void k_zero_counter(unsigned long long* c) { *c = 0ull; }

// Per-lane access tally of the counted instantiation.
struct UTally {
  u32 lines = 0, steps = 0, kmers = 0, searches = 0;
};

// The state of one work item being walked by one lane.
// Answer slots per lane: 64-byte groups for the 16-byte chunk layout (the
// larger chunks' kernels hold more registers, and a larger group costs them
// a wave of occupancy).
template <int CB>
constexpr u32 answer_slots() {
  return CB == 16 ? 8u : kAnswerSlots;
}

template <int CB, bool COUNT, int FX>
struct Walker {
  using F = Fx<FX>;
  static constexpr u32 COLS = Layout<CB>::COLS;
  using Ans = AnswerBufferT<answer_slots<CB>()>;
  // the item: k-mers [j, jend) of a row of rowk k-mers starting at character cbase
  u64 cbase = 0, obase = 0;
  u32 j = 0, jend = 0, rowk = 0;
  u32 jv = 0;  // a run item's first k-mer after its leading gap of N windows (0: none)
  bool live = true;
  // the current step
  u32 mode = M_IDLE, purpose = P_KMER;
  u32 l = 0, r = 0, i = 0, len = 0;  // search state (M_GSCAN: l = chunk, r = last offset)
  // top-aligned window of the searched string (while purpose has kWinPending:
  // the string's row-relative start, the window not loaded yet)
  u64 win = 0;
  bool prev = false;
  u32 pnode = 0, sc = 0;
  // certified-absent k-mers [clo, chi]; a k-mer at kpos already found at knode
  u32 clo = 1, chi = 0, kpos = kNoPos, knode = 0, cp = 0;
  u32 nN = kNoPos;  // first non-ACGT position (row-relative) at or after j, if >= j; kNoPos = unknown
  // The packed words wi and wi + 1 of the batch (characters 32 wi .. 32 wi
  // + 63), wi = the word of the k-mer at j: they hold both the next streamed
  // character (j + k - 1 <= 32 wi + 62) and the window of a search from j.
  // wn is word wi + 2, loaded when wi moves on by one word and not read
  // before wi moves again (32 k-mers later when streaming), so a wave never
  // waits for it: the load overlaps the steps instead of delaying one.
  // (wi = ~0 - 1: no words yet, as after a change of item; wi + 1 is never a
  // valid word.)
  u64 w0 = 0, w1 = 0, wn = 0, wi = ~0ull - 1;
  MaybeFlags mf;               // prefilter flags of the item's k-mers
  // One step's reads: chunks at columns a and b of one row, or a presearch
  // entry. Only the addresses live across the loads; update() recomputes the
  // rest from the walker's state.
  struct Step {
    const uint4* pa;
    const uint4* pb;
    bool pre;
  };
  struct Cols {
    u32 a, b, ca, cb, c;
    bool pre;
  };

  __device__ __forceinline__ Cols cols(const DevIndex& ix) const {
    Cols q;
    const bool busy = mode != M_IDLE, srch = mode == M_SEARCH, gscan = mode == M_GSCAN;
    q.pre = srch && i == 0 && ix.m && len >= ix.m;
    q.c = srch ? static_cast<u32>(win >> (62 - 2 * i)) & 3u : sc;
    q.a = !busy ? 0u : (srch ? l : (gscan ? l * COLS : (mode == M_EDGE ? l : pnode)));
    q.b = srch ? r + 1u : (gscan || !busy ? q.a : q.a + 1u);
    q.ca = q.a / COLS;
    q.cb = q.b / COLS;
    return q;
  }

  __device__ __forceinline__ void setup(const ReadsDev& rd, u64 it) {
    u64 row = 0;
    uint4 run = make_uint4(0, 0, 0, 0);
    if (F::runs(rd)) {
      run = rd.runs[it];
      row = run.x;
    } else if (FX == FX_SEGMENTS || rd.items) {
      row = rd.items[it];
    } else {
      u64 hi = rd.rows;
      while (hi - row > 1) {
        const u64 mid = (row + hi) >> 1;
        if (rd.work[mid] <= it) row = mid; else hi = mid;
      }
    }
    const u64 o0 = rd.offsets[row];
    rowk = static_cast<u32>(rd.offsets[row + 1] - o0);
    if (F::runs(rd)) {
      j = run.y;
      jv = run.z;
      jend = run.w;
    } else {
      j = static_cast<u32>(it - rd.work[row]) * rd.seg;
      jv = 0;
      jend = rowk - j < rd.seg ? rowk : j + rd.seg;
    }
    cbase = rd.row_starts[row];
    obase = o0 + rd.oshift;
    wi = ~0ull - 1;  // the words held belong to the last row (wn only if that row reached it)
    prev = false;
    clo = 1;
    chi = 0;
    kpos = kNoPos;
    nN = kNoPos;
    if (F::maybe(rd)) mf.reset(F::maybe(rd), obase - rd.oshift + j);
  }

  // k-mers [from, to) of the row of rk k-mers at characters cb, answers at
  // ob: the second half of another lane's item
  // (gap_end: the donor's jv, so a half that starts inside a run item's gap
  // answers the rest of the gap ~0)
  __device__ __forceinline__ void take(const ReadsDev& rd, u64 cb, u64 ob, u32 rk, u32 from, u32 to, u32 gap_end) {
    cbase = cb;
    obase = ob;
    wi = ~0ull - 1;
    rowk = rk;
    j = from;
    jv = gap_end;
    jend = to;
    live = true;
    prev = false;
    clo = 1;
    chi = 0;
    kpos = kNoPos;
    nN = kNoPos;
    if (F::maybe(rd)) mf.reset(F::maybe(rd), obase - rd.oshift + j);
  }

  // a search of `length` characters whose top-aligned window is `window`
  // (with pur | kWinPending: whose row-relative start is `window`)
  __device__ __forceinline__ void begin_search(const DevIndex& ix, UTally& t, u32 pur, u32 length, u64 window) {
    len = length;
    purpose = pur;
    win = window;
    i = 0;
    mode = M_SEARCH;
    if (COUNT) t.searches++;
    if (!(ix.m && length >= ix.m)) {
      l = 0;
      r = ix.C[4] - 1u;
    }
  }

  // the top-aligned window at offset o (< 64) of the held words w0, w1
  // (characters past w1 read as A)
  __device__ __forceinline__ u64 held_window(u32 o) const {
    return o >= 32 ? w1 << (2 * (o - 32)) : (o ? (w0 << (2 * o)) | (w1 >> (64 - 2 * o)) : w0);
  }

  // a search of the string of `length` characters at row position start:
  // its window comes from the held words when they hold the string (a
  // certificate's string mostly lies within a few characters of j), else
  // it is loaded by load_window()
  __device__ __forceinline__ void start_search(const DevIndex& ix, UTally& t, u32 pur, u32 start, u32 length) {
    const u64 at = cbase + start, w = wi << 5;
    if (wi != ~0ull - 1 && at >= w && at + length <= w + 64)
      begin_search(ix, t, pur, length, held_window(static_cast<u32>(at - w)));
    else
      begin_search(ix, t, pur | kWinPending, length, start);
  }

  // the answer for the k-mer at j (through the lane's answer buffer); moves to j + 1
  __device__ __forceinline__ void answer(u64* __restrict__ out, const ReadsDev& rd, Ans& ab, UTally& t,
                                         u64 v) {
    ab.put(out, F::vec(rd), obase + j, v, j + 1 >= jend);
    j++;
    if (COUNT) t.kmers++;
  }

  // v for the n k-mers from j on; moves to j + n
  __device__ __forceinline__ void answer_run(u64* __restrict__ out, const ReadsDev& rd, Ans& ab,
                                             UTally& t, u32 n, u64 v) {
    ab.put_run(out, F::vec(rd), obase + j, n, v, j + n >= jend);
    j += n;
    if (COUNT) t.kmers += n;
  }

  // A length at which a string with a sequencing error is absent from the
  // index with high probability: log4(n) + 2 (a random string of that
  // length occurs among n columns with probability about 1/16), at most k.
  __device__ __forceinline__ static u32 dead_len(const DevIndex& ix) {
    const u32 l = (33u - static_cast<u32>(__clz(ix.n))) / 2u + 2u;
    return l < ix.k ? l : ix.k;
  }

  // a k-mer before j was found absent with a dead string ending at char d
  // (shrt: by a search that died within dead_len characters)
  __device__ __forceinline__ void start_cert(const DevIndex& ix, const ReadsDev& rd, UTally& t, u32 d,
                                             bool shrt) {
    if (!F::cert(rd)) return;
    if (clo <= chi && chi >= j) return;  // a pending range still covers what follows
    if (d < j) return;                   // nothing ahead to certify
    if (d >= rowk) {
      // No k-mer starts at d (d is in the row's last k - 1 characters). The
      // k-mers j .. q, q <= rowk-1, all contain [s, d] for any s >= q (each
      // starts at or before s and ends after d). Far from d their searches
      // are long (each runs to d); q is the last start whose string [q, d]
      // is long enough to be absent from the index by chance, so one short
      // search can answer them all. When the row's last k-mer starts more
      // than dead_len + 2 characters before d, the string starts at
      // d + 1 - (dead_len + 2) instead: as certain, and shorter.
      const u32 q = min(rowk - 1, d + 1 - dead_len(ix));
      if (rowk == 0 || d + 1 < dead_len(ix) || q < j) return;
      const u32 ls = min(ix.k, dead_len(ix) + 2);
      const u32 s = max(q, d + 1 >= ls ? d + 1 - ls : 0u);
      chi = q;
      clo = q + 1;  // empty until the search proves it (P_CERTL then sets clo = j)
      start_search(ix, t, P_CERTL, s, d - s + 1);
      return;
    }
    // The search from d certifies only if it dies, which a string starting
    // with an error does within about dead_len characters. It stops alive
    // after dead_len + 2: that happens when d is past the error (a search
    // that died late because the error sat near its start), and running
    // on to k characters would cost a whole search for one known k-mer.
    cp = d;
    const u32 cap = dead_len(ix) + 2;
    start_search(ix, t, shrt ? P_CERTS : P_CERTP, d, cap < ix.k ? cap : ix.k);
  }

  // The window of a search started since the last step. Searches start at
  // four places (a k-mer from scratch, two certificate kinds, the shared
  // string after a dead certificate); loading their windows here, at one
  // place per iteration, keeps a wave from running the load code once per
  // place that some lane reached.
  __device__ __forceinline__ void load_window(const ReadsDev& rd) {
    if (purpose & kWinPending) {
      win = window2(rd.packed, cbase + static_cast<u32>(win));
      purpose &= ~kWinPending;
    }
  }

  // does the k-mer at j contain a non-ACGT character? nN = the first
  // non-ACGT position at or after j among the characters the item's k-mers
  // hold (up to jend + k - 1), or that bound when there is none. The scan
  // loads four mask words at a time before testing any (one wait per 256
  // characters instead of one per word: a 150-character read is one round
  // trip), and stops at the item's end, not the row's.
  __device__ __forceinline__ bool has_n(const ReadsDev& rd, u32 k) {
    if (nN == kNoPos || nN < j) {
      const u32 lim = jend + k - 1;
      u64 p = cbase + j;
      const u64 end = cbase + lim;
      nN = lim;
      while (p < end) {
        const u64 w0 = p >> 6, wl = (end - 1) >> 6;
        u64 x[4];
#pragma unroll
        for (u32 t = 0; t < 4; t++) x[t] = w0 + t <= wl ? rd.nmask[w0 + t] : 0ull;
        x[0] &= ~0ull >> (p & 63);
        u32 hit = 4;
#pragma unroll
        for (u32 t = 4; t-- > 0;)
          if (x[t]) hit = t;
        if (hit < 4) {
          const u64 q = ((w0 + hit) << 6) + static_cast<u64>(__clzll(x[hit]));
          if (q < end) nN = static_cast<u32>(q - cbase);
          break;
        }
        p = (w0 + 4) << 6;
      }
    }
    return nN <= j + k - 1;
  }

  // Answer what needs no index access, then set up one step.
  // (nk: the number of k-mers of the batch, the length of rd.maybe)
  __device__ __forceinline__ void advance(const DevIndex& ix, const ReadsDev& rd, u64 nk, u64* __restrict__ out,
                                          Ans& ab, UTally& t) {
    const u32 k = ix.k;
    if (j < jv) {  // a run item's gap: answered ~0 before the search
      if (COUNT) t.kmers += (jend < jv ? jend : jv) - j;
      j = jend < jv ? jend : jv;
    }
    while (mode == M_IDLE && j < jend) {
      // the packed words of the k-mer at j, loaded before the non-ACGT scan
      // below waits for its mask words, so that a new item's two loads
      // overlap
      const u64 at = cbase + j;
      const u64 wt = at >> 5;
      if (wt != wi) {
        if (wt == wi + 1) {
          w0 = w1;
          w1 = wn;
        } else {
          w0 = rd.packed[wt];
          w1 = rd.packed[wt + 1];
        }
        // (only when the row reaches that word: the batch's packed array
        // ends one word after its last character's word)
        wn = ((wt + 2) << 5) <= cbase + rowk + k - 2 ? rd.packed[wt + 2] : 0ull;
        wi = wt;
      }
      // a run of k-mers answered ~0 without a search: certified absent,
      // holding a non-ACGT character (every k-mer from j up to that
      // character holds it), or filtered out
      u32 run = 0;
      if (j >= clo && j <= chi) run = chi + 1 - j;
      else if (F::nbreak(rd) && has_n(rd, k)) run = nN + 1 - j;
      else if (F::maybe(rd)) run = mf.zeros(F::maybe(rd), nk, obase - rd.oshift + j);
      const bool known = j == kpos;  // found by a certificate search
      if (run || known) {
        prev = !run;
        pnode = knode;
        answer_run(out, rd, ab, t, run ? (run < jend - j ? run : jend - j) : 1u, run ? kNone : knode);
      } else {
        const u32 o = static_cast<u32>(at & 31);  // j's offset in w0
        if (F::stream(rd) && prev) {
          const u32 q = o + k - 1;  // the new character, at most 62
          sc = static_cast<u32>((q < 32 ? w0 >> (62 - 2 * q) : w1 >> (126 - 2 * q))) & 3u;
          mode = M_STREAM;
        } else {
          // the window of the k-mer at j, from the two words
          begin_search(ix, t, P_KMER, k, held_window(o));
        }
      }
    }
  }

  // The step's addresses (idle walkers read chunk 0 of row 0 and keep their
  // state). A presearch lookup reads the 16-byte block holding its entry; for
  // chunks of 32 or 64 bytes it reads that block again for the other parts.
  __device__ __forceinline__ Step prepare(const DevIndex& ix) const {
    const Cols q = cols(ix);
    const uint4* rowp = row_of(ix, mode == M_GSCAN ? kGroupRow : (mode != M_IDLE ? q.c : 0u));
    Step st;
    st.pre = q.pre;
    st.pa = q.pre ? reinterpret_cast<const uint4*>(ix.pre) + ((win >> (64 - 2 * ix.m)) >> 1)
                  : rowp + static_cast<u64>(q.ca) * Layout<CB>::Q;
    st.pb = q.pre ? st.pa : rowp + static_cast<u64>(q.cb) * Layout<CB>::Q;
    return st;
  }

  __device__ __forceinline__ static void fetch(const Step& st, Chunk<CB>& qa, Chunk<CB>& qb) {
    // a presearch lookup reads its one 16-byte block for every part (the
    // later parts are unused and must not step past the table)
#pragma unroll
    for (u32 t = 0; t < Layout<CB>::Q; t++) {
      // the second chunk only where it differs from the first (a streaming
      // step's j and j + 1 share a chunk but at its end; so does a narrow
      // interval): no second request for the same line
      const uint4 va = st.pa[st.pre ? 0 : t];
      uint4 vb = va;
      if (st.pb != st.pa) vb = st.pb[t];
      qa.w[4 * t] = va.x;
      qa.w[4 * t + 1] = va.y;
      qa.w[4 * t + 2] = va.z;
      qa.w[4 * t + 3] = va.w;
      qb.w[4 * t] = vb.x;
      qb.w[4 * t + 1] = vb.y;
      qb.w[4 * t + 2] = vb.z;
      qb.w[4 * t + 3] = vb.w;
    }
  }

  // Apply the step: the common outcomes by selects, the rare ones in one branch.
  // skip_short: start no certificate after a search that died early (the
  // wave's recent P_CERTS searches mostly ended alive); returns +1 / -1 when
  // a P_CERTS search ends alive / dead, else 0.
  __device__ __forceinline__ int update(const DevIndex& ix, const ReadsDev& rd, u64* __restrict__ out,
                                        Ans& ab, UTally& t, const Chunk<CB>& qa, const Chunk<CB>& qb,
                                        bool skip_short) {
    const u32 k = ix.k;
    const Cols st = cols(ix);
    const u64 pi = st.pre ? win >> (64 - 2 * ix.m) : 0;
    const bool busy = mode != M_IDLE, srch = mode == M_SEARCH, gscan = mode == M_GSCAN;
    const bool edge = mode == M_EDGE, strm = mode == M_STREAM;
    const u32 oa = st.a - st.ca * COLS, ob = st.b - st.cb * COLS;
    const u32 ra = rank_chunk<CB>(qa, oa), rb = rank_chunk<CB>(qb, ob);
    const u32 base = c_of(ix, st.c);
    if (COUNT) {
      if (st.pre) t.lines++;
      else if (srch) t.lines += chunk_line<CB>(st.a) == chunk_line<CB>(st.b) ? 1u : 2u;
      else if (edge) t.lines += chunk_line<CB>(st.a) != chunk_line<CB>(pnode) ? 1u : 0u;
      else if (busy) t.lines++;
      if (strm || (srch && !st.pre)) t.steps++;
    }
    // search: the narrowed interval
    const u32 sl = st.pre ? ((pi & 1) ? qa.w[2] : qa.w[0]) : base + ra;
    const u32 sr = st.pre ? ((pi & 1) ? qa.w[3] : qa.w[1]) : base + rb - 1u;
    const u32 si = st.pre ? ix.m : i + 1;
    const bool sdead = srch && sl > sr, salive = srch && sl <= sr && si >= len;
    // stream / edge: the narrowing of [a, a]
    const bool found = (strm || edge) && rb > ra;
    const u32 node = srch ? sl : base + ra;
    const bool kdone = (srch && purpose == P_KMER && (sdead || salive)) || found || (edge && !found);
    const bool kfound = salive || found;
    if (srch) {
      l = sl;
      r = sr;
      i = si;
    }
    bool emit = kdone;
    u64 ev = kfound ? static_cast<u64>(node) : kNone;
    if (kdone) {
      prev = kfound;
      pnode = node;
    }
    const bool settled = kdone || (srch && (sdead || salive)) || (strm && !found);
    mode = settled ? M_IDLE : mode;
    // rarer outcomes: certificates, group scans
    const bool kfail = kdone && !kfound;
    const bool cdone = srch && purpose != P_KMER && (sdead || salive);
    const int cev = cdone && purpose == P_CERTS ? (salive ? 1 : -1) : 0;
    if (emit) answer(out, rd, ab, t, ev);
    if (kfail || cdone || gscan || (strm && !found)) {
      if (kfail) {
        // a k-mer search of the k-mer at j died after i characters; a
        // streaming or edge step failed on the k-mer's last character
        const bool shrt = srch && i < dead_len(ix);
        if (!(shrt && skip_short)) start_cert(ix, rd, t, srch ? j - 1 + i - 1 : j - 1 + k - 1, shrt);
      } else if (cdone) {
        if (purpose == P_CERTP || purpose == P_CERTS) {
          if (salive) {
            if (len == k) {  // the k-mer at cp is present; answered when j reaches it
              kpos = cp;
              knode = l;
            }
          } else {
            const u32 lo = cp + i > k ? cp + i - k : 0u;
            clo = lo > j ? lo : j;
            chi = cp;
            if (lo > j && cp - lo + 2 <= k) start_search(ix, t, P_CERTL, lo - 1, cp - lo + 2);  // [lo-1, cp]
          }
        } else if (sdead) {  // P_CERTL: the k-mers j..lo-1 all contain [lo-1, cp]
          clo = j;
        }
      } else if (strm) {
        // no c-edge at j: find j's group start g (column 0 starts a group)
        mode = M_GSCAN;
        l = st.ca;
        r = oa;
      } else {  // M_GSCAN
        if (COUNT) t.lines++;
        const int p = last_one_upto<CB>(qa, r);
        if (p >= 0 || l == 0) {
          const u32 g = p >= 0 ? l * COLS + static_cast<u32>(p) : 0u;
          if (g == pnode) {  // j starts its group and has no c-edge
            mode = M_IDLE;
            prev = false;
            answer(out, rd, ab, t, kNone);  // (rare: a second call site)
            start_cert(ix, rd, t, j - 1 + k - 1, false);
          } else {
            mode = M_EDGE;
            l = g;
          }
        } else {
          l--;
          r = COLS - 1;
        }
      }
    }
    return cev;
  }
};

// The wave's pool of claimed work items (the same in every lane).
struct ItemPool {
  u64 next = 0, end = 0;
};

// Lanes whose walker needs work take the next items from the wave's pool,
// which claims kItemBatch items at a time from the global counter.
template <class W>
__device__ __forceinline__ void refill(W& w, ItemPool& pool, unsigned long long* item_ctr, u64 total_items,
                                       const ReadsDev& rd, u32 lane) {
  const bool want = w.live && w.mode == M_IDLE && w.j >= w.jend;
  const u64 wm = __ballot(want);
  if (!wm) return;
  const u64 nw = static_cast<u64>(__popcll(wm));
  const u64 rank = static_cast<u64>(__popcll(wm & ((1ull << lane) - 1)));
  const u64 avail = pool.end - pool.next;
  u64 it;
  if (avail >= nw) {
    it = pool.next + rank;
    pool.next += nw;
  } else {
    const u32 leader = static_cast<u32>(__ffsll(static_cast<unsigned long long>(wm)) - 1);
    u64 nb = 0;
    if (lane == leader) nb = atomicAdd(item_ctr, static_cast<unsigned long long>(kItemBatch));
    nb = __shfl(nb, static_cast<int>(leader));
    it = rank < avail ? pool.next + rank : nb + (rank - avail);
    pool.next = nb + (nw - avail);
    pool.end = nb + kItemBatch;
  }
  if (want) {
    if (it >= total_items) w.live = false;
    else w.setup(rd, it);
  }
}

// position of the n-th (from 0) set bit of m, which has more than n
__device__ __forceinline__ u32 select_bit(u64 m, u32 n) {
  u32 pos = 0;
#pragma unroll
  for (u32 wd = 32; wd; wd >>= 1) {
    const u32 c = static_cast<u32>(__popcll(m & ((1ull << wd) - 1)));
    if (n >= c) {
      n -= c;
      m >>= wd;
      pos += wd;
    }
  }
  return pos;
}

// Once the work items run out, lanes of a wave that have nothing left take
// the second half of the k-mers left to other lanes (pairs in lane order).
// A lane's last item would otherwise run alone while the rest of its wave
// idles. The half starts with a search of about k - m steps, so a split pays
// when the donor's remaining work is longer than that:
// * a streaming lane spends about one step per k-mer: it gives from
//   kSplitMin k-mers on (on the 100 Mbp read batch the kernel took 4.33 ms
//   without splits, 4.21 ms with kSplitMin 32 or 48, 4.25 ms with 16);
// * a lane searching its k-mer from scratch (the previous one was absent and
//   no certificate covered this one: an error in a row's first characters,
//   a certificate search that ran into a chance match) may need a whole
//   search for each k-mer left: it gives from kSplitSearchMin on. These lanes
//   set the length of a wave's tail. With them the batch's lane-iterations
//   fall from 209 to 171 per read (151 to 159 busy) and the kernel from 3.61
//   to 3.41 ms; 2, 4 and 8 measure alike.
constexpr u32 kSplitMin = 48;
constexpr u32 kSplitSearchMin = 4;

template <class W>
__device__ __forceinline__ void donate(W& w, const ReadsDev& rd, u32 lane) {
  const bool idle = !w.live && w.mode == M_IDLE;
  const u64 im = __ballot(idle);
  if (!im) return;
  const u32 rem = w.live && w.j < w.jend ? w.jend - w.j - 1 : 0u;  // after the k-mer at j
  const bool searching = !w.prev && w.mode == M_SEARCH && w.purpose == P_KMER;
  const bool give = rem >= kSplitMin || (searching && rem >= kSplitSearchMin);
  const u64 gm = __ballot(give);
  if (!gm) return;
  const u32 ni = static_cast<u32>(__popcll(im)), ng = static_cast<u32>(__popcll(gm));
  const u32 np = ni < ng ? ni : ng;
  const u64 below = (1ull << lane) - 1;
  const u32 grank = static_cast<u32>(__popcll(gm & below)), irank = static_cast<u32>(__popcll(im & below));
  const u32 mid = w.j + 1 + (rem >> 1);
  const bool recv = idle && irank < np;
  const int src = static_cast<int>(recv ? select_bit(gm, irank) : lane);
  const u64 cb = __shfl(w.cbase, src);
  const u64 ob = __shfl(w.obase, src);
  const u32 rk = __shfl(w.rowk, src);
  const u32 md = __shfl(mid, src);
  const u32 je = __shfl(w.jend, src);
  const u32 jv = __shfl(w.jv, src);
  if (give && grank < np) w.jend = mid;
  if (recv) w.take(rd, cb, ob, rk, md, je, jv);
}

// Each lane walks P work items at once; every iteration issues all P steps'
// reads before using any, so P dependent chains are in flight per lane.
// Six waves per SIMD for 16-byte chunks (80 VGPRs; the scalar registers
// allow no more): unbounded, the compiler took 82-84 and the kernel ran at
// five. For 32-byte chunks the same bound spills VGPRs to scratch, and five
// waves without spills measure faster (3.25 against 3.37 ms per 1M reads).
template <int CB, bool COUNT, int P, int FX>
__global__ __launch_bounds__(256, CB == 16 ? 6 : 1) void k_reads_uniform(DevIndex ix, ReadsDev rd,
                                                       unsigned long long* __restrict__ item_ctr,
                                                       u64* __restrict__ out, unsigned long long* ctr) {
  using W = Walker<CB, COUNT, FX>;
  static_assert(P == 1, "one LDS answer buffer per lane");
  __shared__ u64 obuf[answer_slots<CB>() * kAnswerStride];
  typename W::Ans ab{obuf + threadIdx.x};
  const u32 lane = threadIdx.x & 63;
  const u64 total_items = Fx<FX>::runs(rd) ? min(static_cast<u64>(*rd.run_count), rd.run_cap) : (rd.rows ? rd.work[rd.rows] : 0);
  const u64 nk = Fx<FX>::maybe(rd) && rd.rows ? rd.offsets[rd.rows] : 0;
  W w[P];
  UTally t;
  ItemPool pool;
  // The wave's tally of recent P_CERTS outcomes (alive minus dead, within
  // +-64). A search that dies early either started at a sequencing error
  // in the first characters of its k-mer (it then ran past the error, and
  // the certificate search from its dead end finds a present string) or
  // ran into a random string (the certificate search dies and certifies).
  // Read batches mostly give the first, negative-heavy batches the second.
  int cscore = 0;
  for (;;) {
    bool active = false;
#pragma unroll
    for (int p = 0; p < P; p++) {
      refill(w[p], pool, item_ctr, total_items, rd, lane);
      donate(w[p], rd, lane);
      w[p].advance(ix, rd, nk, out, ab, t);
      w[p].load_window(rd);
      active = active || w[p].mode != M_IDLE || w[p].live;
    }
    if (!__any(active)) break;
    typename W::Step st[P];
    Chunk<CB> qa[P], qb[P];
#pragma unroll
    for (int p = 0; p < P; p++) st[p] = w[p].prepare(ix);
#pragma unroll
    for (int p = 0; p < P; p++) W::fetch(st[p], qa[p], qb[p]);
#pragma unroll
    for (int p = 0; p < P; p++) {
      const int ev = w[p].update(ix, rd, out, ab, t, qa[p], qb[p], cscore > 0);
      cscore += __popcll(__ballot(ev > 0)) - __popcll(__ballot(ev < 0));
      cscore = cscore > 64 ? 64 : (cscore < -64 ? -64 : cscore);
    }
  }
  if (COUNT) {
    u64 v[4] = {t.lines, t.steps, t.kmers, t.searches};
#pragma unroll
    for (int q = 0; q < 4; q++) {
      u64 x = v[q];
      for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
      if (lane == 0) atomicAdd(ctr + q, static_cast<unsigned long long>(x));
    }
  }
}

// Work items per lane. Two (occupancy 5 instead of 7, two dependent chains per
// lane) measured slower on the 100 Mbp read batch: 6.6 ms against 5.8 ms.
constexpr int kWalkersPerLane = 1;

template <int CB, bool COUNT, int FX>
static int launch_uniform(const sbwt_index* ix, const ReadsDev& rd, u64 bound, u64* out,
                          unsigned long long* item_ctr, hipStream_t s) {
  const u32 bt = kAnswerStride;  // the answer buffer's LDS layout assumes this block size
  constexpr int P = kWalkersPerLane;
  int per = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_reads_uniform<CB, COUNT, P, FX>, bt, 0) != hipSuccess ||
      per < 1)
    per = 4;
  u64 grid = (bound + bt - 1) / bt;
  const u64 cap = static_cast<u64>(ix->num_cus ? ix->num_cus : 256) * static_cast<u64>(per);
  if (grid > cap) grid = cap;
  if (grid == 0) grid = 1;
  unsigned long long* ctr = item_ctr ? item_ctr : ix->d_item_ctr;
  k_zero_counter<<<1, 1, 0, s>>>(ctr);
  k_reads_uniform<CB, COUNT, P, FX><<<grid, bt, 0, s>>>(dev_view(ix), rd, ctr, out, ix->d_counters);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

template <int CB>
static int launch_uniform_fx(const sbwt_index* ix, const ReadsDev& rd, u64 bound, u64* out,
                             unsigned long long* item_ctr, bool count, int fx, hipStream_t s) {
  switch (fx) {
    case FX_SEGMENTS: return count ? launch_uniform<CB, true, FX_SEGMENTS>(ix, rd, bound, out, item_ctr, s)
                                   : launch_uniform<CB, false, FX_SEGMENTS>(ix, rd, bound, out, item_ctr, s);
    case FX_RUNS: return count ? launch_uniform<CB, true, FX_RUNS>(ix, rd, bound, out, item_ctr, s)
                               : launch_uniform<CB, false, FX_RUNS>(ix, rd, bound, out, item_ctr, s);
    default: return count ? launch_uniform<CB, true, FX_ANY>(ix, rd, bound, out, item_ctr, s)
                          : launch_uniform<CB, false, FX_ANY>(ix, rd, bound, out, item_ctr, s);
  }
}

int search_reads_persistent(const sbwt_index* ix, const u64* packed, const u64* nmask, const u64* rs,
                            const u64* off, const u64* work, const u32* items, u64 rows, u64 bound, u32 seg,
                            const u8* maybe, u64* out, bool nbreak, bool stream, bool cert, bool count,
                            const RunList& runs, unsigned long long* item_ctr, hipStream_t s) {
  const uintptr_t oa = reinterpret_cast<uintptr_t>(out) & 15;
  const u32 oshift = oa == 8 ? 1u : 0u;
  const ReadsDev rd{packed, nmask, rs, off, work, items, maybe, rows, seg,
                    nbreak ? 1u : 0u, stream ? 1u : 0u, cert ? 1u : 0u,
                    oa == 0 || oa == 8 ? 1u : 0u, oshift, runs.runs, runs.count, runs.cap};
  out = reinterpret_cast<u64*>(reinterpret_cast<uintptr_t>(out) - 8 * oshift);
  const bool fixed = stream && cert && !maybe && rd.vec_out;
  const int fx = !fixed ? FX_ANY : (runs.runs ? FX_RUNS : (nbreak && items ? FX_SEGMENTS : FX_ANY));
  switch (ix->layout) {
    case 16: return launch_uniform_fx<16>(ix, rd, bound, out, item_ctr, count, fx, s);
    case 32: return launch_uniform_fx<32>(ix, rd, bound, out, item_ctr, count, fx, s);
    default: return launch_uniform_fx<64>(ix, rd, bound, out, item_ctr, count, fx, s);
  }
}

}  // namespace sbwt
