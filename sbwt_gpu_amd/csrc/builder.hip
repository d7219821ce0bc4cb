// GPU SBWT construction (no reference counterpart: the reference reads four
// bit vectors made by an external tool, include/matrix_sbwt.hpp:70-110).
//
// Definition followed (Alanko, Puglisi, Vuohtoniemi, "Small searchable
// k-spectra via subset rank queries on the spectral Burrows-Wheeler
// transform"): the padded k-spectrum R is the k-mer set K plus, for every
// source k-mer x (no k-mer of K ends with x[0..k-2]), the nodes $^{k-i}x[0..i-1],
// i = 1..k-1, plus the root $^k. R is sorted colexicographically ('$' < A).
// Node j carries the set {c : suffix_{k-1}(R_j)c in R} if it is the first of
// its suffix group (the nodes sharing the last k-1 characters), else {}.
//
// Encoding: a node is (key, len): key holds character p of the padded label
// in bits 2p..2p+1 ('$' codes as 0) and len is the number of real
// characters, so colex order is (key, len) order. Every node d but the root
// has one incoming edge, labelled by its last character, from the first
// node of the group whose suffix is d's first k-1 characters, which is
// lower_bound(R, ((key_d mod 4^(k-1)) << 2, len_d - 1)). Each node therefore
// sets exactly one bit, and C[4] = n.
//
// Sorting K and the padded nodes: an MSD bucket partition of this file's own
// (8 key bits per level, placement from a per-block histogram and one
// exclusive scan, no global atomics) splits the keys until every bucket
// holds at most `chunk` items (2^28 by default); each bucket is then sorted
// and deduplicated by rocPRIM calls of at most that size. A window count of
// 3e9 (a 3 Gbp input) therefore never reaches a rocPRIM call whose item
// count exceeds 2^31.
//
// Always-on consistency checks between the stages (each failure returns
// SBWT_ERR_INTERNAL naming the stage):
//   * windows: the valid windows the partition placed = an independent count;
//   * sort:    the sum of a 64-bit mix of every item is the same before the
//              partition and after the bucket sorts (a lost, duplicated or
//              altered item changes it), and every bucket is non-decreasing;
//   * unique:  K (and the padded nodes) strictly increasing; its size = the
//              number of runs of the sorted keys; the mix sum over K = the
//              mix sum over the first item of every run;
//   * K:       up to 2^22 windows sampled evenly from the sequences are found
//              in K by binary search;
//   * merge:   every slot of R written once and R strictly increasing;
//   * rows:    C[4] = n (each node but the root sets exactly one edge bit);
//   * index:   the sampled k-mers are all found by the finished index's search.
#include <vector>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>
#include <rocprim/device/device_select.hpp>

#include "index.hpp"

namespace sbwt {

__host__ __device__ __forceinline__ u64 lowmask64(u32 bits) { return bits >= 64 ? ~0ull : ((1ull << bits) - 1ull); }

// Reverses the order of the 32 two-bit groups of x.
__device__ __forceinline__ u64 rev_pairs(u64 x) {
  x = ((x >> 2) & 0x3333333333333333ull) | ((x & 0x3333333333333333ull) << 2);
  x = ((x >> 4) & 0x0F0F0F0F0F0F0F0Full) | ((x & 0x0F0F0F0F0F0F0F0Full) << 4);
  return __builtin_bswap64(x);
}

__device__ __forceinline__ u64 window2(const u64* __restrict__ w, u64 pos) {
  const u64 b = pos << 1;
  return (w[b >> 6] << (b & 63)) | ((w[(b >> 6) + 1] >> (63 - (b & 63))) >> 1);
}
__device__ __forceinline__ u64 window1(const u64* __restrict__ w, u64 pos) {
  return (w[pos >> 6] << (pos & 63)) | ((w[(pos >> 6) + 1] >> (63 - (pos & 63))) >> 1);
}

// splitmix64's output function: the per-item mix of the multiset checks.
__device__ __forceinline__ u64 item_mix(u64 key, u32 len) {
  u64 x = key + static_cast<u64>(len) * 0x9E3779B97F4A7C15ull;
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

// Sum of v over the wave, added to *dst by lane 0. Every lane must call it.
__device__ __forceinline__ void wave_add(unsigned long long* dst, u64 v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(static_cast<unsigned long long>(v), o, 64);
  if ((threadIdx.x & 63u) == 0 && v) atomicAdd(dst, static_cast<unsigned long long>(v));
}

__device__ __forceinline__ bool node_less(u64 ka, u32 la, u64 kb, u32 lb) { return ka < kb || (ka == kb && la < lb); }

// ---- item sources of a partition --------------------------------------------

// Window p of the concatenated sequences: valid if it holds no non-ACGT
// character and no sequence starts inside (p, p+k); key = its colex key.
struct WindowSrc {
  const u64* packed;
  const u64* nmask;
  const u64* bounds;
  u32 k;
  __device__ bool get(u64 p, u64& key, u32& len) const {
    const u64 bad = window1(nmask, p) >> (64 - k);
    const u64 crossing = (window1(bounds, p) << 1) >> (64 - k + 1 > 63 ? 63 : 64 - k + 1);
    len = 0;
    if (bad != 0 || (k > 1 && crossing != 0)) return false;
    key = rev_pairs(window2(packed, p)) & lowmask64(2 * k);
    return true;
  }
  // The window's k characters top-aligned (char 0 in bits 63:62).
  __device__ u64 top_aligned(u64 p) const { return window2(packed, p) & ~lowmask64(64 - 2 * k); }
};

// Items of an array of keys (and lengths, or none: length 0).
struct ArraySrc {
  const u64* keys;
  const u8* lens;
  __device__ bool get(u64 i, u64& key, u32& len) const {
    key = keys[i];
    len = lens ? lens[i] : 0u;
    return true;
  }
};

// ---- MSD bucket partition ------------------------------------------------------

constexpr u32 kPartThreads = 256;
constexpr u64 kPartTile = 65536;  // items per block of a partition pass
constexpr u32 kMaxBuckets = 256;

// Bucket of an item: `bits` key bits from `shift`, or its length (by_len).
__device__ __forceinline__ u32 bucket_of(u64 key, u32 len, u32 shift, u32 bits, bool by_len) {
  return by_len ? len : static_cast<u32>(key >> shift) & ((1u << bits) - 1u);
}

// Pass 1: per-block bucket counts, stored bucket-major (hist[b * nblocks +
// block]) so that one exclusive scan gives every block's start in every
// bucket. `mix` (level 0 only) gets the mix sum of the valid items.
template <class Src>
__global__ __launch_bounds__(kPartThreads) void k_part_hist(Src src, u64 n, u32 shift, u32 bits, bool by_len,
                                                            u64 nblocks, u64* __restrict__ hist,
                                                            unsigned long long* __restrict__ mix) {
  __shared__ u32 h[kMaxBuckets];
  const u32 nb = 1u << bits;
  for (u32 t = threadIdx.x; t < nb; t += blockDim.x) h[t] = 0;
  __syncthreads();
  const u64 lo = blockIdx.x * kPartTile;
  const u64 hi = lo + kPartTile < n ? lo + kPartTile : n;
  u64 acc = 0;
  for (u64 i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    u64 key;
    u32 len;
    if (src.get(i, key, len)) {
      atomicAdd(&h[bucket_of(key, len, shift, bits, by_len)], 1u);
      acc += item_mix(key, len);
    }
  }
  if (mix) wave_add(mix, acc);
  __syncthreads();
  for (u32 t = threadIdx.x; t < nb; t += blockDim.x) hist[t * nblocks + blockIdx.x] = h[t];
}

// Pass 2: each block places its items from its scanned starts (the order
// inside a bucket is arbitrary; the buckets are sorted afterwards).
template <class Src>
__global__ __launch_bounds__(kPartThreads) void k_part_scatter(Src src, u64 n, u32 shift, u32 bits, bool by_len,
                                                               u64 nblocks, const u64* __restrict__ offs,
                                                               u64* __restrict__ okeys, u8* __restrict__ olens) {
  __shared__ u32 cur[kMaxBuckets];
  __shared__ u64 base[kMaxBuckets];
  const u32 nb = 1u << bits;
  for (u32 t = threadIdx.x; t < nb; t += blockDim.x) {
    cur[t] = 0;
    base[t] = offs[t * nblocks + blockIdx.x];
  }
  __syncthreads();
  const u64 lo = blockIdx.x * kPartTile;
  const u64 hi = lo + kPartTile < n ? lo + kPartTile : n;
  for (u64 i = lo + threadIdx.x; i < hi; i += blockDim.x) {
    u64 key;
    u32 len;
    if (src.get(i, key, len)) {
      const u32 b = bucket_of(key, len, shift, bits, by_len);
      const u64 pos = base[b] + atomicAdd(&cur[b], 1u);
      okeys[pos] = key;
      if (olens) olens[pos] = static_cast<u8>(len);
    }
  }
}

// starts[b] = first slot of bucket b, starts[nb] = the total.
__global__ void k_bucket_starts(const u64* __restrict__ offs, const u64* __restrict__ hist, u64 nblocks, u32 nb,
                                u64* __restrict__ starts) {
  const u32 b = threadIdx.x;
  if (b < nb) starts[b] = offs[b * nblocks];
  if (b == nb - 1) starts[nb] = offs[b * nblocks + nblocks - 1] + hist[b * nblocks + nblocks - 1];
}

// Valid windows, counted apart from the partition.
__global__ void k_count_valid(WindowSrc src, u64 n, unsigned long long* __restrict__ count) {
  u64 c = 0;
  const u64 stride = gridDim.x * static_cast<u64>(blockDim.x);
  for (u64 p = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; p < n; p += stride) {
    u64 key;
    u32 len;
    c += src.get(p, key, len) ? 1 : 0;
  }
  wave_add(count, c);
}

// Check accumulators (device u64 words).
enum Acc : u32 {
  kMixSrc = 0,    // mix sum of the items entering the partition
  kMixSorted,     // mix sum of the sorted buckets
  kMixFirst,      // mix sum of the first item of every run of the sorted buckets
  kRuns,          // runs of equal items in the sorted buckets
  kUnsorted,      // adjacent pairs out of order inside a sorted bucket
  kMixOut,        // mix sum of the deduplicated output
  kOutBad,        // adjacent output pairs not strictly increasing, or unwritten slots
  kValid,         // independent count of the valid windows
  kMisses,        // sampled windows absent from K, or sampled k-mers the index does not find
  kSampled,       // sampled valid windows
  kAccWords
};

// Sorted bucket: mix sums, runs, order.
__global__ void k_seg_check(const u64* __restrict__ keys, const u8* __restrict__ lens, u64 n,
                            unsigned long long* __restrict__ acc) {
  u64 ms = 0, mf = 0, runs = 0, bad = 0;
  const u64 stride = gridDim.x * static_cast<u64>(blockDim.x);
  for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < n; i += stride) {
    const u64 key = keys[i];
    const u32 len = lens ? lens[i] : 0u;
    const u64 m = item_mix(key, len);
    ms += m;
    bool first = i == 0;
    if (!first) {
      const u64 pk = keys[i - 1];
      const u32 pl = lens ? lens[i - 1] : 0u;
      first = pk != key || pl != len;
      bad += node_less(key, len, pk, pl) ? 1 : 0;
    }
    if (first) {
      mf += m;
      runs++;
    }
  }
  wave_add(acc + kMixSorted, ms);
  wave_add(acc + kMixFirst, mf);
  wave_add(acc + kRuns, runs);
  wave_add(acc + kUnsorted, bad);
}

// Deduplicated output: strictly increasing, mix sum.
__global__ void k_out_check(const u64* __restrict__ keys, const u8* __restrict__ lens, u64 n,
                            unsigned long long* __restrict__ acc) {
  u64 ms = 0, bad = 0;
  const u64 stride = gridDim.x * static_cast<u64>(blockDim.x);
  for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < n; i += stride) {
    const u64 key = keys[i];
    const u32 len = lens ? lens[i] : 0u;
    ms += item_mix(key, len);
    if (i > 0) bad += node_less(keys[i - 1], lens ? lens[i - 1] : 0u, key, len) ? 0 : 1;
  }
  wave_add(acc + kMixOut, ms);
  wave_add(acc + kOutBad, bad);
}

__global__ void k_first_of_run(const u64* __restrict__ key, const u8* __restrict__ len, u64 n, u8* __restrict__ keep) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  keep[i] = (i == 0 || key[i] != key[i - 1] || len[i] != len[i - 1]) ? 1 : 0;
}

__device__ __forceinline__ u64 lower_bound_u64(const u64* __restrict__ a, u64 n, u64 v) {
  u64 lo = 0, hi = n;
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if (a[mid] < v) lo = mid + 1; else hi = mid;
  }
  return lo;
}
__device__ __forceinline__ u64 upper_bound_u64(const u64* __restrict__ a, u64 n, u64 v) {
  u64 lo = 0, hi = n;
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if (a[mid] <= v) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Windows 0, stride, 2*stride, ...: the valid ones are looked up in K and
// their top-aligned k-mers kept (compacted, any order) for the final check.
__global__ void k_sample_windows(WindowSrc src, u64 windows, u64 stride, u64 samples, const u64* __restrict__ K,
                                 u64 nk, u64* __restrict__ kmers, unsigned long long* __restrict__ acc) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  u64 miss = 0;
  if (i < samples && i * stride < windows) {
    const u64 p = i * stride;
    u64 key;
    u32 len;
    if (src.get(p, key, len)) {
      const u64 at = lower_bound_u64(K, nk, key);
      miss = (at < nk && K[at] == key) ? 0 : 1;
      const unsigned long long slot = atomicAdd(acc + kSampled, 1ull);
      kmers[slot] = src.top_aligned(p);
    }
  }
  wave_add(acc + kMisses, miss);
}

__global__ void k_count_absent(const u64* __restrict__ ans, u64 n, unsigned long long* __restrict__ acc) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  wave_add(acc + kMisses, (i < n && ans[i] == kNone) ? 1 : 0);
}

__global__ void k_mark_bounds(const u64* __restrict__ starts, u64 nseq, u64 total, u64* __restrict__ bounds) {
  const u64 s = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x + 1;
  if (s >= nseq) return;
  const u64 q = starts[s];
  if (q < total) atomicOr(reinterpret_cast<unsigned long long*>(bounds + (q >> 6)), 1ull << (63 - (q & 63)));
}

// Source k-mers: no y in K with (key_y >> 2) == (key_x mod 4^(k-1)).
__global__ void k_sources(const u64* __restrict__ K, u64 nk, u32 k, u8* __restrict__ is_src) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (i >= nk) return;
  const u64 want = K[i] & lowmask64(2 * (k - 1));
  const u64 lo = lower_bound_u64(K, nk, want << 2);
  is_src[i] = (lo < nk && (K[lo] >> 2) == want) ? 0 : 1;
}

// Padded nodes of source s: len = 1..k-1, key = (x mod 4^len) << 2(k-len);
// slot 0 of the output is the root.
__global__ void k_padded(const u64* __restrict__ src, u64 nsrc, u32 k, u64* __restrict__ pkey, u8* __restrict__ plen) {
  const u64 t = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  const u64 per = k - 1;
  if (t == 0) {
    pkey[0] = 0;
    plen[0] = 0;
  }
  if (per == 0 || t >= nsrc * per) return;
  const u64 s = t / per;
  const u32 len = static_cast<u32>(t % per) + 1;
  pkey[1 + t] = (src[s] & lowmask64(2 * len)) << (2 * (k - len));
  plen[1 + t] = static_cast<u8>(len);
}

// Merge K (all len k) and P (sorted by (key, len), all len < k) into R.
__global__ void k_scatter_K(const u64* __restrict__ K, u64 nk, const u64* __restrict__ pkey, u64 np, u32 k,
                            u64* __restrict__ rkey, u8* __restrict__ rlen) {
  const u64 a = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (a >= nk) return;
  const u64 v = K[a];
  const u64 pos = a + upper_bound_u64(pkey, np, v);  // padded nodes with key <= v precede
  rkey[pos] = v;
  rlen[pos] = static_cast<u8>(k);
}
__global__ void k_scatter_P(const u64* __restrict__ pkey, const u8* __restrict__ plen, u64 np,
                            const u64* __restrict__ K, u64 nk, u64* __restrict__ rkey, u8* __restrict__ rlen) {
  const u64 b = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (b >= np) return;
  const u64 v = pkey[b];
  const u64 pos = b + lower_bound_u64(K, nk, v);  // k-mers with key < v precede
  rkey[pos] = v;
  rlen[pos] = plen[b];
}

// R: every slot written (lengths start as 0xFF) and strictly increasing.
__global__ void k_check_R(const u64* __restrict__ rkey, const u8* __restrict__ rlen, u64 n,
                          unsigned long long* __restrict__ acc) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  u64 bad = 0;
  if (i < n) {
    bad = rlen[i] == 0xFF ? 1 : 0;
    if (i > 0 && !node_less(rkey[i - 1], rlen[i - 1], rkey[i], rlen[i])) bad = 1;
  }
  wave_add(acc + kOutBad, bad);
}

// Bucket table over the top `bb` key bits: bucket[b] = first node with key
// >= b << shift.
__global__ void k_buckets(const u64* __restrict__ rkey, u64 n, u32 shift, u64 nb, u64* __restrict__ bucket) {
  const u64 b = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (b > nb) return;
  bucket[b] = b == nb ? n : lower_bound_u64(rkey, n, b << shift);
}

// Edge bits: node d (d > 0) sets bit `last char of d` at its source column;
// group marks from adjacent comparisons.
__global__ void k_edges(const u64* __restrict__ rkey, const u8* __restrict__ rlen, u64 n, u32 k,
                        const u64* __restrict__ bucket, u32 shift, u64* __restrict__ rows, u64 nwords) {
  const u64 d = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (d >= n) return;
  const u64 key = rkey[d];
  const u32 len = rlen[d];
  // group mark
  {
    bool first = d == 0;
    if (!first) {
      const u64 pk = rkey[d - 1];
      const u32 pl = rlen[d - 1];
      const u32 sl = len < k - 1 ? len : k - 1, psl = pl < k - 1 ? pl : k - 1;
      first = !((pk >> 2) == (key >> 2) && psl == sl);
    }
    if (first)
      atomicOr(reinterpret_cast<unsigned long long*>(rows + 4 * nwords + (d >> 6)), 1ull << (63 - (d & 63)));
  }
  if (len == 0) return;  // the root has no incoming edge
  const u32 c = static_cast<u32>(key >> (2 * (k - 1))) & 3u;
  const u64 tk = (key & lowmask64(2 * (k - 1))) << 2;
  const u32 tl = len - 1;
  const u64 b = tk >> shift;
  u64 lo = bucket[b], hi = bucket[b + 1];
  while (lo < hi) {
    const u64 mid = (lo + hi) >> 1;
    if (node_less(rkey[mid], rlen[mid], tk, tl)) lo = mid + 1; else hi = mid;
  }
  if (lo < n)  // always true for a well-formed R; keeps a malformed one in bounds
    atomicOr(reinterpret_cast<unsigned long long*>(rows + c * nwords + (lo >> 6)), 1ull << (63 - (lo & 63)));
}

// ---- host side -------------------------------------------------------------------

// Device memory that frees itself.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

static int dalloc(DevBuf& b, size_t bytes) {
  b.reset();
  if (hipMalloc(&b.p, bytes ? bytes : 1) != hipSuccess) {
    b.p = nullptr;
    return fail(SBWT_ERR_HIP, "out of device memory (%llu bytes) in the SBWT builder", (unsigned long long)bytes);
  }
  b.bytes = bytes;
  return SBWT_OK;
}

// Grows b to at least `bytes` (keeps it when large enough). Work queued on
// s may still read the old buffer, so s drains before it is freed.
static int dgrow(DevBuf& b, size_t bytes, hipStream_t s) {
  if (b.p && b.bytes >= bytes) return SBWT_OK;
  if (b.p) SBWT_HIP(hipStreamSynchronize(s));
  return dalloc(b, bytes);
}

static inline u32 blocks(u64 n, u32 bt = 256) { return static_cast<u32>((n + bt - 1) / bt); }
static inline u32 check_grid(u64 n) {
  const u64 g = (n + 255) / 256;
  return static_cast<u32>(g < 4096 ? (g ? g : 1) : 4096);
}

static u64 g_sort_chunk = 1ull << 28;  // largest item count of one sort / select / unique call

// One sort-and-deduplicate job: K (keys only, len 0) or the padded nodes
// (keys and lengths). Items come from a source into bucket buffers a/b
// (each `cap` items); the deduplicated, sorted output goes to out/olen.
struct SortJob {
  hipStream_t s;
  bool pairs;
  u64 chunk;
  u64 *ka = nullptr, *kb = nullptr;
  u8 *la = nullptr, *lb = nullptr;
  u64* out = nullptr;
  u8* olen = nullptr;
  u64 nout = 0;
  unsigned long long* acc = nullptr;
  DevBuf tmp, hist, offs, starts, count;
  const char* stage = "";
};

static int ensure_tmp(SortJob& J, size_t bytes) { return dgrow(J.tmp, bytes, J.s); }

// Partition plan of a source: per-block histograms, scan, bucket bounds.
template <class Src>
static int part_plan(SortJob& J, const Src& src, u64 n, u32 shift, u32 bits, bool by_len,
                     unsigned long long* mix, std::vector<u64>& bstart) {
  const u32 nb = 1u << bits;
  const u64 nblocks = (n + kPartTile - 1) / kPartTile;
  SBWT_TRY(dgrow(J.hist, nb * nblocks * 8, J.s));
  SBWT_TRY(dgrow(J.offs, nb * nblocks * 8, J.s));
  SBWT_TRY(dgrow(J.starts, (kMaxBuckets + 1) * 8, J.s));
  k_part_hist<Src><<<static_cast<u32>(nblocks), kPartThreads, 0, J.s>>>(src, n, shift, bits, by_len, nblocks,
                                                                        J.hist.as<u64>(), mix);
  SBWT_HIP(hipGetLastError());
  size_t tb = 0;
  SBWT_HIP(rocprim::exclusive_scan(nullptr, tb, J.hist.as<u64>(), J.offs.as<u64>(), u64(0), nb * nblocks,
                                   rocprim::plus<u64>(), J.s));
  SBWT_TRY(ensure_tmp(J, tb));
  SBWT_HIP(rocprim::exclusive_scan(J.tmp.p, tb, J.hist.as<u64>(), J.offs.as<u64>(), u64(0), nb * nblocks,
                                   rocprim::plus<u64>(), J.s));
  k_bucket_starts<<<1, kMaxBuckets, 0, J.s>>>(J.offs.as<u64>(), J.hist.as<u64>(), nblocks, nb, J.starts.as<u64>());
  SBWT_HIP(hipGetLastError());
  bstart.assign(nb + 1, 0);
  SBWT_HIP(hipMemcpyAsync(bstart.data(), J.starts.p, (nb + 1) * 8, hipMemcpyDeviceToHost, J.s));
  SBWT_HIP(hipStreamSynchronize(J.s));
  return SBWT_OK;
}

template <class Src>
static int part_scatter(SortJob& J, const Src& src, u64 n, u32 shift, u32 bits, bool by_len, u64* okeys, u8* olens) {
  const u64 nblocks = (n + kPartTile - 1) / kPartTile;
  k_part_scatter<Src><<<static_cast<u32>(nblocks), kPartThreads, 0, J.s>>>(src, n, shift, bits, by_len, nblocks,
                                                                           J.offs.as<u64>(), okeys, olens);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

static int read_count(SortJob& J, u64* v) {
  SBWT_HIP(hipMemcpyAsync(v, J.count.p, 8, hipMemcpyDeviceToHost, J.s));
  SBWT_HIP(hipStreamSynchronize(J.s));
  return SBWT_OK;
}

// Appends the distinct items of a sorted bucket (keys, lens; n <= chunk) to
// the output; `flags` is n bytes of scratch.
static int dedup_append(SortJob& J, const u64* keys, const u8* lens, u64 n, u8* flags) {
  k_seg_check<<<check_grid(n), 256, 0, J.s>>>(keys, lens, n, J.acc);
  SBWT_HIP(hipGetLastError());
  u64 got = 0;
  if (!J.pairs) {
    size_t tb = 0;
    SBWT_HIP(rocprim::unique(nullptr, tb, keys, J.out + J.nout, J.count.as<u64>(), n, rocprim::equal_to<u64>(), J.s));
    SBWT_TRY(ensure_tmp(J, tb));
    SBWT_HIP(rocprim::unique(J.tmp.p, tb, keys, J.out + J.nout, J.count.as<u64>(), n, rocprim::equal_to<u64>(), J.s));
    SBWT_TRY(read_count(J, &got));
  } else {
    k_first_of_run<<<blocks(n), 256, 0, J.s>>>(keys, lens, n, flags);
    SBWT_HIP(hipGetLastError());
    size_t tb = 0, tb8 = 0;
    SBWT_HIP(rocprim::select(nullptr, tb, keys, flags, J.out + J.nout, J.count.as<u64>(), n, J.s));
    SBWT_HIP(rocprim::select(nullptr, tb8, lens, flags, J.olen + J.nout, J.count.as<u64>(), n, J.s));
    SBWT_TRY(ensure_tmp(J, tb > tb8 ? tb : tb8));
    SBWT_HIP(rocprim::select(J.tmp.p, tb, keys, flags, J.out + J.nout, J.count.as<u64>(), n, J.s));
    SBWT_HIP(rocprim::select(J.tmp.p, tb8, lens, flags, J.olen + J.nout, J.count.as<u64>(), n, J.s));
    SBWT_TRY(read_count(J, &got));
  }
  J.nout += got;
  return SBWT_OK;
}

// A bucket whose items are all equal (no key bits left, and for pairs
// equal lengths): one output item.
static int emit_equal(SortJob& J, const u64* keys, const u8* lens, u64 n) {
  k_seg_check<<<check_grid(n), 256, 0, J.s>>>(keys, lens, n, J.acc);
  SBWT_HIP(hipGetLastError());
  SBWT_HIP(hipMemcpyAsync(J.out + J.nout, keys, 8, hipMemcpyDeviceToDevice, J.s));
  if (J.pairs) SBWT_HIP(hipMemcpyAsync(J.olen + J.nout, lens, 1, hipMemcpyDeviceToDevice, J.s));
  J.nout += 1;
  return SBWT_OK;
}

// Sorts and deduplicates the n items at keys/lens, whose keys agree on every
// bit at or above hibit (hibit < 0: the lengths agree too); alt/altl is
// scratch of the same size. Buckets larger than the chunk are split again.
static int sort_bucket(SortJob& J, u64* keys, u8* lens, u64* alt, u8* altl, u64 n, int hibit) {
  if (n == 0) return SBWT_OK;
  const bool all_equal = hibit < 0 || (hibit == 0 && !J.pairs);
  if (all_equal) return emit_equal(J, keys, lens, n);
  if (n > J.chunk) {
    const bool by_len = hibit == 0;  // pairs only: the keys agree, split by length
    const u32 bits = by_len ? 6u : static_cast<u32>(hibit < 8 ? hibit : 8);
    const u32 shift = by_len ? 0u : static_cast<u32>(hibit) - bits;
    std::vector<u64> bs;
    ArraySrc src{keys, J.pairs ? lens : nullptr};
    SBWT_TRY(part_plan(J, src, n, shift, bits, by_len, nullptr, bs));
    if (bs.back() != n) return fail(SBWT_ERR_INTERNAL, "SBWT builder, %s: a partition placed %llu of %llu items",
                                    J.stage, (unsigned long long)bs.back(), (unsigned long long)n);
    SBWT_TRY(part_scatter(J, src, n, shift, bits, by_len, alt, J.pairs ? altl : nullptr));
    const int sub = by_len ? -1 : static_cast<int>(shift);
    for (u32 b = 0; b + 1 < bs.size(); b++)
      SBWT_TRY(sort_bucket(J, alt + bs[b], altl ? altl + bs[b] : nullptr, keys + bs[b],
                           lens ? lens + bs[b] : nullptr, bs[b + 1] - bs[b], sub));
    return SBWT_OK;
  }
  size_t tb = 0;
  if (!J.pairs) {
    SBWT_HIP(rocprim::radix_sort_keys(nullptr, tb, keys, alt, n, 0, hibit, J.s));
    SBWT_TRY(ensure_tmp(J, tb));
    SBWT_HIP(rocprim::radix_sort_keys(J.tmp.p, tb, keys, alt, n, 0, hibit, J.s));
    return dedup_append(J, alt, nullptr, n, nullptr);
  }
  // pairs: stable passes, by length then by key -> (key, len) order
  size_t tb2 = 0;
  SBWT_HIP(rocprim::radix_sort_pairs(nullptr, tb, lens, altl, keys, alt, n, 0, 6, J.s));
  if (hibit > 0) SBWT_HIP(rocprim::radix_sort_pairs(nullptr, tb2, alt, keys, altl, lens, n, 0, hibit, J.s));
  SBWT_TRY(ensure_tmp(J, tb > tb2 ? tb : tb2));
  SBWT_HIP(rocprim::radix_sort_pairs(J.tmp.p, tb, lens, altl, keys, alt, n, 0, 6, J.s));
  if (hibit == 0) return dedup_append(J, alt, altl, n, lens);
  SBWT_HIP(rocprim::radix_sort_pairs(J.tmp.p, tb2, alt, keys, altl, lens, n, 0, hibit, J.s));
  return dedup_append(J, keys, lens, n, altl);
}

// The whole job over a source of n candidate items with `total` valid ones:
// level-0 partition on the top key bits into a, then sort_bucket per bucket.
template <class Src>
static int sort_unique(SortJob& J, const Src& src, u64 n, u32 keybits, u64* out_count) {
  const u32 bits = keybits < 8 ? keybits : 8;
  const u32 shift = keybits - bits;
  std::vector<u64> bs;
  SBWT_TRY(part_plan(J, src, n, shift, bits, false, J.acc + kMixSrc, bs));
  const u64 total = bs.back();
  SBWT_TRY(part_scatter(J, src, n, shift, bits, false, J.ka, J.pairs ? J.la : nullptr));
  for (u32 b = 0; b + 1 < bs.size(); b++)
    SBWT_TRY(sort_bucket(J, J.ka + bs[b], J.pairs ? J.la + bs[b] : nullptr, J.kb + bs[b],
                         J.pairs ? J.lb + bs[b] : nullptr, bs[b + 1] - bs[b], static_cast<int>(shift)));
  if (J.nout) {
    k_out_check<<<check_grid(J.nout), 256, 0, J.s>>>(J.out, J.pairs ? J.olen : nullptr, J.nout, J.acc);
    SBWT_HIP(hipGetLastError());
  }
  u64 a[kAccWords];
  SBWT_HIP(hipMemcpyAsync(a, J.acc, sizeof a, hipMemcpyDeviceToHost, J.s));
  SBWT_HIP(hipStreamSynchronize(J.s));
  if (a[kMixSorted] != a[kMixSrc] || a[kUnsorted] != 0)
    return fail(SBWT_ERR_INTERNAL,
                "SBWT builder, %s: the sorted buckets are not a sorted permutation of the %llu input items "
                "(mix sums %016llx / %016llx, %llu adjacent pairs out of order)",
                J.stage, (unsigned long long)total, (unsigned long long)a[kMixSrc],
                (unsigned long long)a[kMixSorted], (unsigned long long)a[kUnsorted]);
  if (a[kRuns] != J.nout || a[kMixOut] != a[kMixFirst] || a[kOutBad] != 0)
    return fail(SBWT_ERR_INTERNAL,
                "SBWT builder, %s: deduplication kept %llu items for %llu runs (mix sums %016llx / %016llx, "
                "%llu pairs not strictly increasing)",
                J.stage, (unsigned long long)J.nout, (unsigned long long)a[kRuns], (unsigned long long)a[kMixFirst],
                (unsigned long long)a[kMixOut], (unsigned long long)a[kOutBad]);
  *out_count = total;
  return SBWT_OK;
}

// rocprim::select over n items in calls of at most `chunk` items.
static int select_chunked(SortJob& J, const u64* in, const u8* flags, u64* out, u64 n, u64* selected) {
  u64 done = 0;
  for (u64 off = 0; off < n; off += J.chunk) {
    const u64 len = n - off < J.chunk ? n - off : J.chunk;
    size_t tb = 0;
    SBWT_HIP(rocprim::select(nullptr, tb, in + off, flags + off, out + done, J.count.as<u64>(), len, J.s));
    SBWT_TRY(ensure_tmp(J, tb));
    SBWT_HIP(rocprim::select(J.tmp.p, tb, in + off, flags + off, out + done, J.count.as<u64>(), len, J.s));
    u64 got = 0;
    SBWT_TRY(read_count(J, &got));
    done += got;
  }
  *selected = done;
  return SBWT_OK;
}

constexpr u64 kMaxSamples = 1ull << 22;

int build_sbwt(const u8* d_seq, const u64* d_starts, u64 nseq, u32 k, const sbwt_config* cfg, hipStream_t s,
               sbwt_index** out) {
  u64 total = 0;
  SBWT_HIP(hipMemcpyAsync(&total, d_starts + nseq, sizeof(u64), hipMemcpyDeviceToHost, s));
  SBWT_HIP(hipStreamSynchronize(s));

  SortJob J;
  J.s = s;
  J.chunk = g_sort_chunk;
  DevBuf accbuf;
  SBWT_TRY(dalloc(accbuf, kAccWords * 8));
  SBWT_TRY(dalloc(J.count, 8));
  J.acc = accbuf.as<unsigned long long>();
  SBWT_HIP(hipMemsetAsync(J.acc, 0, kAccWords * 8, s));

  // 1. pack + non-ACGT mask, sequence-start marks
  DevBuf packed, nmask, bounds;
  const u64 pw = sbwt_packed_words(total), mw = sbwt_nmask_words(total);
  SBWT_TRY(dalloc(packed, pw * 8));
  SBWT_TRY(dalloc(nmask, mw * 8));
  SBWT_TRY(dalloc(bounds, mw * 8));
  SBWT_TRY(sbwt_pack(d_seq, total, packed.as<u64>(), nmask.as<u64>(), s));
  SBWT_HIP(hipMemsetAsync(bounds.p, 0, mw * 8, s));
  if (nseq > 1) k_mark_bounds<<<blocks(nseq), 256, 0, s>>>(d_starts, nseq, total, bounds.as<u64>());
  const WindowSrc win{packed.as<u64>(), nmask.as<u64>(), bounds.as<u64>(), k};

  // 2. K = the distinct valid window keys, sorted
  const u64 windows = total >= k ? total - k + 1 : 0;
  DevBuf K, ka, kb;
  u64 nk = 0, nvalid = 0, nsample = 0;
  DevBuf sample;
  SBWT_TRY(dalloc(sample, 8 * (windows < kMaxSamples ? windows : kMaxSamples)));
  if (windows) {
    k_count_valid<<<check_grid(windows), 256, 0, s>>>(win, windows, J.acc + kValid);
    SBWT_HIP(hipGetLastError());
    // the valid count sizes the buffers before the partition writes them
    u64 a[kAccWords];
    SBWT_HIP(hipMemcpyAsync(a, J.acc, sizeof a, hipMemcpyDeviceToHost, s));
    SBWT_HIP(hipStreamSynchronize(s));
    nvalid = a[kValid];
    SBWT_TRY(dalloc(ka, nvalid * 8));
    SBWT_TRY(dalloc(kb, nvalid * 8));
    SBWT_TRY(dalloc(K, nvalid * 8));
    J.ka = ka.as<u64>();
    J.kb = kb.as<u64>();
    J.out = K.as<u64>();
    J.pairs = false;
    J.stage = "k-mer set";
    u64 placed = 0;
    SBWT_TRY(sort_unique(J, win, windows, 2 * k, &placed));
    if (placed != nvalid)
      return fail(SBWT_ERR_INTERNAL, "SBWT builder, valid windows: the partition placed %llu, a separate count "
                  "found %llu", (unsigned long long)placed, (unsigned long long)nvalid);
    nk = J.nout;
    SBWT_HIP(hipStreamSynchronize(s));
    ka.reset();
    kb.reset();
    // sampled windows must be in K
    const u64 ns = windows < kMaxSamples ? windows : kMaxSamples;
    const u64 stride = (windows + ns - 1) / ns;
    k_sample_windows<<<blocks(ns), 256, 0, s>>>(win, windows, stride, ns, K.as<u64>(), nk, sample.as<u64>(), J.acc);
    SBWT_HIP(hipGetLastError());
    SBWT_HIP(hipMemcpyAsync(a, J.acc, sizeof a, hipMemcpyDeviceToHost, s));
    SBWT_HIP(hipStreamSynchronize(s));
    if (a[kMisses] != 0)
      return fail(SBWT_ERR_INTERNAL, "SBWT builder, k-mer set: %llu of %llu sampled sequence k-mers are missing "
                  "from K", (unsigned long long)a[kMisses], (unsigned long long)a[kSampled]);
    nsample = a[kSampled];
  }
  SBWT_HIP(hipStreamSynchronize(s));
  packed.reset();
  nmask.reset();
  bounds.reset();
  const u64* dK = K.as<u64>();

  // 3. sources and padded nodes (sorted by (key, len), deduplicated)
  DevBuf issrc, srcs;
  u64 nsrc = 0;
  if (nk) {
    SBWT_TRY(dalloc(issrc, nk));
    SBWT_TRY(dalloc(srcs, nk * 8));
    k_sources<<<blocks(nk), 256, 0, s>>>(dK, nk, k, issrc.as<u8>());
    SBWT_HIP(hipGetLastError());
    SBWT_TRY(select_chunked(J, dK, issrc.as<u8>(), srcs.as<u64>(), nk, &nsrc));
  }
  SBWT_HIP(hipStreamSynchronize(s));
  issrc.reset();
  const u64 np0 = 1 + nsrc * (k - 1);
  DevBuf pk, pl, pka, pla, pkb, plb, P, PL;
  SBWT_TRY(dalloc(pk, np0 * 8));
  SBWT_TRY(dalloc(pl, np0));
  k_padded<<<blocks(np0), 256, 0, s>>>(srcs.as<u64>(), nsrc, k, pk.as<u64>(), pl.as<u8>());
  SBWT_HIP(hipGetLastError());
  SBWT_HIP(hipStreamSynchronize(s));
  srcs.reset();
  SBWT_TRY(dalloc(pka, np0 * 8));
  SBWT_TRY(dalloc(pla, np0));
  SBWT_TRY(dalloc(pkb, np0 * 8));
  SBWT_TRY(dalloc(plb, np0));
  SBWT_TRY(dalloc(P, np0 * 8));
  SBWT_TRY(dalloc(PL, np0));
  SBWT_HIP(hipMemsetAsync(J.acc, 0, kMisses * 8, s));  // the sort checks start again; the sample counters stay
  J.ka = pka.as<u64>();
  J.la = pla.as<u8>();
  J.kb = pkb.as<u64>();
  J.lb = plb.as<u8>();
  J.out = P.as<u64>();
  J.olen = PL.as<u8>();
  J.nout = 0;
  J.pairs = true;
  J.stage = "padded nodes";
  u64 placed = 0;
  SBWT_TRY(sort_unique(J, ArraySrc{pk.as<u64>(), pl.as<u8>()}, np0, 2 * k, &placed));
  if (placed != np0)
    return fail(SBWT_ERR_INTERNAL, "SBWT builder, padded nodes: the partition placed %llu of %llu",
                (unsigned long long)placed, (unsigned long long)np0);
  const u64 np = J.nout;
  SBWT_HIP(hipStreamSynchronize(s));
  pk.reset();
  pl.reset();
  pka.reset();
  pla.reset();
  pkb.reset();
  plb.reset();
  J.tmp.reset();
  J.hist.reset();
  J.offs.reset();

  // 4. merge into R
  const u64 n = nk + np;
  SBWT_TRY(check_shape(n, k));
  DevBuf rkey, rlen;
  SBWT_TRY(dalloc(rkey, n * 8));
  SBWT_TRY(dalloc(rlen, n));
  SBWT_HIP(hipMemsetAsync(rlen.p, 0xFF, n, s));
  SBWT_HIP(hipMemsetAsync(J.acc + kOutBad, 0, 8, s));
  if (nk) k_scatter_K<<<blocks(nk), 256, 0, s>>>(dK, nk, P.as<u64>(), np, k, rkey.as<u64>(), rlen.as<u8>());
  k_scatter_P<<<blocks(np), 256, 0, s>>>(P.as<u64>(), PL.as<u8>(), np, dK, nk, rkey.as<u64>(), rlen.as<u8>());
  k_check_R<<<blocks(n), 256, 0, s>>>(rkey.as<u64>(), rlen.as<u8>(), n, J.acc);
  SBWT_HIP(hipGetLastError());
  {
    u64 bad = 0;
    SBWT_HIP(hipMemcpyAsync(&bad, J.acc + kOutBad, 8, hipMemcpyDeviceToHost, s));
    SBWT_HIP(hipStreamSynchronize(s));
    if (bad)
      return fail(SBWT_ERR_INTERNAL, "SBWT builder, merge: %llu of %llu slots of the padded k-spectrum are "
                  "unwritten or out of order", (unsigned long long)bad, (unsigned long long)n);
  }
  K.reset();
  P.reset();
  PL.reset();

  // 5. rows: edge bits + group marks
  sbwt_index* ix = new_index(cfg, k, n);
  ix->has_groups = 1;
  ix->num_kmers = nk;
  if (hipMalloc(&ix->d_rows, kRows * ix->nwords * sizeof(u64)) != hipSuccess) {
    destroy_index(ix);
    return fail(SBWT_ERR_HIP, "out of device memory for %llu-column rows", (unsigned long long)n);
  }
  int st = SBWT_OK;
  do {
    if (hipMemsetAsync(ix->d_rows, 0, kRows * ix->nwords * sizeof(u64), s) != hipSuccess) {
      st = fail(SBWT_ERR_HIP, "row clear failed");
      break;
    }
    const u32 kb2 = 2 * k < 24 ? 2 * k : 24;  // bucket on the top kb2 key bits
    const u32 shift = 2 * k - kb2;
    const u64 nb = 1ull << kb2;
    DevBuf bucket;
    if ((st = dalloc(bucket, (nb + 1) * 8)) != SBWT_OK) break;
    k_buckets<<<blocks(nb + 1), 256, 0, s>>>(rkey.as<u64>(), n, shift, nb, bucket.as<u64>());
    k_edges<<<blocks(n), 256, 0, s>>>(rkey.as<u64>(), rlen.as<u8>(), n, k, bucket.as<u64>(), shift, ix->d_rows,
                                      ix->nwords);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      st = fail(SBWT_ERR_HIP, "edge kernel failed");
      break;
    }
  } while (0);
  rkey.reset();
  rlen.reset();
  if (st == SBWT_OK) st = finish_index(ix, s);  // checks C[4] = n
  // 6. the finished index finds every sampled k-mer
  if (st == SBWT_OK && nsample) {
    DevBuf ans;
    st = dalloc(ans, nsample * 8);
    // the counted instance, so that the plain kernel's dispatches are the
    // caller's own; its counts are cleared afterwards
    if (st == SBWT_OK) st = sbwt_search_kmers_flags(ix, sample.as<u64>(), nsample, ans.as<u64>(), SBWT_COUNT, s);
    if (st == SBWT_OK && hipMemsetAsync(ix->d_counters, 0, kCounters * sizeof(unsigned long long), s) != hipSuccess)
      st = fail(SBWT_ERR_HIP, "counter reset failed");
    if (st == SBWT_OK) {
      u64 miss = 0;
      k_count_absent<<<blocks(nsample), 256, 0, s>>>(ans.as<u64>(), nsample, J.acc);
      if (hipGetLastError() != hipSuccess ||
          hipMemcpyAsync(&miss, J.acc + kMisses, 8, hipMemcpyDeviceToHost, s) != hipSuccess ||
          hipStreamSynchronize(s) != hipSuccess)
        st = fail(SBWT_ERR_HIP, "the index check after the build failed to run");
      else if (miss)
        st = fail(SBWT_ERR_INTERNAL, "SBWT builder, index: the built index does not find %llu of %llu sampled "
                  "sequence k-mers", (unsigned long long)miss, (unsigned long long)nsample);
    }
  }
  if (st != SBWT_OK) {
    destroy_index(ix);
    return st;
  }
  *out = ix;
  return SBWT_OK;
}

}  // namespace sbwt

using namespace sbwt;

extern "C" int sbwt_index_build(const uint8_t* d_seq, const uint64_t* d_starts, uint64_t nseq, uint32_t k,
                                const sbwt_config* cfg, void* stream, sbwt_index** out) {
  clear_error();
  SBWT_REQUIRE(d_starts && out, "d_starts and out must not be NULL");
  SBWT_REQUIRE(k >= 1 && k <= 32, "k = %u: must be in [1, 32]", k);
  *out = nullptr;
  const int st = build_sbwt(d_seq, d_starts, nseq, k, cfg, as_stream(stream), out);
  if (st != SBWT_OK) (void)hipStreamSynchronize(as_stream(stream));  // nothing still reads freed buffers
  return st;
}

extern "C" int sbwt_build_set_sort_chunk(uint64_t max_items) {
  clear_error();
  SBWT_REQUIRE(max_items == 0 || max_items >= 1, "max_items must be 0 or positive");
  g_sort_chunk = max_items ? max_items : (1ull << 28);
  return SBWT_OK;
}
