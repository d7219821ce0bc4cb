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
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

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

// Colex key of the k-mer starting at character p.
struct KeyAt {
  const u64* packed;
  u32 k;
  __device__ u64 operator()(u64 p) const { return rev_pairs(window2(packed, p)) & lowmask64(2 * k); }
};

// The k-mer at p is valid if it has no non-ACGT character and no sequence
// starts inside (p, p+k).
struct ValidAt {
  const u64* nmask;
  const u64* bounds;
  u32 k;
  __device__ u8 operator()(u64 p) const {
    const u64 bad = window1(nmask, p) >> (64 - k);
    const u64 crossing = (window1(bounds, p) << 1) >> (64 - k + 1 > 63 ? 63 : 64 - k + 1);
    return (bad == 0 && (k == 1 || crossing == 0)) ? 1 : 0;
  }
};

__global__ void k_mark_bounds(const u64* __restrict__ starts, u64 nseq, u64 total, u64* __restrict__ bounds) {
  const u64 s = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x + 1;
  if (s >= nseq) return;
  const u64 q = starts[s];
  if (q < total) atomicOr(reinterpret_cast<unsigned long long*>(bounds + (q >> 6)), 1ull << (63 - (q & 63)));
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

// Drop duplicate (key, len) pairs of a sorted list: flag the first of each run.
__global__ void k_first_of_run(const u64* __restrict__ key, const u8* __restrict__ len, u64 n, u8* __restrict__ keep) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  keep[i] = (i == 0 || key[i] != key[i - 1] || len[i] != len[i - 1]) ? 1 : 0;
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

// Bucket table over the top `bb` key bits: bucket[b] = first node with key
// >= b << shift.
__global__ void k_buckets(const u64* __restrict__ rkey, u64 n, u32 shift, u64 nb, u64* __restrict__ bucket) {
  const u64 b = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (b > nb) return;
  bucket[b] = b == nb ? n : lower_bound_u64(rkey, n, b << shift);
}

__device__ __forceinline__ bool node_less(u64 ka, u32 la, u64 kb, u32 lb) { return ka < kb || (ka == kb && la < lb); }

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

// Device memory that frees itself.
struct DevBuf {
  void* p = nullptr;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { reset(); }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
  }
  template <class T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

static int dalloc(DevBuf& b, size_t bytes) {
  b.reset();
  if (hipMalloc(&b.p, bytes ? bytes : 1) != hipSuccess)
    return fail(SBWT_ERR_HIP, "out of device memory (%llu bytes) in the SBWT builder", (unsigned long long)bytes);
  return SBWT_OK;
}

static inline u32 blocks(u64 n, u32 bt = 256) { return static_cast<u32>((n + bt - 1) / bt); }

int build_sbwt(const u8* d_seq, const u64* d_starts, u64 nseq, u32 k, const sbwt_config* cfg, hipStream_t s,
               sbwt_index** out) {
  u64 total = 0;
  SBWT_HIP(hipMemcpyAsync(&total, d_starts + nseq, sizeof(u64), hipMemcpyDeviceToHost, s));
  SBWT_HIP(hipStreamSynchronize(s));

  // 1. pack + non-ACGT mask, sequence-start marks
  DevBuf packed, nmask, bounds;
  const u64 pw = sbwt_packed_words(total), mw = sbwt_nmask_words(total);
  SBWT_TRY(dalloc(packed, pw * 8));
  SBWT_TRY(dalloc(nmask, mw * 8));
  SBWT_TRY(dalloc(bounds, mw * 8));
  SBWT_TRY(sbwt_pack(d_seq, total, packed.as<u64>(), nmask.as<u64>(), s));
  SBWT_HIP(hipMemsetAsync(bounds.p, 0, mw * 8, s));
  if (nseq > 1) k_mark_bounds<<<blocks(nseq), 256, 0, s>>>(d_starts, nseq, total, bounds.as<u64>());

  // 2. valid k-mer keys, compacted
  const u64 windows = total >= k ? total - k + 1 : 0;
  DevBuf keys, keys2, count, tmp;
  SBWT_TRY(dalloc(count, 8));
  SBWT_TRY(dalloc(keys, windows * 8));
  u64 nk = 0;
  if (windows) {
    auto cnt = rocprim::make_counting_iterator<u64>(0);
    auto kin = rocprim::make_transform_iterator(cnt, KeyAt{packed.as<u64>(), k});
    auto fin = rocprim::make_transform_iterator(cnt, ValidAt{nmask.as<u64>(), bounds.as<u64>(), k});
    size_t tb = 0;
    SBWT_HIP(rocprim::select(nullptr, tb, kin, fin, keys.as<u64>(), count.as<u64>(), windows, s));
    SBWT_TRY(dalloc(tmp, tb));
    SBWT_HIP(rocprim::select(tmp.p, tb, kin, fin, keys.as<u64>(), count.as<u64>(), windows, s));
    SBWT_HIP(hipMemcpyAsync(&nk, count.p, 8, hipMemcpyDeviceToHost, s));
    SBWT_HIP(hipStreamSynchronize(s));
  }
  packed.reset();
  nmask.reset();
  bounds.reset();

  // 3. sort + unique -> K
  if (nk) {
    SBWT_TRY(dalloc(keys2, nk * 8));
    size_t tb = 0;
    SBWT_HIP(rocprim::radix_sort_keys(nullptr, tb, keys.as<u64>(), keys2.as<u64>(), nk, 0, 2 * k, s));
    SBWT_TRY(dalloc(tmp, tb));
    SBWT_HIP(rocprim::radix_sort_keys(tmp.p, tb, keys.as<u64>(), keys2.as<u64>(), nk, 0, 2 * k, s));
    tb = 0;
    SBWT_HIP(rocprim::unique(nullptr, tb, keys2.as<u64>(), keys.as<u64>(), count.as<u64>(), nk,
                             rocprim::equal_to<u64>(), s));
    SBWT_TRY(dalloc(tmp, tb));
    SBWT_HIP(rocprim::unique(tmp.p, tb, keys2.as<u64>(), keys.as<u64>(), count.as<u64>(), nk,
                             rocprim::equal_to<u64>(), s));
    SBWT_HIP(hipMemcpyAsync(&nk, count.p, 8, hipMemcpyDeviceToHost, s));
    SBWT_HIP(hipStreamSynchronize(s));
  }
  keys2.reset();
  const u64* K = keys.as<u64>();

  // 4. sources and padded nodes (sorted by (key, len), deduplicated)
  DevBuf issrc, srcs;
  u64 nsrc = 0;
  if (nk) {
    SBWT_TRY(dalloc(issrc, nk));
    SBWT_TRY(dalloc(srcs, nk * 8));
    k_sources<<<blocks(nk), 256, 0, s>>>(K, nk, k, issrc.as<u8>());
    size_t tb = 0;
    SBWT_HIP(rocprim::select(nullptr, tb, K, issrc.as<u8>(), srcs.as<u64>(), count.as<u64>(), nk, s));
    SBWT_TRY(dalloc(tmp, tb));
    SBWT_HIP(rocprim::select(tmp.p, tb, K, issrc.as<u8>(), srcs.as<u64>(), count.as<u64>(), nk, s));
    SBWT_HIP(hipMemcpyAsync(&nsrc, count.p, 8, hipMemcpyDeviceToHost, s));
    SBWT_HIP(hipStreamSynchronize(s));
  }
  issrc.reset();
  const u64 np0 = 1 + nsrc * (k - 1);
  DevBuf pk, pl, pk2, pl2, keep;
  SBWT_TRY(dalloc(pk, np0 * 8));
  SBWT_TRY(dalloc(pl, np0));
  SBWT_TRY(dalloc(pk2, np0 * 8));
  SBWT_TRY(dalloc(pl2, np0));
  k_padded<<<blocks(np0), 256, 0, s>>>(srcs.as<u64>(), nsrc, k, pk.as<u64>(), pl.as<u8>());
  srcs.reset();
  {
    // stable LSD passes: by len, then by key -> (key, len) order
    size_t tb = 0, tb2 = 0;
    SBWT_HIP(rocprim::radix_sort_pairs(nullptr, tb, pl.as<u8>(), pl2.as<u8>(), pk.as<u64>(), pk2.as<u64>(), np0, 0, 8, s));
    SBWT_HIP(rocprim::radix_sort_pairs(nullptr, tb2, pk2.as<u64>(), pk.as<u64>(), pl2.as<u8>(), pl.as<u8>(), np0, 0,
                                       2 * k, s));
    SBWT_TRY(dalloc(tmp, tb > tb2 ? tb : tb2));
    SBWT_HIP(rocprim::radix_sort_pairs(tmp.p, tb, pl.as<u8>(), pl2.as<u8>(), pk.as<u64>(), pk2.as<u64>(), np0, 0, 8, s));
    SBWT_HIP(rocprim::radix_sort_pairs(tmp.p, tb2, pk2.as<u64>(), pk.as<u64>(), pl2.as<u8>(), pl.as<u8>(), np0, 0,
                                       2 * k, s));
  }
  u64 np = 0;
  SBWT_TRY(dalloc(keep, np0));
  k_first_of_run<<<blocks(np0), 256, 0, s>>>(pk.as<u64>(), pl.as<u8>(), np0, keep.as<u8>());
  {
    size_t tb = 0, tb8 = 0;
    SBWT_HIP(rocprim::select(nullptr, tb, pk.as<u64>(), keep.as<u8>(), pk2.as<u64>(), count.as<u64>(), np0, s));
    SBWT_HIP(rocprim::select(nullptr, tb8, pl.as<u8>(), keep.as<u8>(), pl2.as<u8>(), count.as<u64>(), np0, s));
    if (tb8 > tb) tb = tb8;
    SBWT_TRY(dalloc(tmp, tb));
    SBWT_HIP(rocprim::select(tmp.p, tb, pk.as<u64>(), keep.as<u8>(), pk2.as<u64>(), count.as<u64>(), np0, s));
    SBWT_HIP(rocprim::select(tmp.p, tb, pl.as<u8>(), keep.as<u8>(), pl2.as<u8>(), count.as<u64>(), np0, s));
    SBWT_HIP(hipMemcpyAsync(&np, count.p, 8, hipMemcpyDeviceToHost, s));
    SBWT_HIP(hipStreamSynchronize(s));
  }
  pk.reset();
  pl.reset();
  keep.reset();
  tmp.reset();

  // 5. merge into R
  const u64 n = nk + np;
  SBWT_TRY(check_shape(n, k));
  DevBuf rkey, rlen;
  SBWT_TRY(dalloc(rkey, n * 8));
  SBWT_TRY(dalloc(rlen, n));
  if (nk) k_scatter_K<<<blocks(nk), 256, 0, s>>>(K, nk, pk2.as<u64>(), np, k, rkey.as<u64>(), rlen.as<u8>());
  k_scatter_P<<<blocks(np), 256, 0, s>>>(pk2.as<u64>(), pl2.as<u8>(), np, K, nk, rkey.as<u64>(), rlen.as<u8>());
  SBWT_HIP(hipGetLastError());
  keys.reset();
  pk2.reset();
  pl2.reset();

  // 6. rows: edge bits + group marks
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
    const u32 kb = 2 * k < 24 ? 2 * k : 24;  // bucket on the top kb key bits
    const u32 shift = 2 * k - kb;
    const u64 nb = 1ull << kb;
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
  if (st == SBWT_OK) st = finish_index(ix, s);
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
  return build_sbwt(d_seq, d_starts, nseq, k, cfg, as_stream(stream), out);
}
