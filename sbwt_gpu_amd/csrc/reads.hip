// Query ingest kernels of libsbwt_gpu_amd: 2-bit packing with a non-ACGT
// mask, per-row k-mer offsets, k-mer start positions, and the work prefix
// of the read search.
//
// Reference counterparts: create_kmer_string (include/gpu/create_kmer_string.hpp:22-42),
// create_positions (include/gpu/positions_kernel.hpp:18-36) and the host-side
// position counting of KmerSearch::read_file (include/epicseq/kmer_search.hpp:319-342).
#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "index.hpp"

namespace sbwt {

// The reference packing table (create_kmer_string.hpp:20) on four bytes at
// once (byte 0 first): C/c->1, G/g->2, T/t->3, every other byte -> 0.
// Returns the four 2-bit codes MSB-first in 8 bits and sets `bad` to the four
// non-ACGT flags MSB-first in 4 bits. Branch-free: a chain of per-byte
// selects compiles to a branch per character.
__device__ __forceinline__ u32 code4(u32 w, u32& bad) {
  const u32 lc = w | 0x20202020u;  // lower case
  // 0x80 in each byte equal to a, c, g or t (an exact zero-byte test of lc ^ v)
  u32 ok = 0;
  const u32 letters[4] = {0x61u, 0x63u, 0x67u, 0x74u};
#pragma unroll
  for (int i = 0; i < 4; i++) {
    const u32 y = lc ^ (letters[i] * 0x01010101u);
    ok |= ~(((y & 0x7F7F7F7Fu) + 0x7F7F7F7Fu) | y) & 0x80808080u;
  }
  // bits 2:1 of a, c, g, t are 00, 01, 11, 10: a Gray code of 0, 1, 2, 3
  const u32 x = (lc >> 1) & 0x03030303u;
  const u32 code = (x ^ ((x >> 1) & 0x01010101u)) & ((ok >> 7) * 3u);
  // gather byte i's field to bits 7-2i:6-2i (codes) and bit 3-i (flags); the
  // products' partial terms never overlap, so nothing carries
  bad = ((((~ok & 0x80808080u) >> 7) * 0x08040201u) >> 24) & 0xFu;
  return (code * 0x40100401u) >> 24;
}

// Thread q packs characters [32q, 32q+32): one packed word (char 0 in bits
// 63:62) and one 32-bit half of the mask (MSB-first, so the half of mask
// word q/2 that holds them is u32 q^1 in little-endian memory). Characters
// past n read as byte 0 (not ACGT).
__global__ __launch_bounds__(256) void k_pack(const u8* __restrict__ chars, u64 n, u64* __restrict__ packed,
                                              u64 pwords, u32* __restrict__ mask32, u64 mhalves, u64 threads) {
  const u64 q = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (q >= threads) return;
  const u64 first = q * 32;
  u32 w[8];
  // the same misalignment for every thread: a batch cut out of a larger
  // buffer starts anywhere
  const u32 off = static_cast<u32>(reinterpret_cast<uintptr_t>(chars) & 15);
  const uintptr_t end16 = (reinterpret_cast<uintptr_t>(chars + n) + 15) & ~static_cast<uintptr_t>(15);
  if (first + 32 <= n && off == 0) {
    const uint4* p = reinterpret_cast<const uint4*>(chars + first);
    const uint4 v0 = p[0], v1 = p[1];
    w[0] = v0.x; w[1] = v0.y; w[2] = v0.z; w[3] = v0.w;
    w[4] = v1.x; w[5] = v1.y; w[6] = v1.z; w[7] = v1.w;
  } else if (first + 32 <= n && reinterpret_cast<uintptr_t>(chars + first - off) + 48 <= end16) {
    // three aligned 16-byte loads (the last one ends in the 16-byte block
    // that holds the batch's last byte) and byte funnel shifts
    const uint4* p = reinterpret_cast<const uint4*>(chars + first - off);
    const uint4 v0 = p[0], v1 = p[1], v2 = p[2];
    const u32 W[12] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w, v2.x, v2.y, v2.z, v2.w};
    const u32 wo = off >> 2, bo = off & 3;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const u32 lo = wo == 0 ? W[i] : (wo == 1 ? W[i + 1] : (wo == 2 ? W[i + 2] : W[i + 3]));
      const u32 hi = wo == 0 ? W[i + 1] : (wo == 1 ? W[i + 2] : (wo == 2 ? W[i + 3] : W[i + 4]));
      w[i] = bo ? __builtin_amdgcn_alignbyte(hi, lo, bo) : lo;
    }
  } else {
#pragma unroll
    for (int i = 0; i < 8; i++) {
      u32 v = 0;
#pragma unroll
      for (int t = 0; t < 4; t++) {
        const u64 at = first + 4 * i + t;
        v |= (at < n ? static_cast<u32>(chars[at]) : 0u) << (8 * t);
      }
      w[i] = v;
    }
  }
  u32 hi = 0, lo = 0, half = 0;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    u32 b0, b1;
    hi |= code4(w[i], b0) << (24 - 8 * i);
    lo |= code4(w[i + 4], b1) << (24 - 8 * i);
    half |= b0 << (28 - 4 * i);
    half |= b1 << (12 - 4 * i);
  }
  if (q < pwords) packed[q] = (static_cast<u64>(hi) << 32) | lo;
  if (mask32 && (q ^ 1) < mhalves) mask32[q ^ 1] = half;
}

struct KmerCount {
  const u64* rs;
  u64 rows;
  u32 k;
  __host__ __device__ u64 operator()(u64 r) const {
    if (r >= rows) return 0;
    const u64 len = rs[r + 1] - rs[r];
    return len >= k ? len - k + 1 : 0;
  }
};

struct WorkCount {
  const u64* off;
  u64 rows;
  u32 seg;
  __host__ __device__ u64 operator()(u64 r) const {
    if (r >= rows) return 0;
    const u64 nk = off[r + 1] - off[r];
    return (nk + seg - 1) / seg;
  }
};

template <class F>
static int scan_rows(F f, u64 rows, u64* out, hipStream_t s) {
  auto it = rocprim::make_transform_iterator(rocprim::make_counting_iterator<u64>(0), f);
  size_t tmp = 0;
  SBWT_HIP(rocprim::exclusive_scan(nullptr, tmp, it, out, u64(0), rows + 1, rocprim::plus<u64>(), s));
  void* d_tmp = nullptr;
  SBWT_HIP(hipMallocAsync(&d_tmp, tmp ? tmp : 1, s));
  hipError_t e = rocprim::exclusive_scan(d_tmp, tmp, it, out, u64(0), rows + 1, rocprim::plus<u64>(), s);
  (void)hipFreeAsync(d_tmp, s);
  if (e != hipSuccess) return fail(SBWT_ERR_HIP, "row scan failed: %s", hipGetErrorString(e));
  return SBWT_OK;
}

// One 64-lane block per row writes its consecutive start offsets.
__global__ __launch_bounds__(64) void k_positions(const u64* __restrict__ rs, const u64* __restrict__ off, u64 rows,
                                                  u64* __restrict__ out) {
  for (u64 r = blockIdx.x; r < rows; r += gridDim.x) {
    const u64 nk = off[r + 1] - off[r];
    const u64 base = rs[r];
    u64* o = out + off[r];
    for (u64 i = threadIdx.x; i < nk; i += blockDim.x) o[i] = base + i;
  }
}

// items[w] = the row of work item w, for every w in [work[r], work[r+1]).
__global__ __launch_bounds__(256) void k_items(const u64* __restrict__ work, u64 rows, u32* __restrict__ items) {
  const u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (r >= rows) return;
  for (u64 w = work[r]; w < work[r + 1]; w++) items[w] = static_cast<u32>(r);
}

// The k-mer starting at each position of the packed string, top-aligned
// (the extraction of include/gpu/search_kernel.hpp:66-68, kept separate so
// a caller can hold k-mers instead of positions).
__global__ __launch_bounds__(256) void k_extract(const u64* __restrict__ packed, const u64* __restrict__ pos,
                                                 u64 count, u64* __restrict__ kmers) {
  const u64 q = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (q >= count) return;
  const u64 b = pos[q] << 1;
  kmers[q] = (packed[b >> 6] << (b & 63)) | ((packed[(b >> 6) + 1] >> (63 - (b & 63))) >> 1);
}

}  // namespace sbwt

using namespace sbwt;

extern "C" {

uint64_t sbwt_packed_words(uint64_t n_chars) { return n_chars / 32 + 2; }
uint64_t sbwt_nmask_words(uint64_t n_chars) { return n_chars / 64 + 2; }

int sbwt_pack(const uint8_t* d_chars, uint64_t n_chars, uint64_t* d_packed, uint64_t* d_nmask, void* stream) {
  clear_error();
  SBWT_REQUIRE(d_packed, "d_packed must not be NULL");
  SBWT_REQUIRE(d_chars || n_chars == 0, "d_chars must not be NULL");
  const u64 pw = sbwt_packed_words(n_chars);
  const u64 mh = 2 * sbwt_nmask_words(n_chars);
  const u64 threads = n_chars / 32 + 4;  // covers every packed word and mask half
  const u32 bt = 256;
  k_pack<<<(threads + bt - 1) / bt, bt, 0, as_stream(stream)>>>(d_chars, n_chars, d_packed, pw,
                                                               reinterpret_cast<u32*>(d_nmask), mh, threads);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

int sbwt_kmer_offsets(const uint64_t* d_row_starts, uint64_t rows, uint32_t k, uint64_t* d_offsets, void* stream) {
  clear_error();
  SBWT_REQUIRE(d_row_starts && d_offsets, "device buffers must not be NULL");
  SBWT_REQUIRE(k >= 1 && k <= 32, "k = %u: must be in [1, 32]", k);
  return scan_rows(KmerCount{d_row_starts, rows, k}, rows, d_offsets, as_stream(stream));
}

int sbwt_positions(const uint64_t* d_row_starts, const uint64_t* d_offsets, uint64_t rows, uint32_t k,
                   uint64_t* d_positions, void* stream) {
  clear_error();
  (void)k;
  if (rows == 0) return SBWT_OK;
  SBWT_REQUIRE(d_row_starts && d_offsets && d_positions, "device buffers must not be NULL");
  const u64 grid = rows < (1u << 20) ? rows : (1u << 20);
  k_positions<<<grid, 64, 0, as_stream(stream)>>>(d_row_starts, d_offsets, rows, d_positions);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

int sbwt_extract_kmers(const uint64_t* d_packed, const uint64_t* d_positions, uint64_t count, uint64_t* d_kmers,
                       void* stream) {
  clear_error();
  if (count == 0) return SBWT_OK;
  SBWT_REQUIRE(d_packed && d_positions && d_kmers, "device buffers must not be NULL");
  k_extract<<<(count + 255) / 256, 256, 0, as_stream(stream)>>>(d_packed, d_positions, count, d_kmers);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

uint64_t sbwt_reads_work_bound(uint64_t rows, uint64_t total_kmers, uint32_t seg) {
  if (seg == 0) seg = 256;
  return rows + total_kmers / seg + 1;
}

int sbwt_reads_work(const uint64_t* d_offsets, uint64_t rows, uint32_t seg, uint64_t* d_work, uint32_t* d_items,
                    void* stream) {
  clear_error();
  SBWT_REQUIRE(d_offsets && d_work, "device buffers must not be NULL");
  SBWT_REQUIRE(rows < 0xFFFFFFFFull, "rows = %llu: at most 2^32 - 2 rows per call", (unsigned long long)rows);
  if (seg == 0) seg = 256;
  SBWT_TRY(scan_rows(WorkCount{d_offsets, rows, seg}, rows, d_work, as_stream(stream)));
  if (d_items && rows) {
    k_items<<<(rows + 255) / 256, 256, 0, as_stream(stream)>>>(d_work, rows, d_items);
    SBWT_HIP(hipGetLastError());
  }
  return SBWT_OK;
}

}  // extern "C"
