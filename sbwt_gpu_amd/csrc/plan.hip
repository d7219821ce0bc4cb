// Read-search plans: preallocated scratch for batches of rows, so that the
// whole query path of a batch (packing, k-mer offsets, work prefix, item
// map, search) allocates nothing, never synchronizes with the host and can
// be captured in a HIP graph. Replaces the per-call allocations of the
// reference's query pipeline (KmerData::allocate_memory,
// include/epicseq/kmer_data.hpp:56-69, and DeviceStartPositions,
// include/epicseq/start_positions.hpp:30-73).
#include <mutex>

#include <rocprim/device/device_scan.hpp>
#include <rocprim/iterator/counting_iterator.hpp>
#include <rocprim/iterator/transform_iterator.hpp>

#include "index.hpp"

struct sbwt_reads_plan {
  sbwt::u64 max_rows = 0, max_chars = 0;
  sbwt::u32 seg = 256;
  sbwt::u64* d_packed = nullptr;
  sbwt::u64* d_nmask = nullptr;
  sbwt::u64* d_offsets = nullptr;
  sbwt::u64* d_work = nullptr;
  sbwt::u32* d_items = nullptr;
  void* d_scan_tmp = nullptr;
  size_t scan_bytes = 0;
  // run items (SBWT_READS_RUN_ITEMS): capacity for any batch of the plan at k >= 16
  uint4* d_runs = nullptr;
  unsigned long long* d_run_count = nullptr;
  sbwt::u64 run_cap = 0;
  // the uniform-step kernel's work-item counter: one per plan, so searches
  // on different plans may overlap (on different streams) on one index
  unsigned long long* d_item_ctr = nullptr;
};

namespace sbwt {

namespace {

struct PlanKmerCount {
  const u64* rs;
  u64 rows;
  u32 k;
  __host__ __device__ u64 operator()(u64 r) const {
    if (r >= rows) return 0;
    const u64 len = rs[r + 1] - rs[r];
    return len >= k ? len - k + 1 : 0;
  }
};

struct PlanWorkCount {
  const u64* off;
  u64 rows;
  u32 seg;
  __host__ __device__ u64 operator()(u64 r) const {
    if (r >= rows) return 0;
    return (off[r + 1] - off[r] + seg - 1) / seg;
  }
};

__global__ __launch_bounds__(256) void k_plan_items(const u64* __restrict__ work, u64 rows, u32* __restrict__ items) {
  const u64 r = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (r >= rows) return;
  for (u64 w = work[r]; w < work[r + 1]; w++) items[w] = static_cast<u32>(r);
}

template <class F>
size_t scan_bytes_for(u64 n) {
  auto it = rocprim::make_transform_iterator(rocprim::make_counting_iterator<u64>(0), F{nullptr, 0, 1});
  size_t b = 0;
  if (rocprim::exclusive_scan(nullptr, b, it, static_cast<u64*>(nullptr), u64(0), n, rocprim::plus<u64>()) !=
      hipSuccess)
    return 0;
  return b;
}

template <class F>
int plan_scan(sbwt_reads_plan* p, F f, u64 rows, u64* out, hipStream_t s) {
  auto it = rocprim::make_transform_iterator(rocprim::make_counting_iterator<u64>(0), f);
  size_t b = p->scan_bytes;
  SBWT_HIP(rocprim::exclusive_scan(p->d_scan_tmp, b, it, out, u64(0), rows + 1, rocprim::plus<u64>(), s));
  return SBWT_OK;
}

}  // namespace

}  // namespace sbwt

using namespace sbwt;

extern "C" {

int sbwt_reads_plan_create(uint64_t max_rows, uint64_t max_chars, uint32_t seg, sbwt_reads_plan** out) {
  clear_error();
  SBWT_REQUIRE(out, "out must not be NULL");
  SBWT_REQUIRE(max_rows < 0xFFFFFFFFull, "max_rows = %llu: at most 2^32 - 2 rows", (unsigned long long)max_rows);
  *out = nullptr;
  if (seg == 0) seg = 256;
  auto* p = new sbwt_reads_plan();
  p->max_rows = max_rows;
  p->max_chars = max_chars;
  p->seg = seg;
  const u64 bound = sbwt_reads_work_bound(max_rows, max_chars, seg);
  size_t b1 = scan_bytes_for<PlanKmerCount>(max_rows + 1), b2 = scan_bytes_for<PlanWorkCount>(max_rows + 1);
  p->scan_bytes = (b1 > b2 ? b1 : b2) + 256;
  if (hipMalloc(&p->d_packed, sbwt_packed_words(max_chars) * 8) != hipSuccess ||
      hipMalloc(&p->d_nmask, sbwt_nmask_words(max_chars) * 8) != hipSuccess ||
      hipMalloc(&p->d_offsets, (max_rows + 1) * 8) != hipSuccess ||
      hipMalloc(&p->d_work, (max_rows + 1) * 8) != hipSuccess ||
      hipMalloc(&p->d_items, bound * 4) != hipSuccess || hipMalloc(&p->d_scan_tmp, p->scan_bytes) != hipSuccess ||
      hipMalloc(&p->d_runs, (p->run_cap = run_items_bound(max_rows, max_chars, seg)) * sizeof(uint4)) != hipSuccess ||
      hipMalloc(&p->d_run_count, sizeof(unsigned long long)) != hipSuccess ||
      hipMalloc(&p->d_item_ctr, sizeof(unsigned long long)) != hipSuccess) {
    sbwt_reads_plan_free(p);
    return fail(SBWT_ERR_HIP, "out of device memory for a plan of %llu rows, %llu characters",
                (unsigned long long)max_rows, (unsigned long long)max_chars);
  }
  *out = p;
  return SBWT_OK;
}

int sbwt_reads_plan_free(sbwt_reads_plan* p) {
  if (!p) return SBWT_OK;
  for (void* q : {static_cast<void*>(p->d_packed), static_cast<void*>(p->d_nmask), static_cast<void*>(p->d_offsets),
                  static_cast<void*>(p->d_work), static_cast<void*>(p->d_items), p->d_scan_tmp,
                  static_cast<void*>(p->d_runs), static_cast<void*>(p->d_run_count),
                  static_cast<void*>(p->d_item_ctr)})
    if (q) (void)hipFree(q);
  delete p;
  return SBWT_OK;
}

}  // extern "C"

namespace sbwt {

// The whole query path of one batch on the scratch of p (a plan, or a view of
// the static scratch): allocates nothing.
static int plan_search_impl(sbwt_reads_plan* p, const sbwt_index* ix, const uint8_t* d_chars, uint64_t n_chars,
                            const uint64_t* d_row_starts, uint64_t rows, uint64_t* d_out, uint64_t* d_offsets_out,
                            uint32_t flags, void* stream) {
  SBWT_REQUIRE(rows <= p->max_rows && n_chars <= p->max_chars,
               "batch of %llu rows / %llu characters exceeds the plan (%llu / %llu)", (unsigned long long)rows,
               (unsigned long long)n_chars, (unsigned long long)p->max_rows, (unsigned long long)p->max_chars);
  if (rows == 0) return SBWT_OK;
  SBWT_REQUIRE(d_chars && d_row_starts && d_out, "device buffers must not be NULL");
  hipStream_t s = as_stream(stream);
  u64* off = d_offsets_out ? d_offsets_out : p->d_offsets;
  SBWT_TRY(sbwt_pack(d_chars, n_chars, p->d_packed, p->d_nmask, stream));
  SBWT_TRY(plan_scan(p, PlanKmerCount{d_row_starts, rows, ix->k}, rows, off, s));
  if ((flags & SBWT_READS_RUN_ITEMS) && !(flags & SBWT_READS_N_AS_A) && ix->k >= kRunItemsMinK &&
      n_chars < 0xFFFFFFC0ull) {
    // the rows' N-free runs, each with the gap of N windows before it, as
    // work items; the same pass fills every answer with ~0 first
    SBWT_TRY(run_items(p->d_nmask, d_row_starts, off, rows, ix->k, p->seg, p->d_runs, p->run_cap, p->d_run_count,
                       d_out, s));
    return search_reads_runs(ix, p->d_packed, d_row_starts, off, rows, p->seg, nullptr, d_out, flags,
                             RunList{p->d_runs, p->d_run_count, p->run_cap}, p->d_item_ctr, s);
  }
  SBWT_TRY(plan_scan(p, PlanWorkCount{off, rows, p->seg}, rows, p->d_work, s));
  k_plan_items<<<(rows + 255) / 256, 256, 0, s>>>(p->d_work, rows, p->d_items);
  SBWT_HIP(hipGetLastError());
  const u64 bound = sbwt_reads_work_bound(rows, n_chars, p->seg);
  return search_reads_counted_on(ix, p->d_packed, p->d_nmask, d_row_starts, off, p->d_work, p->d_items, rows,
                                 bound, p->seg, nullptr, d_out, flags, p->d_item_ctr, s);
}

}  // namespace sbwt

extern "C" int sbwt_reads_plan_search(sbwt_reads_plan* p, const sbwt_index* ix, const uint8_t* d_chars,
                                      uint64_t n_chars, const uint64_t* d_row_starts, uint64_t rows, uint64_t* d_out,
                                      uint64_t* d_offsets_out, uint32_t flags, void* stream) {
  clear_error();
  SBWT_REQUIRE(p && ix, "plan and index must not be NULL");
  return plan_search_impl(p, ix, d_chars, n_chars, d_row_starts, rows, d_out, d_offsets_out, flags, stream);
}

// ---- the allocation-free variant -----------------------------------------
//
// The same query path on a file-scope device array instead of a plan, for
// callers that cannot allocate (code under HIP graph capture, serving
// loops). Each GPU holds its own copy of the array until the process ends;
// in host code its name is not a device address, so its address is taken
// once per device with hipGetSymbolAddress and kept. Every call on a GPU
// shares that GPU's array: calls that overlap on one GPU race.

namespace sbwt {
namespace {

constexpr u64 kStaticRows = SBWT_STATIC_MAX_ROWS;
constexpr u64 kStaticChars = SBWT_STATIC_MAX_CHARS;
constexpr u32 kStaticMinSeg = 32;
constexpr u64 kScanReserve = 1ull << 20;  // rocPRIM's scan temporaries (checked on first use)

constexpr u64 align256(u64 b) { return (b + 255) / 256 * 256; }
constexpr u64 kPackedBytes = align256((kStaticChars / 32 + 2) * 8);
constexpr u64 kMaskBytes = align256((kStaticChars / 64 + 2) * 8);
constexpr u64 kRowBytes = align256((kStaticRows + 1) * 8);
constexpr u64 kItemBytes = align256((kStaticRows + kStaticChars / kStaticMinSeg + 1) * 4);
constexpr u64 kRunCap = 2 * kStaticRows + kStaticChars / (kRunItemsMinK + 1) + kStaticChars / kStaticMinSeg + 64;
constexpr u64 kRunBytes = align256(kRunCap * 16);
constexpr u64 kScratchBytes =
    kPackedBytes + kMaskBytes + 2 * kRowBytes + kItemBytes + kScanReserve + kRunBytes + 2 * 256;

__device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The calling device's copy of `scratch` (set up on the device's first call).
int scratch_on_device(void** out) {
  int dev = 0;
  SBWT_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(SBWT_ERR_UNSUPPORTED, "device %d: at most %d devices", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    const size_t need = scan_bytes_for<PlanKmerCount>(kStaticRows + 1) > scan_bytes_for<PlanWorkCount>(kStaticRows + 1)
                            ? scan_bytes_for<PlanKmerCount>(kStaticRows + 1)
                            : scan_bytes_for<PlanWorkCount>(kStaticRows + 1);
    if (need > kScanReserve)
      return fail(SBWT_ERR_UNSUPPORTED, "the scan needs %zu bytes of temporaries, the static scratch holds %llu", need,
                  (unsigned long long)kScanReserve);
    void* ptr = nullptr;
    SBWT_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return SBWT_OK;
}

}  // namespace
}  // namespace sbwt

extern "C" int sbwt_reads_search_static(const sbwt_index* ix, const uint8_t* d_chars, uint64_t n_chars,
                                        const uint64_t* d_row_starts, uint64_t rows, uint32_t seg, uint64_t* d_out,
                                        uint64_t* d_offsets_out, uint32_t flags, void* stream) {
  clear_error();
  SBWT_REQUIRE(ix, "index must not be NULL");
  if (seg == 0) seg = 256;
  SBWT_REQUIRE(seg >= kStaticMinSeg && seg <= 4096, "seg = %u: must be in [%u, 4096] (0 = 256)", seg, kStaticMinSeg);
  SBWT_REQUIRE(rows <= kStaticRows && n_chars <= kStaticChars,
               "batch of %llu rows / %llu characters exceeds the static scratch (at most %llu rows, %llu characters "
               "per call)",
               (unsigned long long)rows, (unsigned long long)n_chars, (unsigned long long)kStaticRows,
               (unsigned long long)kStaticChars);
  void* base = nullptr;
  SBWT_TRY(scratch_on_device(&base));
  u8* b = static_cast<u8*>(base);
  sbwt_reads_plan v;  // a view of the static scratch: nothing here is freed
  v.max_rows = kStaticRows;
  v.max_chars = kStaticChars;
  v.seg = seg;
  v.d_packed = reinterpret_cast<u64*>(b);
  b += kPackedBytes;
  v.d_nmask = reinterpret_cast<u64*>(b);
  b += kMaskBytes;
  v.d_offsets = reinterpret_cast<u64*>(b);
  b += kRowBytes;
  v.d_work = reinterpret_cast<u64*>(b);
  b += kRowBytes;
  v.d_items = reinterpret_cast<u32*>(b);
  b += kItemBytes;
  v.d_scan_tmp = b;
  v.scan_bytes = kScanReserve;
  b += kScanReserve;
  v.d_runs = reinterpret_cast<uint4*>(b);
  v.run_cap = kRunCap;
  b += kRunBytes;
  v.d_run_count = reinterpret_cast<unsigned long long*>(b);
  b += 256;
  v.d_item_ctr = reinterpret_cast<unsigned long long*>(b);
  return plan_search_impl(&v, ix, d_chars, n_chars, d_row_starts, rows, d_out, d_offsets_out, flags, stream);
}
