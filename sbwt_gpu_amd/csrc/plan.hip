// Read-search plans: preallocated scratch for batches of rows, so that the
// whole query path of a batch (packing, k-mer offsets, work prefix, item
// map, search) allocates nothing, never synchronizes with the host and can
// be captured in a HIP graph. Replaces the per-call allocations of the
// reference's query pipeline (KmerData::allocate_memory,
// include/epicseq/kmer_data.hpp:56-69, and DeviceStartPositions,
// include/epicseq/start_positions.hpp:30-73).
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
      hipMalloc(&p->d_run_count, sizeof(unsigned long long)) != hipSuccess) {
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
                  static_cast<void*>(p->d_runs), static_cast<void*>(p->d_run_count)})
    if (q) (void)hipFree(q);
  delete p;
  return SBWT_OK;
}

int sbwt_reads_plan_search(sbwt_reads_plan* p, const sbwt_index* ix, const uint8_t* d_chars, uint64_t n_chars,
                           const uint64_t* d_row_starts, uint64_t rows, uint64_t* d_out, uint64_t* d_offsets_out,
                           uint32_t flags, void* stream) {
  clear_error();
  SBWT_REQUIRE(p && ix, "plan and index must not be NULL");
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
                             RunList{p->d_runs, p->d_run_count, p->run_cap}, s);
  }
  SBWT_TRY(plan_scan(p, PlanWorkCount{off, rows, p->seg}, rows, p->d_work, s));
  k_plan_items<<<(rows + 255) / 256, 256, 0, s>>>(p->d_work, rows, p->d_items);
  SBWT_HIP(hipGetLastError());
  const u64 bound = sbwt_reads_work_bound(rows, n_chars, p->seg);
  return sbwt_search_reads(ix, p->d_packed, p->d_nmask, d_row_starts, off, p->d_work, p->d_items, rows, bound,
                           p->seg, d_out, flags, stream);
}

}  // extern "C"
