// Index lifecycle of libsbwt_gpu_amd: upload of the four SBWT rows (and the
// suffix-group starts), the interleaved rank-chunk layout, the C array, the
// presearch table, and the bit-vector file format.
//
// Reference counterparts: epic::MatrixSBWT::create (include/matrix_sbwt.hpp:32-200),
// epic::RankDataStructures (include/rank_data_structures.hpp), epic::BitVector
// (include/bit_vector.hpp:65-128), epic::PreSearch (include/presearch.hpp:24-130)
// and the presearch kernel (include/gpu/presearch_kernel.hpp:19-37).
#include <cstdarg>
#include <fstream>
#include <thread>
#include <vector>

#include <rocprim/device/device_scan.hpp>

#include "index.hpp"

namespace sbwt {

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

void clear_error() { g_last_error.clear(); }

DevIndex dev_view(const sbwt_index* ix) {
  DevIndex v{};
  v.lines = ix->d_lines;
  v.row_q = ix->row_q;
  v.pre = ix->d_pre;
  for (int c = 0; c < 5; c++) v.C[c] = static_cast<u32>(ix->C[c]);
  v.n = static_cast<u32>(ix->n);
  v.k = ix->k;
  v.m = ix->m;
  v.has_groups = ix->has_groups;
  return v;
}

void destroy_index(sbwt_index* ix) {
  if (!ix) return;
  if (ix->d_rows) (void)hipFree(ix->d_rows);
  if (ix->d_lines) (void)hipFree(ix->d_lines);
  if (ix->d_pre) (void)hipFree(ix->d_pre);
  if (ix->d_counters) (void)hipFree(ix->d_counters);
  if (ix->d_item_ctr) (void)hipFree(ix->d_item_ctr);
  if (ix->d_select) (void)hipFree(ix->d_select);
  delete ix;
}

// ---- layout build ------------------------------------------------------

// 32 columns starting at column 32*h of an MSB-first row, returned LSB-first.
__device__ __forceinline__ u32 half_lsb(const u64* __restrict__ row, u64 nwords, u64 h) {
  const u64 w = h >> 1;
  if (w >= nwords) return 0u;
  const u64 v = row[w];
  const u32 half = (h & 1) ? static_cast<u32>(v) : static_cast<u32>(v >> 32);
  return __brev(half);
}

// Pass 1: ones per chunk (rows are zero past n).
__global__ void k_chunk_counts(const u64* __restrict__ rows, u64 nwords, u64 nchunks, u32 pw,
                               u32* __restrict__ counts) {
  const u64 ch = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  const u32 r = blockIdx.y;
  if (ch >= nchunks) return;
  const u64* row = rows + r * nwords;
  u32 cnt = 0;
  for (u32 t = 0; t < pw; t++) cnt += __popc(half_lsb(row, nwords, ch * pw + t));
  counts[r * nchunks + ch] = cnt;
}

// Pass 2: chunk = {prefix count, payload words}.
__global__ void k_chunk_write(const u64* __restrict__ rows, u64 nwords, u64 nchunks, u32 pw,
                              const u32* __restrict__ prefix, u32* __restrict__ lines, u64 row_words) {
  const u64 ch = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  const u32 r = blockIdx.y;
  if (ch >= nchunks) return;
  const u64* row = rows + r * nwords;
  u32* out = lines + r * row_words + ch * (pw + 1);
  out[0] = prefix[r * nchunks + ch];
  for (u32 t = 0; t < pw; t++) out[1 + t] = half_lsb(row, nwords, ch * pw + t);
}

// rank(n) of each row = the C array increments.
template <int CB>
__global__ void k_row_totals(DevIndex ix, u32* out) {
  const u32 c = threadIdx.x;
  if (c < 4) out[c] = rank_dev<CB>(ix, c, ix.n);
}

template <int CB>
static int row_totals(const sbwt_index* ix, u32* d_out, hipStream_t s) {
  k_row_totals<CB><<<1, 64, 0, s>>>(dev_view(ix), d_out);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

int build_layout(sbwt_index* ix, hipStream_t s) {
  const u32 cb = ix->layout;
  if (cb != 16 && cb != 32 && cb != 64)
    return fail(SBWT_ERR_ARG, "unknown rank layout %u (use 16, 32 or 64)", cb);
  const u32 pw = cb / 4 - 1;
  const u32 cols = pw * 32;
  const u64 nchunks = chunks_for(ix->n, cols);
  // rows start on 128-byte lines, so chunk ch of any row lies in line ch*cb/128
  const u64 row_q = ((nchunks * (cb / 16) + 7) / 8) * 8;
  if (ix->d_lines) {
    SBWT_HIP(hipFree(ix->d_lines));
    ix->d_lines = nullptr;
  }
  SBWT_HIP(hipMalloc(&ix->d_lines, kRows * row_q * sizeof(uint4)));
  SBWT_HIP(hipMemsetAsync(ix->d_lines, 0, kRows * row_q * sizeof(uint4), s));
  ix->row_q = row_q;

  u32 *d_counts = nullptr, *d_prefix = nullptr;
  void* d_tmp = nullptr;
  size_t tmp_bytes = 0;
  int st = SBWT_OK;
  do {
    if (hipMalloc(&d_counts, kRows * nchunks * sizeof(u32)) != hipSuccess ||
        hipMalloc(&d_prefix, kRows * nchunks * sizeof(u32)) != hipSuccess) {
      st = fail(SBWT_ERR_HIP, "out of device memory for the rank layout (%llu chunks)",
                (unsigned long long)nchunks);
      break;
    }
    dim3 grid(static_cast<u32>((nchunks + 255) / 256), kRows);
    k_chunk_counts<<<grid, 256, 0, s>>>(ix->d_rows, ix->nwords, nchunks, pw, d_counts);
    if (hipGetLastError() != hipSuccess) { st = fail(SBWT_ERR_HIP, "chunk count launch failed"); break; }
    if (rocprim::exclusive_scan(nullptr, tmp_bytes, d_counts, d_prefix, 0u, nchunks, rocprim::plus<u32>(), s) !=
            hipSuccess ||
        hipMalloc(&d_tmp, tmp_bytes ? tmp_bytes : 1) != hipSuccess) {
      st = fail(SBWT_ERR_HIP, "scan setup failed");
      break;
    }
    for (u32 r = 0; r < kRows && st == SBWT_OK; r++)
      if (rocprim::exclusive_scan(d_tmp, tmp_bytes, d_counts + r * nchunks, d_prefix + r * nchunks, 0u, nchunks,
                                  rocprim::plus<u32>(), s) != hipSuccess)
        st = fail(SBWT_ERR_HIP, "chunk prefix scan failed");
    if (st != SBWT_OK) break;
    k_chunk_write<<<grid, 256, 0, s>>>(ix->d_rows, ix->nwords, nchunks, pw, d_prefix,
                                       reinterpret_cast<u32*>(ix->d_lines), row_q * 4);
    if (hipGetLastError() != hipSuccess) { st = fail(SBWT_ERR_HIP, "chunk write launch failed"); break; }
    if (hipStreamSynchronize(s) != hipSuccess) st = fail(SBWT_ERR_HIP, "layout build failed");
  } while (0);
  if (d_tmp) (void)hipFree(d_tmp);
  if (d_counts) (void)hipFree(d_counts);
  if (d_prefix) (void)hipFree(d_prefix);
  return st;
}

// ---- presearch -----------------------------------------------------------

// One thread per m-mer: the first character gives [C[c], C[c+1]-1], each
// further character narrows it, with no early exit, so an empty interval
// keeps l = r + 1 (include/gpu/presearch_kernel.hpp:19-37).
template <int CB>
__global__ void k_presearch(DevIndex ix, u32 m, u64 count, uint2* __restrict__ out) {
  // grid-stride: a dispatch holds fewer than 2^32 work-items, and m >= 16 has 2^32 entries or more
  const u64 stride = gridDim.x * static_cast<u64>(blockDim.x);
  for (u64 idx = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; idx < count; idx += stride) {
    u32 c = static_cast<u32>(idx >> (2 * (m - 1))) & 3u;
    u32 l = c_of(ix, c);
    u32 r = (c == 3 ? ix.C[4] : c_of(ix, c + 1)) - 1u;
    for (int sh = 2 * static_cast<int>(m) - 4; sh >= 0; sh -= 2) {
      c = static_cast<u32>(idx >> sh) & 3u;
      narrow<CB>(ix, c, l, r);
    }
    out[idx] = make_uint2(l, r);
  }
}

static int run_presearch(sbwt_index* ix, u32 m, hipStream_t s) {
  // a table of the same depth is rewritten in place: freeing and
  // reallocating tens of GB costs the HIP runtime seconds
  if (ix->d_pre && ix->pre_alloc_m != m) {
    SBWT_HIP(hipFree(ix->d_pre));
    ix->d_pre = nullptr;
  }
  ix->m = 0;
  if (m == 0) return SBWT_OK;
  const u64 count = 1ull << (2 * m);
  if (!ix->d_pre) {
    SBWT_HIP(hipMalloc(&ix->d_pre, count * sizeof(uint2)));
    ix->pre_alloc_m = m;
  } else {
    SBWT_HIP(hipDeviceSynchronize());  // searches still reading the old table end first, as under hipFree
  }
  DevIndex v = dev_view(ix);
  const u32 bt = 256;
  const u64 want = (count + bt - 1) / bt;
  const u64 grid = want < (1ull << 22) ? want : (1ull << 22);
  switch (ix->layout) {
    case 16: k_presearch<16><<<grid, bt, 0, s>>>(v, m, count, ix->d_pre); break;
    case 32: k_presearch<32><<<grid, bt, 0, s>>>(v, m, count, ix->d_pre); break;
    default: k_presearch<64><<<grid, bt, 0, s>>>(v, m, count, ix->d_pre); break;
  }
  SBWT_HIP(hipGetLastError());
  SBWT_HIP(hipStreamSynchronize(s));
  ix->m = m;
  return SBWT_OK;
}

// ---- finishing an index ----------------------------------------------------

// Clears the bits past n of every row, so that rank(n) counts exactly the
// n columns whatever the caller's padding held.
// Column 0 (the root) always starts a suffix group; setting its mark
// guarantees that the backward scan for a group start ends.
__global__ void k_clear_tail(u64* rows, u64 nwords, u64 n, u32 mark_root) {
  const u64 w = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  const u32 r = blockIdx.y;
  if (w >= nwords) return;
  u64* row = rows + r * nwords;
  const u64 first = w * 64;
  if (first >= n) row[w] = 0;
  else if (first + 64 > n) row[w] &= ~(~0ull >> (n - first));
  if (mark_root && r == kGroupRow && w == 0) row[0] |= 1ull << 63;
}

int finish_index(sbwt_index* ix, hipStream_t s, bool verify) {
  SBWT_HIP(hipMalloc(&ix->d_counters, kCounters * sizeof(unsigned long long)));
  SBWT_HIP(hipMalloc(&ix->d_item_ctr, sizeof(unsigned long long)));
  SBWT_HIP(hipDeviceGetAttribute(&ix->num_cus, hipDeviceAttributeMultiprocessorCount, ix->device));
  SBWT_HIP(hipMemsetAsync(ix->d_counters, 0, kCounters * sizeof(unsigned long long), s));
  {
    dim3 grid(static_cast<u32>((ix->nwords + 255) / 256), kRows);
    k_clear_tail<<<grid, 256, 0, s>>>(ix->d_rows, ix->nwords, ix->n, ix->has_groups);
    SBWT_HIP(hipGetLastError());
  }
  SBWT_TRY(build_layout(ix, s));
  u32* d_tot = nullptr;
  SBWT_HIP(hipMalloc(&d_tot, 4 * sizeof(u32)));
  ix->C[4] = ix->n;  // provisional, for the view
  int st;
  switch (ix->layout) {
    case 16: st = row_totals<16>(ix, d_tot, s); break;
    case 32: st = row_totals<32>(ix, d_tot, s); break;
    default: st = row_totals<64>(ix, d_tot, s); break;
  }
  u32 tot[4] = {0, 0, 0, 0};
  if (st == SBWT_OK && hipMemcpyAsync(tot, d_tot, sizeof tot, hipMemcpyDeviceToHost, s) == hipSuccess &&
      hipStreamSynchronize(s) == hipSuccess) {
  } else if (st == SBWT_OK) {
    st = fail(SBWT_ERR_HIP, "reading the row totals failed");
  }
  (void)hipFree(d_tot);
  if (st != SBWT_OK) return st;
  // matrix_sbwt.hpp:150-155
  ix->C[0] = 1;
  for (int c = 0; c < 4; c++) ix->C[c + 1] = ix->C[c] + tot[c];
  if (ix->C[4] != ix->n)
    return fail(SBWT_ERR_FORMAT,
                "not an SBWT: C[4] = 1 + total ones = %llu but the rows have n = %llu columns",
                (unsigned long long)ix->C[4], (unsigned long long)ix->n);
  // suffix groups: a given row is checked against the edges; a missing one
  // is derived from the labels, so every index can stream
  if (ix->has_groups) {
    SBWT_TRY(verify ? verify_groups(ix, s) : check_groups(ix, s));
  } else {
    SBWT_TRY(derive_groups(ix, ix->d_rows + kGroupRow * ix->nwords, s));
    ix->has_groups = 1;
    SBWT_TRY(build_layout(ix, s));
  }
  u32 m = ix->cfg.presearch_depth;
  if (m > ix->k) m = ix->k;
  return run_presearch(ix, m, s);
}

static int check_config(const sbwt_config* cfg) {
  if (!cfg) return SBWT_OK;
  if (cfg->presearch_depth > kMaxPresearch)
    return fail(SBWT_ERR_ARG, "presearch depth %u exceeds %u", cfg->presearch_depth, kMaxPresearch);
  if (cfg->layout != 0 && cfg->layout != 16 && cfg->layout != 32 && cfg->layout != 64)
    return fail(SBWT_ERR_ARG, "unknown rank layout %u (use 16, 32 or 64)", cfg->layout);
  return SBWT_OK;
}

sbwt_index* new_index(const sbwt_config* cfg, u32 k, u64 n) {
  sbwt_index* ix = new sbwt_index();
  sbwt_config_init(&ix->cfg);
  if (cfg) ix->cfg = *cfg;
  if (ix->cfg.layout == 0) ix->cfg.layout = SBWT_LAYOUT_L16;
  ix->layout = ix->cfg.layout;
  ix->k = k;
  ix->n = n;
  ix->nwords = words_for(n);
  (void)hipGetDevice(&ix->device);
  return ix;
}

int check_shape(u64 n, u32 k) {
  if (k < 1 || k > 32) return fail(SBWT_ERR_ARG, "k = %u: must be in [1, 32]", k);
  if (n == 0) return fail(SBWT_ERR_FORMAT, "an SBWT has at least the root column (n = 0)");
  if (n > 0xFFFFFFFEull)
    return fail(SBWT_ERR_UNSUPPORTED, "n = %llu columns: this build supports n < 2^32 - 1",
                (unsigned long long)n);
  return SBWT_OK;
}

}  // namespace sbwt

using namespace sbwt;

extern "C" {

int sbwt_config_init(sbwt_config* cfg) {
  if (!cfg) return fail(SBWT_ERR_ARG, "cfg is NULL");
  *cfg = sbwt_config{};
  cfg->presearch_depth = 12;  // include/parameters.hpp:50
  cfg->layout = SBWT_LAYOUT_L16;
  cfg->block_threads = 0;
  cfg->segment_kmers = 0;
  cfg->variant = 0;
  return SBWT_OK;
}

const char* sbwt_last_error(void) { return g_last_error.c_str(); }

int sbwt_index_from_rows(const uint64_t* const rows[4], const uint64_t* group, uint64_t n, uint32_t k,
                         const sbwt_config* cfg, sbwt_index** out) {
  clear_error();
  SBWT_REQUIRE(rows && out, "rows and out must not be NULL");
  for (int c = 0; c < 4; c++) SBWT_REQUIRE(rows[c], "row %d is NULL", c);
  SBWT_TRY(check_shape(n, k));
  SBWT_TRY(check_config(cfg));
  *out = nullptr;
  sbwt_index* ix = new_index(cfg, k, n);
  const u64 used = (n + 63) / 64;
  if (hipMalloc(&ix->d_rows, kRows * ix->nwords * sizeof(u64)) != hipSuccess ||
      hipMemset(ix->d_rows, 0, kRows * ix->nwords * sizeof(u64)) != hipSuccess) {
    destroy_index(ix);
    return fail(SBWT_ERR_HIP, "out of device memory for %llu-column rows", (unsigned long long)n);
  }
  for (int c = 0; c < 5; c++) {
    const uint64_t* src = c < 4 ? rows[c] : group;
    if (!src) continue;
    if (hipMemcpy(ix->d_rows + c * ix->nwords, src, used * sizeof(u64), hipMemcpyHostToDevice) != hipSuccess) {
      destroy_index(ix);
      return fail(SBWT_ERR_HIP, "row upload failed");
    }
  }
  ix->has_groups = group ? 1u : 0u;
  int st = finish_index(ix, nullptr, /*verify_groups=*/true);
  if (st != SBWT_OK) {
    destroy_index(ix);
    return st;
  }
  *out = ix;
  return SBWT_OK;
}

// include/bit_vector.hpp:65-109: [u64 little-endian bit count][ceil(n/8) bytes]
static int read_bv_header(const char* path, u64* n) {
  std::ifstream f(path, std::ios::binary | std::ios::ate);
  if (!f.is_open()) return fail(SBWT_ERR_IO, "the file %s can not be opened", path);
  const u64 size = static_cast<u64>(f.tellg());
  if (size < 8) return fail(SBWT_ERR_IO, "the file %s is too small to hold a bit vector", path);
  f.seekg(0);
  unsigned char h[8];
  f.read(reinterpret_cast<char*>(h), 8);
  u64 v = 0;
  for (int i = 0; i < 8; i++) v |= static_cast<u64>(h[i]) << (8 * i);
  if (v == 0) return fail(SBWT_ERR_FORMAT, "the file %s holds a bit vector of size 0", path);
  if (size - 8 < (v + 7) / 8)
    return fail(SBWT_ERR_IO, "the file %s declares %llu bits but holds only %llu data bytes", path,
                (unsigned long long)v, (unsigned long long)(size - 8));
  *n = v;
  return SBWT_OK;
}

// Reads ceil(n/8) bytes and turns them into MSB-first words: bit i is file
// byte 8+(i>>3), bit 7-(i&7) (include/bit_vector.hpp:111-128 with the byte
// swap of include/rank_data_structures.hpp:98).
static bool read_bv_words(const char* path, u64 n, u64* words, u64 nwords) {
  std::ifstream f(path, std::ios::binary);
  if (!f.is_open()) return false;
  f.seekg(8);
  std::fill(words, words + nwords, 0ull);
  const u64 nbytes = (n + 7) / 8;
  f.read(reinterpret_cast<char*>(words), static_cast<std::streamsize>(nbytes));
  if (static_cast<u64>(f.gcount()) != nbytes) return false;
  const u64 used = (n + 63) / 64;
  for (u64 w = 0; w < used; w++) words[w] = __builtin_bswap64(words[w]);
  if (n & 63) words[n >> 6] &= ~(~0ull >> (n & 63));
  return true;
}

int sbwt_index_from_files(const char* const paths[4], const char* group_path, uint32_t k, const sbwt_config* cfg,
                          sbwt_index** out) {
  clear_error();
  SBWT_REQUIRE(paths && out, "paths and out must not be NULL");
  u64 n[5] = {0, 0, 0, 0, 0};
  for (int c = 0; c < 4; c++) {
    SBWT_REQUIRE(paths[c], "path %d is NULL", c);
    SBWT_TRY(read_bv_header(paths[c], &n[c]));
  }
  if (group_path) SBWT_TRY(read_bv_header(group_path, &n[4]));
  // include/matrix_sbwt.hpp:52-55
  for (int c = 1; c < 4; c++)
    if (n[c] != n[0])
      return fail(SBWT_ERR_FORMAT, "bit vectors differ in length: %s has %llu bits, %s has %llu", paths[0],
                  (unsigned long long)n[0], paths[c], (unsigned long long)n[c]);
  if (group_path && n[4] != n[0])
    return fail(SBWT_ERR_FORMAT, "group vector %s has %llu bits, rows have %llu", group_path,
                (unsigned long long)n[4], (unsigned long long)n[0]);
  const u64 nw = words_for(n[0]);
  std::vector<std::vector<u64>> rows(5);
  bool ok[5] = {true, true, true, true, true};
  std::vector<std::thread> th;
  const int nrows = group_path ? 5 : 4;
  for (int c = 0; c < nrows; c++) {
    rows[c].resize(nw ? nw : 1);
    th.emplace_back([&, c] { ok[c] = read_bv_words(c < 4 ? paths[c] : group_path, n[0], rows[c].data(), nw); });
  }
  for (auto& t : th) t.join();
  for (int c = 0; c < nrows; c++)
    if (!ok[c]) return fail(SBWT_ERR_IO, "reading %s failed", c < 4 ? paths[c] : group_path);
  const uint64_t* rp[4] = {rows[0].data(), rows[1].data(), rows[2].data(), rows[3].data()};
  return sbwt_index_from_rows(rp, group_path ? rows[4].data() : nullptr, n[0], k, cfg, out);
}

int sbwt_index_free(sbwt_index* ix) {
  clear_error();
  destroy_index(ix);
  return SBWT_OK;
}

int sbwt_index_get_info(const sbwt_index* ix, sbwt_index_info* info) {
  SBWT_REQUIRE(ix && info, "index and info must not be NULL");
  *info = sbwt_index_info{};
  info->n = ix->n;
  for (int c = 0; c < 5; c++) info->C[c] = ix->C[c];
  info->k = ix->k;
  info->presearch_depth = ix->m;
  info->layout = ix->layout;
  info->has_groups = ix->has_groups;
  info->device_bytes = kRows * ix->nwords * 8 + kRows * ix->row_q * 16 + (ix->m ? (8ull << (2 * ix->m)) : 0);
  return SBWT_OK;
}

int sbwt_index_get_rows(const sbwt_index* ix, uint64_t* const rows[4], uint64_t* group, uint64_t nwords) {
  clear_error();
  SBWT_REQUIRE(ix && rows, "index and rows must not be NULL");
  const u64 used = (ix->n + 63) / 64;
  SBWT_REQUIRE(nwords >= used, "nwords = %llu < ceil(n/64) = %llu", (unsigned long long)nwords,
               (unsigned long long)used);
  const u64 cp = nwords < ix->nwords ? nwords : ix->nwords;
  for (int c = 0; c < 5; c++) {
    uint64_t* dst = c < 4 ? rows[c] : group;
    if (!dst) continue;
    SBWT_HIP(hipMemcpy(dst, ix->d_rows + c * ix->nwords, cp * sizeof(u64), hipMemcpyDeviceToHost));
    for (u64 w = cp; w < nwords; w++) dst[w] = 0;
  }
  return SBWT_OK;
}

int sbwt_index_write_files(const sbwt_index* ix, const char* const paths[4], const char* group_path) {
  clear_error();
  SBWT_REQUIRE(ix && paths, "index and paths must not be NULL");
  std::vector<u64> w(ix->nwords ? ix->nwords : 1);
  for (int c = 0; c < 5; c++) {
    const char* p = c < 4 ? paths[c] : group_path;
    if (!p) continue;
    if (c == 4 && !ix->has_groups) return fail(SBWT_ERR_ARG, "the index has no suffix-group row");
    SBWT_HIP(hipMemcpy(w.data(), ix->d_rows + c * ix->nwords, ix->nwords * sizeof(u64), hipMemcpyDeviceToHost));
    std::ofstream f(p, std::ios::binary);
    if (!f.is_open()) return fail(SBWT_ERR_IO, "the file %s can not be created", p);
    unsigned char h[8];
    for (int i = 0; i < 8; i++) h[i] = static_cast<unsigned char>(ix->n >> (8 * i));
    f.write(reinterpret_cast<const char*>(h), 8);
    const u64 nbytes = (ix->n + 7) / 8;
    std::vector<u64> be((nbytes + 7) / 8);
    for (size_t i = 0; i < be.size(); i++) be[i] = __builtin_bswap64(w[i]);
    f.write(reinterpret_cast<const char*>(be.data()), static_cast<std::streamsize>(nbytes));
    if (!f.good()) return fail(SBWT_ERR_IO, "writing %s failed", p);
  }
  return SBWT_OK;
}

int sbwt_presearch(sbwt_index* ix, uint32_t m, void* stream) {
  clear_error();
  SBWT_REQUIRE(ix, "index is NULL");
  SBWT_REQUIRE(m <= kMaxPresearch && m <= ix->k, "presearch depth %u must be <= min(k, %u)", m, kMaxPresearch);
  return run_presearch(ix, m, as_stream(stream));
}

int sbwt_presearch_get(const sbwt_index* ix, uint64_t* left, uint64_t* right) {
  clear_error();
  SBWT_REQUIRE(ix && left && right, "arguments must not be NULL");
  if (ix->m == 0) return SBWT_OK;
  const u64 count = 1ull << (2 * ix->m);
  std::vector<uint2> t(count);
  SBWT_HIP(hipMemcpy(t.data(), ix->d_pre, count * sizeof(uint2), hipMemcpyDeviceToHost));
  for (u64 i = 0; i < count; i++) {
    left[i] = t[i].x;
    right[i] = t[i].y;
  }
  return SBWT_OK;
}

int sbwt_counters(const sbwt_index* ix, uint64_t out[4], int reset, void* stream) {
  clear_error();
  SBWT_REQUIRE(ix && out, "index and out must not be NULL");
  hipStream_t s = as_stream(stream);
  SBWT_HIP(hipMemcpyAsync(out, ix->d_counters, kCounters * sizeof(u64), hipMemcpyDeviceToHost, s));
  if (reset) SBWT_HIP(hipMemsetAsync(ix->d_counters, 0, kCounters * sizeof(u64), s));
  SBWT_HIP(hipStreamSynchronize(s));
  return SBWT_OK;
}

int sbwt_index_set_layout(sbwt_index* ix, uint32_t layout, void* stream) {
  clear_error();
  SBWT_REQUIRE(ix, "index is NULL");
  SBWT_REQUIRE(layout == 16 || layout == 32 || layout == 64, "unknown rank layout %u", layout);
  ix->layout = layout;
  ix->cfg.layout = layout;
  return build_layout(ix, as_stream(stream));
}

}  // extern "C"
