// Index images: checkpoint and resume of a device-resident index.
//
// No reference counterpart: the reference rebuilds its index from the four
// bit-vector files at every start (epic::MatrixSBWT::create,
// include/matrix_sbwt.hpp:32-200, reading include/bit_vector.hpp:65-109).
// An image is one file: a 128-byte header and the five rows (A, C, G, T and
// the suffix-group starts) exactly as they sit in HBM, so that loading is
// one stream of large copies. The rank lines and the presearch table are
// not stored: the GPU derives them from the rows faster than a disk reads
// them (a 3 Gbp index: 1.9 GB of rows; its lines take about 2 GB and its
// depth-17 table 137 GB).
//
// Both directions stage through two pinned 64 MiB buffers, so the disk and
// the PCIe copies overlap. A checksum of the rows, computed on the device
// at save time and again after loading, catches a damaged file.
#include <fcntl.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstring>

#include "index.hpp"

namespace sbwt {

namespace {

constexpr char kImageMagic[8] = {'S', 'B', 'W', 'T', 'I', 'M', 'G', '1'};
constexpr u32 kImageVersion = 1;
constexpr u64 kStage = 64ull << 20;

// The file header, little-endian as the host writes it.
struct ImageHeader {
  char magic[8];
  u32 version;
  u32 k;
  u64 n;
  u64 nwords;  // words per row
  u64 C[5];
  u32 has_groups;
  u32 presearch_depth;  // of the saved index
  sbwt_config cfg;      // with the saved index's layout
  u64 num_kmers;
  u64 checksum;  // of the rows (image_checksum)
};
static_assert(sizeof(ImageHeader) == 128, "the image header is 128 bytes");

// splitmix64's output function (Steele, Lea and Flood, "Fast splittable
// pseudorandom number generators", OOPSLA 2014)
__device__ __forceinline__ u64 mix(u64 x) {
  x ^= x >> 30;
  x *= 0xBF58476D1CE4E5B9ull;
  x ^= x >> 27;
  x *= 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// sum over words of mix(word + index * golden ratio), mod 2^64: a change of
// any word, or an exchange of two words, changes the sum
__global__ __launch_bounds__(256) void k_checksum(const u64* __restrict__ p, u64 n, unsigned long long* out) {
  u64 acc = 0;
  const u64 stride = static_cast<u64>(gridDim.x) * blockDim.x;
  for (u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x; i < n; i += stride)
    acc += mix(p[i] + i * 0x9E3779B97F4A7C15ull);
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, static_cast<unsigned long long>(acc));
}

__global__ void k_zero(unsigned long long* p) { *p = 0; }

// The checksum of n words at p (synchronous, the null stream).
static int words_checksum(const u64* p, u64 n, u64* out) {
  unsigned long long* d = nullptr;
  SBWT_HIP(hipMalloc(&d, sizeof *d));
  const u64 want = (n + 255) / 256;
  k_zero<<<1, 1>>>(d);
  if (n) k_checksum<<<static_cast<u32>(want < 16384 ? want : 16384), 256>>>(p, n, d);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(out, d, sizeof *out, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(SBWT_ERR_HIP, "a checksum of the index failed: %s", hipGetErrorString(e));
  return SBWT_OK;
}

int rows_checksum(const sbwt_index* ix, u64* out) { return words_checksum(ix->d_rows, kRows * ix->nwords, out); }

// Two pinned staging buffers and the events of their copies.
struct Staging {
  char* buf[2] = {nullptr, nullptr};
  hipEvent_t done[2] = {nullptr, nullptr};
  hipStream_t s = nullptr;
  int fd = -1;
  ~Staging() {
    for (int i = 0; i < 2; i++) {
      if (buf[i]) (void)hipHostFree(buf[i]);
      if (done[i]) (void)hipEventDestroy(done[i]);
    }
    if (s) (void)hipStreamDestroy(s);
    if (fd >= 0) close(fd);
  }
  int init() {
    SBWT_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (int i = 0; i < 2; i++) {
      SBWT_HIP(hipHostMalloc(reinterpret_cast<void**>(&buf[i]), kStage, hipHostMallocDefault));
      SBWT_HIP(hipEventCreateWithFlags(&done[i], hipEventDisableTiming));
    }
    return SBWT_OK;
  }
};

int write_all(int fd, const void* p, u64 n, const char* path) {
  const char* c = static_cast<const char*>(p);
  for (u64 o = 0; o < n;) {
    const ssize_t w = write(fd, c + o, n - o);
    if (w <= 0) return fail(SBWT_ERR_IO, "writing %s failed", path);
    o += static_cast<u64>(w);
  }
  return SBWT_OK;
}

int read_all(int fd, void* p, u64 n, const char* path) {
  char* c = static_cast<char*>(p);
  for (u64 o = 0; o < n;) {
    const ssize_t r = read(fd, c + o, n - o);
    if (r <= 0) return fail(SBWT_ERR_IO, "%s is shorter than its header says", path);
    o += static_cast<u64>(r);
  }
  return SBWT_OK;
}

// device -> file, chunk i + 1 copying while chunk i is written
int stream_out(Staging& st, const char* d, u64 bytes, const char* path) {
  const u64 chunks = (bytes + kStage - 1) / kStage;
  for (u64 i = 0; i <= chunks; i++) {
    if (i < chunks) {
      const u64 len = i + 1 < chunks ? kStage : bytes - i * kStage;
      SBWT_HIP(hipMemcpyAsync(st.buf[i & 1], d + i * kStage, len, hipMemcpyDeviceToHost, st.s));
      SBWT_HIP(hipEventRecord(st.done[i & 1], st.s));
    }
    if (i > 0) {
      const u64 j = i - 1;
      SBWT_HIP(hipEventSynchronize(st.done[j & 1]));
      SBWT_TRY(write_all(st.fd, st.buf[j & 1], j + 1 < chunks ? kStage : bytes - j * kStage, path));
    }
  }
  return SBWT_OK;
}

// file -> device, chunk i + 1 read while chunk i copies
int stream_in(Staging& st, char* d, u64 bytes, const char* path) {
  const u64 chunks = (bytes + kStage - 1) / kStage;
  for (u64 i = 0; i < chunks; i++) {
    const u64 len = i + 1 < chunks ? kStage : bytes - i * kStage;
    if (i >= 2) SBWT_HIP(hipEventSynchronize(st.done[i & 1]));  // chunk i - 2 has left this buffer
    SBWT_TRY(read_all(st.fd, st.buf[i & 1], len, path));
    SBWT_HIP(hipMemcpyAsync(d + i * kStage, st.buf[i & 1], len, hipMemcpyHostToDevice, st.s));
    SBWT_HIP(hipEventRecord(st.done[i & 1], st.s));
  }
  SBWT_HIP(hipStreamSynchronize(st.s));
  return SBWT_OK;
}

}  // namespace

}  // namespace sbwt

using namespace sbwt;

extern "C" {

int sbwt_index_save(const sbwt_index* ix, const char* path) {
  clear_error();
  SBWT_REQUIRE(ix && path, "index and path must not be NULL");
  SBWT_HIP(hipDeviceSynchronize());
  ImageHeader h{};
  memcpy(h.magic, kImageMagic, sizeof h.magic);
  h.version = kImageVersion;
  h.k = ix->k;
  h.n = ix->n;
  h.nwords = ix->nwords;
  for (int c = 0; c < 5; c++) h.C[c] = ix->C[c];
  h.has_groups = ix->has_groups;
  h.presearch_depth = ix->m;
  h.cfg = ix->cfg;
  h.cfg.layout = ix->layout;
  h.cfg.presearch_depth = ix->m;
  h.num_kmers = ix->num_kmers;
  SBWT_TRY(rows_checksum(ix, &h.checksum));
  Staging st;
  SBWT_TRY(st.init());
  st.fd = open(path, O_WRONLY | O_CREAT | O_TRUNC, 0644);
  if (st.fd < 0) return fail(SBWT_ERR_IO, "the file %s can not be created", path);
  SBWT_TRY(write_all(st.fd, &h, sizeof h, path));
  SBWT_TRY(stream_out(st, reinterpret_cast<const char*>(ix->d_rows), kRows * ix->nwords * 8, path));
  if (close(st.fd) != 0) {
    st.fd = -1;
    return fail(SBWT_ERR_IO, "closing %s failed", path);
  }
  st.fd = -1;
  return SBWT_OK;
}

int sbwt_index_load(const char* path, const sbwt_config* cfg, sbwt_index** out) {
  clear_error();
  SBWT_REQUIRE(path && out, "path and out must not be NULL");
  *out = nullptr;
  Staging st;
  st.fd = open(path, O_RDONLY);
  if (st.fd < 0) return fail(SBWT_ERR_IO, "the file %s can not be opened", path);
  struct stat sb {};
  if (fstat(st.fd, &sb) != 0) return fail(SBWT_ERR_IO, "the size of %s is unknown", path);
  ImageHeader h{};
  if (static_cast<u64>(sb.st_size) < sizeof h) return fail(SBWT_ERR_FORMAT, "%s is not an index image", path);
  SBWT_TRY(read_all(st.fd, &h, sizeof h, path));
  if (memcmp(h.magic, kImageMagic, sizeof h.magic) != 0)
    return fail(SBWT_ERR_FORMAT, "%s is not an index image (no SBWTIMG1 header)", path);
  if (h.version != kImageVersion)
    return fail(SBWT_ERR_UNSUPPORTED, "%s is an image of version %u; this build reads version %u", path, h.version,
                kImageVersion);
  SBWT_TRY(check_shape(h.n, h.k));
  if (h.nwords != words_for(h.n) || h.has_groups > 1)
    return fail(SBWT_ERR_FORMAT, "%s: the header's row size does not fit n = %llu", path,
                (unsigned long long)h.n);
  const u64 bytes = kRows * h.nwords * 8;
  if (static_cast<u64>(sb.st_size) != sizeof h + bytes)
    return fail(SBWT_ERR_FORMAT, "%s holds %llu bytes; its header says %llu", path,
                (unsigned long long)sb.st_size, (unsigned long long)(sizeof h + bytes));
  // the saved configuration, unless the caller gives one
  sbwt_config use = cfg ? *cfg : h.cfg;
  if (use.presearch_depth > kMaxPresearch)
    return fail(SBWT_ERR_ARG, "presearch depth %u exceeds %u", use.presearch_depth, kMaxPresearch);
  if (use.layout != 0 && use.layout != 16 && use.layout != 32 && use.layout != 64)
    return fail(SBWT_ERR_FORMAT, "unknown rank layout %u", use.layout);
  SBWT_TRY(st.init());
  sbwt_index* ix = new_index(&use, h.k, h.n);
  ix->has_groups = h.has_groups;
  ix->num_kmers = h.num_kmers;
  int rc = SBWT_OK;
  u64 sum = 0;
  if (hipMalloc(&ix->d_rows, bytes) != hipSuccess)
    rc = fail(SBWT_ERR_HIP, "out of device memory for %llu-column rows", (unsigned long long)h.n);
  if (rc == SBWT_OK) rc = stream_in(st, reinterpret_cast<char*>(ix->d_rows), bytes, path);
  if (rc == SBWT_OK) rc = rows_checksum(ix, &sum);
  if (rc == SBWT_OK && sum != h.checksum)
    rc = fail(SBWT_ERR_FORMAT, "%s is damaged: the rows' checksum differs from the header's", path);
  if (rc == SBWT_OK) rc = finish_index(ix, nullptr);
  for (int c = 0; rc == SBWT_OK && c < 5; c++)
    if (ix->C[c] != h.C[c]) rc = fail(SBWT_ERR_FORMAT, "%s: the rows give C[%d] = %llu, the header %llu", path, c,
                                      (unsigned long long)ix->C[c], (unsigned long long)h.C[c]);
  if (rc != SBWT_OK) {
    destroy_index(ix);
    return rc;
  }
  *out = ix;
  return SBWT_OK;
}

int sbwt_index_digest(const sbwt_index* ix, uint64_t out[3]) {
  clear_error();
  SBWT_REQUIRE(ix && out, "index and out must not be NULL");
  SBWT_HIP(hipDeviceSynchronize());  // searches still running on other streams end first
  SBWT_TRY(rows_checksum(ix, &out[0]));
  SBWT_TRY(words_checksum(reinterpret_cast<const u64*>(ix->d_lines), kRows * ix->row_q * 2, &out[1]));
  out[2] = 0;
  if (ix->m && ix->d_pre) SBWT_TRY(words_checksum(reinterpret_cast<const u64*>(ix->d_pre), 1ull << (2 * ix->m), &out[2]));
  return SBWT_OK;
}

}  // extern "C"
