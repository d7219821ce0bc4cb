// Binary results on the device: the found/absent bitmap of a batch's
// answers (bit q set iff answer q is a node, not ~0).
//
// No reference counterpart: the reference prints every answer as text
// (ResultsPersister::save_results_to_file,
// include/epicseq/results_persister.hpp:82-96), 2-20 bytes per k-mer; the
// bitmap is one bit per k-mer.
//
// One wave writes one 64-bit word per step: each lane loads one answer
// (a wave reads 512 consecutive bytes) and a ballot gathers the 64
// found bits. The kernel is HBM bound: 8 bytes read and 1/8 byte written
// per k-mer.
#include "common.hpp"

namespace sbwt {

namespace {

__global__ __launch_bounds__(256) void k_found_bitmap(const u64* __restrict__ res, u64 count, u32 shift,
                                                      u64 nwords, u64* __restrict__ bits,
                                                      unsigned long long* __restrict__ found) {
  const u32 lane = threadIdx.x & 63;
  const u64 waves = static_cast<u64>(gridDim.x) * (blockDim.x >> 6);
  u64 ones = 0;
  for (u64 w = blockIdx.x * static_cast<u64>(blockDim.x >> 6) + (threadIdx.x >> 6); w < nwords; w += waves) {
    // word w holds bits [64w, 64w + 64) of shift zero bits followed by the answers
    const u64 b = w * 64 + lane;
    const bool hit = b >= shift && b - shift < count && res[b - shift] != kNone;
    const u64 word = __ballot(hit);
    if (lane == 0) bits[w] = word;
    ones += __popcll(word);
  }
  if (found && lane == 0 && ones) atomicAdd(found, static_cast<unsigned long long>(ones));
}

__global__ void k_zero_word(unsigned long long* p) { *p = 0; }

}  // namespace

}  // namespace sbwt

using namespace sbwt;

extern "C" {

int sbwt_found_bitmap(const uint64_t* d_results, uint64_t count, uint32_t bit_offset, uint64_t* d_bitmap,
                      uint64_t* d_found, void* stream) {
  clear_error();
  SBWT_REQUIRE(bit_offset < 64, "bit_offset %u must be below 64", bit_offset);
  const u64 nwords = (bit_offset + count + 63) / 64;
  SBWT_REQUIRE((d_results || count == 0) && (d_bitmap || nwords == 0), "device buffers must not be NULL");
  hipStream_t s = as_stream(stream);
  auto* found = reinterpret_cast<unsigned long long*>(d_found);
  if (found) {
    k_zero_word<<<1, 1, 0, s>>>(found);
    SBWT_HIP(hipGetLastError());
  }
  if (nwords == 0) return SBWT_OK;
  const u64 want = (nwords + 3) / 4;
  const u32 grid = static_cast<u32>(want < 8192 ? want : 8192);
  k_found_bitmap<<<grid, 256, 0, s>>>(d_results, count, bit_offset, nwords, d_bitmap, found);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

}  // extern "C"
