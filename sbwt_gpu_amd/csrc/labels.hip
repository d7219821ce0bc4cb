// Node labels and suffix groups from the four SBWT rows (no reference
// counterpart: the reference reads no group row and has no enumeration).
//
// Every node d > 0 has exactly one incoming edge. An edge labelled c leaves
// column i (bit i of row c) and enters node C[c] + rank_c(i), so one pass
// over the set bits of the four rows gives pred[d], the source column of
// d's edge (pred[0] = 0, the root). The last character of node d's label is
// the c with C[c] <= d < C[c+1], and the label of pred[d] ends with the
// first k-1 characters of d's label, so walking pred spells d's label from
// its end; reaching the root after i steps means the label is $^{k-i}
// followed by i real characters.
//
// Suffix groups: node d starts a group unless its last k-1 characters
// (with '$' padding) equal those of node d-1. Both labels are spelt
// backwards in lockstep until they differ; adjacent nodes of a colex order
// share about log4(n) characters, so a walk stops after about that many
// steps.
#include "index.hpp"

namespace sbwt {

__device__ __forceinline__ u32 last_char(const DevIndex& ix, u32 d) {
  return d >= ix.C[3] ? 3u : (d >= ix.C[2] ? 2u : (d >= ix.C[1] ? 1u : 0u));
}

// One thread per 64-column word of a row: pred of the nodes its 1-bits enter.
template <int CB>
__global__ void k_pred(DevIndex ix, const u64* __restrict__ rows, u64 nwords, u32* __restrict__ pred) {
  const u64 w = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  const u32 c = blockIdx.y;
  if (w == 0 && c == 0) pred[0] = 0;
  if (w >= nwords || w * 64 >= ix.n) return;
  u64 x = rows[c * nwords + w];
  if (!x) return;
  u32 node = c_of(ix, c) + rank_dev<CB>(ix, c, static_cast<u32>(w * 64));
  const u32 col0 = static_cast<u32>(w * 64);
  while (x) {
    const u32 b = static_cast<u32>(__clzll(x));
    pred[node++] = col0 + b;
    x &= ~(1ull << (63 - b));
  }
}

// Label of node j as a top-aligned k-mer, or ~0 for a '$'-padded node.
__global__ void k_labels_pred(DevIndex ix, const u32* __restrict__ pred, u64* __restrict__ out,
                              u8* __restrict__ full) {
  const u64 j = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (j >= ix.n) return;
  u32 cur = static_cast<u32>(j);
  u64 code = 0;
  bool ok = true;
  for (u32 step = 0; step < ix.k; step++) {
    if (cur == 0) {
      ok = false;
      break;
    }
    code |= static_cast<u64>(last_char(ix, cur)) << (62 - 2 * (ix.k - 1 - step));
    if (step + 1 < ix.k) cur = pred[cur];
  }
  out[j] = ok ? code : kNone;
  full[j] = ok ? 1 : 0;
}

// Whether nodes a < b differ in their last k-1 characters.
__device__ __forceinline__ bool suffix_differs(const DevIndex& ix, const u32* __restrict__ pred, u32 a, u32 b) {
  for (u32 s = 0; s + 1 < ix.k; s++) {
    const bool ra = a == 0, rb = b == 0;
    if (ra || rb) return ra != rb;  // both at the root: the rest is '$' on both sides
    if (last_char(ix, a) != last_char(ix, b)) return true;
    a = pred[a];
    b = pred[b];
  }
  return false;
}

// One lane per node, 64 nodes per wave = one MSB-first word of the group row.
__global__ __launch_bounds__(256) void k_group_marks(DevIndex ix, const u32* __restrict__ pred, u64 nwords,
                                                     u64* __restrict__ grp) {
  const u64 d = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  bool start = false;
  if (d < ix.n) start = d == 0 || suffix_differs(ix, pred, static_cast<u32>(d - 1), static_cast<u32>(d));
  const u64 mask = __ballot(start);
  const u64 w = d >> 6;
  if ((threadIdx.x & 63u) == 0 && w < nwords) grp[w] = __brevll(mask);
}

// A group row given with the index: every column with an outgoing edge, and
// column 0, must start a group (only a group's first node has edges).
__global__ void k_check_groups(const u64* __restrict__ rows, u64 nwords, unsigned long long* __restrict__ bad) {
  const u64 w = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (w >= nwords) return;
  const u64 edges = rows[w] | rows[nwords + w] | rows[2 * nwords + w] | rows[3 * nwords + w];
  const u64 g = rows[4 * nwords + w];
  u64 miss = edges & ~g;
  if (w == 0) miss |= ~g & (1ull << 63);
  if (miss) atomicAdd(bad, static_cast<unsigned long long>(__popcll(miss)));
}

int build_pred(const sbwt_index* ix, u32** out, hipStream_t s) {
  *out = nullptr;
  u32* pred = nullptr;
  if (hipMalloc(&pred, (ix->n ? ix->n : 1) * sizeof(u32)) != hipSuccess)
    return fail(SBWT_ERR_HIP, "out of device memory for %llu predecessors", (unsigned long long)ix->n);
  dim3 grid(static_cast<u32>((ix->nwords + 255) / 256), 4);
  DevIndex v = dev_view(ix);
  switch (ix->layout) {
    case 16: k_pred<16><<<grid, 256, 0, s>>>(v, ix->d_rows, ix->nwords, pred); break;
    case 32: k_pred<32><<<grid, 256, 0, s>>>(v, ix->d_rows, ix->nwords, pred); break;
    default: k_pred<64><<<grid, 256, 0, s>>>(v, ix->d_rows, ix->nwords, pred); break;
  }
  if (hipGetLastError() != hipSuccess) {
    (void)hipFree(pred);
    return fail(SBWT_ERR_HIP, "predecessor kernel launch failed");
  }
  *out = pred;
  return SBWT_OK;
}

int derive_groups(const sbwt_index* ix, u64* dst, hipStream_t s) {
  u32* pred = nullptr;
  SBWT_TRY(build_pred(ix, &pred, s));
  const u64 lanes = (ix->n + 63) / 64 * 64;
  k_group_marks<<<static_cast<u32>((lanes + 255) / 256), 256, 0, s>>>(dev_view(ix), pred, ix->nwords, dst);
  const hipError_t e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(pred);
  if (e != hipSuccess || e2 != hipSuccess) return fail(SBWT_ERR_HIP, "suffix-group derivation failed");
  return SBWT_OK;
}

__global__ void k_count_diff(const u64* __restrict__ a, const u64* __restrict__ b, u64 nwords,
                             unsigned long long* __restrict__ bad) {
  const u64 w = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (w < nwords && a[w] != b[w]) atomicAdd(bad, static_cast<unsigned long long>(__popcll(a[w] ^ b[w])));
}

int verify_groups(const sbwt_index* ix, hipStream_t s) {
  u64* mine = nullptr;
  unsigned long long* d_bad = nullptr;
  SBWT_HIP(hipMalloc(&mine, ix->nwords * sizeof(u64)));
  if (hipMalloc(&d_bad, sizeof(unsigned long long)) != hipSuccess) {
    (void)hipFree(mine);
    return fail(SBWT_ERR_HIP, "out of device memory for the group check");
  }
  unsigned long long bad = 0;
  int st = SBWT_OK;
  if (hipMemsetAsync(mine, 0, ix->nwords * sizeof(u64), s) != hipSuccess ||
      hipMemsetAsync(d_bad, 0, sizeof bad, s) != hipSuccess)
    st = fail(SBWT_ERR_HIP, "group check setup failed");
  if (st == SBWT_OK) st = derive_groups(ix, mine, s);
  if (st == SBWT_OK) {
    k_count_diff<<<static_cast<u32>((ix->nwords + 255) / 256), 256, 0, s>>>(
        mine, ix->d_rows + kGroupRow * ix->nwords, ix->nwords, d_bad);
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      st = fail(SBWT_ERR_HIP, "group check failed to run");
  }
  (void)hipFree(mine);
  (void)hipFree(d_bad);
  if (st == SBWT_OK && bad)
    st = fail(SBWT_ERR_FORMAT,
              "the suffix-group row does not belong to these rows: it differs from the row derived from their "
              "labels in %llu columns",
              bad);
  return st;
}

int check_groups(const sbwt_index* ix, hipStream_t s) {
  unsigned long long* d_bad = nullptr;
  SBWT_HIP(hipMalloc(&d_bad, sizeof(unsigned long long)));
  unsigned long long bad = 0;
  hipError_t e = hipMemsetAsync(d_bad, 0, sizeof bad, s);
  if (e == hipSuccess) {
    k_check_groups<<<static_cast<u32>((ix->nwords + 255) / 256), 256, 0, s>>>(ix->d_rows, ix->nwords, d_bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&bad, d_bad, sizeof bad, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d_bad);
  if (e != hipSuccess) return fail(SBWT_ERR_HIP, "group row check failed: %s", hipGetErrorString(e));
  if (bad)
    return fail(SBWT_ERR_FORMAT,
                "the suffix-group row does not belong to these rows: %llu columns with outgoing edges "
                "(or column 0) are not marked as group starts",
                bad);
  return SBWT_OK;
}

int node_labels(sbwt_index* ix, u64* d_labels, u8* d_full, hipStream_t s) {
  u32* pred = nullptr;
  SBWT_TRY(build_pred(ix, &pred, s));
  k_labels_pred<<<static_cast<u32>((ix->n + 255) / 256), 256, 0, s>>>(dev_view(ix), pred, d_labels, d_full);
  const hipError_t e = hipGetLastError();
  const hipError_t e2 = hipStreamSynchronize(s);  // the walk reads pred until it ends
  (void)hipFree(pred);
  if (e != hipSuccess || e2 != hipSuccess) return fail(SBWT_ERR_HIP, "label kernel failed");
  return SBWT_OK;
}

}  // namespace sbwt
