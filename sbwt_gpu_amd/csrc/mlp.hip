// Learned negative prefilter: a small MLP over one-hot 2-bit k-mers
// (4k inputs -> 128 -> 64 -> 1, ReLU, logistic output), bf16 weights with
// FP32 accumulation, trained on the GPU by the kernels below, and a blocked
// Bloom filter over the model's false negatives so that the filter never
// rejects a k-mer of the index. No reference counterpart.
//
// Inference runs on the matrix cores: a wave takes 16 k-mers at a time and
// computes each layer transposed, D^T = W^T X^T, with 16x16x32 bf16 MFMAs,
// its weight fragments held in registers for the whole launch. The fp32
// accumulator fragment of layer 1 (lane l holds rows 4(l/16)..4(l/16)+3 of
// column l%16) becomes layer 2's B operand in place, the contraction slots
// of both operands permuted alike, so activations never leave registers.
//
// Training (forward, backward, Adam) uses plain FP32 vector kernels over a
// mini-batch; gradients are taken through the bf16-rounded weights the
// forward uses, and the master weights stay FP32.
#include <cmath>
#include <vector>

#include "common.hpp"

namespace sbwt {

constexpr u32 kH1 = 128, kH2 = 64, kMaxIn = 128;  // inputs padded to 128 (k <= 32)
constexpr u32 kGradBlocks = 256;                  // blocks of the gradient partial sums

using bf16 = u16;
typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

// Round-to-nearest-even fp32 -> bf16 (inputs are finite).
__host__ __device__ __forceinline__ bf16 to_bf16(float f) {
  u32 u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);
  return static_cast<bf16>(u >> 16);
}
__host__ __device__ __forceinline__ float from_bf16(bf16 h) {
  const u32 u = static_cast<u32>(h) << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

// Parameter offsets in the flat fp32 vector: W1[in][128], b1[128],
// W2[128][64], b2[64], w3[64], b3.
struct MlpShape {
  u32 k, in;
  u64 oW1, ob1, oW2, ob2, ow3, ob3, total;
  explicit MlpShape(u32 k_ = 31) : k(k_), in(4 * k_) {
    oW1 = 0;
    ob1 = oW1 + static_cast<u64>(in) * kH1;
    oW2 = ob1 + kH1;
    ob2 = oW2 + static_cast<u64>(kH1) * kH2;
    ow3 = ob2 + kH2;
    ob3 = ow3 + kH2;
    total = ob3 + 1;
  }
};

}  // namespace sbwt

struct sbwt_mlp {
  sbwt::MlpShape shape{31};
  sbwt::u64 max_batch = 0;
  sbwt::u64 step = 0;
  float* d_params = nullptr;  // fp32 master
  float* d_m = nullptr;       // Adam moments
  float* d_v = nullptr;
  float* d_grad = nullptr;
  sbwt::u16* d_W1T = nullptr;  // bf16 W1^T [128][128] (inputs padded with zeros)
  sbwt::u16* d_W2T = nullptr;  // bf16 W2^T [64][128]
  sbwt::u16* d_wb = nullptr;   // bf16 copies in the flat layout (forward of training)
  float* d_h1 = nullptr;       // [batch][128]
  float* d_h2 = nullptr;       // [batch][64]
  float* d_dz = nullptr;       // [batch]
  float* d_dh2 = nullptr;      // [batch][64]
  float* d_dh1 = nullptr;      // [batch][128]
  float* d_loss = nullptr;     // [1]
  float* d_part = nullptr;     // [kGradBlocks][total] gradient partial sums
};

struct sbwt_bloom {
  sbwt::u64 blocks = 0;  // 128-byte blocks (16 words each)
  sbwt::u32 hashes = 0;
  sbwt::u32 k = 32;
  sbwt::u64 keymask = ~0ull;  // the top 2k bits
  sbwt::u64* d_bits = nullptr;
};

namespace sbwt {

// ---- bf16 copies -----------------------------------------------------------

__global__ void k_mlp_cast(const float* __restrict__ p, u64 total, u32 in, u64 oW1, u64 oW2, bf16* __restrict__ wb,
                           bf16* __restrict__ W1T, bf16* __restrict__ W2T) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (i < total) wb[i] = to_bf16(p[i]);
  if (i < static_cast<u64>(kH1) * kMaxIn) {  // W1T[u][x]
    const u32 u = static_cast<u32>(i / kMaxIn), x = static_cast<u32>(i % kMaxIn);
    W1T[i] = x < in ? to_bf16(p[oW1 + static_cast<u64>(x) * kH1 + u]) : 0;
  }
  if (i < static_cast<u64>(kH2) * kH1) {  // W2T[v][u]
    const u32 v = static_cast<u32>(i / kH1), u = static_cast<u32>(i % kH1);
    W2T[i] = to_bf16(p[oW2 + static_cast<u64>(u) * kH2 + v]);
  }
}

// ---- inference on the matrix cores -------------------------------------------

struct MlpDev {
  const bf16* W1T;
  const bf16* W2T;
  const float* p;  // fp32 params (biases read from here; they stay fp32)
  u64 ob1, ob2, ow3, ob3;
  u32 k;
};

typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef u32 u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f4v mfma32(u32x4 a, u32x4 b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8v, a), __builtin_bit_cast(bf8v, b), c, 0, 0,
                                                  0);
}

// One wave's weight fragments, held in registers for the whole launch. The
// 16x16x32 bf16 MFMA takes from lane l = 16q + c row c of A and column c of
// B, with the 8 contraction slots 8q .. 8q+7. Layer 1 (K block kb = inputs
// 32kb .. 32kb+31): slot 8q + i is input 32kb + 8q + i, i.e. characters
// 8kb + 2q and 8kb + 2q + 1 of the k-mer. Layer 2 (K block j = hidden units
// 32j .. 32j+31): slots 8q .. 8q+3 are units 32j + 4q + i and slots
// 8q+4 .. 8q+7 are units 32j + 16 + 4q + i, which are exactly the rows the
// layer-1 accumulators of blocks 2j and 2j+1 hold in lane l, so layer 1's
// output feeds layer 2 without leaving registers.
struct MlpRegs {
  u32x4 a1[8][4];  // [hidden block nb][input block kb]
  u32x4 a2[4][4];  // [output block mb][hidden block j]
  f4v b1[8];       // b1 of the rows this lane's layer-1 accumulators hold
  f4v b2[4];       // b2 and bf16-rounded w3 of its layer-2 rows
  f4v w3[4];
  float b3;
};

__device__ __forceinline__ void mlp_load(const MlpDev& md, MlpRegs& w) {
  const u32 lane = threadIdx.x & 63, c = lane & 15, q = lane >> 4;
#pragma unroll
  for (u32 nb = 0; nb < 8; nb++)
#pragma unroll
    for (u32 i = 0; i < 4; i++) w.b1[nb][i] = md.p[md.ob1 + nb * 16 + q * 4 + i];
#pragma unroll
  for (u32 mb = 0; mb < 4; mb++)
#pragma unroll
    for (u32 i = 0; i < 4; i++) {
      w.b2[mb][i] = md.p[md.ob2 + mb * 16 + q * 4 + i];
      w.w3[mb][i] = from_bf16(to_bf16(md.p[md.ow3 + mb * 16 + q * 4 + i]));
    }
  w.b3 = md.p[md.ob3];
#pragma unroll
  for (u32 nb = 0; nb < 8; nb++)
#pragma unroll
    for (u32 kb = 0; kb < 4; kb++)
      w.a1[nb][kb] = *reinterpret_cast<const u32x4*>(md.W1T + (nb * 16 + c) * kMaxIn + kb * 32 + q * 8);
#pragma unroll
  for (u32 mb = 0; mb < 4; mb++)
#pragma unroll
    for (u32 j = 0; j < 4; j++) {
      const bf16* r = md.W2T + (mb * 16 + c) * kH1 + j * 32 + q * 4;
      const uint2 lo = *reinterpret_cast<const uint2*>(r), hi = *reinterpret_cast<const uint2*>(r + 16);
      w.a2[mb][j] = u32x4{lo.x, lo.y, hi.x, hi.y};
    }
  // The weights feed only the matrix cores, which read AGPRs directly: pin
  // them there so the accumulators, which the vector ALU reads, get VGPRs.
#pragma unroll
  for (u32 nb = 0; nb < 8; nb++)
#pragma unroll
    for (u32 kb = 0; kb < 4; kb++) asm volatile("" : "+a"(w.a1[nb][kb]));
#pragma unroll
  for (u32 mb = 0; mb < 4; mb++)
#pragma unroll
    for (u32 j = 0; j < 4; j++) asm volatile("" : "+a"(w.a2[mb][j]));
}

typedef __bf16 bf2v __attribute__((ext_vector_type(2)));
typedef short s2v __attribute__((ext_vector_type(2)));

// Round two values to nearest even bf16 (one v_cvt_pk_bf16_f32), then ReLU
// both: a bf16 (or fp32) bit pattern is negative as a signed integer exactly
// when the value is negative, so max(bits, 0) is ReLU (one v_pk_max_i16; and
// rounding then clamping equals clamping then rounding).
__device__ __forceinline__ u32 pack_relu_bf16(float lo, float hi) {
  typedef float f2v __attribute__((ext_vector_type(2)));
  const bf2v v = __builtin_convertvector(f2v{lo, hi}, bf2v);
  const s2v r = __builtin_elementwise_max(__builtin_bit_cast(s2v, v), s2v{0, 0});
  return __builtin_bit_cast(u32, r);
}

__device__ __forceinline__ float relu_bits(float x) {
  return __int_as_float(max(__float_as_int(x), 0));
}

// One-hot B fragments of layer 1 by the 4-bit code of two characters
// (c0 c1): slots 0..3 one-hot of c0, slots 4..7 one-hot of c1, bf16 1.0 at
// the code. 16 entries of 16 bytes in LDS span the 64 banks once, so a
// wave's lookups never conflict. Characters at or past k need no masking:
// their columns of W1^T are zero.
struct OneHotLds {
  u32x4 e[16];
};

__device__ __forceinline__ void onehot_init(OneHotLds& t) {
  if (threadIdx.x < 16) {
    const u32 c0 = threadIdx.x >> 2, c1 = threadIdx.x & 3;
    const u64 o0 = 0x3F80ull << (16 * c0), o1 = 0x3F80ull << (16 * c1);
    t.e[threadIdx.x] =
        u32x4{static_cast<u32>(o0), static_cast<u32>(o0 >> 32), static_cast<u32>(o1), static_cast<u32>(o1 >> 32)};
  }
}

// Input block kb of k-mer km for lane quad q: characters 8kb + 2q, 8kb + 2q + 1
// are the 4 bits at 60 - 16kb - 4q.
__device__ __forceinline__ u32x4 onehot_block(const OneHotLds& t, u64 km, u32 kb, u32 q) {
  const u32 word = kb < 2 ? static_cast<u32>(km >> 32) : static_cast<u32>(km);
  const u32 code = (word >> (28 - 16 * (kb & 1) - 4 * q)) & 15u;
  return t.e[code];
}

// Global index of this wave, as a scalar (the compiler cannot prove
// threadIdx.x >> 6 uniform by itself).
__device__ __forceinline__ u64 wave_index() {
  const u32 w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  return static_cast<u64>(blockIdx.x) * (blockDim.x >> 6) + w;
}

// Logit of this lane's own k-mer: the wave's 64 k-mers as four groups of 16
// (group g = lanes 16g .. 16g+15, one k-mer per column). Software-pipelined:
// the layer-1 MFMAs of group g+1 are issued between the ReLU/bf16 packing and
// layer-2 MFMAs of group g, so the matrix pipeline never drains between
// layers; the epilogue of group g (layer 3, a dot product, on the vector
// ALU) overlaps the next group's matrix work.
__device__ __forceinline__ float mlp64(const MlpRegs& w, const OneHotLds& oh, u64 km) {
  const u32 lane = threadIdx.x & 63, c = lane & 15, q = lane >> 4;
  u64 kg[4];
#pragma unroll
  for (u32 g = 0; g < 4; g++) kg[g] = __shfl(km, static_cast<int>(g * 16 + c));
  f4v a1[2][8];
#pragma unroll
  for (int nb = 0; nb < 8; nb++) a1[0][nb] = w.b1[nb];  // biases enter as the accumulators' start
#pragma unroll
  for (u32 kb = 0; kb < 4; kb++) {
    const u32x4 b = onehot_block(oh, kg[0], kb, q);
#pragma unroll
    for (u32 nb = 0; nb < 8; nb++) a1[0][nb] = mfma32(w.a1[nb][kb], b, a1[0][nb]);
  }
  float zp[4];
#pragma unroll
  for (u32 g = 0; g < 4; g++) {
    const u32 cur = g & 1, nxt = cur ^ 1;
    if (g + 1 < 4) {
#pragma unroll
      for (int nb = 0; nb < 8; nb++) a1[nxt][nb] = w.b1[nb];
    }
    f4v acc2[4];
#pragma unroll
    for (int mb = 0; mb < 4; mb++) acc2[mb] = w.b2[mb];
#pragma unroll
    for (u32 j = 0; j < 4; j++) {
      if (g + 1 < 4) {
        const u32x4 b = onehot_block(oh, kg[g + 1], j, q);
#pragma unroll
        for (u32 nb = 0; nb < 8; nb++) a1[nxt][nb] = mfma32(w.a1[nb][j], b, a1[nxt][nb]);
      }
      const f4v x0 = a1[cur][2 * j], x1 = a1[cur][2 * j + 1];
      const u32x4 h = u32x4{pack_relu_bf16(x0[0], x0[1]), pack_relu_bf16(x0[2], x0[3]),
                            pack_relu_bf16(x1[0], x1[1]), pack_relu_bf16(x1[2], x1[3])};
#pragma unroll
      for (u32 mb = 0; mb < 4; mb++) acc2[mb] = mfma32(w.a2[mb][j], h, acc2[mb]);
    }
    float z = 0.f;
#pragma unroll
    for (u32 mb = 0; mb < 4; mb++)
#pragma unroll
      for (u32 i = 0; i < 4; i++) z = fmaf(relu_bits(acc2[mb][i]), w.w3[mb][i], z);
    zp[g] = z;
  }
  float mine = 0.f;
#pragma unroll
  for (u32 g = 0; g < 4; g++) {
    float z = zp[g];
    z += __shfl_xor(z, 16);
    z += __shfl_xor(z, 32);
    mine = q == g ? z : mine;
  }
  return mine + w.b3;
}

__global__ __launch_bounds__(256) void k_mlp_logits(MlpDev md, const u64* __restrict__ kmers, u64 count,
                                                    float* __restrict__ out) {
  __shared__ OneHotLds oh;
  onehot_init(oh);
  MlpRegs w;
  mlp_load(md, w);
  __syncthreads();
  const u32 lane = threadIdx.x & 63;
  const u64 wave = wave_index();  // wave-uniform (scalar): scalar loop control
  const u64 waves = (gridDim.x * static_cast<u64>(blockDim.x)) >> 6;
  const u64 groups = (count + 63) / 64;
  auto load = [&](u64 g) -> u64 {  // unconditional, clamped: lanes past the end store nothing
    const u64 i = g * 64 + lane;
    return kmers[i < count ? i : count - 1];
  };
  // Full groups, unrolled by two with the k-mer slots swapping roles and no
  // exit between a load and its use: each load lands in a register whose
  // old value is dead and is in flight while the other slot computes;
  // stores are unconditional, so each wait is exact.
  const u64 full = count / 64;
  u64 g = wave;
  u64 ka = load(g), kb = 0;
  for (; g + waves < full; g += 2 * waves) {
    kb = load(g + waves);
    out[g * 64 + lane] = mlp64(w, oh, ka);
    ka = load(g + 2 * waves);
    out[(g + waves) * 64 + lane] = mlp64(w, oh, kb);
  }
  if (g < full) out[g * 64 + lane] = mlp64(w, oh, ka);
  if (full < groups && wave == full % waves) {  // the last, partial group
    const float z = mlp64(w, oh, load(full));
    if (full * 64 + lane < count) out[full * 64 + lane] = z;
  }
}

// The same logits with FP32 vector arithmetic, one thread per k-mer (the
// cross-check of the matrix-core kernel; identical bf16 rounding of h1).
__global__ void k_mlp_logits_ref(const float* __restrict__ p, const bf16* __restrict__ wb, MlpShape sh,
                                 const u64* __restrict__ kmers, u64 count, float* __restrict__ out) {
  const u64 q = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (q >= count) return;
  const u64 km = kmers[q];
  float h1[kH1];
  for (u32 u = 0; u < kH1; u++) h1[u] = 0.f;
  for (u32 t = 0; t < sh.k; t++) {
    const u32 x = 4 * t + (static_cast<u32>(km >> (62 - 2 * t)) & 3u);
    for (u32 u = 0; u < kH1; u++) h1[u] += from_bf16(wb[sh.oW1 + static_cast<u64>(x) * kH1 + u]);
  }
  for (u32 u = 0; u < kH1; u++) {
    const float v = h1[u] + p[sh.ob1 + u];
    h1[u] = from_bf16(to_bf16(v > 0.f ? v : 0.f));
  }
  float z = p[sh.ob3];
  for (u32 v = 0; v < kH2; v++) {
    float a = 0.f;
    for (u32 u = 0; u < kH1; u++) a += h1[u] * from_bf16(wb[sh.oW2 + static_cast<u64>(u) * kH2 + v]);
    a += p[sh.ob2 + v];
    z += (a > 0.f ? a : 0.f) * from_bf16(wb[sh.ow3 + v]);
  }
  out[q] = z;
}

// ---- training: forward / backward over a mini-batch (FP32 vector) ------------

__global__ void k_tr_h1(const float* __restrict__ p, const bf16* __restrict__ wb, MlpShape sh,
                        const u64* __restrict__ kmers, u64 B, float* __restrict__ h1) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (i >= B * kH1) return;
  const u64 s = i / kH1;
  const u32 u = static_cast<u32>(i % kH1);
  const u64 km = kmers[s];
  float a = p[sh.ob1 + u];
  for (u32 t = 0; t < sh.k; t++) {
    const u32 x = 4 * t + (static_cast<u32>(km >> (62 - 2 * t)) & 3u);
    a += from_bf16(wb[sh.oW1 + static_cast<u64>(x) * kH1 + u]);
  }
  h1[i] = a > 0.f ? a : 0.f;
}

__global__ void k_tr_h2(const float* __restrict__ p, const bf16* __restrict__ wb, MlpShape sh,
                        const float* __restrict__ h1, u64 B, float* __restrict__ h2) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (i >= B * kH2) return;
  const u64 s = i / kH2;
  const u32 v = static_cast<u32>(i % kH2);
  float a = p[sh.ob2 + v];
  for (u32 u = 0; u < kH1; u++) a += h1[s * kH1 + u] * from_bf16(wb[sh.oW2 + static_cast<u64>(u) * kH2 + v]);
  h2[i] = a > 0.f ? a : 0.f;
}

// z, mean BCE loss, dz = (sigmoid(z) - y) / B
__global__ void k_tr_out(const float* __restrict__ p, const bf16* __restrict__ wb, MlpShape sh,
                         const float* __restrict__ h2, const float* __restrict__ y, u64 B, float* __restrict__ dz,
                         float* __restrict__ loss) {
  const u64 s = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  float l = 0.f;
  if (s < B) {
    float z = p[sh.ob3];
    for (u32 v = 0; v < kH2; v++) z += h2[s * kH2 + v] * from_bf16(wb[sh.ow3 + v]);
    const float pr = 1.f / (1.f + expf(-z));
    dz[s] = (pr - y[s]) / static_cast<float>(B);
    // log(1 + e^z) - y z, written stably
    l = (z > 0.f ? z + log1pf(expf(-z)) : log1pf(expf(z))) - y[s] * z;
  }
  for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off);
  if ((threadIdx.x & 63) == 0) atomicAdd(loss, l / static_cast<float>(B));
}

__global__ void k_tr_dh2(const bf16* __restrict__ wb, MlpShape sh, const float* __restrict__ h2,
                         const float* __restrict__ dz, u64 B, float* __restrict__ dh2) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (i >= B * kH2) return;
  const u64 s = i / kH2;
  const u32 v = static_cast<u32>(i % kH2);
  dh2[i] = h2[i] > 0.f ? dz[s] * from_bf16(wb[sh.ow3 + v]) : 0.f;
}

__global__ void k_tr_dh1(const bf16* __restrict__ wb, MlpShape sh, const float* __restrict__ h1,
                         const float* __restrict__ dh2, u64 B, float* __restrict__ dh1) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (i >= B * kH1) return;
  const u64 s = i / kH1;
  const u32 u = static_cast<u32>(i % kH1);
  if (h1[i] <= 0.f) {
    dh1[i] = 0.f;
    return;
  }
  float a = 0.f;
  for (u32 v = 0; v < kH2; v++) a += dh2[s * kH2 + v] * from_bf16(wb[sh.oW2 + static_cast<u64>(u) * kH2 + v]);
  dh1[i] = a;
}

// Gradient partial sums: each block takes a contiguous range of samples and
// accumulates every parameter's gradient over it in registers, thread
// (u = tid % 128, half = tid / 128) owning W2[u][32 half .. 32 half + 31],
// W1[4t + c][u] for t = half, half + 2, ..., and (threads < 65) b2, w3, b3.
// Samples are staged through LDS in chunks; k_tr_grad_reduce then sums the
// per-block partials in a fixed order (deterministic gradients).
constexpr u32 kGradChunk = 32;

__global__ __launch_bounds__(256) void k_tr_grad_partial(MlpShape sh, const u64* __restrict__ kmers,
                                                         const float* __restrict__ h1, const float* __restrict__ h2,
                                                         const float* __restrict__ dz, const float* __restrict__ dh2,
                                                         const float* __restrict__ dh1, u64 B, u64 per_block,
                                                         float* __restrict__ part) {
  __shared__ float s_dh2[kGradChunk][kH2];
  __shared__ float s_h2[kGradChunk][kH2];
  __shared__ float s_dz[kGradChunk];
  __shared__ u64 s_km[kGradChunk];
  const u32 tid = threadIdx.x, u = tid & (kH1 - 1), half = tid >> 7;
  float gw1[16][4], gw2[32];
#pragma unroll
  for (int j = 0; j < 16; j++) gw1[j][0] = gw1[j][1] = gw1[j][2] = gw1[j][3] = 0.f;
#pragma unroll
  for (int j = 0; j < 32; j++) gw2[j] = 0.f;
  float gb1 = 0.f, gb2 = 0.f, gw3 = 0.f, gb3 = 0.f;
  const u64 s0 = blockIdx.x * per_block;
  const u64 s1 = s0 + per_block < B ? s0 + per_block : B;
  for (u64 c0 = s0; c0 < s1; c0 += kGradChunk) {
    const u32 n = static_cast<u32>(s1 - c0 < kGradChunk ? s1 - c0 : kGradChunk);
    __syncthreads();
    for (u32 i = tid; i < n * kH2; i += blockDim.x) {
      s_dh2[i / kH2][i % kH2] = dh2[c0 * kH2 + i];
      s_h2[i / kH2][i % kH2] = h2[c0 * kH2 + i];
    }
    if (tid < n) {
      s_dz[tid] = dz[c0 + tid];
      s_km[tid] = kmers[c0 + tid];
    }
    __syncthreads();
    for (u32 j = 0; j < n; j++) {
      const u64 s = c0 + j;
      const float a = h1[s * kH1 + u], d = dh1[s * kH1 + u];
#pragma unroll
      for (int jj = 0; jj < 32; jj++) gw2[jj] += a * s_dh2[j][half * 32 + jj];
      gb1 += d;
      const u64 km = s_km[j];
#pragma unroll
      for (int jj = 0; jj < 16; jj++) {
        const u32 t = 2 * jj + half;
        const u32 c = static_cast<u32>(km >> (62 - 2 * t)) & 3u;
        gw1[jj][0] += c == 0 ? d : 0.f;
        gw1[jj][1] += c == 1 ? d : 0.f;
        gw1[jj][2] += c == 2 ? d : 0.f;
        gw1[jj][3] += c == 3 ? d : 0.f;
      }
      if (tid < kH2) {
        gb2 += s_dh2[j][tid];
        gw3 += s_dz[j] * s_h2[j][tid];
      } else if (tid == kH2) {
        gb3 += s_dz[j];
      }
    }
  }
  float* out = part + blockIdx.x * sh.total;
#pragma unroll
  for (int jj = 0; jj < 32; jj++) out[sh.oW2 + static_cast<u64>(u) * kH2 + half * 32 + jj] = gw2[jj];
#pragma unroll
  for (int jj = 0; jj < 16; jj++) {
    const u32 t = 2 * jj + half;
    if (t < sh.k) {
#pragma unroll
      for (int c = 0; c < 4; c++) out[sh.oW1 + static_cast<u64>(4 * t + c) * kH1 + u] = gw1[jj][c];
    }
  }
  if (half == 0) out[sh.ob1 + u] = gb1;
  if (tid < kH2) {
    out[sh.ob2 + tid] = gb2;
    out[sh.ow3 + tid] = gw3;
  } else if (tid == kH2) {
    out[sh.ob3] = gb3;
  }
}

__global__ void k_tr_grad_reduce(const float* __restrict__ part, u32 blocks, u64 total, float* __restrict__ g) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (i >= total) return;
  float x = 0.f;
  for (u32 b = 0; b < blocks; b++) x += part[b * total + i];
  g[i] = x;
}

__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ g, u64 n, float lr, float b1c, float b2c) {
  const u64 i = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = 0.9f * m[i] + 0.1f * gi;
  const float vi = 0.999f * v[i] + 0.001f * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / b1c) / (sqrtf(vi / b2c) + 1e-8f);
}

static int mlp_cast(sbwt_mlp* nn, hipStream_t s) {
  const u64 n = nn->shape.total > kH1 * kMaxIn ? nn->shape.total : kH1 * kMaxIn;
  k_mlp_cast<<<(n + 255) / 256, 256, 0, s>>>(nn->d_params, nn->shape.total, nn->shape.in, nn->shape.oW1,
                                            nn->shape.oW2, nn->d_wb, nn->d_W1T, nn->d_W2T);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

static int mlp_grad(sbwt_mlp* nn, const u64* kmers, const float* y, u64 B, hipStream_t s) {
  const MlpShape& sh = nn->shape;
  const u32 bt = 256;
  SBWT_HIP(hipMemsetAsync(nn->d_loss, 0, sizeof(float), s));
  k_tr_h1<<<(B * kH1 + bt - 1) / bt, bt, 0, s>>>(nn->d_params, nn->d_wb, sh, kmers, B, nn->d_h1);
  k_tr_h2<<<(B * kH2 + bt - 1) / bt, bt, 0, s>>>(nn->d_params, nn->d_wb, sh, nn->d_h1, B, nn->d_h2);
  k_tr_out<<<(B + bt - 1) / bt, bt, 0, s>>>(nn->d_params, nn->d_wb, sh, nn->d_h2, y, B, nn->d_dz, nn->d_loss);
  k_tr_dh2<<<(B * kH2 + bt - 1) / bt, bt, 0, s>>>(nn->d_wb, sh, nn->d_h2, nn->d_dz, B, nn->d_dh2);
  k_tr_dh1<<<(B * kH1 + bt - 1) / bt, bt, 0, s>>>(nn->d_wb, sh, nn->d_h1, nn->d_dh2, B, nn->d_dh1);
  const u64 per_block = (B + kGradBlocks - 1) / kGradBlocks;
  const u32 blocks = static_cast<u32>((B + per_block - 1) / per_block);
  k_tr_grad_partial<<<blocks, bt, 0, s>>>(sh, kmers, nn->d_h1, nn->d_h2, nn->d_dz, nn->d_dh2, nn->d_dh1, B,
                                           per_block, nn->d_part);
  k_tr_grad_reduce<<<(sh.total + bt - 1) / bt, bt, 0, s>>>(nn->d_part, blocks, sh.total, nn->d_grad);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

// ---- Bloom filter: blocks chosen by the key's minimizer ---------------------
// A key's bits all lie in one 16-byte pair of 64-bit words (a register-blocked
// filter, Putze, Sanders, Singler 2007, with 128-bit blocks) inside one
// 128-byte block. The block is chosen by the key's minimizer, so the
// consecutive k-mers of a read, which mostly share it, share one cache line;
// the pair and the bits are chosen by the key itself:
//   minimizer = the m-mer (m = min(k, 16)) of the key whose code c has the
//               smallest o(c) = x ^ (x >> 16), x = c * 0x9E3779 mod 2^32
//               (the leftmost one on ties)
//   block     = fastrange(mix64(minimizer ^ kBloomBlockSeed), blocks)
//   h         = mix64(key ^ golden); pair = h >> 61 (8 pairs per block)
//   bit i     = bits 7i..7i+6 of mix64(h), over the pair's 128 bits (i < hashes <= 9)
// A lookup is one 16-byte load and two mask compares.
constexpr u32 kBloomBlockWords = 16;  // 128 bytes
constexpr u64 kBloomBlockSeed = 0x632BE59BD9B4E019ull;

__device__ __host__ __forceinline__ u64 mix64(u64 x) {
  x ^= x >> 30;
  x *= 0xbf58476d1ce4e5b9ull;
  x ^= x >> 27;
  x *= 0x94d049bb133111ebull;
  x ^= x >> 31;
  return x;
}

// The minimizer of a top-aligned key of k characters (its 2m-bit code). With
// m = 16 a window is one 32-bit word, so the next window's x is 4x plus the
// new character times the multiplier (mod 2^32, the leaving character shifts
// out): one full-rate 24-bit multiply per window. For k < 16 the one window
// is the k-mer itself.
constexpr u32 kMinimizerMul = 0x9E3779u;  // odd, 24 bits

__device__ __forceinline__ u32 kmer_minimizer(u64 key, u32 k) {
  if (k < 16) return static_cast<u32>(key >> (64 - 2 * k));
  const u32 lo = static_cast<u32>(key);
  u32 x = static_cast<u32>(key >> 32) * kMinimizerMul;
  u32 bo = x ^ (x >> 16), bi = 0;
#pragma unroll
  for (u32 i = 1; i <= 16; i++) {
    if (i + 16 > k) break;  // k is the same in every lane
    x = (x << 2) + ((lo >> (32 - 2 * i)) & 3u) * kMinimizerMul;
    const u32 o = x ^ (x >> 16);
    bi = o < bo ? i : bi;
    bo = o < bo ? o : bo;
  }
  return static_cast<u32>(key >> (32 - 2 * bi));
}

// The index of the key's pair (its first word) and the pair's two masks.
__device__ __forceinline__ u64 bloom_locate(u64 blocks, u32 hashes, u32 k, u64 key, u64& mlo, u64& mhi) {
  const u64 hb = mix64(static_cast<u64>(kmer_minimizer(key, k)) ^ kBloomBlockSeed);
  const u64 blk = static_cast<u64>((static_cast<unsigned __int128>(hb) * blocks) >> 64);
  const u64 h = mix64(key ^ 0x9E3779B97F4A7C15ull);
  u64 h2 = mix64(h);
  mlo = 0;
  mhi = 0;
  for (u32 i = 0; i < hashes; i++, h2 >>= 7) {
    const u32 b = static_cast<u32>(h2) & 127u;
    const u64 bit = 1ull << (b & 63u);
    mlo |= b < 64 ? bit : 0ull;
    mhi |= b < 64 ? 0ull : bit;
  }
  return blk * kBloomBlockWords + (h >> 61) * 2;
}

__global__ void k_bloom_insert(u64* __restrict__ bits, u64 blocks, u32 hashes, u32 k, u64 keymask,
                               const u64* __restrict__ keys, u64 count) {
  const u64 q = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (q >= count) return;
  u64 mlo, mhi;
  const u64 w = bloom_locate(blocks, hashes, k, keys[q] & keymask, mlo, mhi);
  if (mlo) atomicOr(reinterpret_cast<unsigned long long*>(bits + w), mlo);
  if (mhi) atomicOr(reinterpret_cast<unsigned long long*>(bits + w + 1), mhi);
}

// the pair's 16 bytes hold both masks
__device__ __forceinline__ bool bloom_match(uint4 v, u64 mlo, u64 mhi) {
  const u64 lo = static_cast<u64>(v.x) | static_cast<u64>(v.y) << 32;
  const u64 hi = static_cast<u64>(v.z) | static_cast<u64>(v.w) << 32;
  return (lo & mlo) == mlo && (hi & mhi) == mhi;
}

__device__ __forceinline__ bool bloom_has(const u64* __restrict__ bits, u64 blocks, u32 hashes, u32 k, u64 key) {
  u64 mlo, mhi;
  const u64 w = bloom_locate(blocks, hashes, k, key, mlo, mhi);
  return bloom_match(*reinterpret_cast<const uint4*>(bits + w), mlo, mhi);
}

__global__ void k_bloom_query(const u64* __restrict__ bits, u64 blocks, u32 hashes, u32 k, u64 keymask,
                              const u64* __restrict__ keys, u64 count, u8* __restrict__ out) {
  const u64 q = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (q >= count) return;
  out[q] = bits && blocks ? (bloom_has(bits, blocks, hashes, k, keys[q] & keymask) ? 1 : 0) : 0;
}

// out[offsets[r] + j] = the filter's answer for the k-mer at character j of
// row r, read straight from the packed rows: one wave per row (grid-stride),
// its lanes on consecutive k-mers, which mostly share a block.
__global__ __launch_bounds__(64) void k_bloom_query_reads(const u64* __restrict__ bits, u64 blocks, u32 hashes,
                                                          u32 k, u64 keymask, const u64* __restrict__ packed,
                                                          const u64* __restrict__ row_starts,
                                                          const u64* __restrict__ offsets, u64 rows,
                                                          u8* __restrict__ out) {
  for (u64 r = blockIdx.x; r < rows; r += gridDim.x) {
    const u64 o0 = offsets[r], nk = offsets[r + 1] - o0, rs = row_starts[r];
    for (u64 j = threadIdx.x; j < nk; j += 64) {
      const u64 b = (rs + j) << 1;  // the window's first bit in the packed string
      const u64 win = (packed[b >> 6] << (b & 63)) | ((packed[(b >> 6) + 1] >> (63 - (b & 63))) >> 1);
      out[o0 + j] = bloom_has(bits, blocks, hashes, k, win & keymask) ? 1 : 0;
    }
  }
}

// maybe[q] = logit(k-mer q) >= threshold. Each wave takes 64 k-mers per step;
// the k-mers two steps ahead are loaded while the matrix cores run. Every
// load is issued by every lane in straight-line code (k-mer indexes are
// clamped), so the wait before each use covers exactly the loads it needs and
// never the ones still in flight.
__global__ __launch_bounds__(256) void k_prefilter(MlpDev md, float threshold, const u64* __restrict__ kmers,
                                                   u64 count, u8* __restrict__ maybe) {
  __shared__ OneHotLds oh;
  onehot_init(oh);
  MlpRegs w;
  mlp_load(md, w);
  __syncthreads();
  const u32 lane = threadIdx.x & 63;
  const u64 wave = wave_index();  // wave-uniform (scalar): scalar loop control
  const u64 waves = (gridDim.x * static_cast<u64>(blockDim.x)) >> 6;
  const u64 groups = (count + 63) / 64;
  // lanes past the end take the last k-mer (their results are not stored);
  // no select on the loaded value, which would wait for it at once
  auto load = [&](u64 g) -> u64 {
    const u64 i = g * 64 + lane;
    return kmers[i < count ? i : count - 1];
  };
  // Full groups (all 64 lanes valid): one step stores group g's decisions
  // (logits zc) while the matrix cores take group g + waves (kn), and kc's
  // register receives group g + 2 waves. Unrolled by two with the slots
  // swapping roles and no exit between a load and its use (so no register
  // copy of a load in flight is needed and the compiler cannot sink a load
  // to its use); loads, matrix work and stores are unconditional (loads
  // clamp, surplus logits are dropped), so each wait covers exactly what it
  // needs.
  const u64 full = count / 64;
  auto step = [&](u64& kc, float& zc, u64 kn, float& zn, u64 g) {
    const float zcur = zc;
    kc = load(g + 2 * waves);
    zn = mlp64(w, oh, kn);
    maybe[g * 64 + lane] = zcur >= threshold ? 1 : 0;
  };
  u64 g = wave;
  u64 ka = load(g), kb = load(g + waves);
  float za = mlp64(w, oh, ka), zb = 0.f;
  for (; g + waves < full; g += 2 * waves) {
    step(ka, za, kb, zb, g);
    step(kb, zb, ka, za, g + waves);
  }
  if (g < full) maybe[g * 64 + lane] = za >= threshold ? 1 : 0;  // one full group left: its logits are in za
  // the last, partial group
  if (full < groups && wave == full % waves) {
    const u64 idx = full * 64 + lane;
    const u64 kc = load(full);
    const float zc = mlp64(w, oh, kc);
    if (idx < count) maybe[idx] = zc >= threshold ? 1 : 0;
  }
}

// maybe[q] |= bloom(k-mer q) for the k-mers the model rejected. A pass of its
// own at full occupancy: inside the matrix-core kernel (one wave per SIMD)
// the minimizer of every k-mer cost more than its 16-byte probe. k-mers in
// sequence order share blocks, so their probes share lines.
__global__ void k_bloom_or(const u64* __restrict__ bits, u64 blocks, u32 hashes, u32 k, u64 keymask,
                           const u64* __restrict__ kmers, u64 count, u8* __restrict__ maybe) {
  const u64 q = blockIdx.x * static_cast<u64>(blockDim.x) + threadIdx.x;
  if (q >= count || maybe[q]) return;
  maybe[q] = bloom_has(bits, blocks, hashes, k, kmers[q] & keymask) ? 1 : 0;
}

// False-positive rate of a register-blocked filter (128-bit blocks) at
// bits_per_key with h hashes: the load of a block is Poisson(128 /
// bits_per_key) keys, and a non-key is accepted when all h of its bits are set.
static double bloom_fpr(double bits_per_key, u32 h) {
  const double lam = 128.0 / bits_per_key;
  const double per = std::log1p(-1.0 / 128.0) * h;  // ln P(a given bit unset by one key)
  double pmf = std::exp(-lam), s = 0.0;
  const int nmax = static_cast<int>(lam + 12.0 * std::sqrt(lam) + 24.0);
  for (int n = 0; n <= nmax; n++) {
    s += pmf * std::pow(1.0 - std::exp(per * n), h);
    pmf *= lam / (n + 1);
  }
  return s;
}

// Keys that share a minimizer land in one block together (about 8 consecutive
// k-mers of a sequence each), which the Poisson model does not see; genome
// k-mers (100 Mbp-scale synthetic genomes, k = 31) need about this many more
// bits per key for the model's rate.
constexpr double kBloomClumpMargin = 1.2;

// The smallest bits per key (on a grid of 1/20 of the unblocked filter's
// -ln p / ln^2 2) and the hash count (1..9) at which the model reaches
// false-positive rate p; bits_per_key includes the clump margin.
static void bloom_size(double p, double& bits_per_key, u32& hashes) {
  const double base = -std::log(p) / (M_LN2 * M_LN2);
  bits_per_key = 0.0;
  hashes = 1;
  for (u32 h = 1; h <= 9; h++)
    for (int j = 0; j <= 200; j++) {
      const double bpk = base * (1.0 + 0.05 * j);
      if (bloom_fpr(bpk, h) <= p) {
        if (bits_per_key == 0.0 || bpk < bits_per_key) {
          bits_per_key = bpk;
          hashes = h;
        }
        break;
      }
    }
  if (bits_per_key == 0.0) {  // p beyond reach of 9 hashes in 128-bit blocks
    bits_per_key = base * 11.0;
    hashes = 9;
  }
  bits_per_key *= kBloomClumpMargin;
}

static MlpDev mlp_dev(const sbwt_mlp* nn) {
  MlpDev md{};
  md.W1T = nn->d_W1T;
  md.W2T = nn->d_W2T;
  md.p = nn->d_params;
  md.ob1 = nn->shape.ob1;
  md.ob2 = nn->shape.ob2;
  md.ow3 = nn->shape.ow3;
  md.ob3 = nn->shape.ob3;
  md.k = nn->shape.k;
  return md;
}

// Blocks of 256 threads (4 waves, one group of 64 k-mers each per step) for
// a grid-stride kernel: no more than fit on the device at once.
template <class K>
static u64 resident_grid(K kernel, u64 groups) {
  int dev = 0, cus = 256, per = 1;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, 256, 0) != hipSuccess || per < 1) per = 1;
  const u64 want = (groups + 3) / 4;
  const u64 cap = static_cast<u64>(cus) * static_cast<u64>(per);
  return want < cap ? (want ? want : 1) : cap;
}

}  // namespace sbwt

using namespace sbwt;

extern "C" {

int sbwt_mlp_create(uint32_t k, uint64_t seed, uint64_t max_batch, sbwt_mlp** out) {
  clear_error();
  SBWT_REQUIRE(out, "out must not be NULL");
  SBWT_REQUIRE(k >= 1 && k <= 32, "k = %u: must be in [1, 32]", k);
  SBWT_REQUIRE(max_batch > 0, "max_batch must be positive");
  *out = nullptr;
  auto* nn = new sbwt_mlp();
  nn->shape = MlpShape(k);
  nn->max_batch = max_batch;
  const u64 P = nn->shape.total;
  bool ok = hipMalloc(&nn->d_params, P * 4) == hipSuccess && hipMalloc(&nn->d_m, P * 4) == hipSuccess &&
            hipMalloc(&nn->d_v, P * 4) == hipSuccess && hipMalloc(&nn->d_grad, P * 4) == hipSuccess &&
            hipMalloc(&nn->d_W1T, kH1 * kMaxIn * 2) == hipSuccess && hipMalloc(&nn->d_W2T, kH2 * kH1 * 2) == hipSuccess &&
            hipMalloc(&nn->d_wb, P * 2) == hipSuccess && hipMalloc(&nn->d_h1, max_batch * kH1 * 4) == hipSuccess &&
            hipMalloc(&nn->d_h2, max_batch * kH2 * 4) == hipSuccess && hipMalloc(&nn->d_dz, max_batch * 4) == hipSuccess &&
            hipMalloc(&nn->d_dh2, max_batch * kH2 * 4) == hipSuccess &&
            hipMalloc(&nn->d_dh1, max_batch * kH1 * 4) == hipSuccess && hipMalloc(&nn->d_loss, 4) == hipSuccess &&
            hipMalloc(&nn->d_part, static_cast<u64>(kGradBlocks) * P * 4) == hipSuccess;
  if (!ok) {
    sbwt_mlp_free(nn);
    return fail(SBWT_ERR_HIP, "out of device memory for the MLP (batch %llu)", (unsigned long long)max_batch);
  }
  // He-style initialisation from a fixed seed (host side, splitmix64 + Box-Muller)
  std::vector<float> h(P, 0.f);
  u64 st = seed * 0x9E3779B97F4A7C15ull + 1;
  auto next = [&]() {
    st += 0x9E3779B97F4A7C15ull;
    u64 z = st;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
  };
  auto gauss = [&]() {
    const double u1 = ((next() >> 11) + 1.0) / 9007199254740993.0, u2 = (next() >> 11) / 9007199254740992.0;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  };
  const MlpShape& sh = nn->shape;
  for (u64 i = 0; i < static_cast<u64>(sh.in) * kH1; i++) h[sh.oW1 + i] = static_cast<float>(gauss() * std::sqrt(2.0 / sh.k));
  for (u64 i = 0; i < static_cast<u64>(kH1) * kH2; i++) h[sh.oW2 + i] = static_cast<float>(gauss() * std::sqrt(2.0 / kH1));
  for (u64 i = 0; i < kH2; i++) h[sh.ow3 + i] = static_cast<float>(gauss() * std::sqrt(1.0 / kH2));
  int st2 = sbwt_mlp_set_params(nn, h.data(), P);
  if (st2 != SBWT_OK) {
    sbwt_mlp_free(nn);
    return st2;
  }
  *out = nn;
  return SBWT_OK;
}

int sbwt_mlp_free(sbwt_mlp* nn) {
  if (!nn) return SBWT_OK;
  for (void* q : {static_cast<void*>(nn->d_params), static_cast<void*>(nn->d_m), static_cast<void*>(nn->d_v),
                  static_cast<void*>(nn->d_grad), static_cast<void*>(nn->d_W1T), static_cast<void*>(nn->d_W2T),
                  static_cast<void*>(nn->d_wb), static_cast<void*>(nn->d_h1), static_cast<void*>(nn->d_h2),
                  static_cast<void*>(nn->d_dz), static_cast<void*>(nn->d_dh2), static_cast<void*>(nn->d_dh1),
                  static_cast<void*>(nn->d_loss), static_cast<void*>(nn->d_part)})
    if (q) (void)hipFree(q);
  delete nn;
  return SBWT_OK;
}

uint64_t sbwt_mlp_param_count(const sbwt_mlp* nn) { return nn ? nn->shape.total : 0; }

int sbwt_mlp_set_params(sbwt_mlp* nn, const float* params, uint64_t count) {
  clear_error();
  SBWT_REQUIRE(nn && params && count == nn->shape.total, "expected %llu parameters",
               (unsigned long long)(nn ? nn->shape.total : 0));
  SBWT_HIP(hipMemcpy(nn->d_params, params, count * 4, hipMemcpyHostToDevice));
  SBWT_HIP(hipMemset(nn->d_m, 0, count * 4));
  SBWT_HIP(hipMemset(nn->d_v, 0, count * 4));
  nn->step = 0;
  SBWT_TRY(mlp_cast(nn, nullptr));
  SBWT_HIP(hipDeviceSynchronize());
  return SBWT_OK;
}

int sbwt_mlp_get_params(const sbwt_mlp* nn, float* params, uint64_t count) {
  clear_error();
  SBWT_REQUIRE(nn && params && count == nn->shape.total, "expected %llu parameters",
               (unsigned long long)(nn ? nn->shape.total : 0));
  SBWT_HIP(hipMemcpy(params, nn->d_params, count * 4, hipMemcpyDeviceToHost));
  return SBWT_OK;
}

int sbwt_mlp_grad(sbwt_mlp* nn, const uint64_t* d_kmers, const float* d_labels, uint64_t batch, float* d_grad,
                  float* d_loss, void* stream) {
  clear_error();
  SBWT_REQUIRE(nn && d_kmers && d_labels, "arguments must not be NULL");
  SBWT_REQUIRE(batch > 0 && batch <= nn->max_batch, "batch %llu outside (0, %llu]", (unsigned long long)batch,
               (unsigned long long)nn->max_batch);
  hipStream_t s = as_stream(stream);
  SBWT_TRY(mlp_grad(nn, d_kmers, d_labels, batch, s));
  if (d_grad) SBWT_HIP(hipMemcpyAsync(d_grad, nn->d_grad, nn->shape.total * 4, hipMemcpyDeviceToDevice, s));
  if (d_loss) SBWT_HIP(hipMemcpyAsync(d_loss, nn->d_loss, 4, hipMemcpyDeviceToDevice, s));
  return SBWT_OK;
}

int sbwt_mlp_train_step(sbwt_mlp* nn, const uint64_t* d_kmers, const float* d_labels, uint64_t batch, float lr,
                        float* d_loss, void* stream) {
  clear_error();
  SBWT_REQUIRE(nn && d_kmers && d_labels, "arguments must not be NULL");
  SBWT_REQUIRE(batch > 0 && batch <= nn->max_batch, "batch %llu outside (0, %llu]", (unsigned long long)batch,
               (unsigned long long)nn->max_batch);
  hipStream_t s = as_stream(stream);
  SBWT_TRY(mlp_grad(nn, d_kmers, d_labels, batch, s));
  nn->step++;
  const float b1c = 1.f - std::pow(0.9f, static_cast<float>(nn->step));
  const float b2c = 1.f - std::pow(0.999f, static_cast<float>(nn->step));
  const u64 P = nn->shape.total;
  k_adam<<<(P + 255) / 256, 256, 0, s>>>(nn->d_params, nn->d_m, nn->d_v, nn->d_grad, P, lr, b1c, b2c);
  SBWT_HIP(hipGetLastError());
  SBWT_TRY(mlp_cast(nn, s));
  if (d_loss) SBWT_HIP(hipMemcpyAsync(d_loss, nn->d_loss, 4, hipMemcpyDeviceToDevice, s));
  return SBWT_OK;
}

int sbwt_mlp_logits(const sbwt_mlp* nn, const uint64_t* d_kmers, uint64_t count, float* d_logits, int reference,
                    void* stream) {
  clear_error();
  SBWT_REQUIRE(nn, "model is NULL");
  if (count == 0) return SBWT_OK;
  SBWT_REQUIRE(d_kmers && d_logits, "device buffers must not be NULL");
  hipStream_t s = as_stream(stream);
  if (reference) {
    k_mlp_logits_ref<<<(count + 127) / 128, 128, 0, s>>>(nn->d_params, nn->d_wb, nn->shape, d_kmers, count, d_logits);
  } else {
    const u64 groups = (count + 63) / 64;
    k_mlp_logits<<<resident_grid(k_mlp_logits, groups), 256, 0, s>>>(mlp_dev(nn), d_kmers, count, d_logits);
  }
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

int sbwt_bloom_create(uint32_t k, uint64_t expected, double fpr, sbwt_bloom** out) {
  clear_error();
  SBWT_REQUIRE(out && fpr > 0 && fpr < 1, "out must not be NULL and 0 < fpr < 1");
  SBWT_REQUIRE(k >= 1 && k <= 32, "k = %u: must be in [1, 32]", k);
  *out = nullptr;
  auto* b = new sbwt_bloom();
  b->k = k;
  b->keymask = k == 32 ? ~0ull : ~((1ull << (64 - 2 * k)) - 1);
  double bits_per_key = 0.0;
  bloom_size(fpr, bits_per_key, b->hashes);
  const double block_bits = 64.0 * kBloomBlockWords;
  b->blocks = expected ? static_cast<u64>(std::ceil(static_cast<double>(expected) * bits_per_key / block_bits)) : 0;
  const u64 bytes = b->blocks * kBloomBlockWords * 8;
  if (b->blocks) {
    if (hipMalloc(&b->d_bits, bytes) != hipSuccess || hipMemset(b->d_bits, 0, bytes) != hipSuccess) {
      delete b;
      return fail(SBWT_ERR_HIP, "out of device memory for a Bloom filter of %llu blocks", (unsigned long long)b->blocks);
    }
  }
  *out = b;
  return SBWT_OK;
}

int sbwt_bloom_free(sbwt_bloom* b) {
  if (!b) return SBWT_OK;
  if (b->d_bits) (void)hipFree(b->d_bits);
  delete b;
  return SBWT_OK;
}

uint64_t sbwt_bloom_bytes(const sbwt_bloom* b) { return b ? b->blocks * kBloomBlockWords * 8 : 0; }

uint32_t sbwt_bloom_hashes(const sbwt_bloom* b) { return b ? b->hashes : 0; }

int sbwt_bloom_get_words(const sbwt_bloom* b, uint64_t* words, uint64_t count) {
  clear_error();
  SBWT_REQUIRE(b && words && count == b->blocks * kBloomBlockWords, "expected %llu words",
               (unsigned long long)(b ? b->blocks * kBloomBlockWords : 0));
  if (count) SBWT_HIP(hipMemcpy(words, b->d_bits, count * 8, hipMemcpyDeviceToHost));
  return SBWT_OK;
}

int sbwt_bloom_insert(sbwt_bloom* b, const uint64_t* d_keys, uint64_t count, void* stream) {
  clear_error();
  SBWT_REQUIRE(b, "filter is NULL");
  if (count == 0) return SBWT_OK;
  SBWT_REQUIRE(b->blocks, "the filter was created for 0 items");
  k_bloom_insert<<<(count + 255) / 256, 256, 0, as_stream(stream)>>>(b->d_bits, b->blocks, b->hashes, b->k,
                                                                    b->keymask, d_keys, count);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

int sbwt_bloom_query(const sbwt_bloom* b, const uint64_t* d_keys, uint64_t count, uint8_t* d_out, void* stream) {
  clear_error();
  SBWT_REQUIRE(b, "filter is NULL");
  if (count == 0) return SBWT_OK;
  k_bloom_query<<<(count + 255) / 256, 256, 0, as_stream(stream)>>>(b->d_bits, b->blocks, b->hashes, b->k,
                                                                   b->keymask, d_keys, count, d_out);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

int sbwt_bloom_query_reads(const sbwt_bloom* b, const uint64_t* d_packed, const uint64_t* d_row_starts,
                           const uint64_t* d_offsets, uint64_t rows, uint8_t* d_out, void* stream) {
  clear_error();
  SBWT_REQUIRE(b, "filter is NULL");
  if (rows == 0) return SBWT_OK;
  SBWT_REQUIRE(d_packed && d_row_starts && d_offsets && d_out, "device buffers must not be NULL");
  SBWT_REQUIRE(b->blocks, "the filter was created for 0 items");
  const u64 grid = rows < (1u << 20) ? rows : (1u << 20);
  k_bloom_query_reads<<<grid, 64, 0, as_stream(stream)>>>(b->d_bits, b->blocks, b->hashes, b->k, b->keymask,
                                                        d_packed, d_row_starts, d_offsets, rows, d_out);
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

int sbwt_prefilter(const sbwt_mlp* nn, float threshold, const sbwt_bloom* b, const uint64_t* d_kmers, uint64_t count,
                   uint8_t* d_maybe, void* stream) {
  clear_error();
  SBWT_REQUIRE(nn, "model is NULL");
  if (count == 0) return SBWT_OK;
  SBWT_REQUIRE(d_kmers && d_maybe, "device buffers must not be NULL");
  SBWT_REQUIRE(!b || b->k == nn->shape.k, "the Bloom filter is for k = %u, the model for k = %u", b ? b->k : 0,
               nn->shape.k);
  const u64 groups = (count + 63) / 64;
  k_prefilter<<<resident_grid(k_prefilter, groups), 256, 0, as_stream(stream)>>>(mlp_dev(nn), threshold, d_kmers,
                                                                                  count, d_maybe);
  if (b && b->blocks) {
    SBWT_HIP(hipGetLastError());
    k_bloom_or<<<(count + 255) / 256, 256, 0, as_stream(stream)>>>(b->d_bits, b->blocks, b->hashes, b->k,
                                                                  b->keymask, d_kmers, count, d_maybe);
  }
  SBWT_HIP(hipGetLastError());
  return SBWT_OK;
}

}  // extern "C"
