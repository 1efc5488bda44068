// Learned initial density on CDNA4: feature kernel, the three-layer network on the bf16 matrix
// cores (v_mfma_f32_16x16x32_bf16 / 16x16x16_bf16, fp32 accumulation), and its training step
// (loss, back-propagation, weight gradients, Adam) written as explicit kernels.
//
// Matrix-core formulation.  Activations are kept transposed: a layer computes Z^T = W^T X^T with
// the weights as the A operand (rows = output features) and the activations as the B operand
// (columns = directed pairs).  A 16x16 result tile leaves lane l holding features 4(l/16)+r of
// pair l%16, which is exactly a B operand of the next layer if that layer's k index is permuted:
// k-slot (kt, g, j) of the next layer stands for feature 16 (2 kt + j/4) + 4 g + j%4.  The
// weight fragments are pre-arranged in that order (pack_kernel), so hidden activations never
// leave registers between layers.  Weight gradients need the pair dimension as k; the training
// kernel writes its activations through LDS into a chunk-major layout [pair/8][feature][8] from
// which the split-k gradient kernel reads 16-byte operands.
//
// Every reduction has a fixed order (no floating-point atomics), so training is reproducible
// bit for bit for a given seed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>

#include "pm3_guess.hpp"

using namespace pm3guess;

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef short i16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));

#define HIP_OK(x)                                            \
  do {                                                       \
    hipError_t e_ = (x);                                     \
    if (e_ != hipSuccess) {                                  \
      err = std::string(#x) + ": " + hipGetErrorString(e_);  \
      return false;                                          \
    }                                                        \
  } while (0)

// Fragment buffer (u32x4 units): forward A operands W1^T [8 ot][3 kt][64], W2^T [8][4][64],
// W3^T [4][64]; backward A operands W2 [8][4][64] and W3 (16x16x16, u32x2) [8][64]; then the
// biases b2 (NH) and b3 (NO) as floats.
constexpr int FR_F1 = 0, FR_F2 = FR_F1 + 8 * 3 * 64, FR_F3 = FR_F2 + 8 * 4 * 64, FR_G2 = FR_F3 + 4 * 64,
              FR_G3 = FR_G2 + 8 * 4 * 64, FR_END = FR_G3 + 8 * 64 / 2;
constexpr int FR_FWD = FR_G2;                      // forward fragments [0, FR_FWD)
constexpr size_t FRAG_BYTES = (size_t)FR_END * 16 + (NH + NO) * 4;
constexpr int NTILES_DW = 48 + 64 + 8 + 8 + 1;    // dW1, dW2, dW3, db2, db3 tiles of 16 x 16
constexpr int DW_SPLIT = 32;

__device__ inline f32x4 mma32(u32x4 a, u32x4 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0, 0,
                                                 0);
}
__device__ inline f32x4 mma16(u32x2 a, u32x2 b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(__builtin_bit_cast(i16x4, a), __builtin_bit_cast(i16x4, b), c, 0, 0,
                                                   0);
}
__device__ inline u32x2 pack4(f32x4 v) {
  u32x2 r;
  r.x = (uint32_t)f32_to_bf16(v.x) | ((uint32_t)f32_to_bf16(v.y) << 16);
  r.y = (uint32_t)f32_to_bf16(v.z) | ((uint32_t)f32_to_bf16(v.w) << 16);
  return r;
}
__device__ inline u32x4 join(u32x2 lo, u32x2 hi) {
  u32x4 r;
  r.x = lo.x; r.y = lo.y; r.z = hi.x; r.w = hi.y;
  return r;
}
__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float dsilu(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
__device__ inline f32x4 silu4(f32x4 v) { return f32x4{silu(v.x), silu(v.y), silu(v.z), silu(v.w)}; }
__device__ inline f32x4 zero4() { return f32x4{0.0f, 0.0f, 0.0f, 0.0f}; }

// k index of slot (kt, g, j) for layers fed from registers (see the header comment).
__host__ __device__ inline int kperm(int kt, int g, int j) { return 16 * (2 * kt + j / 4) + 4 * g + (j % 4); }

// ------------------------------------------------------------------------------------------
// Features of every directed pair.
__global__ __launch_bounds__(256) void feature_kernel(PairGeom G, const int* gsrc, long long npairs, uint16_t* X) {
  for (long long j = (long long)blockIdx.x * 256 + threadIdx.x; j < npairs; j += (long long)gridDim.x * 256) {
    uint16_t f[NF];
    pair_features(G, gsrc[j], G.gdst[j], f);
    u32x4* dst = (u32x4*)(X + j * NF);
#pragma unroll
    for (int c = 0; c < NF / 8; ++c) {
      u32x4 v;
      v.x = (uint32_t)f[8 * c] | ((uint32_t)f[8 * c + 1] << 16);
      v.y = (uint32_t)f[8 * c + 2] | ((uint32_t)f[8 * c + 3] << 16);
      v.z = (uint32_t)f[8 * c + 4] | ((uint32_t)f[8 * c + 5] << 16);
      v.w = (uint32_t)f[8 * c + 6] | ((uint32_t)f[8 * c + 7] << 16);
      dst[c] = v;
    }
  }
}

// ------------------------------------------------------------------------------------------
// Fragments from the fp32 master parameters.
__global__ __launch_bounds__(256) void pack_kernel(const float* theta, u32x4* frag) {
  const int t = blockIdx.x * 256 + threadIdx.x;
  auto w1 = [&](int k, int i) { return theta[OFF_W1 + k * NH + i]; };
  auto w2 = [&](int k, int i) { return theta[OFF_W2 + k * NH + i]; };
  auto w3 = [&](int k, int i) { return theta[OFF_W3 + k * NO + i]; };
  auto put8 = [&](int idx, const float* v) {
    u32x4 r;
    r.x = (uint32_t)f32_to_bf16(v[0]) | ((uint32_t)f32_to_bf16(v[1]) << 16);
    r.y = (uint32_t)f32_to_bf16(v[2]) | ((uint32_t)f32_to_bf16(v[3]) << 16);
    r.z = (uint32_t)f32_to_bf16(v[4]) | ((uint32_t)f32_to_bf16(v[5]) << 16);
    r.w = (uint32_t)f32_to_bf16(v[6]) | ((uint32_t)f32_to_bf16(v[7]) << 16);
    frag[idx] = r;
  };
  float v[8];
  if (t < FR_F2) {  // W1^T: A[i = 16 ot + l%16][k = 32 kt + 8 g + j]
    const int l = t % 64, kt = (t / 64) % 3, ot = t / 192, i = 16 * ot + l % 16, g = l / 16;
    for (int j = 0; j < 8; ++j) v[j] = w1(32 * kt + 8 * g + j, i);
    put8(t, v);
  } else if (t < FR_F3) {  // W2^T: A[i = h2][k = kperm h1]
    const int u = t - FR_F2, l = u % 64, kt = (u / 64) % 4, ot = u / 256, i = 16 * ot + l % 16, g = l / 16;
    for (int j = 0; j < 8; ++j) v[j] = w2(kperm(kt, g, j), i);
    put8(t, v);
  } else if (t < FR_G2) {  // W3^T: A[i = o][k = kperm h2]
    const int u = t - FR_F3, l = u % 64, kt = u / 64, i = l % 16, g = l / 16;
    for (int j = 0; j < 8; ++j) v[j] = w3(kperm(kt, g, j), i);
    put8(t, v);
  } else if (t < FR_G3) {  // W2 (backward): A[i = h1][k = kperm h2] = W2[h1][h2]
    const int u = t - FR_G2, l = u % 64, kt = (u / 64) % 4, ot = u / 256, i = 16 * ot + l % 16, g = l / 16;
    for (int j = 0; j < 8; ++j) v[j] = w2(i, kperm(kt, g, j));
    put8(t, v);
  } else if (t < FR_END) {  // W3 (backward, 16x16x16): A[i = h2][k = o = 4 g + j], two lanes per u32x4
    const int u = t - FR_G3;
    u32x2* g3 = (u32x2*)(frag + FR_G3);
    for (int half = 0; half < 2; ++half) {
      const int e = 2 * u + half, l = e % 64, ot = e / 64, i = 16 * ot + l % 16, g = l / 16;
      u32x2 r;
      r.x = (uint32_t)f32_to_bf16(w3(i, 4 * g)) | ((uint32_t)f32_to_bf16(w3(i, 4 * g + 1)) << 16);
      r.y = (uint32_t)f32_to_bf16(w3(i, 4 * g + 2)) | ((uint32_t)f32_to_bf16(w3(i, 4 * g + 3)) << 16);
      g3[e] = r;
    }
  } else if (t < FR_END + NH + NO) {
    float* bias = (float*)(frag + FR_END);
    const int u = t - FR_END;
    bias[u] = u < NH ? theta[OFF_B2 + u] : theta[OFF_B3 + (u - NH)];
  }
}

// ------------------------------------------------------------------------------------------
// Forward pass: one wave per 32 directed pairs (two 16-pair tiles sharing each weight fragment).
// Pair q of the launch reads input row idx[q] (or q without idx); y is written for q < n.
__device__ inline void load_x(const uint16_t* X, long long src, int g, u32x4 (&xb)[3]) {
#pragma unroll
  for (int kt = 0; kt < 3; ++kt)
    xb[kt] = src >= 0 ? *(const u32x4*)(X + src * NF + 32 * kt + 8 * g) : u32x4{0u, 0u, 0u, 0u};
}

__global__ __launch_bounds__(256) void mlp_fwd_kernel(const uint16_t* X, const int* idx, int n, int ntile,
                                                      const u32x4* frag, float* Y) {
  __shared__ u32x4 sF[FR_FWD];
  __shared__ float sB[NH + NO];
  for (int i = threadIdx.x; i < FR_FWD; i += 256) sF[i] = frag[i];
  for (int i = threadIdx.x; i < NH + NO; i += 256) sB[i] = ((const float*)(frag + FR_END))[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
  for (int wt = blockIdx.x * 4 + wave; wt < ntile; wt += gridDim.x * 4) {
    u32x4 xb[2][3];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int q = wt * 32 + 16 * t + c;
      const long long src = q < n ? (idx ? (long long)idx[q] : (long long)q) : -1;
      load_x(X, src, g, xb[t]);
    }
    u32x2 h1[2][8];
#pragma unroll
    for (int ot = 0; ot < 8; ++ot) {
      f32x4 a0 = zero4(), a1 = zero4();
#pragma unroll
      for (int kt = 0; kt < 3; ++kt) {
        const u32x4 A = sF[FR_F1 + (ot * 3 + kt) * 64 + lane];
        a0 = mma32(A, xb[0][kt], a0);
        a1 = mma32(A, xb[1][kt], a1);
      }
      h1[0][ot] = pack4(silu4(a0));
      h1[1][ot] = pack4(silu4(a1));
    }
    u32x2 h2[2][8];
#pragma unroll
    for (int ot = 0; ot < 8; ++ot) {
      const f32x4 b = {sB[16 * ot + 4 * g], sB[16 * ot + 4 * g + 1], sB[16 * ot + 4 * g + 2], sB[16 * ot + 4 * g + 3]};
      f32x4 a0 = b, a1 = b;
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) {
        const u32x4 A = sF[FR_F2 + (ot * 4 + kt) * 64 + lane];
        a0 = mma32(A, join(h1[0][2 * kt], h1[0][2 * kt + 1]), a0);
        a1 = mma32(A, join(h1[1][2 * kt], h1[1][2 * kt + 1]), a1);
      }
      h2[0][ot] = pack4(silu4(a0));
      h2[1][ot] = pack4(silu4(a1));
    }
    const f32x4 b3 = {sB[NH + 4 * g], sB[NH + 4 * g + 1], sB[NH + 4 * g + 2], sB[NH + 4 * g + 3]};
    f32x4 y0 = b3, y1 = b3;
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      const u32x4 A = sF[FR_F3 + kt * 64 + lane];
      y0 = mma32(A, join(h2[0][2 * kt], h2[0][2 * kt + 1]), y0);
      y1 = mma32(A, join(h2[1][2 * kt], h2[1][2 * kt + 1]), y1);
    }
    const int q0 = wt * 32 + c, q1 = q0 + 16;
    if (q0 < n) *(f32x4*)(Y + (size_t)q0 * NO + 4 * g) = y0;
    if (q1 < n) *(f32x4*)(Y + (size_t)q1 * NO + 4 * g) = y1;
  }
}

// ------------------------------------------------------------------------------------------
// Training: forward again, back-propagate dy to dZ2 and dZ1, and write the chunk-major
// activations [pair/8][feature][8] (bf16) that the weight-gradient kernel consumes.
// One wave per 16 pairs, eight waves per workgroup; dynamic LDS holds all fragments plus a
// 4 KB staging tile per wave.
constexpr int BWD_WAVES = 8;
constexpr size_t BWD_LDS = (size_t)FR_END * 16 + (NH + NO) * 4 + (size_t)BWD_WAVES * NH * 16 * 2;

struct ActOut {
  u32x4 *xt, *h1t, *dz1t, *h2t, *dz2t, *dyt;
};

// Staging of a wave's 16-pair tile: S[f * 16 + pair] holds bf16 bits; flush copies it to the
// chunk-major output (two 8-pair chunks per feature).  The workgroup barriers order the writes
// of the fill against the reads of the flush and the next fill.
__device__ inline void fill_z(uint16_t* S, const u32x2 (&v)[8], int g, int c) {
#pragma unroll
  for (int ot = 0; ot < 8; ++ot) {
    const int f = 16 * ot + 4 * g;
    S[(f + 0) * 16 + c] = (uint16_t)(v[ot].x & 0xFFFFu);
    S[(f + 1) * 16 + c] = (uint16_t)(v[ot].x >> 16);
    S[(f + 2) * 16 + c] = (uint16_t)(v[ot].y & 0xFFFFu);
    S[(f + 3) * 16 + c] = (uint16_t)(v[ot].y >> 16);
  }
}

__device__ inline void flush_stage(const uint16_t* S, int nfeat, int q0, u32x4* dst) {
  __syncthreads();
  if (dst) {
    const int lane = threadIdx.x & 63;
    for (int i = lane; i < 2 * nfeat; i += 64) {
      const int ch = i / nfeat, f = i % nfeat;
      dst[(size_t)(q0 / 8 + ch) * nfeat + f] = *(const u32x4*)(S + f * 16 + 8 * ch);
    }
  }
  __syncthreads();
}

__global__ __launch_bounds__(64 * BWD_WAVES) void mlp_bwd_kernel(const uint16_t* X, const int* idx, int n, int ntile16,
                                                                 const u32x4* frag, const float* dY, ActOut act) {
  extern __shared__ u32x4 lds_b[];
  u32x4* sF = lds_b;
  float* sB = (float*)(lds_b + FR_END);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
  uint16_t* S = (uint16_t*)(sB + NH + NO) + wave * NH * 16;
  for (int i = threadIdx.x; i < FR_END; i += 64 * BWD_WAVES) sF[i] = frag[i];
  for (int i = threadIdx.x; i < NH + NO; i += 64 * BWD_WAVES) sB[i] = ((const float*)(frag + FR_END))[i];
  __syncthreads();
  const u32x2* sG3 = (const u32x2*)(sF + FR_G3);
  // every wave runs the same number of iterations so the workgroup barriers in flush_stage match
  const int per = gridDim.x * BWD_WAVES;
  const int iters = (ntile16 + per - 1) / per;
  for (int it = 0; it < iters; ++it) {
    const int tile = it * per + blockIdx.x * BWD_WAVES + wave;
    const bool live = tile < ntile16;
    const int q0 = (live ? tile : 0) * 16, q = q0 + c;
    const long long src = (live && q < n) ? (idx ? (long long)idx[q] : (long long)q) : -1;
    u32x4 xb[3];
    load_x(X, src, g, xb);
    f32x4 z1[8], z2[8];
    u32x2 h1[8], h2[8];
#pragma unroll
    for (int ot = 0; ot < 8; ++ot) {
      f32x4 a = zero4();
#pragma unroll
      for (int kt = 0; kt < 3; ++kt) a = mma32(sF[FR_F1 + (ot * 3 + kt) * 64 + lane], xb[kt], a);
      z1[ot] = a;
      h1[ot] = pack4(silu4(a));
    }
#pragma unroll
    for (int ot = 0; ot < 8; ++ot) {
      f32x4 a = {sB[16 * ot + 4 * g], sB[16 * ot + 4 * g + 1], sB[16 * ot + 4 * g + 2], sB[16 * ot + 4 * g + 3]};
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) a = mma32(sF[FR_F2 + (ot * 4 + kt) * 64 + lane], join(h1[2 * kt], h1[2 * kt + 1]), a);
      z2[ot] = a;
      h2[ot] = pack4(silu4(a));
    }
    const f32x4 dy = src >= 0 ? *(const f32x4*)(dY + (size_t)q * NO + 4 * g) : zero4();
    const u32x2 dyb = pack4(dy);
    u32x2 dz2[8], dz1[8];
#pragma unroll
    for (int ot = 0; ot < 8; ++ot) {
      const f32x4 d = mma16(sG3[ot * 64 + lane], dyb, zero4());
      const f32x4 s = {dsilu(z2[ot].x), dsilu(z2[ot].y), dsilu(z2[ot].z), dsilu(z2[ot].w)};
      dz2[ot] = pack4(d * s);
    }
#pragma unroll
    for (int ot = 0; ot < 8; ++ot) {
      f32x4 d = zero4();
#pragma unroll
      for (int kt = 0; kt < 4; ++kt) d = mma32(sF[FR_G2 + (ot * 4 + kt) * 64 + lane], join(dz2[2 * kt], dz2[2 * kt + 1]), d);
      const f32x4 s = {dsilu(z1[ot].x), dsilu(z1[ot].y), dsilu(z1[ot].z), dsilu(z1[ot].w)};
      dz1[ot] = pack4(d * s);
    }
    // staging: lane (c, g) owns features 16 ot + 4 g + r of pair c (Z layout) or 32 kt + 8 g + j (x)
#pragma unroll
    for (int kt = 0; kt < 3; ++kt) {
      const uint32_t w[4] = {xb[kt].x, xb[kt].y, xb[kt].z, xb[kt].w};
#pragma unroll
      for (int j = 0; j < 8; ++j) S[(32 * kt + 8 * g + j) * 16 + c] = (uint16_t)((w[j / 2] >> (16 * (j % 2))) & 0xFFFFu);
    }
    flush_stage(S, NF, q0, live ? act.xt : nullptr);
    fill_z(S, h1, g, c);
    flush_stage(S, NH, q0, live ? act.h1t : nullptr);
    fill_z(S, dz1, g, c);
    flush_stage(S, NH, q0, live ? act.dz1t : nullptr);
    fill_z(S, h2, g, c);
    flush_stage(S, NH, q0, live ? act.h2t : nullptr);
    fill_z(S, dz2, g, c);
    flush_stage(S, NH, q0, live ? act.dz2t : nullptr);
    S[(4 * g + 0) * 16 + c] = (uint16_t)(dyb.x & 0xFFFFu);
    S[(4 * g + 1) * 16 + c] = (uint16_t)(dyb.x >> 16);
    S[(4 * g + 2) * 16 + c] = (uint16_t)(dyb.y & 0xFFFFu);
    S[(4 * g + 3) * 16 + c] = (uint16_t)(dyb.y >> 16);
    flush_stage(S, NO, q0, live ? act.dyt : nullptr);
  }
}

// Weight gradients as split-k products over the pair dimension.  Tile t of NTILES_DW: 0-47 dW1
// (6 x 8 tiles of NF x NH), 48-111 dW2 (8 x 8), 112-119 dW3 (8 x 1), 120-127 db2 (row 0 of
// ones x dZ2), 128 db3.  partial[split][tile][16 x 16] (row-major), reduced in fixed order.
__global__ __launch_bounds__(256) void dw_kernel(ActOut act, int nks, float* partial) {
  __shared__ f32x4 red[4][64];
  const int tile = blockIdx.x, split = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = lane >> 4, c = lane & 15;
  const u32x4 *A, *B;
  int FA = 0, FB = 0, rt = 0, ct = 0;
  bool ones = false;
  if (tile < 48) { A = act.xt; FA = NF; rt = tile / 8; B = act.dz1t; FB = NH; ct = tile % 8; }
  else if (tile < 112) { A = act.h1t; FA = NH; rt = (tile - 48) / 8; B = act.dz2t; FB = NH; ct = (tile - 48) % 8; }
  else if (tile < 120) { A = act.h2t; FA = NH; rt = tile - 112; B = act.dyt; FB = NO; ct = 0; }
  else if (tile < 128) { A = nullptr; ones = true; B = act.dz2t; FB = NH; ct = tile - 120; }
  else { A = nullptr; ones = true; B = act.dyt; FB = NO; ct = 0; }
  const u32x4 one = c == 0 ? u32x4{0x3F803F80u, 0x3F803F80u, 0x3F803F80u, 0x3F803F80u} : u32x4{0u, 0u, 0u, 0u};
  f32x4 acc = zero4();
  for (int ks = split * 4 + wave; ks < nks; ks += gridDim.y * 4) {
    const size_t ch = (size_t)ks * 4 + g;
    const u32x4 a = ones ? one : A[ch * FA + 16 * rt + c];
    const u32x4 b = B[ch * FB + 16 * ct + c];
    acc = mma32(a, b, acc);
  }
  red[wave][lane] = acc;
  __syncthreads();
  if (wave == 0) {
    f32x4 s = red[0][lane];
    s += red[1][lane];
    s += red[2][lane];
    s += red[3][lane];
    float* out = partial + ((size_t)split * NTILES_DW + tile) * 256;
    out[(4 * g + 0) * 16 + c] = s.x;
    out[(4 * g + 1) * 16 + c] = s.y;
    out[(4 * g + 2) * 16 + c] = s.z;
    out[(4 * g + 3) * 16 + c] = s.w;
  }
}

// Tile and position of parameter i in the partial-gradient layout.
__device__ inline void param_slot(int i, int& tile, int& pos) {
  if (i < OFF_W2) {
    const int k = i / NH, h = i % NH;
    tile = (k / 16) * 8 + h / 16; pos = (k % 16) * 16 + h % 16;
  } else if (i < OFF_B2) {
    const int u = i - OFF_W2, k = u / NH, h = u % NH;
    tile = 48 + (k / 16) * 8 + h / 16; pos = (k % 16) * 16 + h % 16;
  } else if (i < OFF_W3) {
    const int h = i - OFF_B2;
    tile = 120 + h / 16; pos = h % 16;
  } else if (i < OFF_B3) {
    const int u = i - OFF_W3, k = u / NO, o = u % NO;
    tile = 112 + k / 16; pos = (k % 16) * 16 + o;
  } else {
    tile = 128; pos = i - OFF_B3;
  }
}

__global__ __launch_bounds__(256) void grad_reduce_kernel(const float* partial, int nsplit, float* grad) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= NPARAM) return;
  int tile, pos;
  param_slot(i, tile, pos);
  float s = 0.0f;
  for (int k = 0; k < nsplit; ++k) s += partial[((size_t)k * NTILES_DW + tile) * 256 + pos];
  grad[i] = s;
}

__global__ __launch_bounds__(256) void adam_kernel(float* theta, float* m, float* v, const float* grad, float lr, float b1,
                                                   float b2, float eps, float c1, float c2) {
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= NPARAM) return;
  const float gi = grad[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  theta[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------------------------------------------------
// Training set and loss.
struct DataDev {
  const int *m_atom0, *m_nat, *m_pair0, *m_npair;
  const int8_t* a_elem;
  const int *a_pbeg, *a_pcnt;
  const float* a_tgt;
  const int *p_src, *p_dst, *p_part;
  const float4* p_u;
  const float* p_tgt;
};

__global__ __launch_bounds__(256) void expand_kernel(const int* smol, const int* soff, DataDev D, int* idx) {
  const int i = blockIdx.x, m = smol[i];
  const int p0 = D.m_pair0[m], np = D.m_npair[m], o = soff[i];
  for (int t = threadIdx.x; t < np; t += 256) idx[o + t] = p0 + t;
}

constexpr int LOSS_MAX_ATOMS = 256;

// One workgroup per molecule of the step: on-site residuals, pair residuals, the output
// gradient dy (scaled by `scale`) and the molecule's loss (scaled) in loss_mol[i].
__global__ __launch_bounds__(256) void loss_kernel(const int* smol, const int* soff, DataDev D, const float* Y, float* dY,
                                                   double scale, double* loss_mol) {
  __shared__ float R[LOSS_MAX_ATOMS][16];
  __shared__ double red[4];
  const int i = blockIdx.x, m = smol[i];
  const int a0 = D.m_atom0[m], nat = D.m_nat[m], p0 = D.m_pair0[m], np = D.m_npair[m], o = soff[i];
  double loss = 0.0;
  for (int a = threadIdx.x; a < nat; a += 256) {
    const int ga = a0 + a, e = D.a_elem[ga], na = e == 0 ? 1 : 4;
    double P[4][4];
    for (int r = 0; r < 4; ++r)
      for (int s = 0; s < 4; ++s) P[r][s] = 0.0;
    P[0][0] = sad_s(e);
    for (int k = 1; k < na; ++k) P[k][k] = sad_p(e);
    const int pb = D.a_pbeg[ga], pc = D.a_pcnt[ga];
    for (int t = 0; t < pc; ++t) {
      const int j = pb + t;
      const float4 u = D.p_u[j];
      double Mm[4][4];
      onsite_message(Y + (size_t)(o + (j - p0)) * NO, u.x, u.y, u.z, Mm);
      for (int r = 0; r < na; ++r)
        for (int s = 0; s < na; ++s) P[r][s] += Mm[r][s];
    }
    for (int r = 0; r < 4; ++r)
      for (int s = 0; s < 4; ++s) {
        const double d = (r < na && s < na) ? P[r][s] - (double)D.a_tgt[(size_t)ga * 16 + r * 4 + s] : 0.0;
        R[a][r * 4 + s] = (float)d;
        loss += d * d;
      }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < np; t += 256) {
    const int j = p0 + t, jp = D.p_part[j];
    const int a = D.p_src[j], b = D.p_dst[j];
    const int na = D.a_elem[a] == 0 ? 1 : 4, nb = D.a_elem[b] == 0 ? 1 : 4;
    const float4 u = D.p_u[j], up = D.p_u[jp];
    double Bab[4][4], Bba[4][4];
    pair_block(Y + (size_t)(o + t) * NO, u.x, u.y, u.z, Bab);
    pair_block(Y + (size_t)(o + (jp - p0)) * NO, up.x, up.y, up.z, Bba);
    double g[NO];
    for (int k = 0; k < NO; ++k) g[k] = 0.0;
    const double ux[3] = {u.x, u.y, u.z};
    for (int r = 0; r < na; ++r)
      for (int s = 0; s < nb; ++s) {
        const double d = 0.5 * (Bab[r][s] + Bba[s][r]) - (double)D.p_tgt[(size_t)j * 16 + r * 4 + s];
        loss += d * d;
        const double r2 = 2.0 * d;
        if (r == 0 && s == 0) g[0] += r2;
        else if (r == 0) g[1] += r2 * ux[s - 1];
        else if (s == 0) g[2] += r2 * ux[r - 1];
        else {
          const double uu = ux[r - 1] * ux[s - 1];
          g[3] += r2 * uu;
          g[4] += r2 * ((r == s ? 1.0 : 0.0) - uu);
        }
      }
    const int la = a - a0;
    for (int r = 0; r < na; ++r)
      for (int s = 0; s < na; ++s) {
        const double r2 = 2.0 * (double)R[la][r * 4 + s];
        if (r == 0 && s == 0) g[5] += r2;
        else if (r == 0) g[6] += r2 * ux[s - 1];
        else if (s == 0) g[6] += r2 * ux[r - 1];
        else {
          const double uu = ux[r - 1] * ux[s - 1];
          g[7] += r2 * uu;
          g[8] += r2 * ((r == s ? 1.0 : 0.0) - uu);
        }
      }
    if (dY) {
      float* out = dY + (size_t)(o + t) * NO;
      for (int k = 0; k < NO; ++k) out[k] = (float)(g[k] * scale);
    }
  }
  // fixed-order block reduction
  for (int off = 32; off > 0; off >>= 1) loss += __shfl_down(loss, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = loss;
  __syncthreads();
  if (threadIdx.x == 0) loss_mol[i] = (red[0] + red[1] + red[2] + red[3]) * scale;
}

__global__ void loss_sum_kernel(const double* loss_mol, int nm, double* out) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s = 0.0;
  for (int k = 0; k < nm; ++k) s += loss_mol[k];
  *out = s;
}

// Training samples of one solved batch: features, unit vectors and target blocks of every
// directed pair, on-site targets of every atom.
struct ExtractJob {
  int atom0, nat, ngp, nao;
  long long gp0, out_mat;
  int datom0, dpair0;
};

struct DataOut {
  int8_t* a_elem;
  int *a_pbeg, *a_pcnt;
  float* a_tgt;
  int *p_src, *p_dst, *p_part;
  float4* p_u;
  float* p_tgt;
  uint16_t* p_x;
};

__global__ __launch_bounds__(256) void extract_kernel(const ExtractJob* jobs, PairGeom G, const int* aoff, const int* gsrc,
                                                      const int* gpart, const double* outP, DataOut O) {
  const ExtractJob J = jobs[blockIdx.x];
  const double* P = outP + J.out_mat;
  const int n = J.nao;
  for (int a = threadIdx.x; a < J.nat; a += 256) {
    const int ga = J.atom0 + a, da = J.datom0 + a, e = G.elem[ga], na = e == 0 ? 1 : 4, o = aoff[ga];
    O.a_elem[da] = (int8_t)e;
    O.a_pbeg[da] = J.dpair0 + (int)(G.gbeg[ga] - J.gp0);
    O.a_pcnt[da] = G.gcnt[ga];
    for (int r = 0; r < 4; ++r)
      for (int s = 0; s < 4; ++s)
        O.a_tgt[(size_t)da * 16 + r * 4 + s] = (r < na && s < na) ? (float)P[(size_t)(o + r) * n + o + s] : 0.0f;
  }
  for (int t = threadIdx.x; t < J.ngp; t += 256) {
    const long long j = J.gp0 + t;
    const int a = gsrc[j], b = G.gdst[j], dj = J.dpair0 + t;
    O.p_src[dj] = J.datom0 + (a - J.atom0);
    O.p_dst[dj] = J.datom0 + (b - J.atom0);
    O.p_part[dj] = J.dpair0 + (int)(gpart[j] - J.gp0);
    const double dx = G.x[b] - G.x[a], dy = G.y[b] - G.y[a], dz = G.z[b] - G.z[a];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    O.p_u[dj] = make_float4((float)(dx / r), (float)(dy / r), (float)(dz / r), 0.0f);
    const int na = G.elem[a] == 0 ? 1 : 4, nb = G.elem[b] == 0 ? 1 : 4, oa = aoff[a], ob = aoff[b];
    for (int rr = 0; rr < 4; ++rr)
      for (int s = 0; s < 4; ++s)
        O.p_tgt[(size_t)dj * 16 + rr * 4 + s] = (rr < na && s < nb) ? (float)P[(size_t)(oa + rr) * n + ob + s] : 0.0f;
    uint16_t f[NF];
    pair_features(G, a, b, f);
    for (int k = 0; k < NF; ++k) O.p_x[(size_t)dj * NF + k] = f[k];
  }
}

// ------------------------------------------------------------------------------------------
// Host helpers.
uint64_t splitmix(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform01(uint64_t& s) { return ((splitmix(s) >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
double normal(uint64_t& s) {
  const double u1 = uniform01(s), u2 = uniform01(s);
  return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * 3.14159265358979323846 * u2);
}

template <class T>
bool dalloc(T*& p, size_t count, std::string& err) {
  p = nullptr;
  hipError_t e = hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(count, 1));
  if (e != hipSuccess) {
    err = std::string("hipMalloc: ") + hipGetErrorString(e);
    p = nullptr;
    return false;
  }
  return true;
}
template <class T>
void dfree(T*& p) {
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

}  // namespace

namespace pm3host {

struct GuessModel {
  float *theta = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *partial = nullptr;
  u32x4* frag = nullptr;
  int step = 0;
  ~GuessModel() {
    dfree(theta); dfree(m); dfree(v); dfree(grad); dfree(partial); dfree(frag);
  }
};

static bool model_alloc(GuessModel* M, std::string& err) {
  if (!dalloc(M->theta, NPARAM, err) || !dalloc(M->m, NPARAM, err) || !dalloc(M->v, NPARAM, err) ||
      !dalloc(M->grad, NPARAM, err) || !dalloc(M->partial, (size_t)DW_SPLIT * NTILES_DW * 256, err))
    return false;
  unsigned char* fb = nullptr;
  if (!dalloc(fb, FRAG_BYTES, err)) return false;
  M->frag = (u32x4*)fb;
  HIP_OK(hipMemset(M->m, 0, sizeof(float) * NPARAM));
  HIP_OK(hipMemset(M->v, 0, sizeof(float) * NPARAM));
  return true;
}

static bool model_pack(const GuessModel* M, std::string& err) {
  const int nthreads = FR_END + NH + NO;
  hipLaunchKernelGGL(pack_kernel, dim3((nthreads + 255) / 256), dim3(256), 0, 0, M->theta, M->frag);
  HIP_OK(hipGetLastError());
  return true;
}

bool guess_model_set_params(GuessModel* M, const float* in, std::string& err) {
  HIP_OK(hipMemcpy(M->theta, in, sizeof(float) * NPARAM, hipMemcpyHostToDevice));
  HIP_OK(hipMemset(M->m, 0, sizeof(float) * NPARAM));
  HIP_OK(hipMemset(M->v, 0, sizeof(float) * NPARAM));
  if (!model_pack(M, err)) return false;
  HIP_OK(hipDeviceSynchronize());
  return true;
}

bool guess_model_create(uint64_t seed, GuessModel*& out, std::string& err) {
  out = nullptr;
  GuessModel* M = new GuessModel();
  if (!model_alloc(M, err)) { delete M; return false; }
  std::vector<float> th(NPARAM, 0.0f);
  uint64_t s = seed * 0x2545F4914F6CDD1Dull + 12345;
  // W1: rows 0..92 carry features, 93-94 are constant zero, row 95 is the first bias (zero)
  for (int k = 0; k < 93; ++k)
    for (int h = 0; h < NH; ++h) th[OFF_W1 + k * NH + h] = (float)(normal(s) / std::sqrt(93.0));
  for (int k = 0; k < NH * NH; ++k) th[OFF_W2 + k] = (float)(normal(s) / std::sqrt((double)NH));
  for (int k = 0; k < NH * NO; ++k) th[OFF_W3 + k] = (float)(0.1 * normal(s) / std::sqrt((double)NH));
  if (!guess_model_set_params(M, th.data(), err)) { delete M; return false; }
  out = M;
  return true;
}

struct ModelFileHeader {
  char magic[8];
  int32_t nf, nh, no, nparam, steps, reserved[3];
};

bool guess_model_load(const char* path, GuessModel*& out, std::string& err) {
  out = nullptr;
  FILE* f = std::fopen(path, "rb");
  if (!f) { err = std::string("cannot open ") + path; return false; }
  ModelFileHeader h{};
  std::vector<float> th(NPARAM);
  const bool ok = std::fread(&h, sizeof(h), 1, f) == 1 && std::memcmp(h.magic, "PM3GSV1", 8) == 0 && h.nf == NF &&
                  h.nh == NH && h.no == NO && h.nparam == NPARAM &&
                  std::fread(th.data(), sizeof(float), NPARAM, f) == (size_t)NPARAM;
  std::fclose(f);
  if (!ok) { err = std::string("not a guess-model file of this network: ") + path; return false; }
  for (float t : th)
    if (!std::isfinite(t)) { err = "non-finite parameter in model file"; return false; }
  GuessModel* M = new GuessModel();
  if (!model_alloc(M, err) || !guess_model_set_params(M, th.data(), err)) { delete M; return false; }
  M->step = h.steps;
  out = M;
  return true;
}

bool guess_model_params(const GuessModel* M, float* outp, std::string& err) {
  HIP_OK(hipMemcpy(outp, M->theta, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  return true;
}

bool guess_model_save(const GuessModel* M, const char* path, std::string& err) {
  std::vector<float> th(NPARAM);
  if (!guess_model_params(M, th.data(), err)) return false;
  ModelFileHeader h{};
  std::memcpy(h.magic, "PM3GSV1", 8);
  h.nf = NF; h.nh = NH; h.no = NO; h.nparam = NPARAM; h.steps = M->step;
  FILE* f = std::fopen(path, "wb");
  if (!f) { err = std::string("cannot write ") + path; return false; }
  const bool ok = std::fwrite(&h, sizeof(h), 1, f) == 1 && std::fwrite(th.data(), sizeof(float), NPARAM, f) == (size_t)NPARAM;
  std::fclose(f);
  if (!ok) { err = std::string("short write to ") + path; return false; }
  return true;
}

int guess_model_steps(const GuessModel* M) { return M ? M->step : 0; }
void guess_model_destroy(GuessModel* M) { delete M; }

static int fwd_blocks(int ntile) { return std::max(1, std::min((ntile + 3) / 4, 2048)); }

bool guess_infer(const GuessModel* M, const BatchGuessView& v, std::string& err, hipStream_t stream,
                 hipEvent_t between) {
  if (v.ngp_total == 0) {
    if (between) HIP_OK(hipEventRecord(between, stream));
    return true;
  }
  PairGeom G{v.ax, v.ay, v.az, v.elem, v.gbeg, v.gcnt, v.gdst};
  const int fb = (int)std::min<long long>((v.ngp_total + 255) / 256, 16384);
  hipLaunchKernelGGL(feature_kernel, dim3(fb), dim3(256), 0, stream, G, v.gsrc, v.ngp_total, v.X);
  if (between) HIP_OK(hipEventRecord(between, stream));
  const int n = (int)v.ngp_total, ntile = (n + 31) / 32;
  hipLaunchKernelGGL(mlp_fwd_kernel, dim3(fwd_blocks(ntile)), dim3(256), 0, stream, v.X, (const int*)nullptr, n, ntile,
                     (const u32x4*)M->frag, v.Y);
  HIP_OK(hipGetLastError());
  return true;
}

// ------------------------------------------------------------------------------------------
struct GuessData {
  int *m_atom0 = nullptr, *m_nat = nullptr, *m_pair0 = nullptr, *m_npair = nullptr;
  int8_t* a_elem = nullptr;
  int *a_pbeg = nullptr, *a_pcnt = nullptr;
  float* a_tgt = nullptr;
  int *p_src = nullptr, *p_dst = nullptr, *p_part = nullptr;
  float4* p_u = nullptr;
  float* p_tgt = nullptr;
  uint16_t* p_x = nullptr;
  long long nmol = 0, natom = 0, npair = 0, cap_mol = 0, cap_atom = 0, cap_pair = 0;
  std::vector<int> h_nat, h_npair, h_atom0, h_pair0;
  ~GuessData() {
    dfree(m_atom0); dfree(m_nat); dfree(m_pair0); dfree(m_npair); dfree(a_elem); dfree(a_pbeg); dfree(a_pcnt);
    dfree(a_tgt); dfree(p_src); dfree(p_dst); dfree(p_part); dfree(p_u); dfree(p_tgt); dfree(p_x);
  }
  DataDev dev() const {
    return DataDev{m_atom0, m_nat, m_pair0, m_npair, a_elem, a_pbeg, a_pcnt, a_tgt, p_src, p_dst, p_part, p_u, p_tgt};
  }
};

template <class T>
static bool grow(T*& p, long long used, long long cap_new, int per, std::string& err) {
  T* q = nullptr;
  if (!dalloc(q, (size_t)cap_new * per, err)) return false;
  if (p && used > 0) HIP_OK(hipMemcpy(q, p, sizeof(T) * (size_t)used * per, hipMemcpyDeviceToDevice));
  dfree(p);
  p = q;
  return true;
}

static bool data_reserve(GuessData* D, long long nm, long long na, long long np, std::string& err) {
  if (nm > D->cap_mol) {
    const long long c = std::max(nm, 2 * D->cap_mol);
    if (!grow(D->m_atom0, D->nmol, c, 1, err) || !grow(D->m_nat, D->nmol, c, 1, err) ||
        !grow(D->m_pair0, D->nmol, c, 1, err) || !grow(D->m_npair, D->nmol, c, 1, err))
      return false;
    D->cap_mol = c;
  }
  if (na > D->cap_atom) {
    const long long c = std::max(na, 2 * D->cap_atom);
    if (!grow(D->a_elem, D->natom, c, 1, err) || !grow(D->a_pbeg, D->natom, c, 1, err) ||
        !grow(D->a_pcnt, D->natom, c, 1, err) || !grow(D->a_tgt, D->natom, c, 16, err))
      return false;
    D->cap_atom = c;
  }
  if (np > D->cap_pair) {
    const long long c = std::max(np, 2 * D->cap_pair);
    if (!grow(D->p_src, D->npair, c, 1, err) || !grow(D->p_dst, D->npair, c, 1, err) ||
        !grow(D->p_part, D->npair, c, 1, err) || !grow(D->p_u, D->npair, c, 1, err) ||
        !grow(D->p_tgt, D->npair, c, 16, err) || !grow(D->p_x, D->npair, c, NF, err))
      return false;
    D->cap_pair = c;
  }
  return true;
}

bool guess_data_create(GuessData*& out, std::string& err) {
  (void)err;
  out = new GuessData();
  return true;
}

void guess_data_destroy(GuessData* d) { delete d; }

void guess_data_stats(const GuessData* d, long long& n_mol, long long& n_atoms, long long& n_pairs) {
  n_mol = d ? d->nmol : 0;
  n_atoms = d ? d->natom : 0;
  n_pairs = d ? d->npair : 0;
}

bool guess_data_add(GuessData* D, Batch* b, long long& added, std::string& err) {
  added = 0;
  BatchGuessView v;
  if (!batch_guess_view(b, v, err)) return false;
  if (!v.out_P) { err = "the batch was created without density output"; return false; }
  std::vector<ExtractJob> jobs;
  long long na = D->natom, np = D->npair;
  for (int k = 0; k < v.nmol; ++k) {
    if (!v.solved[k] || !v.eligible[k] || v.nat[k] > LOSS_MAX_ATOMS || v.out_mat[k] < 0) continue;
    ExtractJob J;
    J.atom0 = v.atom0[k]; J.nat = v.nat[k]; J.ngp = v.ngp[k]; J.nao = v.nao[k];
    J.gp0 = v.gp0[k]; J.out_mat = v.out_mat[k];
    J.datom0 = (int)na; J.dpair0 = (int)np;
    na += J.nat;
    np += J.ngp;
    jobs.push_back(J);
  }
  if (jobs.empty()) return true;
  if (np >= (1LL << 31) || na >= (1LL << 31)) { err = "training set exceeds 2^31 pairs"; return false; }
  if (!data_reserve(D, D->nmol + (long long)jobs.size(), na, np, err)) return false;
  ExtractJob* d_jobs = nullptr;
  if (!dalloc(d_jobs, jobs.size(), err)) return false;
  HIP_OK(hipMemcpy(d_jobs, jobs.data(), sizeof(ExtractJob) * jobs.size(), hipMemcpyHostToDevice));
  PairGeom G{v.ax, v.ay, v.az, v.elem, v.gbeg, v.gcnt, v.gdst};
  DataOut O{D->a_elem, D->a_pbeg, D->a_pcnt, D->a_tgt, D->p_src, D->p_dst, D->p_part, D->p_u, D->p_tgt, D->p_x};
  hipLaunchKernelGGL(extract_kernel, dim3((unsigned)jobs.size()), dim3(256), 0, 0, d_jobs, G, v.aoff, v.gsrc, v.gpart,
                     v.out_P, O);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  dfree(d_jobs);
  if (e != hipSuccess) { err = std::string("extract_kernel: ") + hipGetErrorString(e); return false; }
  std::vector<int> ma, mn, mp, mq;
  for (const ExtractJob& J : jobs) {
    ma.push_back(J.datom0); mn.push_back(J.nat); mp.push_back(J.dpair0); mq.push_back(J.ngp);
  }
  const size_t m0 = (size_t)D->nmol, nm = jobs.size();
  HIP_OK(hipMemcpy(D->m_atom0 + m0, ma.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(D->m_nat + m0, mn.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(D->m_pair0 + m0, mp.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(D->m_npair + m0, mq.data(), sizeof(int) * nm, hipMemcpyHostToDevice));
  D->h_atom0.insert(D->h_atom0.end(), ma.begin(), ma.end());
  D->h_nat.insert(D->h_nat.end(), mn.begin(), mn.end());
  D->h_pair0.insert(D->h_pair0.end(), mp.begin(), mp.end());
  D->h_npair.insert(D->h_npair.end(), mq.begin(), mq.end());
  D->nmol += (long long)nm;
  D->natom = na;
  D->npair = np;
  added = (long long)nm;
  return true;
}

// Per-step scratch for training and evaluation.
struct StepBufs {
  int *smol = nullptr, *soff = nullptr, *idx = nullptr;
  float *Y = nullptr, *dY = nullptr;
  double *loss_mol = nullptr, *loss_out = nullptr;
  u32x4 *xt = nullptr, *h1t = nullptr, *dz1t = nullptr, *h2t = nullptr, *dz2t = nullptr, *dyt = nullptr;
  long long cap_pairs = 0;
  int cap_mol = 0;
  ~StepBufs() {
    dfree(smol); dfree(soff); dfree(idx); dfree(Y); dfree(dY); dfree(loss_mol); dfree(loss_out);
    dfree(xt); dfree(h1t); dfree(dz1t); dfree(h2t); dfree(dz2t); dfree(dyt);
  }
  bool alloc(int nm, long long npairs, int total_steps, bool train, std::string& err) {
    cap_mol = nm;
    cap_pairs = (npairs + 31) / 32 * 32;
    const size_t nch = (size_t)cap_pairs / 8;
    if (!dalloc(smol, (size_t)nm * total_steps, err) || !dalloc(soff, (size_t)nm * total_steps, err) ||
        !dalloc(idx, (size_t)cap_pairs, err) || !dalloc(Y, (size_t)cap_pairs * NO, err) ||
        !dalloc(dY, (size_t)cap_pairs * NO, err) || !dalloc(loss_mol, (size_t)nm, err) ||
        !dalloc(loss_out, (size_t)std::max(total_steps, 1), err))
      return false;
    if (train && (!dalloc(xt, nch * NF, err) || !dalloc(h1t, nch * NH, err) || !dalloc(dz1t, nch * NH, err) ||
                  !dalloc(h2t, nch * NH, err) || !dalloc(dz2t, nch * NH, err) || !dalloc(dyt, nch * NO, err)))
      return false;
    return true;
  }
};

// Molecule lists of `nsteps` steps of `bm` molecules drawn without replacement per epoch.
static void plan_steps(const GuessData* D, int bm, int nsteps, uint64_t seed, bool shuffle, std::vector<int>& smol,
                       std::vector<int>& soff, std::vector<long long>& spairs, std::vector<long long>& satoms) {
  const int nm = (int)D->nmol;
  std::vector<int> perm(nm);
  std::iota(perm.begin(), perm.end(), 0);
  uint64_t s = seed * 0x9E3779B97F4A7C15ull + 7;
  int pos = nm;
  smol.assign((size_t)bm * nsteps, 0);
  soff.assign((size_t)bm * nsteps, 0);
  spairs.assign(nsteps, 0);
  satoms.assign(nsteps, 0);
  for (int st = 0; st < nsteps; ++st) {
    if (pos + bm > nm) {
      if (shuffle)
        for (int i = nm - 1; i > 0; --i) std::swap(perm[i], perm[(int)(splitmix(s) % (uint64_t)(i + 1))]);
      pos = 0;
    }
    long long off = 0, at = 0;
    for (int k = 0; k < bm; ++k) {
      const int m = perm[pos + k];
      smol[(size_t)st * bm + k] = m;
      soff[(size_t)st * bm + k] = (int)off;
      off += D->h_npair[m];
      at += D->h_nat[m];
    }
    spairs[st] = off;
    satoms[st] = at;
    pos += bm;
  }
}

bool guess_train(GuessModel* M, GuessData* D, const GuessTrainOptions& o, float* losses, double* seconds,
                 std::string& err) {
  if (D->nmol < 1) { err = "empty training set"; return false; }
  const int bm = (int)std::min<long long>(o.batch_molecules, D->nmol);
  if (bm < 1 || o.steps < 1) { err = "steps and batch_molecules must be >= 1"; return false; }
  std::vector<int> smol, soff;
  std::vector<long long> spairs, satoms;
  plan_steps(D, bm, o.steps, o.seed, true, smol, soff, spairs, satoms);
  const long long maxp = *std::max_element(spairs.begin(), spairs.end());
  StepBufs S;
  if (!S.alloc(bm, maxp, o.steps, true, err)) return false;
  HIP_OK(hipMemcpy(S.smol, smol.data(), sizeof(int) * smol.size(), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(S.soff, soff.data(), sizeof(int) * soff.size(), hipMemcpyHostToDevice));
  HIP_OK(hipFuncSetAttribute((const void*)mlp_bwd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)BWD_LDS));
  const DataDev DD = D->dev();
  ActOut act{S.xt, S.h1t, S.dz1t, S.h2t, S.dz2t, S.dyt};
  const auto t0 = std::chrono::steady_clock::now();
  for (int st = 0; st < o.steps; ++st) {
    const int* sm = S.smol + (size_t)st * bm;
    const int* so = S.soff + (size_t)st * bm;
    const int n = (int)spairs[st];
    const int npad = (n + 31) / 32 * 32;
    hipLaunchKernelGGL(expand_kernel, dim3(bm), dim3(256), 0, 0, sm, so, DD, S.idx);
    const int ntile = npad / 32;
    hipLaunchKernelGGL(mlp_fwd_kernel, dim3(fwd_blocks(ntile)), dim3(256), 0, 0, (const uint16_t*)D->p_x,
                       (const int*)S.idx, n, ntile, (const u32x4*)M->frag, S.Y);
    const double scale = 1.0 / (double)std::max<long long>(satoms[st], 1);
    hipLaunchKernelGGL(loss_kernel, dim3(bm), dim3(256), 0, 0, sm, so, DD, (const float*)S.Y, S.dY, scale, S.loss_mol);
    hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(64), 0, 0, (const double*)S.loss_mol, bm, S.loss_out + st);
    const int nt16 = npad / 16;
    const int bb = std::max(1, std::min((nt16 + BWD_WAVES - 1) / BWD_WAVES, 1024));
    hipLaunchKernelGGL(mlp_bwd_kernel, dim3(bb), dim3(64 * BWD_WAVES), BWD_LDS, 0, (const uint16_t*)D->p_x,
                       (const int*)S.idx, n, nt16, (const u32x4*)M->frag, (const float*)S.dY, act);
    const int nks = npad / 32;
    const int nsplit = std::max(1, std::min(DW_SPLIT, (nks + 3) / 4));
    hipLaunchKernelGGL(dw_kernel, dim3(NTILES_DW, nsplit), dim3(256), 0, 0, act, nks, M->partial);
    hipLaunchKernelGGL(grad_reduce_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, (const float*)M->partial, nsplit,
                       M->grad);
    float lr = o.lr;
    if (o.decay_steps > 0) {
      const double t = std::min(1.0, (double)M->step / (double)o.decay_steps);
      lr = (float)(o.lr_final + 0.5 * (o.lr - o.lr_final) * (1.0 + std::cos(3.14159265358979323846 * t)));
    }
    M->step += 1;
    const float c1 = 1.0f - std::pow(o.beta1, (float)M->step), c2 = 1.0f - std::pow(o.beta2, (float)M->step);
    hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, M->theta, M->m, M->v,
                       (const float*)M->grad, lr, o.beta1, o.beta2, o.eps, c1, c2);
    if (!model_pack(M, err)) return false;
    HIP_OK(hipGetLastError());
  }
  HIP_OK(hipDeviceSynchronize());
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  if (losses) {
    std::vector<double> l(o.steps);
    HIP_OK(hipMemcpy(l.data(), S.loss_out, sizeof(double) * o.steps, hipMemcpyDeviceToHost));
    for (int k = 0; k < o.steps; ++k) losses[k] = (float)l[k];
  }
  return true;
}

// Gradient (NPARAM floats) and loss of the whole set as one step, without an update.
bool guess_grad(const GuessModel* M, GuessData* D, float* grad, double& loss, std::string& err) {
  if (D->nmol < 1) { err = "empty data set"; return false; }
  const int bm = (int)D->nmol;
  std::vector<int> smol(bm), soff(bm);
  long long off = 0, atoms = 0;
  for (int k = 0; k < bm; ++k) {
    smol[k] = k;
    soff[k] = (int)off;
    off += D->h_npair[k];
    atoms += D->h_nat[k];
  }
  if (off >= (1LL << 31)) { err = "data set too large for one step"; return false; }
  StepBufs S;
  if (!S.alloc(bm, off, 1, true, err)) return false;
  HIP_OK(hipMemcpy(S.smol, smol.data(), sizeof(int) * bm, hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(S.soff, soff.data(), sizeof(int) * bm, hipMemcpyHostToDevice));
  HIP_OK(hipFuncSetAttribute((const void*)mlp_bwd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)BWD_LDS));
  const DataDev DD = D->dev();
  ActOut act{S.xt, S.h1t, S.dz1t, S.h2t, S.dz2t, S.dyt};
  const int n = (int)off, npad = (n + 31) / 32 * 32, ntile = npad / 32;
  hipLaunchKernelGGL(expand_kernel, dim3(bm), dim3(256), 0, 0, S.smol, S.soff, DD, S.idx);
  hipLaunchKernelGGL(mlp_fwd_kernel, dim3(fwd_blocks(ntile)), dim3(256), 0, 0, (const uint16_t*)D->p_x,
                     (const int*)S.idx, n, ntile, (const u32x4*)M->frag, S.Y);
  const double scale = 1.0 / (double)std::max<long long>(atoms, 1);
  hipLaunchKernelGGL(loss_kernel, dim3(bm), dim3(256), 0, 0, S.smol, S.soff, DD, (const float*)S.Y, S.dY, scale,
                     S.loss_mol);
  hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(64), 0, 0, (const double*)S.loss_mol, bm, S.loss_out);
  const int nt16 = npad / 16;
  const int bb = std::max(1, std::min((nt16 + BWD_WAVES - 1) / BWD_WAVES, 1024));
  hipLaunchKernelGGL(mlp_bwd_kernel, dim3(bb), dim3(64 * BWD_WAVES), BWD_LDS, 0, (const uint16_t*)D->p_x,
                     (const int*)S.idx, n, nt16, (const u32x4*)M->frag, (const float*)S.dY, act);
  const int nks = npad / 32;
  const int nsplit = std::max(1, std::min(DW_SPLIT, (nks + 3) / 4));
  hipLaunchKernelGGL(dw_kernel, dim3(NTILES_DW, nsplit), dim3(256), 0, 0, act, nks, M->partial);
  hipLaunchKernelGGL(grad_reduce_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, (const float*)M->partial, nsplit,
                     M->grad);
  HIP_OK(hipGetLastError());
  HIP_OK(hipMemcpy(grad, M->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(&loss, S.loss_out, sizeof(double), hipMemcpyDeviceToHost));
  return true;
}

bool guess_eval(const GuessModel* M, GuessData* D, double& loss, std::string& err) {
  loss = 0.0;
  if (D->nmol < 1) { err = "empty data set"; return false; }
  const int bm = (int)std::min<long long>(1024, D->nmol);
  const int nsteps = (int)((D->nmol + bm - 1) / bm);
  // sequential molecule order; the last step may be short
  std::vector<int> smol((size_t)bm * nsteps, 0), soff((size_t)bm * nsteps, 0), cnt(nsteps, 0);
  std::vector<long long> spairs(nsteps, 0);
  long long total_atoms = 0;
  for (int st = 0; st < nsteps; ++st) {
    long long off = 0;
    for (int k = 0; k < bm; ++k) {
      const long long m = (long long)st * bm + k;
      if (m >= D->nmol) break;
      smol[(size_t)st * bm + k] = (int)m;
      soff[(size_t)st * bm + k] = (int)off;
      off += D->h_npair[m];
      total_atoms += D->h_nat[m];
      cnt[st] += 1;
    }
    spairs[st] = off;
  }
  const long long maxp = *std::max_element(spairs.begin(), spairs.end());
  StepBufs S;
  if (!S.alloc(bm, maxp, nsteps, false, err)) return false;
  HIP_OK(hipMemcpy(S.smol, smol.data(), sizeof(int) * smol.size(), hipMemcpyHostToDevice));
  HIP_OK(hipMemcpy(S.soff, soff.data(), sizeof(int) * soff.size(), hipMemcpyHostToDevice));
  const DataDev DD = D->dev();
  const double scale = 1.0 / (double)std::max<long long>(total_atoms, 1);
  for (int st = 0; st < nsteps; ++st) {
    const int* sm = S.smol + (size_t)st * bm;
    const int* so = S.soff + (size_t)st * bm;
    const int n = (int)spairs[st], ntile = (n + 31) / 32;
    hipLaunchKernelGGL(expand_kernel, dim3(cnt[st]), dim3(256), 0, 0, sm, so, DD, S.idx);
    hipLaunchKernelGGL(mlp_fwd_kernel, dim3(fwd_blocks(ntile)), dim3(256), 0, 0, (const uint16_t*)D->p_x,
                       (const int*)S.idx, n, ntile, (const u32x4*)M->frag, S.Y);
    hipLaunchKernelGGL(loss_kernel, dim3(cnt[st]), dim3(256), 0, 0, sm, so, DD, (const float*)S.Y, (float*)nullptr,
                       scale, S.loss_mol);
    hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(64), 0, 0, (const double*)S.loss_mol, cnt[st], S.loss_out + st);
    HIP_OK(hipGetLastError());
  }
  std::vector<double> l(nsteps);
  HIP_OK(hipMemcpy(l.data(), S.loss_out, sizeof(double) * nsteps, hipMemcpyDeviceToHost));
  for (double x : l) loss += x;
  return true;
}

bool guess_mlp(const GuessModel* M, int n, const uint16_t* x, float* y, const float* dy, float* grad, std::string& err) {
  if (n < 1) { err = "n must be >= 1"; return false; }
  const int npad = (n + 31) / 32 * 32;
  uint16_t* dx = nullptr;
  float *dyv = nullptr, *dyd = nullptr;
  StepBufs S;
  bool ok = dalloc(dx, (size_t)npad * NF, err) && dalloc(dyv, (size_t)npad * NO, err) &&
            dalloc(dyd, (size_t)npad * NO, err) && S.alloc(1, npad, 1, dy != nullptr, err);
  auto cleanup = [&]() { dfree(dx); dfree(dyv); dfree(dyd); };
  if (!ok) { cleanup(); return false; }
  hipError_t e = hipMemcpy(dx, x, sizeof(uint16_t) * (size_t)n * NF, hipMemcpyHostToDevice);
  const int ntile = npad / 32;
  if (e == hipSuccess) {
    hipLaunchKernelGGL(mlp_fwd_kernel, dim3(fwd_blocks(ntile)), dim3(256), 0, 0, (const uint16_t*)dx, (const int*)nullptr,
                       n, ntile, (const u32x4*)M->frag, dyv);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(y, dyv, sizeof(float) * (size_t)n * NO, hipMemcpyDeviceToHost);
  if (e == hipSuccess && dy) {
    e = hipMemcpy(dyd, dy, sizeof(float) * (size_t)n * NO, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipFuncSetAttribute((const void*)mlp_bwd_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                                 (int)BWD_LDS);
    if (e == hipSuccess) {
      ActOut act{S.xt, S.h1t, S.dz1t, S.h2t, S.dz2t, S.dyt};
      const int nt16 = npad / 16;
      const int bb = std::max(1, std::min((nt16 + BWD_WAVES - 1) / BWD_WAVES, 1024));
      hipLaunchKernelGGL(mlp_bwd_kernel, dim3(bb), dim3(64 * BWD_WAVES), BWD_LDS, 0, (const uint16_t*)dx,
                         (const int*)nullptr, n, nt16, (const u32x4*)M->frag, (const float*)dyd, act);
      const int nks = npad / 32;
      const int nsplit = std::max(1, std::min(DW_SPLIT, (nks + 3) / 4));
      hipLaunchKernelGGL(dw_kernel, dim3(NTILES_DW, nsplit), dim3(256), 0, 0, act, nks, M->partial);
      hipLaunchKernelGGL(grad_reduce_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, 0, (const float*)M->partial,
                         nsplit, M->grad);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(grad, M->grad, sizeof(float) * NPARAM, hipMemcpyDeviceToHost);
  }
  cleanup();
  if (e != hipSuccess) { err = std::string("guess_mlp: ") + hipGetErrorString(e); return false; }
  return true;
}

bool guess_batch_features(Batch* b, std::vector<int>& src, std::vector<int>& dst, std::vector<uint16_t>& x,
                          std::vector<int>& caller, std::string& err) {
  BatchGuessView v;
  if (!batch_guess_view(b, v, err)) return false;
  caller = v.caller_atom;
  const long long n = v.ngp_total;
  src.assign(n, 0);
  dst.assign(n, 0);
  x.assign((size_t)n * NF, 0);
  if (n == 0) return true;
  PairGeom G{v.ax, v.ay, v.az, v.elem, v.gbeg, v.gcnt, v.gdst};
  const int fb = (int)std::min<long long>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(feature_kernel, dim3(fb), dim3(256), 0, 0, G, v.gsrc, n, v.X);
  HIP_OK(hipGetLastError());
  HIP_OK(hipDeviceSynchronize());
  HIP_OK(hipMemcpy(x.data(), v.X, sizeof(uint16_t) * x.size(), hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(src.data(), v.gsrc, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIP_OK(hipMemcpy(dst.data(), v.gdst, sizeof(int) * n, hipMemcpyDeviceToHost));
  return true;
}

}  // namespace pm3host
