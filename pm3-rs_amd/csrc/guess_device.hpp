// Learned initial density: pair features, the network's sizes and the equivariant assembly of
// density blocks from the network's outputs.  Shared by the SCF kernel (assembly), the feature
// and MLP kernels and the training kernels.
//
// A directed pair (a -> b) exists for every ordered pair of distinct atoms of a molecule closer
// than 4 Angstrom.  Its network input (NF bf16 values) is:
//   [0, 9)    one-hot of the ordered element pair 3 e(a) + e(b), e: H 0, C 1, O 2
//   [9, 41)   32 Gaussians exp(-((r - mu_k) / 0.125)^2) of the distance r (Angstrom),
//             mu_k = 0.5 + 3.5 k / 31
//   [41, 65)  environment of a: for every other neighbour c of a (c != b) with element e(c),
//             fc(r_ac) exp(-((r_ac - nu_m) / 0.5)^2) added to slot 41 + 8 e(c) + m,
//             nu_m = 1 + 3 m / 7, fc(r) = (cos(pi r / 4) + 1) / 2
//   [65, 89)  the same for b (neighbours c != a)
//   [89, 93)  angular sums  sum_c fc(r_ac) cos(t_bac), sum_c fc(r_ac) cos^2(t_bac), and the same
//             about b with cos(t_abc)
//   93, 94    zero;  95  one (carries the first layer's bias)
// Features are evaluated in double, rounded to float and then to bf16 (round to nearest even).
//
// Network: Z1 = W1^T x (W1: NF x NH, its row 95 is the first bias), H1 = silu(Z1) (bf16),
// Z2 = W2^T H1 + b2, H2 = silu(Z2) (bf16), y = W3^T H2 + b3 (NO outputs, 9 used).  Weights are
// bf16, accumulation and biases fp32.
//
// Assembly, with u the unit vector from a to b:
//   pair block  B(a->b)[s][s] = y0, [s][p_k] = y1 u_k, [p_k][s] = y2 u_k,
//               [p_k][p_l] = y3 u_k u_l + y4 (delta_kl - u_k u_l)
//   message     M(a->b) to a's on-site block: [s][s] += y5, [s][p_k] = [p_k][s] += y6 u_k,
//               [p_k][p_l] += y7 u_k u_l + y8 (delta_kl - u_k u_l)
//   P_aa = superposition block of a + sum_b M(a->b);  P_ab = (B(a->b) + B(b->a)^T) / 2 for pairs
//   inside the cutoff, zero otherwise.  The guess is rotation-equivariant by construction.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace pm3guess {

constexpr int NF = 96;   // input features
constexpr int NH = 128;  // hidden width
constexpr int NO = 16;   // outputs (9 used)
constexpr int NW1 = NF * NH, NW2 = NH * NH, NB2 = NH, NW3 = NH * NO, NB3 = NO;
constexpr int OFF_W1 = 0, OFF_W2 = OFF_W1 + NW1, OFF_B2 = OFF_W2 + NW2, OFF_W3 = OFF_B2 + NB2, OFF_B3 = OFF_W3 + NW3;
constexpr int NPARAM = OFF_B3 + NB3;  // 30864
constexpr double CUT_A = 4.0;
constexpr double BOHR_A = 0.529177210903;  // src/constants.rs:23
constexpr double PI = 3.14159265358979323846;

__host__ __device__ inline int elem_index(int z) { return z == 1 ? 0 : z == 6 ? 1 : z == 8 ? 2 : -1; }

__host__ __device__ inline uint16_t f32_to_bf16(float f) {
  const uint32_t u = __builtin_bit_cast(uint32_t, f);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
__host__ __device__ inline float bf16_to_f32(uint16_t h) { return __builtin_bit_cast(float, (uint32_t)h << 16); }

// Superposition occupations of the learned-guess elements (H, C, O): s and per-p populations.
__host__ __device__ inline double sad_s(int e) { return e == 0 ? 1.0 : 2.0; }
__host__ __device__ inline double sad_p(int e) { return e == 0 ? 0.0 : e == 1 ? 2.0 / 3.0 : 4.0 / 3.0; }

// Geometry of one batch as the feature code reads it: positions (Bohr) and element index per
// global atom; the directed pairs of every source atom are contiguous, [gbeg, gbeg + gcnt), and
// gdst holds their destination as a global atom index.
struct PairGeom {
  const double* x;
  const double* y;
  const double* z;
  const int8_t* elem;
  const int* gbeg;
  const int* gcnt;
  const int* gdst;
};

__device__ inline void add_env(double (&env)[24], double (&ang)[2], const PairGeom& G, int src, int skip,
                               double ux, double uy, double uz) {
  const double xs = G.x[src], ys = G.y[src], zs = G.z[src];
  const int b0 = G.gbeg[src], nb = G.gcnt[src];
  for (int t = 0; t < nb; ++t) {
    const int c = G.gdst[b0 + t];
    if (c == skip) continue;
    const double dx = G.x[c] - xs, dy = G.y[c] - ys, dz = G.z[c] - zs;
    const double rb = sqrt(dx * dx + dy * dy + dz * dz);
    const double r = rb * BOHR_A;
    const double fc = 0.5 * (cos(PI * r / CUT_A) + 1.0);
    const double cs = (dx * ux + dy * uy + dz * uz) / rb;
    ang[0] += fc * cs;
    ang[1] += fc * cs * cs;
    const int e = G.elem[c];
#pragma unroll
    for (int m = 0; m < 8; ++m) {
      const double t0 = (r - (1.0 + 3.0 * m / 7.0)) / 0.5;
      const double g = fc * exp(-t0 * t0);
      if (e == 0) env[m] += g;
      else if (e == 1) env[8 + m] += g;
      else env[16 + m] += g;
    }
  }
}

// Network input of the directed pair (a -> b), a and b global atom indices, as NF bf16 values.
__device__ inline void pair_features(const PairGeom& G, int a, int b, uint16_t (&out)[NF]) {
  const double dx = G.x[b] - G.x[a], dy = G.y[b] - G.y[a], dz = G.z[b] - G.z[a];
  const double rb = sqrt(dx * dx + dy * dy + dz * dz);
  const double ux = dx / rb, uy = dy / rb, uz = dz / rb;
  const double r = rb * BOHR_A;
  const int ta = 3 * G.elem[a] + G.elem[b];
#pragma unroll
  for (int k = 0; k < 9; ++k) out[k] = f32_to_bf16(k == ta ? 1.0f : 0.0f);
#pragma unroll
  for (int k = 0; k < 32; ++k) {
    const double t0 = (r - (0.5 + 3.5 * k / 31.0)) / 0.125;
    out[9 + k] = f32_to_bf16((float)exp(-t0 * t0));
  }
  double env[24], ang[2];
#pragma unroll
  for (int k = 0; k < 24; ++k) env[k] = 0.0;
  ang[0] = ang[1] = 0.0;
  add_env(env, ang, G, a, b, ux, uy, uz);
#pragma unroll
  for (int k = 0; k < 24; ++k) out[41 + k] = f32_to_bf16((float)env[k]);
  out[89] = f32_to_bf16((float)ang[0]);
  out[90] = f32_to_bf16((float)ang[1]);
#pragma unroll
  for (int k = 0; k < 24; ++k) env[k] = 0.0;
  ang[0] = ang[1] = 0.0;
  add_env(env, ang, G, b, a, -ux, -uy, -uz);
#pragma unroll
  for (int k = 0; k < 24; ++k) out[65 + k] = f32_to_bf16((float)env[k]);
  out[91] = f32_to_bf16((float)ang[0]);
  out[92] = f32_to_bf16((float)ang[1]);
  out[93] = 0;
  out[94] = 0;
  out[95] = f32_to_bf16(1.0f);
}

// Pair block B(a->b) (4 x 4, rows a's AOs s px py pz, columns b's) from outputs y0..y4.
__host__ __device__ inline void pair_block(const float* y, double ux, double uy, double uz, double (&B)[4][4]) {
  const double u[3] = {ux, uy, uz};
  B[0][0] = y[0];
  for (int k = 0; k < 3; ++k) {
    B[0][1 + k] = y[1] * u[k];
    B[1 + k][0] = y[2] * u[k];
    for (int l = 0; l < 3; ++l) {
      const double uu = u[k] * u[l];
      B[1 + k][1 + l] = y[3] * uu + y[4] * ((k == l ? 1.0 : 0.0) - uu);
    }
  }
}

// On-site message M(a->b) (4 x 4) from outputs y5..y8.
__host__ __device__ inline void onsite_message(const float* y, double ux, double uy, double uz, double (&M)[4][4]) {
  const double u[3] = {ux, uy, uz};
  M[0][0] = y[5];
  for (int k = 0; k < 3; ++k) {
    M[0][1 + k] = y[6] * u[k];
    M[1 + k][0] = y[6] * u[k];
    for (int l = 0; l < 3; ++l) {
      const double uu = u[k] * u[l];
      M[1 + k][1 + l] = y[7] * uu + y[8] * ((k == l ? 1.0 : 0.0) - uu);
    }
  }
}

}  // namespace pm3guess
