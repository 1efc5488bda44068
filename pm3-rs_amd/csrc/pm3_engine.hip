// Batched PM3 RHF SCF on CDNA4: one persistent launch, one workgroup per molecule.
//
// Workgroups pull molecule indices from an atomic queue (largest first) and run the whole
// single-point for that molecule without leaving the kernel: pair integrals and H_core, the
// superposition guess, then the SCF loop -- Fock build G(P), A-DIIS/CDIIS extrapolation,
// eigensolve, aufbau density -- until the molecule's own convergence test passes, followed by
// the stability search when its frontier gap is small.  Each molecule therefore carries its own
// convergence mask (the loop exit) and iteration count, and a finished workgroup immediately
// takes the next molecule.
//
// Reference path restated (file:line in the reference):
//   H_core and pair cache ....... src/hamiltonian.rs:191-344
//   Fock build .................. src/fock.rs:93-193 (two-centre), :244-332 (one-centre)
//   SCF loop .................... src/scf.rs:1179-1316 (rhf_loop)
//   DIIS history and solvers .... src/scf.rs:1065-1173, :1501-1693
//   stability search ............ src/scf.rs:599-715
//   guesses ..................... src/scf.rs:855-935
//
// Layout (per workgroup scratch, all doubles, n x n matrices row-major with leading dimension
// np = nao rounded up to 16, atoms sorted heavy-first inside each molecule so that the pair
// tables split into homogeneous XX / XH / HH blocks stored structure-of-arrays):
//   H F P Pn Fu T C Cn V Pb Cb | DIIS F[8] E[8] D[8] | W_xx[100][nxx] W_xh[10][nxh] W_hh[nhh] | vectors
//
// Eigensolver: cyclic two-sided Jacobi with the round-robin parallel ordering, the symmetric
// matrix packed in LDS, warm-started in the previous iteration's orbital basis
// (A = C^T F C is nearly diagonal once the SCF settles, so few sweeps are needed).  Dense
// products (C^T F C, C V, C_occ C_occ^T, F P) run on the FP64 matrix cores
// (v_mfma_f64_16x16x4_f64).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <memory>
#include <mutex>
#include <numeric>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "pm3_guess.hpp"
#include "pm3_host.hpp"

using namespace pm3dev;

namespace {

// Threads per workgroup of the SCF kernel.  The main build is 512 (one workgroup per CU); the
// same device code is also compiled for 256 and 128 (pm3_scf_nt256.hip, pm3_scf_nt128.hip: two
// and four workgroups per CU) for molecules small enough to share a CU.
#ifndef PM3_NT
#define PM3_NT 512
#endif
constexpr int NT = PM3_NT;
constexpr int NW = NT / 64;   // waves per workgroup
constexpr int MAXD = 8;       // maximum DIIS depth (src/scf.rs:309)
constexpr int RED_K = 3 * MAXD + 3;  // most values one block_sum reduces (the DIIS history pass)
constexpr int LDS_BYTES_MAX = 160 * 1024;
constexpr int STATIC_LDS_BYTES = 4096;  // upper bound on the kernel's static __shared__ use
constexpr int LARGE_MAX_NAO = 16384;    // largest molecule of the host-driven path (12 n^2 doubles)
// Early SCF iterations need eigenvectors only to a fraction of the current DIIS error: Jacobi
// stops at max(1e-14 |A|, EIG_TOL_FACTOR * |F P - P F|), which reaches full precision as the
// SCF converges, so the converged state is unchanged.
constexpr double EIG_TOL_FACTOR = 1.0e-2;
constexpr double EIG_REL = 1.0e-2;  // cap of the loose tolerance relative to the starting off-diagonal
constexpr double EIG_SKIP = 0.0;  // threshold Jacobi: skip rotations of |a_pq| <= EIG_SKIP x tolerance
constexpr double SP2_MARGIN = 0.05;  // margin of the Lanczos purification bounds, relative to the width
constexpr int PH_LAST = 12;   // phase timers: slots [0, 12) phases, [12] last stamp, [13, 16) counters
constexpr int PH_SLOTS = 16;

__constant__ TableConst c_tab;

struct MolDesc {
  int atom0, nat, nheavy, nao, nocc, np;
  int nbeta, uhf;              // beta electrons; uhf = 1: the unrestricted path (nocc is then n_alpha)
  int ngp, learn;              // directed pairs of the learned guess; learn = all atoms H, C or O
  long long out_mat, out_vec;  // offsets into the density/coefficient and vector outputs, -1 if none
  long long h_off, w_off;      // per-molecule H_core (np x np) and pair tables in the molecule store
  long long gp0;               // first directed pair
};

struct MolOut {
  double e_elec, e_core, last_dp, first_e, homo, lumo;
  double dip[3];
  int iters, converged, paths, status;
  int unrestricted;  // the returned state is unrestricted (spin density in out_S)
  int guess_used;  // 0 superposition (learned not requested or not applicable), 1 learned, 2 learned rejected
};

struct KArgs {
  const MolDesc* mols;
  const int* order;
  int nmol;
  int* queue;
  const int* aslot;
  const double* ax;
  const double* ay;
  const double* az;
  const int* aoff;
  double* scratch;
  long long stride;  // doubles per workgroup
  int np_max;
  long long w_cap;
  MolOut* out;
  double* out_P;
  double* out_C;
  double* out_eps;
  double* out_q;  // Mulliken charge per atom (sorted atom order)
  double* out_g;  // gradient per atom (sorted atom order, eV/Bohr), when requested
  double* out_S;     // spin density P_a - P_b per unrestricted molecule (out_P layout), or null
  double* out_Cb;    // beta orbitals (out_C layout), eigenvector route only, or null
  double* out_epsb;  // beta orbital energies (out_eps layout), or null
  long long ustride; // doubles of each workgroup's slab that hold the beta channel (0: none)
  double* store;  // per-molecule H_core and pair tables written by pair_kernel
  const long long* pair_base;  // first global pair of each queue position (nmol + 1 entries)
  long long npairs_total;
  unsigned long long* prof;    // optional per-phase wall-clock sums (10 ns ticks), may be null
  int lds_doubles;             // dynamic LDS per workgroup, in doubles
  int nmol_pairs;              // queue positions covered by pair_base: batched, then large molecules
  double eig_tol_factor;       // eigensolver stops at max(1e-14 |A|, min(factor * |[F,P]|, eig_rel * off0))
  double eig_rel;
  double eig_skip;
  int purify;                  // 1: densities by purification, frontier levels by Lanczos (no orbitals)
  int warm;                    // 1: each molecule's first SCF path starts from its density in out_P
  int sp2_lds;                 // purification operand placements allowed: 0 HBM only, 1 + LDS pair, 2 + LDS X,
                               // 3 + LDS X overwritten in place, 4 + leading rows of X in LDS, 5 + the packed
                               // tiled image in LDS, 6 + that image split between LDS and HBM
  int sp2_tight_n;             // smallest molecule (AOs) whose purification uses Lanczos spectral bounds
  int sp2_tile16_n;            // purification products on 16 x 16 wave tiles up to this size (AOs), 32 x 32 above
  double sp2_margin;           // margin of those bounds relative to the estimated width (fresh estimate)
  // learned guess: directed pairs (global indices, see batch_create) and network outputs
  const int8_t* elem8;
  const int* gbeg;
  const int* gcnt;
  const int* gsrc;
  const int* gdst;
  const int* gpart;
  const float* gY;
  // options
  int max_scf;
  double e_tol, p_tol;
  int accel;
  double adiis_switch;
  int guess, stability, scf_memory_mb;
  int use_diis;           // Pm3Options::use_diis (the unrestricted rung's CDIIS needs it apart from accel)
  double shift, damp;     // Pm3Options::level_shift_ev, ::damping: non-zero steers the first path
  int has_field;          // a uniform external field f (eV / Bohr / e) in H_core, core energy and gradient
  double field[3];
};

typedef double double4_t __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------------------------------------
// Workgroup reductions.
// DPP move of a double: both halves through the same data-parallel-primitive lane pattern.
template <int CTRL> __device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_value(double v, int j);
// Sum over the 64 lanes, the same value on every lane (full waves only): four DPP steps inside
// each 16-lane row (quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror, row_mirror), then
// the four row sums read by v_readlane.  VALU latency only, where a butterfly of __shfl_xor would
// chain six round trips through the LDS crossbar.
__device__ inline double wave_sum(double v) {
  v += dpp_f64<0xB1>(v);
  v += dpp_f64<0x4E>(v);
  v += dpp_f64<0x141>(v);
  v += dpp_f64<0x140>(v);
  return (lane_value(v, 0) + lane_value(v, 16)) + (lane_value(v, 32) + lane_value(v, 48));
}
// Lane j's value of v, given to every lane of the wave (j wave-uniform): v_readlane, a register
// read, where __shfl would go through the LDS crossbar (ds_bpermute).
__device__ __forceinline__ double lane_value(double v, int j) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
  return __hiloint2double(hi, lo);
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// red holds NW * K doubles (the SCF kernels declare NW * RED_K).
template <int K> __device__ inline void block_sum(double (&v)[K], double* red) {
  static_assert(K <= RED_K, "reduction buffer holds RED_K values per wave");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  __syncthreads();
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) red[wave * K + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int w = 0; w < NW; ++w) s += red[w * K + k];
    v[k] = s;
  }
}

// block_sum for callers that have passed a workgroup barrier since every wave last read red (so
// red may be rewritten at once): one barrier instead of two.
template <int K> __device__ inline void block_sum_fenced(double (&v)[K], double* red) {
  static_assert(K <= RED_K, "reduction buffer holds RED_K values per wave");
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < K; ++k) red[wave * K + k] = v[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double s = 0.0;
    for (int w = 0; w < NW; ++w) s += red[w * K + k];
    v[k] = s;
  }
}

__device__ inline double block_max(double v, double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  v = wave_max(v);
  __syncthreads();
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  for (int w = 0; w < NW; ++w) s = fmax(s, red[w]);
  return s;
}

// Row-major upper-triangle pair index p -> (a, b), a < b < nat.
__device__ inline void tri_decode(int p, int nat, int& a, int& b) {
  const double tn = 2.0 * nat - 1.0;
  int aa = (int)floor((tn - sqrt(fmax(tn * tn - 8.0 * p, 0.0))) * 0.5);
  aa = max(0, min(aa, nat - 2));
  while (aa > 0 && aa * (2 * nat - aa - 1) / 2 > p) --aa;
  while (aa < nat - 2 && (aa + 1) * (2 * nat - aa - 2) / 2 <= p) ++aa;
  a = aa;
  b = p - aa * (2 * nat - aa - 1) / 2 + aa + 1;
}
__device__ inline int tri_index(int a, int b, int nat) { return a * (2 * nat - a - 1) / 2 + (b - a - 1); }

// Packed upper-triangle index of (i, j) for the LDS copy of a symmetric matrix.
__device__ inline int sym_idx(int i, int j) { return i <= j ? j * (j + 1) / 2 + i : i * (i + 1) / 2 + j; }

// Packed triangle of an n x n matrix, diagonal first: e < n is (e, e); then the strict upper
// triangle column by column, e = n + j (j - 1) / 2 + i for i < j.
__device__ inline void tri_unpack(int e, int n, int& i, int& j) {
  if (e < n) {
    i = j = e;
    return;
  }
  const int f = e - n;
  int c = (int)((1.0f + sqrtf(1.0f + 8.0f * (float)f)) * 0.5f);
  while (c * (c - 1) / 2 > f) --c;
  while ((c + 1) * c / 2 <= f) ++c;
  j = c;
  i = f - c * (c - 1) / 2;
}

// ------------------------------------------------------------------------------------------
// Dense product on the FP64 matrix cores: out(i, j) = sum_{k<K} A(i, k) B(k, j) for i < M, j < N.
// Each wave owns 16 x 16 output tiles; per step the 64 lanes feed one 16x4 slice of A (lane l holds
// A[l % 16][l / 16]) and one 4x16 slice of B (lane l holds B[l / 16][l % 16]) into
// v_mfma_f64_16x16x4_f64.  Loaders return 0 outside the logical shape.
// C = A B on the FP64 matrix cores by one workgroup: each wave owns 32 x 32 output blocks (four
// 16 x 16 accumulators, so consecutive MFMAs are independent) and loads 16 k-values ahead.
// SYM: the product is known to be symmetric (M == N); only blocks on or above the diagonal are
// computed and each value is stored at (i, j) and (j, i).
template <bool SYM = false, class LA, class LB, class ST>
__device__ inline void wg_gemm(int M, int N, int K, LA la, LB lb, ST st) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int bm = (M + 31) >> 5, bn = (N + 31) >> 5;
  const int li = lane & 15, lk = lane >> 4;
  const int nblk = SYM ? bm * (bm + 1) / 2 : bm * bn;
  for (int t = wave; t < nblk; t += NW) {
    int bi, bj;
    if (SYM) {  // upper-triangle block t -> (bi, bj), bi <= bj, row by row
      bi = 0;
      int rest = t;
      while (rest >= bm - bi) { rest -= bm - bi; ++bi; }
      bj = bi + rest;
    } else {
      bi = t / bn;
      bj = t % bn;
    }
    const int i0 = bi << 5, j0 = bj << 5;
    const int ia = i0 + li, ib = ia + 16, ja = j0 + li, jb = ja + 16;
    const bool va = ia < M, vb = ib < M, wa = ja < N, wb = jb < N;
    double4_t c00 = {0.0, 0.0, 0.0, 0.0}, c01 = c00, c10 = c00, c11 = c00;
    for (int k0 = 0; k0 < K; k0 += 16) {
      double a0[4], a1[4], b0[4], b1[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = k0 + 4 * u + lk;
        const bool kk = k < K;
        a0[u] = (kk && va) ? la(ia, k) : 0.0;
        a1[u] = (kk && vb) ? la(ib, k) : 0.0;
        b0[u] = (kk && wa) ? lb(k, ja) : 0.0;
        b1[u] = (kk && wb) ? lb(k, jb) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], b0[u], c00, 0, 0, 0);
        c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], b1[u], c01, 0, 0, 0);
        c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], b0[u], c10, 0, 0, 0);
        c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], b1[u], c11, 0, 0, 0);
      }
    }
    // result layout: item r of lane l holds row 4 r + l / 16, column l % 16 of its tile
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int ra = i0 + 4 * r + lk, rb = ra + 16;
      if (ra < M && ja < N && (!SYM || bi < bj || ra <= ja)) { st(ra, ja, c00[r]); if (SYM && ra != ja) st(ja, ra, c00[r]); }
      if (ra < M && jb < N && (!SYM || bi < bj || ra <= jb)) { st(ra, jb, c01[r]); if (SYM && ra != jb) st(jb, ra, c01[r]); }
      if (rb < M && ja < N && (!SYM || bi < bj || rb <= ja)) { st(rb, ja, c10[r]); if (SYM && rb != ja) st(ja, rb, c10[r]); }
      if (rb < M && jb < N && (!SYM || bi < bj || rb <= jb)) { st(rb, jb, c11[r]); if (SYM && rb != jb) st(jb, rb, c11[r]); }
    }
  }
}

// One wave's output tile of wg_gemm_pad at (i0, j0): T = 32 fills c[0..3] (c00, c01, c10, c11),
// T = 16 fills c[0].
template <int T>
__device__ inline void pad_tile(int np, int kp, const double* S, const double* Bm, int i0, int j0, int li, int lk,
                                double4_t (&c)[T == 32 ? 4 : 1]) {
  if constexpr (T == 16) {
    const int ia = min(i0 + li, np - 1), ja = min(j0 + li, np - 1);
    c[0] = double4_t{0.0, 0.0, 0.0, 0.0};
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t r = (size_t)(4 * u + lk) * np;
      a[u] = S[r + ia];
      b[u] = Bm[r + ja];
    }
    for (int k0 = 0; k0 < kp; k0 += 16) {
      const int kn = k0 + 16 < kp ? k0 + 16 : k0;  // after the last chunk: a harmless reload
      double na[4], nb[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const size_t r = (size_t)(kn + 4 * u + lk) * np;
        na[u] = S[r + ia];
        nb[u] = Bm[r + ja];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) c[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], c[0], 0, 0, 0);
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a[u] = na[u];
        b[u] = nb[u];
      }
    }
  } else {
    const int ia = min(i0 + li, np - 1), ib = min(i0 + li + 16, np - 1);
    const int ja = min(j0 + li, np - 1), jb = min(j0 + li + 16, np - 1);
    double4_t c00 = {0.0, 0.0, 0.0, 0.0}, c01 = c00, c10 = c00, c11 = c00;
    double a0[4], a1[4], b0[4], b1[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t r = (size_t)(4 * u + lk) * np;
      a0[u] = S[r + ia];
      a1[u] = S[r + ib];
      b0[u] = Bm[r + ja];
      b1[u] = Bm[r + jb];
    }
    for (int k0 = 0; k0 < kp; k0 += 16) {
      const int kn = k0 + 16 < kp ? k0 + 16 : k0;  // after the last chunk: a harmless reload
      double na0[4], na1[4], nb0[4], nb1[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const size_t r = (size_t)(kn + 4 * u + lk) * np;
        na0[u] = S[r + ia];
        na1[u] = S[r + ib];
        nb0[u] = Bm[r + ja];
        nb1[u] = Bm[r + jb];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        c00 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], b0[u], c00, 0, 0, 0);
        c01 = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[u], b1[u], c01, 0, 0, 0);
        c10 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], b0[u], c10, 0, 0, 0);
        c11 = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[u], b1[u], c11, 0, 0, 0);
      }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        a0[u] = na0[u];
        a1[u] = na1[u];
        b0[u] = nb0[u];
        b1[u] = nb1[u];
      }
    }
    c[0] = c00;
    c[1] = c01;
    c[2] = c10;
    c[3] = c11;
  }
}

// Block (bi, bj) of tile t: row-major over the bm x bm grid, or over its upper triangle (SYM).
template <bool SYM>
__device__ inline void pad_block(int t, int bm, int& bi, int& bj) {
  if (SYM) {
    bi = 0;
    int rest = t;
    while (rest >= bm - bi) { rest -= bm - bi; ++bi; }
    bj = bi + rest;
  } else {
    bi = t / bm;
    bj = t % bm;
  }
}

// Stores one tile's results: item r of lane l holds row 4 r + l / 16, column l % 16 of each 16 x 16
// quarter; SYM stores the upper part of diagonal blocks and mirrors every element.
template <bool SYM, int T, class ST>
__device__ inline void pad_store(int n, int bi, int bj, int li, int lk, const double4_t (&c)[T == 32 ? 4 : 1], ST& st) {
  const int i0 = bi * T, j0 = bj * T;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
#pragma unroll
    for (int q = 0; q < (T == 32 ? 4 : 1); ++q) {
      const int di = (q >> 1) * 16, dj = (q & 1) * 16;
      const int ra = i0 + di + 4 * r + lk, ca = j0 + dj + li;
      if (ra < n && ca < n && (!SYM || bi < bj || ra <= ca)) {
        st(ra, ca, c[q][r]);
        if (SYM && ra != ca) st(ca, ra, c[q][r]);
      }
    }
  }
}

// C = S B for n x n operands held np-strided with zero padding up to np rows and columns, S
// symmetric (read as S[k][i], so that both operands are read along rows: 16 consecutive doubles
// per lane group).  No bounds tests on the loads (padding supplies the zeros; rows beyond np are
// clamped and their results dropped), and the next 16-wide k chunk is loaded while the MFMAs of
// the current one run.  Same arithmetic order as wg_gemm, so the results are bitwise equal.
// T is the output tile of one wave: 32 x 32 (four 16 x 16 MFMA chains sharing their loads) or
// 16 x 16 (one chain; four times the tiles, so small n keeps more of the eight waves busy).  Each
// element is the same chain of MFMAs in the same k order either way, so both give the same bits.
template <bool SYM = false, int T = 32, class ST>
__device__ inline void wg_gemm_pad(int n, int np, const double* S, const double* Bm, ST st) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int bm = (n + T - 1) / T;
  const int nblk = SYM ? bm * (bm + 1) / 2 : bm * bm;
  const int kp = (n + 15) & ~15;  // <= np
  for (int t = wave; t < nblk; t += NW) {
    int bi, bj;
    pad_block<SYM>(t, bm, bi, bj);
    double4_t c[T == 32 ? 4 : 1];
    pad_tile<T>(np, kp, S, Bm, bi * T, bj * T, li, lk, c);
    pad_store<SYM, T>(n, bi, bj, li, lk, c, st);
  }
}

// Tiles a wave holds at most in wg_gemm_sym_hold, and whether a size fits that budget.
constexpr int HOLD_T16 = 5, HOLD_T32 = 2;
__host__ __device__ constexpr bool hold_fits(int n, int T) {
  return ((n + T - 1) / T) * ((n + T - 1) / T + 1) / 2 <= (T == 16 ? HOLD_T16 : HOLD_T32) * NW;
}

// The symmetric product X X of wg_gemm_pad<true, T>, with every wave's tiles held in registers
// until all waves have finished reading X, so that the store may overwrite X itself (one LDS image
// instead of two).  Each element and its mirror belong to one wave, so the stores do not race.
// Requires hold_fits(n, T).  Same MFMA chains, so the same bits as wg_gemm_pad.
template <int T, class ST>
__device__ inline void wg_gemm_sym_hold(int n, int np, const double* X, ST st) {
  constexpr int H = T == 16 ? HOLD_T16 : HOLD_T32;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int bm = (n + T - 1) / T;
  const int nblk = bm * (bm + 1) / 2;
  const int kp = (n + 15) & ~15;
  double4_t c[H][T == 32 ? 4 : 1];
#pragma unroll
  for (int h = 0; h < H; ++h) {
    const int t = wave + h * NW;
    if (t < nblk) {
      int bi, bj;
      pad_block<true>(t, bm, bi, bj);
      pad_tile<T>(np, kp, X, X, bi * T, bj * T, li, lk, c[h]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int h = 0; h < H; ++h) {
    const int t = wave + h * NW;
    if (t < nblk) {
      int bi, bj;
      pad_block<true>(t, bm, bi, bj);
      pad_store<true, T>(n, bi, bj, li, lk, c[h], st);
    }
  }
}

// pad_tile with the k rows [0, kL) of the operand read from an LDS copy (Sl, leading dimension
// ldl) and the rows from kL on from HBM (Sh, ldh); kL is a multiple of 16, so each 16-row chunk
// has one source.  The same MFMA chains, so the same bits as pad_tile on either copy.
template <int T>
__device__ inline void pad_tile_split(int kL, const double* Sl, int ldl, const double* Sh, int ldh, int kp, int i0,
                                      int j0, int li, int lk, double4_t (&c)[T == 32 ? 4 : 1]) {
  constexpr int NC = T == 32 ? 2 : 1;  // 16-wide column groups per operand
  int il[NC], ih[NC], jl[NC], jh[NC];
#pragma unroll
  for (int g = 0; g < NC; ++g) {
    il[g] = min(i0 + li + 16 * g, ldl - 1);
    ih[g] = min(i0 + li + 16 * g, ldh - 1);
    jl[g] = min(j0 + li + 16 * g, ldl - 1);
    jh[g] = min(j0 + li + 16 * g, ldh - 1);
  }
  double4_t acc[NC][NC];
#pragma unroll
  for (int x = 0; x < NC; ++x)
#pragma unroll
    for (int y = 0; y < NC; ++y) acc[x][y] = double4_t{0.0, 0.0, 0.0, 0.0};
  double a[NC][4], b[NC][4];
  auto load = [&](int k0, double (&ra)[NC][4], double (&rb)[NC][4]) {
    const bool in_lds = k0 < kL;
    const double* S = in_lds ? Sl : Sh;
    const int ld = in_lds ? ldl : ldh;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const size_t r = (size_t)(k0 + 4 * u + lk) * ld;
#pragma unroll
      for (int g = 0; g < NC; ++g) {
        ra[g][u] = S[r + (in_lds ? il[g] : ih[g])];
        rb[g][u] = S[r + (in_lds ? jl[g] : jh[g])];
      }
    }
  };
  load(0, a, b);
  for (int k0 = 0; k0 < kp; k0 += 16) {
    const int kn = k0 + 16 < kp ? k0 + 16 : k0;  // after the last chunk: a harmless reload
    double na[NC][4], nb[NC][4];
    load(kn, na, nb);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int x = 0; x < NC; ++x)
#pragma unroll
        for (int y = 0; y < NC; ++y) acc[x][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[x][u], b[y][u], acc[x][y], 0, 0, 0);
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int g = 0; g < NC; ++g) {
        a[g][u] = na[g][u];
        b[g][u] = nb[g][u];
      }
  }
#pragma unroll
  for (int x = 0; x < NC; ++x)
#pragma unroll
    for (int y = 0; y < NC; ++y) c[x * NC + y] = acc[x][y];
}

// The symmetric product X X of wg_gemm_pad<true, T> with X's first kL rows read from LDS.
template <int T, class ST>
__device__ inline void wg_gemm_sym_split(int n, int kL, const double* Xl, int ldl, const double* Xh, int ldh, ST st) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int bm = (n + T - 1) / T;
  const int nblk = bm * (bm + 1) / 2;
  const int kp = (n + 15) & ~15;
  for (int t = wave; t < nblk; t += NW) {
    int bi, bj;
    pad_block<true>(t, bm, bi, bj);
    double4_t c[T == 32 ? 4 : 1];
    pad_tile_split<T>(kL, Xl, ldl, Xh, ldh, kp, bi * T, bj * T, li, lk, c);
    pad_store<true, T>(n, bi, bj, li, lk, c, st);
  }
}

// ------------------------------------------------------------------------------------------
// One-centre two-electron integral (ab|cd) on one atom (src/fock.rs:244-282).
__device__ inline double oc_int(int a, int b, int c, int d, const ElemConst& e) {
  if (a == b && c == d) {
    if (a == 0 && c == 0) return e.gss;
    if (a == 0 || c == 0) return e.gsp;
    return a == c ? e.gpp : e.gp2;
  }
  const int ba = min(a, b), bb = max(a, b), kc = min(c, d), kd = max(c, d);
  if (ba == 0 && bb != 0 && kc == 0 && kd != 0 && bb == kd) return e.hsp;
  if (ba != 0 && bb != 0 && ba != bb && ba == kc && bb == kd) return 0.5 * (e.gpp - e.gp2);
  return 0.0;
}

struct Mol {
  int atom0, nat, nh, nH, n, np, nocc;
  int nbeta;       // beta electrons (= nocc for a restricted molecule)
  double wt, xs;   // occupation of an aufbau orbital (2 restricted, 1 per spin) and its inverse
  int ngp, learn;
  long long gp0;
  int nxx, nxh, nhh;
  double* Wxx;
  double* Wxh;
  double* Whh;
  double* H;     // H_core of this molecule (np x np, molecule store)
  double* S1;    // [10][npairs] attraction of a's AOs to b's core, per pair (molecule store)
  double* S2;    // [10][npairs] attraction of b's AOs to a's core
  double* ecc;   // [npairs] core-core energy per pair
  double* slot;  // per-pair Coulomb partial sums of the Fock build (workgroup scratch)
  int lds_doubles;  // dynamic LDS of the launch, in doubles
  double eig_rel;   // loose-tolerance cap relative to the starting off-diagonal
  double eig_skip;  // rotations of elements below eig_skip x tolerance are skipped
  int sp2_lds;      // purification operand placements allowed (KArgs::sp2_lds)
  int sp2_tight_n;  // KArgs::sp2_tight_n
  int sp2_tile16_n;  // KArgs::sp2_tile16_n
  double sp2_margin;  // KArgs::sp2_margin
  const double* prev;  // this molecule's converged density of the batch's last solve (n x n, engine order), or null
};

// ------------------------------------------------------------------------------------------
// Pair integrals, core-core energy and H_core (src/hamiltonian.rs:191-344, src/repulsion.rs:13-27).
// One work-item per atom pair: W (packed two-electron table), the two electron-core attraction
// blocks and the core-core energy go to per-pair slots of the molecule store; the resonance block
// 1/2 (beta_mu + beta_nu) S_mu,nu is written straight into H_core.
__device__ inline void pair_integrals(const KArgs& A, const Mol& M, int p) {
  const int np = M.np, nat = M.nat;
  const long long npairs = (long long)nat * (nat - 1) / 2;
  int a, b;
  tri_decode(p, nat, a, b);
  const int ga = M.atom0 + a, gb = M.atom0 + b;
  const int sa = A.aslot[ga], sb = A.aslot[gb];
  const ElemConst& ea = c_tab.el[sa];
  const ElemConst& eb = c_tab.el[sb];
  const double dx = A.ax[gb] - A.ax[ga], dy = A.ay[gb] - A.ay[ga], dz = A.az[gb] - A.az[ga];
  const double r = sqrt(dx * dx + dy * dy + dz * dz);
  M.ecc[p] = core_core_pair<double>(c_tab, sa, sb, r);
  const double ux = dx / r, uy = dy / r, uz = dz / r;
  // Two-electron integrals and electron-core attraction (src/integrals.rs:167-230).
  const double ddx = ux * r, ddy = uy * r, ddz = uz * r;
  const double rr = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
  const double inv = 1.0 / rr;
  double R[3][3];
  rot_to_x(-(ddx * inv), -(ddy * inv), -(ddz * inv), R);
  double e1b[10], e2a[10];
  ecore(ea, eb, ddx, ddy, ddz, e1b);
  ecore(eb, ea, -ddx, -ddy, -ddz, e2a);
  const int na = ea.norb, nb = eb.norb;
  for (int k = 0; k < (na == 4 ? 10 : 1); ++k) M.S1[k * npairs + p] = e1b[k];
  for (int k = 0; k < (nb == 4 ? 10 : 1); ++k) M.S2[k * npairs + p] = e2a[k];
  if (na == 4 && nb == 4) {
    double ri[22];
    local_xx(ea, eb, rr, ri);
    const int kx = tri_index(a, b, M.nh);
    double* W = M.Wxx;
    const int nxx = M.nxx;
    assemble_xx(ri, R, [&](int idx, double v) { W[(long long)idx * nxx + kx] = v; });
  } else if (na == 4) {
    double ri[4], w[10];
    local_xh(ea.dd, ea.qq, ea.rho0, ea.rho1, ea.rho2, eb.rho0, rr, ri);
    assemble_xh(ri, R, w);
    const int kx = a * M.nH + (b - M.nh);
#pragma unroll
    for (int k = 0; k < 10; ++k) M.Wxh[(long long)k * M.nxh + kx] = w[k];
  } else {
    const double s = ea.rho0 + eb.rho0;
    M.Whh[tri_index(a - M.nh, b - M.nh, M.nH)] = rsq(rr * rr + s * s) * EV;
  }
  // Resonance 1/2 (beta_mu + beta_nu) S_mu,nu (src/hamiltonian.rs:322-331).
  double S[4][4];
  overlap_block(ea, eb, ux, uy, uz, r, S);
  const int oa = A.aoff[ga], ob = A.aoff[gb];
  for (int i = 0; i < na; ++i) {
    const double bi = i == 0 ? ea.betas : ea.betap;
    for (int j = 0; j < nb; ++j) {
      const double bj = j == 0 ? eb.betas : eb.betap;
      const double v = 0.5 * (bi + bj) * S[i][j];
      M.H[(oa + i) * np + ob + j] = v;
      M.H[(ob + j) * np + oa + i] = v;
    }
  }
}

// The external field's one-electron operator +M.f on atom a's block (field_terms, src/dipole.rs:
// 156-181, dipole_matrix :93-125 at the coordinate origin): f.R_a on every diagonal element, f_x dd
// on the s-p_x pair of elements; it enters H_core after U and before the pair terms
// (src/hamiltonian.rs:224-233).
__device__ inline double field_term(const KArgs& A, const ElemConst& e, int ga, int mu, int nu) {
  if (mu == nu) return A.field[0] * A.ax[ga] + A.field[1] * A.ay[ga] + A.field[2] * A.az[ga];
  if (mu == 0) return A.field[nu - 1] * e.dd;
  if (nu == 0) return A.field[mu - 1] * e.dd;
  return 0.0;
}

// The field's nuclear term -sum_a Z_a R_a.f (src/dipole.rs:176-179), summed by work-items t0, t0 + st.
__device__ inline double field_nuclear(const KArgs& A, const Mol& M, int t0, int st) {
  double v = 0.0;
  if (!A.has_field) return v;
  for (int a = t0; a < M.nat; a += st) {
    const int ga = M.atom0 + a;
    v -= c_tab.el[A.aslot[ga]].tore * (A.field[0] * A.ax[ga] + A.field[1] * A.ay[ga] + A.field[2] * A.az[ga]);
  }
  return v;
}

// H_core diagonal blocks (U plus the attraction to every other core, summed in atom order) and the
// total core-core energy, from the per-pair slots.  One workgroup per molecule.
// Diagonal atom blocks of H_core: U plus the electron-core attractions summed over partners in a
// fixed order (src/hamiltonian.rs:310-320).  Work-items t0, t0 + st, ...
__device__ inline void core_diag(const KArgs& A, const Mol& M, int t0, int st) {
  const int np = M.np, nat = M.nat;
  const long long npairs = (long long)nat * (nat - 1) / 2;
  for (int e = t0; e < nat * 16; e += st) {
    const int a = e >> 4, mu = (e >> 2) & 3, nu = e & 3;
    const int ga = M.atom0 + a;
    const ElemConst& ea = c_tab.el[A.aslot[ga]];
    if (mu < ea.norb && nu < ea.norb) {
      const int oa = A.aoff[ga];
      const int q = pk(mu, nu);
      double s = 0.0;
      for (int b = 0; b < nat; ++b) {
        if (b == a) continue;
        s += b > a ? M.S1[q * npairs + tri_index(a, b, nat)] : M.S2[q * npairs + tri_index(b, a, nat)];
      }
      double u = mu != nu ? 0.0 : (mu == 0 ? ea.uss : ea.upp);
      if (A.has_field) u += field_term(A, ea, ga, mu, nu);
      M.H[(oa + mu) * np + oa + nu] = u + s;
    }
  }
}

__device__ double finish_core(const KArgs& A, const Mol& M, double* red) {
  const int tid = threadIdx.x;
  const long long npairs = (long long)M.nat * (M.nat - 1) / 2;
  core_diag(A, M, tid, NT);
  double v[1] = {field_nuclear(A, M, tid, NT)};
  for (long long p = tid; p < npairs; p += NT) v[0] += M.ecc[p];
  block_sum<1>(v, red);
  __syncthreads();
  return v[0];
}

// ------------------------------------------------------------------------------------------
// F = H + G(P), the RHF Fock build (src/fock.rs:285-332, :488-500), in three phases: off-diagonal
// atom blocks owned by their pair (exchange K) with per-pair Coulomb partials; per-atom Coulomb
// sums J; then each diagonal block adds its one-centre term and J.  The phases take a work-item
// range so that one workgroup (batched molecules, J in LDS) or a whole grid (large molecules, one
// launch per phase) runs the same code.
// Phase 1 of the Fock build: off-diagonal atom blocks (exchange K, owned by their pair) and the
// per-pair Coulomb partial sums in M.slot.  Work-items t0, t0 + st, ... (a workgroup or a grid).
// SW: every two-centre term of a pair is weighted by its switch (the CPHF distance cutoff of
// response_fock_rhf, src/fock.rs:106-159 with sw): sw[kx] for heavy-heavy pairs, sw[nxx + kx]
// heavy-hydrogen, sw[nxx + nxh + kx] hydrogen-hydrogen, applied where the reference applies it.
template <bool SW = false>
__device__ inline void fock_pairs(const KArgs& A, const Mol& M, const double* H, const double* P, const double* X,
                                  double xf, double* F, int t0, int st, const double* sw = nullptr) {
  const int np = M.np;
  double* JX = M.slot;                          // [20][nxx]: J_a (0..9), J_b (10..19) per XX pair
  double* JH = M.slot + 20LL * M.nxx;           // [11][nxh]: J_a (0..9), J_b (10) per XH pair
  // heavy-heavy pairs
  for (int kx = t0; kx < M.nxx; kx += st) {
    int a, b;
    tri_decode(kx, M.nh, a, b);
    const int oa = A.aoff[M.atom0 + a], ob = A.aoff[M.atom0 + b];
    double da[10], db[10], pab[4][4];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
#pragma unroll
      for (int nu = 0; nu <= mu; ++nu) {
        const double f = mu == nu ? 1.0 : 2.0;
        da[pk(mu, nu)] = f * P[(oa + mu) * np + oa + nu];
        db[pk(mu, nu)] = f * P[(ob + mu) * np + ob + nu];
      }
#pragma unroll
      for (int la = 0; la < 4; ++la) pab[mu][la] = X[(oa + mu) * np + ob + la];
    }
    double ja[10], jb[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, K[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) K[i][j] = 0.0;
    const double* W = M.Wxx + kx;
    const long long ld = M.nxx;
    const double swp = SW ? sw[kx] : 1.0;
    const double xfp = SW ? swp * xf : xf;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
      for (int nu = 0; nu <= mu; ++nu) {
        const int p = pk(mu, nu);
        double w[10];
#pragma unroll
        for (int q = 0; q < 10; ++q) w[q] = W[(long long)(p * 10 + q) * ld];
        double acc = 0.0;
        const double dap = SW ? swp * da[p] : da[p];
#pragma unroll
        for (int q = 0; q < 10; ++q) {
          acc += db[q] * w[q];
          jb[q] += dap * w[q];
        }
        ja[p] = SW ? swp * acc : acc;
        // exchange: K[mu][la] -= xf sum_si X(nu, si) (mu nu|la si), and the (nu, mu) term
#pragma unroll
        for (int la = 0; la < 4; ++la) {
          double s1 = 0.0, s2 = 0.0;
#pragma unroll
          for (int si = 0; si < 4; ++si) {
            s1 += pab[nu][si] * w[pk(la, si)];
            s2 += pab[mu][si] * w[pk(la, si)];
          }
          K[mu][la] -= xfp * s1;
          if (mu != nu) K[nu][la] -= xfp * s2;
        }
      }
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
      for (int la = 0; la < 4; ++la) {
        const double v = H[(oa + mu) * np + ob + la] + K[mu][la];
        F[(oa + mu) * np + ob + la] = v;
        F[(ob + la) * np + oa + mu] = v;
      }
#pragma unroll
    for (int q = 0; q < 10; ++q) {
      JX[(long long)q * M.nxx + kx] = ja[q];
      JX[(long long)(10 + q) * M.nxx + kx] = jb[q];
    }
  }
  // heavy-hydrogen pairs
  for (int kx = t0; kx < M.nxh; kx += st) {
    const int a = kx / M.nH, b = M.nh + kx % M.nH;
    const int oa = A.aoff[M.atom0 + a], ob = A.aoff[M.atom0 + b];
    const double pbb = P[ob * np + ob];
    const double swp = SW ? sw[M.nxx + kx] : 1.0;
    double jbv = 0.0, pa[4];
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) pa[nu] = X[(oa + nu) * np + ob];
    double w[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) w[q] = M.Wxh[(long long)q * M.nxh + kx];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
      for (int nu = 0; nu <= mu; ++nu) {
        const double f = mu == nu ? 1.0 : 2.0;
        if (SW) jbv += swp * (f * P[(oa + mu) * np + oa + nu]) * w[pk(mu, nu)];
        else jbv += f * P[(oa + mu) * np + oa + nu] * w[pk(mu, nu)];
      }
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      double k = 0.0;
#pragma unroll
      for (int nu = 0; nu < 4; ++nu) k += pa[nu] * w[pk(mu, nu)];
      const double v = H[(oa + mu) * np + ob] - (SW ? swp * xf : xf) * k;
      F[(oa + mu) * np + ob] = v;
      F[ob * np + oa + mu] = v;
    }
#pragma unroll
    for (int q = 0; q < 10; ++q) JH[(long long)q * M.nxh + kx] = SW ? swp * (w[q] * pbb) : w[q] * pbb;
    JH[10LL * M.nxh + kx] = jbv;
  }
  // hydrogen-hydrogen pairs
  for (int kx = t0; kx < M.nhh; kx += st) {
    int a, b;
    tri_decode(kx, M.nH, a, b);
    a += M.nh;
    b += M.nh;
    const int oa = A.aoff[M.atom0 + a], ob = A.aoff[M.atom0 + b];
    const double w = M.Whh[kx];
    const double v = H[oa * np + ob] - (SW ? sw[M.nxx + M.nxh + kx] * xf : xf) * X[oa * np + ob] * w;
    F[oa * np + ob] = v;
    F[ob * np + oa] = v;
  }
}

// Phase 2: Coulomb sums per atom (J[nat][10]), partners in a fixed order.
__device__ inline void fock_jsum(const KArgs& A, const Mol& M, const double* P, double* J, int t0, int st) {
  const int np = M.np;
  const double* JX = M.slot;
  const double* JH = M.slot + 20LL * M.nxx;
  for (int e = t0; e < M.nh * 10; e += st) {
    const int a = e / 10, q = e % 10;
    double s = 0.0;
#pragma unroll 8
    for (int b = 0; b < M.nh; ++b) {
      const double x = b > a ? JX[(long long)q * M.nxx + tri_index(a, b, M.nh)]
                             : b < a ? JX[(long long)(10 + q) * M.nxx + tri_index(b, a, M.nh)] : 0.0;
      if (b != a) s += x;
    }
    const double* jh = JH + (long long)q * M.nxh + a * M.nH;
#pragma unroll 8
    for (int h = 0; h < M.nH; ++h) s += jh[h];
    J[a * 10 + q] = s;
  }
  for (int h = t0; h < M.nH; h += st) {
    double s = 0.0;
#pragma unroll 8
    for (int a = 0; a < M.nh; ++a) s += JH[10LL * M.nxh + a * M.nH + h];
#pragma unroll 8
    for (int g = 0; g < M.nH; ++g) {
      if (g == h) continue;
      const int lo = g < h ? g : h, hi = g < h ? h : g;
      const int og = A.aoff[M.atom0 + M.nh + g];
      s += M.Whh[tri_index(lo, hi, M.nH)] * P[og * np + og];
    }
    J[(M.nh + h) * 10] = s;
  }
}

// Phase 3: diagonal blocks, H + one-centre terms + J.  Each element with mu <= nu is computed once
// and stored on both sides, so that F is exactly symmetric (the (nu, mu) sum visits the same terms
// in another order).
__device__ inline void fock_diag(const KArgs& A, const Mol& M, const double* H, const double* P, const double* X,
                                 double xf, double* F, const double* J, int t0, int st) {
  const int np = M.np;
  for (int e = t0; e < M.nat * 16; e += st) {
    const int a = e >> 4, mu = (e >> 2) & 3, nu = e & 3;
    const int ga = M.atom0 + a;
    const ElemConst& ea = c_tab.el[A.aslot[ga]];
    const int nn = ea.norb;
    if (mu < nn && nu < nn && mu <= nu) {
      const int o = A.aoff[ga];
      double acc = 0.0;
      for (int la = 0; la < nn; ++la)
        for (int si = 0; si < nn; ++si) {
          const double pls = P[(o + la) * np + o + si];
          acc += pls * oc_int(mu, nu, la, si, ea);
          acc -= (xf * X[(o + la) * np + o + si]) * oc_int(mu, la, nu, si, ea);
        }
      const double v = H[(o + mu) * np + o + nu] + acc + J[a * 10 + pk(mu, nu)];
      F[(o + mu) * np + o + nu] = v;
      F[(o + nu) * np + o + mu] = v;
    }
  }
}

// F = H + J(P) - xf K(X) by one workgroup, J in LDS: the spin Fock build (build_fock_spin,
// src/fock.rs:285-332) with Coulomb from the total density P and exchange from the same-spin
// density X (xf = 1), or the RHF build (src/fock.rs:488-500) with X = P and xf = 1/2.
__device__ __noinline__ void build_fock_x(const KArgs& A, const Mol& M, const double* H, const double* P,
                                          const double* X, double xf, double* F, double* lds) {
  const int tid = threadIdx.x;
  fock_pairs(A, M, H, P, X, xf, F, tid, NT);
  __syncthreads();
  fock_jsum(A, M, P, lds, tid, NT);
  __syncthreads();
  fock_diag(A, M, H, P, X, xf, F, lds, tid, NT);
  __syncthreads();
}

__device__ inline void build_fock(const KArgs& A, const Mol& M, const double* H, const double* P, double* F, double* lds) {
  build_fock_x(A, M, H, P, P, 0.5, F, lds);
}

// ------------------------------------------------------------------------------------------
// Symmetric eigensolver: A (n x n, row-major ld np) -> eigenvalues lam[0..n) (unsorted, in the
// order of V's columns) and V (column-major, ld np) with A = V diag(lam) V^T.
// Cyclic Jacobi, round-robin (circle) ordering: in round r slot k pairs players
// ((r + k) mod (nn-1), nn-1) for k = 0 and ((r + k) mod (nn-1), (nn - 1 - k + r) mod (nn-1))
// otherwise, so every pair meets once per sweep and the m = nn/2 rotations of a round are
// disjoint.  Golub & Van Loan Alg. 8.5.2 for the rotation angle.
__device__ inline void rr_pair(int k, int r, int nn, int& p, int& q) {
  const int m1 = nn - 1;
  p = (r + k) % m1;
  q = (k == 0) ? m1 : (m1 - k + r) % m1;
}

// LDS layout of the eigensolver for an n x n problem (in doubles): packed upper triangle of A,
// then per-slot rotation (c, s, t), then the slot -> (p, q) schedule of the current round (two ints
// per slot), then, when room allows, the eigenvector matrix V (column-major, ld n).
__device__ __host__ inline int jac_lds_base(int n) {
  const int m = (n + 1) >> 1;
  return n * (n + 1) / 2 + 4 * m;
}
__device__ __host__ inline int jac_lds_with_v(int n) { return jac_lds_base(n) + n * n; }
// Eigenvectors in global memory with deferred updates: the rotations of JR rounds are buffered
// (c, s, p | q << 16) and applied to V in chunks of JCH rows staged through LDS.
constexpr int JR = 16, JCH = 32;  // JCH is a multiple of the 8 waves of a workgroup
__device__ __host__ inline int jac_lds_deferred(int n) {
  const int m = (n + (n & 1)) >> 1;
  return jac_lds_base(n) + 3 * m * JR + JCH * n;
}

// Dynamic LDS of the SCF kernels.  The eigensolver indexes it directly (not through a pointer
// parameter) so every access compiles to ds_read / ds_write rather than flat operations.
extern __shared__ double g_lds[];

constexpr int JB = 10;  // upper bound on the triangle blocks one thread owns (m(m+1)/2 <= JB * NT)

// V (eigenvectors, column-major, ld ldv) is in LDS at g_lds + vbase when VL, else at Vg.
template <bool VL>
__device__ __noinline__ int jacobi(const double* Ain, int n, int np, double* Vg, int ldv, int vbase, double* lam,
                                   double* red, int max_sweeps, double tol_abs, double rel, bool defer,
                                   double skip_frac) {
  const int tid = threadIdx.x;
  const int nn = n + (n & 1), m = nn >> 1;
  const int npk = n * (n + 1) / 2;
  // deferred eigenvector updates (VL == false): ring of JR rounds after the schedule, then the
  // staging chunk
  const bool dv = !VL && defer;
  double* ring_c = g_lds + jac_lds_base(n);
  double* ring_s = ring_c + JR * m;
  int* ring_pq = (int*)(ring_s + JR * m);
  double* chunk = ring_c + 3 * m * JR;
  int nbuf = 0;
  auto flush = [&](int nb) {
    for (int r0 = 0; r0 < n; r0 += JCH) {
      const int rows = min(JCH, n - r0);
      for (int e = tid; e < rows * n; e += NT) {
        const int i = e % rows, k = e / rows;
        chunk[i * n + k] = Vg[k * ldv + r0 + i];
      }
      __syncthreads();
      // wave w owns rows w, w + NW, ...; each lane applies its rotations of a round to all of the
      // wave's rows (independent chains), then the wave moves to the next round
      const int lane = tid & 63, w = tid >> 6;
      for (int j = 0; j < nb; ++j) {
        for (int kk = lane; kk < m; kk += 64) {
          const double sv = ring_s[j * m + kk];
          const int pqv = ring_pq[j * m + kk];
          const int p = pqv & 0xffff, q = pqv >> 16;
          if (sv == 0.0 || p >= n || q >= n) continue;
          const double cv = ring_c[j * m + kk];
#pragma unroll
          for (int u = 0; u < JCH / NW; ++u) {
            const int row = w + u * NW;
            if (row < rows) {
              double* v = chunk + row * n;
              const double a = v[p], b = v[q];
              v[p] = cv * a - sv * b;
              v[q] = sv * a + cv * b;
            }
          }
        }
        // the next round may pair columns this round's other lanes wrote
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      }
      __syncthreads();
      for (int e = tid; e < rows * n; e += NT) {
        const int i = e % rows, k = e / rows;
        Vg[k * ldv + r0 + i] = chunk[i * n + k];
      }
      __syncthreads();
    }
  };
  // g_lds: [0, npk) packed A | [npk, npk+3m) rotation (c, s, t) | [npk+3m, npk+4m) pair schedule
  const int cso = npk, pqo = npk + 3 * m;
  int* pq = (int*)(g_lds + pqo);
  for (int j = 0; j < n; ++j)
    for (int i = tid; i <= j; i += NT) g_lds[j * (j + 1) / 2 + i] = Ain[i * np + j];
  for (int e = tid; e < n * ldv; e += NT) {
    const int c = e / ldv, r = e - c * ldv;
    const double v = (c == r) ? 1.0 : 0.0;
    if (VL) g_lds[vbase + e] = v; else Vg[e] = v;
  }
  // this thread's blocks (K, L), K <= L, of the m x m slot grid, packed as K | L << 16
  int kl[JB];
  const int nblk = m * (m + 1) / 2;
  {
    int L = 0, base = 0;  // walk the triangle row by row: row L holds K = 0..L
#pragma unroll
    for (int j = 0; j < JB; ++j) {
      const int bidx = tid + j * NT;
      kl[j] = -1;
      if (bidx < nblk) {
        while (base + L + 1 <= bidx) { base += L + 1; ++L; }
        kl[j] = (bidx - base) | (L << 16);
      }
    }
  }
  __syncthreads();
  double dmax = 0.0;
  for (int i = tid; i < n; i += NT) dmax = fmax(dmax, fabs(g_lds[i * (i + 1) / 2 + i]));
  const double scale = fmax(block_max(dmax, red), 1.0);
  const double tol_floor = 1.0e-14 * scale;
  double tol = tol_floor, skip = 1.0e-300;
  // eigenvector update mapping: row i = tid % n, slots k = tid / n, tid / n + rstep, ...
  const int rstep = NT / max(n, 1);
  const int vi = tid % max(n, 1), vk0 = tid / max(n, 1);
  int sweep = 0;
  const int sweeps = n < 2 ? 0 : max_sweeps;
  for (; sweep < sweeps; ++sweep) {
    {  // largest off-diagonal: each thread walks one contiguous chunk of the packed triangle
      const int chunk = (npk + NT - 1) / NT;
      const int i0 = tid * chunk;
      double off = 0.0;
      if (i0 < npk) {
        int j = (int)((sqrtf(8.0f * (float)i0 + 1.0f) - 1.0f) * 0.5f);
        while (j * (j + 1) / 2 > i0) --j;
        while ((j + 1) * (j + 2) / 2 <= i0) ++j;
        int i = i0 - j * (j + 1) / 2;
        const int iend = min(i0 + chunk, npk);
        for (int idx = i0; idx < iend; ++idx) {
          if (i != j) off = fmax(off, fabs(g_lds[idx]));
          if (++i > j) { i = 0; ++j; }
        }
      }
      const double offmax = block_max(off, red);
      // a loose tolerance never exceeds a fixed fraction of the starting off-diagonal, so every
      // call that has anything to rotate does rotate (no stalled, falsely converged SCF)
      if (sweep == 0) {
        tol = fmax(tol_floor, fmin(tol_abs, rel * offmax));
        skip = fmax(1.0e-300, skip_frac * tol);
      }
      if (offmax <= tol) break;
    }
    for (int r = 0; r < nn - 1; ++r) {
      for (int k = tid; k < m; k += NT) {
        int p, q;
        rr_pair(k, r, nn, p, q);
        double c = 1.0, s = 0.0, t = 0.0;
        if (p < n && q < n) {
          const double apq = g_lds[sym_idx(p, q)];
          const double app = g_lds[sym_idx(p, p)], aqq = g_lds[sym_idx(q, q)];
          // threshold Jacobi: elements already below the stopping tolerance are not rotated
          if (fabs(apq) > skip && fabs(apq) > 1.0e-18 * (fabs(app) + fabs(aqq))) {
            const double tau = (aqq - app) / (2.0 * apq);
            t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        g_lds[cso + 3 * k] = c;
        g_lds[cso + 3 * k + 1] = s;
        g_lds[cso + 3 * k + 2] = t;
        pq[2 * k] = p;
        pq[2 * k + 1] = q;
        if (dv) {
          ring_c[nbuf * m + k] = c;
          ring_s[nbuf * m + k] = s;
          ring_pq[nbuf * m + k] = p | (q << 16);
        }
      }
      __syncthreads();
      // two-sided update of every block: J_K^T A_KL J_L (the diagonal blocks become diagonal),
      // four owned blocks at a time so their LDS reads overlap
#pragma unroll
      for (int j0 = 0; j0 < JB; j0 += 4) {
        double x[4][4];
        int ix[4][4];
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int j = j0 + jj;
          ix[jj][0] = ix[jj][1] = ix[jj][2] = ix[jj][3] = -1;
          if (j >= JB || kl[j] < 0) continue;
          const int K = kl[j] & 0xffff, L = kl[j] >> 16;
          // blocks whose two rotations are both identities keep their values
          if (g_lds[cso + 3 * K + 1] == 0.0 && g_lds[cso + 3 * L + 1] == 0.0) continue;
          const int p = pq[2 * K], q = pq[2 * K + 1];
          if (K == L) {
            if (p < n && q < n) ix[jj][0] = sym_idx(p, q);
          } else {
            const int u = pq[2 * L], v = pq[2 * L + 1];
            if (p < n && u < n) ix[jj][0] = sym_idx(p, u);
            if (p < n && v < n) ix[jj][1] = sym_idx(p, v);
            if (q < n && u < n) ix[jj][2] = sym_idx(q, u);
            if (q < n && v < n) ix[jj][3] = sym_idx(q, v);
          }
#pragma unroll
          for (int w = 0; w < 4; ++w) x[jj][w] = ix[jj][w] >= 0 ? g_lds[ix[jj][w]] : 0.0;
        }
#pragma unroll
        for (int jj = 0; jj < 4; ++jj) {
          const int j = j0 + jj;
          if (j >= JB || kl[j] < 0) continue;
          const int K = kl[j] & 0xffff, L = kl[j] >> 16;
          if (K == L) {
            const double tK = g_lds[cso + 3 * K + 2];
            if (ix[jj][0] >= 0 && tK != 0.0) {
              const int p = pq[2 * K], q = pq[2 * K + 1];
              g_lds[sym_idx(p, p)] -= tK * x[jj][0];
              g_lds[sym_idx(q, q)] += tK * x[jj][0];
              g_lds[ix[jj][0]] = 0.0;
            }
          } else {
            const double cK = g_lds[cso + 3 * K], sK = g_lds[cso + 3 * K + 1];
            const double cL = g_lds[cso + 3 * L], sL = g_lds[cso + 3 * L + 1];
            if (sK == 0.0 && sL == 0.0) continue;
            const double xpu = cK * x[jj][0] - sK * x[jj][2], xqu = sK * x[jj][0] + cK * x[jj][2];
            const double xpv = cK * x[jj][1] - sK * x[jj][3], xqv = sK * x[jj][1] + cK * x[jj][3];
            if (ix[jj][0] >= 0) g_lds[ix[jj][0]] = cL * xpu - sL * xpv;
            if (ix[jj][1] >= 0) g_lds[ix[jj][1]] = sL * xpu + cL * xpv;
            if (ix[jj][2] >= 0) g_lds[ix[jj][2]] = cL * xqu - sL * xqv;
            if (ix[jj][3] >= 0) g_lds[ix[jj][3]] = sL * xqu + cL * xqv;
          }
        }
      }
      // V <- V J: thread owns row vi, slots vk0, vk0 + rstep, ...
      if (vk0 < rstep && !dv) {
        for (int k0 = vk0; k0 < m; k0 += 8 * rstep) {
          double vp[8], vq[8], c8[8], s8[8];
          int op[8], oq[8];
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            const int k = k0 + u * rstep;
            op[u] = -1;
            if (k >= m) continue;
            const double sv = g_lds[cso + 3 * k + 1];
            const int p = pq[2 * k], q = pq[2 * k + 1];
            if (sv == 0.0 || p >= n || q >= n) continue;
            s8[u] = sv;
            c8[u] = g_lds[cso + 3 * k];
            op[u] = p * ldv + vi;
            oq[u] = q * ldv + vi;
            vp[u] = VL ? g_lds[vbase + op[u]] : Vg[op[u]];
            vq[u] = VL ? g_lds[vbase + oq[u]] : Vg[oq[u]];
          }
#pragma unroll
          for (int u = 0; u < 8; ++u) {
            if (op[u] < 0) continue;
            const double a = c8[u] * vp[u] - s8[u] * vq[u], b = s8[u] * vp[u] + c8[u] * vq[u];
            if (VL) { g_lds[vbase + op[u]] = a; g_lds[vbase + oq[u]] = b; }
            else { Vg[op[u]] = a; Vg[oq[u]] = b; }
          }
        }
      }
      __syncthreads();
      if (dv && ++nbuf == JR) {
        flush(nbuf);
        nbuf = 0;
      }
    }
  }
  if (dv && nbuf > 0) flush(nbuf);
  for (int i = tid; i < n; i += NT) lam[i] = g_lds[i * (i + 1) / 2 + i];
  __syncthreads();
  return sweep;
}

// ------------------------------------------------------------------------------------------
// DIIS history (src/scf.rs:1065-1173) kept in LDS as logical slots oldest -> newest.
struct Diis {
  int len, depth, keep_d, newslot;
  double en;  // |E| of the newest entry
  int slot[MAXD];
  double b[MAXD][MAXD];
  double mm[MAXD][MAXD];
  double coef[MAXD];
  int ok;
  int mode;      // coefficient rule of the newest step: 0 none, 1 CDIIS, 2 A-DIIS
  int mm_stale;  // some <D_i, F_j> entries were skipped (only A-DIIS reads them)
};

// Pulay CDIIS coefficients from the (scale-normalised) bordered Gram system with a pivot guard
// (src/scf.rs:1544-1630).  Single thread (device) or host.
__host__ __device__ __noinline__ bool cdiis_coeffs(const Diis& d, double* c) {
  const int n = d.len;
  if (n < 2) return false;
  double scale = 0.0;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) scale = fmax(scale, fabs(d.b[i][j]));
  if (!(scale > 0.0) || !isfinite(scale)) return false;
  const int dim = n + 1;
  double m[MAXD + 1][MAXD + 1], rhs[MAXD + 1], x[MAXD + 1];
  for (int i = 0; i < dim; ++i) {
    rhs[i] = 0.0;
    for (int j = 0; j < dim; ++j) m[i][j] = 0.0;
  }
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) m[i][j] = d.b[i][j] / scale;
    m[i][n] = -1.0;
    m[n][i] = -1.0;
  }
  rhs[n] = -1.0;
  for (int col = 0; col < dim; ++col) {
    int piv = col;
    double best = fabs(m[col][col]);
    for (int row = col + 1; row < dim; ++row)
      if (fabs(m[row][col]) > best) { best = fabs(m[row][col]); piv = row; }
    if (best < 1.0e-12) return false;
    if (piv != col) {
      for (int j = 0; j < dim; ++j) { double t = m[col][j]; m[col][j] = m[piv][j]; m[piv][j] = t; }
      double t = rhs[col]; rhs[col] = rhs[piv]; rhs[piv] = t;
    }
    for (int row = col + 1; row < dim; ++row) {
      const double f = m[row][col] / m[col][col];
      if (f == 0.0) continue;
      for (int j = col; j < dim; ++j) m[row][j] -= f * m[col][j];
      rhs[row] -= f * rhs[col];
    }
  }
  for (int col = dim - 1; col >= 0; --col) {
    double s = rhs[col];
    for (int j = col + 1; j < dim; ++j) s -= m[col][j] * x[j];
    x[col] = s / m[col][col];
  }
  for (int i = 0; i < n; ++i) c[i] = x[i];
  return true;
}

// A-DIIS (Hu & Yang 2010): projected gradient on the probability simplex, 400 steps at most
// (src/scf.rs:1637-1693).  Single thread (device) or host.
__host__ __device__ __noinline__ void adiis_coeffs(const Diis& D, double* c) {
  const int n = D.len, last = n - 1;
  const double mnn = D.mm[last][last];
  double d[MAXD], s[MAXD][MAXD];
  for (int i = 0; i < n; ++i) d[i] = D.mm[i][last] - mnn;
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) s[i][j] = D.mm[i][j] - D.mm[i][last] - D.mm[last][j] + mnn;
  double L = 1.0e-12;
  for (int i = 0; i < n; ++i) {
    double row = 0.0;
    for (int j = 0; j < n; ++j) row += fabs(s[i][j] + s[j][i]);
    L = fmax(L, row);
  }
  const double lr = 1.0 / L;
  for (int i = 0; i < n; ++i) c[i] = 0.0;
  c[n - 1] = 1.0;
  for (int it = 0; it < 400; ++it) {
    double tr[MAXD], u[MAXD];
    for (int k = 0; k < n; ++k) {
      double acc = 2.0 * d[k];
      for (int j = 0; j < n; ++j) acc += (s[k][j] + s[j][k]) * c[j];
      tr[k] = c[k] - lr * acc;
      u[k] = tr[k];
    }
    // sort u descending (insertion sort, n <= 8) and find the simplex threshold
    for (int i = 1; i < n; ++i) {
      const double x = u[i];
      int j = i - 1;
      while (j >= 0 && u[j] < x) { u[j + 1] = u[j]; --j; }
      u[j + 1] = x;
    }
    double css = 0.0, theta = 0.0;
    for (int j = 0; j < n; ++j) {
      css += u[j];
      const double t = (css - 1.0) / (double)(j + 1);
      if (u[j] - t > 0.0) theta = t;
    }
    double delta = 0.0;
    for (int i = 0; i < n; ++i) {
      const double pr = fmax(tr[i] - theta, 0.0);
      delta += fabs(pr - c[i]);
      c[i] = pr;
    }
    if (delta < 1.0e-12) break;
  }
}

// cdiis_coeffs on one wave: lane r holds row r of the bordered system in registers (static
// indices only, so nothing goes to scratch).  Pivot search (first largest), row exchange,
// elimination (row r updated with the same operations in the same order) and back substitution
// follow the serial routine, so the coefficients are bitwise those of cdiis_coeffs.
// Returns the same success flag on every lane.
__device__ bool cdiis_wave(const Diis& D, double* c) {
  constexpr int DM = MAXD + 1;
  const int lane = threadIdx.x & 63;
  const int n = D.len;
  if (n < 2) return false;
  double sc = 0.0;
  if (lane < n)
    for (int j = 0; j < n; ++j) sc = fmax(sc, fabs(D.b[lane][j]));
  double scale = 0.0;
  for (int i = 0; i < n; ++i) scale = fmax(scale, lane_value(sc, i));
  if (!(scale > 0.0) || !isfinite(scale)) return false;
  const int dim = n + 1;
  double mr[DM];
  double rhs = 0.0;
#pragma unroll
  for (int j = 0; j < DM; ++j) {
    double v = 0.0;
    if (lane < n) {
      if (j < n) v = D.b[lane][j] / scale;
      else if (j == n) v = -1.0;
    } else if (lane == n && j < n) {
      v = -1.0;
    }
    mr[j] = v;
  }
  if (lane == n) rhs = -1.0;
  for (int col = 0; col < dim; ++col) {
    double mc = 0.0;
#pragma unroll
    for (int j = 0; j < DM; ++j)
      if (j == col) mc = mr[j];
    const double val = (lane >= col && lane < dim) ? fabs(mc) : -1.0;
    double best = -1.0;
    for (int r = col; r < dim; ++r) best = fmax(best, lane_value(val, r));
    if (best < 1.0e-12) return false;
    const unsigned long long hit = __ballot(lane >= col && lane < dim && val == best);
    const int piv = __ffsll((long long)hit) - 1;
    if (piv != col) {
      const int partner = lane == col ? piv : lane == piv ? col : lane;
#pragma unroll
      for (int j = 0; j < DM; ++j) mr[j] = __shfl(mr[j], partner, 64);
      rhs = __shfl(rhs, partner, 64);
    }
    double prow[DM];
#pragma unroll
    for (int j = 0; j < DM; ++j) prow[j] = lane_value(mr[j], col);
    const double prhs = lane_value(rhs, col);
    double pc = 0.0, rc = 0.0;
#pragma unroll
    for (int j = 0; j < DM; ++j)
      if (j == col) {
        pc = prow[j];
        rc = mr[j];
      }
    if (lane > col && lane < dim) {
      const double f = rc / pc;
      if (f != 0.0) {
#pragma unroll
        for (int j = 0; j < DM; ++j)
          if (j >= col && j < dim) mr[j] -= f * prow[j];
        rhs -= f * prhs;
      }
    }
  }
  double x[DM];
#pragma unroll
  for (int j = 0; j < DM; ++j) x[j] = 0.0;
  for (int col = dim - 1; col >= 0; --col) {
    double sv = rhs, dg = 1.0;
#pragma unroll
    for (int j = 0; j < DM; ++j) {
      if (j > col && j < dim) sv -= mr[j] * x[j];
      if (j == col) dg = mr[j];
    }
    const double xc = lane_value(sv / dg, col);
#pragma unroll
    for (int j = 0; j < DM; ++j)
      if (j == col) x[j] = xc;
  }
  if (lane < n) {
#pragma unroll
    for (int j = 0; j < DM; ++j)
      if (j == lane) c[lane] = x[j];
  }
  return true;
}

// adiis_coeffs on one wave: lane k < len owns c[k].  Every lane keeps the serial routine's order
// of operations (its own row of the gradient summed in j order; the stable descending sort as a
// rank, then prefix sums over the sorted values in order; the change summed in index order), so
// the coefficients are bitwise those of adiis_coeffs.  srt: 8 doubles of LDS.
__device__ void adiis_wave(Diis& D, double* srt) {
  const int lane = threadIdx.x & 63;
  const int n = D.len, last = n - 1;
  const int k = lane < n ? lane : 0;
  const double mnn = D.mm[last][last];
  const double dk = D.mm[k][last] - mnn;
  double srow[MAXD];
  double row = 0.0;
#pragma unroll
  for (int j = 0; j < MAXD; ++j) {
    srow[j] = 0.0;
    if (j < n) {
      const double skj = D.mm[k][j] - D.mm[k][last] - D.mm[last][j] + mnn;
      const double sjk = D.mm[j][k] - D.mm[j][last] - D.mm[last][k] + mnn;
      srow[j] = skj + sjk;
      row += fabs(srow[j]);
    }
  }
  double L = 1.0e-12;
  for (int i = 0; i < n; ++i) L = fmax(L, lane_value(row, i));
  const double lr = 1.0 / L;
  double c = (lane == n - 1) ? 1.0 : 0.0;
  for (int it = 0; it < 400; ++it) {
    double acc = 2.0 * dk;
#pragma unroll
    for (int j = 0; j < MAXD; ++j)
      if (j < n) acc += srow[j] * lane_value(c, j);
    const double tr = c - lr * acc;
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double uj = lane_value(tr, j);
      rank += (uj > tr || (uj == tr && j < lane)) ? 1 : 0;
    }
    if (lane < n) srt[rank] = tr;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    double css = 0.0;
    for (int i = 0; i <= k; ++i) css += srt[i];
    const double t = (css - 1.0) / (double)(k + 1);
    const bool cond = lane < n && srt[k] - t > 0.0;
    const unsigned long long m = __ballot(cond);
    const double theta = m ? lane_value(t, 63 - __clzll((long long)m)) : 0.0;
    const double pr = fmax(tr - theta, 0.0);
    const double dlt = fabs(pr - c);
    if (lane < n) c = pr;
    double delta = 0.0;
    for (int i = 0; i < n; ++i) delta += lane_value(dlt, i);
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
    if (delta < 1.0e-12) break;
  }
  if (lane < n) D.coef[lane] = c;
}

// ------------------------------------------------------------------------------------------
struct Bufs {
  double *H, *F, *P, *Pn, *Fu, *T, *C, *Cn, *V, *Pb, *Cb, *Fh, *Eh, *Dh, *eps, *epsb, *lam;
  double *fs, *ft;  // frontier-level seed s of the last purification, and X s
  double* bnd;  // LDS: cached spectral-bound estimates of the purification ([0] lo, [1] hi, [2] age),
                // [3] 1 once a purification of this path has written the frontier seed
  int* perm;
  unsigned long long* ph;  // LDS: [0..11] per-phase wall-clock sums, [12] last stamp, [13..15] counters; null when off
};

// Phase stamp: adds the wall-clock ticks since the previous stamp to phase k (thread 0 only).
// Phases: 0 H_core finish, 1 Fock build, 2 energy + DIIS history and Gram sums, 3 dense products
// (F P and the eigensolver's warm-start transform), 4 eigensolve or purification, 5 eigenvector
// sort and back-transform or frontier levels, 6 density change and convergence test, 7 other
// (guess, stability search, epilogue), 8 DIIS coefficient solve, 9 DIIS extrapolation.
__device__ inline void stamp(const Bufs& B, int k) {
  if (B.ph && threadIdx.x == 0) {
    const unsigned long long now = wall_clock64();
    B.ph[k] += now - B.ph[PH_LAST];
    B.ph[PH_LAST] = now;
  }
}

// Eigenvectors of Fu -> Cn (row-major, columns sorted by ascending eigenvalue) and eps.
// warm != 0: Jacobi on C^T Fu C and Cn = C V; otherwise Jacobi on Fu itself.  Sweeps stop once
// the largest off-diagonal element is below max(1e-14 |A|, tol_abs).
__device__ __noinline__ void eigensolve(const Mol& M, const Bufs& B, const double* Fu, bool warm, double* lds, double* red,
                                        double tol_abs = 0.0) {
  const int n = M.n, np = M.np, tid = threadIdx.x;
  const double* Amat = Fu;
  if (warm) {
    const double* C = B.C;
    double* T = B.T;
    // T = Fu C ; A = C^T T  (A written into Cn as scratch)
    wg_gemm(n, n, n, [&](int i, int k) { return Fu[i * np + k]; }, [&](int k, int j) { return C[k * np + j]; },
            [&](int i, int j, double v) { T[i * np + j] = v; });
    __syncthreads();
    double* Am = B.Cn;
    wg_gemm(n, n, n, [&](int i, int k) { return C[k * np + i]; }, [&](int k, int j) { return T[k * np + j]; },
            [&](int i, int j, double v) { Am[i * np + j] = v; });
    __syncthreads();
    Amat = Am;
  }
  stamp(B, 3);
  const bool v_in_lds = jac_lds_with_v(n) <= M.lds_doubles;
  double* V = v_in_lds ? lds + jac_lds_base(n) : B.V;
  const int ldv = v_in_lds ? n : np;
  const int nsw = v_in_lds ? jacobi<true>(Amat, n, np, nullptr, ldv, jac_lds_base(n), B.lam, red, 60, tol_abs, M.eig_rel, false,
                                          M.eig_skip)
                           : jacobi<false>(Amat, n, np, B.V, ldv, 0, B.lam, red, 60, tol_abs, M.eig_rel,
                                           jac_lds_deferred(n) <= M.lds_doubles, M.eig_skip);
  if (B.ph && threadIdx.x == 0) {
    B.ph[PH_LAST + 1] += (unsigned long long)nsw;
    B.ph[PH_LAST + 2] += 1ull;
    B.ph[PH_LAST + 3] += (unsigned long long)nsw * (unsigned long long)(n + (n & 1) - 1);
  }
  stamp(B, 4);
  // ascending order, ties by index
  for (int k = tid; k < n; k += NT) {
    const double lk = B.lam[k];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double lj = B.lam[j];
      rank += (lj < lk || (lj == lk && j < k)) ? 1 : 0;
    }
    B.perm[rank] = k;
  }
  __syncthreads();
  for (int j = tid; j < n; j += NT) B.eps[j] = B.lam[B.perm[j]];
  const int* perm = B.perm;
  double* Cn = B.Cn;
  if (warm) {
    const double* C = B.C;
    wg_gemm(n, n, n, [&](int i, int k) { return C[i * np + k]; }, [&](int k, int j) { return V[perm[j] * ldv + k]; },
            [&](int i, int j, double v) { Cn[i * np + j] = v; });
  } else {
    for (int e = tid; e < n * n; e += NT) {
      const int i = e / n, j = e % n;
      Cn[i * np + j] = V[perm[j] * ldv + i];
    }
  }
  __syncthreads();
  stamp(B, 5);
}

// P = wt C_occ C_occ^T from Cn into Pdst (wt = 2 restricted, 1 for one spin).
__device__ void density(const Mol& M, const double* Cn, double* Pdst) {
  const int n = M.n, np = M.np;
  const double wt = M.wt;
  wg_gemm<true>(n, n, M.nocc, [&](int i, int k) { return Cn[i * np + k]; }, [&](int k, int j) { return Cn[j * np + k]; },
                [&](int i, int j, double v) { Pdst[i * np + j] = wt * v; });
  __syncthreads();
}

// Lowest and highest Gershgorin bound of the symmetric matrix F (columns read coalesced).
__device__ inline void gershgorin(const double* F, int n, int np, double* red, double& lo, double& hi) {
  double l = 1.0e300, h = -1.0e300;
  for (int i = threadIdx.x; i < n; i += NT) {
    double r = 0.0;
    for (int j = 0; j < n; ++j) r += j == i ? 0.0 : fabs(F[j * np + i]);
    const double c = F[i * np + i];
    l = fmin(l, c - r);
    h = fmax(h, c + r);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    l = fmin(l, __shfl_xor(l, o, 64));
    h = fmax(h, __shfl_xor(h, o, 64));
  }
  __syncthreads();
  if (lane == 0) {
    red[2 * wave] = l;
    red[2 * wave + 1] = h;
  }
  __syncthreads();
  lo = red[0];
  hi = red[1];
  for (int w = 1; w < NW; ++w) {
    lo = fmin(lo, red[2 * w]);
    hi = fmax(hi, red[2 * w + 1]);
  }
  __syncthreads();
}

constexpr int SP2_MAX = 90;
constexpr int SP2_EXTRA = 2;
constexpr double SP2_TAIL = 1.0e-7;

// LDS image of an n x n operand for the products: kp = n rounded up to 16 rows (zero beyond n) of
// sp2_ld(n) doubles, a stride of 16 mod 32 so that the four k rows one load touches fall in
// different banks.
__host__ __device__ constexpr int sp2_kp(int n) { return (n + 15) & ~15; }
__host__ __device__ constexpr int sp2_ld(int n) { return (sp2_kp(n) & 31) == 0 ? sp2_kp(n) + 16 : sp2_kp(n); }
__host__ __device__ constexpr int sp2_img(int n) { return sp2_kp(n) * sp2_ld(n); }
// Rows of X that the LM = 4 placement keeps in LDS: as many 16-row chunks as lds doubles allow.
__host__ __device__ constexpr int sp2_rows(int n, int lds) {
  return ((lds / sp2_ld(n)) & ~15) < sp2_kp(n) ? ((lds / sp2_ld(n)) & ~15) : sp2_kp(n);
}

// Fixed generic vector of the frontier seed (cheap to evaluate: one per row and thread).
__host__ __device__ inline double seed_r(int j) {
  const double x = 0.6180339887498949 * (double)j + 0.1;
  return 1.0 + 0.5 * (x - floor(x));
}

// Frontier seed s = (X - X^2) r from the purification iterate X at the start of its stopping tail.
// X = p(Fu) for a polynomial p that is increasing on [0, 1] and splits the spectrum at the gap, so
// x (1 - x) is largest at the levels next to the gap (the HOMO among the occupied levels, the LUMO
// among the virtual ones): s is dominated by the frontier directions.  One thread per row; X is
// symmetric, so row i is read as column i (consecutive threads, consecutive addresses).
__device__ void frontier_seed(const double* X, int ld, int n, const Bufs& B) {
  const int tid = threadIdx.x;
  double t = 0.0;
  if (tid < n)
    for (int j = 0; j < n; ++j) t += X[j * ld + tid] * seed_r(j);
  if (tid < n) B.ft[tid] = t;
  __syncthreads();
  double u = 0.0;
  if (tid < n)
    for (int j = 0; j < n; ++j) u += X[j * ld + tid] * B.ft[j];
  __syncthreads();
  if (tid < n) B.fs[tid] = t - u;
  if (tid == 0 && B.bnd) B.bnd[3] = 1.0;
  __syncthreads();
}

// The purification loop for one placement of the operands: LM = 2 both X and the product in LDS
// (ping-pong of two images), LM = 3 one LDS image that each product overwrites in place (its tiles
// held in registers until every wave has read X, wg_gemm_sym_hold), LM = 1 X in LDS and the
// product to HBM (B.T, copied back into the image after each step), LM = 0 both in HBM (B.Cn / B.T),
// LM = 4 both in HBM with X's first sp2_rows(n) rows also kept in LDS (refreshed after each
// product), for sizes whose image does not fit.
template <int LM>
__device__ int sp2_run(const Mol& M, const Bufs& B, const double* Fu, double* Pdst, double* red, double lo, double hi,
                       bool seed) {
  const int n = M.n, np = M.np, nocc = M.nocc, tid = threadIdx.x;
  const int img = sp2_img(n), ldl = sp2_ld(n);
  const double s = 1.0 / (hi - lo);
  constexpr bool HB = LM == 0 || LM == 4;  // X and its product in HBM
  double* X = HB ? B.Cn : g_lds;
  double* Y = LM == 2 ? g_lds + img : LM == 3 ? g_lds : B.T;
  const int ldx = HB ? np : ldl, ldy = (LM == 2 || LM == 3) ? ldl : np;
  const int kL = LM == 4 ? sp2_rows(n, M.lds_doubles) : 0;
  // LM = 4: rows [0, kL) of the HBM iterate copied into the LDS image (zeros beyond n)
  auto refresh = [&](const double* Xh) {
    for (int e = tid; e < kL * ldl; e += NT) {
      const int r = e / ldl, c = e % ldl;
      g_lds[e] = (r < n && c < n) ? Xh[r * np + c] : 0.0;
    }
    __syncthreads();
  };
  double v[2] = {0.0, 0.0};
  if (!HB) {
    for (int e = tid; e < (LM == 2 ? 2 * img : img); e += NT) {
      const int r = (e % img) / ldl, c = e % ldl;
      double x = 0.0;
      if (e < img && r < n && c < n) {
        x = ((r == c ? hi : 0.0) - Fu[r * np + c]) * s;
        v[1] += x * x;
        if (r == c) v[0] += x;
      }
      g_lds[e] = x;
    }
  } else {
    for (int e = tid; e < n * n; e += NT) {
      const int i = e / n, j = e % n;
      const double x = ((i == j ? hi : 0.0) - Fu[i * np + j]) * s;
      X[i * np + j] = x;
      v[1] += x * x;
      if (i == j) v[0] += x;
    }
  }
  block_sum<2>(v, red);
  if (LM == 4) refresh(X);
  double tr = v[0], t2 = v[1];
  const double occ = (double)nocc;
  int tail = -1, k = 0;
  for (; k < SP2_MAX; ++k) {
    // every eigenvalue of X in [0, 1] keeps tr X^2 <= tr X; more means a bound was violated
    if (t2 > tr + 1.0e-9 * fmax(1.0, tr)) return -2;
    if (tail < 0 && tr - t2 < SP2_TAIL) {
      tail = SP2_EXTRA;
      if (seed) frontier_seed(X, ldx, n, B);
    }
    if (tail == 0) break;
    const bool sq = fabs(t2 - occ) < fabs(2.0 * tr - t2 - occ);
    const double* Xc = X;
    double* Yc = Y;
    double w[2] = {0.0, 0.0};
    auto store = [&](int i, int j, double x2) {
      const double y = sq ? x2 : 2.0 * Xc[i * ldx + j] - x2;
      Yc[i * ldy + j] = y;
      w[1] += y * y;
      if (i == j) w[0] += y;
    };
    if constexpr (LM == 3) {
      if (n <= M.sp2_tile16_n) wg_gemm_sym_hold<16>(n, ldx, Xc, store);
      else wg_gemm_sym_hold<32>(n, ldx, Xc, store);
    } else if constexpr (LM == 4) {
      if (n <= M.sp2_tile16_n) wg_gemm_sym_split<16>(n, kL, g_lds, ldl, Xc, np, store);
      else wg_gemm_sym_split<32>(n, kL, g_lds, ldl, Xc, np, store);
    } else {
      if (n <= M.sp2_tile16_n) wg_gemm_pad<true, 16>(n, ldx, Xc, Xc, store);
      else wg_gemm_pad<true, 32>(n, ldx, Xc, Xc, store);
    }
    block_sum<2>(w, red);
    tr = w[0];
    t2 = w[1];
    if (LM == 1) {
      for (int e = tid; e < n * n; e += NT) {
        const int i = e / n, j = e % n;
        X[i * ldl + j] = Y[i * np + j];
      }
      __syncthreads();
    } else if (LM != 3) {
      double* t = X;
      X = Y;
      Y = t;
      if (LM == 4) refresh(X);
    }
    if (tail > 0) --tail;
  }
  if (tail != 0) return -1;
  if (fabs(tr - occ) > 1.0e-6) return -2;
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e % n;
    Pdst[i * np + j] = M.wt * X[i * ldx + j];
  }
  __syncthreads();
  return k;
}

// ---- Packed tiled image of a symmetric matrix (the purification's LM = 5 placement) -----------
// Only the upper-triangle 16 x 16 blocks are kept, row-major over the block rows (bm = n / 16
// rounded up): block (br, bc), br <= bc, starts at tb_index(bm, br, bc) * 256 doubles, and its
// element (r, c) sits at r * 16 + (c ^ r).  The XOR swizzle puts the 16 elements of a block row
// and the 16 elements of a block column each in 16 different bank pairs, so a lane group reads
// either direction, and writes its product tile, without bank conflicts.  Diagonal blocks hold
// both triangles; X(r, c) with r's block below c's is read from the (c, r) copy.  Half the LDS of
// a square image and no mirrored stores.
__host__ __device__ constexpr int tb_blocks(int n) { return ((n + 15) >> 4) * (((n + 15) >> 4) + 1) / 2; }
__device__ __forceinline__ int tb_index(int bm, int br, int bc) { return br * bm - (br * (br - 1)) / 2 + (bc - br); }
__device__ __forceinline__ int tb_at(int bm, int r, int c) {
  int br = r >> 4, bc = c >> 4, lr = r & 15, lc = c & 15;
  if (br > bc) {
    const int t = br; br = bc; bc = t;
    const int u = lr; lr = lc; lc = u;
  }
  return (tb_index(bm, br, bc) << 8) + (lr << 4) + (lc ^ lr);
}
// Product tiles a wave holds at most in tb_sym_step (registers: 4 doubles each).
constexpr int TB_HOLD = NT == 512 ? 10 : 9;
__host__ __device__ constexpr bool tb_fits(int n, int lds_doubles) {
  return tb_blocks(n) * 256 <= lds_doubles && tb_blocks(n) <= TB_HOLD * NW;
}

// The packed tiled image as a whole: blocks [0, nl) at L (LDS) and, when the image does not fit,
// blocks [nl, nblk) at H (the workgroup's HBM slab, where L2 keeps them).  Every access branches on
// the block index (the same for a whole wave), so an LDS block is read with LDS instructions.
struct TbImg {
  double* L;
  double* H;
  int nl;
};
template <bool SPLIT> __device__ __forceinline__ double tb_get(const TbImg& im, int at) {
  const int blk = at >> 8;
  if (!SPLIT || blk < im.nl) return im.L[at];
  return im.H[at - (im.nl << 8)];
}
template <bool SPLIT> __device__ __forceinline__ void tb_put(const TbImg& im, int at, double v) {
  const int blk = at >> 8;
  if (!SPLIT || blk < im.nl)
    im.L[at] = v;
  else
    im.H[at - (im.nl << 8)] = v;
}

// One 16 x 16 tile (bi, bj) of X X on the FP64 matrix cores: out(i, j) = sum_k X(k, i) X(k, j),
// each 16-row block of k as four v_mfma_f64_16x16x4_f64 steps, the next block's operands loaded
// while the current one's MFMAs run.  Lane l: A value X(k, 16 bi + l % 16), B value
// X(k, 16 bj + l % 16) with k = 16 kb + 4 u + l / 16.
template <bool SPLIT>
__device__ inline double4_t tb_tile(int bm, const TbImg& im, int bi, int bj, int li, int lk) {
  int offd[4], offt[4];  // this lane's element of a directly stored / transposed operand block
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const int lr = 4 * u + lk;
    offd[u] = (lr << 4) + (li ^ lr);
    offt[u] = (li << 4) + (lr ^ li);
  }
  auto load = [&](int kb, int ob, double (&v)[4]) {
    const bool direct = kb <= ob;
    const int blk = direct ? tb_index(bm, kb, ob) : tb_index(bm, ob, kb);
    if (!SPLIT || blk < im.nl) {
      const double* base = im.L + (blk << 8);
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = base[direct ? offd[u] : offt[u]];
    } else {
      const double* base = im.H + ((blk - im.nl) << 8);
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = base[direct ? offd[u] : offt[u]];
    }
  };
  double4_t c = {0.0, 0.0, 0.0, 0.0};
  double a[4], b[4];
  load(0, bi, a);
  load(0, bj, b);
  for (int kb = 0; kb < bm; ++kb) {
    const int kn = kb + 1 < bm ? kb + 1 : kb;  // after the last block: a harmless reload
    double na[4], nb[4];
    load(kn, bi, na);
    load(kn, bj, nb);
#pragma unroll
    for (int u = 0; u < 4; ++u) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], b[u], c, 0, 0, 0);
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      a[u] = na[u];
      b[u] = nb[u];
    }
  }
  return c;
}

// X <- X^2 (sq) or 2 X - X^2 in place on the packed tiled image: every wave computes its upper
// tiles into registers, all waves meet once X has been read, then each tile is stored over its own
// block (one writer per element, no mirror).  w0 / w1 receive this thread's share of tr X and
// |X|_F^2 of the new X (off-diagonal blocks count twice).  Without SPLIT this requires tb_fits;
// with it, the tiles beyond the TB_HOLD a wave holds go through T (HBM, same layout) before the
// meeting, and the image may extend into HBM.
template <bool SPLIT>
__device__ inline void tb_sym_step(int n, int bm, const TbImg& im, bool sq, double& w0, double& w1, double* T) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int nblk = bm * (bm + 1) / 2;
  double4_t c[TB_HOLD];
  int bi_[TB_HOLD], bj_[TB_HOLD];
#pragma unroll
  for (int h = 0; h < TB_HOLD; ++h) {
    const int t = wave + h * NW;
    bi_[h] = bj_[h] = 0;
    if (t < nblk) {
      pad_block<true>(t, bm, bi_[h], bj_[h]);
      c[h] = tb_tile<SPLIT>(bm, im, bi_[h], bj_[h], li, lk);
    }
  }
  if (SPLIT) {
    for (int t = wave + TB_HOLD * NW; t < nblk; t += NW) {
      int bi, bj;
      pad_block<true>(t, bm, bi, bj);
      const double4_t x2 = tb_tile<SPLIT>(bm, im, bi, bj, li, lk);
      double* o = T + ((t - TB_HOLD * NW) << 8);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int lr = 4 * r + lk;
        o[(lr << 4) + (li ^ lr)] = x2[r];
      }
    }
  }
  __syncthreads();
  auto upd = [&](int bi, int bj, int r, double x2) {
    const int lr = 4 * r + lk;
    const int at = (tb_index(bm, bi, bj) << 8) + (lr << 4) + (li ^ lr);
    const double y = sq ? x2 : 2.0 * tb_get<SPLIT>(im, at) - x2;
    tb_put<SPLIT>(im, at, y);
    w1 += (bi < bj ? 2.0 : 1.0) * y * y;
    if (bi == bj && lr == li) w0 += y;
  };
#pragma unroll
  for (int h = 0; h < TB_HOLD; ++h) {
    const int t = wave + h * NW;
    if (t < nblk) {
#pragma unroll
      for (int r = 0; r < 4; ++r) upd(bi_[h], bj_[h], r, c[h][r]);
    }
  }
  if (SPLIT) {
    for (int t = wave + TB_HOLD * NW; t < nblk; t += NW) {  // this lane's own entries of T
      int bi, bj;
      pad_block<true>(t, bm, bi, bj);
      const double* o = T + ((t - TB_HOLD * NW) << 8);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int lr = 4 * r + lk;
        upd(bi, bj, r, o[(lr << 4) + (li ^ lr)]);
      }
    }
  }
}

// frontier_seed on the packed tiled image.
template <bool SPLIT>
__device__ void frontier_seed_tb(const TbImg& im, int bm, int n, const Bufs& B) {
  const int tid = threadIdx.x;
  double t = 0.0;
  if (tid < n)
    for (int j = 0; j < n; ++j) t += tb_get<SPLIT>(im, tb_at(bm, j, tid)) * seed_r(j);
  if (tid < n) B.ft[tid] = t;
  __syncthreads();
  double u = 0.0;
  if (tid < n)
    for (int j = 0; j < n; ++j) u += tb_get<SPLIT>(im, tb_at(bm, j, tid)) * B.ft[j];
  __syncthreads();
  if (tid < n) B.fs[tid] = t - u;
  if (tid == 0 && B.bnd) B.bnd[3] = 1.0;
  __syncthreads();
}

// Blocks of the packed tiled image that LDS holds for the split placement.
__host__ __device__ constexpr int tb_lds_blocks(int n, int lds_doubles) {
  return tb_blocks(n) < lds_doubles / 256 ? tb_blocks(n) : lds_doubles / 256;
}

// The purification loop of sp2_run on the packed tiled image (the same stopping rules; the trace
// sums run in another order, so a step choice can differ only where two of them tie to rounding).
// SPLIT: the blocks beyond what LDS holds live in B.Cn, and the product tiles beyond a wave's hold
// pass through B.T (the image of a 194-AO molecule is 91 blocks, 182 KiB).
template <bool SPLIT>
__device__ int sp2_run_tb(const Mol& M, const Bufs& B, const double* Fu, double* Pdst, double* red, double lo,
                          double hi, bool seed) {
  const int n = M.n, np = M.np, nocc = M.nocc, tid = threadIdx.x;
  const int bm = (n + 15) >> 4, nimg = tb_blocks(n) * 256;
  const TbImg im{g_lds, B.Cn, SPLIT ? tb_lds_blocks(n, M.lds_doubles) : tb_blocks(n)};
  const double s = 1.0 / (hi - lo);
  double v[2] = {0.0, 0.0};
  for (int e = tid; e < nimg; e += NT) {
    const int blk = e >> 8;
    int br = 0;
    while (tb_index(bm, br + 1, br + 1) <= blk && br + 1 < bm) ++br;
    const int bc = br + (blk - tb_index(bm, br, br));
    const int lr = (e >> 4) & 15, lc = (e & 15) ^ lr;
    const int r = (br << 4) + lr, c = (bc << 4) + lc;
    double x = 0.0;
    if (r < n && c < n) {
      x = ((r == c ? hi : 0.0) - Fu[r * np + c]) * s;
      v[1] += (br < bc ? 2.0 : 1.0) * x * x;
      if (r == c) v[0] += x;
    }
    tb_put<SPLIT>(im, e, x);
  }
  block_sum<2>(v, red);
  double tr = v[0], t2 = v[1];
  const double occ = (double)nocc;
  int tail = -1, k = 0;
  for (; k < SP2_MAX; ++k) {
    if (t2 > tr + 1.0e-9 * fmax(1.0, tr)) return -2;
    if (tail < 0 && tr - t2 < SP2_TAIL) {
      tail = SP2_EXTRA;
      if (seed) frontier_seed_tb<SPLIT>(im, bm, n, B);
    }
    if (tail == 0) break;
    const bool sq = fabs(t2 - occ) < fabs(2.0 * tr - t2 - occ);
    double w[2] = {0.0, 0.0};
    tb_sym_step<SPLIT>(n, bm, im, sq, w[0], w[1], B.T);
    block_sum_fenced<2>(w, red);  // tb_sym_step's barrier follows every wave's last read of red
    tr = w[0];
    t2 = w[1];
    if (tail > 0) --tail;
  }
  if (tail != 0) return -1;
  if (fabs(tr - occ) > 1.0e-6) return -2;
  for (int e = tid; e < n * n; e += NT) {
    const int i = e / n, j = e % n;
    Pdst[i * np + j] = M.wt * tb_get<SPLIT>(im, tb_at(bm, i, j));
  }
  __syncthreads();
  return k;
}

__host__ __device__ constexpr int lanczos_lds(int n);
__device__ __noinline__ void lanczos_extremes(const double* Ag, int n, int np, double* v, double& lo_est, double& hi_est);
constexpr double SP2_MARGIN_AGED = 0.03;  // extra margin when the estimate is from an earlier SCF step
constexpr int SP2_REFRESH = 4;         // estimate again every this many steps (and the first two)
constexpr double FRONTIER_SEED_DP = 1.0e-4;  // RMS density change below which purifications seed the frontier runs

__device__ inline int sp2_dispatch(const Mol& M, const Bufs& B, const double* Fu, double* Pdst, double* red, double lo,
                                   double hi, bool seed) {
  const int img = sp2_img(M.n);
  if (M.sp2_lds >= 5 && tb_fits(M.n, M.lds_doubles)) return sp2_run_tb<false>(M, B, Fu, Pdst, red, lo, hi, seed);
  if (M.sp2_lds >= 6) return sp2_run_tb<true>(M, B, Fu, Pdst, red, lo, hi, seed);
  if (M.sp2_lds >= 1 && 2 * img <= M.lds_doubles) return sp2_run<2>(M, B, Fu, Pdst, red, lo, hi, seed);
  if (M.sp2_lds >= 3 && img <= M.lds_doubles && hold_fits(M.n, M.n <= M.sp2_tile16_n ? 16 : 32))
    return sp2_run<3>(M, B, Fu, Pdst, red, lo, hi, seed);
  if (M.sp2_lds >= 2 && img <= M.lds_doubles) return sp2_run<1>(M, B, Fu, Pdst, red, lo, hi, seed);
  if (M.sp2_lds >= 4 && sp2_rows(M.n, M.lds_doubles) >= 16) return sp2_run<4>(M, B, Fu, Pdst, red, lo, hi, seed);
  return sp2_run<0>(M, B, Fu, Pdst, red, lo, hi, seed);
}

// Aufbau density of the symmetric matrix Fu by purification instead of eigenvectors: the
// trace-correcting second-order spectral projection of Niklasson (Phys. Rev. B 66, 155115, 2002).
// X0 = (e_max I - Fu) / (e_max - e_min) maps the spectrum into [0, 1] in reverse order; every
// step replaces X by X^2 or by 2X - X^2, whichever brings tr(X) nearer to n_occ, which drives the
// n_occ lowest eigenvalues of Fu to 1 and the others to 0.  P = 2 X is then the projector
// 2 C_occ C_occ^T of the eigenvector route whenever the HOMO-LUMO gap is open.  tr(X^2) = |X|_F^2
// for symmetric X, so the trace and the norm of each new X are summed in the product's store.  The
// iteration stops SP2_EXTRA products after the idempotency error tr(X) - tr(X^2) has fallen below
// SP2_TAIL (each further product squares the distance of the eigenvalues on one side from 0 or 1).
// Bounds: Gershgorin, or with `tight` (n >= M.sp2_tight_n) the Ritz values of a short Lanczos run
// widened by a margin and clipped to Gershgorin -- a narrower map costs fewer products (about 14
// instead of 20).  The estimate is kept in B.bnd across the SCF steps of one path (the extremes of
// the spectrum move little) and taken again at the first two steps, every SP2_REFRESH steps and
// after a miss.  A bound that misses an eigenvalue shows as tr X^2 > tr X or a wrong final trace;
// the step is then redone with Gershgorin bounds.  The operands live in LDS when they fit
// (sp2_run); P goes to Pdst.  Returns the number of products, or -1 when the error has not fallen
// within SP2_MAX products (frontier orbitals degenerate or nearly so); the caller then diagonalises.
__device__ __noinline__ int sp2_density(const Mol& M, const Bufs& B, const double* Fu, double* Pdst, double* red,
                                        bool tight, bool seed = false) {
  const int n = M.n, np = M.np, nocc = M.nocc, tid = threadIdx.x;
  if (nocc == 0 || nocc == n) {
    for (int e = tid; e < n * n; e += NT) {
      const int i = e / n, j = e % n;
      Pdst[i * np + j] = (nocc == n && i == j) ? M.wt : 0.0;
    }
    __syncthreads();
    return 0;
  }
  double glo, ghi;
  gershgorin(Fu, n, np, red, glo, ghi);
  if (!(ghi - glo > 1.0e-12 * fmax(1.0, fabs(ghi)))) return -1;
  if (tight && n >= M.sp2_tight_n && lanczos_lds(n) <= M.lds_doubles) {
    const double age = B.bnd[2];
    double le = B.bnd[0], he = B.bnd[1], mf = M.sp2_margin + SP2_MARGIN_AGED;
    __syncthreads();  // every thread has read the cache before it changes
    if (age < 2.0 || fmod(age, (double)SP2_REFRESH) == 0.0) {
      lanczos_extremes(Fu, n, np, g_lds, le, he);
      mf = M.sp2_margin;
      if (tid == 0) {
        B.bnd[0] = le;
        B.bnd[1] = he;
      }
    }
    if (tid == 0) B.bnd[2] = age < 0.0 ? 1.0 : age + 1.0;
    const double m = mf * (he - le);
    const double lo = fmax(glo, le - m), hi = fmin(ghi, he + m);
    int r = -1;
    if (hi - lo < ghi - glo && hi - lo > 1.0e-12 * fmax(1.0, fabs(hi))) r = sp2_dispatch(M, B, Fu, Pdst, red, lo, hi, seed);
    if (r < 0 && tid == 0) B.bnd[2] = 0.0;  // a miss: estimate afresh next time
    __syncthreads();
    if (r >= 0) return r;
  }
  const int r = sp2_dispatch(M, B, Fu, Pdst, red, glo, ghi, seed);
  return r >= 0 ? r : -1;
}

// Number of eigenvalues below x of the symmetric tridiagonal matrix (diagonal a, off-diagonal b):
// sign changes of the Sturm sequence q_0 = 1, q_i = (a_i - x) q_{i-1} - b_{i-1}^2 q_{i-2}
// (q_i = det(T_i - x I)); only one multiply-add on the dependent chain per element.  The pair
// (q_{i-1}, q_i) is rescaled by a select every fourth element: over four elements |q| changes by
// far less than the 1e100 headroom on either side, for matrix entries up to about 1e20.
__device__ inline int sturm_below(const double* a, const double* b, int m, double x) {
  double q0 = 1.0, q1 = a[0] - x;
  q1 = q1 == 0.0 ? -1.0e-300 : q1;
  int cnt = q1 < 0.0;
  for (int i = 1; i < m; ++i) {
    double q2 = fma(a[i] - x, q1, -(b[i - 1] * b[i - 1]) * q0);
    q2 = q2 == 0.0 ? (q1 < 0.0 ? 1.0e-300 : -1.0e-300) : q2;
    cnt += (q2 < 0.0) != (q1 < 0.0);
    q0 = q1;
    q1 = q2;
    if ((i & 3) == 0) {
      const double aq = fabs(q1);
      const double f = aq > 1.0e100 ? 1.0e-100 : aq < 1.0e-100 && fabs(q0) < 1.0e-100 ? 1.0e100 : 1.0;
      q0 *= f;
      q1 *= f;
    }
  }
  return cnt;
}

// Largest eigenvalue of a symmetric tridiagonal matrix inside the bracket [lo, hi] by
// multisection: the 64 lanes of the calling wave test 64 points of the bracket per round.
__device__ inline double tridiag_max(const double* a, const double* b, int m, double lo, double hi) {
  const int lane = threadIdx.x & 63;
  const double scale = fmax(fabs(lo), fabs(hi));
  for (int round = 0; round < 24 && hi - lo > 2.0e-16 * scale; ++round) {
    const double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
    const bool above = sturm_below(a, b, m, x) == m;  // x lies above the largest eigenvalue
    const unsigned long long mask = __ballot(above);
    const int first = mask ? __ffsll((long long)mask) - 1 : 64;
    const double nhi = first < 64 ? lo + (hi - lo) * (double)(first + 1) / 65.0 : hi;
    const double nlo = first > 0 ? lo + (hi - lo) * (double)first / 65.0 : lo;
    lo = nlo;
    hi = nhi;
  }
  return 0.5 * (lo + hi);
}

// Residual b[m-1] |s[m-1]| of the top Ritz pair (theta, s), |s| = 1, of the Lanczos tridiagonal:
// s by the forward three-term recurrence from s_0 = 1, which for the largest eigenvalue (above every
// eigenvalue of the leading submatrices) follows the dominant, sign-stable solution.
// The reciprocals 1 / b[i] do not depend on the recurrence, so its chain is multiply-adds only.
__host__ __device__ inline double ritz_residual(const double* a, const double* b, int m, double theta) {
  double s0 = 0.0, s1 = 1.0, ss = 1.0;
  for (int i = 0; i + 1 < m; ++i) {
    const double rb = 1.0 / b[i];
    const double s2 = ((theta - a[i]) * s1 - (i > 0 ? b[i - 1] * s0 : 0.0)) * rb;
    ss += s2 * s2;
    s0 = s1;
    s1 = s2;
    const double f = fabs(s1) > 1.0e100 ? 1.0e-100 : 1.0;
    s0 *= f;
    s1 *= f;
    ss *= f * f;
  }
  return fabs(b[m - 1] * s1) / sqrt(ss);
}

// LDS doubles of one Lanczos run of frontier_levels: q, q_prev, w, the product's partial sums
// (8 slices), the tridiagonal (diagonal, off-diagonal) and reduction slots; the run's matrix comes
// first, as a full square (lanczos_lds_square) or a packed lower triangle (lanczos_lds_packed)
// when it fits, else it stays in HBM (lanczos_lds).
__host__ __device__ constexpr int lanczos_cap(int n) { return 2 * n + 16; }
__host__ __device__ constexpr int lanczos_lds(int n) { return 11 * n + 2 * (lanczos_cap(n) + 2) + 32; }
__host__ __device__ constexpr int lanczos_lds_packed(int n) { return n * (n + 1) / 2 + lanczos_lds(n); }
__host__ __device__ constexpr int lanczos_lds_square(int n) { return n * n + lanczos_lds(n); }
constexpr double LZ_TOL = 1.0e-10;  // Ritz residual (eV) at which a frontier level is taken

// One Lanczos step (three-term recurrence, no reorthogonalisation) on the symmetric n x n matrix A:
// the LDS matrix La -- full square (ld n) when sq, else packed lower triangle (row i at
// i (i + 1) / 2) -- or, when La is null, the np-strided square Ag in HBM.  On entry w holds the next
// basis vector before its normalisation by wn and q the previous one; on exit q = w / wn, w the
// new residual, (alpha, beta) the new tridiagonal entries.  The product runs lanes along rows (64
// rows per wave), each wave over one slice of the columns, the slices' partial sums meeting in LDS:
// no cross-lane reductions inside the product.  v: q, w, part[8][n] at offsets 0, n, 3n; sh: 16
// reduction slots.  Three barriers.
__device__ __forceinline__ void lanczos_step(const double* La, bool sq, const double* Ag, int n, int np, double* v,
                                             double* sh, double wn, double bsub, double& alpha, double& beta) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  double* q = v;
  double* w = q + n;
  double* part = w + 2 * n;
  const int nrb = (n + 63) >> 6, nsl = NW / nrb, rb = wave % nrb, sl = wave / nrb;
  const int row = rb * 64 + lane, slw = (n + nsl - 1) / nsl, c0 = sl * slw, c1 = min(n, c0 + slw);
  if (sl < nsl && row < n) {
    // partial (A w)_row over columns [c0, c1) (A symmetric: A[row][j] read as A[j][row])
    double s0 = 0.0, s1 = 0.0;
    if (La && sq) {
      int j = c0;
      for (; j + 1 < c1; j += 2) {
        s0 += La[j * n + row] * w[j];
        s1 += La[(j + 1) * n + row] * w[j + 1];
      }
      if (j < c1) s0 += La[j * n + row] * w[j];
    } else if (La) {
      for (int j = c0; j < c1; ++j) s0 += La[j <= row ? row * (row + 1) / 2 + j : j * (j + 1) / 2 + row] * w[j];
    } else {
      int j = c0;
      for (; j + 1 < c1; j += 2) {
        s0 += Ag[(size_t)j * np + row] * w[j];
        s1 += Ag[(size_t)(j + 1) * np + row] * w[j + 1];
      }
      if (j < c1) s0 += Ag[(size_t)j * np + row] * w[j];
    }
    part[sl * n + row] = s0 + s1;
  }
  __syncthreads();
  // q_k = w / wn; w = A q_k - beta_{k-1} q_{k-1}; alpha_k = q_k . w
  double av = 0.0;
  for (int i = tid; i < n; i += NT) {
    double pa = 0.0;
    for (int c = 0; c < nsl; ++c) pa += part[c * n + i];
    const double qn = w[i] / wn;
    const double x = pa / wn - bsub * q[i];
    q[i] = qn;
    w[i] = x;
    av += qn * x;
  }
  av = wave_sum(av);
  if (lane == 0) sh[wave] = av;
  __syncthreads();
  alpha = 0.0;
  for (int x = 0; x < NW; ++x) alpha += sh[x];
  double bv = 0.0;
  for (int i = tid; i < n; i += NT) {
    const double x = w[i] - alpha * q[i];
    w[i] = x;
    bv += x * x;
  }
  bv = wave_sum(bv);
  if (lane == 0) sh[8 + wave] = bv;
  __syncthreads();
  double b2 = 0.0;
  for (int x = 0; x < NW; ++x) b2 += sh[8 + x];
  beta = sqrt(b2);
}

// Start vector of the Lanczos runs in w (its norm returned), q zeroed: a fixed spread over all
// basis directions, or, with a seed (side 1: t, side 2: s - t), the normalised seed plus 1e-3 of
// the normalised spread, so that no direction is missing from the start (a seed that is zero or
// not finite falls back to the spread alone).
__device__ inline double lanczos_start(int n, double* v, const double* s = nullptr, const double* t = nullptr,
                                       int side = 0) {
  const int tid = threadIdx.x, lane = tid & 63;
  double* q = v;
  double* w = q + n;
  double ca = 0.0, cg = 1.0;
  if (side != 0) {
    double a2 = 0.0, g2 = 0.0;
    for (int i = lane; i < n; i += 64) {
      const double x = side == 1 ? t[i] : s[i] - t[i];
      const double g = 1.0 + 0.5 * sin(1.7 * (double)i + 0.3);
      a2 += x * x;
      g2 += g * g;
    }
    a2 = wave_sum(a2);  // every wave sums the same vectors
    g2 = wave_sum(g2);
    if (a2 > 1.0e-280 && a2 < 1.0e300) {
      ca = 1.0 / sqrt(a2);
      cg = 1.0e-3 / sqrt(g2);
    }
  }
  for (int i = tid; i < n; i += NT) {
    const double g = 1.0 + 0.5 * sin(1.7 * (double)i + 0.3);
    w[i] = ca != 0.0 ? ca * (side == 1 ? t[i] : s[i] - t[i]) + cg * g : g;
    q[i] = 0.0;
  }
  __syncthreads();
  double nn = 0.0;
  for (int i = lane; i < n; i += 64) nn += w[i] * w[i];
  return sqrt(wave_sum(nn));
}

// Largest eigenvalue of the symmetric matrix A (n x n, n <= 256) by a plain Lanczos run (extreme
// Ritz values stay accurate without reorthogonalisation, C. Paige 1976), steps as lanczos_step.
// Every eighth step from the sixteenth, wave 0 takes the top eigenvalue of the tridiagonal by
// multisection (the previous value is a lower bound: by interlacing it never decreases) and the
// run stops when the Ritz residual is below LZ_TOL or the value has stalled over 8 and 16 steps (a
// converged value gains ghost copies, whose near-degeneracy spoils the residual estimate).
// A seeded start (side != 0, lanczos_start) is checked from the twelfth step on, every fourth.
__device__ __forceinline__ double lanczos_top(const double* La, bool sq, const double* Ag, int n, int np, double scale,
                                              double* v, const double* s = nullptr, const double* t = nullptr,
                                              int side = 0, unsigned long long* steps = nullptr) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int cap = lanczos_cap(n);
  double* al = v + 11 * n;
  double* be = al + (cap + 2);
  double* sh = be + (cap + 2);  // [0, 16) reductions, [16] value, [17] stop
  double wn = lanczos_start(n, v, s, t, side);
  const int first = side != 0 ? 12 : 16, every = side != 0 ? 3 : 7;
  double bsub = 0.0, th1 = -1.0e300, th2 = -1.0e300, top = 0.0;
  for (int k = 0; k < cap; ++k) {
    double alpha, beta;
    lanczos_step(La, sq, Ag, n, np, v, sh, wn, bsub, alpha, beta);
    const int mk = k + 1;
    const bool breakdown = beta <= 1.0e-13 * scale;
    wn = beta;
    bsub = beta;
    const bool check = mk == cap || breakdown || (mk >= first && (mk & every) == 0);
    if (tid == 0) {
      al[k] = alpha;
      be[k] = beta;
    }
    if (!check) continue;
    __syncthreads();
    if (wave == 0) {
      double ghi = -1.0e300, glo = 1.0e300;
      for (int i = lane; i < mk; i += 64) {
        const double r = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i + 1 < mk ? fabs(be[i]) : 0.0);
        ghi = fmax(ghi, al[i] + r);
        glo = fmin(glo, al[i] - r);
      }
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        ghi = fmax(ghi, __shfl_xor(ghi, o, 64));
        glo = fmin(glo, __shfl_xor(glo, o, 64));
      }
      const double sc = fmax(fabs(ghi), fabs(glo));
      double blo = fmax(glo, th1) - 1.0e-14 * sc, bhi = ghi + 1.0e-14 * sc;
      if (th2 > -1.0e299) {
        // the values rise by shrinking steps: try a narrow bracket above the last one first
        const double x = th1 + fmax(4.0 * (th1 - th2), 1.0e-9 * sc);
        if (x < bhi && sturm_below(al, be, mk, x) == mk) bhi = x;
      }
      const double th = tridiag_max(al, be, mk, blo, bhi);
      const bool stall = th - th1 <= 1.0e-13 * scale && th - th2 <= 1.0e-12 * scale;
      const bool stop = mk == cap || breakdown || stall || ritz_residual(al, be, mk, th) <= LZ_TOL;
      th2 = th1;
      th1 = th;
      if (lane == 0) {
        sh[16] = th;
        sh[17] = stop ? 1.0 : 0.0;
      }
    }
    __syncthreads();
    top = sh[16];
    const bool stop = sh[17] != 0.0;
    __syncthreads();
    if (stop) {
      if (steps && tid == 0) *steps += (unsigned long long)mk;
      break;
    }
  }
  return top;
}

// lanczos_top run by one wave, with no workgroup barrier: lane i holds elements i and i + 64 (R = 2,
// n <= 128) of the basis vector q and the residual w in registers; the product reads the operator
// L (leading dimension ld: LDS or an HBM slab matrix) down its columns, w_j given to the wave by the
// lane that holds it.  The tridiagonal al / be (lanczos_cap(n) + 2 doubles each) is the wave's own
// LDS; lane 0 writes it and the whole wave reads it at the checks.  Start vector, recurrence,
// checks and stopping rules are lanczos_top's; only the product's summation order differs (two
// running sums over the columns instead of per-slice sums).
template <int R>
__device__ __noinline__ double lanczos_top_wave(const double* L, int ld, int n, double scale, double* al, double* be,
                                                const double* s, const double* t, int side,
                                                unsigned long long* steps) {
  const int lane = threadIdx.x & 63;
  const int cap = lanczos_cap(n);
  bool live[R];
  int row[R];
  double w[R], q[R];
#pragma unroll
  for (int h = 0; h < R; ++h) {
    const int i = lane + 64 * h;
    live[h] = i < n;
    row[h] = live[h] ? i : 0;  // lanes past n read row 0 and keep zeros
    q[h] = 0.0;
  }
  {
    double x[R], g[R], xx = 0.0, gg = 0.0;
#pragma unroll
    for (int h = 0; h < R; ++h) {
      const int i = lane + 64 * h;
      g[h] = 1.0 + 0.5 * sin(1.7 * (double)i + 0.3);
      x[h] = side == 0 || !live[h] ? 0.0 : side == 1 ? t[i] : s[i] - t[i];
      xx += x[h] * x[h];
      gg += live[h] ? g[h] * g[h] : 0.0;
    }
    double ca = 0.0, cg = 1.0;
    if (side != 0) {
      const double a2 = wave_sum(xx);
      const double g2 = wave_sum(gg);
      if (a2 > 1.0e-280 && a2 < 1.0e300) {
        ca = 1.0 / sqrt(a2);
        cg = 1.0e-3 / sqrt(g2);
      }
    }
#pragma unroll
    for (int h = 0; h < R; ++h) w[h] = !live[h] ? 0.0 : ca != 0.0 ? ca * x[h] + cg * g[h] : g[h];
  }
  double ww = 0.0;
#pragma unroll
  for (int h = 0; h < R; ++h) ww += w[h] * w[h];
  double wn = sqrt(wave_sum(ww));
  const int first = side != 0 ? 12 : 16, every = side != 0 ? 3 : 7;
  double bsub = 0.0, th1 = -1.0e300, th2 = -1.0e300, top = 0.0;
  for (int k = 0; k < cap; ++k) {
    double s0[R], s1[R];
#pragma unroll
    for (int h = 0; h < R; ++h) s0[h] = s1[h] = 0.0;
#pragma unroll
    for (int hj = 0; hj < R; ++hj) {  // columns j held by w[hj]
      const int j0 = 64 * hj, j1 = min(n, j0 + 64);
      int j = j0;
      for (; j + 1 < j1; j += 2) {
        const double wa = lane_value(w[hj], j - j0), wb = lane_value(w[hj], j + 1 - j0);
#pragma unroll
        for (int h = 0; h < R; ++h) {
          s0[h] += L[(size_t)j * ld + row[h]] * wa;
          s1[h] += L[(size_t)(j + 1) * ld + row[h]] * wb;
        }
      }
      if (j < j1) {
        const double wa = lane_value(w[hj], j - j0);
#pragma unroll
        for (int h = 0; h < R; ++h) s0[h] += L[(size_t)j * ld + row[h]] * wa;
      }
    }
    // q_k = w / wn; w = A q_k - beta_{k-1} q_{k-1}; alpha_k = q_k . w; w -= alpha_k q_k
    double qx = 0.0, x[R], qn[R];
#pragma unroll
    for (int h = 0; h < R; ++h) {
      qn[h] = w[h] / wn;
      x[h] = live[h] ? (s0[h] + s1[h]) / wn - bsub * q[h] : 0.0;
      q[h] = qn[h];
      qx += qn[h] * x[h];
    }
    const double alpha = wave_sum(qx);
    double bb = 0.0;
#pragma unroll
    for (int h = 0; h < R; ++h) {
      w[h] = x[h] - alpha * q[h];
      bb += w[h] * w[h];
    }
    const double beta = sqrt(wave_sum(bb));
    const int mk = k + 1;
    const bool breakdown = beta <= 1.0e-13 * scale;
    wn = beta;
    bsub = beta;
    const bool check = mk == cap || breakdown || (mk >= first && (mk & every) == 0);
    if (lane == 0) {
      al[k] = alpha;
      be[k] = beta;
    }
    if (!check) continue;
    // lane 0's LDS stores come before the wave's loads (a wave's LDS operations run in order)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    double ghi = -1.0e300, glo = 1.0e300;
    for (int i = lane; i < mk; i += 64) {
      const double rr = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i + 1 < mk ? fabs(be[i]) : 0.0);
      ghi = fmax(ghi, al[i] + rr);
      glo = fmin(glo, al[i] - rr);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ghi = fmax(ghi, __shfl_xor(ghi, o, 64));
      glo = fmin(glo, __shfl_xor(glo, o, 64));
    }
    const double sc = fmax(fabs(ghi), fabs(glo));
    double blo = fmax(glo, th1) - 1.0e-14 * sc, bhi = ghi + 1.0e-14 * sc;
    if (th2 > -1.0e299) {
      const double xb = th1 + fmax(4.0 * (th1 - th2), 1.0e-9 * sc);
      if (xb < bhi && sturm_below(al, be, mk, xb) == mk) bhi = xb;
    }
    const double th = tridiag_max(al, be, mk, blo, bhi);
    const bool stall = th - th1 <= 1.0e-13 * scale && th - th2 <= 1.0e-12 * scale;
    const bool stop = mk == cap || breakdown || stall || ritz_residual(al, be, mk, th) <= LZ_TOL;
    th2 = th1;
    th1 = th;
    top = th;
    if (stop) {
      if (steps && lane == 0) atomicAdd(steps, (unsigned long long)mk);
      break;
    }
  }
  return top;
}
// Estimates of the lowest and highest eigenvalue of the symmetric np-strided matrix Ag from a
// fixed number of Lanczos steps (Ritz values: lo_est >= lambda_min, hi_est <= lambda_max); LDS
// v as lanczos_lds(n).
constexpr int SP2_LZ_STEPS = 12;
__device__ __noinline__ void lanczos_extremes(const double* Ag, int n, int np, double* v, double& lo_est, double& hi_est) {
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int steps = min(n, SP2_LZ_STEPS);
  double* al = v + 11 * n;
  double* be = al + (steps + 2);
  double* nal = be + (steps + 2);
  double* sh = nal + (steps + 2);
  double wn = lanczos_start(n, v);
  double bsub = 0.0, scale = 0.0;
  int m = 0;
  for (int k = 0; k < steps; ++k) {
    double alpha, beta;
    lanczos_step(nullptr, false, Ag, n, np, v, sh, wn, bsub, alpha, beta);
    if (tid == 0) {
      al[k] = alpha;
      nal[k] = -alpha;
      be[k] = beta;
    }
    m = k + 1;
    scale = fmax(scale, fabs(alpha) + beta);
    if (beta <= 1.0e-13 * scale) break;
    wn = beta;
    bsub = beta;
  }
  __syncthreads();
  if (wave == 0) {
    double ghi = -1.0e300, glo = 1.0e300;
    for (int i = lane; i < m; i += 64) {
      const double r = (i > 0 ? fabs(be[i - 1]) : 0.0) + (i + 1 < m ? fabs(be[i]) : 0.0);
      ghi = fmax(ghi, al[i] + r);
      glo = fmin(glo, al[i] - r);
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      ghi = fmax(ghi, __shfl_xor(ghi, o, 64));
      glo = fmin(glo, __shfl_xor(glo, o, 64));
    }
    const double sc = fmax(fabs(ghi), fabs(glo)) * 1.0e-14 + 1.0e-300;
    const double hi = tridiag_max(al, be, m, glo - sc, ghi + sc);
    const double lo = -tridiag_max(nal, be, m, -ghi - sc, -glo + sc);
    if (lane == 0) {
      sh[16] = lo;
      sh[17] = hi;
    }
  }
  __syncthreads();
  lo_est = sh[16];
  hi_est = sh[17];
  __syncthreads();
}

// HOMO and LUMO of Fu given the projector X = P / 2 onto its n_occ lowest eigenvectors, without
// diagonalising Fu: the HOMO is the largest eigenvalue of A_o = X Fu X + c (I - X) and the LUMO the
// smallest of A_v = (I - X) Fu (I - X) + c' X, with c below and c' above the spectrum, so that the
// complement of each subspace sits at the far end; one Lanczos run each (lanczos_top), A_o written
// straight into LDS (packed) and A_v to HBM, then staged into LDS for its run when it fits.
// Scratch: B.T (Fu X), B.Cn (A_o when it does not fit), B.V (-A_v).  Writes eps[nocc - 1], eps[nocc].
__device__ __noinline__ void frontier_levels(const Mol& M, const Bufs& B, const double* Fu, const double* P, double* lds,
                                             double* red) {
  const int n = M.n, np = M.np, nocc = M.nocc, tid = threadIdx.x;
  if (n == 0) return;
  if (lanczos_lds(n) > M.lds_doubles) {  // the host sizes the launch for this; never taken
    if (tid == 0) {
      if (nocc > 0) B.eps[nocc - 1] = __builtin_nan("");
      if (nocc < n) B.eps[nocc] = __builtin_nan("");
    }
    __syncthreads();
    return;
  }
  // up to 128 AOs both runs go at once, one wave each (one or two vector elements per lane), with
  // their operators square in LDS as far as they fit and in the slab (HBM) otherwise; above, one
  // run after the other over the whole workgroup
  const int cap2 = lanczos_cap(n) + 2;
  const bool waves = n <= 128;
  const bool wv_both = waves && 2 * n * n + 4 * cap2 <= M.lds_doubles;
  const bool wv_one = waves && !wv_both && n * n + 4 * cap2 <= M.lds_doubles;
  const bool square = wv_both || wv_one || (!waves && lanczos_lds_square(n) <= M.lds_doubles);
  const bool packed = square || (!waves && lanczos_lds_packed(n) <= M.lds_doubles);
  const bool av_lds = wv_both;
  double* Lp = packed ? g_lds : nullptr;
  double* Lv = g_lds + n * n;  // -A_v of the one-wave runs
  double* vec = square ? g_lds + n * n : packed ? g_lds + n * (n + 1) / 2 : g_lds;
  // (with n_occ = 0 or n_occ = n one side is empty and the other operator is Fu itself)
  double lo, hi;
  gershgorin(Fu, n, np, red, lo, hi);
  const double c_lo = lo - 1.0 - 0.1 * (hi - lo), c_hi = hi + 1.0 + 0.1 * (hi - lo);
  const double scale = fmax(fabs(c_lo), fabs(c_hi));
  double* G = B.T;
  double* Ao = B.Cn;
  double* Av = B.V;
  const double xs = M.xs;  // X = P / wt
  wg_gemm_pad(n, np, Fu, P, [&](int i, int j, double v) { G[i * np + j] = xs * v; });
  __syncthreads();
  wg_gemm_pad<true>(n, np, P, G, [&](int i, int j, double sv2) {
                  const double sv = xs * sv2;
                  const double x = xs * P[i * np + j];
                  const double id = i == j ? 1.0 : 0.0;
                  const double ao = sv + c_lo * (id - x);
                  if (square) {
                    Lp[i * n + j] = ao;
                  } else if (packed) {
                    if (j <= i) Lp[i * (i + 1) / 2 + j] = ao;
                  } else {
                    Ao[i * np + j] = ao;
                  }
                  const double av = -(Fu[i * np + j] - G[i * np + j] - G[j * np + i] + sv + c_hi * x);
                  if (av_lds) Lv[i * n + j] = av;
                  else Av[i * np + j] = av;
                });
  __syncthreads();
  // seeds from the last purification: t = X s, HOMO start t, LUMO start s - t (X = P / 2)
  const int seeded = (B.bnd && B.bnd[3] != 0.0) ? 1 : 0;
  if (seeded) {
    if (tid < n) {
      double t = 0.0;
      for (int j = 0; j < n; ++j) t += xs * P[j * np + tid] * B.fs[j];
      B.ft[tid] = t;
    }
    __syncthreads();
  }
  unsigned long long* lsteps = B.ph ? &B.ph[10] : nullptr;
  if (B.ph && tid == 0) B.ph[11] += (unsigned long long)((nocc > 0) + (nocc < n));
  if (waves) {
    const int wave = tid >> 6;
    double* tri = g_lds + (wv_both ? 2 * n * n : wv_one ? n * n : 0) + 2 * cap2 * wave;
    const double* Lo = square ? Lp : Ao;
    const double* Lw = av_lds ? Lv : Av;
    const int ldo = square ? n : np, ldw = av_lds ? n : np;
    if (wave == 0 && nocc > 0) {
      const double homo = n <= 64 ? lanczos_top_wave<1>(Lo, ldo, n, scale, tri, tri + cap2, B.fs, B.ft, seeded, lsteps)
                                  : lanczos_top_wave<2>(Lo, ldo, n, scale, tri, tri + cap2, B.fs, B.ft, seeded, lsteps);
      if ((tid & 63) == 0) B.eps[nocc - 1] = homo;
    } else if (wave == 1 && nocc < n) {
      const int side = seeded ? 2 : 0;
      const double lumo = -(n <= 64 ? lanczos_top_wave<1>(Lw, ldw, n, scale, tri, tri + cap2, B.fs, B.ft, side, lsteps)
                                    : lanczos_top_wave<2>(Lw, ldw, n, scale, tri, tri + cap2, B.fs, B.ft, side, lsteps));
      if ((tid & 63) == 0) B.eps[nocc] = lumo;
    }
    __syncthreads();
    return;
  }
  double homo = 0.0, lumo = 0.0;
  if (nocc > 0) homo = lanczos_top(Lp, square, Ao, n, np, scale, vec, B.fs, B.ft, seeded, lsteps);
  if (nocc < n) {
    if (square) {
      for (int e = tid; e < n * n; e += NT) Lp[e] = Av[(e / n) * np + e % n];
      __syncthreads();
    } else if (packed) {
      for (int e = tid; e < n * (n + 1) / 2; e += NT) {
        int i = (int)((sqrt(8.0 * (double)e + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > e) --i;
        while ((i + 1) * (i + 2) / 2 <= e) ++i;
        Lp[e] = Av[i * np + (e - i * (i + 1) / 2)];
      }
      __syncthreads();
    }
    lumo = -lanczos_top(Lp, square, Av, n, np, scale, vec, B.fs, B.ft, seeded ? 2 : 0, lsteps);
  }
  if (tid == 0) {
    if (nocc > 0) B.eps[nocc - 1] = homo;
    if (nocc < n) B.eps[nocc] = lumo;
  }
  __syncthreads();
}

// Superposition density (src/scf.rs:855-881) into Pdst (assumed zero).
__device__ inline void sad_into(const KArgs& A, const Mol& M, double* Pdst) {
  const int np = M.np;
  for (int a = threadIdx.x; a < M.nat; a += NT) {
    const int ga = M.atom0 + a;
    const ElemConst& e = c_tab.el[A.aslot[ga]];
    const int o = A.aoff[ga];
    Pdst[o * np + o] = e.ios;
    if (e.norb == 4)
      for (int k = 1; k < 4; ++k) Pdst[(o + k) * np + o + k] = e.iop / 3.0;
  }
  __syncthreads();
}

// Learned density (definitions in guess_device.hpp) into Pdst (assumed zero): on-site blocks are
// the superposition block plus the messages of the atom's directed pairs in their stored order;
// each pair block is (B(a->b) + B(b->a)^T) / 2, written once from the pair with a < b.
__device__ void learned_into(const KArgs& A, const Mol& M, double* Pdst) {
  const int np = M.np;
  for (int a = threadIdx.x; a < M.nat; a += NT) {
    const int ga = M.atom0 + a, e = A.elem8[ga], o = A.aoff[ga], na = e == 0 ? 1 : 4;
    double Pa[4][4];
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) Pa[r][c] = 0.0;
    Pa[0][0] = pm3guess::sad_s(e);
    for (int k = 1; k < na; ++k) Pa[k][k] = pm3guess::sad_p(e);
    const int j0 = A.gbeg[ga], nb = A.gcnt[ga];
    for (int t = 0; t < nb; ++t) {
      const int j = j0 + t, c = A.gdst[j];
      const double dx = A.ax[c] - A.ax[ga], dy = A.ay[c] - A.ay[ga], dz = A.az[c] - A.az[ga];
      const double r = sqrt(dx * dx + dy * dy + dz * dz);
      double Mm[4][4];
      pm3guess::onsite_message(A.gY + (size_t)j * pm3guess::NO, dx / r, dy / r, dz / r, Mm);
      for (int p = 0; p < na; ++p)
        for (int q = 0; q < na; ++q) Pa[p][q] += Mm[p][q];
    }
    for (int p = 0; p < na; ++p)
      for (int q = 0; q < na; ++q) Pdst[(o + p) * np + o + q] = Pa[p][q];
  }
  for (int t = threadIdx.x; t < M.ngp; t += NT) {
    const long long j = M.gp0 + t;
    const int a = A.gsrc[j], b = A.gdst[j];
    if (a > b) continue;
    const int jp = A.gpart[j];
    const double dx = A.ax[b] - A.ax[a], dy = A.ay[b] - A.ay[a], dz = A.az[b] - A.az[a];
    const double r = sqrt(dx * dx + dy * dy + dz * dz);
    const double ux = dx / r, uy = dy / r, uz = dz / r;
    double Bab[4][4], Bba[4][4];
    pm3guess::pair_block(A.gY + (size_t)j * pm3guess::NO, ux, uy, uz, Bab);
    pm3guess::pair_block(A.gY + (size_t)jp * pm3guess::NO, -ux, -uy, -uz, Bba);
    const int na = A.elem8[a] == 0 ? 1 : 4, nb = A.elem8[b] == 0 ? 1 : 4, oa = A.aoff[a], ob = A.aoff[b];
    for (int p = 0; p < na; ++p)
      for (int q = 0; q < nb; ++q) {
        const double v = 0.5 * (Bab[p][q] + Bba[q][p]);
        Pdst[(oa + p) * np + ob + q] = v;
        Pdst[(ob + q) * np + oa + p] = v;
      }
  }
  __syncthreads();
}

// |F(P) P - P F(P)|_F^2 for the density in B.P (F into B.F, F P into B.T): the first-iteration
// DIIS error norm of an SCF started from B.P.
__device__ double commutator2(const KArgs& A, const Mol& M, const Bufs& B, double* lds, double* red) {
  const int n = M.n, np = M.np;
  build_fock(A, M, M.H, B.P, B.F, lds);
  wg_gemm_pad(n, np, B.F, B.P, [&](int i, int j, double v) { B.T[i * np + j] = v; });
  __syncthreads();
  double v[1] = {0.0};
  for (int e = threadIdx.x; e < n * n; e += NT) {
    const int i = e / n, j = e % n;
    const double d = B.T[i * np + j] - B.T[j * np + i];
    v[0] += d * d;
  }
  block_sum<1>(v, red);
  return v[0];
}

// The symmetry-broken guess's tilt (src/scf.rs:894-916), thread 0: a fifth of an electron from
// p_z to p_x on the first atom with the largest p population (heavy atoms keep their caller order).
__device__ inline void sb_tilt(const KArgs& A, const Mol& M, double* P) {
  const int np = M.np;
  int target = -1;
  double best = 0.0;
  for (int a = 0; a < M.nat; ++a) {
    const ElemConst& e = c_tab.el[A.aslot[M.atom0 + a]];
    if (e.norb == 4 && e.iop > best) { best = e.iop; target = a; }
  }
  if (target >= 0) {
    const int o = A.aoff[M.atom0 + target];
    const double tilt = fmin(0.2, P[(o + 3) * np + o + 3]);
    P[(o + 3) * np + o + 3] -= tilt;
    P[(o + 1) * np + o + 1] += tilt;
  }
}

// Initial density for a guess mode (src/scf.rs:855-935): 1 SAD, 2 core, 3 symmetry-broken,
// 4 learned with the superposition fallback.  The learned density is kept only when its
// first-iteration commutator norm is not larger than the superposition's; it then only builds
// the first Fock matrix, whose aufbau density (orbitals left in B.C for the warm start) starts the
// SCF, so that only idempotent densities enter the DIIS history.  Returns the guess used:
// 0 superposition (or not learned), 1 learned (one pre-iteration done), 2 learned rejected.
__device__ __noinline__ int initial_guess(const KArgs& A, const Mol& M, const Bufs& B, int guess, double* lds,
                                          double* red, bool tight, bool warm = false) {
  const int np = M.np, tid = threadIdx.x;
  for (int e = tid; e < np * np; e += NT) B.P[e] = 0.0;
  __syncthreads();
  if (warm && M.prev) {
    // the molecule's converged density of the previous solve (previous geometry)
    const int n = M.n;
    for (int e = tid; e < n * n; e += NT) B.P[(e / n) * np + e % n] = M.prev[e];
    __syncthreads();
    return 0;
  }
  if (guess == 2) {
    if (A.purify && sp2_density(M, B, M.H, B.P, red, tight) >= 0) return 0;
    eigensolve(M, B, M.H, false, lds, red);
    density(M, B.Cn, B.P);
    return 0;
  }
  if (guess == 4 && M.learn && A.gY) {
    sad_into(A, M, B.P);
    const double cs = commutator2(A, M, B, lds, red);
    for (int e = tid; e < np * np; e += NT) B.P[e] = 0.0;
    __syncthreads();
    learned_into(A, M, B.P);
    const double cl = commutator2(A, M, B, lds, red);  // leaves F(P_learned) in B.F
    if (cl <= cs) {
      if (A.purify) {
        if (sp2_density(M, B, B.F, B.P, red, tight) < 0) {
          eigensolve(M, B, B.F, false, lds, red);
          density(M, B.Cn, B.P);
        }
        return 1;
      }
      eigensolve(M, B, B.F, false, lds, red, A.eig_tol_factor * sqrt(cl));
      density(M, B.Cn, B.P);
      for (int e = tid; e < np * np; e += NT) B.C[e] = B.Cn[e];
      __syncthreads();
      return 1;
    }
    for (int e = tid; e < np * np; e += NT) B.P[e] = 0.0;
    __syncthreads();
    sad_into(A, M, B.P);
    return 2;
  }
  sad_into(A, M, B.P);
  if (guess == 3 && tid == 0) sb_tilt(A, M, B.P);
  __syncthreads();
  return 0;
}

__device__ inline double dot_nn(const double* X, const double* Y, int np, double* red) {
  double v[1] = {0.0};
  for (int e = threadIdx.x; e < np * np; e += NT) v[0] += X[e] * Y[e];
  block_sum<1>(v, red);
  return v[0];
}

// The DIIS history pass of rhf_loop: the new entry in slot D.newslot, its energy, and the Gram
// and A-DIIS entries it adds, with D's bookkeeping (length, mode, error norm) on thread 0; returns
// the electronic energy P.(H + F) / 2 on every thread.  A function of its own so that its sums
// and loads do not compete for registers with the SCF loop's state.
__device__ __noinline__ double diis_history_pass(const KArgs& A, const Mol& M, const Bufs& B, Diis& D, int accel,
                                                  bool keep_d, double* red) {
  const int n = M.n, np = M.np, tid = threadIdx.x;
  const int ns = D.newslot;
  const long long ms = (long long)np * np;
  double* Fn = B.Fh + ns * ms;
  double* En = B.Eh + ns * ms;
  double* Dn = B.Dh + ns * ms;
  // One pass over the packed triangle: the new history entry (F and D symmetric, E
  // antisymmetric with zero diagonal; diagonal weight 1, off-diagonal weight 2 in the inner
  // products), the energy P.(H + F) / 2, and the Gram updates <E_i,E_new>, <D_new,F_j>,
  // <D_i,F_new> for the existing entries plus <E_new,E_new>, <D_new,F_new> (index len).  The
  // <D,F> entries feed only A-DIIS: they are skipped while the previous error is well below
  // the A-DIIS switch, and recomputed in full should A-DIIS be needed later.
  const int npk = n * (n + 1) / 2;
  const int len = D.len;
  const bool want_mm = keep_d && accel == 2 && !D.mm_stale && D.en > 0.25 * A.adiis_switch;
  double v[3 * MAXD + 3];
#pragma unroll
  for (int k = 0; k < 3 * MAXD + 3; ++k) v[k] = 0.0;
  // one packed element: the new entry's values and the older entries' E values are loaded
  // before this element's stores (the history slots are never the new slot)
  struct El {
    double en, fn, pn, hn;
  };
  auto elem = [&](int e, El& x) {
    int i, j;
    tri_unpack(e, n, i, j);  // i <= j; consecutive e run down column j: read (j, i) of symmetric F, P, H
    x.en = (i == j) ? 0.0 : B.T[i * np + j] - B.T[j * np + i];
    x.fn = B.F[j * np + i];
    x.pn = B.P[j * np + i];
    x.hn = M.H[j * np + i];
  };
  auto put = [&](int e, const El& x) {
    En[e] = x.en;
    Fn[e] = x.fn;
    if (keep_d) Dn[e] = x.pn;
  };
  if (want_mm) {  // A-DIIS may be needed: the <D, F> entries as well
    for (int e = tid; e < npk; e += NT) {
      El x;
      elem(e, x);
      double eh[MAXD];
#pragma unroll
      for (int q = 0; q < MAXD; ++q)
        if (q < len) eh[q] = B.Eh[(long long)D.slot[q] * ms + e];
      put(e, x);
      const double wgt = e < n ? 1.0 : 2.0;
      const double dn = keep_d ? x.pn : 0.0;
      const double enw = wgt * x.en, fnw = wgt * x.fn, dnw = wgt * dn;
      v[3 * MAXD + 2] += wgt * x.pn * (x.hn + x.fn);
      v[3 * MAXD] += enw * x.en;
      v[3 * MAXD + 1] += dnw * x.fn;
#pragma unroll
      for (int q = 0; q < MAXD; ++q) {
        if (q < len) {
          const long long so = (long long)D.slot[q] * ms;
          v[q] += eh[q] * enw;
          v[MAXD + q] += dnw * B.Fh[so + e];
          v[2 * MAXD + q] += B.Dh[so + e] * fnw;
        }
      }
    }
    block_sum<3 * MAXD + 3>(v, red);
  } else {
    // the common case, E history only: ten sums (E dots, E.E, energy), which leaves the
    // registers for two elements' loads in flight at once; each sum takes the same terms in the
    // same order as above, so the results are the same bits
    double u[MAXD + 2];
#pragma unroll
    for (int k = 0; k < MAXD + 2; ++k) u[k] = 0.0;
    auto acc = [&](int e, const El& x, const double (&eh)[MAXD]) {
      const double wgt = e < n ? 1.0 : 2.0;
      const double enw = wgt * x.en;
      u[MAXD + 1] += wgt * x.pn * (x.hn + x.fn);
      u[MAXD] += enw * x.en;
#pragma unroll
      for (int q = 0; q < MAXD; ++q)
        if (q < len) u[q] += eh[q] * enw;
    };
    int e = tid;
    for (; e + NT < npk; e += 2 * NT) {
      El x0, x1;
      elem(e, x0);
      elem(e + NT, x1);
      double eh0[MAXD], eh1[MAXD];
#pragma unroll
      for (int q = 0; q < MAXD; ++q)
        if (q < len) {
          const long long so = (long long)D.slot[q] * ms;
          eh0[q] = B.Eh[so + e];
          eh1[q] = B.Eh[so + e + NT];
        }
      put(e, x0);
      put(e + NT, x1);
      acc(e, x0, eh0);
      acc(e + NT, x1, eh1);
    }
    if (e < npk) {
      El x0;
      elem(e, x0);
      double eh0[MAXD];
#pragma unroll
      for (int q = 0; q < MAXD; ++q)
        if (q < len) eh0[q] = B.Eh[(long long)D.slot[q] * ms + e];
      put(e, x0);
      acc(e, x0, eh0);
    }
    block_sum<MAXD + 2>(u, red);
#pragma unroll
    for (int q = 0; q < MAXD; ++q) v[q] = u[q];
    v[3 * MAXD] = u[MAXD];
    v[3 * MAXD + 2] = u[MAXD + 1];
  }
  const double e_elec = 0.5 * v[3 * MAXD + 2];
  stamp(B, 2);
  if (tid == 0) {
    for (int i = 0; i < len; ++i) {
      D.b[i][len] = v[i];
      D.b[len][i] = v[i];
      D.mm[len][i] = v[MAXD + i];
      D.mm[i][len] = v[2 * MAXD + i];
    }
    D.b[len][len] = v[3 * MAXD];
    D.mm[len][len] = v[3 * MAXD + 1];
    D.slot[len] = ns;
    D.len = len + 1;
    if (keep_d && accel == 2 && !want_mm) D.mm_stale = 1;
    const double en = sqrt(v[3 * MAXD]);
    D.en = en;
    if (accel == 2 && en > A.adiis_switch) {
      D.mode = (D.len >= 2 && D.keep_d) ? 2 : 0;
      D.ok = D.mode == 2;
    } else {
      D.mode = 1;
    }
  }
  return e_elec;
}

// sum_i c_i F_i over the kept history into both triangles of Fu (rhf_loop's extrapolation).
__device__ __noinline__ void diis_extrapolate(const Mol& M, const Bufs& B, const Diis& D) {
  const int n = M.n, np = M.np, tid = threadIdx.x;
  const long long ms = (long long)np * np;
  const int npk = n * (n + 1) / 2;
  const int L = D.len;
  for (int e = tid; e < npk; e += NT) {
    double s = 0.0;
    for (int i = 0; i < L; ++i) s += D.coef[i] * B.Fh[(long long)D.slot[i] * ms + e];
    int i, j;
    tri_unpack(e, n, i, j);
    B.Fu[i * np + j] = s;
    B.Fu[j * np + i] = s;
  }
  __syncthreads();
}

// T = F P on the matrix cores (the DIIS error's product), then a barrier.
__device__ __noinline__ void fp_product(const Mol& M, const Bufs& B) {
  const int n = M.n, np = M.np;
  wg_gemm_pad(n, np, B.F, B.P, [&](int i, int j, double v) { B.T[i * np + j] = v; });
  __syncthreads();
}

// The density change of an SCF step and the new density in one pass: returns the RMS change of
// Pn against P, then P <- Pn (damped: (1 - damp) Pn + damp P), and C <- Cn on the eigenvector
// route; block_sum's barriers order the stores before any later read of P.
__device__ __noinline__ double density_step(const KArgs& A, const Mol& M, const Bufs& B, double damp, double* red) {
  const int n = M.n, np = M.np, tid = threadIdx.x;
  double dv[1] = {0.0};
  for (int e = tid; e < np * np; e += NT) {
    const double pn = B.Pn[e], po = B.P[e];
    const double d = pn - po;
    dv[0] += d * d;
    B.P[e] = damp > 0.0 ? (1.0 - damp) * pn + damp * po : pn;
    if (!A.purify) B.C[e] = B.Cn[e];
  }
  block_sum<1>(dv, red);
  return sqrt(dv[0] / (double)max(n * n, 1));
}

struct PathResult {
  double e_elec, last_dp;
  int iters, converged, guess_used;
};

// One restricted SCF path (rhf_loop, src/scf.rs:1179-1316).  Leaves P, eps, Cn (orbitals of the
// last extrapolated Fock) in B; with A.purify each density is the purified aufbau projector of
// the extrapolated Fock matrix and only eps[nocc - 1], eps[nocc] (its frontier levels) are set.
//
// urung: the unrestricted iteration of a closed shell (uhf_loop, src/scf.rs:1319-1496, with
// n_alpha = n_beta and a spin-symmetric guess), which the stability search runs as its last two
// rungs (src/scf.rs:662-673).  Both spin densities start equal (the scaled guess is P/2 each) and
// every step treats them alike, so they stay equal bit for bit and the path is carried here on the
// total density P = 2 P_a.  What differs from the restricted loop: CDIIS only (accel 0 or 1, no
// A-DIIS; the paired Gram of [E_a; E_b] is half the restricted one, and the scale-normalised
// coefficients are the same), a history depth for four matrices per slot, the level shift
// F + shift (I - P/2) on the extrapolated Fock matrix (src/scf.rs:337-350, :1413-1418), the sum of
// the two spin RMS changes (= the RMS change of P), and damping P <- (1 - damp) P_new + damp P
// (:1439-1448).  The undamped aufbau projector of the last step stays in B.Pn for the frontier
// levels; their eigenvalues are those of the shifted matrix, as the reference reports them.
__device__ __noinline__ PathResult rhf_loop(const KArgs& A, const Mol& M, const Bufs& B, int guess, int accel, double* lds,
                               double* red, Diis& D, bool tight = true, bool warm = false, double shift = 0.0,
                               double damp = 0.0, bool urung = false) {
  const int n = M.n, np = M.np, tid = threadIdx.x;
  if (tid == 0) {
    B.bnd[2] = -1.0;  // no spectral-bound estimate for this path yet
    B.bnd[3] = 0.0;   // nor a frontier seed
  }
  __syncthreads();
  const int gused = initial_guess(A, M, B, guess, lds, red, tight, warm);
  // DIIS depth from the memory budget (src/scf.rs:309-327, :1205-1228)
  const long long n2 = (long long)n * n * 8;
  const long long budget = (long long)A.scf_memory_mb * 1024 * 1024;
  if (urung && accel == 2) accel = 1;
  int keep_d = accel == 2;
  if (keep_d && !(A.scf_memory_mb == 0 || 2 * 3 * n2 <= budget)) keep_d = 0;
  if (accel == 2 && !keep_d) accel = 1;
  const long long per = (urung ? 4 : keep_d ? 3 : 2) * n2;
  int depth = MAXD;
  if (A.scf_memory_mb != 0 && per != 0) depth = (int)max(2LL, min((long long)MAXD, budget / per));
  if (tid == 0) {
    D.len = 0;
    D.depth = depth;
    D.keep_d = keep_d;
    D.mm_stale = 0;
    D.en = 1.0e300;
  }
  __syncthreads();
  PathResult R{0.0, 1.0e300, 0, 0, gused};
  double e_old = 0.0;
  bool have_c = gused == 1 && !A.purify;
  double eig_tol = 0.0;  // per-iteration eigensolver tolerance (DIIS error norm scaled)
  const double* Flast = B.F;  // the matrix of the last aufbau step
  for (int it = 0; it < A.max_scf; ++it) {
    R.iters = it + 1;
    stamp(B, 7);
    build_fock(A, M, M.H, B.P, B.F, lds);
    stamp(B, 1);
    double e_elec = 0.0;
    const double* Fuse = B.F;
    if (accel == 0) {
      double ev[1] = {0.0};
      for (int e = tid; e < np * np; e += NT) ev[0] += B.P[e] * (M.H[e] + B.F[e]);
      block_sum<1>(ev, red);
      e_elec = 0.5 * ev[0];
    } else {
      // choose the slot of the new entry (drop the oldest when full)
      if (tid == 0) {
        if (D.len == D.depth) {
          const int s0 = D.slot[0];
          for (int i = 0; i + 1 < D.len; ++i) {
            D.slot[i] = D.slot[i + 1];
            for (int j = 0; j + 1 < D.len; ++j) {
              D.b[i][j] = D.b[i + 1][j + 1];
              D.mm[i][j] = D.mm[i + 1][j + 1];
            }
          }
          D.len -= 1;
          D.newslot = s0;
        } else {
          D.newslot = D.len;
        }
      }
      __syncthreads();
      const int ns = D.newslot;
      const long long ms = (long long)np * np;
      double* Fn = B.Fh + ns * ms;
      double* En = B.Eh + ns * ms;
      double* Dn = B.Dh + ns * ms;
      // E = F P - P F (src/scf.rs:1039-1054) via T = F P
      stamp(B, 2);
      fp_product(M, B);
      stamp(B, 3);
      const int npk = n * (n + 1) / 2;
      e_elec = diis_history_pass(A, M, B, D, accel, keep_d, red);
      __syncthreads();
      if (D.mode == 2 && D.mm_stale) {
        // A-DIIS after skipped steps: every <D_i, F_j> again, summed as in the pass above
        const int L = D.len;
        for (int a = 0; a < L; ++a) {
          double w[MAXD];
#pragma unroll
          for (int q = 0; q < MAXD; ++q) w[q] = 0.0;
          const double* Da = B.Dh + (long long)D.slot[a] * ms;
          for (int e = tid; e < npk; e += NT) {
            const double dw = (e < n ? 1.0 : 2.0) * Da[e];
#pragma unroll
            for (int q = 0; q < MAXD; ++q)
              if (q < L) w[q] += dw * B.Fh[(long long)D.slot[q] * ms + e];
          }
          block_sum<MAXD>(w, red);
          if (tid == 0)
            for (int q = 0; q < L; ++q) D.mm[a][q] = w[q];
          __syncthreads();
        }
        if (tid == 0) D.mm_stale = 0;
      }
      __syncthreads();
      if (D.mode != 0) {  // coefficients on wave 0
        if (tid < 64) {
          if (D.mode == 2) {
            adiis_wave(D, red);
          } else {
            const bool ok = cdiis_wave(D, D.coef);
            if (tid == 0) D.ok = ok;
          }
        }
        __syncthreads();
      }
      stamp(B, 8);
      if (D.ok) {
        diis_extrapolate(M, B, D);
        Fuse = B.Fu;
      } else {
        Fuse = B.F;  // the newest entry's Fock matrix
      }
      stamp(B, 9);
    }
    stamp(B, 2);
    if (shift != 0.0) {  // level shift on the virtual space, F + shift (I - P / 2)
      for (int e = tid; e < n * n; e += NT) {
        const int i = e / n, j = e % n;
        B.Fu[i * np + j] = Fuse[i * np + j] + shift * ((i == j ? 1.0 : 0.0) - 0.5 * B.P[i * np + j]);
      }
      __syncthreads();
      Fuse = B.Fu;
    }
    Flast = Fuse;
    if (A.purify) {
      // the frontier seed is only needed from the path's last purification: take it once the SCF is
      // close to converged
      const int ns = sp2_density(M, B, Fuse, B.Pn, red, tight, R.last_dp < FRONTIER_SEED_DP);
      if (ns < 0) {  // no open gap: diagonalise
        eigensolve(M, B, Fuse, false, lds, red);
        density(M, B.Cn, B.Pn);
      }
      if (B.ph && tid == 0) {
        B.ph[PH_LAST + 1] += (unsigned long long)max(ns, 0);
        B.ph[PH_LAST + 2] += 1ull;
        B.ph[PH_LAST + 3] += ns < 0 ? 1ull : 0ull;
      }
      stamp(B, 4);
    } else {
      eig_tol = accel != 0 ? A.eig_tol_factor * D.en : 0.0;
      eigensolve(M, B, Fuse, have_c, lds, red, eig_tol);
      density(M, B.Cn, B.Pn);
    }
    const double dp = density_step(A, M, B, damp, red);
    const double de = fabs(e_elec - e_old);
    R.last_dp = dp;
    have_c = !A.purify;
    stamp(B, 6);
    e_old = e_elec;
    if (it > 0 && de < A.e_tol && dp < A.p_tol) {
      R.converged = 1;
      break;
    }
  }
  if (A.purify) {
    frontier_levels(M, B, Flast, damp > 0.0 ? B.Pn : B.P, lds, red);
    stamp(B, 5);
  }
  build_fock(A, M, M.H, B.P, B.F, lds);
  double ev[1] = {0.0};
  for (int e = tid; e < np * np; e += NT) ev[0] += B.P[e] * (M.H[e] + B.F[e]);
  block_sum<1>(ev, red);
  R.e_elec = 0.5 * ev[0];
  if (gused == 1) R.iters += 1;  // the learned pre-iteration
  return R;
}

// ------------------------------------------------------------------------------------------
// The unrestricted path (uhf_loop, src/scf.rs:1319-1496) for s/p elements.  Each spin channel is a
// Bufs view (alpha: the workgroup slab; beta: the slab's beta part) with its own density, Fock
// matrices, history and orbitals; scratch (T, V, eigen work) is shared, the channels run one after
// the other.  A channel's Mol has n_occ = its electron count and weight-1 densities.
__device__ inline Mol spin_mol(const Mol& M, int nocc) {
  Mol S = M;
  S.nocc = nocc;
  S.wt = 1.0;
  S.xs = 1.0;
  return S;
}

// Aufbau density of the channel's (extrapolated, shifted) Fock matrix Fu into Pdst: purified, or
// from eigenvectors (then eps and Cn hold the channel's orbitals).
__device__ inline int channel_aufbau(const KArgs& A, const Mol& Mc, const Bufs& Bc, const double* Fu, double* Pdst,
                                     double* lds, double* red, bool tight, bool seed, bool warm) {
  if (A.purify) {
    const int ns = sp2_density(Mc, Bc, Fu, Pdst, red, tight, seed);
    if (ns < 0) {
      eigensolve(Mc, Bc, Fu, false, lds, red);
      density(Mc, Bc.Cn, Pdst);
    }
    return ns;
  }
  eigensolve(Mc, Bc, Fu, warm, lds, red);
  density(Mc, Bc.Cn, Pdst);
  return 0;
}

// One history entry of a channel: E = F P - P F (through T = F P) and F, packed triangles.
__device__ inline void uhf_entry(const Mol& M, const Bufs& B, const double* F, const double* P, double* Fn,
                                 double* En) {
  const int n = M.n, np = M.np, tid = threadIdx.x;
  wg_gemm_pad(n, np, F, P, [&](int i, int j, double v) { B.T[i * np + j] = v; });
  __syncthreads();
  for (int e = tid; e < n * (n + 1) / 2; e += NT) {
    int i, j;
    tri_unpack(e, n, i, j);
    En[e] = (i == j) ? 0.0 : B.T[i * np + j] - B.T[j * np + i];
    Fn[e] = F[j * np + i];
  }
  __syncthreads();
}

// Fu = the combination sum_i c_i F_i of a channel's history (unpacked, symmetric).
__device__ inline void uhf_combine(const Mol& M, const Bufs& B, const Diis& D) {
  const int n = M.n, np = M.np;
  const long long ms = (long long)np * np;
  for (int e = threadIdx.x; e < n * (n + 1) / 2; e += NT) {
    double v = 0.0;
    for (int i = 0; i < D.len; ++i) v += D.coef[i] * B.Fh[(long long)D.slot[i] * ms + e];
    int i, j;
    tri_unpack(e, n, i, j);
    B.Fu[i * np + j] = v;
    B.Fu[j * np + i] = v;
  }
  __syncthreads();
}

// out = F + shift (I - P) for the channel (level_shift, src/scf.rs:337-350).
__device__ inline void uhf_shift(const Mol& M, const double* F, const double* P, double* out, double shift) {
  const int n = M.n, np = M.np;
  for (int e = threadIdx.x; e < n * n; e += NT) {
    const int i = e / n, j = e % n;
    out[i * np + j] = F[i * np + j] + shift * ((i == j ? 1.0 : 0.0) - P[i * np + j]);
  }
  __syncthreads();
}

// (P_t . H + P_a . F_a + P_b . F_b) / 2 over the padded matrices.
__device__ inline double uhf_energy(const Mol& M, const double* Pt, const double* Pa, const double* Fa,
                                    const double* Pb, const double* Fb, double* red) {
  double v[1] = {0.0};
  for (int e = threadIdx.x; e < M.np * M.np; e += NT) v[0] += Pt[e] * M.H[e] + Pa[e] * Fa[e] + Pb[e] * Fb[e];
  block_sum<1>(v, red);
  return 0.5 * v[0];
}

// One unrestricted SCF path, uhf_loop (src/scf.rs:1319-1496): the guess of uhf_guess
// (:937-980: core-Hamiltonian orbitals filled to each depth, or the superposition / symmetry-broken
// density scaled by n_a / N and n_b / N); per iteration the two spin Fock matrices from
// (P_a + P_b, P_sigma), the energy (P_t . H + P_a . F_a + P_b . F_b) / 2, paired CDIIS on the stacked
// error [E_a; E_b] only when accel = 1 (CDIIS requested and use_diis), depth for four matrices per
// slot; the level shift F + shift (I - P_sigma); aufbau densities of weight 1; the sum of the two
// RMS density changes; damping.  Leaves P_a, P_b in Ba.P, Bb.P, each channel's frontier levels (or
// orbitals) in its eps, and the total density in Pt.
__device__ PathResult uhf_loop(const KArgs& A, const Mol& M, const Bufs& Ba, const Bufs& Bb, double* Pt, int guess,
                               int accel, double shift, double damp, double* lds, double* red, Diis& D, bool tight) {
  const int n = M.n, np = M.np, tid = threadIdx.x;
  const int na = M.nocc, nb = M.nbeta;
  const Mol Ma = spin_mol(M, na), Mb = spin_mol(M, nb);
  if (tid == 0) {
    Ba.bnd[2] = -1.0;
    Ba.bnd[3] = 0.0;
    Bb.bnd[2] = -1.0;
    Bb.bnd[3] = 0.0;
  }
  for (int e = tid; e < np * np; e += NT) {
    Ba.P[e] = 0.0;
    Bb.P[e] = 0.0;
    Pt[e] = 0.0;
  }
  __syncthreads();
  if (guess == 2) {
    channel_aufbau(A, Ma, Ba, M.H, Ba.P, lds, red, tight, false, false);
    channel_aufbau(A, Mb, Bb, M.H, Bb.P, lds, red, tight, false, false);
  } else {
    sad_into(A, M, Pt);
    if (guess == 3 && tid == 0) sb_tilt(A, M, Pt);
    __syncthreads();
    const double ne = (double)(na + nb);
    const double fa = ne > 0.0 ? na / ne : 0.5, fb = ne > 0.0 ? nb / ne : 0.5;
    for (int e = tid; e < np * np; e += NT) {
      Ba.P[e] = Pt[e] * fa;
      Bb.P[e] = Pt[e] * fb;
    }
    __syncthreads();
  }
  const bool cdiis = accel == 1;
  {
    const long long per = 4LL * n * n * 8, budget = (long long)A.scf_memory_mb * 1024 * 1024;
    int depth = MAXD;
    if (A.scf_memory_mb != 0 && per != 0) depth = (int)max(2LL, min((long long)MAXD, budget / per));
    if (tid == 0) {
      D.len = 0;
      D.depth = depth;
      D.keep_d = 0;
      D.mm_stale = 0;
      D.en = 1.0e300;
      D.ok = 0;
    }
    __syncthreads();
  }
  PathResult R{0.0, 1.0e300, 0, 0, 0};
  double e_old = 0.0;
  bool have_c = false;
  const double* Fla = Ba.F;  // the matrices of the last aufbau step
  const double* Flb = Bb.F;
  for (int it = 0; it < A.max_scf; ++it) {
    R.iters = it + 1;
    for (int e = tid; e < np * np; e += NT) Pt[e] = Ba.P[e] + Bb.P[e];
    __syncthreads();
    build_fock_x(A, M, M.H, Pt, Ba.P, 1.0, Ba.F, lds);
    build_fock_x(A, M, M.H, Pt, Bb.P, 1.0, Bb.F, lds);
    const double e_elec = uhf_energy(M, Pt, Ba.P, Ba.F, Bb.P, Bb.F, red);
    const double* Fa = Ba.F;
    const double* Fb = Bb.F;
    if (cdiis) {
      if (tid == 0) {
        if (D.len == D.depth) {
          const int s0 = D.slot[0];
          for (int i = 0; i + 1 < D.len; ++i) {
            D.slot[i] = D.slot[i + 1];
            for (int j = 0; j + 1 < D.len; ++j) D.b[i][j] = D.b[i + 1][j + 1];
          }
          D.len -= 1;
          D.newslot = s0;
        } else {
          D.newslot = D.len;
        }
      }
      __syncthreads();
      const int ns = D.newslot, len = D.len;
      const long long ms = (long long)np * np;
      uhf_entry(M, Ba, Ba.F, Ba.P, Ba.Fh + ns * ms, Ba.Eh + ns * ms);
      uhf_entry(M, Bb, Bb.F, Bb.P, Bb.Fh + ns * ms, Bb.Eh + ns * ms);
      // Gram row of the stacked error: <E_a,i, E_a,new> + <E_b,i, E_b,new> (off-diagonal weight 2)
      double v[MAXD + 1];
#pragma unroll
      for (int q = 0; q <= MAXD; ++q) v[q] = 0.0;
      const double* Ea = Ba.Eh + ns * ms;
      const double* Eb = Bb.Eh + ns * ms;
      for (int e = tid; e < n * (n + 1) / 2; e += NT) {
        const double w = e < n ? 1.0 : 2.0;
        const double ea = Ea[e], eb = Eb[e];
        v[MAXD] += w * (ea * ea + eb * eb);
#pragma unroll
        for (int q = 0; q < MAXD; ++q)
          if (q < len) {
            const long long so = (long long)D.slot[q] * ms;
            v[q] += w * (Ba.Eh[so + e] * ea + Bb.Eh[so + e] * eb);
          }
      }
      block_sum<MAXD + 1>(v, red);
      if (tid == 0) {
        for (int i = 0; i < len; ++i) {
          D.b[i][len] = v[i];
          D.b[len][i] = v[i];
        }
        D.b[len][len] = v[MAXD];
        D.slot[len] = ns;
        D.len = len + 1;
        D.en = sqrt(v[MAXD]);
        D.mode = 1;
      }
      __syncthreads();
      if (tid < 64) {
        const bool ok = cdiis_wave(D, D.coef);
        if (tid == 0) D.ok = ok;
      }
      __syncthreads();
      if (D.ok) {
        uhf_combine(M, Ba, D);
        uhf_combine(M, Bb, D);
        Fa = Ba.Fu;
        Fb = Bb.Fu;
      }
    }
    if (shift != 0.0) {
      uhf_shift(M, Fa, Ba.P, Ba.Fu, shift);
      uhf_shift(M, Fb, Bb.P, Bb.Fu, shift);
      Fa = Ba.Fu;
      Fb = Bb.Fu;
    }
    Fla = Fa;
    Flb = Fb;
    const bool seed = R.last_dp < FRONTIER_SEED_DP;
    channel_aufbau(A, Ma, Ba, Fa, Ba.Pn, lds, red, tight, seed, have_c);
    channel_aufbau(A, Mb, Bb, Fb, Bb.Pn, lds, red, tight, seed, have_c);
    double dv[2] = {0.0, 0.0};
    for (int e = tid; e < np * np; e += NT) {
      const double da = Ba.Pn[e] - Ba.P[e], db = Bb.Pn[e] - Bb.P[e];
      dv[0] += da * da;
      dv[1] += db * db;
    }
    block_sum<2>(dv, red);
    const double nn = (double)max(n * n, 1);
    const double dp = sqrt(dv[0] / nn) + sqrt(dv[1] / nn);
    const double de = fabs(e_elec - e_old);
    R.last_dp = dp;
    for (int e = tid; e < np * np; e += NT) {
      Ba.P[e] = damp > 0.0 ? (1.0 - damp) * Ba.Pn[e] + damp * Ba.P[e] : Ba.Pn[e];
      Bb.P[e] = damp > 0.0 ? (1.0 - damp) * Bb.Pn[e] + damp * Bb.P[e] : Bb.Pn[e];
      if (!A.purify) {
        Ba.C[e] = Ba.Cn[e];
        Bb.C[e] = Bb.Cn[e];
      }
    }
    __syncthreads();
    have_c = !A.purify;
    e_old = e_elec;
    if (it > 0 && de < A.e_tol && dp < A.p_tol) {
      R.converged = 1;
      break;
    }
  }
  if (A.purify) {
    frontier_levels(Ma, Ba, Fla, damp > 0.0 ? Ba.Pn : Ba.P, lds, red);
    frontier_levels(Mb, Bb, Flb, damp > 0.0 ? Bb.Pn : Bb.P, lds, red);
  }
  for (int e = tid; e < np * np; e += NT) Pt[e] = Ba.P[e] + Bb.P[e];
  __syncthreads();
  build_fock_x(A, M, M.H, Pt, Ba.P, 1.0, Ba.F, lds);
  build_fock_x(A, M, M.H, Pt, Bb.P, 1.0, Bb.F, lds);
  R.e_elec = uhf_energy(M, Pt, Ba.P, Ba.F, Bb.P, Bb.F, red);
  return R;
}

// HOMO (max over channels) and LUMO (min) of an unrestricted state (src/scf.rs:106-124, 797-820);
// +-1e300 for an empty side.
__device__ inline void uhf_frontier(const Mol& M, const Bufs& Ba, const Bufs& Bb, double& homo, double& lumo) {
  const int n = M.n, na = M.nocc, nb = M.nbeta;
  homo = -1.0e300;
  lumo = 1.0e300;
  if (na > 0) homo = fmax(homo, Ba.eps[na - 1]);
  if (nb > 0) homo = fmax(homo, Bb.eps[nb - 1]);
  if (na < n) lumo = fmin(lumo, Ba.eps[na]);
  if (nb < n) lumo = fmin(lumo, Bb.eps[nb]);
}

// A purified unrestricted path whose channel ended with its HOMO above its LUMO (a missed spectral
// bound): run again with Gershgorin bounds.
__device__ inline bool uhf_aufbau_broken(const KArgs& A, const Mol& M, const Bufs& Ba, const Bufs& Bb) {
  if (!A.purify) return false;
  auto bad = [&](const Bufs& B, int occ) {
    if (occ <= 0 || occ >= M.n) return false;
    const double h = B.eps[occ - 1], l = B.eps[occ];
    return h > l + 1.0e-9 * fmax(1.0, fabs(l));
  };
  return bad(Ba, M.nocc) || bad(Bb, M.nbeta);
}

// With purified densities: a final HOMO above the LUMO means some eigenvalue ended on the wrong
// side of the purification (a spectral bound that missed it); the path is then run again with
// Gershgorin bounds throughout.
__device__ inline bool aufbau_broken(const KArgs& A, const Mol& M, const Bufs& B) {
  const int n = M.n, nocc = M.nocc;
  if (!A.purify || nocc <= 0 || nocc >= n) return false;
  const double h = B.eps[nocc - 1], l = B.eps[nocc];
  return h > l + 1.0e-9 * fmax(1.0, fabs(l));
}

__device__ inline Bufs slab(const KArgs& A, double*& Wbase) {
  double* base = A.scratch + (long long)blockIdx.x * A.stride;
  const long long m2 = (long long)A.np_max * A.np_max;
  Bufs B;
  B.ph = nullptr;   // phase timers off unless the kernel sets them
  B.bnd = nullptr;  // set by the kernel to its LDS slots
  B.H = base;
  B.F = B.H + m2;
  B.P = B.F + m2;
  B.Pn = B.P + m2;
  B.Fu = B.Pn + m2;
  B.T = B.Fu + m2;
  B.C = B.T + m2;
  B.Cn = B.C + m2;
  B.V = B.Cn + m2;
  B.Pb = B.V + m2;
  B.Cb = B.Pb + m2;
  B.Fh = B.Cb + m2;
  B.Eh = B.Fh + MAXD * m2;
  B.Dh = B.Eh + MAXD * m2;
  Wbase = B.Dh + MAXD * m2;
  B.eps = Wbase + A.w_cap;
  B.epsb = B.eps + A.np_max;
  B.lam = B.epsb + A.np_max;
  B.perm = (int*)(B.lam + A.np_max);
  B.fs = B.lam + 2 * A.np_max;
  B.ft = B.fs + A.np_max;
  return B;
}

// The beta channel's part of the workgroup slab (its last A.ustride doubles): P_total, F, P, Pn,
// Fu, C, Cn, the incumbent P and C of the stability search, the history F[8] and E[8], then eps,
// the incumbent's eps and the frontier seeds.  Scratch matrices (T, V) and the eigensolver's work
// vectors are the alpha channel's.
__device__ inline Bufs beta_slab(const KArgs& A, const Bufs& Ba, double*& Pt) {
  double* x = A.scratch + (long long)blockIdx.x * A.stride + (A.stride - A.ustride);
  const long long m2 = (long long)A.np_max * A.np_max;
  Bufs B = Ba;
  Pt = x;
  B.F = x + m2;
  B.P = B.F + m2;
  B.Pn = B.P + m2;
  B.Fu = B.Pn + m2;
  B.C = B.Fu + m2;
  B.Cn = B.C + m2;
  B.Pb = B.Cn + m2;
  B.Cb = B.Pb + m2;
  B.Fh = B.Cb + m2;
  B.Eh = B.Fh + MAXD * m2;
  B.eps = B.Eh + MAXD * m2;
  B.epsb = B.eps + A.np_max;
  B.fs = B.epsb + A.np_max;
  B.ft = B.fs + A.np_max;
  return B;
}

__host__ __device__ inline Mol mol_of(const KArgs& A, const MolDesc& md, double* Wbase) {
  Mol M;
  M.atom0 = md.atom0;
  M.prev = A.warm && md.out_mat >= 0 ? A.out_P + md.out_mat : nullptr;
  M.nat = md.nat;
  M.nh = md.nheavy;
  M.nH = md.nat - md.nheavy;
  M.n = md.nao;
  M.np = md.np;
  M.nocc = md.nocc;
  M.nbeta = md.nbeta;
  M.wt = 2.0;
  M.xs = 0.5;
  M.ngp = md.ngp;
  M.learn = md.learn;
  M.gp0 = md.gp0;
  M.nxx = M.nh * (M.nh - 1) / 2;
  M.nxh = M.nh * M.nH;
  M.nhh = M.nH * (M.nH - 1) / 2;
  const long long npairs = (long long)md.nat * (md.nat - 1) / 2;
  M.H = A.store + md.h_off;
  M.Wxx = A.store + md.w_off;
  M.Wxh = M.Wxx + 100LL * M.nxx;
  M.Whh = M.Wxh + 10LL * M.nxh;
  M.S1 = M.Whh + M.nhh;
  M.S2 = M.S1 + 10 * npairs;
  M.ecc = M.S2 + 10 * npairs;
  M.slot = Wbase;
  M.lds_doubles = A.lds_doubles;
  M.eig_rel = A.eig_rel;
  M.eig_skip = A.eig_skip;
  M.sp2_lds = A.sp2_lds;
  M.sp2_tight_n = A.sp2_tight_n;
  M.sp2_tile16_n = A.sp2_tile16_n;
  M.sp2_margin = A.sp2_margin;
  return M;
}

__device__ inline void zero_mats(const Mol& M, const Bufs& B) {
  const long long mm = (long long)M.np * M.np;
  for (long long e = threadIdx.x; e < mm; e += NT) {
    B.F[e] = 0.0; B.P[e] = 0.0; B.Pn[e] = 0.0; B.Fu[e] = 0.0;
    B.T[e] = 0.0; B.C[e] = 0.0; B.Cn[e] = 0.0;
  }
}

#ifndef PM3_KERNEL_VARIANT  // kernels launched by the main build only
// All atom pairs of all molecules, one work-item each (grid-stride).
__global__ __launch_bounds__(256) void pair_kernel(KArgs A) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < A.npairs_total; g += stride) {
    int lo = 0, hi = A.nmol_pairs;  // largest queue position with pair_base <= g
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (A.pair_base[mid] <= g) lo = mid; else hi = mid;
    }
    const Mol M = mol_of(A, A.mols[A.order[lo]], nullptr);
    pair_integrals(A, M, (int)(g - A.pair_base[lo]));
  }
}


// ------------------------------------------------------------------------------------------
// Analytic gradient at the converged density (closed_form_gradient, src/gradient.rs:201-276;
// electronic part src/gradient.rs:423-489, core-core src/repulsion.rs:30-49).  One work-item per
// pair evaluates the pair's integrals in dual numbers seeded on d = R_b - R_a and contracts the
// derivatives with the density on the fly:
//   dE/dR_b = sum P_ab (beta_a + beta_b) dS + sum P_aa de1b + sum P_bb de2a
//           + sum (P_aa P_bb - 1/2 P_ab P_ab) dW + dE_core
// The per-pair force goes to a slot; grad_gather sums them per atom in a fixed order.
// With a spin density S (unrestricted result) the exchange term is the spin-resolved one of
// electronic_gradient_fixed_density_spin (src/gradient.rs:523-599): -(P_a P_a + P_b P_b) on the
// off-diagonal block with P_a = (P + S) / 2, P_b = (P - S) / 2.
__device__ inline void pair_gradient(const KArgs& A, const Mol& M, const double* P, const double* S, int n, int p,
                                     double* out) {
  const int nat = M.nat;
  int a, b;
  tri_decode(p, nat, a, b);
  const int ga = M.atom0 + a, gb = M.atom0 + b;
  const int sa = A.aslot[ga], sb = A.aslot[gb];
  const ElemConst& ea = c_tab.el[sa];
  const ElemConst& eb = c_tab.el[sb];
  const int oa = A.aoff[ga], ob = A.aoff[gb];
  const int na = ea.norb, nb = eb.norb;
  const D3 dx = mk(A.ax[gb] - A.ax[ga], 1.0, 0.0, 0.0);
  const D3 dy = mk(A.ay[gb] - A.ay[ga], 0.0, 1.0, 0.0);
  const D3 dz = mk(A.az[gb] - A.az[ga], 0.0, 0.0, 1.0);
  const D3 r = sqrtx(dx * dx + dy * dy + dz * dz);
  const D3 inv = recip(r);
  double f0 = 0.0, f1 = 0.0, f2 = 0.0;
  auto acc = [&](double c, const D3& v) { f0 += c * v.d0; f1 += c * v.d1; f2 += c * v.d2; };
  // core-core
  acc(1.0, core_core_pair<D3>(c_tab, sa, sb, r));
  // density blocks (P_aa, P_bb, P_ab) of the pair
  double paa[4][4], pbb[4][4], pab[4][4], sab[4][4];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      paa[i][j] = (i < na && j < na) ? P[(oa + i) * n + oa + j] : 0.0;
      pbb[i][j] = (i < nb && j < nb) ? P[(ob + i) * n + ob + j] : 0.0;
      pab[i][j] = (i < na && j < nb) ? P[(oa + i) * n + ob + j] : 0.0;
      sab[i][j] = (S && i < na && j < nb) ? S[(oa + i) * n + ob + j] : 0.0;
    }
  // electron-core attraction blocks
  {
    D3 e1b[10], e2a[10];
    ecore(ea, eb, dx, dy, dz, e1b);
    ecore(eb, ea, -dx, -dy, -dz, e2a);
    for (int mu = 0; mu < na; ++mu)
      for (int nu = 0; nu <= mu; ++nu) acc(mu == nu ? paa[mu][mu] : paa[mu][nu] + paa[nu][mu], e1b[pk(mu, nu)]);
    for (int mu = 0; mu < nb; ++mu)
      for (int nu = 0; nu <= mu; ++nu) acc(mu == nu ? pbb[mu][mu] : pbb[mu][nu] + pbb[nu][mu], e2a[pk(mu, nu)]);
  }
  // resonance
  {
    D3 S[4][4];
    overlap_block(ea, eb, dx * inv, dy * inv, dz * inv, r, S);
    for (int i = 0; i < na; ++i) {
      const double bi = i == 0 ? ea.betas : ea.betap;
      for (int j = 0; j < nb; ++j) acc(pab[i][j] * (bi + (j == 0 ? eb.betas : eb.betap)), S[i][j]);
    }
  }
  // two-electron integrals
  D3 R[3][3];
  rot_to_x(-(dx * inv), -(dy * inv), -(dz * inv), R);
  // coefficient of the packed element (mu nu | la si): both orderings of each index pair
  auto coef = [&](int mu, int nu, int la, int si) {
    double c = 0.0;
    const int m2 = mu == nu ? 1 : 2, l2 = la == si ? 1 : 2;
    for (int x = 0; x < m2; ++x)
      for (int y = 0; y < l2; ++y) {
        const int u = x ? nu : mu, v = x ? mu : nu, s = y ? si : la, t = y ? la : si;
        if (S) {
          const double a1 = 0.5 * (pab[u][s] + sab[u][s]), a2 = 0.5 * (pab[v][t] + sab[v][t]);
          const double b1 = 0.5 * (pab[u][s] - sab[u][s]), b2 = 0.5 * (pab[v][t] - sab[v][t]);
          c += paa[u][v] * pbb[s][t] - (a1 * a2 + b1 * b2);
        } else {
          c += paa[u][v] * pbb[s][t] - 0.5 * pab[u][s] * pab[v][t];
        }
      }
    return c;
  };
  if (na == 4 && nb == 4) {
    D3 ri[22];
    local_xx(ea, eb, r, ri);
    // idx = pk(kk, ll) * 10 + pk(mm, nn); unpack from the loop order of assemble_xx
    assemble_xx(ri, R, [&](int idx, const D3& v) {
      const int pp = idx / 10, qq = idx % 10;
      const int k1 = pp >= 6 ? 3 : (pp >= 3 ? 2 : (pp >= 1 ? 1 : 0));
      const int l1 = pp - k1 * (k1 + 1) / 2;
      const int m1 = qq >= 6 ? 3 : (qq >= 3 ? 2 : (qq >= 1 ? 1 : 0));
      const int n1 = qq - m1 * (m1 + 1) / 2;
      acc(coef(k1, l1, m1, n1), v);
    });
  } else if (na == 4) {
    D3 ri[4], w[10];
    local_xh(ea.dd, ea.qq, ea.rho0, ea.rho1, ea.rho2, eb.rho0, r, ri);
    assemble_xh(ri, R, w);
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = 0; nu <= mu; ++nu) acc(coef(mu, nu, 0, 0), w[pk(mu, nu)]);
  } else {
    const double s = ea.rho0 + eb.rho0;
    acc(coef(0, 0, 0, 0), rsq(r * r + s * s) * EV);
  }
  out[0] = f0;
  out[1] = f1;
  out[2] = f2;
}

// Per-pair forces for every pair of every molecule (grid-stride); the densities are the
// converged ones in the output buffer (sorted AO order, dense n x n).
__global__ __launch_bounds__(256) void grad_pair_kernel(KArgs A) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x; g < A.npairs_total; g += stride) {
    int lo = 0, hi = A.nmol_pairs;
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (A.pair_base[mid] <= g) lo = mid; else hi = mid;
    }
    const MolDesc md = A.mols[A.order[lo]];
    const Mol M = mol_of(A, md, nullptr);
    const int p = (int)(g - A.pair_base[lo]);
    double f[3];
    const double* S = (md.uhf && A.out_S) ? A.out_S + md.out_mat : nullptr;
    pair_gradient(A, M, A.out_P + md.out_mat, S, md.nao, p, f);
    const long long npairs = (long long)md.nat * (md.nat - 1) / 2;
    M.S1[p] = f[0];  // the electron-core slots are free once H_core is assembled
    M.S1[npairs + p] = f[1];
    M.S1[2 * npairs + p] = f[2];
  }
}

// g_x = sum_{b > x} (-f_xb) + sum_{a < x} f_ax, one workgroup per molecule.
__global__ __launch_bounds__(256) void grad_gather_kernel(KArgs A) {
  const MolDesc md = A.mols[A.order[blockIdx.x]];
  const Mol M = mol_of(A, md, nullptr);
  const int nat = md.nat;
  const long long npairs = (long long)nat * (nat - 1) / 2;
  for (int e = threadIdx.x; e < 3 * nat; e += blockDim.x) {
    const int x = e / 3, c = e % 3;
    double s = 0.0;
    for (int y = 0; y < nat; ++y) {
      if (y == x) continue;
      s += y > x ? -M.S1[c * npairs + tri_index(x, y, nat)] : M.S1[c * npairs + tri_index(y, x, nat)];
    }
    // the field pulls the net charge: -q_x f (add_field_gradient, src/gradient.rs:296-301)
    if (A.has_field) s -= A.field[c] * A.out_q[md.atom0 + x];
    A.out_g[3LL * (md.atom0 + x) + c] = s;
  }
}

// The same sums for molecules of the host-driven path (queue positions first + blockIdx.x): one
// wave per atom x, lanes over the partner atoms, then a fixed shuffle tree; blockIdx.y strides the
// atoms in groups of four.
__global__ __launch_bounds__(256) void grad_gather_wide_kernel(KArgs A, int first) {
  const MolDesc md = A.mols[A.order[first + blockIdx.x]];
  const Mol M = mol_of(A, md, nullptr);
  const int nat = md.nat, lane = threadIdx.x & 63;
  const long long npairs = (long long)nat * (nat - 1) / 2;
  for (int x = blockIdx.y * 4 + (threadIdx.x >> 6); x < nat; x += gridDim.y * 4) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int y = lane; y < nat; y += 64) {
      if (y == x) continue;
      const long long t = y > x ? tri_index(x, y, nat) : tri_index(y, x, nat);
      const double sg = y > x ? -1.0 : 1.0;
      s0 += sg * M.S1[t];
      s1 += sg * M.S1[npairs + t];
      s2 += sg * M.S1[2 * npairs + t];
    }
    for (int off = 32; off > 0; off >>= 1) {
      s0 += __shfl_down(s0, off, 64);
      s1 += __shfl_down(s1, off, 64);
      s2 += __shfl_down(s2, off, 64);
    }
    if (lane == 0) {
      const double q = A.has_field ? A.out_q[md.atom0 + x] : 0.0;
      A.out_g[3LL * (md.atom0 + x)] = s0 - A.field[0] * q;
      A.out_g[3LL * (md.atom0 + x) + 1] = s1 - A.field[1] * q;
      A.out_g[3LL * (md.atom0 + x) + 2] = s2 - A.field[2] * q;
    }
  }
}

#endif  // PM3_KERNEL_VARIANT

// Mulliken charges (into out_q) and the dipole about the centre of mass, atomic units
// (src/scf.rs:762-791, src/dipole.rs:64-140).  One workgroup; P has leading dimension M.np.
__device__ inline void charges_dipole(const KArgs& A, const Mol& M, const double* P, double* red, double (&dip)[3]) {
  const int tid = threadIdx.x, np = M.np;
  double com[4] = {0.0, 0.0, 0.0, 0.0};
  for (int a = tid; a < M.nat; a += NT) {
    const int ga = M.atom0 + a;
    const double ms = c_tab.el[A.aslot[ga]].mass;
    com[0] += ms * A.ax[ga]; com[1] += ms * A.ay[ga]; com[2] += ms * A.az[ga]; com[3] += ms;
  }
  block_sum<4>(com, red);
  const double cx = com[3] > 0.0 ? com[0] / com[3] : 0.0, cy = com[3] > 0.0 ? com[1] / com[3] : 0.0,
               cz = com[3] > 0.0 ? com[2] / com[3] : 0.0;
  dip[0] = dip[1] = dip[2] = 0.0;
  for (int a = tid; a < M.nat; a += NT) {
    const int ga = M.atom0 + a;
    const ElemConst& e = c_tab.el[A.aslot[ga]];
    const int o = A.aoff[ga];
    double pop = 0.0;
    for (int mu = 0; mu < e.norb; ++mu) pop += P[(o + mu) * np + o + mu];
    A.out_q[ga] = e.tore - pop;
    const double disp[3] = {A.ax[ga] - cx, A.ay[ga] - cy, A.az[ga] - cz};
    for (int ax = 0; ax < 3; ++ax) {
      double el = pop * disp[ax];
      if (e.norb == 4) el += (P[o * np + o + ax + 1] + P[(o + ax + 1) * np + o]) * e.dd;
      dip[ax] += disp[ax] * e.tore - el;
    }
  }
  block_sum<3>(dip, red);
}

// An unrestricted molecule of the batch (run_pm3's UHF branch, src/scf.rs:521-715): the first
// path (steered by the level shift and damping when either is set), then, unless steered, the
// stability search when the frontier gap over both channels is below 6 eV (or always): all seven
// rungs of src/scf.rs:647-673 as unrestricted paths, the lowest converged energy kept (1e-6 eV).
// Writes the result record, the total and spin densities, and the orbitals on the eigenvector route.
__device__ __noinline__ void uhf_molecule(const KArgs& A, const Mol& M, const MolDesc& md, int mi, const Bufs& B,
                                          double ecore, double* lds, double* red, Diis& D, double* bnd_b) {
  const int tid = threadIdx.x, n = M.n, np = M.np;
  double* Pt;
  Bufs Bb = beta_slab(A, B, Pt);
  Bb.bnd = bnd_b;
  for (long long e = tid; e < (long long)np * np; e += NT) {
    Bb.F[e] = 0.0; Bb.P[e] = 0.0; Bb.Pn[e] = 0.0; Bb.Fu[e] = 0.0; Bb.C[e] = 0.0; Bb.Cn[e] = 0.0; Pt[e] = 0.0;
  }
  __syncthreads();
  const bool steered = A.shift != 0.0 || A.damp != 0.0;
  const int g0 = A.guess == 4 ? 1 : A.guess;  // the learned guess covers restricted molecules
  const int cd = A.accel == 1 ? 1 : 0;        // uhf_loop extrapolates only when CDIIS is asked for
  const double s0 = steered ? A.shift : 0.0, d0 = steered ? A.damp : 0.0;
  PathResult best = uhf_loop(A, M, B, Bb, Pt, g0, cd, s0, d0, lds, red, D, true);
  if (uhf_aufbau_broken(A, M, B, Bb)) best = uhf_loop(A, M, B, Bb, Pt, g0, cd, s0, d0, lds, red, D, false);
  const double first_e = best.e_elec;
  int paths = 1;
  double homo, lumo;
  uhf_frontier(M, B, Bb, homo, lumo);
  const double gap = (homo > -1.0e299 && lumo < 1.0e299) ? lumo - homo : 1.0e300;
  const bool search = !steered && (A.stability == 2 || (A.stability == 1 && gap < 6.0));
  auto keep = [&]() {  // the current state becomes the incumbent
    for (int e = tid; e < np * np; e += NT) {
      B.Pb[e] = B.P[e];
      B.Cb[e] = B.C[e];
      Bb.Pb[e] = Bb.P[e];
      Bb.Cb[e] = Bb.C[e];
    }
    for (int j = tid; j < n; j += NT) {
      B.epsb[j] = B.eps[j];
      Bb.epsb[j] = Bb.eps[j];
    }
    __syncthreads();
  };
  if (search) {
    keep();
    const int primary = A.guess == 2 ? 2 : 1, other = primary == 2 ? 1 : 2;
    const int rg[7][2] = {{primary, 0}, {other, cd}, {other, 0}, {3, cd}, {3, 0}, {3, 0}, {primary, A.use_diis ? 1 : 0}};
    const double rs[7][2] = {{0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0}, {-2.0, 0.0}, {0.0, 0.85}};
    for (int k = 0; k < 7; ++k) {
      ++paths;
      PathResult c = uhf_loop(A, M, B, Bb, Pt, rg[k][0], rg[k][1], rs[k][0], rs[k][1], lds, red, D, true);
      if (uhf_aufbau_broken(A, M, B, Bb))
        c = uhf_loop(A, M, B, Bb, Pt, rg[k][0], rg[k][1], rs[k][0], rs[k][1], lds, red, D, false);
      if (!c.converged) continue;
      if (!best.converged || c.e_elec < best.e_elec - 1.0e-6) {
        best = c;
        keep();
      }
    }
    for (int e = tid; e < np * np; e += NT) {
      B.P[e] = B.Pb[e];
      B.C[e] = B.Cb[e];
      Bb.P[e] = Bb.Pb[e];
      Bb.C[e] = Bb.Cb[e];
    }
    for (int j = tid; j < n; j += NT) {
      B.eps[j] = B.epsb[j];
      Bb.eps[j] = Bb.epsb[j];
    }
    __syncthreads();
  }
  for (int e = tid; e < np * np; e += NT) Pt[e] = B.P[e] + Bb.P[e];
  __syncthreads();
  double dip[3];
  charges_dipole(A, M, Pt, red, dip);
  uhf_frontier(M, B, Bb, homo, lumo);
  if (tid == 0) {
    MolOut o;
    o.dip[0] = dip[0]; o.dip[1] = dip[1]; o.dip[2] = dip[2];
    o.e_elec = best.e_elec;
    o.e_core = ecore;
    o.last_dp = best.last_dp;
    o.first_e = first_e;
    o.homo = homo > -1.0e299 ? homo : 0.0;
    o.lumo = lumo < 1.0e299 ? lumo : 0.0;
    o.iters = best.iters;
    o.converged = best.converged;
    o.paths = paths;
    o.status = best.converged ? 0 : 3;
    o.unrestricted = 1;
    o.guess_used = 0;
    A.out[mi] = o;
  }
  if (md.out_mat >= 0) {
    for (int e = tid; e < n * n; e += NT) {
      const int i = e / n, j = e % n;
      A.out_P[md.out_mat + e] = Pt[i * np + j];
      if (A.out_S) A.out_S[md.out_mat + e] = B.P[i * np + j] - Bb.P[i * np + j];
      if (!A.purify) {
        A.out_C[md.out_mat + e] = B.C[i * np + j];
        if (A.out_Cb) A.out_Cb[md.out_mat + e] = Bb.C[i * np + j];
      }
    }
    if (!A.purify)
      for (int j = tid; j < n; j += NT) {
        A.out_eps[md.out_vec + j] = B.eps[j];
        if (A.out_epsb) A.out_epsb[md.out_vec + j] = Bb.eps[j];
      }
  }
  __syncthreads();
}

// Launch bounds of 512 threads in every build: the register budget is then 256 VGPRs per lane
// (two waves per SIMD), so the 256- and 128-thread builds fit two and four workgroups per CU.
#ifndef PM3_SCF_KERNEL
#define PM3_SCF_KERNEL scf_batch_kernel
#endif
__global__ __launch_bounds__(512) void PM3_SCF_KERNEL(KArgs A) {
  double* lds = g_lds;
  __shared__ double red[NW * RED_K];
  __shared__ Diis D;
  __shared__ int s_mol;
  __shared__ unsigned long long s_ph[PH_SLOTS];
  __shared__ double s_bnd[4];
  __shared__ double s_bnd_b[4];
  const int tid = threadIdx.x;
  double* Wbase;
  Bufs B = slab(A, Wbase);
  B.ph = A.prof ? s_ph : nullptr;
  B.bnd = s_bnd;
  if (tid < PH_SLOTS) s_ph[tid] = 0;
  if (tid == 0) s_ph[PH_LAST] = wall_clock64();

  for (;;) {
    if (tid == 0) s_mol = atomicAdd(A.queue, 1);
    if (tid == 0 && A.prof) {
      for (int k = 0; k < PH_SLOTS; ++k) {
        if (k == PH_LAST) continue;
        atomicAdd(&A.prof[k], s_ph[k]);
        s_ph[k] = 0;
      }
    }
    __syncthreads();
    const int qi = s_mol;
    __syncthreads();
    if (qi >= A.nmol) break;
    const int mi = A.order[qi];
    const MolDesc md = A.mols[mi];
    Mol M = mol_of(A, md, Wbase);
    // a warm start only from a density whose solve converged (A.out still holds that solve's)
    if (M.prev && A.out[mi].converged == 0) M.prev = nullptr;
    // Matrices use the molecule's own leading dimension np inside the workgroup's slab; the
    // padding beyond nao stays zero for the whole solve.
    zero_mats(M, B);
    const double ecore = finish_core(A, M, red);
    stamp(B, 0);
    if (md.uhf) {
      uhf_molecule(A, M, md, mi, B, ecore, lds, red, D, s_bnd_b);
      continue;
    }
    PathResult best = rhf_loop(A, M, B, A.guess, A.accel, lds, red, D, true, true);
    if (aufbau_broken(A, M, B)) best = rhf_loop(A, M, B, A.guess, A.accel, lds, red, D, false, true);
    const int guess_used = best.guess_used;
    const int nocc = M.nocc, n = M.n, np = M.np;
    const double first_e = best.e_elec;
    int paths = 1;
    auto gap = [&]() { return (nocc > 0 && nocc < n) ? B.eps[nocc] - B.eps[nocc - 1] : 1.0e300; };
    // the stability search (src/scf.rs:599-715): never on a steered path; for a closed shell on the
    // restricted path only the two unrestricted rungs run (a rung runs when its unrestricted flag
    // equals closed_shell, :676), each spin-pure by construction here and so read as the
    // restricted state (:694-708)
    const bool steered = A.shift != 0.0 || A.damp != 0.0;
    const bool search = !steered && (A.stability == 2 || (A.stability == 1 && gap() < 6.0));
    if (search) {
      // keep the incumbent
      for (int e = tid; e < np * np; e += NT) {
        B.Pb[e] = B.P[e];
        B.Cb[e] = B.C[e];
      }
      for (int j = tid; j < n; j += NT) B.epsb[j] = B.eps[j];
      __syncthreads();
      const int primary = A.guess == 2 ? 2 : 1;
      const int rg[2][2] = {{3, 0}, {primary, A.use_diis ? 1 : 0}};
      const double rs[2][2] = {{-2.0, 0.0}, {0.0, 0.85}};
      for (int k = 0; k < 2; ++k) {
        ++paths;
        PathResult c = rhf_loop(A, M, B, rg[k][0], rg[k][1], lds, red, D, true, false, rs[k][0], rs[k][1], true);
        if (aufbau_broken(A, M, B))
          c = rhf_loop(A, M, B, rg[k][0], rg[k][1], lds, red, D, false, false, rs[k][0], rs[k][1], true);
        if (!c.converged) continue;
        if (!best.converged || c.e_elec < best.e_elec - 1.0e-6) {
          best = c;
          for (int e = tid; e < np * np; e += NT) {
            B.Pb[e] = B.P[e];
            B.Cb[e] = B.C[e];
          }
          for (int j = tid; j < n; j += NT) B.epsb[j] = B.eps[j];
          __syncthreads();
        }
      }
      for (int e = tid; e < np * np; e += NT) {
        B.P[e] = B.Pb[e];
        B.C[e] = B.Cb[e];
      }
      for (int j = tid; j < n; j += NT) B.eps[j] = B.epsb[j];
      __syncthreads();
    }
    double dip[3];
    charges_dipole(A, M, B.P, red, dip);
    if (tid == 0) {
      MolOut o;
      o.dip[0] = dip[0]; o.dip[1] = dip[1]; o.dip[2] = dip[2];
      o.e_elec = best.e_elec;
      o.e_core = ecore;
      o.last_dp = best.last_dp;
      o.first_e = first_e;
      o.homo = nocc > 0 ? B.eps[nocc - 1] : 0.0;
      o.lumo = nocc < n ? B.eps[nocc] : 0.0;
      o.iters = best.iters;
      o.converged = best.converged;
      o.paths = paths;
      o.status = best.converged ? 0 : 3;
      o.unrestricted = 0;
      o.guess_used = guess_used;
      A.out[mi] = o;
    }
    stamp(B, 7);
    if (md.out_mat >= 0) {
      for (int e = tid; e < n * n; e += NT) {
        const int i = e / n, j = e % n;
        A.out_P[md.out_mat + e] = B.P[i * np + j];
        if (!A.purify) A.out_C[md.out_mat + e] = B.C[i * np + j];
      }
      if (!A.purify)
        for (int j = tid; j < n; j += NT) A.out_eps[md.out_vec + j] = B.eps[j];
    }
    __syncthreads();
  }
}

#ifndef PM3_KERNEL_VARIANT  // kernels launched by the main build only
// Single stages of the pipeline on molecule 0 of a batch, for kernel-level parity checks:
// mode 0 H_core (+ core-core energy at out[n*n]); mode 1 F(P) for P given in `in`;
// mode 2 eigenpairs of `in` (cold Jacobi): out = eps[n] then C[n*n];
// mode 3 out = in[0:n*n] . in[n*n:2n*n] on the matrix cores;
// mode 4 warm eigensolve of in[0:n*n] in the basis in[n*n:2n*n]: out = eps[n] then C[n*n];
// mode 5 learned initial density (network outputs already computed): out = P[n*n], then the
// squared first-iteration commutator norms of the learned and the superposition densities;
// mode 6 purified aufbau density of `in` (n_occ of molecule 0): out = P[n*n], then the number of
// products (-1: not converged), HOMO and LUMO from the Lanczos runs.
// Matrices in `in`/`out` are dense n x n row-major in the engine's (heavy-first) AO order.
__global__ __launch_bounds__(NT) void stage_kernel(KArgs A, int mode, const double* in, double* out) {
  double* lds = g_lds;
  __shared__ double red[NW * RED_K];
  const int tid = threadIdx.x;
  double* Wbase;
  __shared__ double s_bnd[4];
  Bufs B = slab(A, Wbase);
  B.bnd = s_bnd;
  if (tid == 0) {
    s_bnd[2] = -1.0;
    s_bnd[3] = 0.0;
  }
  const Mol M = mol_of(A, A.mols[A.order[0]], Wbase);
  const int n = M.n, np = M.np;
  zero_mats(M, B);
  __syncthreads();
  if (mode == 0 || mode == 1) {
    const double ecore = finish_core(A, M, red);
    if (mode == 0) {
      for (int e = tid; e < n * n; e += NT) out[e] = M.H[(e / n) * np + e % n];
      if (tid == 0) out[n * n] = ecore;
    } else {
      for (int e = tid; e < n * n; e += NT) B.P[(e / n) * np + e % n] = in[e];
      __syncthreads();
      build_fock(A, M, M.H, B.P, B.F, lds);
      for (int e = tid; e < n * n; e += NT) out[e] = B.F[(e / n) * np + e % n];
    }
  } else if (mode == 2 || mode == 4) {
    for (int e = tid; e < n * n; e += NT) {
      B.Fu[(e / n) * np + e % n] = in[e];
      if (mode == 4) B.C[(e / n) * np + e % n] = in[n * n + e];
    }
    __syncthreads();
    eigensolve(M, B, B.Fu, mode == 4, lds, red);
    for (int j = tid; j < n; j += NT) out[j] = B.eps[j];
    for (int e = tid; e < n * n; e += NT) out[n + e] = B.Cn[(e / n) * np + e % n];
  } else if (mode == 3) {
    const double* X = in;
    const double* Y = in + n * n;
    wg_gemm(n, n, n, [&](int i, int k) { return X[i * n + k]; }, [&](int k, int j) { return Y[k * n + j]; },
            [&](int i, int j, double v) { out[i * n + j] = v; });
  } else if (mode == 5 && M.learn && A.gY) {
    (void)finish_core(A, M, red);
    learned_into(A, M, B.P);
    const double cl = commutator2(A, M, B, lds, red);
    for (int e = tid; e < n * n; e += NT) out[e] = B.P[(e / n) * np + e % n];
    __syncthreads();
    for (int e = tid; e < np * np; e += NT) B.P[e] = 0.0;
    __syncthreads();
    sad_into(A, M, B.P);
    const double cs = commutator2(A, M, B, lds, red);
    if (tid == 0) {
      out[n * n] = cl;
      out[n * n + 1] = cs;
    }
  } else if (mode == 6) {
    for (int e = tid; e < n * n; e += NT) B.Fu[(e / n) * np + e % n] = in[e];
    __syncthreads();
    const int ns = sp2_density(M, B, B.Fu, B.P, red, true, true);
    if (ns >= 0) frontier_levels(M, B, B.Fu, B.P, lds, red);
    for (int e = tid; e < n * n; e += NT) out[e] = B.P[(e / n) * np + e % n];
    if (tid == 0) {
      out[n * n] = (double)ns;
      out[n * n + 1] = (ns >= 0 && M.nocc > 0) ? B.eps[M.nocc - 1] : 0.0;
      out[n * n + 2] = (ns >= 0 && M.nocc < n) ? B.eps[M.nocc] : 0.0;
    }
  }
}


// ------------------------------------------------------------------------------------------
// Large molecules (more AOs than one workgroup's LDS holds, e.g. 512 waters = 3072 AOs): the SCF
// loop runs on the host and every step is a grid-wide kernel over the one molecule -- the same
// integral, Fock-phase and H_core code as the batched kernel -- while the dense products go to
// rocBLAS and the eigensolve to rocSOLVER.  Matrices keep the molecule's leading dimension np.
constexpr int LG_BLOCKS = 1024;
constexpr int LG_THREADS = 256;
constexpr int LG_MAXDOT = 3 * MAXD + 2;
constexpr double LARGE_SP2_MARGIN = 0.15;  // least relative margin of the large path's Lanczos bounds
constexpr int LG_SPLIT_MAX = 4;            // largest q of the q x q block-lower purification products
constexpr int LARGE_STEPS_N = 1024;        // up to this size the host enqueues several purification steps ...
constexpr int LARGE_STEPS_PER_LOOK = 8;    // ... this many, between looks at the device state, each
                                           // step a single fused product and update kernel
constexpr int LARGE_LZ_FUSED_N = 512;      // up to this size a Lanczos step is one one-block kernel
constexpr int LZ_VEC_THREADS = 1024;       // threads of the one-block Lanczos kernels

struct Dots {
  const double* x[LG_MAXDOT];
  const double* y[LG_MAXDOT];
  int k;
};

__device__ inline int gtid() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline int gstride() { return gridDim.x * blockDim.x; }

// Block partials of K dot products over m elements (red[block * LG_MAXDOT + k]); the host adds
// the blocks in order, so the sums are reproducible.
__global__ __launch_bounds__(LG_THREADS) void lg_dots_kernel(Dots D, long long m, double* red) {
  __shared__ double part[LG_MAXDOT][LG_THREADS / 64];
  double v[LG_MAXDOT];
#pragma unroll
  for (int k = 0; k < LG_MAXDOT; ++k) v[k] = 0.0;
  for (long long e = gtid(); e < m; e += gstride()) {
#pragma unroll
    for (int k = 0; k < LG_MAXDOT; ++k)
      if (k < D.k) v[k] += D.x[k][e] * D.y[k][e];
  }
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < LG_MAXDOT; ++k) {
    if (k >= D.k) break;
    const double t = wave_sum(v[k]);
    if ((threadIdx.x & 63) == 0) part[k][w] = t;
  }
  __syncthreads();
  if (threadIdx.x < D.k) {
    double t = 0.0;
    for (int q = 0; q < LG_THREADS / 64; ++q) t += part[threadIdx.x][q];
    red[(long long)blockIdx.x * LG_MAXDOT + threadIdx.x] = t;
  }
}

// Purification of the large path (the batched kernel's sp2_density, with rocBLAS products).
// Gershgorin bounds of the symmetric np-strided A, one wave per row: per-block lowest disc bottom
// and highest disc top in red[block * LG_MAXDOT + 0 / 1].
__global__ __launch_bounds__(LG_THREADS) void lg_gershgorin_kernel(const double* A, int n, int np, double* red) {
  __shared__ double part[2][LG_THREADS / 64];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  double lo = 1.0e300, hi = -1.0e300;
  for (int i = blockIdx.x * (LG_THREADS / 64) + w; i < n; i += gridDim.x * (LG_THREADS / 64)) {
    const double* row = A + (size_t)i * np;
    double r = 0.0;
    for (int j = lane; j < n; j += 64) r += j == i ? 0.0 : fabs(row[j]);
    r = wave_sum(r);
    lo = fmin(lo, row[i] - r);
    hi = fmax(hi, row[i] + r);
  }
  if (lane == 0) {
    part[0][w] = lo;
    part[1][w] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < LG_THREADS / 64; ++q) {
      lo = fmin(lo, part[0][q]);
      hi = fmax(hi, part[1][q]);
    }
    red[(size_t)blockIdx.x * LG_MAXDOT] = fmin(lo, part[0][0]);
    red[(size_t)blockIdx.x * LG_MAXDOT + 1] = fmax(hi, part[1][0]);
  }
}

// Per-block partials of tr X and |X|_F^2 of the n x n block-stride loop (red columns 0 and 1).
__device__ inline void lg_trace_partials(double tr, double fr, double* red) {
  __shared__ double part[2][LG_THREADS / 64];
  tr = wave_sum(tr);
  fr = wave_sum(fr);
  if ((threadIdx.x & 63) == 0) {
    part[0][threadIdx.x >> 6] = tr;
    part[1][threadIdx.x >> 6] = fr;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, c = 0.0;
    for (int q = 0; q < LG_THREADS / 64; ++q) {
      a += part[0][q];
      c += part[1][q];
    }
    red[(size_t)blockIdx.x * LG_MAXDOT] = a;
    red[(size_t)blockIdx.x * LG_MAXDOT + 1] = c;
  }
}

// X0 = (hi I - A) / (hi - lo), with its trace and norm partials.
__global__ __launch_bounds__(LG_THREADS) void lg_sp2_init_kernel(const double* A, double* X, int n, int np, double lo,
                                                                 double hi, double* red) {
  const double s = 1.0 / (hi - lo);
  double tr = 0.0, fr = 0.0;
  for (long long e = gtid(); e < (long long)n * n; e += gstride()) {
    const int i = (int)(e / n), j = (int)(e % n);
    const double x = ((i == j ? hi : 0.0) - A[(size_t)i * np + j]) * s;
    X[(size_t)i * np + j] = x;
    fr += x * x;
    if (i == j) tr += x;
  }
  lg_trace_partials(tr, fr, red);
}

// The two frontier operators from F, P (X = P / 2), G = X F and H = X F X (row-major):
// A_o = (H + H^T) / 2 + c_lo (I - X) and A_v = -(F - G - G^T + (H + H^T) / 2 + c_hi X).
__global__ __launch_bounds__(LG_THREADS) void lg_frontier_ops_kernel(const double* F, const double* P, const double* G,
                                                                     const double* H, int n, int np, double c_lo,
                                                                     double c_hi, double* Ao, double* Av) {
  for (long long e = gtid(); e < (long long)n * n; e += gstride()) {
    const int i = (int)(e / n), j = (int)(e % n);
    const size_t ij = (size_t)i * np + j, ji = (size_t)j * np + i;
    const double x = 0.5 * P[ij], h = 0.5 * (H[ij] + H[ji]), id = i == j ? 1.0 : 0.0;
    Ao[ij] = h + c_lo * (id - x);
    Av[ij] = -(F[ij] - G[ij] - G[ji] + h + c_hi * x);
  }
}

// X <- Y or 2X - Y, with Y = X X (only its row-major lower triangle is read, so X stays exactly
// symmetric), then the trace and norm partials of the new X.
__global__ __launch_bounds__(LG_THREADS) void lg_sp2_update_kernel(const double* Y, double* X, int n, int np, int sq,
                                                                   double* red) {
  double tr = 0.0, fr = 0.0;
  for (long long e = gtid(); e < (long long)n * n; e += gstride()) {
    const int i = (int)(e / n), j = (int)(e % n);
    const double y = i >= j ? Y[(size_t)i * np + j] : Y[(size_t)j * np + i];
    const double x = sq ? y : 2.0 * X[(size_t)i * np + j] - y;
    X[(size_t)i * np + j] = x;
    fr += x * x;
    if (i == j) tr += x;
  }
  lg_trace_partials(tr, fr, red);
}

// Device-side control of the purification loop (large_purify_run): the state sp2_density keeps in
// registers -- tr X and tr X^2 of the current iterate, the tail counter, the steps done, the next
// step's choice and the outcome -- in a small device array that one block updates after every
// step, so that the host looks at it only every few steps.  Status: 0 running, 1 converged (X is
// the projector), 2 a bound was violated (tr X^2 > tr X), 3 no convergence or a wrong trace.
constexpr int SP2S_TR = 0, SP2S_T2 = 1, SP2S_TAIL = 2, SP2S_STATUS = 3, SP2S_SQ = 4, SP2S_K = 5, SP2S_SEED = 6,
              SP2S_SEEDED = 7, SP2S_N = 8;

__device__ inline void sp2_decide(double* st, double tr, double t2, double occ, int init, int want_seed);

// After X0 (init) or after a step: tr and tr X^2 from the block partials (the halving tree of
// lg_finish_kernel), the bookkeeping of the step just taken, then the decisions at the top of the
// next step of sp2_density's loop (bound check, tail start with the frontier seed, stop, X^2 or
// 2X - X^2).  A finished state is left as it is.  Row r of the partials starts at red[r * rstride];
// thread t adds rows t, t + LG_BLOCKS, ... in that order before the tree.
__global__ __launch_bounds__(LG_BLOCKS) void lg_sp2_state_kernel(const double* red, double* st, double occ, int init,
                                                                 int want_seed, int nrows, int rstride) {
  __shared__ double t[LG_BLOCKS];
  if (!init && st[SP2S_STATUS] != 0.0) return;  // the same value for every thread of the block
  double v[2];
  for (int q = 0; q < 2; ++q) {
    double a = 0.0;
    for (int r = threadIdx.x; r < nrows; r += LG_BLOCKS) a += red[(size_t)r * rstride + q];
    t[threadIdx.x] = a;
    __syncthreads();
    for (int h = LG_BLOCKS / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) t[threadIdx.x] += t[threadIdx.x + h];
      __syncthreads();
    }
    v[q] = t[0];
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  sp2_decide(st, v[0], v[1], occ, init, want_seed);
}

// lg_sp2_state_kernel's bookkeeping and decisions, by one thread.
__device__ inline void sp2_decide(double* st, double tr, double t2, double occ, int init, int want_seed) {
  int tail = init ? -1 : (int)st[SP2S_TAIL];
  const int k = init ? 0 : (int)st[SP2S_K] + 1;
  if (!init && tail > 0) --tail;
  int status = 0, seed = 0;
  if (init) st[SP2S_SEEDED] = 0.0;
  if (k >= SP2_MAX) {
    status = (tail == 0 && fabs(tr - occ) <= 1.0e-6) ? 1 : 3;
  } else if (t2 > tr + 1.0e-9 * fmax(1.0, tr)) {
    status = 2;
  } else {
    if (tail < 0 && tr - t2 < SP2_TAIL) {
      tail = SP2_EXTRA;
      seed = want_seed;
    }
    if (tail == 0) status = fabs(tr - occ) > 1.0e-6 ? 3 : 1;
  }
  st[SP2S_TR] = tr;
  st[SP2S_T2] = t2;
  st[SP2S_TAIL] = (double)tail;
  st[SP2S_K] = (double)k;
  st[SP2S_STATUS] = (double)status;
  st[SP2S_SEED] = (double)seed;
  st[SP2S_SQ] = fabs(t2 - occ) < fabs(2.0 * tr - t2 - occ) ? 1.0 : 0.0;
  if (seed) st[SP2S_SEEDED] = 1.0;
}

// lg_sp2_update_kernel with the step's choice read from the state; nothing once it has finished.
__global__ __launch_bounds__(LG_THREADS) void lg_sp2_step_kernel(const double* Y, double* X, int n, int np,
                                                                 const double* st, double* red) {
  if (st[SP2S_STATUS] != 0.0) return;
  const bool sq = st[SP2S_SQ] != 0.0;
  double tr = 0.0, fr = 0.0;
  for (long long e = gtid(); e < (long long)n * n; e += gstride()) {
    const int i = (int)(e / n), j = (int)(e % n);
    const double y = i >= j ? Y[(size_t)i * np + j] : Y[(size_t)j * np + i];
    const double x = sq ? y : 2.0 * X[(size_t)i * np + j] - y;
    X[(size_t)i * np + j] = x;
    fr += x * x;
    if (i == j) tr += x;
  }
  lg_trace_partials(tr, fr, red);
}

// One purification step on the matrix cores for matrices up to LARGE_STEPS_N AOs, product and
// update in one launch: Xn = X X, or 2 X - X X, as the state chose, from X into the other buffer.
// Each wave owns 16 x 16 tiles of the upper triangle (blocks stride over them) and stores each
// element and its mirror; loads are bounds-checked, so no padding is read.  Block b leaves the
// trace and norm partials of its tiles in red[b] (the state kernel sums the grid's rows).
__global__ __launch_bounds__(64) void lg_sp2_fused_kernel(const double* X, double* Xn, int n, int np, const double* st,
                                                          double* red) {
  if (st[SP2S_STATUS] != 0.0) return;
  const bool sq = st[SP2S_SQ] != 0.0;
  const int lane = threadIdx.x, li = lane & 15, lk = lane >> 4;
  const int bm = (n + 15) >> 4, ntile = bm * (bm + 1) / 2;
  double tr = 0.0, fr = 0.0;
  for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
    int bi, bj;
    pad_block<true>(t, bm, bi, bj);
    const int ia = bi * 16 + li, ja = bj * 16 + li;
    double4_t c = {0.0, 0.0, 0.0, 0.0};
    // 64 k values per round: all 32 loads of a round are issued before its 16 MFMAs
    constexpr int FK = 16;
    for (int k0 = 0; k0 < n; k0 += 4 * FK) {
      double a[FK], bv[FK];
#pragma unroll
      for (int u = 0; u < FK; ++u) {
        const int k = k0 + 4 * u + lk;
        // A[i][k] = X[k][i] (X symmetric) and B[k][j]: lane l holds A[l % 16][l / 16], B[l / 16][l % 16]
        a[u] = (k < n && ia < n) ? X[(size_t)k * np + ia] : 0.0;
        bv[u] = (k < n && ja < n) ? X[(size_t)k * np + ja] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < FK; ++u) c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u], bv[u], c, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {  // item r: row 4 r + l / 16, column l % 16 of the tile
      const int i = bi * 16 + 4 * r + lk, j = ja;
      if (i < n && j < n && (bi < bj || i <= j)) {
        const double y = sq ? c[r] : 2.0 * X[(size_t)i * np + j] - c[r];
        Xn[(size_t)i * np + j] = y;
        if (i != j) Xn[(size_t)j * np + i] = y;
        fr += (i == j ? 1.0 : 2.0) * y * y;
        if (i == j) tr += y;
      }
    }
  }
  tr = wave_sum(tr);
  fr = wave_sum(fr);
  if (lane == 0) {
    red[(size_t)blockIdx.x * LG_MAXDOT] = tr;
    red[(size_t)blockIdx.x * LG_MAXDOT + 1] = fr;
  }
}

// One purification step of a large matrix on the matrix cores, product and update in one launch:
// Xn = X X, or 2 X - X X, as the state chose.  X is symmetric, so tile (I, J) of X X is the product
// of X's row panels I and J (X[I,:] X[J,:]^T), and only the tiles of the lower triangle (I >= J)
// are computed: about half the flops of the whole square.  A workgroup of four waves computes a
// 64 x 64 tile, each wave a 32 x 32 quarter as 2 x 2 MFMA blocks.  32-column slabs of both panels
// pass through LDS, double-buffered: a slab's global loads are issued before the MFMAs of the
// previous one and stored after them.  In LDS a panel row is 16 slots of 16 bytes, slot c of row r
// at c ^ (r % 16), so the row-strided fragment reads and the slab stores meet no bank conflict.
// The k index of the 16 x 16 x 4 MFMA is permuted inside a slab: lane group g = lane / 16 takes
// columns 8 g .. 8 g + 7, two per 16-byte read, in both operands, so each column enters once.  The
// store writes each element of the lower triangle and its mirror, so Xn is exactly symmetric.
// Loads beyond n read as zero, so padding rows and columns are never read.  Workgroup w computes
// the tiles map[x * chunk + p] with x = w % 8 and p = w / 8, w / 8 + g / 8, ... (g the grid, a
// multiple of 8; one tile each when g = 8 chunk): workgroups go to the eight XCDs in turn, so each
// XCD takes a contiguous run of the map, which lists the tiles by 8 x 8 super-blocks; the tiles an
// XCD has in flight then share few panels, and its L2 serves the rest.  Workgroup w leaves the
// trace and norm partials of its tiles, summed in that order, in red[2 w], red[2 w + 1] (zeros when
// it has none).
constexpr int SY_T = 64, SY_THREADS = 256;

// Slot of 16-byte column chunk c of panel row `row` in a slab of K columns (K / 2 chunks a row):
// the chunk index is XORed with the row's place among the rows that share a 256-byte span of banks,
// so 16 consecutive rows at one chunk hit 16 different bank groups.
template <int K> __device__ __forceinline__ int sy_slot(int row, int c) {
  constexpr int CH = K / 2, RW = 16 / CH;
  return row * CH + (c ^ ((row / RW) % CH));
}

__host__ __device__ inline int sy_tiles(int n) {
  const int nt = (n + SY_T - 1) / SY_T;
  return nt * (nt + 1) / 2;
}

template <int K>
__global__ __launch_bounds__(SY_THREADS) void lg_sp2_sym_kernel(const double* X, double* Xn, int n, int np,
                                                               const double* st, double* red, const int* map,
                                                               int chunk) {
  static_assert(K == 8 || K == 16 || K == 32, "slab width");
  constexpr int NT = SY_THREADS;
  constexpr int CH = K / 2, LQ = SY_T * CH / NT, LC = CH / 4;  // chunks: a row, a thread's loads, a lane's
  if (st[SP2S_STATUS] != 0.0) return;
  __shared__ double2 slab[2][2][SY_T * CH];  // [buffer][panel I, panel J][slot]
  const bool sq = st[SP2S_SQ] != 0.0;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int li = lane & 15, lg = lane >> 4;
  const int wr = (wv >> 1) * 32, wc = (wv & 1) * 32;  // the wave's quarter of the tile
  const int ntile = sy_tiles(n), nk = (n + K - 1) / K;
  const int per = (int)(gridDim.x >> 3);  // workgroups of each XCD
  double tr = 0.0, fr = 0.0;
  for (int pos = (int)(blockIdx.x >> 3); pos < chunk; pos += per) {
    const int slot = (int)(blockIdx.x & 7) * chunk + pos;
    if (slot >= ntile) break;
    const int I = map[slot] >> 16, J = map[slot] & 0xffff;
    const bool diag = I == J;  // one panel serves both operands
    const int np_ = diag ? 1 : 2;
    const int r0[2] = {I * SY_T, J * SY_T};
    // chunk q * 256 + tid of a slab: panel row id / CH, 16-byte column chunk id % CH.  The loads
    // are unconditional (at a clamped address inside X's n x n part), so that all of a slab's
    // loads are in flight over the previous slab's MFMAs; values beyond n are zeroed as they are
    // stored to LDS.
    double2 g[2][LQ];
    const int kmax = (n - 1) & ~1;  // the last even column: kmax + 1 < np
    auto fetch = [&](int k0) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        if (p >= np_) break;
#pragma unroll
        for (int q = 0; q < LQ; ++q) {
          const int id = q * NT + tid, row = r0[p] + id / CH, k = k0 + 2 * (id % CH);
          g[p][q] = *(const double2*)(X + (size_t)min(row, n - 1) * np + min(k, kmax));
        }
      }
    };
    auto put = [&](int buf, int k0) {
#pragma unroll
      for (int p = 0; p < 2; ++p) {
        if (p >= np_) break;
#pragma unroll
        for (int q = 0; q < LQ; ++q) {
          const int id = q * NT + tid, row = r0[p] + id / CH, k = k0 + 2 * (id % CH);
          double2 v = g[p][q];
          const bool in = row < n && k < n;
          v.x = in ? v.x : 0.0;
          v.y = in && k + 1 < n ? v.y : 0.0;
          slab[buf][p][sy_slot<K>(id / CH, id % CH)] = v;
        }
      }
    };
    double4_t acc[2][2];
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v) acc[u][v] = double4_t{0.0, 0.0, 0.0, 0.0};
    fetch(0);
    put(0, 0);
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int buf = kt & 1;
      if (kt + 1 < nk) fetch((kt + 1) * K);
      const double2* A = slab[buf][0];
      const double2* B = slab[buf][np_ - 1];
#pragma unroll
      for (int s2 = 0; s2 < LC; ++s2) {
        const int c = lg * LC + s2;
        double2 a[2], b[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          a[u] = A[sy_slot<K>(wr + 16 * u + li, c)];
          b[u] = B[sy_slot<K>(wc + 16 * u + li, c)];
        }
        // A[i][k] = X[I row i][k], B[k][j] = X[J row j][k]: lane l holds A[l % 16][l / 16], B[l / 16][l % 16]
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int v = 0; v < 2; ++v)
              acc[u][v] = __builtin_amdgcn_mfma_f64_16x16x4f64(h ? a[u].y : a[u].x, h ? b[v].y : b[v].x, acc[u][v], 0,
                                                               0, 0);
      }
      if (kt + 1 < nk) put(buf ^ 1, (kt + 1) * K);
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int v = 0; v < 2; ++v)
#pragma unroll
        for (int r = 0; r < 4; ++r) {  // item r: row 4 r + l / 16, column l % 16 of the block
          const int i = I * SY_T + wr + 16 * u + 4 * r + lg, j = J * SY_T + wc + 16 * v + li;
          if (i < n && j < n && (!diag || i >= j)) {
            const double y = sq ? acc[u][v][r] : 2.0 * X[(size_t)i * np + j] - acc[u][v][r];
            Xn[(size_t)i * np + j] = y;
            if (i != j) Xn[(size_t)j * np + i] = y;
            fr += (i == j ? 1.0 : 2.0) * y * y;
            if (i == j) tr += y;
          }
        }
  }
  tr = wave_sum(tr);
  fr = wave_sum(fr);
  double* part = (double*)&slab[0][0][0];  // every slab read has passed the loop's last barrier
  if (lane == 0) {
    part[2 * wv] = tr;
    part[2 * wv + 1] = fr;
  }
  __syncthreads();
  if (tid == 0) {
    double a = 0.0, c = 0.0;
    for (int q = 0; q < SY_THREADS / 64; ++q) {
      a += part[2 * q];
      c += part[2 * q + 1];
    }
    red[2 * (size_t)blockIdx.x] = a;
    red[2 * (size_t)blockIdx.x + 1] = c;
  }
}

// A whole Lanczos step for matrices up to LARGE_LZ_FUSED_N AOs in one block: w = A q (A symmetric,
// np-strided, rows read along their length), then the vector work of lg_lanczos_vec_kernel.
__global__ __launch_bounds__(LZ_VEC_THREADS) void lg_lanczos_fused_kernel(const double* A, int np, double* q, double* qp,
                                                                          double* w, int n, double* al, double* be,
                                                                          int k) {
  __shared__ double part[LZ_VEC_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = wv; i < n; i += LZ_VEC_THREADS / 64) {
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += A[(size_t)i * np + j] * q[j];
    s = wave_sum(s);
    if (lane == 0) w[i] = s;
  }
  __syncthreads();
  auto block_total = [&](double x) {
    x = wave_sum(x);
    __syncthreads();
    if (lane == 0) part[wv] = x;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < LZ_VEC_THREADS / 64; ++i) s += part[i];
    return s;
  };
  const double bprev = k > 0 ? be[k - 1] : 0.0;
  double a = 0.0;
  for (int i = tid; i < n; i += LZ_VEC_THREADS) {
    const double x = w[i] - bprev * qp[i];
    w[i] = x;
    a += q[i] * x;
  }
  const double alpha = block_total(a);
  double b2 = 0.0;
  for (int i = tid; i < n; i += LZ_VEC_THREADS) {
    const double x = w[i] - alpha * q[i];
    w[i] = x;
    b2 += x * x;
  }
  const double beta = sqrt(block_total(b2));
  if (tid == 0) {
    al[k] = alpha;
    be[k] = beta;
  }
  const double ib = beta > 0.0 ? 1.0 / beta : 0.0;
  __syncthreads();  // every thread has read q in the products above before it is overwritten
  for (int i = tid; i < n; i += LZ_VEC_THREADS) {
    qp[i] = q[i];
    q[i] = w[i] * ib;
  }
}

// The frontier seed's products when the state asks for one (the tail has just started): out = X v
// for the symmetric np-strided X, one wave per row with lanes along it; diff: s = t - u.
__global__ __launch_bounds__(LG_THREADS) void lg_seed_gemv_kernel(const double* X, const double* v, double* out, int n,
                                                                  int np, const double* st) {
  if (st[SP2S_SEED] == 0.0 || st[SP2S_STATUS] != 0.0) return;
  const int lane = threadIdx.x & 63;
  for (int i = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6); i < n; i += (int)((gridDim.x * blockDim.x) >> 6)) {
    double s = 0.0;
    for (int j = lane; j < n; j += 64) s += X[(size_t)i * np + j] * v[j];
    s = wave_sum(s);
    if (lane == 0) out[i] = s;
  }
}
__global__ __launch_bounds__(LG_THREADS) void lg_seed_diff_kernel(const double* t, const double* u, double* sv, int n,
                                                                  const double* st) {
  if (st[SP2S_SEED] == 0.0 || st[SP2S_STATUS] != 0.0) return;
  for (long long i = gtid(); i < n; i += gstride()) sv[i] = t[i] - u[i];
}

// Start of the bound-estimating Lanczos run: q_i = 1 + sin(1.7 i + 0.3) / 2, normalised, and qp = 0.
__global__ __launch_bounds__(LZ_VEC_THREADS) void lg_extremes_start_kernel(double* q, double* qp, int n) {
  __shared__ double part[LZ_VEC_THREADS / 64];
  const int tid = threadIdx.x;
  double s = 0.0;
  for (int i = tid; i < n; i += LZ_VEC_THREADS) {
    const double x = 1.0 + 0.5 * sin(1.7 * (double)i + 0.3);
    q[i] = x;
    qp[i] = 0.0;
    s += x * x;
  }
  s = wave_sum(s);
  if ((tid & 63) == 0) part[tid >> 6] = s;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < LZ_VEC_THREADS / 64; ++i) t += part[i];
  const double inv = 1.0 / sqrt(t);
  for (int i = tid; i < n; i += LZ_VEC_THREADS) q[i] *= inv;
}

// The vector work of one Lanczos step after w = A q (lanczos_step): w -= beta_{k-1} qp,
// alpha = q.w, w -= alpha q, beta = |w|, then qp = q and q = w / beta (q = 0 after a breakdown to
// zero); alpha and beta go to al[k], be[k] on the device.  One block, fixed reduction order.
__global__ __launch_bounds__(LZ_VEC_THREADS) void lg_lanczos_vec_kernel(double* q, double* qp, double* w, int n,
                                                                        double* al, double* be, int k) {
  __shared__ double part[LZ_VEC_THREADS / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  auto block_total = [&](double x) {
    x = wave_sum(x);
    __syncthreads();
    if (lane == 0) part[wv] = x;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < LZ_VEC_THREADS / 64; ++i) s += part[i];
    return s;
  };
  const double bprev = k > 0 ? be[k - 1] : 0.0;
  double a = 0.0;
  for (int i = tid; i < n; i += LZ_VEC_THREADS) {
    const double x = w[i] - bprev * qp[i];
    w[i] = x;
    a += q[i] * x;
  }
  const double alpha = block_total(a);
  double b2 = 0.0;
  for (int i = tid; i < n; i += LZ_VEC_THREADS) {
    const double x = w[i] - alpha * q[i];
    w[i] = x;
    b2 += x * x;
  }
  const double beta = sqrt(block_total(b2));
  if (tid == 0) {
    al[k] = alpha;
    be[k] = beta;
  }
  const double ib = beta > 0.0 ? 1.0 / beta : 0.0;
  for (int i = tid; i < n; i += LZ_VEC_THREADS) {
    qp[i] = q[i];
    q[i] = w[i] * ib;
  }
}

// Pdst = 2 X on the n x n block.
__global__ __launch_bounds__(LG_THREADS) void lg_twice_kernel(const double* X, double* Pdst, int n, int np) {
  for (long long e = gtid(); e < (long long)n * n; e += gstride()) {
    const size_t o = (size_t)(e / n) * np + (size_t)(e % n);
    Pdst[o] = 2.0 * X[o];
  }
}

// Sums of the first k columns of the block partials into out[0..k), so that the host reads k
// values instead of every block's: one thread per block partial and a fixed halving tree.
__global__ __launch_bounds__(LG_BLOCKS) void lg_finish_kernel(const double* red, int k, double* out) {
  __shared__ double t[LG_BLOCKS];
  for (int q = 0; q < k; ++q) {
    t[threadIdx.x] = red[(size_t)threadIdx.x * LG_MAXDOT + q];
    __syncthreads();
    for (int h = LG_BLOCKS / 2; h > 0; h >>= 1) {
      if ((int)threadIdx.x < h) t[threadIdx.x] += t[threadIdx.x + h];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = t[0];
    __syncthreads();
  }
}

// Block partials of sum (X - Y)^2 (red[block]).
__global__ __launch_bounds__(LG_THREADS) void lg_diff2_kernel(const double* X, const double* Y, long long m, double* red) {
  __shared__ double part[LG_THREADS / 64];
  double v = 0.0;
  for (long long e = gtid(); e < m; e += gstride()) {
    const double d = X[e] - Y[e];
    v += d * d;
  }
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < LG_THREADS / 64; ++q) t += part[q];
    red[blockIdx.x] = t;
  }
}

// Diagonal blocks of H_core and block partials of the core-core energy (red[block]).
__global__ __launch_bounds__(LG_THREADS) void lg_core_kernel(KArgs A, MolDesc md, double* red) {
  __shared__ double part[LG_THREADS / 64];
  const Mol M = mol_of(A, md, nullptr);
  core_diag(A, M, gtid(), gstride());
  const long long npairs = (long long)M.nat * (M.nat - 1) / 2;
  double v = field_nuclear(A, M, gtid(), gstride());
  for (long long p = gtid(); p < npairs; p += gstride()) v += M.ecc[p];
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int q = 0; q < LG_THREADS / 64; ++q) t += part[q];
    red[blockIdx.x] = t;
  }
}

// Level shift on the virtual space, out = F + shift (I - P / 2) on the n x n block (in place when
// out == F); the padding is left alone.
__global__ __launch_bounds__(LG_THREADS) void lg_shift_kernel(const double* F, const double* P, double* out, int n,
                                                              int np, double shift) {
  for (long long e = gtid(); e < (long long)n * n; e += gstride()) {
    const int i = (int)(e / n), j = (int)(e % n);
    out[(long long)i * np + j] = F[(long long)i * np + j] + shift * ((i == j ? 1.0 : 0.0) - 0.5 * P[(long long)i * np + j]);
  }
}

// Density damping P <- (1 - damp) Pn + damp P over all m elements (zero padding stays zero).
__global__ __launch_bounds__(LG_THREADS) void lg_damp_kernel(const double* Pn, double* P, long long m, double damp) {
  for (long long e = gtid(); e < m; e += gstride()) P[e] = (1.0 - damp) * Pn[e] + damp * P[e];
}

// Superposition density on a zeroed P; `tilt_atom` >= 0 moves min(0.2, P_zz) from p_z to p_x on
// that atom (the symmetry-broken guess, src/scf.rs:894-916).
__global__ __launch_bounds__(LG_THREADS) void lg_sad_kernel(KArgs A, MolDesc md, double* P, int tilt_atom) {
  const int np = md.np;
  for (int a = gtid(); a < md.nat; a += gstride()) {
    const int ga = md.atom0 + a;
    const ElemConst& e = c_tab.el[A.aslot[ga]];
    const int o = A.aoff[ga];
    P[(long long)o * np + o] = e.ios;
    if (e.norb == 4) {
      double pp[3] = {e.iop / 3.0, e.iop / 3.0, e.iop / 3.0};
      if (a == tilt_atom) {
        const double t = fmin(0.2, pp[2]);
        pp[2] -= t;
        pp[0] += t;
      }
      for (int k = 1; k < 4; ++k) P[(long long)(o + k) * np + o + k] = pp[k - 1];
    }
  }
}

__global__ __launch_bounds__(LG_THREADS) void lg_fock_pairs_kernel(KArgs A, MolDesc md, const double* P, double* F,
                                                                   double* slot) {
  const Mol M = mol_of(A, md, slot);
  fock_pairs(A, M, M.H, P, P, 0.5, F, gtid(), gstride());
}

// The sums of fock_jsum with one wave per output (heavy atom and component, or hydrogen): lanes
// over the partner atoms, then a fixed shuffle tree, so the whole grid works on a single molecule's
// few thousand outputs and the sums stay reproducible.
// Phase 2 of the Fock build over a grid, one wave per output (J[a][q] of a heavy atom, J[h] of a
// hydrogen), partners summed lane-strided then by a fixed shuffle tree.  SW weights the
// hydrogen-hydrogen Coulomb terms by their pair switch (the other pairs' partials carry theirs).
template <bool SW>
__device__ inline void fock_jsum_wide(const KArgs& A, const Mol& M, const double* P, double* J, const double* sw,
                                      int wave0, int nwaves) {
  const int np = M.np, lane = threadIdx.x & 63;
  const double* JX = M.slot;
  const double* JH = M.slot + 20LL * M.nxx;
  const int nitems = M.nh * 10 + M.nH;
  for (int e = wave0; e < nitems; e += nwaves) {
    double s = 0.0;
    if (e < M.nh * 10) {
      const int a = e / 10, q = e % 10;
      for (int b = lane; b < M.nh; b += 64) {
        if (b == a) continue;
        s += b > a ? JX[(long long)q * M.nxx + tri_index(a, b, M.nh)]
                   : JX[(long long)(10 + q) * M.nxx + tri_index(b, a, M.nh)];
      }
      const double* jh = JH + (long long)q * M.nxh + (long long)a * M.nH;
      for (int h = lane; h < M.nH; h += 64) s += jh[h];
    } else {
      const int h = e - M.nh * 10;
      for (int a = lane; a < M.nh; a += 64) s += JH[10LL * M.nxh + (long long)a * M.nH + h];
      for (int g = lane; g < M.nH; g += 64) {
        if (g == h) continue;
        const int lo = g < h ? g : h, hi = g < h ? h : g;
        const int og = A.aoff[M.atom0 + M.nh + g];
        const int t = tri_index(lo, hi, M.nH);
        if (SW) s += sw[M.nxx + M.nxh + t] * (M.Whh[t] * P[(long long)og * np + og]);
        else s += M.Whh[t] * P[(long long)og * np + og];
      }
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) J[e < M.nh * 10 ? e : (M.nh + (e - M.nh * 10)) * 10] = s;
  }
}

__global__ __launch_bounds__(LG_THREADS) void lg_fock_jsum_wide_kernel(KArgs A, MolDesc md, const double* P, double* J,
                                                                       double* slot) {
  const Mol M = mol_of(A, md, slot);
  fock_jsum_wide<false>(A, M, P, J, nullptr, (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6),
                        (int)((gridDim.x * blockDim.x) >> 6));
}

__global__ __launch_bounds__(LG_THREADS) void lg_fock_diag_kernel(KArgs A, MolDesc md, const double* P, double* F,
                                                                  const double* J) {
  const Mol M = mol_of(A, md, nullptr);
  fock_diag(A, M, M.H, P, P, 0.5, F, J, gtid(), gstride());
}

// The CPHF response operator G(dP) = J(dP) - K(dP / 2) with no H_core (response_fock_rhf,
// src/fock.rs:444-485) for many right-hand sides at once: blockIdx.y is the right-hand side r,
// its density dP_r and its output at r * m2 (leading dimension np), its Coulomb partials at
// r * w_cap of `slots` and its per-atom sums at r * 10 nat of J.  Hz is a zero np x np matrix (the
// Fock phases add H).  The same three phases as the Fock build, each a grid over the pairs or
// atoms of every right-hand side; SW applies the per-pair switch (null: every pair at 1).
template <bool SW>
__global__ __launch_bounds__(LG_THREADS) void rf_pairs_kernel(KArgs A, MolDesc md, const double* dP, double* G,
                                                              const double* Hz, const double* sw, double* slots,
                                                              long long m2, long long wcap) {
  const long long r = blockIdx.y;
  const Mol M = mol_of(A, md, slots + r * wcap);
  fock_pairs<SW>(A, M, Hz, dP + r * m2, dP + r * m2, 0.5, G + r * m2, (int)(blockIdx.x * blockDim.x + threadIdx.x),
                 (int)(gridDim.x * blockDim.x), sw);
}
template <bool SW>
__global__ __launch_bounds__(LG_THREADS) void rf_jsum_kernel(KArgs A, MolDesc md, const double* dP, double* J,
                                                             const double* sw, double* slots, long long m2,
                                                             long long wcap) {
  const long long r = blockIdx.y;
  const Mol M = mol_of(A, md, slots + r * wcap);
  fock_jsum_wide<SW>(A, M, dP + r * m2, J + r * 10LL * md.nat, sw, (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6),
                     (int)((gridDim.x * blockDim.x) >> 6));
}
__global__ __launch_bounds__(LG_THREADS) void rf_diag_kernel(KArgs A, MolDesc md, const double* dP, double* G,
                                                             const double* Hz, const double* J, long long m2) {
  const long long r = blockIdx.y;
  const Mol M = mol_of(A, md, nullptr);
  fock_diag(A, M, Hz, dP + r * m2, dP + r * m2, 0.5, G + r * m2, J + r * 10LL * md.nat,
            (int)(blockIdx.x * blockDim.x + threadIdx.x), (int)(gridDim.x * blockDim.x));
}

// DIIS entry of one iteration: E = T - T^T with T = F P (commutator, src/scf.rs:1039-1054), and
// copies of F and (for A-DIIS) P into the history slots.
__global__ __launch_bounds__(LG_THREADS) void lg_commutator_kernel(const double* T, const double* F, const double* P,
                                                                   double* En, double* Fn, double* Dn, int n, int np,
                                                                   int keep_d) {
  const long long m = (long long)np * np;
  for (long long e = gtid(); e < m; e += gstride()) {
    const int i = (int)(e / np), j = (int)(e % np);
    En[e] = (i < n && j < n) ? T[(long long)i * np + j] - T[(long long)j * np + i] : 0.0;
    Fn[e] = F[e];
    if (keep_d) Dn[e] = P[e];
  }
}

struct Combine {
  double c[MAXD];
  int slot[MAXD];
  int len;
};

__global__ __launch_bounds__(LG_THREADS) void lg_combine_kernel(const double* Fh, long long ms, Combine C, double* Fu) {
  for (long long e = gtid(); e < ms; e += gstride()) {
    double s = 0.0;
    for (int i = 0; i < C.len; ++i) s += C.c[i] * Fh[(long long)C.slot[i] * ms + e];
    Fu[e] = s;
  }
}

// Row-major n x n eigenvector matrix (columns = eigenvectors) from the column-major solver output.
__global__ __launch_bounds__(LG_THREADS) void lg_vectors_kernel(const double* V, double* C, int n, int np) {
  const long long m = (long long)n * n;
  for (long long e = gtid(); e < m; e += gstride()) {
    const int i = (int)(e / n), j = (int)(e % n);
    C[e] = V[(long long)j * np + i];
  }
}

// n x n dense copy out of a leading-dimension-np matrix.
__global__ __launch_bounds__(LG_THREADS) void lg_pack_kernel(const double* X, double* out, int n, int np) {
  const long long m = (long long)n * n;
  for (long long e = gtid(); e < m; e += gstride()) out[e] = X[(e / n) * np + e % n];
}

__global__ __launch_bounds__(NT) void lg_epilogue_kernel(KArgs A, MolDesc md, const double* P, double* dip_out) {
  __shared__ double red[NW * 4];
  const Mol M = mol_of(A, md, nullptr);
  double dip[3];
  charges_dipole(A, M, P, red, dip);
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) dip_out[k] = dip[k];
}

#endif  // PM3_KERNEL_VARIANT

#define HIP_OK(x)                                                  \
  do {                                                             \
    hipError_t e_ = (x);                                           \
    if (e_ != hipSuccess) {                                        \
      err = std::string(#x) + ": " + hipGetErrorString(e_);        \
      return false;                                                \
    }                                                              \
  } while (0)

}  // namespace

// Entry points of the SCF kernel's narrower builds (one translation unit each; their device code
// objects, element tables and kernels are their own).  KArgs crosses as bytes: both sides compile
// the same definition.
#define PM3_VCAT2(a, b) a##_##b
#define PM3_VCAT(a, b) PM3_VCAT2(a, b)
namespace pm3variant {
#ifdef PM3_KERNEL_VARIANT
size_t PM3_VCAT(PM3_KERNEL_VARIANT, kargs_bytes)() { return sizeof(KArgs); }
hipError_t PM3_VCAT(PM3_KERNEL_VARIANT, table)(const void* table, size_t bytes) {
  return hipMemcpyToSymbol(HIP_SYMBOL(c_tab), table, bytes);
}
hipError_t PM3_VCAT(PM3_KERNEL_VARIANT, prepare)(int lds_bytes) {
  return hipFuncSetAttribute((const void*)PM3_SCF_KERNEL, hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
}
hipError_t PM3_VCAT(PM3_KERNEL_VARIANT, launch)(const void* args, int nwg, int lds_bytes, hipStream_t s) {
  KArgs a;
  std::memcpy(&a, args, sizeof(KArgs));
  hipLaunchKernelGGL(PM3_SCF_KERNEL, dim3(nwg), dim3(NT), lds_bytes, s, a);
  return hipGetLastError();
}
#else
size_t nt256_kargs_bytes();
hipError_t nt256_table(const void* table, size_t bytes);
hipError_t nt256_prepare(int lds_bytes);
hipError_t nt256_launch(const void* args, int nwg, int lds_bytes, hipStream_t s);
size_t nt128_kargs_bytes();
hipError_t nt128_table(const void* table, size_t bytes);
hipError_t nt128_prepare(int lds_bytes);
hipError_t nt128_launch(const void* args, int nwg, int lds_bytes, hipStream_t s);
#endif
}  // namespace pm3variant

#ifndef PM3_KERNEL_VARIANT

namespace pm3host {

// Synchronous copy and fill on one stream: the stream's earlier work is ordered before them and
// they are complete on return (the null stream is never used, so batches on other host threads
// proceed concurrently).
static hipError_t scopy(hipStream_t s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
  return e == hipSuccess ? hipStreamSynchronize(s) : e;
}
static hipError_t smemset(hipStream_t s, void* dst, int value, size_t bytes) {
  hipError_t e = hipMemsetAsync(dst, value, bytes, s);
  return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

// Device-resident batch: inputs uploaded and scratch allocated once; solve() is the launch only.
// Every device operation of a batch goes to its own stream (created non-blocking at the default
// priority), so calls from several host threads do not serialise on the null stream.
struct LargeBufs {
  double *F = nullptr, *P = nullptr, *Pn = nullptr, *Fu = nullptr, *T = nullptr, *V = nullptr, *w = nullptr;
  double *ework = nullptr, *Pb = nullptr, *Vb = nullptr, *wb = nullptr, *Fh = nullptr, *Eh = nullptr, *Dh = nullptr;
  double* Pd = nullptr;  // a damped path's final density while its undamped projector gives the frontier levels
  double *J = nullptr, *slot = nullptr, *red = nullptr, *dip = nullptr, *fin = nullptr;
  double* lz = nullptr;                          // Lanczos vectors (3 np)
  double* fs = nullptr;                          // frontier seed s and X s (np each), then ...
  double* fr = nullptr;                          // ... the generic start vector r (np_max)
  double** bp = nullptr;                         // device pointer arrays of the block-lower products
  bool seeded = false;                           // a purification of this path has written the seed
  bool last_eig = false;                         // the path's last density came from the eigensolver
  double bnd_lo = 0.0, bnd_hi = 0.0;            // cached spectral-bound estimate of the purification
  int bnd_age = -1;
  bool no_tight = false;                         // Gershgorin bounds only (a path being rerun)
  int* info = nullptr;
  long long m2 = 0;  // elements per matrix (np_max^2)
  int depth = 0;     // DIIS history slots allocated
  double* sp2s = nullptr;  // device state of the purification loop (SP2S_* slots)
  double* lzab = nullptr;  // Lanczos coefficients alpha (first half) and beta of the current run
  // device time of the purification products and of the Fock builds in the last solve: (start,
  // end) event pairs recorded as they are launched and read once the solve has synchronised
  std::vector<hipEvent_t> pev, fev;
  int npev = 0, nfev = 0;
  double product_ms = 0.0, fock_ms = 0.0;
  long long products = 0, fock_builds = 0, product_flop = 0;
  int sp2_steps = 8;  // steps of the last converged purification (when the tile kernel first looks)
  int* sy_map = nullptr;  // the tile kernel's tile list (I << 16 | J) for sy_map_n AOs
  int sy_map_n = 0, sy_map_cap = 0;
  long long lanczos_steps = 0, lanczos_runs = 0;  // frontier-level Lanczos of the last solve
};

// The SCF launch of one workgroup width: its molecules are order_all[off, off + n), pulled from its
// own queue counter by nwg persistent workgroups of `lds` bytes of dynamic LDS, each with a slab of
// `stride` doubles of its own scratch region.
struct ScfClass {
  int n = 0, off = 0, nwg = 0, lds = 0;
  int np_max = 16;
  long long w_cap = 1, stride = 0, ustride = 0, scr_off = 0;
  bool any_uhf = false;
};
constexpr int SCF_WIDTH[3] = {512, 256, 128};  // threads per workgroup of each class
constexpr int SCF_PER_CU[3] = {1, 2, 4};       // workgroups resident per CU (256 VGPRs per lane each)

struct Batch {
  int nmol = 0, nsolve = 0, nwg = 0, lds_need = 0;
  ScfClass cls[3];                  // 512-, 256- and 128-thread launches
  double* scr_base = nullptr;       // the classes' scratch regions start here (offsets scr_off)
  hipStream_t aux[2] = {nullptr, nullptr};  // the narrower launches run beside the widest one
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  std::vector<MolDesc> desc;
  std::vector<int> solve;
  std::vector<int> large;      // molecules solved by the host-driven multi-kernel path
  std::vector<int> order_all;  // queue order of the pair kernels: solve (largest first), then large
  LargeBufs lg;
  rocblas_handle blas = nullptr;
  // divide-and-conquer: a stream, solver handle and completion event for every size bucket after
  // the first (which runs on stream and blas), so the buckets' batched eigensolves overlap
  struct DcLane {
    hipStream_t s = nullptr;
    rocblas_handle h = nullptr;
    hipEvent_t done = nullptr;
  };
  std::vector<DcLane> dc_lanes;
  hipEvent_t dc_fork = nullptr;
  EngineOptions opt{};
  TableConst table{};
  std::vector<double> h_red;
  std::vector<std::vector<int>> ao_map, atom_map;  // sorted -> caller indices
  std::vector<int> pre_status;
  std::vector<int8_t> cls_of;          // SCF class (workgroup width) of each molecule
  std::vector<int> pre_z;              // the refused atomic number of a status-2 / 4 molecule
  std::vector<double> pre_req, pre_lim;  // pair-cache size and ceiling (MiB) of a status-5 molecule
  bool want_gradient = false;
  bool keep_density = false;  // the caller asked for densities: copy them back at each fetch
  bool warm_start = false;    // solves after the first start from each molecule's last density
  bool have_prev = false;     // out_P / out hold densities carried in from another batch
  bool purify = false;  // densities by purification; orbitals are then not produced
  bool any_uhf = false; // some molecule takes the unrestricted path (the slab then holds a beta part)
  long long mat_total = 0, vec_total = 0, atoms_total = 0;
  KArgs args{};
  void* bufs[128] = {nullptr};
  int nbufs = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  double last_kernel_ms = 0.0, last_pair_ms = 0.0, last_guess_ms = 0.0;
  int pair_blocks = 1, pair_block = 256;
  bool large_syrk = false;
  int large_split = 0;
  int large_steps_n = 0;  // the fused purification kernel up to this many AOs
  int sy_k = 16;          // slab width of the tile kernel (8, 16 or 32 columns)
  // learned guess: directed pairs of the eligible molecules and their network inputs/outputs
  const GuessModel* guess_model = nullptr;
  bool learned = false, solved_once = false;
  long long ngp_total = 0;
  long long gp_cap = 0;          // directed pairs the device pair lists and network buffers hold
  std::vector<int> eligible;
  std::vector<int8_t> h_elem8;   // guess element index of every atom (engine order), -1 outside H, C, O
  uint16_t* d_X = nullptr;
  float* d_Y = nullptr;
  hipEvent_t evg = nullptr, evm = nullptr;
  double last_feature_ms = 0.0, last_network_ms = 0.0;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  // allocation-free mode: every buffer is carved from this device block, the stream is borrowed,
  // and no events are created (no timing)
  unsigned char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  ~Batch() {
    if (stream) (void)hipStreamSynchronize(stream);
    for (int i = 0; i < nbufs; ++i)
      if (bufs[i]) (void)hipFree(bufs[i]);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev2) (void)hipEventDestroy(ev2);
    if (evg) (void)hipEventDestroy(evg);
    if (evm) (void)hipEventDestroy(evm);
    for (DcLane& l : dc_lanes) {
      if (l.s) (void)hipStreamSynchronize(l.s);
      if (l.h) (void)rocblas_destroy_handle(l.h);
      if (l.done) (void)hipEventDestroy(l.done);
      if (l.s) (void)hipStreamDestroy(l.s);
    }
    if (dc_fork) (void)hipEventDestroy(dc_fork);
    if (blas) (void)rocblas_destroy_handle(blas);
    for (hipEvent_t e : lg.pev) (void)hipEventDestroy(e);
    for (hipEvent_t e : lg.fev) (void)hipEventDestroy(e);
    for (int i = 0; i < 2; ++i) {
      if (aux[i]) (void)hipStreamSynchronize(aux[i]);
      if (ev_join[i]) (void)hipEventDestroy(ev_join[i]);
      if (aux[i]) (void)hipStreamDestroy(aux[i]);
    }
    if (ev_fork) (void)hipEventDestroy(ev_fork);
    if (stream && own_stream) (void)hipStreamDestroy(stream);
  }
  bool ensure_guess_buffers(std::string& err) {
    if (d_X || gp_cap == 0) return true;
    return alloc(d_X, (size_t)gp_cap * pm3guess::NF, err) && alloc(d_Y, (size_t)gp_cap * pm3guess::NO, err) &&
           set_y();
  }
  // frees one buffer of this batch (not in the allocation-free mode) after the stream drains
  template <class T> bool release(T*& p, std::string& err) {
    if (!p) return true;
    if (arena) {
      err = "internal: release in the allocation-free mode";
      return false;
    }
    for (int i = 0; i < nbufs; ++i)
      if (bufs[i] == (void*)p) {
        hipError_t e = hipStreamSynchronize(stream);
        if (e == hipSuccess) e = hipFree(bufs[i]);
        if (e != hipSuccess) {
          err = std::string("hipFree: ") + hipGetErrorString(e);
          return false;
        }
        bufs[i] = bufs[--nbufs];
        bufs[nbufs] = nullptr;
        p = nullptr;
        return true;
      }
    err = "internal: release of a buffer the batch does not own";
    return false;
  }
  bool set_y() {
    args.gY = d_Y;
    return true;
  }
  template <class T> bool alloc(T*& p, size_t count, std::string& err) {
    if (arena) {
      const size_t bytes = (sizeof(T) * std::max<size_t>(count, 1) + 255) & ~(size_t)255;
      if (arena_used + bytes > arena_cap) {
        err = "the molecule's buffers need more than the " + std::to_string(arena_cap >> 20) +
              " MiB of the allocation-free scratch block";
        return false;
      }
      p = reinterpret_cast<T*>(arena + arena_used);
      arena_used += bytes;
      return true;
    }
    if (nbufs >= (int)(sizeof(bufs) / sizeof(bufs[0]))) {
      err = "internal: too many batch buffers";
      return false;
    }
    hipError_t e = hipMalloc((void**)&p, sizeof(T) * std::max<size_t>(count, 1));
    if (e != hipSuccess) {
      err = std::string("hipMalloc: ") + hipGetErrorString(e);
      return false;
    }
    bufs[nbufs++] = p;
    return true;
  }
};

// The element table in constant memory, once per device (the table is the same for every call).
static bool upload_table(const Params& P, std::string& err) {
  static std::mutex mu;
  static bool uploaded[64] = {false};
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lock(mu);
  if (dev < 0 || dev >= 64 || !uploaded[dev]) {
    HIP_OK(hipMemcpyToSymbol(HIP_SYMBOL(c_tab), &P.table, sizeof(TableConst)));
    // the narrower SCF builds hold their own copies of the table
    HIP_OK(pm3variant::nt256_table(&P.table, sizeof(TableConst)));
    HIP_OK(pm3variant::nt128_table(&P.table, sizeof(TableConst)));
    if (dev >= 0 && dev < 64) uploaded[dev] = true;
  }
  return true;
}


// ------------------------------------------------------------------------------------------
// Host-driven SCF for large molecules (see the lg_* kernels).  The loop mirrors rhf_loop step for
// step: Fock build, energy, DIIS entry and coefficients (the same solver code, on the host),
// eigensolve, aufbau density, convergence test, then the stability search of scf_batch_kernel.
#define BLAS_OK(x)                                                    \
  do {                                                                \
    rocblas_status s_ = (x);                                          \
    if (s_ != rocblas_status_success) {                               \
      err = std::string(#x) + ": " + rocblas_status_to_string(s_);    \
      return false;                                                   \
    }                                                                 \
  } while (0)

static int diis_depth_mats(int n, int mats, int scf_memory_mb) {
  const long long n2 = (long long)n * n * 8, per = mats * n2;
  const long long budget = (long long)scf_memory_mb * 1024 * 1024;
  if (scf_memory_mb == 0 || per == 0) return MAXD;
  return (int)std::max(2LL, std::min((long long)MAXD, budget / per));
}

static int diis_depth_for(int n, int keep_d, int scf_memory_mb) {
  const long long n2 = (long long)n * n * 8, per = (keep_d ? 3 : 2) * n2;
  const long long budget = (long long)scf_memory_mb * 1024 * 1024;
  if (scf_memory_mb == 0 || per == 0) return MAXD;
  return (int)std::max(2LL, std::min((long long)MAXD, budget / per));
}

static bool large_setup(Batch& b, std::string& err) {
  int np = 16, nat = 1, depth = 2;
  long long slot = 1;
  for (int k : b.large) {
    const MolDesc& d = b.desc[k];
    np = std::max(np, d.np);
    nat = std::max(nat, d.nat);
    const long long nh = d.nheavy, nH = d.nat - d.nheavy;
    slot = std::max(slot, 20 * (nh * (nh - 1) / 2) + 11 * nh * nH);
    const int keep_d = b.opt.accel == 2 &&
                       (b.opt.scf_memory_mb == 0 || 2LL * 3 * d.nao * d.nao * 8 <= (long long)b.opt.scf_memory_mb * 1024 * 1024);
    depth = std::max(depth, diis_depth_for(d.nao, keep_d ? 1 : 0, b.opt.scf_memory_mb));
  }
  LargeBufs& L = b.lg;
  L.m2 = (long long)np * np;
  L.depth = depth;
  if (!b.alloc(L.F, L.m2, err) || !b.alloc(L.P, L.m2, err) || !b.alloc(L.Pn, L.m2, err) || !b.alloc(L.Fu, L.m2, err) ||
      !b.alloc(L.T, L.m2, err) || !b.alloc(L.V, L.m2, err) || !b.alloc(L.w, np, err) || !b.alloc(L.ework, np, err) ||
      !b.alloc(L.Pb, L.m2, err) || !b.alloc(L.Vb, L.m2, err) || !b.alloc(L.wb, np, err) || !b.alloc(L.Pd, L.m2, err) ||
      !b.alloc(L.Fh, L.m2 * depth, err) || !b.alloc(L.Eh, L.m2 * depth, err) || !b.alloc(L.Dh, L.m2 * depth, err) ||
      !b.alloc(L.J, 10LL * nat, err) || !b.alloc(L.slot, slot, err) ||
      !b.alloc(L.red, (size_t)LG_BLOCKS * LG_MAXDOT, err) || !b.alloc(L.dip, 3, err) || !b.alloc(L.info, 1, err) ||
      !b.alloc(L.fin, LG_MAXDOT, err) || !b.alloc(L.lz, 3LL * np, err) || !b.alloc(L.fs, 3LL * np, err) ||
      !b.alloc(L.bp, (size_t)LG_SPLIT_MAX * 3 * (LG_SPLIT_MAX * (LG_SPLIT_MAX + 1) / 2), err) ||
      !b.alloc(L.sp2s, SP2S_N, err) || !b.alloc(L.lzab, 2LL * (2 * np + 16), err))
    return false;
  {
    std::vector<double> r(np);
    for (int j = 0; j < np; ++j) r[j] = seed_r(j);
    L.fr = L.fs + 2 * np;
    HIP_OK(scopy(b.stream, L.fr, r.data(), sizeof(double) * np, hipMemcpyHostToDevice));
  }
  b.h_red.resize((size_t)LG_BLOCKS * LG_MAXDOT);
  BLAS_OK(rocblas_create_handle(&b.blas));
  BLAS_OK(rocblas_set_stream(b.blas, b.stream));
  return true;
}

// The next (start, end) pair of an event pool, created on first use.
static hipError_t pool_pair(std::vector<hipEvent_t>& pool, int& used, hipEvent_t& e0, hipEvent_t& e1) {
  while ((int)pool.size() < 2 * (used + 1)) {
    hipEvent_t e = nullptr;
    const hipError_t r = hipEventCreate(&e);
    if (r != hipSuccess) return r;
    pool.push_back(e);
  }
  e0 = pool[2 * (size_t)used];
  e1 = pool[2 * (size_t)used + 1];
  ++used;
  return hipSuccess;
}

// Summed device time (ms) of the first `used` pairs of a pool; the events must have completed.
static bool pool_ms(const std::vector<hipEvent_t>& pool, int used, double& ms, std::string& err) {
  ms = 0.0;
  for (int i = 0; i < used; ++i) {
    float t = 0.0f;
    HIP_OK(hipEventElapsedTime(&t, pool[2 * (size_t)i], pool[2 * (size_t)i + 1]));
    ms += t;
  }
  return true;
}

// Sum of the first k columns of the per-block partials, blocks in order.
static bool large_reduce(Batch& b, int k, double* out, std::string& err) {
  HIP_OK(scopy(b.stream, b.h_red.data(), b.lg.red, sizeof(double) * (size_t)LG_BLOCKS * LG_MAXDOT, hipMemcpyDeviceToHost));
  for (int q = 0; q < k; ++q) out[q] = 0.0;
  for (int blk = 0; blk < LG_BLOCKS; ++blk)
    for (int q = 0; q < k; ++q) out[q] += b.h_red[(size_t)blk * LG_MAXDOT + q];
  return true;
}

static bool large_dots(Batch& b, const Dots& D, long long m, double* out, std::string& err) {
  hipLaunchKernelGGL(lg_dots_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, D, m, b.lg.red);
  HIP_OK(hipGetLastError());
  return large_reduce(b, D.k, out, err);
}

// Block partials written as red[block] (one column): sum them.
static bool large_sum1(Batch& b, double& out, std::string& err) {
  std::vector<double> h(LG_BLOCKS);
  HIP_OK(scopy(b.stream, h.data(), b.lg.red, sizeof(double) * LG_BLOCKS, hipMemcpyDeviceToHost));
  out = 0.0;
  for (double v : h) out += v;
  return true;
}

static bool large_fock(Batch& b, const MolDesc& d, const double* P, double* F, std::string& err) {
  LargeBufs& L = b.lg;
  hipEvent_t f0, f1;
  HIP_OK(pool_pair(L.fev, L.nfev, f0, f1));
  HIP_OK(hipEventRecord(f0, b.stream));
  hipLaunchKernelGGL(lg_fock_pairs_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, b.args, d, P, F, b.lg.slot);
  hipLaunchKernelGGL(lg_fock_jsum_wide_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, b.args, d, P, b.lg.J,
                     b.lg.slot);
  hipLaunchKernelGGL(lg_fock_diag_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, b.args, d, P, F,
                     (const double*)b.lg.J);
  HIP_OK(hipGetLastError());
  HIP_OK(hipEventRecord(f1, b.stream));
  ++L.fock_builds;
  return true;
}

static bool large_energy(Batch& b, const MolDesc& d, const double* P, const double* F, double& e, std::string& err) {
  const Mol M = mol_of(b.args, d, nullptr);
  Dots D{};
  D.k = 2;
  D.x[0] = P; D.y[0] = M.H;
  D.x[1] = P; D.y[1] = F;
  double v[2];
  if (!large_dots(b, D, (long long)d.np * d.np, v, err)) return false;
  e = 0.5 * (v[0] + v[1]);
  return true;
}

// Eigenpairs of the symmetric matrix A (np leading dimension) into L.V (column-major, ascending
// eigenvalues in L.w), then P = 2 C_occ C_occ^T into Pdst.
static bool large_eigen_density(Batch& b, const MolDesc& d, const double* A, double* Pdst, std::string& err) {
  LargeBufs& L = b.lg;
  const int n = d.nao, np = d.np;
  HIP_OK(hipMemcpyAsync(L.V, A, sizeof(double) * (size_t)np * np, hipMemcpyDeviceToDevice, b.stream));
  BLAS_OK(rocsolver_dsyevd(b.blas, rocblas_evect_original, rocblas_fill_upper, n, L.V, np, L.w, L.ework, L.info));
  int info = 0;
  HIP_OK(scopy(b.stream, &info, L.info, sizeof(int), hipMemcpyDeviceToHost));
  if (info != 0) {
    err = "rocsolver_dsyevd did not converge (info " + std::to_string(info) + ")";
    return false;
  }
  const double two = 2.0, zero = 0.0;
  BLAS_OK(rocblas_dgemm(b.blas, rocblas_operation_none, rocblas_operation_transpose, n, n, d.nocc, &two, L.V, np, L.V,
                        np, &zero, Pdst, np));
  return true;
}

// Largest eigenvalue of the symmetric tridiagonal (a, b) of size m, by bisection on the Sturm count.
static double host_tridiag_max(const std::vector<double>& a, const std::vector<double>& b, int m) {
  double lo = 1.0e300, hi = -1.0e300;
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  for (int it = 0; it < 200 && hi - lo > 1.0e-15 * std::max(std::fabs(lo), std::fabs(hi)); ++it) {
    const double x = 0.5 * (lo + hi);
    int below = 0;
    double d = 1.0;
    for (int i = 0; i < m; ++i) {
      d = a[i] - x - (i > 0 ? b[i - 1] * b[i - 1] / d : 0.0);
      if (d == 0.0) d = -1.0e-300;
      below += d < 0.0;
    }
    if (below == m) hi = x; else lo = x;
  }
  return 0.5 * (lo + hi);
}

// Ritz estimates of the extreme eigenvalues of the symmetric np-strided A: SP2_LZ_STEPS plain
// Lanczos steps with rocBLAS level-1/2 calls (start vector as lanczos_start).
static bool large_extremes(Batch& b, const MolDesc& d, const double* A, double& lo_est, double& hi_est, std::string& err) {
  LargeBufs& L = b.lg;
  const int n = d.nao, np = d.np;
  double* q = L.lz;
  double* qp = L.lz + np;
  double* w = L.lz + 2 * np;
  hipLaunchKernelGGL(lg_extremes_start_kernel, dim3(1), dim3(LZ_VEC_THREADS), 0, b.stream, q, qp, n);
  const double one = 1.0, zero = 0.0;
  const int steps = std::min(n, SP2_LZ_STEPS);
  // all steps enqueued at once (alpha, beta stay on the device), read back together
  double* dal = L.lzab;
  double* dbe = L.lzab + (2 * np + 16);
  for (int k = 0; k < steps; ++k) {
    if (n <= LARGE_LZ_FUSED_N) {
      hipLaunchKernelGGL(lg_lanczos_fused_kernel, dim3(1), dim3(LZ_VEC_THREADS), 0, b.stream, A, np, q, qp, w, n, dal,
                         dbe, k);
      continue;
    }
    // A symmetric: A^T q, whose kernel reads the column-major columns along rows of A (coalesced)
    BLAS_OK(rocblas_dgemv(b.blas, rocblas_operation_transpose, n, n, &one, A, np, q, 1, &zero, w, 1));
    hipLaunchKernelGGL(lg_lanczos_vec_kernel, dim3(1), dim3(LZ_VEC_THREADS), 0, b.stream, q, qp, w, n, dal, dbe, k);
  }
  HIP_OK(hipGetLastError());
  std::vector<double> ha(steps), hb(steps);
  HIP_OK(hipMemcpyAsync(ha.data(), dal, sizeof(double) * steps, hipMemcpyDeviceToHost, b.stream));
  HIP_OK(hipMemcpyAsync(hb.data(), dbe, sizeof(double) * steps, hipMemcpyDeviceToHost, b.stream));
  HIP_OK(hipStreamSynchronize(b.stream));
  std::vector<double> al, be, nal;
  for (int k = 0; k < steps; ++k) {  // up to and including a breakdown, as the one-step loop stopped
    al.push_back(ha[k]);
    nal.push_back(-ha[k]);
    be.push_back(hb[k]);
    if (hb[k] <= 1.0e-13 * (std::fabs(ha[k]) + hb[k])) break;
  }
  const int m = (int)al.size();
  hi_est = host_tridiag_max(al, be, m);
  lo_est = -host_tridiag_max(nal, be, m);
  return true;
}

// Aufbau density of the symmetric A into Pdst by purification (the steps, stopping rule and
// checks of sp2_density), X in L.V and the product in L.T.  Bounds from Gershgorin, or from a
// cached Lanczos estimate as in sp2_density (refreshed at the first two steps of a path, every
// SP2_REFRESH steps and after a miss).  False in `ok` (nothing written) when the frontier gap is
// closed or a check fails with Gershgorin bounds: the caller then diagonalises.
static bool large_purify_run(Batch& b, const MolDesc& d, const double* A, double* Pdst, double lo, double hi, bool seed,
                             bool& ok, std::string& err);
static bool large_purify(Batch& b, const MolDesc& d, const double* A, double* Pdst, bool seed, bool& ok,
                         std::string& err) {
  LargeBufs& L = b.lg;
  const int n = d.nao, np = d.np, nocc = d.nocc;
  ok = false;
  if (nocc <= 0 || nocc >= n) return true;
  hipLaunchKernelGGL(lg_gershgorin_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, A, n, np, L.red);
  HIP_OK(hipGetLastError());
  HIP_OK(scopy(b.stream, b.h_red.data(), L.red, sizeof(double) * (size_t)LG_BLOCKS * LG_MAXDOT, hipMemcpyDeviceToHost));
  double glo = 1.0e300, ghi = -1.0e300;
  for (int blk = 0; blk < LG_BLOCKS; ++blk) {
    glo = std::min(glo, b.h_red[(size_t)blk * LG_MAXDOT]);
    ghi = std::max(ghi, b.h_red[(size_t)blk * LG_MAXDOT + 1]);
  }
  if (!(ghi - glo > 1.0e-12 * std::max(1.0, std::fabs(ghi)))) return true;
  if (n >= b.args.sp2_tight_n && !L.no_tight) {
    const int age = L.bnd_age;
    // (a few Lanczos steps see less of a large spectrum: a wider margin here)
    const double m0 = b.args.sp2_margin < 0.0 ? b.args.sp2_margin : std::max(b.args.sp2_margin, LARGE_SP2_MARGIN);
    double mf = m0 + SP2_MARGIN_AGED;
    if (age < 2 || age % SP2_REFRESH == 0) {
      if (!large_extremes(b, d, A, L.bnd_lo, L.bnd_hi, err)) return false;
      mf = m0;
    }
    L.bnd_age = age < 0 ? 1 : age + 1;
    const double m = mf * (L.bnd_hi - L.bnd_lo);
    const double lo = std::max(glo, L.bnd_lo - m), hi = std::min(ghi, L.bnd_hi + m);
    if (hi - lo < ghi - glo && hi - lo > 1.0e-12 * std::max(1.0, std::fabs(hi))) {
      if (!large_purify_run(b, d, A, Pdst, lo, hi, seed, ok, err)) return false;
      if (ok) return true;
      L.bnd_age = 0;  // a miss: estimate afresh next time
    }
  }
  return large_purify_run(b, d, A, Pdst, glo, ghi, seed, ok, err);
}

// The tile list of lg_sp2_sym_kernel for n AOs: the lower-triangle 64 x 64 tiles (I << 16 | J),
// grouped by 8 x 8 super-blocks taken row by row, each super-block's tiles row by row.
static std::vector<int> sy_map_host(int n) {
  const int nt = (n + SY_T - 1) / SY_T, ns = (nt + 7) / 8;
  std::vector<int> m;
  m.reserve(sy_tiles(n));
  for (int P = 0; P < ns; ++P)
    for (int Q = 0; Q <= P; ++Q)
      for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
          const int I = 8 * P + i, J = 8 * Q + j;
          if (I < nt && J <= I) m.push_back(I << 16 | J);
        }
  return m;
}

// The tile kernel's launch for one step from X into Xn (slab width k: 16 or 32 columns).
static void sy_launch(int k, int grid, hipStream_t s, const double* X, double* Xn, int n, int np, const double* st,
                      double* red, const int* map, int chunk) {
  if (k == 8)
    hipLaunchKernelGGL(lg_sp2_sym_kernel<8>, dim3(grid), dim3(SY_THREADS), 0, s, X, Xn, n, np, st, red, map, chunk);
  else if (k == 32)
    hipLaunchKernelGGL(lg_sp2_sym_kernel<32>, dim3(grid), dim3(SY_THREADS), 0, s, X, Xn, n, np, st, red, map, chunk);
  else
    hipLaunchKernelGGL(lg_sp2_sym_kernel<16>, dim3(grid), dim3(SY_THREADS), 0, s, X, Xn, n, np, st, red, map, chunk);
}

static bool sy_tile_map(Batch& b, int n, std::string& err) {
  LargeBufs& L = b.lg;
  const std::vector<int> m = sy_map_host(n);
  if ((int)m.size() > L.sy_map_cap) {
    if (!b.alloc(L.sy_map, (long long)m.size(), err)) return false;
    L.sy_map_cap = (int)m.size();
  }
  HIP_OK(scopy(b.stream, L.sy_map, m.data(), sizeof(int) * m.size(), hipMemcpyHostToDevice));
  L.sy_map_n = n;
  return true;
}

static bool large_purify_run(Batch& b, const MolDesc& d, const double* A, double* Pdst, double lo, double hi, bool seed,
                             bool& ok, std::string& err) {
  LargeBufs& L = b.lg;
  const int n = d.nao, np = d.np, nocc = d.nocc;
  ok = false;
  // q x q block-lower products: X X restricted to the blocks (I, J), I >= J, of size bs = n / q, as
  // q rounds of one batched dgemm each (round K: Y_IJ += X_IK X_KJ); the update reads only the lower
  // triangle, which these blocks cover.  (q + 1) / (2 q) of the full product's flops.
  const bool fused = n <= b.large_steps_n;  // one kernel per step, X ping-ponging between L.V and L.T
  // the lower-triangle tile kernel (lg_sp2_sym_kernel): also one kernel per step, ping-ponging
  const bool own = !fused && b.large_split == 1;
  const bool pp = fused || own;
  const int q = (!pp && !b.large_syrk && b.large_split >= 2 && b.large_split <= LG_SPLIT_MAX &&
                 n % b.large_split == 0 && n / b.large_split >= 64) ? b.large_split : 0;
  const int bs = q ? n / q : 0, nb = q * (q + 1) / 2;
  if (q) {
    std::vector<double*> h((size_t)q * 3 * nb);
    for (int K = 0; K < q; ++K) {
      int e = 0;
      for (int I = 0; I < q; ++I)
        for (int J = 0; J <= I; ++J, ++e) {
          // row-major Y_IJ = X_IK X_KJ through column-major dgemm: A = X_KJ, B = X_IK
          h[((size_t)K * 3 + 0) * nb + e] = L.V + (size_t)K * bs * np + (size_t)J * bs;
          h[((size_t)K * 3 + 1) * nb + e] = L.V + (size_t)I * bs * np + (size_t)K * bs;
          h[((size_t)K * 3 + 2) * nb + e] = L.T + (size_t)I * bs * np + (size_t)J * bs;
        }
    }
    HIP_OK(scopy(b.stream, L.bp, h.data(), sizeof(double*) * h.size(), hipMemcpyHostToDevice));
  }
  // The loop of sp2_density with its decisions on the device (lg_sp2_state_kernel): the host
  // enqueues `per` steps at a time -- product, update, state -- and then reads the state.  Steps
  // enqueued after the loop has stopped leave X alone; only their products run (the fused and
  // tile kernels return at once).  Small matrices take several steps per look (their products
  // cost less than the host's round trip); the tile kernel's first look comes two steps before
  // the step count of the last purification, then one step per look; the rest look every step.
  const int per0 = fused ? LARGE_STEPS_PER_LOOK : own ? std::max(1, L.sp2_steps - 2) : 1;
  // one workgroup per tile (the dispatcher balances them over the CUs), two partials each in L.red;
  // above about 10 000 AOs the grid is capped by L.red and workgroups take several tiles
  const int schunk = own ? (sy_tiles(n) + 7) / 8 : 0;
  const int sgrid = std::min(8 * schunk, (LG_BLOCKS * LG_MAXDOT / 2) & ~7);
  if (own && L.sy_map_n != n) {
    if (!sy_tile_map(b, n, err)) return false;
  }
  const double occ = (double)nocc, one = 1.0, zero = 0.0;
  hipLaunchKernelGGL(lg_sp2_init_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, A, L.V, n, np, lo, hi, L.red);
  hipLaunchKernelGGL(lg_sp2_state_kernel, dim3(1), dim3(LG_BLOCKS), 0, b.stream, (const double*)L.red, L.sp2s, occ, 1,
                     seed ? 1 : 0, LG_BLOCKS, LG_MAXDOT);
  HIP_OK(hipGetLastError());
  double st[SP2S_N] = {0.0};  // read after each group of steps (a state finished at X0 makes them no-ops)
  const int gs = std::max(1, std::min(LG_BLOCKS, (n + 3) / 4));
  hipEvent_t f0 = nullptr, f1 = nullptr;  // the fused path times its whole loop (products and updates)
  const int fbm = (n + 15) / 16, fgrid = std::min(LG_BLOCKS, fbm * (fbm + 1) / 2);
  if (fused) {
    HIP_OK(pool_pair(L.pev, L.npev, f0, f1));
    HIP_OK(hipEventRecord(f0, b.stream));
  }
  for (int enq = 0; st[SP2S_STATUS] == 0.0 && enq < SP2_MAX;) {
    const int per = enq == 0 || !own ? per0 : 1;
    for (int s = 0; s < per && enq < SP2_MAX; ++s, ++enq) {
      // the iterate of this step: the fused and tile paths alternate buffers (if the loop has
      // already stopped, the step does nothing and this pointer is not read)
      const double* Xc = pp && (enq & 1) ? L.T : L.V;
      if (seed) {  // s = (X - X^2) r when the tail has just started (frontier_seed)
        double* sv = L.fs;
        double* t = L.fs + np;
        double* u = L.lz;
        hipLaunchKernelGGL(lg_seed_gemv_kernel, dim3(gs), dim3(LG_THREADS), 0, b.stream, Xc,
                           (const double*)L.fr, t, n, np, (const double*)L.sp2s);
        hipLaunchKernelGGL(lg_seed_gemv_kernel, dim3(gs), dim3(LG_THREADS), 0, b.stream, Xc,
                           (const double*)t, u, n, np, (const double*)L.sp2s);
        hipLaunchKernelGGL(lg_seed_diff_kernel, dim3(std::max(1, std::min(LG_BLOCKS, (n + LG_THREADS - 1) / LG_THREADS))),
                           dim3(LG_THREADS), 0, b.stream, (const double*)t, (const double*)u, sv, n,
                           (const double*)L.sp2s);
      }
      if (fused) {
        hipLaunchKernelGGL(lg_sp2_fused_kernel, dim3(fgrid), dim3(64), 0, b.stream, Xc, (enq & 1) ? L.V : L.T, n, np,
                           (const double*)L.sp2s, L.red);
        hipLaunchKernelGGL(lg_sp2_state_kernel, dim3(1), dim3(LG_BLOCKS), 0, b.stream, (const double*)L.red, L.sp2s,
                           occ, 0, seed ? 1 : 0, fgrid, LG_MAXDOT);
        HIP_OK(hipGetLastError());
        ++L.products;
        const long long bm = (n + 15) / 16;
        L.product_flop += bm * (bm + 1) / 2 * 2LL * 256 * ((n + 63) / 64 * 64);
        continue;
      }
      if (own) {
        hipEvent_t p0, p1;
        HIP_OK(pool_pair(L.pev, L.npev, p0, p1));
        HIP_OK(hipEventRecord(p0, b.stream));
        sy_launch(b.sy_k, sgrid, b.stream, Xc, (enq & 1) ? L.V : L.T, n, np, (const double*)L.sp2s, L.red,
                  (const int*)L.sy_map, schunk);
        HIP_OK(hipEventRecord(p1, b.stream));
        hipLaunchKernelGGL(lg_sp2_state_kernel, dim3(1), dim3(LG_BLOCKS), 0, b.stream, (const double*)L.red, L.sp2s,
                           occ, 0, seed ? 1 : 0, sgrid, 2);
        HIP_OK(hipGetLastError());
        ++L.products;
        L.product_flop += 2LL * sy_tiles(n) * SY_T * SY_T * ((n + b.sy_k - 1) / b.sy_k * b.sy_k);
        continue;
      }
      hipEvent_t p0, p1;
      HIP_OK(pool_pair(L.pev, L.npev, p0, p1));
      HIP_OK(hipEventRecord(p0, b.stream));
      // the block-lower products, or the whole square by dgemm, or one triangle by dsyrk (upper in
      // column-major order = the row-major lower triangle): the update reads only the row-major
      // lower triangle, so X stays exactly symmetric
      if (q) {
        for (int K = 0; K < q; ++K) {
          const double beta = K == 0 ? 0.0 : 1.0;
          BLAS_OK(rocblas_dgemm_batched(b.blas, rocblas_operation_none, rocblas_operation_none, bs, bs, bs, &one,
                                        (const double* const*)(L.bp + ((size_t)K * 3 + 0) * nb), np,
                                        (const double* const*)(L.bp + ((size_t)K * 3 + 1) * nb), np, &beta,
                                        L.bp + ((size_t)K * 3 + 2) * nb, np, nb));
        }
      } else if (b.large_syrk)
        BLAS_OK(rocblas_dsyrk(b.blas, rocblas_fill_upper, rocblas_operation_none, n, n, &one, L.V, np, &zero, L.T, np));
      else
        BLAS_OK(rocblas_dgemm(b.blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, L.V, np, L.V, np,
                              &zero, L.T, np));
      HIP_OK(hipEventRecord(p1, b.stream));
      hipLaunchKernelGGL(lg_sp2_step_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, (const double*)L.T, L.V, n,
                         np, (const double*)L.sp2s, L.red);
      hipLaunchKernelGGL(lg_sp2_state_kernel, dim3(1), dim3(LG_BLOCKS), 0, b.stream, (const double*)L.red, L.sp2s, occ,
                         0, seed ? 1 : 0, LG_BLOCKS, LG_MAXDOT);
      HIP_OK(hipGetLastError());
      ++L.products;
      L.product_flop += q ? 2LL * bs * (long long)bs * bs * nb * q : (b.large_syrk ? 1LL : 2LL) * n * (long long)n * n;
    }
    HIP_OK(scopy(b.stream, st, L.sp2s, sizeof(st), hipMemcpyDeviceToHost));
  }
  if (fused) HIP_OK(hipEventRecord(f1, b.stream));
  if (st[SP2S_SEEDED] != 0.0) L.seeded = true;
  if (st[SP2S_STATUS] == 1.0) L.sp2_steps = (int)st[SP2S_K];
  if (st[SP2S_STATUS] != 1.0) return true;  // a bound violated, or no convergence: ok stays false
  const double* Xf = pp && ((int)st[SP2S_K] & 1) ? L.T : L.V;  // after K steps
  hipLaunchKernelGGL(lg_twice_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, Xf, Pdst, n, np);
  HIP_OK(hipGetLastError());
  ok = true;
  return true;
}

// Eigenvalues only of the symmetric A (ascending, into L.w), for the frontier levels of a
// purified path.
static bool large_levels(Batch& b, const MolDesc& d, const double* A, std::string& err) {
  LargeBufs& L = b.lg;
  const int n = d.nao, np = d.np;
  HIP_OK(hipMemcpyAsync(L.V, A, sizeof(double) * (size_t)np * np, hipMemcpyDeviceToDevice, b.stream));
  BLAS_OK(rocsolver_dsyevd(b.blas, rocblas_evect_none, rocblas_fill_upper, n, L.V, np, L.w, L.ework, L.info));
  int info = 0;
  HIP_OK(scopy(b.stream, &info, L.info, sizeof(int), hipMemcpyDeviceToHost));
  if (info != 0) {
    err = "rocsolver_dsyevd did not converge (info " + std::to_string(info) + ")";
    return false;
  }
  return true;
}

// The aufbau density of A into Pdst: purified when the batch returns no orbitals, else (or when
// purification declines) from the eigenvectors.
static bool large_density(Batch& b, const MolDesc& d, const double* A, double* Pdst, std::string& err,
                          bool seed = false) {
  b.lg.last_eig = false;
  if (b.purify) {
    bool ok = false;
    if (!large_purify(b, d, A, Pdst, seed, ok, err)) return false;
    if (ok) return true;
  }
  b.lg.last_eig = true;
  return large_eigen_density(b, d, A, Pdst, err);
}

// Largest eigenvalue of the symmetric np-strided A by plain Lanczos from the normalised device
// vector st (rocBLAS level-1/2 calls on L.lz), with frontier_levels' stopping tests: a Ritz
// residual below LZ_TOL, a value stalled over two checks, breakdown or the step cap; checks from
// step 8 every fourth for a seeded start, else from 16 every eighth.
static bool large_lanczos_top(Batch& b, const MolDesc& d, const double* A, const double* st, bool seeded, double scale,
                              double& top, std::string& err) {
  LargeBufs& L = b.lg;
  const int n = d.nao, np = d.np;
  double* q = L.lz;
  double* qp = L.lz + np;
  double* w = L.lz + 2 * np;
  HIP_OK(hipMemcpyAsync(q, st, sizeof(double) * n, hipMemcpyDeviceToDevice, b.stream));
  HIP_OK(hipMemsetAsync(qp, 0, sizeof(double) * n, b.stream));
  double th1 = -1.0e300, th2 = -1.0e300;
  const double one = 1.0, zero = 0.0;
  const int cap = 2 * n + 16, first = seeded ? 8 : 16, every = seeded ? 4 : 8;
  top = 0.0;
  // steps are enqueued up to the next check (alpha and beta stay on the device) and then read
  // back; the checks see exactly the coefficients of the one-step loop, so they stop at the same
  // step (steps enqueued past a breakdown are not used)
  double* dal = L.lzab;
  double* dbe = L.lzab + (2 * np + 16);
  std::vector<double> al(cap), be(cap);
  int done = 0, seen = 0;
  while (done < cap) {
    const int next = std::min(cap, done < first ? first : done + every);
    for (; done < next; ++done) {
      if (n <= LARGE_LZ_FUSED_N) {
        hipLaunchKernelGGL(lg_lanczos_fused_kernel, dim3(1), dim3(LZ_VEC_THREADS), 0, b.stream, A, np, q, qp, w, n, dal,
                           dbe, done);
        continue;
      }
      BLAS_OK(rocblas_dgemv(b.blas, rocblas_operation_transpose, n, n, &one, A, np, q, 1, &zero, w, 1));
      hipLaunchKernelGGL(lg_lanczos_vec_kernel, dim3(1), dim3(LZ_VEC_THREADS), 0, b.stream, q, qp, w, n, dal, dbe, done);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(al.data() + seen, dal + seen, sizeof(double) * (done - seen), hipMemcpyDeviceToHost, b.stream));
    HIP_OK(hipMemcpyAsync(be.data() + seen, dbe + seen, sizeof(double) * (done - seen), hipMemcpyDeviceToHost, b.stream));
    HIP_OK(hipStreamSynchronize(b.stream));
    for (int mk = seen + 1; mk <= done; ++mk) {
      const bool breakdown = be[mk - 1] <= 1.0e-13 * scale;
      if (mk == cap || breakdown || (mk >= first && mk % every == 0)) {
        const double th = host_tridiag_max(al, be, mk);
        const bool stall = th - th1 <= 1.0e-13 * scale && th - th2 <= 1.0e-12 * scale;
        top = th;
        if (mk == cap || breakdown || stall || ritz_residual(al.data(), be.data(), mk, th) <= LZ_TOL) {
          L.lanczos_steps += mk;
          ++L.lanczos_runs;
          return true;
        }
        th2 = th1;
        th1 = th;
      }
    }
    seen = done;
  }
  return true;
}

// v <- v / |v| + 1e-3 r / |r| with the generic vector r of L.fr (so that no direction is missing),
// or r / |r| when v is not to be used, is zero or is not finite (lanczos_start).
static bool large_start(Batch& b, const MolDesc& d, double* v, bool use, std::string& err) {
  LargeBufs& L = b.lg;
  const int n = d.nao;
  const double* r = L.fr;
  double rn = 0.0, vn = 0.0;
  BLAS_OK(rocblas_dnrm2(b.blas, n, r, 1, &rn));
  if (use) BLAS_OK(rocblas_dnrm2(b.blas, n, v, 1, &vn));
  if (use && vn > 1.0e-140 && vn < 1.0e140) {
    const double iv = 1.0 / vn, cr = 1.0e-3 / rn;
    BLAS_OK(rocblas_dscal(b.blas, n, &iv, v, 1));
    BLAS_OK(rocblas_daxpy(b.blas, n, &cr, r, 1, v, 1));
  } else {
    const double ir = 1.0 / rn;
    BLAS_OK(rocblas_dcopy(b.blas, n, r, 1, v, 1));
    BLAS_OK(rocblas_dscal(b.blas, n, &ir, v, 1));
  }
  double nn = 0.0;
  BLAS_OK(rocblas_dnrm2(b.blas, n, v, 1, &nn));
  const double in = 1.0 / nn;
  BLAS_OK(rocblas_dscal(b.blas, n, &in, v, 1));
  return true;
}

// Frontier levels of the last step without eigenvalues: the HOMO as the largest eigenvalue of
// A_o = X F X + c (I - X) and the LUMO as minus the largest of A_v = -((I - X) F (I - X) + c' X),
// X = P / 2, c and c' beyond the Gershgorin interval of F (frontier_levels); A_o and A_v from two
// dgemm products and a grid-wide kernel, then one Lanczos run each, seeded from the path's last
// purification when it left a seed.  HOMO and LUMO go to L.w[nocc - 1], L.w[nocc]; `aufbau` is
// false when HOMO > LUMO (an eigenvalue on the wrong side of the projector).
static bool large_levels_lanczos(Batch& b, const MolDesc& d, const double* Fu, bool& aufbau, std::string& err) {
  LargeBufs& L = b.lg;
  const int n = d.nao, np = d.np, nocc = d.nocc;
  aufbau = true;
  hipLaunchKernelGGL(lg_gershgorin_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, Fu, n, np, L.red);
  HIP_OK(hipGetLastError());
  HIP_OK(scopy(b.stream, b.h_red.data(), L.red, sizeof(double) * (size_t)LG_BLOCKS * LG_MAXDOT, hipMemcpyDeviceToHost));
  double glo = 1.0e300, ghi = -1.0e300;
  for (int blk = 0; blk < LG_BLOCKS; ++blk) {
    glo = std::min(glo, b.h_red[(size_t)blk * LG_MAXDOT]);
    ghi = std::max(ghi, b.h_red[(size_t)blk * LG_MAXDOT + 1]);
  }
  const double c_lo = glo - 1.0 - 0.1 * (ghi - glo), c_hi = ghi + 1.0 + 0.1 * (ghi - glo);
  const double scale = std::max(std::fabs(c_lo), std::fabs(c_hi));
  // row-major products through column-major dgemm: dgemm(A = a, B = c) leaves c a (row-major)
  const double half = 0.5, zero = 0.0;
  BLAS_OK(rocblas_dgemm(b.blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &half, Fu, np, L.P, np, &zero,
                        L.T, np));  // G = X F
  BLAS_OK(rocblas_dgemm(b.blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &half, L.P, np, L.T, np, &zero,
                        L.V, np));  // X F X
  double* Ao = L.Pn;
  double* Av = L.Eh;
  hipLaunchKernelGGL(lg_frontier_ops_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, Fu, (const double*)L.P,
                     (const double*)L.T, (const double*)L.V, n, np, c_lo, c_hi, Ao, Av);
  HIP_OK(hipGetLastError());
  // starts: X s for the HOMO, s - X s for the LUMO (s the seed), or the generic vector
  double* sl = L.fs;       // s, then the LUMO start
  double* sh = L.fs + np;  // X s, then the HOMO start
  const bool seeded = L.seeded;
  if (seeded) {
    const double mone = -1.0;
    BLAS_OK(rocblas_dgemv(b.blas, rocblas_operation_transpose, n, n, &half, L.P, np, sl, 1, &zero, sh, 1));
    BLAS_OK(rocblas_daxpy(b.blas, n, &mone, sh, 1, sl, 1));
  }
  if (!large_start(b, d, sh, seeded, err) || !large_start(b, d, sl, seeded, err)) return false;
  double homo = 0.0, lumo = 0.0;
  if (!large_lanczos_top(b, d, Ao, sh, seeded, scale, homo, err)) return false;
  if (!large_lanczos_top(b, d, Av, sl, seeded, scale, lumo, err)) return false;
  lumo = -lumo;
  aufbau = homo <= lumo;
  const double lv[2] = {homo, lumo};
  HIP_OK(scopy(b.stream, L.w + nocc - 1, lv, sizeof(lv), hipMemcpyHostToDevice));
  return true;
}

struct LargePath {
  double e_elec = 0.0, last_dp = 1.0e300;
  int iters = 0, converged = 0;
  bool aufbau = true;  // purified path: the final density is the aufbau projector of its Fock matrix
};

// One path of the host-driven SCF; shift, damp and urung as in the batched kernel's rhf_loop (the
// unrestricted iteration of a closed shell, carried on the total density).
static bool large_rhf(Batch& b, const MolDesc& d, int guess, int accel, LargePath& R, std::string& err,
                      double shift = 0.0, double damp = 0.0, bool urung = false) {
  LargeBufs& L = b.lg;
  const int n = d.nao, np = d.np;
  const long long ms = (long long)np * np;
  const Mol M = mol_of(b.args, d, nullptr);
  const EngineOptions& o = b.opt;
  HIP_OK(hipMemsetAsync(L.P, 0, sizeof(double) * ms, b.stream));
  HIP_OK(hipMemsetAsync(L.Pn, 0, sizeof(double) * ms, b.stream));
  L.bnd_age = -1;  // no spectral-bound estimate for this path yet
  L.seeded = false;  // nor a frontier seed
  if (guess == 2) {
    if (!large_density(b, d, M.H, L.P, err)) return false;
  } else {
    int tilt = -1;
    if (guess == 3) {
      double best = 0.0;
      for (int a = 0; a < d.nat; ++a) {
        int slot = 0;
        HIP_OK(scopy(b.stream, &slot, b.args.aslot + d.atom0 + a, sizeof(int), hipMemcpyDeviceToHost));
        const ElemConst& e = b.table.el[slot];
        if (e.norb == 4 && e.iop > best) { best = e.iop; tilt = a; }
      }
    }
    hipLaunchKernelGGL(lg_sad_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, b.args, d, L.P, tilt);
    HIP_OK(hipGetLastError());
  }
  if (urung && accel == 2) accel = 1;
  int keep_d = accel == 2;
  if (keep_d && !(o.scf_memory_mb == 0 || 2LL * 3 * n * n * 8 <= (long long)o.scf_memory_mb * 1024 * 1024)) keep_d = 0;
  if (accel == 2 && !keep_d) accel = 1;
  Diis D{};
  D.len = 0;
  D.depth = std::min(urung ? diis_depth_mats(n, 4, o.scf_memory_mb) : diis_depth_for(n, keep_d, o.scf_memory_mb),
                     L.depth);
  D.keep_d = keep_d;
  R = LargePath();
  const double* Flast = nullptr;  // the matrix of the last aufbau step
  double e_old = 0.0;
  double* P = L.P;
  double* Pn = L.Pn;
  for (int it = 0; it < o.max_scf; ++it) {
    R.iters = it + 1;
    if (!large_fock(b, d, P, L.F, err)) return false;
    double e_elec = 0.0;
    if (!large_energy(b, d, P, L.F, e_elec, err)) return false;
    const double* Fuse = L.F;
    if (accel != 0) {
      if (D.len == D.depth) {
        const int s0 = D.slot[0];
        for (int i = 0; i + 1 < D.len; ++i) {
          D.slot[i] = D.slot[i + 1];
          for (int j = 0; j + 1 < D.len; ++j) {
            D.b[i][j] = D.b[i + 1][j + 1];
            D.mm[i][j] = D.mm[i + 1][j + 1];
          }
        }
        D.len -= 1;
        D.newslot = s0;
      } else {
        D.newslot = D.len;
      }
      const int ns = D.newslot;
      double* Fn = L.Fh + ns * ms;
      double* En = L.Eh + ns * ms;
      double* Dn = L.Dh + ns * ms;
      const double one = 1.0, zero = 0.0;
      // T = F P (row-major; rocBLAS is column-major and both factors are symmetric)
      BLAS_OK(rocblas_dgemm(b.blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, P, np, L.F, np, &zero,
                            L.T, np));
      hipLaunchKernelGGL(lg_commutator_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, (const double*)L.T,
                         (const double*)L.F, (const double*)P, En, Fn, Dn, n, np, keep_d);
      HIP_OK(hipGetLastError());
      const int len = D.len;
      Dots G{};
      int k = 0;
      int ie[MAXD], idf[MAXD], ifd[MAXD];
      for (int i = 0; i < len; ++i) {
        const long long so = (long long)D.slot[i] * ms;
        ie[i] = k; G.x[k] = L.Eh + so; G.y[k] = En; ++k;
        if (keep_d) {
          idf[i] = k; G.x[k] = Dn; G.y[k] = L.Fh + so; ++k;
          ifd[i] = k; G.x[k] = L.Dh + so; G.y[k] = Fn; ++k;
        }
      }
      const int iee = k;
      G.x[k] = En; G.y[k] = En; ++k;
      const int idd = k;
      if (keep_d) { G.x[k] = Dn; G.y[k] = Fn; ++k; }
      G.k = k;
      double v[LG_MAXDOT];
      if (!large_dots(b, G, ms, v, err)) return false;
      for (int i = 0; i < len; ++i) {
        D.b[i][len] = v[ie[i]];
        D.b[len][i] = v[ie[i]];
        D.mm[len][i] = keep_d ? v[idf[i]] : 0.0;
        D.mm[i][len] = keep_d ? v[ifd[i]] : 0.0;
      }
      D.b[len][len] = v[iee];
      D.mm[len][len] = keep_d ? v[idd] : 0.0;
      D.slot[len] = ns;
      D.len = len + 1;
      const double en = std::sqrt(v[iee]);
      bool ok;
      if (accel == 2 && en > o.adiis_switch) {
        ok = D.len >= 2 && D.keep_d;
        if (ok) adiis_coeffs(D, D.coef);
      } else {
        ok = cdiis_coeffs(D, D.coef);
      }
      if (ok) {
        Combine C{};
        C.len = D.len;
        for (int i = 0; i < D.len; ++i) {
          C.c[i] = D.coef[i];
          C.slot[i] = D.slot[i];
        }
        hipLaunchKernelGGL(lg_combine_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, (const double*)L.Fh, ms, C, L.Fu);
        HIP_OK(hipGetLastError());
        Fuse = L.Fu;
      } else {
        Fuse = Fn;
      }
    }
    if (shift != 0.0) {
      hipLaunchKernelGGL(lg_shift_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, Fuse, (const double*)P, L.Fu,
                         n, np, shift);
      HIP_OK(hipGetLastError());
      Fuse = L.Fu;
    }
    if (!large_density(b, d, Fuse, Pn, err, R.last_dp < FRONTIER_SEED_DP)) return false;
    Flast = Fuse;
    hipLaunchKernelGGL(lg_diff2_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, (const double*)Pn, (const double*)P, ms,
                       L.red);
    double s2 = 0.0;
    if (!large_sum1(b, s2, err)) return false;
    const double dp = std::sqrt(s2 / (double)std::max(n * n, 1));
    const double de = std::fabs(e_elec - e_old);
    R.last_dp = dp;
    if (damp > 0.0) {  // P stays L.P (damped), Pn stays L.Pn (the last aufbau projector)
      hipLaunchKernelGGL(lg_damp_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, (const double*)Pn, P, ms, damp);
      HIP_OK(hipGetLastError());
    } else {
      std::swap(P, Pn);
    }
    e_old = e_elec;
    if (it > 0 && de < o.e_tol && dp < o.p_tol) {
      R.converged = 1;
      break;
    }
  }
  if (P != L.P) {  // keep the final density in L.P
    HIP_OK(hipMemcpyAsync(L.P, P, sizeof(double) * ms, hipMemcpyDeviceToDevice, b.stream));
  }
  if (b.purify && Flast && !L.last_eig) {
    // frontier levels of the last step by Lanczos, and the aufbau test HOMO <= LUMO (a last step
    // taken by the eigensolver left every eigenvalue in L.w already); with an empty side, the
    // eigenvalues of the last Fock matrix
    bool ok = true;
    if (damp > 0.0) {  // the levels from the undamped projector, then the damped density back
      HIP_OK(hipMemcpyAsync(L.Pd, L.P, sizeof(double) * ms, hipMemcpyDeviceToDevice, b.stream));
      HIP_OK(hipMemcpyAsync(L.P, L.Pn, sizeof(double) * ms, hipMemcpyDeviceToDevice, b.stream));
    }
    if (d.nocc > 0 && d.nocc < n) {
      if (!large_levels_lanczos(b, d, Flast, ok, err)) return false;
    } else if (!large_levels(b, d, Flast, err)) {
      return false;
    }
    if (damp > 0.0) HIP_OK(hipMemcpyAsync(L.P, L.Pd, sizeof(double) * ms, hipMemcpyDeviceToDevice, b.stream));
    R.aufbau = ok;
  }
  if (!large_fock(b, d, L.P, L.F, err)) return false;
  return large_energy(b, d, L.P, L.F, R.e_elec, err);
}

static bool large_solve(Batch& b, int k, std::string& err) {
  const MolDesc& d = b.desc[k];
  LargeBufs& L = b.lg;
  const int n = d.nao, np = d.np, nocc = d.nocc;
  const long long ms = (long long)np * np;
  HIP_OK(hipMemsetAsync(L.F, 0, sizeof(double) * ms, b.stream));
  hipLaunchKernelGGL(lg_core_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, b.args, d, L.red);
  HIP_OK(hipGetLastError());
  double ecore = 0.0;
  if (!large_sum1(b, ecore, err)) return false;
  const int guess = b.opt.guess == 4 ? 1 : b.opt.guess;  // the learned guess covers batched molecules
  LargePath best;
  L.no_tight = false;
  if (!large_rhf(b, d, guess, b.opt.accel, best, err)) return false;
  if (!best.aufbau) {  // an eigenvalue ended on the wrong side: the path again with Gershgorin bounds
    L.no_tight = true;
    if (!large_rhf(b, d, guess, b.opt.accel, best, err)) return false;
    L.no_tight = false;
  }
  std::vector<double> w(n);
  HIP_OK(scopy(b.stream, w.data(), L.w, sizeof(double) * n, hipMemcpyDeviceToHost));
  const double first_e = best.e_elec;
  int paths = 1;
  const double gap = (nocc > 0 && nocc < n) ? w[nocc] - w[nocc - 1] : 1.0e300;
  // the stability search of the batched kernel: not when steered; a closed shell runs the two
  // unrestricted rungs (src/scf.rs:647-708)
  const bool steered = b.opt.level_shift != 0.0 || b.opt.damping != 0.0;
  const bool search = !steered && (b.opt.stability == 2 || (b.opt.stability == 1 && gap < 6.0));
  if (search) {
    HIP_OK(hipMemcpyAsync(L.Pb, L.P, sizeof(double) * ms, hipMemcpyDeviceToDevice, b.stream));
    HIP_OK(hipMemcpyAsync(L.Vb, L.V, sizeof(double) * ms, hipMemcpyDeviceToDevice, b.stream));
    HIP_OK(hipMemcpyAsync(L.wb, L.w, sizeof(double) * n, hipMemcpyDeviceToDevice, b.stream));
    const int primary = guess == 2 ? 2 : 1;
    const int rg[2][2] = {{3, 0}, {primary, b.opt.use_diis ? 1 : 0}};
    const double rs[2][2] = {{-2.0, 0.0}, {0.0, 0.85}};
    for (int r = 0; r < 2; ++r) {
      ++paths;
      LargePath c;
      if (!large_rhf(b, d, rg[r][0], rg[r][1], c, err, rs[r][0], rs[r][1], true)) return false;
      if (!c.aufbau) {
        L.no_tight = true;
        if (!large_rhf(b, d, rg[r][0], rg[r][1], c, err, rs[r][0], rs[r][1], true)) return false;
        L.no_tight = false;
      }
      if (!c.converged) continue;
      if (!best.converged || c.e_elec < best.e_elec - 1.0e-6) {
        best = c;
        HIP_OK(hipMemcpyAsync(L.Pb, L.P, sizeof(double) * ms, hipMemcpyDeviceToDevice, b.stream));
        HIP_OK(hipMemcpyAsync(L.Vb, L.V, sizeof(double) * ms, hipMemcpyDeviceToDevice, b.stream));
        HIP_OK(hipMemcpyAsync(L.wb, L.w, sizeof(double) * n, hipMemcpyDeviceToDevice, b.stream));
      }
    }
    HIP_OK(hipMemcpyAsync(L.P, L.Pb, sizeof(double) * ms, hipMemcpyDeviceToDevice, b.stream));
    HIP_OK(hipMemcpyAsync(L.V, L.Vb, sizeof(double) * ms, hipMemcpyDeviceToDevice, b.stream));
    HIP_OK(hipMemcpyAsync(L.w, L.wb, sizeof(double) * n, hipMemcpyDeviceToDevice, b.stream));
    HIP_OK(scopy(b.stream, w.data(), L.w, sizeof(double) * n, hipMemcpyDeviceToHost));
  }
  hipLaunchKernelGGL(lg_epilogue_kernel, dim3(1), dim3(NT), 0, b.stream, b.args, d, (const double*)L.P, L.dip);
  HIP_OK(hipGetLastError());
  double dip[3];
  HIP_OK(scopy(b.stream, dip, L.dip, sizeof(dip), hipMemcpyDeviceToHost));
  MolOut mo{};
  mo.e_elec = best.e_elec;
  mo.e_core = ecore;
  mo.last_dp = best.last_dp;
  mo.first_e = first_e;
  mo.homo = nocc > 0 ? w[nocc - 1] : 0.0;
  mo.lumo = nocc < n ? w[nocc] : 0.0;
  for (int x = 0; x < 3; ++x) mo.dip[x] = dip[x];
  mo.iters = best.iters;
  mo.converged = best.converged;
  mo.paths = paths;
  mo.status = best.converged ? 0 : 3;
  mo.guess_used = 0;
  HIP_OK(scopy(b.stream, b.args.out + k, &mo, sizeof(MolOut), hipMemcpyHostToDevice));
  if (d.out_mat >= 0) {
    hipLaunchKernelGGL(lg_pack_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, (const double*)L.P,
                       b.args.out_P + d.out_mat, n, np);
    if (!b.purify) {  // orbitals only on the eigenvector route
      hipLaunchKernelGGL(lg_vectors_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, (const double*)L.V,
                         b.args.out_C + d.out_mat, n, np);
      HIP_OK(hipMemcpyAsync(b.args.out_eps + d.out_vec, L.w, sizeof(double) * n, hipMemcpyDeviceToDevice, b.stream));
    }
    HIP_OK(hipGetLastError());
  }
  return true;
}

// Directed pairs of the learned guess (molecules made of H, C and O only): for every source atom
// in the engine's order, its neighbours inside 4 Angstrom in the caller's atom order, the pairs
// of one molecule contiguous and the molecules in caller order.  Sets each single-workgroup
// molecule's learn flag and pair range; positions are the engine-order atom coordinates (Bohr).
struct GuessPairs {
  std::vector<int> gbeg, gcnt, gsrc, gdst, gpart;
};
static void build_guess_pairs(Batch& b, const double* ax, const double* ay, const double* az, GuessPairs& g) {
  g.gbeg.assign(b.atoms_total, 0);
  g.gcnt.assign(b.atoms_total, 0);
  g.gsrc.clear();
  g.gdst.clear();
  g.gpart.clear();
  b.eligible.assign(b.nmol, 0);
  std::vector<int> mols = b.solve;
  std::sort(mols.begin(), mols.end());
  for (int k : mols) {
    MolDesc& d = b.desc[k];
    bool ok = d.uhf == 0;  // the learned guess covers restricted molecules
    for (int t = 0; t < d.nat; ++t) ok = ok && b.h_elem8[d.atom0 + t] >= 0;
    d.learn = ok ? 1 : 0;
    d.gp0 = (long long)g.gsrc.size();
    d.ngp = 0;
    b.eligible[k] = d.learn;
    if (!ok) continue;
    const std::vector<int>& ord = b.atom_map[k];
    std::vector<int> inv(d.nat);
    for (int t = 0; t < d.nat; ++t) inv[ord[t]] = t;
    std::vector<int> pidx((size_t)d.nat * d.nat, -1);
    for (int sa = 0; sa < d.nat; ++sa) {
      const int ga = d.atom0 + sa;
      g.gbeg[ga] = (int)g.gsrc.size();
      for (int j = 0; j < d.nat; ++j) {
        const int sb = inv[j];
        if (sb == sa) continue;
        const int gb = d.atom0 + sb;
        const double dx = ax[gb] - ax[ga], dy = ay[gb] - ay[ga], dz = az[gb] - az[ga];
        if (!(std::sqrt(dx * dx + dy * dy + dz * dz) * pm3guess::BOHR_A < pm3guess::CUT_A)) continue;
        pidx[(size_t)sa * d.nat + sb] = (int)g.gsrc.size();
        g.gsrc.push_back(ga);
        g.gdst.push_back(gb);
      }
      g.gcnt[ga] = (int)g.gsrc.size() - g.gbeg[ga];
    }
    d.ngp = (int)((long long)g.gsrc.size() - d.gp0);
    g.gpart.resize(g.gsrc.size());
    for (long long j = d.gp0; j < (long long)g.gsrc.size(); ++j)
      g.gpart[j] = pidx[(size_t)(g.gdst[j] - d.atom0) * d.nat + (g.gsrc[j] - d.atom0)];
  }
}

static bool batch_create_in(const Params& P, const std::vector<MolIn>& mols, const EngineOptions& opt,
                            const RunRequest& req, Batch*& out_batch, std::string& err, unsigned char* arena,
                            size_t arena_bytes, hipStream_t arena_stream);

// Atomic numbers with a PM3 parameter block in the reference (the rows of its
// src/data/pm3_parameters.csv, plus the lanthanide sparkles 57-71 and MOPAC's capped bond 102).
// An element there that this engine lacks (d shells, principal quantum number 4 and above, the
// noble gases) is refused as unsupported; any other is a missing element (src/error.rs:19-24).
static bool reference_parameterises(int z) {
  return (z >= 1 && z <= 20) || (z >= 30 && z <= 38) || (z >= 48 && z <= 71) || (z >= 80 && z <= 83) || z == 102;
}

bool batch_create(const Params& P, const std::vector<MolIn>& mols, const EngineOptions& opt, const RunRequest& req,
                  Batch*& out_batch, std::string& err) {
  return batch_create_in(P, mols, opt, req, out_batch, err, nullptr, 0, nullptr);
}

static bool batch_create_in(const Params& P, const std::vector<MolIn>& mols, const EngineOptions& opt,
                            const RunRequest& req, Batch*& out_batch, std::string& err, unsigned char* arena,
                            size_t arena_bytes, hipStream_t arena_stream) {
  out_batch = nullptr;
  if (!upload_table(P, err)) return false;
  Batch* bp = new Batch();
  Batch& b = *bp;
  if (arena) {
    b.arena = arena;
    b.arena_cap = arena_bytes;
    b.stream = arena_stream;
    b.own_stream = false;
  } else if (hipStreamCreateWithFlags(&b.stream, hipStreamNonBlocking) != hipSuccess) {
    err = "hipStreamCreateWithFlags failed";
    delete bp;
    return false;
  }
  const int nmol = (int)mols.size();
  b.nmol = nmol;
  b.desc.resize(nmol);
  b.ao_map.resize(nmol);
  b.atom_map.resize(nmol);
  b.pre_status.assign(nmol, 0);
  b.pre_z.assign(nmol, 0);
  b.pre_req.assign(nmol, 0.0);
  b.pre_lim.assign(nmol, 0.0);
  b.cls_of.assign(nmol, 0);
  const Tuning tune = tuning_table();
  // the pair-cache ceiling of build_core_limited (src/hamiltonian.rs:44-49, 202-211)
  long long cache_limit_mb = opt.integral_memory_mb;
  if (cache_limit_mb <= 0) {
    cache_limit_mb = 4096;
    if (const char* ev = std::getenv("PM3_MAX_PAIR_CACHE_MB")) cache_limit_mb = std::strtoll(ev, nullptr, 10);
  }
  // Without orbitals to return, each SCF density is purified from the Fock matrix instead of
  // built from its eigenvectors (PM3_PURIFY=0 keeps the eigensolver).
  b.purify = !req.want_orbitals;
  if (const char* ev = std::getenv("PM3_PURIFY")) b.purify = b.purify && ev[0] != '0';
  std::vector<int> aslot, aoff;
  std::vector<double> ax, ay, az;
  int np_max = 16;
  long long w_cap = 1;
  for (int k = 0; k < nmol; ++k) {
    const MolIn& m = mols[k];
    MolDesc& d = b.desc[k];
    d = MolDesc{};
    d.out_mat = d.out_vec = -1;
    d.atom0 = (int)aslot.size();
    int& st = b.pre_status[k];
    if (m.n_atoms < 1) { st = 1; continue; }
    double nel = 0.0;
    for (int i = 0; i < m.n_atoms && !st; ++i) {
      const int z = m.z[i];
      if (z < 0 || z >= 128 || P.slot_of_z[z] < 0) {
        st = reference_parameterises(z) ? 4 : 2;
        b.pre_z[k] = z;
      } else {
        nel += P.table.el[P.slot_of_z[z]].tore;
      }
    }
    if (st) continue;
    // electron and spin counts, restricted or unrestricted (src/scf.rs:486-520)
    nel -= m.charge;
    const long ne = std::lround(nel);
    if (std::fabs(nel - (double)ne) > 1.0e-6 || ne < 0 || m.multiplicity < 1) { st = 1; continue; }
    const long nu = m.multiplicity - 1;
    if (ne - nu < 0 || (ne - nu) % 2 != 0) { st = 1; continue; }
    const int n_alpha = (int)((ne + nu) / 2), n_beta = (int)((ne - nu) / 2);
    if (opt.reference == 1 && n_alpha != n_beta) { st = 1; continue; }
    const bool use_uhf = opt.reference == 2 || n_alpha != n_beta;
    {
      // exact size of the reference's pair cache: sum_{i<j} p_i p_j packed doubles and a 72-byte
      // PairIntegral record per pair (pair_cache_bytes, src/hamiltonian.rs:51-71)
      double tot = 0.0, sq = 0.0;
      for (int i = 0; i < m.n_atoms; ++i) {
        const int no = P.table.el[P.slot_of_z[m.z[i]]].norb;
        const double pk_ = std::max(no * (no + 1) / 2, 1);
        tot += pk_;
        sq += pk_ * pk_;
      }
      const double need = (tot * tot - sq) / 2.0 * 8.0 + (double)m.n_atoms * (m.n_atoms - 1) / 2.0 * 72.0;
      if (cache_limit_mb > 0 && need > (double)cache_limit_mb * 1048576.0) {
        st = 5;
        b.pre_req[k] = std::ceil(need / 1048576.0);
        b.pre_lim[k] = (double)cache_limit_mb;
        continue;
      }
    }
    bool coincident = false;
    for (int i = 0; i < m.n_atoms && !coincident; ++i)
      for (int j = 0; j < i; ++j) {
        const double dx = m.xyz[3 * i] - m.xyz[3 * j], dy = m.xyz[3 * i + 1] - m.xyz[3 * j + 1], dz = m.xyz[3 * i + 2] - m.xyz[3 * j + 2];
        if (!(dx * dx + dy * dy + dz * dz > 1.0e-16)) { coincident = true; break; }
      }
    if (coincident) { st = 1; continue; }
    // heavy atoms first (stable), so the pair tables split into XX / XH / HH blocks
    std::vector<int> ord(m.n_atoms);
    std::iota(ord.begin(), ord.end(), 0);
    auto heavy = [&](int i) { return P.table.el[P.slot_of_z[m.z[i]]].norb == 4; };
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return heavy(x) && !heavy(y); });
    std::vector<int> coff(m.n_atoms + 1, 0);
    for (int i = 0; i < m.n_atoms; ++i) coff[i + 1] = coff[i] + P.table.el[P.slot_of_z[m.z[i]]].norb;
    int nao = 0, nh = 0;
    for (int i : ord) {
      const int s = P.slot_of_z[m.z[i]];
      const int no = P.table.el[s].norb;
      nh += no == 4;
      aslot.push_back(s);
      aoff.push_back(nao);
      ax.push_back(m.xyz[3 * i]);
      ay.push_back(m.xyz[3 * i + 1]);
      az.push_back(m.xyz[3 * i + 2]);
      for (int k2 = 0; k2 < no; ++k2) b.ao_map[k].push_back(coff[i] + k2);
      nao += no;
    }
    b.atom_map[k] = ord;
    d.nat = m.n_atoms;
    d.nheavy = nh;
    d.nao = nao;
    d.nocc = n_alpha;
    d.nbeta = n_beta;
    d.uhf = use_uhf ? 1 : 0;
    d.np = (nao + 15) & ~15;
    // eigensolver LDS: packed matrix + rotation schedule, plus the eigenvectors when they fit
    // (with purification the eigensolver only runs when a gap closes: its vectors stay in HBM)
    auto fits = [](int doubles) { return doubles * 8 + STATIC_LDS_BYTES <= LDS_BYTES_MAX; };
    int jl = jac_lds_base(nao);
    if (!b.purify && fits(jac_lds_with_v(nao))) jl = jac_lds_with_v(nao);
    else if (!b.purify && fits(jac_lds_deferred(nao))) jl = jac_lds_deferred(nao);
    if (b.purify)
      jl = std::max(jl, fits(lanczos_lds_square(nao))   ? lanczos_lds_square(nao)
                        : fits(lanczos_lds_packed(nao)) ? lanczos_lds_packed(nao)
                                                        : lanczos_lds(nao));
    const long long lneed = (long long)std::max(jl, m.n_atoms * 16) * 8;
    // molecules that do not fit one workgroup's LDS take the host-driven multi-kernel path
    // and so does a lone restricted molecule above solo_large_nao AOs: on one workgroup it would
    // leave the rest of the GPU idle, on the multi-kernel path each step spreads over every CU
    // (the multi-kernel path starts from the superposition guess: the learned guess covers the
    // molecules of the persistent kernel)
    const bool solo = req.solo_large && !arena && nmol == 1 && !use_uhf && nao > tune.solo_large_nao;
    const bool large = lneed + STATIC_LDS_BYTES > LDS_BYTES_MAX || req.force_large || solo;
    // unrestricted molecules take the single-workgroup path only
    if (d.nocc > nao || (large && nao > LARGE_MAX_NAO) || (large && use_uhf && !req.dc)) {
      st = d.nocc > nao ? 1 : large && use_uhf ? 4 : 5;
      aslot.resize(d.atom0); aoff.resize(d.atom0); ax.resize(d.atom0); ay.resize(d.atom0); az.resize(d.atom0);
      d.nat = 0;
      continue;
    }
    if (!large) {
      b.any_uhf = b.any_uhf || use_uhf;
      b.lds_need = std::max(b.lds_need, (int)lneed);
      np_max = std::max(np_max, d.np);
      const long long npairs = (long long)d.nat * (d.nat - 1) / 2;
      w_cap = std::max(w_cap, 20LL * npairs);  // Fock-build pair slots per workgroup
      // workgroup width: the narrowest whose share of a CU's LDS holds what the molecule cannot
      // do without (the Jacobi fallback's packed matrix, the Lanczos vectors); every other LDS
      // placement adapts to the launch's budget on the device
      const long long lmin =
          8LL * std::max({jac_lds_base(nao), b.purify ? lanczos_lds(nao) : 0, m.n_atoms * 16});
      int c = 0;
      if (nao <= tune.wg128_nao && lmin + STATIC_LDS_BYTES <= LDS_BYTES_MAX / SCF_PER_CU[2]) c = 2;
      else if (nao <= tune.wg256_nao && lmin + STATIC_LDS_BYTES <= LDS_BYTES_MAX / SCF_PER_CU[1]) c = 1;
      b.cls_of[k] = (int8_t)c;
      ScfClass& sc = b.cls[c];
      sc.any_uhf = sc.any_uhf || use_uhf;
      sc.np_max = std::max(sc.np_max, d.np);
      sc.w_cap = std::max(sc.w_cap, 20LL * npairs);
    }
    if (req.want_density || req.want_orbitals || req.want_gradient || req.warm_start) {
      d.out_mat = b.mat_total;
      d.out_vec = b.vec_total;
      b.mat_total += (long long)nao * nao;
      b.vec_total += nao;
    }
    if (large && req.dc) b.any_uhf = b.any_uhf || use_uhf;  // the spin density slot of its gradient
    (large ? b.large : b.solve).push_back(k);
  }
  b.atoms_total = (long long)aslot.size();
  b.nsolve = (int)b.solve.size();
  // One 512-thread workgroup per CU is resident whatever its LDS (its registers decide), so the
  // launch takes all of it: eigenvectors, purification operands and Lanczos matrices then stay in
  // LDS for as many molecules as they fit.
  if (b.nsolve > 0) b.lds_need = LDS_BYTES_MAX - STATIC_LDS_BYTES;
  std::vector<int8_t> elem8(b.atoms_total, (int8_t)-1);
  for (long long t = 0; t < b.atoms_total; ++t) elem8[t] = (int8_t)pm3guess::elem_index(P.table.el[aslot[t]].z);
  b.h_elem8 = elem8;
  GuessPairs gp;
  build_guess_pairs(b, ax.data(), ay.data(), az.data(), gp);
  std::vector<int>&gbeg = gp.gbeg, &gcnt = gp.gcnt, &gsrc = gp.gsrc, &gdst = gp.gdst, &gpart = gp.gpart;
  b.ngp_total = (long long)gsrc.size();
  b.gp_cap = b.ngp_total;
  if (b.ngp_total >= (1LL << 31)) {
    err = "too many directed atom pairs in one batch";
    delete bp;
    return false;
  }
  // by class (widest workgroups first), and inside a class largest molecules first so that its
  // queue ends on short work
  std::stable_sort(b.solve.begin(), b.solve.end(), [&](int x, int y) {
    return b.cls_of[x] != b.cls_of[y] ? b.cls_of[x] < b.cls_of[y] : b.desc[x].nao > b.desc[y].nao;
  });
  for (int c = 0, off = 0; c < 3; ++c) {
    b.cls[c].off = off;
    b.cls[c].n = 0;
    for (int k : b.solve) b.cls[c].n += b.cls_of[k] == c;
    off += b.cls[c].n;
  }
  b.order_all = b.solve;
  b.order_all.insert(b.order_all.end(), b.large.begin(), b.large.end());
  const int nq = (int)b.order_all.size();
  b.opt = opt;
  b.table = P.table;
  // molecule store: H_core (np x np) then W_xx | W_xh | W_hh | S1 | S2 | e_cc per molecule, and the
  // first global pair index of every queue position for the pair kernel
  std::vector<long long> pair_base(nq + 1, 0);
  long long store_total = 0;
  for (int q = 0; q < nq; ++q) {
    MolDesc& d = b.desc[b.order_all[q]];
    const long long nh = d.nheavy, nH = d.nat - d.nheavy, npairs = (long long)d.nat * (d.nat - 1) / 2;
    d.h_off = store_total;
    store_total += ((long long)d.np * d.np + 7) & ~7LL;
    d.w_off = store_total;
    store_total += (100 * nh * (nh - 1) / 2 + 10 * nh * nH + nH * (nH - 1) / 2 + 21 * npairs + 7) & ~7LL;
    pair_base[q + 1] = pair_base[q] + npairs;
  }
  out_batch = bp;
  if (nq == 0) return true;
  int dev = 0;
  hipDeviceProp_t prop;
  HIP_OK(hipGetDevice(&dev));
  HIP_OK(hipGetDeviceProperties(&prop, dev));
  b.nwg = std::max(1, std::min(std::max(b.nsolve, 1), prop.multiProcessorCount));
  // each class's workgroups get slabs sized for that class's largest molecule
  long long scr_total = 0;
  for (int c = 0; c < 3; ++c) {
    ScfClass& sc = b.cls[c];
    const long long cm2 = (long long)sc.np_max * sc.np_max;
    sc.ustride = sc.any_uhf ? ((9 + 2 * MAXD) * cm2 + 4LL * sc.np_max + 31) & ~31LL : 0;
    sc.stride = (((11 + 3 * MAXD) * cm2 + sc.w_cap + 6LL * sc.np_max + 64 + 31) & ~31LL) + sc.ustride;
    sc.nwg = sc.n > 0 ? std::min(sc.n, SCF_PER_CU[c] * prop.multiProcessorCount) : 0;
    sc.lds = LDS_BYTES_MAX / SCF_PER_CU[c] - STATIC_LDS_BYTES;
    sc.scr_off = scr_total;
    scr_total += sc.stride * sc.nwg;
  }
  // the stage kernel and the molecule at the head of the queue use the slab layout of its class
  const int c0 = b.nsolve > 0 ? b.cls_of[b.solve[0]] : 0;
  if (b.nsolve == 0) {
    const long long cm2 = (long long)np_max * np_max;
    b.cls[0].ustride = b.any_uhf ? ((9 + 2 * MAXD) * cm2 + 4LL * np_max + 31) & ~31LL : 0;
    b.cls[0].stride = (((11 + 3 * MAXD) * cm2 + w_cap + 6LL * np_max + 64 + 31) & ~31LL) + b.cls[0].ustride;
    b.cls[0].np_max = np_max;
    b.cls[0].w_cap = w_cap;
    scr_total = b.cls[0].stride;
  }
  const long long ustride = b.cls[c0].ustride;
  const long long stride = b.cls[c0].stride;
  KArgs& a = b.args;
  MolDesc* d_desc; int *d_order, *d_queue, *d_aslot, *d_aoff;
  double *d_ax, *d_ay, *d_az, *d_scr, *d_P, *d_C, *d_eps, *d_q, *d_store, *d_g;
  long long* d_pbase;
  MolOut* d_out;
  if (!b.alloc(d_desc, nmol, err) || !b.alloc(d_order, nq, err) || !b.alloc(d_queue, 3, err) ||
      !b.alloc(d_aslot, b.atoms_total, err) || !b.alloc(d_aoff, b.atoms_total, err) || !b.alloc(d_ax, b.atoms_total, err) ||
      !b.alloc(d_ay, b.atoms_total, err) || !b.alloc(d_az, b.atoms_total, err) || !b.alloc(d_scr, (size_t)scr_total, err) ||
      !b.alloc(d_out, nmol, err) || !b.alloc(d_P, b.mat_total, err) || !b.alloc(d_C, b.mat_total, err) ||
      !b.alloc(d_eps, b.vec_total, err) || !b.alloc(d_q, b.atoms_total, err) ||
      !b.alloc(d_store, (size_t)store_total, err) || !b.alloc(d_pbase, pair_base.size(), err) ||
      !b.alloc(d_g, 3 * (size_t)b.atoms_total, err))
    return false;
  int8_t* d_elem8;
  int *d_gbeg, *d_gcnt, *d_gsrc, *d_gdst, *d_gpart;
  if (!b.alloc(d_elem8, b.atoms_total, err) || !b.alloc(d_gbeg, b.atoms_total, err) ||
      !b.alloc(d_gcnt, b.atoms_total, err) || !b.alloc(d_gsrc, gsrc.size(), err) || !b.alloc(d_gdst, gdst.size(), err) ||
      !b.alloc(d_gpart, gpart.size(), err))
    return false;
  HIP_OK(scopy(b.stream, d_elem8, elem8.data(), elem8.size(), hipMemcpyHostToDevice));
  HIP_OK(scopy(b.stream, d_gbeg, gbeg.data(), sizeof(int) * gbeg.size(), hipMemcpyHostToDevice));
  HIP_OK(scopy(b.stream, d_gcnt, gcnt.data(), sizeof(int) * gcnt.size(), hipMemcpyHostToDevice));
  if (b.ngp_total > 0) {
    HIP_OK(scopy(b.stream, d_gsrc, gsrc.data(), sizeof(int) * gsrc.size(), hipMemcpyHostToDevice));
    HIP_OK(scopy(b.stream, d_gdst, gdst.data(), sizeof(int) * gdst.size(), hipMemcpyHostToDevice));
    HIP_OK(scopy(b.stream, d_gpart, gpart.data(), sizeof(int) * gpart.size(), hipMemcpyHostToDevice));
  }
  b.want_gradient = req.want_gradient;
  b.keep_density = req.want_density;
  b.warm_start = req.warm_start;
  HIP_OK(smemset(b.stream, d_store, 0, sizeof(double) * std::max<long long>(store_total, 1)));
  // result records start zeroed (converged = 0): a warm start never reads a record no solve wrote
  HIP_OK(smemset(b.stream, d_out, 0, sizeof(MolOut) * std::max(nmol, 1)));
  HIP_OK(scopy(b.stream, d_pbase, pair_base.data(), sizeof(long long) * pair_base.size(), hipMemcpyHostToDevice));
  HIP_OK(scopy(b.stream, d_desc, b.desc.data(), sizeof(MolDesc) * nmol, hipMemcpyHostToDevice));
  HIP_OK(scopy(b.stream, d_order, b.order_all.data(), sizeof(int) * nq, hipMemcpyHostToDevice));
  HIP_OK(scopy(b.stream, d_aslot, aslot.data(), sizeof(int) * aslot.size(), hipMemcpyHostToDevice));
  HIP_OK(scopy(b.stream, d_aoff, aoff.data(), sizeof(int) * aoff.size(), hipMemcpyHostToDevice));
  HIP_OK(scopy(b.stream, d_ax, ax.data(), sizeof(double) * ax.size(), hipMemcpyHostToDevice));
  HIP_OK(scopy(b.stream, d_ay, ay.data(), sizeof(double) * ay.size(), hipMemcpyHostToDevice));
  HIP_OK(scopy(b.stream, d_az, az.data(), sizeof(double) * az.size(), hipMemcpyHostToDevice));
  a.mols = d_desc;
  a.order = d_order;
  a.nmol = b.nsolve;
  a.nmol_pairs = nq;
  a.queue = d_queue;
  a.aslot = d_aslot;
  a.ax = d_ax;
  a.ay = d_ay;
  a.az = d_az;
  a.aoff = d_aoff;
  b.scr_base = d_scr;
  a.scratch = d_scr + b.cls[c0].scr_off;
  a.stride = stride;
  a.np_max = b.cls[c0].np_max;
  a.w_cap = b.cls[c0].w_cap;
  a.out = d_out;
  a.out_P = d_P;
  a.out_C = d_C;
  a.out_eps = d_eps;
  a.out_q = d_q;
  a.out_g = d_g;
  a.ustride = ustride;
  a.out_S = nullptr;
  a.out_Cb = nullptr;
  a.out_epsb = nullptr;
  if (b.any_uhf) {
    if (!b.alloc(a.out_S, b.mat_total, err)) return false;
    if (!b.purify && (!b.alloc(a.out_Cb, b.mat_total, err) || !b.alloc(a.out_epsb, b.vec_total, err))) return false;
  }
  a.store = d_store;
  a.pair_base = d_pbase;
  a.npairs_total = pair_base.back();
  a.prof = nullptr;
  a.lds_doubles = b.lds_need / 8;
  a.elem8 = d_elem8;
  a.gbeg = d_gbeg;
  a.gcnt = d_gcnt;
  a.gsrc = d_gsrc;
  a.gdst = d_gdst;
  a.gpart = d_gpart;
  a.gY = nullptr;
  b.guess_model = opt.guess_model;
  b.learned = opt.guess == 4;
  if (b.learned) {
    if (!b.guess_model) {
      err = "learned guess requested without a guess model";
      return false;
    }
    if (!b.ensure_guess_buffers(err)) return false;
  }
  if (const char* ev = std::getenv("PM3_PHASE_TIMERS")) {
    if (ev[0] == '1') {
      unsigned long long* d_prof;
      if (!b.alloc(d_prof, PH_SLOTS, err)) return false;
      a.prof = d_prof;
    }
  }
  b.pair_block = tune.pair_block;
  b.pair_blocks = (int)std::max<long long>(
      1, std::min<long long>((a.npairs_total + b.pair_block - 1) / b.pair_block,
                             (long long)tune.pair_blocks_per_cu * prop.multiProcessorCount));
  b.large_syrk = tune.large_syrk != 0;
  b.large_split = tune.large_split;
  b.large_steps_n = LARGE_STEPS_N;  // the environment may lower it, so that tests reach the larger-matrix kernels
  if (const char* ev = std::getenv("PM3_LARGE_STEPS_N")) b.large_steps_n = std::atoi(ev);
  if (const char* ev = std::getenv("PM3_SY_K")) {  // the slab width of the tile kernel: 8, 16 (default) or 32
    const int k = std::atoi(ev);
    b.sy_k = k == 8 || k == 32 ? k : 16;
  }
  a.max_scf = opt.max_scf;
  a.e_tol = opt.e_tol;
  a.p_tol = opt.p_tol;
  a.accel = opt.accel;
  a.adiis_switch = opt.adiis_switch;
  a.guess = opt.guess;
  a.stability = opt.stability;
  a.scf_memory_mb = opt.scf_memory_mb;
  a.use_diis = opt.use_diis;
  a.shift = opt.level_shift;
  a.damp = opt.damping;
  a.has_field = opt.has_field;
  for (int x = 0; x < 3; ++x) a.field[x] = opt.has_field ? opt.field[x] : 0.0;
  a.eig_tol_factor = EIG_TOL_FACTOR;
  if (const char* ev = std::getenv("PM3_EIG_TOL_FACTOR")) a.eig_tol_factor = std::atof(ev);
  a.eig_rel = EIG_REL;
  if (const char* ev = std::getenv("PM3_EIG_REL")) a.eig_rel = std::atof(ev);
  a.eig_skip = EIG_SKIP;
  a.purify = b.purify ? 1 : 0;
  a.sp2_lds = tune.sp2_lds;
  a.sp2_tight_n = tune.sp2_tight_n;
  a.sp2_tile16_n = tune.sp2_tile16_n;
  a.sp2_margin = SP2_MARGIN;
  if (const char* ev = std::getenv("PM3_SP2_MARGIN")) a.sp2_margin = std::atof(ev);
  if (const char* ev = std::getenv("PM3_SP2_TIGHT_N")) a.sp2_tight_n = std::atoi(ev);
  if (const char* ev = std::getenv("PM3_SP2_LDS")) a.sp2_lds = std::atoi(ev);
  if (const char* ev = std::getenv("PM3_EIG_SKIP")) a.eig_skip = std::atof(ev);
  HIP_OK(hipFuncSetAttribute((const void*)scf_batch_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, b.lds_need));
  if (pm3variant::nt256_kargs_bytes() != sizeof(KArgs) || pm3variant::nt128_kargs_bytes() != sizeof(KArgs)) {
    err = "the SCF kernel builds disagree on their argument layout";
    return false;
  }
  if (b.cls[1].n > 0) HIP_OK(pm3variant::nt256_prepare(b.cls[1].lds));
  if (b.cls[2].n > 0) HIP_OK(pm3variant::nt128_prepare(b.cls[2].lds));
  if (!b.large.empty() && b.arena) {
    err = "the allocation-free entry point serves molecules of the single-workgroup path only";
    return false;
  }
  if (!b.large.empty() && !large_setup(b, err)) return false;
  return true;
}

static BatchGuessView infer_view(Batch& b) {
  BatchGuessView v;
  v.ngp_total = b.ngp_total;
  v.atoms_total = b.atoms_total;
  v.ax = b.args.ax; v.ay = b.args.ay; v.az = b.args.az;
  v.elem = b.args.elem8;
  v.aoff = b.args.aoff;
  v.gbeg = b.args.gbeg; v.gcnt = b.args.gcnt; v.gsrc = b.args.gsrc; v.gdst = b.args.gdst; v.gpart = b.args.gpart;
  v.X = b.d_X;
  v.Y = b.d_Y;
  return v;
}

// The persistent SCF launches, one per workgroup width in use: the widest on the batch's stream,
// the narrower ones on streams of their own that start after the same point of the batch's stream
// and that the batch's stream then waits for, so that they fill the CUs the widest launch's tail
// frees.  Without events (the allocation-free mode) they run one after another on the batch's
// stream.
static bool launch_scf(Batch& b, std::string& err) {
  int used = 0;
  for (const ScfClass& sc : b.cls) used += sc.n > 0;
  const bool fork = used > 1 && b.arena == nullptr;
  if (fork && !b.ev_fork) {
    HIP_OK(hipEventCreateWithFlags(&b.ev_fork, hipEventDisableTiming));
    for (int i = 0; i < 2; ++i) {
      HIP_OK(hipStreamCreateWithFlags(&b.aux[i], hipStreamNonBlocking));
      HIP_OK(hipEventCreateWithFlags(&b.ev_join[i], hipEventDisableTiming));
    }
  }
  if (fork) HIP_OK(hipEventRecord(b.ev_fork, b.stream));
  int naux = 0;
  bool first = true;
  for (int c = 0; c < 3; ++c) {
    const ScfClass& sc = b.cls[c];
    if (sc.n == 0) continue;
    KArgs a = b.args;
    a.order = b.args.order + sc.off;
    a.nmol = sc.n;
    a.queue = b.args.queue + c;
    a.scratch = b.scr_base + sc.scr_off;
    a.stride = sc.stride;
    a.ustride = sc.ustride;
    a.np_max = sc.np_max;
    a.w_cap = sc.w_cap;
    a.lds_doubles = sc.lds / 8;
    hipStream_t s = b.stream;
    if (fork && !first) {
      s = b.aux[naux];
      HIP_OK(hipStreamWaitEvent(s, b.ev_fork, 0));
    }
    if (c == 0) {
      hipLaunchKernelGGL(scf_batch_kernel, dim3(sc.nwg), dim3(SCF_WIDTH[0]), sc.lds, s, a);
      HIP_OK(hipGetLastError());
    } else if (c == 1) {
      HIP_OK(pm3variant::nt256_launch(&a, sc.nwg, sc.lds, s));
    } else {
      HIP_OK(pm3variant::nt128_launch(&a, sc.nwg, sc.lds, s));
    }
    if (s != b.stream) HIP_OK(hipEventRecord(b.ev_join[naux++], s));
    first = false;
  }
  for (int i = 0; i < naux; ++i) HIP_OK(hipStreamWaitEvent(b.stream, b.ev_join[i], 0));
  return true;
}

bool batch_solve(Batch* bp, std::string& err) {
  Batch& b = *bp;
  b.last_kernel_ms = 0.0;
  b.lg.product_ms = b.lg.fock_ms = 0.0;
  b.lg.products = b.lg.fock_builds = b.lg.product_flop = 0;
  b.lg.npev = b.lg.nfev = 0;
  b.lg.lanczos_steps = b.lg.lanczos_runs = 0;
  if (b.order_all.empty()) return true;
  const bool timed = b.arena == nullptr;  // the allocation-free mode creates no events
  if (timed && !b.ev0) {
    HIP_OK(hipEventCreate(&b.ev0));
    HIP_OK(hipEventCreate(&b.ev1));
    HIP_OK(hipEventCreate(&b.ev2));
    HIP_OK(hipEventCreate(&b.evg));
    HIP_OK(hipEventCreate(&b.evm));
  }
  auto rec = [&](hipEvent_t e) { return timed ? hipEventRecord(e, b.stream) : hipSuccess; };
  HIP_OK(hipMemsetAsync(b.args.queue, 0, 3 * sizeof(int), b.stream));
  if (b.args.prof) HIP_OK(hipMemsetAsync(b.args.prof, 0, PH_SLOTS * sizeof(unsigned long long), b.stream));
  HIP_OK(rec(b.ev0));
  hipLaunchKernelGGL(pair_kernel, dim3(b.pair_blocks), dim3(b.pair_block), 0, b.stream, b.args);
  HIP_OK(hipGetLastError());
  HIP_OK(rec(b.evg));
  if (b.learned && !guess_infer(b.guess_model, infer_view(b), err, b.stream, timed ? b.evm : nullptr)) return false;
  if (!b.learned) HIP_OK(rec(b.evm));
  HIP_OK(rec(b.ev1));
  b.args.warm = b.warm_start && (b.solved_once || b.have_prev) ? 1 : 0;
  if (b.nsolve > 0 && !launch_scf(b, err)) return false;
  for (int k : b.large)
    if (!large_solve(b, k, err)) return false;
  if (b.want_gradient) {
    hipLaunchKernelGGL(grad_pair_kernel, dim3(b.pair_blocks), dim3(b.pair_block), 0, b.stream, b.args);
    if (b.nsolve > 0) hipLaunchKernelGGL(grad_gather_kernel, dim3((unsigned)b.nsolve), dim3(256), 0, b.stream, b.args);
    if (!b.large.empty()) {
      int nat = 1;
      for (int k : b.large) nat = std::max(nat, b.desc[k].nat);
      hipLaunchKernelGGL(grad_gather_wide_kernel, dim3((unsigned)b.large.size(), (unsigned)std::min(256, (nat + 3) / 4)),
                         dim3(256), 0, b.stream, b.args, b.nsolve);
    }
    HIP_OK(hipGetLastError());
  }
  if (!timed) {
    HIP_OK(hipStreamSynchronize(b.stream));
    b.solved_once = true;
    return true;
  }
  HIP_OK(hipEventRecord(b.ev2, b.stream));
  HIP_OK(hipEventSynchronize(b.ev2));
  if (!pool_ms(b.lg.pev, b.lg.npev, b.lg.product_ms, err) || !pool_ms(b.lg.fev, b.lg.nfev, b.lg.fock_ms, err)) return false;
  float ms = 0.0f, pms = 0.0f, gms = 0.0f;
  HIP_OK(hipEventElapsedTime(&pms, b.ev0, b.evg));
  HIP_OK(hipEventElapsedTime(&gms, b.evg, b.ev1));
  HIP_OK(hipEventElapsedTime(&ms, b.ev1, b.ev2));
  b.last_pair_ms = pms;
  b.last_guess_ms = gms;
  float fms = 0.0f;
  HIP_OK(hipEventElapsedTime(&fms, b.evg, b.evm));
  b.last_feature_ms = fms;
  b.last_network_ms = gms - fms;
  b.last_kernel_ms = ms;
  b.solved_once = true;
  return true;
}

bool batch_guess_view(Batch* bp, BatchGuessView& v, std::string& err) {
  Batch& b = *bp;
  if (!b.ensure_guess_buffers(err)) return false;
  v = infer_view(b);
  v.nmol = b.nmol;
  v.atom0.assign(b.nmol, 0); v.nat.assign(b.nmol, 0); v.nao.assign(b.nmol, 0); v.ngp.assign(b.nmol, 0);
  v.eligible.assign(b.nmol, 0); v.solved.assign(b.nmol, 0); v.gp0.assign(b.nmol, 0); v.out_mat.assign(b.nmol, -1);
  std::vector<MolOut> ho;
  if (b.solved_once && !b.order_all.empty()) {
    ho.resize(b.nmol);
    HIP_OK(scopy(b.stream, ho.data(), b.args.out, sizeof(MolOut) * b.nmol, hipMemcpyDeviceToHost));
  }
  v.caller_atom.assign(b.atoms_total, 0);
  for (int k : b.order_all) {
    const MolDesc& d = b.desc[k];
    for (int t = 0; t < d.nat; ++t) v.caller_atom[d.atom0 + t] = b.atom_map[k][t];
    v.atom0[k] = d.atom0; v.nat[k] = d.nat; v.nao[k] = d.nao; v.ngp[k] = d.ngp; v.gp0[k] = d.gp0;
    v.out_mat[k] = d.out_mat;
    v.eligible[k] = b.eligible[k];
    v.solved[k] = !ho.empty() && ho[k].status == 0 ? 1 : 0;
  }
  v.out_P = b.mat_total > 0 ? b.args.out_P : nullptr;
  return true;
}

double batch_guess_ms(const Batch* b) { return b ? b->last_guess_ms : 0.0; }
int batch_large_count(const Batch* b) { return b ? (int)b->large.size() : 0; }
int batch_small_count(const Batch* b) { return b ? b->nsolve : 0; }
void batch_large_timing(const Batch* b, double& product_ms, long long& products, long long& product_flop,
                        double& fock_ms, long long& fock_builds, long long& lanczos_steps, long long& lanczos_runs) {
  lanczos_steps = b ? b->lg.lanczos_steps : 0;
  lanczos_runs = b ? b->lg.lanczos_runs : 0;
  product_ms = b ? b->lg.product_ms : 0.0;
  products = b ? b->lg.products : 0;
  product_flop = b ? b->lg.product_flop : 0;
  fock_ms = b ? b->lg.fock_ms : 0.0;
  fock_builds = b ? b->lg.fock_builds : 0;
}

void batch_guess_timing(const Batch* b, double& features_ms, double& network_ms, long long& directed_pairs) {
  features_ms = b ? b->last_feature_ms : 0.0;
  network_ms = b ? b->last_network_ms : 0.0;
  directed_pairs = b ? b->ngp_total : 0;
}

double batch_kernel_ms(const Batch* b) { return b ? b->last_kernel_ms : 0.0; }
double batch_pair_ms(const Batch* b) { return b ? b->last_pair_ms : 0.0; }

bool batch_phase_ticks(const Batch* b, unsigned long long* out, std::string& err) {
  for (int k = 0; k < PH_SLOTS; ++k) out[k] = 0;
  if (!b || !b->args.prof) return true;
  HIP_OK(scopy(b->stream, out, b->args.prof, PH_SLOTS * sizeof(unsigned long long), hipMemcpyDeviceToHost));
  return true;
}

bool batch_fetch(Batch* bp, const Params& P, const std::vector<MolIn>& mols, std::vector<MolResult>& out, std::string& err) {
  Batch& b = *bp;
  out.assign(b.nmol, MolResult());
  for (int k = 0; k < b.nmol; ++k) {
    out[k].status = b.pre_status[k];
    out[k].missing_z = b.pre_z[k];
    out[k].required_mb = b.pre_req[k];
    out[k].limit_mb = b.pre_lim[k];
  }
  if (b.order_all.empty()) return true;
  std::vector<MolOut> ho(b.nmol);
  std::vector<double> hq(b.atoms_total), hg;
  if (b.want_gradient) {
    hg.resize(3 * b.atoms_total);
    HIP_OK(scopy(b.stream, hg.data(), b.args.out_g, sizeof(double) * hg.size(), hipMemcpyDeviceToHost));
  }
  HIP_OK(scopy(b.stream, ho.data(), b.args.out, sizeof(MolOut) * b.nmol, hipMemcpyDeviceToHost));
  HIP_OK(scopy(b.stream, hq.data(), b.args.out_q, sizeof(double) * b.atoms_total, hipMemcpyDeviceToHost));
  std::vector<double> hP, hC, hE, hS, hCb, hEb;
  if (b.mat_total > 0) {
    // densities stay on the device unless asked for (a gradient batch keeps its own there)
    if (b.keep_density) {
      hP.resize(b.mat_total);
      HIP_OK(scopy(b.stream, hP.data(), b.args.out_P, sizeof(double) * b.mat_total, hipMemcpyDeviceToHost));
      if (b.args.out_S) {
        hS.resize(b.mat_total);
        HIP_OK(scopy(b.stream, hS.data(), b.args.out_S, sizeof(double) * b.mat_total, hipMemcpyDeviceToHost));
      }
    }
    if (!b.purify) {
      hC.resize(b.mat_total);
      hE.resize(b.vec_total);
      HIP_OK(scopy(b.stream, hC.data(), b.args.out_C, sizeof(double) * b.mat_total, hipMemcpyDeviceToHost));
      HIP_OK(scopy(b.stream, hE.data(), b.args.out_eps, sizeof(double) * b.vec_total, hipMemcpyDeviceToHost));
      if (b.args.out_Cb) {
        hCb.resize(b.mat_total);
        hEb.resize(b.vec_total);
        HIP_OK(scopy(b.stream, hCb.data(), b.args.out_Cb, sizeof(double) * b.mat_total, hipMemcpyDeviceToHost));
        HIP_OK(scopy(b.stream, hEb.data(), b.args.out_epsb, sizeof(double) * b.vec_total, hipMemcpyDeviceToHost));
      }
    }
  }
  for (int k : b.order_all) {
    MolResult& r = out[k];
    const MolOut& o = ho[k];
    const MolDesc& d = b.desc[k];
    r.status = o.status;
    r.e_elec = o.e_elec;
    r.e_core = o.e_core;
    r.last_change = o.last_dp;
    r.improvement = std::max(o.first_e - o.e_elec, 0.0);
    r.homo = o.homo;
    r.lumo = o.lumo;
    r.iterations = o.iters;
    r.converged = o.converged;
    r.paths = o.paths;
    r.guess_used = o.guess_used;
    r.nao = d.nao;
    r.nocc = d.nocc;
    r.nbeta = d.nbeta;
    r.unrestricted = o.unrestricted;
    for (int x = 0; x < 3; ++x) r.dipole[x] = o.dip[x] * AU_DIPOLE_TO_DEBYE;
    r.charges.assign(d.nat, 0.0);
    for (int t = 0; t < d.nat; ++t) r.charges[b.atom_map[k][t]] = hq[d.atom0 + t];
    if (b.want_gradient) {
      r.gradient.assign(3 * (size_t)d.nat, 0.0);
      for (int t = 0; t < d.nat; ++t)
        for (int c = 0; c < 3; ++c) r.gradient[3 * b.atom_map[k][t] + c] = hg[3 * (d.atom0 + t) + c];
    }
    double eis = 0.0, ehs = 0.0;
    for (int i = 0; i < mols[k].n_atoms; ++i) {
      const int s = P.slot_of_z[mols[k].z[i]];
      eis += P.host[s].e_isol;
      ehs += P.host[s].eheat_ev;
    }
    r.hof_kcal = (o.e_elec + o.e_core - eis + ehs) * EV_TO_KCAL;
    if (d.out_mat >= 0) {
      const int n = d.nao;
      const std::vector<int>& mp = b.ao_map[k];
      if (b.keep_density) {
        r.density.assign((size_t)n * n, 0.0);
        const double* Ps = hP.data() + d.out_mat;
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) r.density[(size_t)mp[i] * n + mp[j]] = Ps[(size_t)i * n + j];
        if (r.unrestricted && !hS.empty()) {
          r.spin_density.assign((size_t)n * n, 0.0);
          const double* Ss = hS.data() + d.out_mat;
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) r.spin_density[(size_t)mp[i] * n + mp[j]] = Ss[(size_t)i * n + j];
        }
      }
      if (!b.purify) {
        r.mo_coeff.assign((size_t)n * n, 0.0);
        r.mo_energies.assign(hE.begin() + d.out_vec, hE.begin() + d.out_vec + n);
        const double* Cs = hC.data() + d.out_mat;
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) r.mo_coeff[(size_t)mp[i] * n + j] = Cs[(size_t)i * n + j];
        if (r.unrestricted && !hCb.empty()) {
          r.mo_coeff_beta.assign((size_t)n * n, 0.0);
          r.mo_energies_beta.assign(hEb.begin() + d.out_vec, hEb.begin() + d.out_vec + n);
          const double* Cb = hCb.data() + d.out_mat;
          for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) r.mo_coeff_beta[(size_t)mp[i] * n + j] = Cb[(size_t)i * n + j];
        }
      }
    }
  }
  return true;
}

void batch_destroy(Batch* b) { delete b; }

int batch_carry_densities(Batch* dp, const Batch* sp, const std::vector<int>& src_of, std::string& err) {
  Batch& d = *dp;
  const Batch& s = *sp;
  if ((int)src_of.size() != d.nmol) {
    err = "one source index per molecule of the batch";
    return 1;
  }
  if (!s.solved_once) return 0;  // nothing to carry yet
  for (int k = 0; k < d.nmol; ++k) {
    const int j = src_of[k];
    if (j < 0 || j >= s.nmol) {
      err = "source index out of range";
      return 1;
    }
    if (d.desc[k].nao != s.desc[j].nao) {
      err = "molecule " + std::to_string(k) + " differs from its source";
      return 1;
    }
  }
  auto copy = [&]() -> bool {
    for (int k = 0; k < d.nmol; ++k) {
      const MolDesc& dk = d.desc[k];
      const MolDesc& sj = s.desc[src_of[k]];
      if (dk.out_mat < 0 || sj.out_mat < 0 || dk.nat == 0) continue;
      HIP_OK(scopy(d.stream, d.args.out_P + dk.out_mat, s.args.out_P + sj.out_mat, sizeof(double) * dk.nao * dk.nao,
                       hipMemcpyDeviceToDevice));
      HIP_OK(scopy(d.stream, d.args.out + k, s.args.out + src_of[k], sizeof(MolOut), hipMemcpyDeviceToDevice));
    }
    return true;
  };
  if (!copy()) return 6;
  d.have_prev = true;
  return 0;
}

int batch_set_positions(Batch* bp, const std::vector<MolIn>& mols, std::string& err) {
  Batch& b = *bp;
  if ((int)mols.size() != b.nmol) {
    err = "molecule count differs from the batch's";
    return 1;
  }
  if (b.atoms_total == 0) return 0;
  std::vector<double> ax(b.atoms_total), ay(b.atoms_total), az(b.atoms_total);
  for (int k = 0; k < b.nmol; ++k) {
    const MolDesc& d = b.desc[k];
    if (d.nat == 0) continue;  // refused at creation: no atoms on the device
    const MolIn& m = mols[k];
    if (m.n_atoms != d.nat) {
      err = "molecule " + std::to_string(k) + ": atom count differs from the batch's";
      return 1;
    }
    for (int i = 0; i < m.n_atoms; ++i) {
      for (int c = 0; c < 3; ++c)
        if (!std::isfinite(m.xyz[3 * i + c])) {
          err = "molecule " + std::to_string(k) + ": positions must be finite";
          return 1;
        }
      for (int j = 0; j < i; ++j) {
        const double dx = m.xyz[3 * i] - m.xyz[3 * j], dy = m.xyz[3 * i + 1] - m.xyz[3 * j + 1],
                     dz = m.xyz[3 * i + 2] - m.xyz[3 * j + 2];
        if (!(dx * dx + dy * dy + dz * dz > 1.0e-16)) {
          err = "molecule " + std::to_string(k) + ": atoms " + std::to_string(j) + " and " + std::to_string(i) +
                " coincide";
          return 1;
        }
      }
    }
    const std::vector<int>& ord = b.atom_map[k];
    for (int t = 0; t < d.nat; ++t) {
      const int i = ord[t];
      ax[d.atom0 + t] = m.xyz[3 * i];
      ay[d.atom0 + t] = m.xyz[3 * i + 1];
      az[d.atom0 + t] = m.xyz[3 * i + 2];
    }
  }
  // the learned guess's neighbour lists follow the new positions: rebuilt on the host, the device
  // lists and the network buffers regrown (by a quarter at least) when the pair count outgrows them
  GuessPairs gp;
  if (b.learned) {
    const std::vector<MolDesc> saved = b.desc;
    build_guess_pairs(b, ax.data(), ay.data(), az.data(), gp);
    const long long ngp = (long long)gp.gsrc.size();
    if (ngp >= (1LL << 31) || (ngp > b.gp_cap && b.arena)) {
      b.desc = saved;
      err = ngp >= (1LL << 31) ? "too many directed atom pairs in one batch"
                               : "the new positions need more directed atom pairs than the allocation-free scratch holds";
      return ngp >= (1LL << 31) ? 1 : 4;
    }
    if (ngp > b.gp_cap) {
      int* src = const_cast<int*>(b.args.gsrc);
      int* dst = const_cast<int*>(b.args.gdst);
      int* part = const_cast<int*>(b.args.gpart);
      const long long cap = std::max(ngp, b.gp_cap + b.gp_cap / 4);
      b.args.gY = nullptr;
      if (!b.release(src, err) || !b.release(dst, err) || !b.release(part, err) || !b.release(b.d_X, err) ||
          !b.release(b.d_Y, err) || !b.alloc(src, (size_t)cap, err) || !b.alloc(dst, (size_t)cap, err) ||
          !b.alloc(part, (size_t)cap, err))
        return 6;
      b.args.gsrc = src;
      b.args.gdst = dst;
      b.args.gpart = part;
      b.gp_cap = cap;
      if (!b.ensure_guess_buffers(err)) return 6;
    }
    b.ngp_total = ngp;
  }
  const size_t bytes = sizeof(double) * (size_t)b.atoms_total;
  auto upload = [&]() -> bool {
    HIP_OK(scopy(b.stream, const_cast<double*>(b.args.ax), ax.data(), bytes, hipMemcpyHostToDevice));
    HIP_OK(scopy(b.stream, const_cast<double*>(b.args.ay), ay.data(), bytes, hipMemcpyHostToDevice));
    HIP_OK(scopy(b.stream, const_cast<double*>(b.args.az), az.data(), bytes, hipMemcpyHostToDevice));
    if (!b.learned) return true;
    const size_t ab = sizeof(int) * (size_t)b.atoms_total, pb = sizeof(int) * gp.gsrc.size();
    HIP_OK(scopy(b.stream, const_cast<int*>(b.args.gbeg), gp.gbeg.data(), ab, hipMemcpyHostToDevice));
    HIP_OK(scopy(b.stream, const_cast<int*>(b.args.gcnt), gp.gcnt.data(), ab, hipMemcpyHostToDevice));
    if (pb > 0) {
      HIP_OK(scopy(b.stream, const_cast<int*>(b.args.gsrc), gp.gsrc.data(), pb, hipMemcpyHostToDevice));
      HIP_OK(scopy(b.stream, const_cast<int*>(b.args.gdst), gp.gdst.data(), pb, hipMemcpyHostToDevice));
      HIP_OK(scopy(b.stream, const_cast<int*>(b.args.gpart), gp.gpart.data(), pb, hipMemcpyHostToDevice));
    }
    HIP_OK(scopy(b.stream, const_cast<MolDesc*>(b.args.mols), b.desc.data(), sizeof(MolDesc) * b.nmol,
                 hipMemcpyHostToDevice));
    return true;
  };
  return upload() ? 0 : 6;
}

int response_fock_run(const Params& P, const MolIn& mol, int nrhs, const double* dp, const double* pair_sw,
                      double* out, std::string& err) {
  std::vector<MolIn> mols{mol};
  EngineOptions opt{1, 0.0, 0.0, 0, 0.1, 1, 0, 512};
  Batch* b = nullptr;
  if (!batch_create(P, mols, opt, RunRequest(), b, err)) {
    batch_destroy(b);
    return 6;
  }
  if (b->pre_status[0]) {
    const int st = b->pre_status[0];
    err = st == 2 || st == 4 ? "no PM3 parameters in this engine for element Z = " + std::to_string(b->pre_z[0])
                             : "the molecule was refused (status " + std::to_string(st) + ")";
    batch_destroy(b);
    return st;
  }
  const MolDesc d = b->desc[0];
  const int n = d.nao, np = d.np, nat = d.nat, nh = d.nheavy, nH = nat - nh;
  const long long m2 = (long long)np * np, nn = (long long)n * n;
  const long long npairs = (long long)nat * (nat - 1) / 2;
  const long long nxx = (long long)nh * (nh - 1) / 2, nxh = (long long)nh * nH;
  const long long wcap = std::max(20 * npairs, 1LL);
  const std::vector<int>& mp = b->ao_map[0];
  const std::vector<int>& am = b->atom_map[0];
  // the per-pair switch from the reference's pair order -- (u, v), u < v over the caller's atoms,
  // row by row (src/hamiltonian.rs:237-240) -- to the engine's heavy-heavy, heavy-hydrogen and
  // hydrogen-hydrogen lists
  std::vector<double> hsw;
  if (pair_sw) {
    auto ref = [&](int u, int v) {
      if (u > v) std::swap(u, v);
      return (long long)u * (2LL * nat - u - 1) / 2 + (v - u - 1);
    };
    hsw.reserve((size_t)std::max(npairs, 1LL));
    for (int a = 0; a < nh; ++a)
      for (int c = a + 1; c < nh; ++c) hsw.push_back(pair_sw[ref(am[a], am[c])]);
    for (int a = 0; a < nh; ++a)
      for (int h = 0; h < nH; ++h) hsw.push_back(pair_sw[ref(am[a], am[nh + h])]);
    for (int g = 0; g < nH; ++g)
      for (int h = g + 1; h < nH; ++h) hsw.push_back(pair_sw[ref(am[nh + g], am[nh + h])]);
    if (hsw.empty()) hsw.push_back(0.0);
  }
  // right-hand sides per launch: gridDim.y, and the Coulomb partials kept within 1 GiB
  const int chunk = (int)std::max(1LL, std::min<long long>({(long long)nrhs, 65535LL, (1LL << 27) / wcap}));
  double *d_in = nullptr, *d_out = nullptr, *d_hz = nullptr, *d_slots = nullptr, *d_J = nullptr, *d_sw = nullptr;
  bool ok = b->alloc(d_in, (size_t)(chunk * m2), err) && b->alloc(d_out, (size_t)(chunk * m2), err) &&
            b->alloc(d_hz, (size_t)m2, err) && b->alloc(d_slots, (size_t)(chunk * wcap), err) &&
            b->alloc(d_J, (size_t)chunk * 10 * std::max(nat, 1), err) &&
            (!pair_sw || b->alloc(d_sw, hsw.size(), err));
  auto dev = [&](hipError_t e) {
    if (e != hipSuccess && ok) {
      err = hipGetErrorString(e);
      ok = false;
    }
    return ok;
  };
  if (ok) dev(smemset(b->stream, d_hz, 0, sizeof(double) * (size_t)m2));
  if (ok && pair_sw) dev(scopy(b->stream, d_sw, hsw.data(), sizeof(double) * hsw.size(), hipMemcpyHostToDevice));
  if (ok) {
    hipLaunchKernelGGL(pair_kernel, dim3(b->pair_blocks), dim3(b->pair_block), 0, b->stream, b->args);
    dev(hipGetLastError());
  }
  const int gp = (int)std::max(1LL, std::min(64LL, (std::max({nxx, nxh, npairs}) + LG_THREADS - 1) / LG_THREADS));
  const int gj = std::max(1, std::min(64, (nh * 10 + nH + 3) / 4));
  const int gd = std::max(1, (nat * 16 + LG_THREADS - 1) / LG_THREADS);
  std::vector<double> hin, hout;
  for (int r0 = 0; ok && r0 < nrhs; r0 += chunk) {
    const int nr = std::min(chunk, nrhs - r0);
    hin.assign((size_t)(nr * m2), 0.0);
    for (int r = 0; r < nr; ++r)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) hin[r * m2 + (long long)i * np + j] = dp[(r0 + r) * nn + (long long)mp[i] * n + mp[j]];
    if (!dev(scopy(b->stream, d_in, hin.data(), sizeof(double) * hin.size(), hipMemcpyHostToDevice))) break;
    if (!dev(smemset(b->stream, d_out, 0, sizeof(double) * hin.size()))) break;
    if (pair_sw) {
      hipLaunchKernelGGL(rf_pairs_kernel<true>, dim3(gp, nr), dim3(LG_THREADS), 0, b->stream, b->args, d, d_in, d_out,
                         d_hz, d_sw, d_slots, m2, wcap);
      hipLaunchKernelGGL(rf_jsum_kernel<true>, dim3(gj, nr), dim3(LG_THREADS), 0, b->stream, b->args, d, d_in, d_J, d_sw,
                         d_slots, m2, wcap);
    } else {
      hipLaunchKernelGGL(rf_pairs_kernel<false>, dim3(gp, nr), dim3(LG_THREADS), 0, b->stream, b->args, d, d_in, d_out,
                         d_hz, d_sw, d_slots, m2, wcap);
      hipLaunchKernelGGL(rf_jsum_kernel<false>, dim3(gj, nr), dim3(LG_THREADS), 0, b->stream, b->args, d, d_in, d_J,
                         d_sw, d_slots, m2, wcap);
    }
    hipLaunchKernelGGL(rf_diag_kernel, dim3(gd, nr), dim3(LG_THREADS), 0, b->stream, b->args, d, d_in, d_out, d_hz, d_J,
                       m2);
    if (!dev(hipGetLastError())) break;
    hout.resize(hin.size());
    if (!dev(scopy(b->stream, hout.data(), d_out, sizeof(double) * hout.size(), hipMemcpyDeviceToHost))) break;
    for (int r = 0; r < nr; ++r)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) out[(r0 + r) * nn + (long long)mp[i] * n + mp[j]] = hout[r * m2 + (long long)i * np + j];
  }
  batch_destroy(b);
  return ok ? 0 : 6;
}

bool stage_run(const Params& P, const MolIn& mol, int mode, const std::vector<double>& in, std::vector<double>& out,
               int& status, std::string& err, const GuessModel* model) {
  std::vector<MolIn> mols{mol};
  EngineOptions opt{1, 0.0, 0.0, 0, 0.1, 1, 0, 512};
  if (mode == 5) {
    if (!model) {
      err = "stage 5 needs a guess model";
      return false;
    }
    opt.guess = 4;
    opt.guess_model = model;
  }
  Batch* b = nullptr;
  status = 0;
  if (!batch_create(P, mols, opt, RunRequest(), b, err)) {
    batch_destroy(b);
    return false;
  }
  if (b->nsolve == 0) {
    status = b->pre_status[0] ? b->pre_status[0] : 5;  // stage hooks run on one workgroup only
    batch_destroy(b);
    return true;
  }
  if (mode == 5 && !b->desc[0].learn) {
    status = 4;  // the learned guess covers H, C and O only
    batch_destroy(b);
    return true;
  }
  const MolDesc& d = b->desc[0];
  const int n = d.nao;
  const std::vector<int>& mp = b->ao_map[0];
  // modes 0, 1, 5 and 6 speak the caller's AO order; the others take generic matrices as given
  std::vector<double> sin(in.size());
  if (mode == 1 || mode == 6) {
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) sin[(size_t)i * n + j] = in[(size_t)mp[i] * n + mp[j]];
  } else {
    sin = in;
  }
  const size_t nout = (size_t)n * n + n + 3;
  double *d_in = nullptr, *d_out = nullptr;
  bool ok = b->alloc(d_in, std::max<size_t>(sin.size(), 1), err) && b->alloc(d_out, nout, err);
  if (ok && !sin.empty()) {
    hipError_t e = scopy(b->stream, d_in, sin.data(), sizeof(double) * sin.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { err = hipGetErrorString(e); ok = false; }
  }
  if (ok) {
    hipLaunchKernelGGL(pair_kernel, dim3(b->pair_blocks), dim3(b->pair_block), 0, b->stream, b->args);
    if (mode == 5 && !guess_infer(model, infer_view(*b), err, b->stream)) ok = false;
  }
  if (ok) {
    hipLaunchKernelGGL(stage_kernel, dim3(1), dim3(NT), b->lds_need, b->stream, b->args, mode, d_in, d_out);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    std::vector<double> so(nout);
    if (e == hipSuccess) e = scopy(b->stream, so.data(), d_out, sizeof(double) * nout, hipMemcpyDeviceToHost);
    if (e != hipSuccess) { err = hipGetErrorString(e); ok = false; }
    if (ok) {
      if (mode == 0 || mode == 1 || mode == 5 || mode == 6) {
        const int extra = mode == 0 ? 1 : mode == 5 ? 2 : mode == 6 ? 3 : 0;
        out.assign((size_t)n * n + extra, 0.0);
        for (int i = 0; i < n; ++i)
          for (int j = 0; j < n; ++j) out[(size_t)mp[i] * n + mp[j]] = so[(size_t)i * n + j];
        for (int k = 0; k < extra; ++k) out[(size_t)n * n + k] = so[(size_t)n * n + k];
      } else {
        out.assign(so.begin(), so.begin() + (size_t)n * n + (mode == 3 ? 0 : n));
      }
    }
  }
  batch_destroy(b);
  return ok;
}

// ------------------------------------------------------------------------------------------
// Allocation-free single points.  The device block below is the scratch of every buffer such a
// call needs (inputs staged from the host, molecule store, workgroup slab, outputs).  Host code
// never passes `scratch` itself to a kernel or a copy: the symbol is a host-side stand-in, and
// the block's device address on each GPU comes from hipGetSymbolAddress, taken once per device
// with that device current.
constexpr size_t NOALLOC_BYTES = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

struct NoallocDevice {
  bool ready = false;
  unsigned char* ptr = nullptr;  // device address of this GPU's copy of `scratch`
  hipStream_t stream = nullptr;  // the stream of this GPU's allocation-free calls
};

static bool noalloc_device(NoallocDevice& out, std::string& err) {
  static std::mutex mu;
  static NoallocDevice dev_tab[64];
  int dev = 0;
  HIP_OK(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) {
    err = "device index out of range";
    return false;
  }
  std::lock_guard<std::mutex> lock(mu);
  NoallocDevice& d = dev_tab[dev];
  if (!d.ready) {
    void* p = nullptr;
    HIP_OK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    HIP_OK(hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking));
    d.ptr = static_cast<unsigned char*>(p);
    d.ready = true;
  }
  out = d;
  return true;
}

int engine_run_noalloc(const Params& P, const MolIn& mol, const EngineOptions& opt, bool want_density, MolResult& out,
                       std::string& err) {
  NoallocDevice d;
  if (!noalloc_device(d, err)) return 6;
  RunRequest req;
  req.want_density = want_density;
  Batch* b = nullptr;
  const std::vector<MolIn> mols{mol};
  if (!batch_create_in(P, mols, opt, req, b, err, d.ptr, NOALLOC_BYTES, d.stream)) {
    const bool short_block = b && b->arena && err.find("scratch block") != std::string::npos;
    delete b;
    return short_block ? 5 : 6;
  }
  std::vector<MolResult> res;
  const bool ok = batch_solve(b, err) && batch_fetch(b, P, mols, res, err);
  delete b;
  if (!ok) return 6;
  out = res[0];
  return 0;
}

bool engine_run(const Params& P, const std::vector<MolIn>& mols, const EngineOptions& opt, const RunRequest& req,
                std::vector<MolResult>& out, std::string& err) {
  Batch* b = nullptr;
  bool ok = batch_create(P, mols, opt, req, b, err) && batch_solve(b, err) && batch_fetch(b, P, mols, out, err);
  batch_destroy(b);
  return ok;
}

// divide-and-conquer SCF (run_dc) over the host-driven path
#include "pm3_dc.hpp"

// One step of the tile kernel on a host matrix (kernel-level hook): y = x x (square) or 2 x - x x,
// with tr y and |y|_F^2 summed from the kernel's partials.  The device copy of x has NaN in its
// padding and y starts as NaN, so a read beyond n or an element left unwritten shows.  With
// reps > 0 the same step is launched reps more times between two events (ms per launch).
// workgroups > 0 caps the grid (rounded down to a multiple of 8, at least 8), so that
// workgroups take several tiles each as they do beyond about 10 000 AOs.
int sym_step_stage(int n, int square, int slab, const double* x, double* y, double* trace, double* norm2, int reps,
                   int workgroups, double* ms, std::string& err) {
  if (n < 1 || n > 16384 || (slab != 8 && slab != 16 && slab != 32) || reps < 0 || workgroups < 0) {
    err = "n must be in [1, 16384], slab 8, 16 or 32, reps >= 0, workgroups >= 0";
    return 2;
  }
  const int np = (n + 15) & ~15, chunk = (sy_tiles(n) + 7) / 8;
  const int grid = workgroups > 0 ? std::min(8 * chunk, std::max(8, workgroups & ~7)) : 8 * chunk;
  const size_t m2 = (size_t)np * np;
  double *dx = nullptr, *dy = nullptr, *dred = nullptr, *dst = nullptr;
  int* dmap = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto done = [&](bool ok) {
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dred);
    (void)hipFree(dst);
    (void)hipFree(dmap);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    return ok ? 0 : 6;
  };
  const std::vector<int> map = sy_map_host(n);
  double st[SP2S_N] = {0.0};
  st[SP2S_SQ] = square ? 1.0 : 0.0;
#define SY_OK(x)                                           \
  do {                                                     \
    hipError_t e_ = (x);                                   \
    if (e_ != hipSuccess) {                                \
      err = std::string(#x) + ": " + hipGetErrorString(e_); \
      return done(false);                                  \
    }                                                      \
  } while (0)
  SY_OK(hipMalloc(&dx, m2 * sizeof(double)));
  SY_OK(hipMalloc(&dy, m2 * sizeof(double)));
  SY_OK(hipMalloc(&dred, 2 * (size_t)grid * sizeof(double)));
  SY_OK(hipMalloc(&dst, SP2S_N * sizeof(double)));
  SY_OK(hipMalloc(&dmap, map.size() * sizeof(int)));
  SY_OK(hipMemset(dx, 0xff, m2 * sizeof(double)));
  SY_OK(hipMemset(dy, 0xff, m2 * sizeof(double)));
  SY_OK(hipMemcpy2D(dx, np * sizeof(double), x, n * sizeof(double), n * sizeof(double), n, hipMemcpyHostToDevice));
  SY_OK(hipMemcpy(dst, st, sizeof(st), hipMemcpyHostToDevice));
  SY_OK(hipMemcpy(dmap, map.data(), map.size() * sizeof(int), hipMemcpyHostToDevice));
  sy_launch(slab, grid, nullptr, dx, dy, n, np, dst, dred, dmap, chunk);
  SY_OK(hipGetLastError());
  SY_OK(hipDeviceSynchronize());
  std::vector<double> red(2 * (size_t)grid);
  SY_OK(hipMemcpy2D(y, n * sizeof(double), dy, np * sizeof(double), n * sizeof(double), n, hipMemcpyDeviceToHost));
  SY_OK(hipMemcpy(red.data(), dred, red.size() * sizeof(double), hipMemcpyDeviceToHost));
  double a = 0.0, c = 0.0;
  for (int w = 0; w < grid; ++w) {
    a += red[2 * w];
    c += red[2 * w + 1];
  }
  if (trace) *trace = a;
  if (norm2) *norm2 = c;
  if (ms) *ms = 0.0;
  if (reps > 0) {
    SY_OK(hipEventCreate(&e0));
    SY_OK(hipEventCreate(&e1));
    SY_OK(hipEventRecord(e0, nullptr));
    for (int r = 0; r < reps; ++r) sy_launch(slab, grid, nullptr, dx, dy, n, np, dst, dred, dmap, chunk);
    SY_OK(hipEventRecord(e1, nullptr));
    SY_OK(hipGetLastError());
    SY_OK(hipEventSynchronize(e1));
    float t = 0.0f;
    SY_OK(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = (double)t / reps;
  }
#undef SY_OK
  return done(true);
}

}  // namespace pm3host
#endif  // PM3_KERNEL_VARIANT
