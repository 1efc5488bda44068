// Device-side NDDO pair physics for PM3 on CDNA4: two-centre two-electron integrals and their
// frame rotation, electron-core attraction, Slater overlap, and core-core repulsion.
//
// Everything is templated on the scalar: `double` for energies, `D3` (value + derivatives with
// respect to the three components of the pair displacement) for the analytic gradient.  The
// element data sit in constant memory (ElemConst, filled once per process by the host).
//
// Formulas follow the reference kernels:
//   rotation to +x .............. src/integrals.rs:25-59
//   local XH / XX multipoles .... src/integrals.rs:289-464
//   packed W assembly ........... src/integrals.rs:466-552
//   electron-core attraction .... src/integrals.rs:234-271
//   overlap A/B integrals ....... src/overlap.rs:763-850
//   overlap local sigma/pi ...... src/overlap.rs:266-475 (closed-form jcall 2,3,4,431,5,6)
//   overlap rotation ............ src/overlap.rs:480-507
//   core-core ................... src/repulsion.rs:57-120
#pragma once
#include <hip/hip_runtime.h>

namespace pm3dev {

constexpr double EV = 27.211386245988;   // Hartree in eV (src/constants.rs:21)
constexpr double A0 = 0.529177210903;    // Bohr in Angstrom (src/constants.rs:24)
constexpr int MAX_SLOTS = 16;

// Element data as the kernels read it (one row per supported element slot).
struct ElemConst {
  double uss, upp, betas, betap;
  double gss, gsp, gpp, gp2, hsp, alp;
  double dd, qq, rho0, rho1, rho2;          // electron-electron multipoles (Bohr)
  double po1, po2, po3, po9, ddp2, ddp3;    // electron-core multipoles (Bohr)
  double zs, zp, tore, ios, iop, mass;
  double gk[2], gl[2], gm[2];
  int z, n_s, norb, ngauss;
};

struct TableConst {
  ElemConst el[MAX_SLOTS];
  double alpb[MAX_SLOTS * MAX_SLOTS];
  double xfac[MAX_SLOTS * MAX_SLOTS];  // 0 where the pair has no diatomic parameters
};

// ------------------------------------------------------------------------------------------
// Dual number over the three displacement components.
struct D3 {
  double v, d0, d1, d2;
};
__host__ __device__ inline D3 mk(double v, double a, double b, double c) { return D3{v, a, b, c}; }
__host__ __device__ inline D3 operator+(D3 a, D3 b) { return mk(a.v + b.v, a.d0 + b.d0, a.d1 + b.d1, a.d2 + b.d2); }
__host__ __device__ inline D3 operator-(D3 a, D3 b) { return mk(a.v - b.v, a.d0 - b.d0, a.d1 - b.d1, a.d2 - b.d2); }
__host__ __device__ inline D3 operator*(D3 a, D3 b) {
  return mk(a.v * b.v, a.d0 * b.v + a.v * b.d0, a.d1 * b.v + a.v * b.d1, a.d2 * b.v + a.v * b.d2);
}
__host__ __device__ inline D3 operator-(D3 a) { return mk(-a.v, -a.d0, -a.d1, -a.d2); }
__host__ __device__ inline D3 operator+(D3 a, double b) { return mk(a.v + b, a.d0, a.d1, a.d2); }
__host__ __device__ inline D3 operator-(D3 a, double b) { return mk(a.v - b, a.d0, a.d1, a.d2); }
__host__ __device__ inline D3 operator*(D3 a, double b) { return mk(a.v * b, a.d0 * b, a.d1 * b, a.d2 * b); }
__host__ __device__ inline D3 operator/(D3 a, double b) { return a * (1.0 / b); }
__host__ __device__ inline D3 chain(D3 a, double v, double f) { return mk(v, a.d0 * f, a.d1 * f, a.d2 * f); }

__device__ inline double val(double x) { return x; }
__device__ inline double val(D3 x) { return x.v; }
__device__ inline double cst(double x, double) { return x; }
__device__ inline D3 cst(double x, D3) { return mk(x, 0, 0, 0); }
__device__ inline double rsq(double x) { return 1.0 / sqrt(x); }  // 1/sqrt(x), as 1/(sqrt x)
__device__ inline D3 rsq(D3 x) {
  double s = sqrt(x.v);
  double r = 1.0 / s;
  return chain(x, r, -0.5 * r / x.v);
}
__device__ inline double sqrtx(double x) { return sqrt(x); }
__device__ inline D3 sqrtx(D3 x) { double s = sqrt(x.v); return chain(x, s, 0.5 / s); }
__device__ inline double recip(double x) { return 1.0 / x; }
__device__ inline D3 recip(D3 x) { double r = 1.0 / x.v; return chain(x, r, -r * r); }
__device__ inline double expx(double x) { return exp(x); }
__device__ inline D3 expx(D3 x) { double e = exp(x.v); return chain(x, e, e); }
__device__ inline double powfx(double x, double p) { return pow(x, p); }
__device__ inline D3 powfx(D3 x, double p) { double v = pow(x.v, p); return chain(x, v, p * pow(x.v, p - 1.0)); }
__device__ inline double absx(double x) { return fabs(x); }
__device__ inline D3 absx(D3 x) { return x.v < 0.0 ? -x : x; }
template <int N, class S> __device__ inline S powi(S x) {
  S r = x;
#pragma unroll
  for (int i = 1; i < N; ++i) r = r * x;
  return r;
}

__device__ inline int pk(int a, int b) { return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a; }

// Rotation taking unit vector v to +x (rows r0, r1, r2).
template <class S> __device__ inline void rot_to_x(S vx, S vy, S vz, S r[3][3]) {
  if (val(vx) < 0.0) {
    S one = cst(1.0, vx);
    S ux = vx - 1.0;
    S f = recip(ux * ux + vy * vy + vz * vz) * (-2.0);
    r[0][0] = one + f * ux * ux; r[0][1] = f * ux * vy;       r[0][2] = f * ux * vz;
    r[1][0] = f * vy * ux;       r[1][1] = one + f * vy * vy; r[1][2] = f * vy * vz;
    r[2][0] = f * vz * ux;       r[2][1] = f * vz * vy;       r[2][2] = one + f * vz * vz;
  } else {
    S qy = vz, qz = -vy, qw = vx + 1.0;
    S inv = recip(sqrtx(qy * qy + qz * qz + qw * qw));
    qy = qy * inv; qz = qz * inv; qw = qw * inv;
    r[0][0] = (qy * qy + qz * qz) * (-2.0) + 1.0; r[0][1] = qz * qw * (-2.0);       r[0][2] = qy * qw * 2.0;
    r[1][0] = qz * qw * 2.0;                      r[1][1] = qz * qz * (-2.0) + 1.0; r[1][2] = qy * qz * 2.0;
    r[2][0] = qy * qw * (-2.0);                   r[2][1] = qy * qz * 2.0;          r[2][2] = qy * qy * (-2.0) + 1.0;
  }
}

// Local (bond-frame) integrals of a heavy centre against an s-only partner.
template <class S>
__device__ inline void local_xh(double dd, double qq, double r0a, double r1a, double r2a, double r0b, S r, S ri[4]) {
  const double ev1 = EV / 2.0, ev2 = EV / 4.0;
  double qa = qq * 2.0;
  double aee = (r0a + r0b) * (r0a + r0b), ade = (r1a + r0b) * (r1a + r0b), aqe = (r2a + r0b) * (r2a + r0b);
  S r2 = r * r;
  S ee = rsq(r2 + aee) * EV;
  S e6 = rsq(r2 + aqe) * ev1;
  S rpd = r + dd, rmd = r - dd, rpq = r + qa, rmq = r - qa;
  ri[0] = ee;
  ri[1] = rsq(rpd * rpd + ade) * ev1 - rsq(rmd * rmd + ade) * ev1;
  ri[2] = ee + rsq(rpq * rpq + aqe) * ev2 + rsq(rmq * rmq + aqe) * ev2 - e6;
  ri[3] = ee + rsq(r2 + qa * qa + aqe) * ev1 - e6;
}

// Local integrals of two heavy centres (22 unique values in the bond frame).
template <class S> __device__ inline void local_xx(const ElemConst& A, const ElemConst& B, S r, S ri[22]) {
  const double ev1 = EV / 2.0, ev2 = EV / 4.0, ev3 = EV / 8.0, ev4 = EV / 16.0;
  const double da = A.dd, db = B.dd, qa = A.qq * 2.0, qb = B.qq * 2.0, qa1 = A.qq, qb1 = B.qq;
  auto sq = [](double x) { return x * x; };
  const double aee = sq(A.rho0 + B.rho0), ade = sq(A.rho1 + B.rho0), aqe = sq(A.rho2 + B.rho0);
  const double aed = sq(A.rho0 + B.rho1), aeq = sq(A.rho0 + B.rho2), axx = sq(A.rho1 + B.rho1);
  const double adq = sq(A.rho1 + B.rho2), aqd = sq(A.rho2 + B.rho1), aqq = sq(A.rho2 + B.rho2);
  // g(x, c) = 1/sqrt(x + c);  h(a, c) = 1/sqrt(a^2 + c)
  auto g = [](S x, double c) { return rsq(x + c); };
  auto h = [](S a, double c) { return rsq(a * a + c); };
  S r2 = r * r;
  S ee = rsq(r2 + aee) * EV;
  S dze = h(r - da, ade) * ev1 - h(r + da, ade) * ev1;
  S e6 = rsq(r2 + aqe) * ev1;
  S qzze = h(r - qa, aqe) * ev2 + h(r + qa, aqe) * ev2 - e6;
  S qxxe = g(r2 + qa * qa, aqe) * ev1 - e6;
  S edz = h(r + db, aed) * ev1 - h(r - db, aed) * ev1;
  S e12 = rsq(r2 + aeq) * ev1;
  S eqzz = h(r - qb, aeq) * ev2 + h(r + qb, aeq) * ev2 - e12;
  S eqxx = g(r2 + qb * qb, aeq) * ev1 - e12;
  S t20 = h(r + da, adq) * ev2, t22 = h(r - da, adq) * ev2;
  S t24 = h(r - db, aqd) * ev2, t26 = h(r + db, aqd) * ev2;
  S t36 = rsq(r2 + aqq) * ev2;
  S t39 = g(r2 + qa * qa, aqq) * ev2, t40 = g(r2 + qb * qb, aqq) * ev2;
  S t42 = h(r - qb, aqq) * ev3, t44 = h(r + qb, aqq) * ev3;
  S t46 = h(r + qa, aqq) * ev3, t48 = h(r - qa, aqq) * ev3;
  ri[0] = ee;
  ri[1] = -dze;
  ri[2] = ee + qzze;
  ri[3] = ee + qxxe;
  ri[4] = -edz;
  ri[5] = h(r + da - db, axx) * ev2 + h(r - da + db, axx) * ev2 - h(r - da - db, axx) * ev2 - h(r + da + db, axx) * ev2;
  ri[6] = g(r2 + (da - db) * (da - db), axx) * ev1 - g(r2 + (da + db) * (da + db), axx) * ev1;
  ri[7] = -edz + h(r + qa - db, aqd) * ev3 - h(r + qa + db, aqd) * ev3 + h(r - qa - db, aqd) * ev3 -
          h(r - qa + db, aqd) * ev3 - t24 + t26;
  S rmdb = r - db, rpdb = r + db, rpda = r + da, rmda = r - da;
  ri[8] = -edz - t24 + g(rmdb * rmdb + qa * qa, aqd) * ev2 + t26 - g(rpdb * rpdb + qa * qa, aqd) * ev2;
  {
    double m1 = (qa1 - db) * (qa1 - db), p1 = (qa1 + db) * (qa1 + db);
    S rpq = r + qa1, rmq = r - qa1;
    ri[9] = g(rpq * rpq + m1, aqd) * ev2 - g(rmq * rmq + m1, aqd) * ev2 - g(rpq * rpq + p1, aqd) * ev2 +
            g(rmq * rmq + p1, aqd) * ev2;
  }
  ri[10] = ee + eqzz;
  ri[11] = ee + eqxx;
  ri[12] = -dze + h(r + da - qb, adq) * ev3 - h(r - da - qb, adq) * ev3 + h(r + da + qb, adq) * ev3 -
           h(r - da + qb, adq) * ev3 + t22 - t20;
  ri[13] = -dze - t20 + g(rpda * rpda + qb * qb, adq) * ev2 + t22 - g(rmda * rmda + qb * qb, adq) * ev2;
  {
    double m1 = (da - qb1) * (da - qb1), p1 = (da + qb1) * (da + qb1);
    S rmq = r - qb1, rpq = r + qb1;
    ri[14] = g(rmq * rmq + m1, adq) * ev2 - g(rpq * rpq + m1, adq) * ev2 - g(rmq * rmq + p1, adq) * ev2 +
             g(rpq * rpq + p1, adq) * ev2;
  }
  ri[15] = ee + eqzz + qzze + h(r + qa - qb, aqq) * ev4 + h(r + qa + qb, aqq) * ev4 + h(r - qa - qb, aqq) * ev4 +
           h(r - qa + qb, aqq) * ev4 - t48 - t46 - t42 - t44 + t36;
  S rmqb = r - qb, rpqb = r + qb, rpqa = r + qa, rmqa = r - qa;
  ri[16] = ee + eqzz + qxxe + g(rmqb * rmqb + qa * qa, aqq) * ev3 + g(rpqb * rpqb + qa * qa, aqq) * ev3 - t42 - t44 -
           t39 + t36;
  ri[17] = ee + eqxx + qzze + g(rpqa * rpqa + qb * qb, aqq) * ev3 + g(rmqa * rmqa + qb * qb, aqq) * ev3 - t46 - t48 -
           t40 + t36;
  S qxxqxx = g(r2 + (qa - qb) * (qa - qb), aqq) * ev3 + g(r2 + (qa + qb) * (qa + qb), aqq) * ev3 - t39 - t40 + t36;
  ri[18] = ee + eqxx + qxxe + qxxqxx;
  {
    double dm = (qa1 - qb1) * (qa1 - qb1), dp = (qa1 + qb1) * (qa1 + qb1);
    S u1 = r + qa1 - qb1, u2 = r + qa1 + qb1, u3 = r - qa1 - qb1, u4 = r - qa1 + qb1;
    S u11 = u1 * u1, u22 = u2 * u2, u33 = u3 * u3, u44 = u4 * u4;
    ri[19] = g(u11 + dm, aqq) * ev3 - g(u22 + dm, aqq) * ev3 - g(u33 + dm, aqq) * ev3 + g(u44 + dm, aqq) * ev3 -
             g(u11 + dp, aqq) * ev3 + g(u22 + dp, aqq) * ev3 + g(u33 + dp, aqq) * ev3 - g(u44 + dp, aqq) * ev3;
  }
  S qxxqyy = g(r2 + qa * qa + qb * qb, aqq) * ev2 - t39 - t40 + t36;
  ri[20] = ee + eqxx + qxxe + qxxqyy;
  ri[21] = (qxxqxx - qxxqyy) * 0.5;
}

// Packed (mu nu | s s) of a heavy/s-only pair in the molecular frame.
template <class S> __device__ inline void assemble_xh(const S ri[4], const S R[3][3], S w[10]) {
  w[0] = ri[0];
#pragma unroll
  for (int k = 0; k < 3; ++k) w[pk(k + 1, 0)] = ri[1] * R[0][k];
#pragma unroll
  for (int k = 0; k < 3; ++k)
#pragma unroll
    for (int l = 0; l <= k; ++l) w[pk(k + 1, l + 1)] = ri[2] * (R[0][k] * R[0][l]) + ri[3] * (R[1][k] * R[1][l] + R[2][k] * R[2][l]);
}

// Packed 10 x 10 (mu nu | lambda sigma) of a heavy/heavy pair in the molecular frame, written
// with a caller-supplied stride so the same code fills registers or a strided global table.
template <class S, class Store> __device__ inline void assemble_xx(const S ri[22], const S R[3][3], Store store) {
  const S* r0 = R[0];
  const S* r1 = R[1];
  const S* r2 = R[2];
  int idx = 0;
#pragma unroll
  for (int kk = 0; kk < 4; ++kk)
#pragma unroll
    for (int ll = 0; ll <= kk; ++ll)
#pragma unroll
      for (int mm = 0; mm < 4; ++mm)
#pragma unroll
        for (int nn = 0; nn <= mm; ++nn) {
          const int k = kk - 1, l = ll - 1, m = mm - 1, n = nn - 1;
          S v;
          if (kk == 0) {
            if (mm == 0) v = ri[0];
            else if (nn == 0) v = ri[4] * r0[m];
            else v = ri[10] * (r0[m] * r0[n]) + ri[11] * (r1[m] * r1[n] + r2[m] * r2[n]);
          } else if (ll == 0) {
            if (mm == 0) v = ri[1] * r0[k];
            else if (nn == 0) v = ri[5] * (r0[k] * r0[m]) + ri[6] * (r1[k] * r1[m] + r2[k] * r2[m]);
            else {
              S t0 = r0[k] * r0[m] * r0[n];
              S t1 = (r1[m] * r1[n] + r2[m] * r2[n]) * r0[k];
              S mix = r1[k] * (r1[n] * r0[m] + r1[m] * r0[n]) + r2[k] * (r2[m] * r0[n] + r2[n] * r0[m]);
              v = ri[12] * t0 + ri[13] * t1 + ri[14] * mix;
            }
          } else if (mm == 0) {
            v = ri[2] * (r0[k] * r0[l]) + ri[3] * (r1[k] * r1[l] + r2[k] * r2[l]);
          } else if (nn == 0) {
            S t0 = r0[k] * r0[l] * r0[m];
            S t1 = (r1[k] * r1[l] + r2[k] * r2[l]) * r0[m];
            S t2 = r1[l] * r1[m] + r2[l] * r2[m];
            v = ri[7] * t0 + ri[8] * t1 + ri[9] * (r0[k] * t2 + r0[l] * (r1[k] * r1[m] + r2[k] * r2[m]));
          } else {
            S acc = ri[15] * (r0[k] * r0[l] * r0[m] * r0[n]);
            acc = acc + ri[16] * ((r1[k] * r1[l] + r2[k] * r2[l]) * r0[m] * r0[n]);
            acc = acc + ri[17] * ((r1[m] * r1[n] + r2[m] * r2[n]) * (r0[k] * r0[l]));
            acc = acc + ri[18] * (r1[k] * r1[l] * r1[m] * r1[n] + r2[k] * r2[l] * r2[m] * r2[n]);
            S mix1 = r0[m] * (r1[l] * r1[n] + r2[l] * r2[n]);
            S mix2 = r0[n] * (r1[l] * r1[m] + r2[l] * r2[m]);
            S v5 = r0[k] * (mix1 + mix2) + r0[l] * (r0[m] * (r1[k] * r1[n] + r2[k] * r2[n]) + r0[n] * (r1[k] * r1[m] + r2[k] * r2[m]));
            acc = acc + ri[19] * v5;
            acc = acc + ri[20] * (r1[k] * r1[l] * r2[m] * r2[n] + r2[k] * r2[l] * r1[m] * r1[n]);
            acc = acc + ri[21] * ((r1[k] * r2[l] + r2[k] * r1[l]) * (r1[m] * r2[n] + r2[m] * r1[n]));
            v = acc;
          }
          store(idx, v);
          ++idx;
        }
}

// Attraction of the AOs on `orb` to the core of `core`; d points from orb to core.  blk is the
// packed 10-vector (or 1 for an s-only centre) of the 4x4 symmetric block.
template <class S> __device__ inline void ecore(const ElemConst& orb, const ElemConst& core, S dx, S dy, S dz, S blk[10]) {
  S r2v = dx * dx + dy * dy + dz * dz;
  S r = sqrtx(r2v);
  const double crho = core.po9;
  const double s = orb.po1 + crho;
  blk[0] = recip(sqrtx(r * r + s * s)) * (-core.tore * EV);
  if (orb.z < 3 || orb.norb < 4) return;
  S inv = recip(r);
  S R[3][3];
  rot_to_x(-(dx * inv), -(dy * inv), -(dz * inv), R);
  S ri[4];
  local_xh(orb.ddp2, orb.ddp3 / sqrt(2.0), orb.po1, orb.po2, orb.po3, crho, r, ri);
  S w[10];
  assemble_xh(ri, R, w);
#pragma unroll
  for (int k = 0; k < 10; ++k) blk[k] = w[k] * (-core.tore);
}

// A_k(x) = int_1^inf t^k e^{-xt} dt and B_k(x) = int_-1^1 t^k e^{-xt} dt, k = 0..6.
template <class S> __device__ inline void aint(S x, S a[7]) {
  if (fabs(val(x)) < 1.0e-12) {
#pragma unroll
    for (int i = 0; i < 7; ++i) a[i] = cst(0.0, x);
    return;
  }
  S inv = recip(x);
  a[0] = expx(-x) * inv;
#pragma unroll
  for (int i = 1; i < 7; ++i) a[i] = a[0] + a[i - 1] * inv * (double)i;
}

template <class S> __device__ inline void bint(S x, S b[7]) {
#pragma unroll
  for (int i = 0; i < 7; ++i) b[i] = cst(0.0, x);
  if (fabs(val(x)) > 3.0) {
    S inv = recip(x);
    S tx = expx(x) * inv;
    S tmx = expx(-x) * inv * (-1.0);
    b[0] = tx + tmx;
#pragma unroll
    for (int i = 1; i < 7; ++i) b[i] = tx * ((i & 1) ? -1.0 : 1.0) + tmx + b[i - 1] * inv * (double)i;
  } else {
    // Term-by-term series of the defining integral; 40 terms reach the last bit for |x| <= 3.
    S term = cst(1.0, x);
    for (int m = 0; m < 40; ++m) {
#pragma unroll
      for (int k = 0; k < 7; ++k)
        if (((k + m) & 1) == 0) b[k] = b[k] + term * (2.0 / (double)(k + m + 1));
      term = term * (-x) * (1.0 / (double)(m + 1));
    }
  }
}

// Local sigma/pi overlaps s111, s211, s121, s221, s222 for atoms ordered so that
// ea.n_s >= eb.n_s.  Returns false when no closed form exists for the pair.
template <class S> __device__ inline bool overlap_local(const ElemConst& ea, const ElemConst& eb, S r, S s[5]) {
#pragma unroll
  for (int i = 0; i < 5; ++i) s[i] = cst(0.0, r);
  const int na = ea.n_s, nb = eb.n_s;
  int jc = 0;
  if (na == 1 && nb == 1) jc = 2;
  else if (na == 2 && nb == 1) jc = 3;
  else if (na == 2 && nb == 2) jc = 4;
  else if (na == 3 && nb == 1) jc = 431;
  else if (na == 3 && nb == 2) jc = 5;
  else if (na == 3 && nb == 3) jc = 6;
  if (jc == 0) return false;
  const double za0 = ea.zs, za1 = ea.zp, zb0 = eb.zs, zb1 = eb.zp;
  double sm = 1.0e300;
  const double sums[4] = {za0 + zb0, za1 + zb0, za0 + zb1, za1 + zb1};
#pragma unroll
  for (int i = 0; i < 4; ++i)
    if (sums[i] > 0.0 && sums[i] < sm) sm = sums[i];
  sm *= 0.5;
  if (sm < 1.0e299 && val(r) * sm > 300.0) return true;  // overlap has underflowed
  S a111[7], b111[7], a211[7], b211[7], a121[7], b121[7], a22[7], b22[7];
  aint(r * (0.5 * (za0 + zb0)), a111);
  bint(r * (0.5 * (za0 - zb0)), b111);
  if (jc != 2) {
    aint(r * (0.5 * (za1 + zb0)), a211);
    bint(r * (0.5 * (za1 - zb0)), b211);
  }
  if (jc == 4 || jc == 5 || jc == 6) {
    aint(r * (0.5 * (za0 + zb1)), a121);
    bint(r * (0.5 * (za0 - zb1)), b121);
    aint(r * (0.5 * (za1 + zb1)), a22);
    bint(r * (0.5 * (za1 - zb1)), b22);
  }
  const double sq3 = sqrt(3.0);
  if (jc == 2) {
    s[0] = powfx(r * r * (za0 * zb0), 1.5) * (a111[2] * b111[0] - b111[2] * a111[0]) / 4.0;
  } else if (jc == 3) {
    S r4 = powi<4>(r);
    s[0] = r4 * (pow(zb0, 1.5) * pow(za0, 2.5)) * (a111[3] * b111[0] - b111[3] * a111[0] + a111[2] * b111[1] - b111[2] * a111[1]) / (sq3 * 8.0);
    s[1] = r4 * (pow(zb0, 1.5) * pow(za1, 2.5)) * (a211[2] * b211[0] - b211[2] * a211[0] + a211[3] * b211[1] - b211[3] * a211[1]) / 8.0;
  } else if (jc == 4) {
    S r5 = powi<5>(r);
    s[0] = r5 * pow(zb0 * za0, 2.5) * (a111[4] * b111[0] + b111[4] * a111[0] - a111[2] * b111[2] * 2.0) / 48.0;
    s[1] = r5 * pow(zb0 * za1, 2.5) *
           (a211[3] * (b211[0] - b211[2]) - a211[1] * (b211[2] - b211[4]) + b211[3] * (a211[0] - a211[2]) - b211[1] * (a211[2] - a211[4])) /
           (16.0 * sq3);
    s[2] = r5 * pow(zb1 * za0, 2.5) *
           (a121[3] * (b121[0] - b121[2]) - a121[1] * (b121[2] - b121[4]) - b121[3] * (a121[0] - a121[2]) + b121[1] * (a121[2] - a121[4])) /
           (16.0 * sq3);
    S w = r5 * pow(zb1 * za1, 2.5) / 16.0;
    s[3] = -(w * (b22[2] * (a22[4] + a22[0]) - a22[2] * (b22[4] + b22[0])));
    s[4] = w * (a22[4] * (b22[0] - b22[2]) - b22[4] * (a22[0] - a22[2]) - a22[2] * b22[0] + b22[2] * a22[0]) * 0.5;
  } else if (jc == 431) {
    S r5 = powi<5>(r);
    s[0] = r5 * (pow(zb0, 1.5) * pow(za0, 3.5)) * (a111[4] * b111[0] + b111[1] * a111[3] * 2.0 - a111[1] * b111[3] * 2.0 - b111[4] * a111[0]) /
           (sqrt(10.0) * 24.0);
    s[1] = r5 * (pow(zb0, 1.5) * pow(za1, 3.5)) *
           (a211[3] * (b211[0] + b211[2]) - a211[1] * (b211[4] + b211[2]) + b211[1] * (a211[2] + a211[4]) - b211[3] * (a211[2] + a211[0])) /
           (8.0 * sqrt(30.0));
  } else if (jc == 5) {
    S r6 = powi<6>(r);
    s[0] = r6 * (pow(zb0, 2.5) * pow(za0, 3.5)) *
           (a111[5] * b111[0] + b111[1] * a111[4] - b111[2] * a111[3] * 2.0 - a111[2] * b111[3] * 2.0 + b111[4] * a111[1] + b111[5] * a111[0]) /
           (sqrt(30.0) * 48.0);
    s[1] = r6 * (pow(zb0, 2.5) * pow(za1, 3.5)) *
           (a211[4] * b211[0] + b211[1] * a211[5] - b211[3] * a211[3] * 2.0 - a211[2] * b211[2] * 2.0 + a211[1] * b211[5] + a211[0] * b211[4]) /
           (48.0 * sqrt(10.0));
    s[2] = r6 * (pow(zb1, 2.5) * pow(za0, 3.5)) *
           ((a121[4] * b121[0] - a121[5] * b121[1]) + (a121[3] * b121[1] - a121[4] * b121[2]) * 2.0 - (a121[1] * b121[3] - a121[2] * b121[4]) * 2.0 -
            (a121[0] * b121[4] - a121[1] * b121[5])) /
           (48.0 * sqrt(10.0));
    s[3] = r6 * (pow(zb1, 2.5) * pow(za1, 3.5)) *
           ((a22[3] * b22[0] - a22[5] * b22[2]) + (a22[2] * b22[1] - a22[4] * b22[3]) - (a22[1] * b22[2] - a22[3] * b22[4]) -
            (a22[0] * b22[3] - a22[2] * b22[5])) /
           (16.0 * sqrt(30.0));
    s[4] = r6 * (pow(zb1, 2.5) * pow(za1, 3.5)) *
           ((a22[5] - a22[3]) * (b22[0] - b22[2]) + (a22[4] - a22[2]) * (b22[1] - b22[3]) - (a22[3] - a22[1]) * (b22[2] - b22[4]) -
            (a22[2] - a22[0]) * (b22[3] - b22[5])) /
           (32.0 * sqrt(30.0));
  } else {  // jc == 6
    S r7 = powi<7>(r);
    s[0] = r7 * pow(zb0 * za0, 3.5) * (a111[6] * b111[0] - b111[2] * a111[4] * 3.0 + a111[2] * b111[4] * 3.0 - a111[0] * b111[6]) / 1440.0;
    s[1] = r7 * pow(zb0 * za1, 3.5) *
           ((a211[5] * b211[0] + a211[6] * b211[1]) + (-a211[4] * b211[1] - a211[5] * b211[2]) - (a211[3] * b211[2] + a211[4] * b211[3]) * 2.0 -
            (-a211[2] * b211[3] - a211[3] * b211[4]) * 2.0 + (a211[1] * b211[4] + a211[2] * b211[5]) + (-a211[0] * b211[5] - a211[1] * b211[6])) /
           (480.0 * sq3);
    s[2] = r7 * pow(zb1 * za0, 3.5) *
           ((a121[5] * b121[0] - a121[6] * b121[1]) + (a121[4] * b121[1] - a121[5] * b121[2]) - (a121[3] * b121[2] - a121[4] * b121[3]) * 2.0 -
            (a121[2] * b121[3] - a121[3] * b121[4]) * 2.0 + (a121[1] * b121[4] - a121[2] * b121[5]) + (a121[0] * b121[5] - a121[1] * b121[6])) /
           (480.0 * sq3);
    s[3] = r7 * (pow(zb1, 3.5) * pow(za1, 3.5)) *
           ((a22[4] * b22[0] - a22[6] * b22[2]) - (a22[2] * b22[2] - a22[4] * b22[4]) * 2.0 + (a22[0] * b22[4] - a22[2] * b22[6])) / 480.0;
    s[4] = r7 * (pow(zb1, 3.5) * pow(za1, 3.5)) *
           ((a22[6] - a22[4]) * (b22[0] - b22[2]) - (a22[4] - a22[2]) * (b22[2] - b22[4]) * 2.0 + (a22[2] - a22[0]) * (b22[4] - b22[6])) / 960.0;
  }
  return true;
}

// Molecular-frame overlap block S[a][b] (4 x 4) between atom i (rows) and atom j (cols).
// (ux, uy, uz) is the unit vector from i to j and r = |R_j - R_i|.  The element with the higher
// principal quantum number is evaluated first and the block transposed back.
template <class S>
__device__ inline bool overlap_block(const ElemConst& ei, const ElemConst& ej, S ux, S uy, S uz, S r, S out[4][4]) {
  const bool swap = ei.n_s < ej.n_s;
  const ElemConst& ea = swap ? ej : ei;
  const ElemConst& eb = swap ? ei : ej;
  S loc[5];
  if (!overlap_local(ea, eb, r, loc)) return false;
  const double sg = swap ? -1.0 : 1.0;
  S R[3][3];
  rot_to_x(ux * sg, uy * sg, uz * sg, R);
  S di[4][4];
  di[0][0] = loc[0];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    di[1 + a][0] = loc[1] * R[0][a];
    di[0][1 + a] = -(loc[2] * R[0][a]);
  }
  S mm[3] = {-loc[3], loc[4], loc[4]};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      S v = R[0][a] * mm[0] * R[0][b];
      v = v + R[1][a] * mm[1] * R[1][b];
      v = v + R[2][a] * mm[2] * R[2][b];
      di[1 + a][1 + b] = v;
    }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) out[a][b] = swap ? di[b][a] : di[a][b];
  return true;
}

// PM3 core-core repulsion of one pair at distance rb (Bohr).
template <class S>
__device__ inline S core_core_pair(const TableConst& T, int si, int sj, S rb) {
  const ElemConst& ei = T.el[si];
  const ElemConst& ej = T.el[sj];
  S ra = rb * A0;
  const double rho = ei.po9 + ej.po9;
  S gab = recip(sqrtx(rb * rb + rho * rho)) * EV;
  const double cp = ei.tore * ej.tore;
  S bare = gab * cp;
  const double xf = T.xfac[si * MAX_SLOTS + sj];
  const bool has_pair = fabs(xf) > 1.0e-5;
  S e;
  if (has_pair) {
    double al = T.alpb[si * MAX_SLOTS + sj];
    if (al < 1.0e-6) al = 1.2;
    e = bare * (expx(ra * -al) * (2.0 * xf) + 1.0);
  } else {
    S exi = expx(ra * -ei.alp), exj = expx(ra * -ej.alp);
    S sc = exi + exj;
    const int zs = ei.z + ej.z;
    if (zs == 8 || zs == 9) {  // MNDO N-H / O-H adjustment
      if (ei.z == 7 || ei.z == 8) sc = sc + (ra - 1.0) * exi;
      if (ej.z == 7 || ej.z == 8) sc = sc + (ra - 1.0) * exj;
    }
    e = absx(sc * bare) + bare;
    for (int t = 0; t < 2; ++t) {
      const ElemConst& eg = t == 0 ? ei : ej;
      for (int k = 0; k < eg.ngauss; ++k) {
        S dl = ra - eg.gm[k];
        S ax = dl * dl * eg.gl[k];
        if (val(ax) <= 25.0) e = e + expx(-ax) * (cp * eg.gk[k]) * recip(ra);
      }
    }
  }
  return e;
}

}  // namespace pm3dev
