// PM3 CPU ORACLE -- test infrastructure only.
//
// A plain FP64 C++ restatement of the reference's molecular RHF PM3 path, used by tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.  It is never
// linked into, loaded by, or called from the product library (pm3-rs_amd/), and the product
// never falls back to it.
//
// Each function cites the reference file:line it restates (paths relative to the reference
// repository root).  The eigensolver is this file's own Householder tridiagonalisation plus
// implicit-shift QL (Golub & Van Loan, Matrix Computations, 4th ed., Alg. 8.3.1 and Sec. 8.3.5),
// standing in for faer 0.24.4's self_adjoint_eigen (third party, not in the reference tree);
// only its contract (ascending eigenvalues, orthonormal eigenvector columns,
// src/linalg.rs:217-263) is pinned.
//
// Parity pins (tests/test_oracle_*.py): pm3-rs's own total energies in
// tools/oracle/INTERNAL_CONSISTENCY.json and heats of formation in
// tools/oracle/ALL_ELEMENTS_RESULTS.json, MOPAC rows in tests/data/mopac_oracle.tsv, the
// water HoF and dipole (src/scf.rs:1899-1904), the O-H overlap (src/overlap.rs:688-705), the
// closed-form (ss|ss) (src/integrals.rs:585-594), the packed-vs-four-index Fock identity
// (src/fock.rs:510-602), and finite differences of this file's own SCF energy for the
// gradient (src/gradient.rs:137-176).  The UHF restatement (uhf_loop, orc_run_pm3) is pinned by
// the 17 open-shell MOPAC rows of tests/data/mopac_oracle.tsv, and the closed-shell stability
// ladder (two unrestricted rungs, spin-purity collapse) by the molecules the reference's own test
// names as rescued by it (tests/mopac_oracle.rs:268-308): butyne, ketene, tetrafluoroethene,
// tetrachloroethene.  Parity unpinned: exact SCF iteration counts and DIIS trajectories.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// src/constants.rs:21-31,56
constexpr double PM3_EV = 27.211386245988;
constexpr double PM3_A0 = 0.529177210903;
constexpr double EV_TO_KCAL = 23.06054783061903;
constexpr double AU_DIPOLE_TO_DEBYE = 2.541746473;

// ---------------------------------------------------------------------------------------
// Forward-mode dual number over the three displacement components (src/dual.rs:78-...).
struct Dual {
  double v;
  double d[3];
};
static inline Dual dmk(double v, double a, double b, double c) { Dual r; r.v = v; r.d[0] = a; r.d[1] = b; r.d[2] = c; return r; }
static inline Dual operator+(Dual a, Dual b) { return dmk(a.v + b.v, a.d[0] + b.d[0], a.d[1] + b.d[1], a.d[2] + b.d[2]); }
static inline Dual operator-(Dual a, Dual b) { return dmk(a.v - b.v, a.d[0] - b.d[0], a.d[1] - b.d[1], a.d[2] - b.d[2]); }
static inline Dual operator*(Dual a, Dual b) {
  return dmk(a.v * b.v, a.d[0] * b.v + a.v * b.d[0], a.d[1] * b.v + a.v * b.d[1], a.d[2] * b.v + a.v * b.d[2]);
}
static inline Dual operator-(Dual a) { return dmk(-a.v, -a.d[0], -a.d[1], -a.d[2]); }
static inline Dual operator+(Dual a, double b) { return dmk(a.v + b, a.d[0], a.d[1], a.d[2]); }
static inline Dual operator-(Dual a, double b) { return dmk(a.v - b, a.d[0], a.d[1], a.d[2]); }
static inline Dual operator*(Dual a, double b) { return dmk(a.v * b, a.d[0] * b, a.d[1] * b, a.d[2] * b); }
static inline Dual operator/(Dual a, double b) { return a * (1.0 / b); }
static inline Dual operator/(Dual a, Dual b) {
  double inv = 1.0 / b.v, inv2 = inv * inv;
  return dmk(a.v * inv, (a.d[0] * b.v - a.v * b.d[0]) * inv2, (a.d[1] * b.v - a.v * b.d[1]) * inv2,
             (a.d[2] * b.v - a.v * b.d[2]) * inv2);
}
static inline Dual scale_d(Dual a, double v, double f) { return dmk(v, a.d[0] * f, a.d[1] * f, a.d[2] * f); }

template <class S> struct Ops;
template <> struct Ops<double> {
  static double cst(double x) { return x; }
  static double val(double x) { return x; }
  static double sqrt_(double x) { return std::sqrt(x); }
  static double exp_(double x) { return std::exp(x); }
  static double recip(double x) { return 1.0 / x; }
  static double powf_(double x, double p) { return std::pow(x, p); }
  static double powi_(double x, int n) { double r = 1.0; for (int i = 0; i < n; ++i) r *= x; return r; }
  static double abs_(double x) { return std::fabs(x); }
};
template <> struct Ops<Dual> {
  static Dual cst(double x) { return dmk(x, 0, 0, 0); }
  static double val(Dual x) { return x.v; }
  static Dual sqrt_(Dual x) { double s = std::sqrt(x.v); return scale_d(x, s, 0.5 / s); }
  static Dual exp_(Dual x) { double e = std::exp(x.v); return scale_d(x, e, e); }
  static Dual recip(Dual x) { double r = 1.0 / x.v; return scale_d(x, r, -r * r); }
  static Dual powf_(Dual x, double p) { double v = std::pow(x.v, p); return scale_d(x, v, p * std::pow(x.v, p - 1.0)); }
  static Dual powi_(Dual x, int n) {
    double r = 1.0; for (int i = 0; i < n - 1; ++i) r *= x.v;
    return scale_d(x, r * x.v, n * r);
  }
  static Dual abs_(Dual x) { return x.v < 0.0 ? -x : x; }
};

// ---------------------------------------------------------------------------------------
// Parameters: src/params.rs:155-475 (element rows), :584-649 (dd, qq, rho1, rho2 with exactly
// five secant steps), :655-733 (po/ddp for main-group sp elements), :527-552 (e_isol
// average-of-configuration coefficients).
struct Elem {
  int z = 0, n_s = 0, n_p = 0, n_orb = 0;
  double uss = 0, upp = 0, zs = 0, zp = 0, betas = 0, betap = 0;
  double gss = 0, gsp = 0, gpp = 0, gp2 = 0, hsp = 0, alp = 0;
  std::vector<std::array<double, 3>> gauss;
  double ios = 0, iop = 0, eheat_ev = 0, mass = 0, tore = 0, e_isol = 0;
  double dd = 0, qq = 0, rho0 = 0, rho1 = 0, rho2 = 0;
  double po[10] = {0};
  double ddp[7] = {0};
  bool has_p() const { return n_orb >= 4; }
};

struct Table {
  std::map<int, Elem> el;
  std::map<std::pair<int, int>, std::pair<double, double>> pair;  // (max z, min z) -> (alpb, xfac)
};

static double multipole_pqn(int z) {  // src/params.rs:572-582
  if (z <= 2) return 1.0;
  if (z <= 10) return 2.0;
  if (z <= 18) return 3.0;
  return 0.0;
}

static double additive_rho1(double hsp_ev, double dd) {  // src/params.rs:605-622
  double hsp = std::max(hsp_ev, 1.0e-7) / PM3_EV;
  auto g = [&](double d) { return 0.5 * d - 0.5 / std::sqrt(4.0 * dd * dd + 1.0 / (d * d)); };
  double gdd1 = std::pow(hsp / (dd * dd), 1.0 / 3.0);
  double d1 = gdd1, d2 = gdd1 + 0.04;
  for (int it = 0; it < 5; ++it) {
    double df = d2 - d1, h1 = g(d1), h2 = g(d2);
    if (std::fabs(h2 - h1) < 1.0e-25) break;
    double d3 = d1 + df * (hsp - h1) / (h2 - h1);
    d1 = d2;
    d2 = d3;
  }
  return 0.5 / d2;
}

static double additive_rho2(double hpp_ev, double qq) {  // src/params.rs:629-649
  double hpp = hpp_ev / PM3_EV;
  auto g = [&](double q) {
    return 0.25 * q - 0.5 / std::sqrt(4.0 * qq * qq + 1.0 / (q * q)) + 0.25 / std::sqrt(8.0 * qq * qq + 1.0 / (q * q));
  };
  double q4 = qq * qq * qq * qq;
  double gqq = std::pow(16.0 * hpp / (48.0 * q4), 0.2);
  double q1 = gqq, q2 = gqq + 0.04;
  for (int it = 0; it < 5; ++it) {
    double qf = q2 - q1, h1 = g(q1), h2 = g(q2);
    if (std::fabs(h2 - h1) < 1.0e-25) break;
    double q3 = q1 + qf * (hpp - h1) / (h2 - h1);
    q1 = q2;
    q2 = q3;
  }
  return 0.5 / q2;
}

static Table build_table() {
  Table t;
#define PM3_ELEMENT(Z, USS, UPP, ZS, ZP, BS, BP, GSS, GSP, GPP, GP2, HSP, ALP, G1K, G1L, G1M, G2K, G2L, G2M, IOS, IOP, NS, NP, EHEAT, MASS, TORE) \
  {                                                                                                                    \
    Elem e;                                                                                                            \
    e.z = Z; e.uss = USS; e.upp = UPP; e.zs = ZS; e.zp = ZP; e.betas = BS; e.betap = BP;                                \
    e.gss = GSS; e.gsp = GSP; e.gpp = GPP; e.gp2 = GP2; e.hsp = HSP; e.alp = ALP;                                       \
    if (G1K != 0.0 || G1L != 0.0) e.gauss.push_back({G1K, G1L, G1M});                                                   \
    if (G2K != 0.0 || G2L != 0.0) e.gauss.push_back({G2K, G2L, G2M});                                                   \
    e.ios = IOS; e.iop = IOP; e.n_s = (int)NS; e.n_p = (int)NP; e.eheat_ev = EHEAT / EV_TO_KCAL; e.mass = MASS;         \
    e.tore = TORE;                                                                                                     \
    t.el[Z] = e;                                                                                                       \
  }
#define PM3_PAIR(ZI, ZJ, ALPB, XFAC) t.pair[{std::max(ZI, ZJ), std::min(ZI, ZJ)}] = {ALPB, XFAC};
#include "../pm3-rs_amd/csrc/pm3_table.inc"
#undef PM3_ELEMENT
#undef PM3_PAIR
  for (auto& kv : t.el) {
    Elem& e = kv.second;
    e.hsp = std::max(e.hsp, 1.0e-7);  // src/params.rs:206
    e.n_orb = e.zp > 0.0 ? 4 : 1;      // src/params.rs:191-199
    if (e.n_orb >= 4) {                // src/params.rs:240-250
      double zp = std::max(e.zp, 0.3);
      double qn = multipole_pqn(e.z);
      e.dd = (2.0 * qn + 1.0) * std::pow(4.0 * e.zs * zp, qn + 0.5) / std::pow(e.zs + zp, 2.0 * qn + 2.0) / std::sqrt(3.0);
      e.qq = std::sqrt((4.0 * qn * qn + 6.0 * qn + 2.0) / 20.0) / zp;
      double hpp = std::max(0.5 * (e.gpp - e.gp2), 0.1);
      e.rho1 = additive_rho1(e.hsp, e.dd);
      e.rho2 = additive_rho2(hpp, e.qq);
    }
    e.rho0 = e.gss > 0.0 ? 0.5 * PM3_EV / e.gss : 0.0;  // src/params.rs:254-261
    // src/params.rs:266-289 and :527-552
    if (e.z == 1) {
      e.e_isol = e.uss;
    } else {
      double ios = e.ios, iop = e.iop;
      long k = (long)iop;
      long l = std::min(k, 6 - k);
      double gssc = std::max(ios - 1.0, 0.0);
      double gspc = ios * iop;
      double hspc = -iop * ios * 0.5;
      double gp2c = (double)((k * (k - 1)) / 2) + (double)(l * (l - 1)) / 4.0;
      double gppc = -((double)(l * (l - 1))) / 4.0;
      e.e_isol = e.uss * ios + e.upp * iop + 0.0 + e.gss * gssc + e.gpp * gppc + e.gsp * gspc + e.gp2 * gp2c + e.hsp * hspc;
    }
    // src/params.rs:694-710
    if (e.n_orb >= 4) {
      e.po[1] = e.rho0; e.po[2] = e.rho1; e.po[3] = e.rho2; e.po[7] = e.rho0; e.po[9] = e.rho0;
      e.ddp[2] = e.dd; e.ddp[3] = e.qq * std::sqrt(2.0);
    } else {
      e.po[1] = e.rho0; e.po[7] = e.rho0; e.po[9] = e.rho0;
    }
  }
  return t;
}

static const Table& table() {
  static Table t = build_table();
  return t;
}

// ---------------------------------------------------------------------------------------
// Two-centre integrals: src/integrals.rs.
static inline int pack(int a, int b) { int h = a >= b ? a : b, l = a >= b ? b : a; return h * (h + 1) / 2 + l; }  // :18-21

template <class S> using M3 = std::array<std::array<S, 3>, 3>;

template <class S> static M3<S> rotation_to_x(S vx, S vy, S vz) {  // src/integrals.rs:25-59
  using O = Ops<S>;
  M3<S> r;
  if (O::val(vx) < 0.0) {
    S one = O::cst(1.0);
    S ux = vx - 1.0, uy = vy, uz = vz;
    S f = O::recip(ux * ux + uy * uy + uz * uz) * (-2.0);
    r[0] = {one + f * ux * ux, f * ux * uy, f * ux * uz};
    r[1] = {f * uy * ux, one + f * uy * uy, f * uy * uz};
    r[2] = {f * uz * ux, f * uz * uy, one + f * uz * uz};
  } else {
    S qy = vz, qz = -vy, qw = vx + 1.0;
    S inv = O::recip(O::sqrt_(qy * qy + qz * qz + qw * qw));
    qy = qy * inv; qz = qz * inv; qw = qw * inv;
    r[0] = {(qy * qy + qz * qz) * (-2.0) + 1.0, qz * qw * (-2.0), qy * qw * 2.0};
    r[1] = {qz * qw * 2.0, qz * qz * (-2.0) + 1.0, qy * qz * 2.0};
    r[2] = {qy * qw * (-2.0), qy * qz * 2.0, qy * qy * (-2.0) + 1.0};
  }
  return r;
}

// Multipole parameters of one centre as used by the local integrals.
struct Centre { double dd, qq, rho0, rho1, rho2; };
static Centre centre_ee(const Elem& e) { return {e.dd, e.qq, e.rho0, e.rho1, e.rho2}; }

template <class S> static std::array<S, 4> local_xh(const Centre& a, double rho0_b, S r) {  // src/integrals.rs:289-309
  using O = Ops<S>;
  const double ev1 = PM3_EV / 2.0, ev2 = PM3_EV / 4.0;
  S da = O::cst(a.dd), qa = O::cst(a.qq * 2.0);
  double aee = (a.rho0 + rho0_b) * (a.rho0 + rho0_b);
  double ade = (a.rho1 + rho0_b) * (a.rho1 + rho0_b);
  double aqe = (a.rho2 + rho0_b) * (a.rho2 + rho0_b);
  S ee = O::recip(O::sqrt_(r * r + aee)) * PM3_EV;
  S ev1dsqr6 = O::recip(O::sqrt_(r * r + aqe)) * ev1;
  std::array<S, 4> ri;
  ri[0] = ee;
  ri[1] = O::recip(O::sqrt_((r + da) * (r + da) + ade)) * ev1 - O::recip(O::sqrt_((r - da) * (r - da) + ade)) * ev1;
  ri[2] = ee + O::recip(O::sqrt_((r + qa) * (r + qa) + aqe)) * ev2 + O::recip(O::sqrt_((r - qa) * (r - qa) + aqe)) * ev2 - ev1dsqr6;
  ri[3] = ee + O::recip(O::sqrt_(r * r + qa * qa + aqe)) * ev1 - ev1dsqr6;
  return ri;
}

template <class S> static std::array<S, 22> local_xx(const Centre& A, const Centre& B, S r) {  // src/integrals.rs:311-464
  using O = Ops<S>;
  const double ev1 = PM3_EV / 2.0, ev2 = PM3_EV / 4.0, ev3 = PM3_EV / 8.0, ev4 = PM3_EV / 16.0;
  S da = O::cst(A.dd), db = O::cst(B.dd), qa = O::cst(A.qq * 2.0), qb = O::cst(B.qq * 2.0);
  S qa1 = O::cst(A.qq), qb1 = O::cst(B.qq);
  auto sq = [](double x) { return x * x; };
  double aee = sq(A.rho0 + B.rho0), ade = sq(A.rho1 + B.rho0), aqe = sq(A.rho2 + B.rho0);
  double aed = sq(A.rho0 + B.rho1), aeq = sq(A.rho0 + B.rho2), axx = sq(A.rho1 + B.rho1);
  double adq = sq(A.rho1 + B.rho2), aqd = sq(A.rho2 + B.rho1), aqq = sq(A.rho2 + B.rho2);
  auto g = [](S x, double c) { return O::recip(O::sqrt_(x + c)); };
  auto sq2 = [](S a, double c) { return O::recip(O::sqrt_(a * a + c)); };

  S ee = O::recip(O::sqrt_(r * r + aee)) * PM3_EV;
  S dze = sq2(r - da, ade) * ev1 - sq2(r + da, ade) * ev1;
  S ev1dsqr6 = O::recip(O::sqrt_(r * r + aqe)) * ev1;
  S qzze = sq2(r - qa, aqe) * ev2 + sq2(r + qa, aqe) * ev2 - ev1dsqr6;
  S qxxe = g(r * r + qa * qa, aqe) * ev1 - ev1dsqr6;
  S edz = sq2(r + db, aed) * ev1 - sq2(r - db, aed) * ev1;
  S ev1dsqr12 = O::recip(O::sqrt_(r * r + aeq)) * ev1;
  S eqzz = sq2(r - qb, aeq) * ev2 + sq2(r + qb, aeq) * ev2 - ev1dsqr12;
  S eqxx = g(r * r + qb * qb, aeq) * ev1 - ev1dsqr12;
  S ev2dsqr20 = sq2(r + da, adq) * ev2;
  S ev2dsqr22 = sq2(r - da, adq) * ev2;
  S ev2dsqr24 = sq2(r - db, aqd) * ev2;
  S ev2dsqr26 = sq2(r + db, aqd) * ev2;
  S ev2dsqr36 = O::recip(O::sqrt_(r * r + aqq)) * ev2;
  S ev2dsqr39 = g(r * r + qa * qa, aqq) * ev2;
  S ev2dsqr40 = g(r * r + qb * qb, aqq) * ev2;
  S ev3dsqr42 = sq2(r - qb, aqq) * ev3;
  S ev3dsqr44 = sq2(r + qb, aqq) * ev3;
  S ev3dsqr46 = sq2(r + qa, aqq) * ev3;
  S ev3dsqr48 = sq2(r - qa, aqq) * ev3;

  std::array<S, 22> ri;
  ri[0] = ee;
  ri[1] = -dze;
  ri[2] = ee + qzze;
  ri[3] = ee + qxxe;
  ri[4] = -edz;
  ri[5] = sq2(r + da - db, axx) * ev2 + sq2(r - da + db, axx) * ev2 - sq2(r - da - db, axx) * ev2 - sq2(r + da + db, axx) * ev2;
  ri[6] = g(r * r + (da - db) * (da - db), axx) * ev1 - g(r * r + (da + db) * (da + db), axx) * ev1;
  ri[7] = -edz + sq2(r + qa - db, aqd) * ev3 - sq2(r + qa + db, aqd) * ev3 + sq2(r - qa - db, aqd) * ev3 -
          sq2(r - qa + db, aqd) * ev3 - ev2dsqr24 + ev2dsqr26;
  ri[8] = -edz - ev2dsqr24 + g((r - db) * (r - db) + qa * qa, aqd) * ev2 + ev2dsqr26 - g((r + db) * (r + db) + qa * qa, aqd) * ev2;
  ri[9] = g((qa1 - db) * (qa1 - db) + (r + qa1) * (r + qa1), aqd) * ev2 - g((qa1 - db) * (qa1 - db) + (r - qa1) * (r - qa1), aqd) * ev2 -
          g((qa1 + db) * (qa1 + db) + (r + qa1) * (r + qa1), aqd) * ev2 + g((qa1 + db) * (qa1 + db) + (r - qa1) * (r - qa1), aqd) * ev2;
  ri[10] = ee + eqzz;
  ri[11] = ee + eqxx;
  ri[12] = -dze + sq2(r + da - qb, adq) * ev3 - sq2(r - da - qb, adq) * ev3 + sq2(r + da + qb, adq) * ev3 -
           sq2(r - da + qb, adq) * ev3 + ev2dsqr22 - ev2dsqr20;
  ri[13] = -dze - ev2dsqr20 + g((r + da) * (r + da) + qb * qb, adq) * ev2 + ev2dsqr22 - g((r - da) * (r - da) + qb * qb, adq) * ev2;
  ri[14] = g((da - qb1) * (da - qb1) + (r - qb1) * (r - qb1), adq) * ev2 - g((da - qb1) * (da - qb1) + (r + qb1) * (r + qb1), adq) * ev2 -
           g((da + qb1) * (da + qb1) + (r - qb1) * (r - qb1), adq) * ev2 + g((da + qb1) * (da + qb1) + (r + qb1) * (r + qb1), adq) * ev2;
  ri[15] = ee + eqzz + qzze + sq2(r + qa - qb, aqq) * ev4 + sq2(r + qa + qb, aqq) * ev4 + sq2(r - qa - qb, aqq) * ev4 +
           sq2(r - qa + qb, aqq) * ev4 - ev3dsqr48 - ev3dsqr46 - ev3dsqr42 - ev3dsqr44 + ev2dsqr36;
  ri[16] = ee + eqzz + qxxe + g((r - qb) * (r - qb) + qa * qa, aqq) * ev3 + g((r + qb) * (r + qb) + qa * qa, aqq) * ev3 - ev3dsqr42 -
           ev3dsqr44 - ev2dsqr39 + ev2dsqr36;
  ri[17] = ee + eqxx + qzze + g((r + qa) * (r + qa) + qb * qb, aqq) * ev3 + g((r - qa) * (r - qa) + qb * qb, aqq) * ev3 - ev3dsqr46 -
           ev3dsqr48 - ev2dsqr40 + ev2dsqr36;
  S qxxqxx = g(r * r + (qa - qb) * (qa - qb), aqq) * ev3 + g(r * r + (qa + qb) * (qa + qb), aqq) * ev3 - ev2dsqr39 - ev2dsqr40 + ev2dsqr36;
  ri[18] = ee + eqxx + qxxe + qxxqxx;
  S dm = (qa1 - qb1) * (qa1 - qb1), dp = (qa1 + qb1) * (qa1 + qb1);
  ri[19] = g((r + qa1 - qb1) * (r + qa1 - qb1) + dm, aqq) * ev3 - g((r + qa1 + qb1) * (r + qa1 + qb1) + dm, aqq) * ev3 -
           g((r - qa1 - qb1) * (r - qa1 - qb1) + dm, aqq) * ev3 + g((r - qa1 + qb1) * (r - qa1 + qb1) + dm, aqq) * ev3 -
           g((r + qa1 - qb1) * (r + qa1 - qb1) + dp, aqq) * ev3 + g((r + qa1 + qb1) * (r + qa1 + qb1) + dp, aqq) * ev3 +
           g((r - qa1 - qb1) * (r - qa1 - qb1) + dp, aqq) * ev3 - g((r - qa1 + qb1) * (r - qa1 + qb1) + dp, aqq) * ev3;
  S qxxqyy = g(r * r + qa * qa + qb * qb, aqq) * ev2 - ev2dsqr39 - ev2dsqr40 + ev2dsqr36;
  ri[20] = ee + eqxx + qxxe + qxxqyy;
  ri[21] = (qxxqxx - qxxqyy) * 0.5;
  return ri;
}

template <class S> static void build_wxh(const std::array<S, 4>& ri, const M3<S>& R, S* w) {  // src/integrals.rs:466-478
  w[pack(0, 0)] = ri[0];
  for (int k = 0; k < 3; ++k) w[pack(k + 1, 0)] = ri[1] * R[0][k];
  for (int k = 0; k < 3; ++k)
    for (int l = 0; l <= k; ++l) {
      S t0 = R[0][k] * R[0][l];
      S t1 = R[1][k] * R[1][l] + R[2][k] * R[2][l];
      w[pack(k + 1, l + 1)] = ri[2] * t0 + ri[3] * t1;
    }
}

template <class S> static void rotate_xx(const std::array<S, 22>& ri, const M3<S>& R, S* w) {  // src/integrals.rs:480-552
  const auto& r0 = R[0];
  const auto& r1 = R[1];
  const auto& r2 = R[2];
  int idx = 0;
  for (int kk = 0; kk < 4; ++kk)
    for (int ll = 0; ll <= kk; ++ll)
      for (int mm = 0; mm < 4; ++mm)
        for (int nn = 0; nn <= mm; ++nn) {
          int k = kk - 1, l = ll - 1, m = mm - 1, n = nn - 1;
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
            S t0 = r0[k] * r0[l];
            S t1 = r1[k] * r1[l] + r2[k] * r2[l];
            v = ri[2] * t0 + ri[3] * t1;
          } else if (nn == 0) {
            S t0 = r0[k] * r0[l] * r0[m];
            S t1 = (r1[k] * r1[l] + r2[k] * r2[l]) * r0[m];
            S t2 = r1[l] * r1[m] + r2[l] * r2[m];
            v = ri[7] * t0 + ri[8] * t1 + ri[9] * (r0[k] * t2 + r0[l] * (r1[k] * r1[m] + r2[k] * r2[m]));
          } else {
            S t0 = r0[k] * r0[l] * r0[m] * r0[n];
            S acc = ri[15] * t0;
            acc = acc + ri[16] * ((r1[k] * r1[l] + r2[k] * r2[l]) * r0[m] * r0[n]);
            acc = acc + ri[17] * ((r1[m] * r1[n] + r2[m] * r2[n]) * (r0[k] * r0[l]));
            acc = acc + ri[18] * (r1[k] * r1[l] * r1[m] * r1[n] + r2[k] * r2[l] * r2[m] * r2[n]);
            S mix1 = r0[m] * (r1[l] * r1[n] + r2[l] * r2[n]);
            S mix2 = r0[n] * (r1[l] * r1[m] + r2[l] * r2[m]);
            S val5 = r0[k] * (mix1 + mix2) + r0[l] * (r0[m] * (r1[k] * r1[n] + r2[k] * r2[n]) + r0[n] * (r1[k] * r1[m] + r2[k] * r2[m]));
            acc = acc + ri[19] * val5;
            acc = acc + ri[20] * (r1[k] * r1[l] * r2[m] * r2[n] + r2[k] * r2[l] * r1[m] * r1[n]);
            S cross = (r1[k] * r2[l] + r2[k] * r1[l]) * (r1[m] * r2[n] + r2[m] * r1[n]);
            acc = acc + ri[21] * cross;
            v = acc;
          }
          w[idx++] = v;
        }
}

// Electron-core attraction of the AOs on `orb` to the core of `core`: src/integrals.rs:234-271.
template <class S> static void core_attraction(const Elem& orb, const Elem& core, const S dvec[3], S blk[4][4]) {
  using O = Ops<S>;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) blk[a][b] = O::cst(0.0);
  S r = O::sqrt_(dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2]);
  double core_rho = core.po[9];
  double s = orb.po[1] + core_rho;
  blk[0][0] = O::recip(O::sqrt_(r * r + s * s)) * (-core.tore * PM3_EV);
  if (orb.z < 3 || !orb.has_p()) return;
  S inv = O::recip(r);
  S x0 = dvec[0] * inv, x1 = dvec[1] * inv, x2 = dvec[2] * inv;
  M3<S> R = rotation_to_x<S>(-x0, -x1, -x2);
  Centre oc{orb.ddp[2], orb.ddp[3] / std::sqrt(2.0), orb.po[1], orb.po[2], orb.po[3]};
  auto local = local_xh<S>(oc, core_rho, r);
  S packed[10];
  build_wxh<S>(local, R, packed);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) blk[a][b] = packed[pack(a, b)] * (-core.tore);
}

template <class S> struct PairInts {
  int na, nb, pa, pb;
  S w[100];
  S e1b[4][4];
  S e2a[4][4];
};

// src/integrals.rs:167-230 (ei is the heavy atom of a mixed pair).
template <class S> static void pair_two_electron(const Elem& ei, const Elem& ej, const S dvec[3], PairInts<S>& out) {
  using O = Ops<S>;
  S r = O::sqrt_(dvec[0] * dvec[0] + dvec[1] * dvec[1] + dvec[2] * dvec[2]);
  S inv = O::recip(r);
  S x0 = dvec[0] * inv, x1 = dvec[1] * inv, x2 = dvec[2] * inv;
  core_attraction<S>(ei, ej, dvec, out.e1b);
  S mdvec[3] = {-dvec[0], -dvec[1], -dvec[2]};
  core_attraction<S>(ej, ei, mdvec, out.e2a);
  M3<S> R = rotation_to_x<S>(-x0, -x1, -x2);
  bool hi = ei.has_p(), hj = ej.has_p();
  if (!hi && !hj) {
    double s = ei.rho0 + ej.rho0;
    out.na = out.nb = out.pa = out.pb = 1;
    out.w[0] = O::recip(O::sqrt_(r * r + s * s)) * PM3_EV;
    return;
  }
  if (hi && !hj) {
    auto ri = local_xh<S>(centre_ee(ei), ej.rho0, r);
    out.na = 4; out.nb = 1; out.pa = 10; out.pb = 1;
    build_wxh<S>(ri, R, out.w);
    return;
  }
  auto ri = local_xx<S>(centre_ee(ei), centre_ee(ej), r);
  out.na = out.nb = 4; out.pa = out.pb = 10;
  rotate_xx<S>(ri, R, out.w);
}

// ---------------------------------------------------------------------------------------
// Overlap: src/overlap.rs:103-127 (f64), :133-159 (dual), :266-475 (locals), :480-507
// (rotation to the molecular frame), :763-850 (A and B auxiliary integrals).
template <class S> static void aintgs(S x, S* a) {  // src/overlap.rs:763-774 (A_0..A_6)
  using O = Ops<S>;
  for (int i = 0; i < 7; ++i) a[i] = O::cst(0.0);
  if (std::fabs(O::val(x)) < 1.0e-12) return;
  S inv = O::recip(x);
  a[0] = O::exp_(-x) * inv;
  for (int i = 1; i < 7; ++i) a[i] = a[0] + a[i - 1] * inv * (double)i;
}

template <class S> static void bintgs(S x, S* b) {  // src/overlap.rs:819-850 (B_0..B_6)
  using O = Ops<S>;
  for (int i = 0; i < 7; ++i) b[i] = O::cst(0.0);
  if (std::fabs(O::val(x)) > 3.0) {
    S inv = O::recip(x);
    S tx = O::exp_(x) * inv;
    S tmx = O::exp_(-x) * inv * (-1.0);
    b[0] = tx + tmx;
    for (int i = 1; i < 7; ++i) {
      double sign = (i % 2 == 0) ? 1.0 : -1.0;
      b[i] = tx * sign + tmx + b[i - 1] * inv * (double)i;
    }
  } else {
    S term = O::cst(1.0);
    for (int m = 0; m < 40; ++m) {
      for (int k = 0; k < 7; ++k)
        if ((k + m) % 2 == 0) b[k] = b[k] + term * (2.0 / (double)(k + m + 1));
      term = term * (-x) * (1.0 / (double)(m + 1));
    }
  }
}

// Returns false when the element pair has no closed-form kernel (n >= 4, He, Ne).
template <class S> static bool overlap_locals(const Elem& ea, const Elem& eb, S r, S s[5]) {
  using O = Ops<S>;
  for (int i = 0; i < 5; ++i) s[i] = O::cst(0.0);
  int jcall = 0;
  int na = ea.n_s, nb = eb.n_s;
  if (na == 1 && nb == 1) jcall = 2;
  else if (na == 2 && nb == 1) jcall = 3;
  else if (na == 2 && nb == 2) jcall = 4;
  else if (na == 3 && nb == 1) jcall = 431;
  else if (na == 3 && nb == 2) jcall = 5;
  else if (na == 3 && nb == 3) jcall = 6;
  if (jcall == 0) return false;
  double za[2] = {ea.zs, ea.zp}, zb[2] = {eb.zs, eb.zp};
  double sums[4] = {za[0] + zb[0], za[1] + zb[0], za[0] + zb[1], za[1] + zb[1]};
  double smallest = INFINITY;
  for (double v : sums)
    if (v > 0.0) smallest = std::min(smallest, v);
  smallest *= 0.5;
  if (std::isfinite(smallest) && O::val(r) * smallest > 300.0) return true;
  S a111[7], b111[7], a211[7], b211[7], a121[7], b121[7], a22[7], b22[7];
  aintgs<S>(r * (0.5 * (za[0] + zb[0])), a111);
  bintgs<S>(r * (0.5 * (za[0] - zb[0])), b111);
  aintgs<S>(r * (0.5 * (za[1] + zb[0])), a211);
  bintgs<S>(r * (0.5 * (za[1] - zb[0])), b211);
  aintgs<S>(r * (0.5 * (za[0] + zb[1])), a121);
  bintgs<S>(r * (0.5 * (za[0] - zb[1])), b121);
  aintgs<S>(r * (0.5 * (za[1] + zb[1])), a22);
  bintgs<S>(r * (0.5 * (za[1] - zb[1])), b22);
  const double sq3 = std::sqrt(3.0);
  S s111 = O::cst(0.0), s211 = O::cst(0.0), s121 = O::cst(0.0), s221 = O::cst(0.0), s222 = O::cst(0.0);
  switch (jcall) {
    case 2:
      s111 = O::powf_(r * r * (za[0] * zb[0]), 1.5) * (a111[2] * b111[0] - b111[2] * a111[0]) / 4.0;
      break;
    case 3:
      s111 = O::powi_(r, 4) * (std::pow(zb[0], 1.5) * std::pow(za[0], 2.5)) *
             (a111[3] * b111[0] - b111[3] * a111[0] + a111[2] * b111[1] - b111[2] * a111[1]) / (sq3 * 8.0);
      s211 = O::powi_(r, 4) * (std::pow(zb[0], 1.5) * std::pow(za[1], 2.5)) *
             (a211[2] * b211[0] - b211[2] * a211[0] + a211[3] * b211[1] - b211[3] * a211[1]) / 8.0;
      break;
    case 4: {
      s111 = O::powi_(r, 5) * std::pow(zb[0] * za[0], 2.5) * (a111[4] * b111[0] + b111[4] * a111[0] - a111[2] * b111[2] * 2.0) / 48.0;
      s211 = O::powi_(r, 5) * std::pow(zb[0] * za[1], 2.5) *
             (a211[3] * (b211[0] - b211[2]) - a211[1] * (b211[2] - b211[4]) + b211[3] * (a211[0] - a211[2]) - b211[1] * (a211[2] - a211[4])) /
             (16.0 * sq3);
      s121 = O::powi_(r, 5) * std::pow(zb[1] * za[0], 2.5) *
             (a121[3] * (b121[0] - b121[2]) - a121[1] * (b121[2] - b121[4]) - b121[3] * (a121[0] - a121[2]) + b121[1] * (a121[2] - a121[4])) /
             (16.0 * sq3);
      S w = O::powi_(r, 5) * std::pow(zb[1] * za[1], 2.5) / 16.0;
      s221 = -(w * (b22[2] * (a22[4] + a22[0]) - a22[2] * (b22[4] + b22[0])));
      s222 = w * (a22[4] * (b22[0] - b22[2]) - b22[4] * (a22[0] - a22[2]) - a22[2] * b22[0] + b22[2] * a22[0]) * 0.5;
      break;
    }
    case 431:
      s111 = O::powi_(r, 5) * (std::pow(zb[0], 1.5) * std::pow(za[0], 3.5)) *
             (a111[4] * b111[0] + b111[1] * a111[3] * 2.0 - a111[1] * b111[3] * 2.0 - b111[4] * a111[0]) / (std::sqrt(10.0) * 24.0);
      s211 = O::powi_(r, 5) * (std::pow(zb[0], 1.5) * std::pow(za[1], 3.5)) *
             (a211[3] * (b211[0] + b211[2]) - a211[1] * (b211[4] + b211[2]) + b211[1] * (a211[2] + a211[4]) - b211[3] * (a211[2] + a211[0])) /
             (8.0 * std::sqrt(30.0));
      break;
    case 5:
      s111 = O::powi_(r, 6) * (std::pow(zb[0], 2.5) * std::pow(za[0], 3.5)) *
             (a111[5] * b111[0] + b111[1] * a111[4] - b111[2] * a111[3] * 2.0 - a111[2] * b111[3] * 2.0 + b111[4] * a111[1] + b111[5] * a111[0]) /
             (std::sqrt(30.0) * 48.0);
      s211 = O::powi_(r, 6) * (std::pow(zb[0], 2.5) * std::pow(za[1], 3.5)) *
             (a211[4] * b211[0] + b211[1] * a211[5] - b211[3] * a211[3] * 2.0 - a211[2] * b211[2] * 2.0 + a211[1] * b211[5] + a211[0] * b211[4]) /
             (48.0 * std::sqrt(10.0));
      s121 = O::powi_(r, 6) * (std::pow(zb[1], 2.5) * std::pow(za[0], 3.5)) *
             ((a121[4] * b121[0] - a121[5] * b121[1]) + (a121[3] * b121[1] - a121[4] * b121[2]) * 2.0 -
              (a121[1] * b121[3] - a121[2] * b121[4]) * 2.0 - (a121[0] * b121[4] - a121[1] * b121[5])) /
             (48.0 * std::sqrt(10.0));
      s221 = O::powi_(r, 6) * (std::pow(zb[1], 2.5) * std::pow(za[1], 3.5)) *
             ((a22[3] * b22[0] - a22[5] * b22[2]) + (a22[2] * b22[1] - a22[4] * b22[3]) - (a22[1] * b22[2] - a22[3] * b22[4]) -
              (a22[0] * b22[3] - a22[2] * b22[5])) /
             (16.0 * std::sqrt(30.0));
      s222 = O::powi_(r, 6) * (std::pow(zb[1], 2.5) * std::pow(za[1], 3.5)) *
             ((a22[5] - a22[3]) * (b22[0] - b22[2]) + (a22[4] - a22[2]) * (b22[1] - b22[3]) - (a22[3] - a22[1]) * (b22[2] - b22[4]) -
              (a22[2] - a22[0]) * (b22[3] - b22[5])) /
             (32.0 * std::sqrt(30.0));
      break;
    case 6:
      s111 = O::powi_(r, 7) * std::pow(zb[0] * za[0], 3.5) *
             (a111[6] * b111[0] - b111[2] * a111[4] * 3.0 + a111[2] * b111[4] * 3.0 - a111[0] * b111[6]) / 1440.0;
      s211 = O::powi_(r, 7) * std::pow(zb[0] * za[1], 3.5) *
             ((a211[5] * b211[0] + a211[6] * b211[1]) + (-a211[4] * b211[1] - a211[5] * b211[2]) - (a211[3] * b211[2] + a211[4] * b211[3]) * 2.0 -
              (-a211[2] * b211[3] - a211[3] * b211[4]) * 2.0 + (a211[1] * b211[4] + a211[2] * b211[5]) + (-a211[0] * b211[5] - a211[1] * b211[6])) /
             (480.0 * sq3);
      s121 = O::powi_(r, 7) * std::pow(zb[1] * za[0], 3.5) *
             ((a121[5] * b121[0] - a121[6] * b121[1]) + (a121[4] * b121[1] - a121[5] * b121[2]) - (a121[3] * b121[2] - a121[4] * b121[3]) * 2.0 -
              (a121[2] * b121[3] - a121[3] * b121[4]) * 2.0 + (a121[1] * b121[4] - a121[2] * b121[5]) + (a121[0] * b121[5] - a121[1] * b121[6])) /
             (480.0 * sq3);
      s221 = O::powi_(r, 7) * (std::pow(zb[1], 3.5) * std::pow(za[1], 3.5)) *
             ((a22[4] * b22[0] - a22[6] * b22[2]) - (a22[2] * b22[2] - a22[4] * b22[4]) * 2.0 + (a22[0] * b22[4] - a22[2] * b22[6])) / 480.0;
      s222 = O::powi_(r, 7) * (std::pow(zb[1], 3.5) * std::pow(za[1], 3.5)) *
             ((a22[6] - a22[4]) * (b22[0] - b22[2]) - (a22[4] - a22[2]) * (b22[2] - b22[4]) * 2.0 + (a22[2] - a22[0]) * (b22[4] - b22[6])) / 960.0;
      break;
  }
  s[0] = s111; s[1] = s211; s[2] = s121; s[3] = s221; s[4] = s222;
  return true;
}

template <class S> static void build_di(const S s[5], const S dir[3], S di[4][4]) {  // src/overlap.rs:480-507
  using O = Ops<S>;
  M3<S> rr = rotation_to_x<S>(dir[0], dir[1], dir[2]);
  S ro[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) ro[a][b] = rr[b][a];
  di[0][0] = s[0];
  for (int a = 0; a < 3; ++a) {
    di[1 + a][0] = s[1] * ro[a][0];
    di[0][1 + a] = -s[2] * ro[a][0];
  }
  S mm[3] = {-s[3], s[4], s[4]};
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      S v = O::cst(0.0);
      for (int k = 0; k < 3; ++k) v = v + ro[a][k] * mm[k] * ro[b][k];
      di[1 + a][1 + b] = v;
    }
}

// f64 diatomic overlap, src/overlap.rs:103-127.
static bool diatom_overlap(const Elem& ei, const double* pi, const Elem& ej, const double* pj, double S_[4][4]) {
  double d[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
  double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) S_[a][b] = (a == b && r < 1.0e-8) ? 1.0 : 0.0;
  if (r < 1.0e-8) return true;
  double x[3] = {d[0] / r, d[1] / r, d[2] / r};
  double loc[5], dir[3], di[4][4];
  bool swap = ei.n_s < ej.n_s;
  bool ok;
  if (!swap) {
    ok = overlap_locals<double>(ei, ej, r, loc);
    dir[0] = x[0]; dir[1] = x[1]; dir[2] = x[2];
  } else {
    ok = overlap_locals<double>(ej, ei, r, loc);
    dir[0] = x[0] * -1.0; dir[1] = x[1] * -1.0; dir[2] = x[2] * -1.0;
  }
  if (!ok) return false;
  build_di<double>(loc, dir, di);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) S_[a][b] = swap ? di[b][a] : di[a][b];
  return true;
}

// Dual diatomic overlap seeded on R_j - R_i, src/overlap.rs:133-159.
static bool diatom_overlap_dual(const Elem& ei, const double* pi, const Elem& ej, const double* pj, Dual S_[4][4]) {
  double d[3] = {pj[0] - pi[0], pj[1] - pi[1], pj[2] - pi[2]};
  bool swap = ei.n_s < ej.n_s;
  const Elem& ea = swap ? ej : ei;
  const Elem& eb = swap ? ei : ej;
  Dual dv[3] = {dmk(d[0], 1, 0, 0), dmk(d[1], 0, 1, 0), dmk(d[2], 0, 0, 1)};
  Dual r = Ops<Dual>::sqrt_(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
  Dual inv = Ops<Dual>::recip(r);
  double sg = swap ? -1.0 : 1.0;
  Dual dir[3] = {dv[0] * inv * sg, dv[1] * inv * sg, dv[2] * inv * sg};
  Dual loc[5], di[4][4];
  if (!overlap_locals<Dual>(ea, eb, r, loc)) return false;
  build_di<Dual>(loc, dir, di);
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) S_[a][b] = swap ? di[b][a] : di[a][b];
  return true;
}

// ---------------------------------------------------------------------------------------
// Core-core repulsion, src/repulsion.rs:57-120.
template <class S> static S pair_core_energy(const Elem& ei, const Elem& ej, S rb) {
  using O = Ops<S>;
  S ra = rb * PM3_A0;
  double rho = ei.po[9] + ej.po[9];
  S gab = O::recip(O::sqrt_(rb * rb + rho * rho)) * PM3_EV;
  double cp = ei.tore * ej.tore;
  S bare = gab * cp;
  const Table& T = table();
  auto it = T.pair.find({std::max(ei.z, ej.z), std::min(ei.z, ej.z)});
  bool has_pair = it != T.pair.end() && std::fabs(it->second.second) > 1.0e-5;
  S e;
  if (has_pair) {
    double alpha = it->second.first < 1.0e-6 ? 1.2 : it->second.first;
    S sc = O::cst(1.0) + O::exp_(ra * -alpha) * (2.0 * it->second.second);
    e = bare * sc;
  } else {
    S exi = O::exp_(ra * -ei.alp), exj = O::exp_(ra * -ej.alp);
    S sc = exi + exj;
    int zs = ei.z + ej.z;
    if (zs == 8 || zs == 9) {
      if (ei.z == 7 || ei.z == 8) sc = sc + (ra - 1.0) * exi;
      if (ej.z == 7 || ej.z == 8) sc = sc + (ra - 1.0) * exj;
    }
    e = O::abs_(sc * bare) + bare;
  }
  if (!has_pair) {
    std::vector<std::array<double, 3>> gs = ei.gauss;
    gs.insert(gs.end(), ej.gauss.begin(), ej.gauss.end());
    for (auto& g : gs) {
      S delta = ra - g[2];
      S ax = delta * delta * g[1];
      if (O::val(ax) <= 25.0) e = e + O::exp_(-ax) * (cp * g[0]) / ra;
    }
  }
  return e;
}

// ---------------------------------------------------------------------------------------
// Molecule, basis and H_core: src/basis.rs:39-62, src/hamiltonian.rs:191-344.
struct Mol {
  int nat = 0;
  std::vector<int> z;
  std::vector<double> xyz;  // Bohr, nat x 3
  std::vector<const Elem*> el;
  std::vector<int> off, norb;
  int nao = 0;
};

enum Status { OK = 0, E_INVALID = 1, E_MISSING_ELEMENT = 2, E_NOT_CONVERGED = 3, E_UNSUPPORTED = 4, E_RESOURCE = 5 };

static int make_mol(int nat, const int* z, const double* xyz, Mol& m) {
  const Table& T = table();
  m.nat = nat;
  m.z.assign(z, z + nat);
  m.xyz.assign(xyz, xyz + 3 * nat);
  m.el.resize(nat);
  m.off.resize(nat);
  m.norb.resize(nat);
  m.nao = 0;
  for (int i = 0; i < nat; ++i) {
    auto it = T.el.find(z[i]);
    if (it == T.el.end()) return E_MISSING_ELEMENT;
    m.el[i] = &it->second;
    m.off[i] = m.nao;
    m.norb[i] = it->second.n_orb;
    m.nao += it->second.n_orb;
  }
  return OK;
}

struct Pair {
  int a, b;  // ordered: heavy first for a mixed pair
  PairInts<double> t;
};

struct Core {
  std::vector<double> h;  // nao x nao
  std::vector<Pair> pairs;
};

static double beta_of(const Elem& e, int orb) { return orb == 0 ? e.betas : e.betap; }

// Bytes of the reference's resident pair cache (pair_cache_bytes, src/hamiltonian.rs:51-71):
// sum_{i<j} p_i p_j doubles of packed W plus one PairIntegral record per atom pair.  The record is
// two usize indices and a PackedTwoElec of four usize dimensions and a Vec<f64> (pointer,
// capacity, length): 72 bytes on a 64-bit target.
static double pair_cache_bytes(const Mol& m) {
  double total = 0.0, squares = 0.0;
  for (int i = 0; i < m.nat; ++i) {
    const double p = std::max(m.norb[i] * (m.norb[i] + 1) / 2, 1);
    total += p;
    squares += p * p;
  }
  const double elements = (total * total - squares) / 2.0;
  const double pairs = (double)m.nat * (double)(m.nat - 1) / 2.0;
  return elements * 8.0 + pairs * 72.0;
}

// The external field's one-electron operator +M.f (field_terms, src/dipole.rs:156-181, with
// dipole_matrix :93-125 at the coordinate origin): every AO of atom A gets f.R_A on the diagonal,
// the s-p_alpha element gets f_alpha dd_A; added after U and before the pair terms
// (src/hamiltonian.rs:224-233).  Returns the nuclear half -sum_A Z_A R_A.f.
static double add_field(const Mol& m, const double* f, std::vector<double>& h) {
  const int n = m.nao;
  double nuclear = 0.0;
  for (int i = 0; i < m.nat; ++i) {
    const Elem& e = *m.el[i];
    const int o = m.off[i];
    const double* R = &m.xyz[3 * i];
    for (int ax = 0; ax < 3; ++ax) {
      if (f[ax] == 0.0) continue;
      for (int mu = 0; mu < m.norb[i]; ++mu) h[(size_t)(o + mu) * n + o + mu] += f[ax] * R[ax];
      if (m.norb[i] >= 4) {
        h[(size_t)o * n + o + ax + 1] += f[ax] * e.dd;
        h[(size_t)(o + ax + 1) * n + o] += f[ax] * e.dd;
      }
    }
    nuclear -= e.tore * (R[0] * f[0] + R[1] * f[1] + R[2] * f[2]);
  }
  return nuclear;
}

static int build_core(const Mol& m, Core& c, const double* field = nullptr, double* field_nuclear = nullptr) {
  int n = m.nao;
  c.h.assign((size_t)n * n, 0.0);
  c.pairs.clear();
  c.pairs.reserve((size_t)m.nat * (m.nat - 1) / 2);
  for (int i = 0; i < m.nat; ++i) {
    const Elem& e = *m.el[i];
    for (int k = 0; k < m.norb[i]; ++k) c.h[(size_t)(m.off[i] + k) * n + m.off[i] + k] = k == 0 ? e.uss : e.upp;
  }
  if (field) {
    const double nuc = add_field(m, field, c.h);
    if (field_nuclear) *field_nuclear = nuc;
  } else if (field_nuclear) {
    *field_nuclear = 0.0;
  }
  for (int u = 0; u < m.nat; ++u)
    for (int v = u + 1; v < m.nat; ++v) {
      const Elem& eu = *m.el[u];
      const Elem& ev = *m.el[v];
      int a = u, b = v;
      if (!(eu.has_p() || !ev.has_p())) { a = v; b = u; }
      const Elem& ea = *m.el[a];
      const Elem& eb = *m.el[b];
      const double* pa = &m.xyz[3 * a];
      const double* pb = &m.xyz[3 * b];
      double d[3] = {pb[0] - pa[0], pb[1] - pa[1], pb[2] - pa[2]};
      double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      if (!(r > 1.0e-8)) return E_INVALID;
      double x[3] = {d[0] / r, d[1] / r, d[2] / r};
      double dv[3] = {x[0] * r, x[1] * r, x[2] * r};
      Pair p;
      p.a = a; p.b = b;
      pair_two_electron<double>(ea, eb, dv, p.t);
      double S_[4][4];
      if (!diatom_overlap(ea, pa, eb, pb, S_)) return E_UNSUPPORTED;
      int oa = m.off[a], ob = m.off[b], na = m.norb[a], nb = m.norb[b];
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < na; ++j) c.h[(size_t)(oa + i) * n + oa + j] += p.t.e1b[i][j];
      for (int i = 0; i < nb; ++i)
        for (int j = 0; j < nb; ++j) c.h[(size_t)(ob + i) * n + ob + j] += p.t.e2a[i][j];
      for (int i = 0; i < na; ++i) {
        double bi = beta_of(ea, i);
        for (int j = 0; j < nb; ++j) {
          double bj = beta_of(eb, j);
          double val = 0.5 * (bi + bj) * S_[i][j];
          c.h[(size_t)(oa + i) * n + ob + j] = val;
          c.h[(size_t)(ob + j) * n + oa + i] = val;
        }
      }
      c.pairs.push_back(p);
    }
  return OK;
}

// ---------------------------------------------------------------------------------------
// Fock build, src/fock.rs:93-193 (packed two-centre), :244-332 (one-centre), :488-500 (RHF).
static double oc_two_electron(int a, int b, int c, int d, const Elem& e) {  // src/fock.rs:244-282
  if (a == b && c == d) {
    if (a == 0 && c == 0) return e.gss;
    if (a == 0 || c == 0) return e.gsp;
    return a == c ? e.gpp : e.gp2;
  }
  int ba = std::min(a, b), bb = std::max(a, b), kc = std::min(c, d), kd = std::max(c, d);
  if (ba == 0 && bb != 0 && kc == 0 && kd != 0 && bb == kd) return e.hsp;
  if (ba != 0 && bb != 0 && ba != bb && ba == kc && bb == kd) return 0.5 * (e.gpp - e.gp2);
  return 0.0;
}

// F = H_core + J(P) - K(Ps): build_fock_spin (src/fock.rs:285-332), Coulomb from the total
// density P, exchange from the same-spin density Ps.
static void build_fock_spin(const Mol& m, const Core& c, const double* P, const double* Ps, double* F) {
  int n = m.nao;
  std::memcpy(F, c.h.data(), sizeof(double) * (size_t)n * n);
  for (int ia = 0; ia < m.nat; ++ia) {
    const Elem& e = *m.el[ia];
    int o = m.off[ia], nn = m.norb[ia];
    for (int mu = 0; mu < nn; ++mu)
      for (int nu = 0; nu < nn; ++nu) {
        double acc = 0.0;
        for (int la = 0; la < nn; ++la)
          for (int si = 0; si < nn; ++si) {
            acc += P[(size_t)(o + la) * n + o + si] * oc_two_electron(mu, nu, la, si, e);
            acc -= Ps[(size_t)(o + la) * n + o + si] * oc_two_electron(mu, la, nu, si, e);
          }
        F[(size_t)(o + mu) * n + o + nu] += acc;
      }
  }
  for (const Pair& p : c.pairs) {
    const auto& t = p.t;
    int oa = m.off[p.a], ob = m.off[p.b], na = t.na, nb = t.nb, pa = t.pa, pb = t.pb;
    double da[10], db[10], ja[10] = {0}, jb[10] = {0}, kab[4][4] = {{0}};
    for (int mu = 0; mu < na; ++mu) {
      da[pack(mu, mu)] = P[(size_t)(oa + mu) * n + oa + mu];
      for (int nu = 0; nu < mu; ++nu) da[pack(mu, nu)] = 2.0 * P[(size_t)(oa + mu) * n + oa + nu];
    }
    for (int mu = 0; mu < nb; ++mu) {
      db[pack(mu, mu)] = P[(size_t)(ob + mu) * n + ob + mu];
      for (int nu = 0; nu < mu; ++nu) db[pack(mu, nu)] = 2.0 * P[(size_t)(ob + mu) * n + ob + nu];
    }
    for (int q = 0; q < pa; ++q) {
      double acc = 0.0;
      for (int s = 0; s < pb; ++s) acc += db[s] * t.w[q * pb + s];
      ja[q] = acc;
      for (int s = 0; s < pb; ++s) jb[s] += da[q] * t.w[q * pb + s];
    }
    for (int mu = 0; mu < na; ++mu)
      for (int nu = 0; nu < na; ++nu) {
        const double* row = &t.w[pack(mu, nu) * pb];
        for (int la = 0; la < nb; ++la) {
          double acc = 0.0;
          for (int si = 0; si < nb; ++si) acc += Ps[(size_t)(oa + nu) * n + ob + si] * row[pack(la, si)];
          kab[mu][la] -= acc;
        }
      }
    for (int mu = 0; mu < na; ++mu)
      for (int nu = 0; nu < na; ++nu) F[(size_t)(oa + mu) * n + oa + nu] += ja[pack(mu, nu)];
    for (int la = 0; la < nb; ++la)
      for (int si = 0; si < nb; ++si) F[(size_t)(ob + la) * n + ob + si] += jb[pack(la, si)];
    for (int mu = 0; mu < na; ++mu)
      for (int la = 0; la < nb; ++la) {
        double v = F[(size_t)(oa + mu) * n + ob + la] + kab[mu][la];
        F[(size_t)(oa + mu) * n + ob + la] = v;
        F[(size_t)(ob + la) * n + oa + mu] = v;
      }
  }
}

// RHF Fock, build_fock (src/fock.rs:488-500): the spin build with Ps = P / 2.
static void build_fock(const Mol& m, const Core& c, const double* P, double* F) {
  std::vector<double> half((size_t)m.nao * m.nao);
  for (size_t k = 0; k < half.size(); ++k) half[k] = 0.5 * P[k];
  build_fock_spin(m, c, P, half.data(), F);
}

// ---------------------------------------------------------------------------------------
// Dense helpers (src/linalg.rs:65-180).
static double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

static void matmul(int n, const double* A, const double* B, double* C) {
  std::fill(C, C + (size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) {
      double a = A[(size_t)i * n + k];
      if (a == 0.0) continue;
      const double* b = B + (size_t)k * n;
      double* c = C + (size_t)i * n;
      for (int j = 0; j < n; ++j) c[j] += a * b[j];
    }
}

// Symmetric eigensolver: Householder tridiagonalisation with accumulated reflections, then
// implicit-shift QL with Wilkinson shifts (Golub & Van Loan Alg. 8.3.1 and Sec. 8.3.5).
// In: A (n x n row-major, symmetric).  Out: w ascending, V (row-major) with eigenvectors as
// columns, V[i*n+k] = component i of eigenvector k.  Contract: src/linalg.rs:217-263.
static void sym_eigen(int n, const double* A, double* w, double* V) {
  std::vector<double> a(A, A + (size_t)n * n), d(n), e(n, 0.0);
  // Column-major view: work on z = a with z[i][j] addressed as a[i*n+j].
  auto Z = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
  // Householder reduction (rows from the bottom), accumulating the transformation.
  for (int i = n - 1; i > 0; --i) {
    int l = i - 1;
    double h = 0.0;
    if (l > 0) {
      double scale = 0.0;
      for (int k = 0; k <= l; ++k) scale += std::fabs(Z(i, k));
      if (scale == 0.0) {
        e[i] = Z(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          Z(i, k) /= scale;
          h += Z(i, k) * Z(i, k);
        }
        double f = Z(i, l);
        double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        Z(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          Z(j, i) = Z(i, j) / h;
          g = 0.0;
          for (int k = 0; k <= j; ++k) g += Z(j, k) * Z(i, k);
          for (int k = j + 1; k <= l; ++k) g += Z(k, j) * Z(i, k);
          e[j] = g / h;
          f += e[j] * Z(i, j);
        }
        double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          f = Z(i, j);
          e[j] = g = e[j] - hh * f;
          for (int k = 0; k <= j; ++k) Z(j, k) -= (f * e[k] + g * Z(i, k));
        }
      }
    } else {
      e[i] = Z(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    int l = i - 1;
    if (d[i] != 0.0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0.0;
        for (int k = 0; k <= l; ++k) g += Z(i, k) * Z(k, j);
        for (int k = 0; k <= l; ++k) Z(k, j) -= g * Z(k, i);
      }
    }
    d[i] = Z(i, i);
    Z(i, i) = 1.0;
    for (int j = 0; j <= l; ++j) Z(j, i) = Z(i, j) = 0.0;
  }
  // Implicit QL on the tridiagonal (d, e), rotations applied to the columns of Z.
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  if (n > 0) e[n - 1] = 0.0;
  for (int l = 0; l < n; ++l) {
    int iter = 0;
    int m;
    do {
      for (m = l; m < n - 1; ++m) {
        double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= 1.0e-16 * dd) break;
      }
      if (m != l) {
        if (++iter > 60) break;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = std::hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i], b = c * e[i];
          e[i + 1] = (r = std::hypot(f, g));
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g = c * r - b;
          for (int k = 0; k < n; ++k) {
            f = Z(k, i + 1);
            Z(k, i + 1) = s * Z(k, i) + c * f;
            Z(k, i) = c * Z(k, i) - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  std::vector<int> order(n);
  for (int i = 0; i < n; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return d[x] < d[y]; });
  for (int k = 0; k < n; ++k) {
    w[k] = d[order[k]];
    for (int i = 0; i < n; ++i) V[(size_t)i * n + k] = Z(i, order[k]);
  }
}

// ---------------------------------------------------------------------------------------
// SCF: src/scf.rs:1179-1316 (rhf_loop), :1065-1173 (AccelHistory), :1501-1693 (solvers),
// :309-327 (depth rule), :855-935 (guesses), :982-992 (density), :1039-1054 (commutator).
struct Options {
  int max_scf;
  double e_tol;
  double p_tol;
  int use_diis;
  int accelerator;  // 0 none, 1 CDIIS, 2 A-DIIS -> CDIIS
  double adiis_switch;
  int guess;        // 0 auto, 1 SAD, 2 core, 3 symmetry-broken
  int stability;    // 0 off, 1 auto, 2 always
  int scf_memory_mb;
  double level_shift_ev;   // src/scf.rs:193-198: read by uhf_loop; non-zero (or damping) steers the first path
  double damping;          // src/scf.rs:199-202
  int reference;           // 0 auto, 1 RHF, 2 UHF (src/scf.rs:508-520)
  int has_field;           // field is Some (src/scf.rs:248-256)
  double field[3];         // eV per Bohr per elementary charge, referenced to the coordinate origin
  int integral_memory_mb;  // pair-cache ceiling in MiB; 0 takes the process default (src/scf.rs:289-297)
};

struct Result {
  double total_ev, electronic_ev, core_ev, hof_kcal;
  double dipole[3], dipole_debye;
  double homo_ev, lumo_ev;
  double last_change, scf_improvement_ev;
  int iterations, converged, scf_paths_tried, n_occ, nao, has_homo, has_lumo;
  int unrestricted, n_beta;
  double required_mb, limit_mb;  // ResourceLimit details (src/error.rs:12-47)
};

static int diis_depth(int nao, int mats, int budget_mb) {
  if (budget_mb == 0) return 8;
  size_t per = (size_t)mats * nao * nao * 8;
  if (per == 0) return 8;
  size_t budget = (size_t)budget_mb * 1024 * 1024;
  size_t d = budget / per;
  return (int)std::min<size_t>(std::max<size_t>(d, 2), 8);
}

static bool solve_bordered_small(std::vector<std::vector<double>> m, std::vector<double> rhs, std::vector<double>& x) {
  int n = (int)m.size();
  for (int col = 0; col < n; ++col) {
    int piv = col;
    double best = std::fabs(m[col][col]);
    for (int row = col + 1; row < n; ++row)
      if (std::fabs(m[row][col]) > best) { best = std::fabs(m[row][col]); piv = row; }
    if (best < 1.0e-12) return false;
    if (piv != col) { std::swap(m[col], m[piv]); std::swap(rhs[col], rhs[piv]); }
    for (int row = col + 1; row < n; ++row) {
      double f = m[row][col] / m[col][col];
      if (f == 0.0) continue;
      for (int j = col; j < n; ++j) m[row][j] -= f * m[col][j];
      rhs[row] -= f * rhs[col];
    }
  }
  x.assign(n, 0.0);
  for (int col = n - 1; col >= 0; --col) {
    double s = rhs[col];
    for (int j = col + 1; j < n; ++j) s -= m[col][j] * x[j];
    x[col] = s / m[col][col];
  }
  return true;
}

static bool diis_coeffs(const std::vector<std::vector<double>>& gram, std::vector<double>& c) {
  int n = (int)gram.size();
  if (n < 2) return false;
  double scale = 0.0;
  for (auto& r : gram)
    for (double v : r) scale = std::max(scale, std::fabs(v));
  if (!std::isfinite(scale) || scale <= 0.0) return false;
  int dim = n + 1;
  std::vector<std::vector<double>> b(dim, std::vector<double>(dim, 0.0));
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) b[i][j] = gram[i][j] / scale;
    b[i][n] = -1.0;
    b[n][i] = -1.0;
  }
  std::vector<double> rhs(dim, 0.0), x;
  rhs[n] = -1.0;
  if (!solve_bordered_small(b, rhs, x)) return false;
  c.assign(x.begin(), x.begin() + n);
  return true;
}

static std::vector<double> simplex_project(const std::vector<double>& v) {  // src/scf.rs:1677-1693
  std::vector<double> u = v;
  std::sort(u.begin(), u.end(), [](double a, double b) { return a > b; });
  double css = 0.0, theta = 0.0;
  for (size_t j = 0; j < u.size(); ++j) {
    css += u[j];
    double t = (css - 1.0) / (double)(j + 1);
    if (u[j] - t > 0.0) theta = t;
  }
  std::vector<double> out(v.size());
  for (size_t i = 0; i < v.size(); ++i) out[i] = std::max(v[i] - theta, 0.0);
  return out;
}

static std::vector<double> solve_adiis(const std::vector<double>& d, const std::vector<std::vector<double>>& s) {
  int n = (int)d.size();
  double L = 1.0e-12;
  for (int i = 0; i < n; ++i) {
    double row = 0.0;
    for (int j = 0; j < n; ++j) row += std::fabs(s[i][j] + s[j][i]);
    L = std::max(L, row);
  }
  double lr = 1.0 / L;
  std::vector<double> c(n, 0.0);
  c[n - 1] = 1.0;
  for (int it = 0; it < 400; ++it) {
    std::vector<double> g(n), trial(n);
    for (int k = 0; k < n; ++k) {
      double acc = 2.0 * d[k];
      for (int j = 0; j < n; ++j) acc += (s[k][j] + s[j][k]) * c[j];
      g[k] = acc;
    }
    for (int i = 0; i < n; ++i) trial[i] = c[i] - lr * g[i];
    std::vector<double> proj = simplex_project(trial);
    double delta = 0.0;
    for (int i = 0; i < n; ++i) delta += std::fabs(proj[i] - c[i]);
    c = proj;
    if (delta < 1.0e-12) break;
  }
  return c;
}

struct History {
  int depth;
  bool keep_d;
  std::vector<std::vector<double>> F, E, D;
  std::vector<std::vector<double>> b, m;
  double push(std::vector<double> f, std::vector<double> e, const std::vector<double>* dens) {
    std::vector<double> brow;
    for (auto& ei : E) brow.push_back(dot(ei, e));
    double en = dot(e, e);
    brow.push_back(en);
    for (size_t i = 0; i < b.size(); ++i) b[i].push_back(brow[i]);
    b.push_back(brow);
    if (keep_d && dens) {
      std::vector<double> mrow;
      for (auto& fj : F) mrow.push_back(dot(*dens, fj));
      mrow.push_back(dot(*dens, f));
      for (size_t i = 0; i < m.size(); ++i) m[i].push_back(dot(D[i], f));
      m.push_back(mrow);
      D.push_back(*dens);
    }
    F.push_back(std::move(f));
    E.push_back(std::move(e));
    while ((int)F.size() > depth) {
      F.erase(F.begin());
      E.erase(E.begin());
      b.erase(b.begin());
      for (auto& r : b) r.erase(r.begin());
      if (keep_d) {
        D.erase(D.begin());
        m.erase(m.begin());
        for (auto& r : m) r.erase(r.begin());
      }
    }
    return en;
  }
  std::vector<double> combine(const std::vector<double>& c) const {
    std::vector<double> out(F[0].size(), 0.0);
    for (size_t i = 0; i < F.size(); ++i)
      for (size_t k = 0; k < out.size(); ++k) out[k] += c[i] * F[i][k];
    return out;
  }
  bool cdiis(std::vector<double>& out) const {
    std::vector<double> c;
    if (!diis_coeffs(b, c)) return false;
    out = combine(c);
    return true;
  }
  bool adiis(std::vector<double>& out) const {
    int n = (int)F.size();
    if (n < 2 || !keep_d) return false;
    int last = n - 1;
    double mnn = m[last][last];
    std::vector<double> d(n);
    std::vector<std::vector<double>> s(n, std::vector<double>(n));
    for (int i = 0; i < n; ++i) d[i] = m[i][last] - mnn;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) s[i][j] = m[i][j] - m[i][last] - m[last][j] + mnn;
    out = combine(solve_adiis(d, s));
    return true;
  }
};

static void sad_density(const Mol& m, std::vector<double>& P) {
  int n = m.nao;
  P.assign((size_t)n * n, 0.0);
  for (int i = 0; i < m.nat; ++i) {
    const Elem& e = *m.el[i];
    int o = m.off[i];
    P[(size_t)o * n + o] = e.ios;
    if (m.norb[i] >= 4)
      for (int k = 1; k < 4; ++k) P[(size_t)(o + k) * n + o + k] = e.iop / 3.0;
  }
}

static void density_from_coeff(int n, const double* C, int nocc, double w, double* P) {
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) {
      double s = 0.0;
      for (int k = 0; k < nocc; ++k) s += C[(size_t)i * n + k] * C[(size_t)j * n + k];
      P[(size_t)i * n + j] = w * s;
    }
}

struct State {
  std::vector<double> P, eps, C;
  double e_elec = 0, last_change = INFINITY;
  int iterations = 0;
  bool converged = false;
};

// P0, when given, is the starting density (n x n, row-major) in place of the guess rule.
static void rhf_loop(const Mol& m, const Core& core, int nocc, const Options& o, int guess, int accel_in, State& st,
                     const double* P0 = nullptr) {
  int n = m.nao;
  std::vector<double> P;
  int pre = 0;
  if (P0) {
    // a caller density (the learned guess) only builds the first Fock matrix; the SCF proper
    // starts from that Fock matrix's aufbau density, counted as one iteration
    std::vector<double> F0((size_t)n * n), w(n), C((size_t)n * n);
    build_fock(m, core, P0, F0.data());
    sym_eigen(n, F0.data(), w.data(), C.data());
    P.assign((size_t)n * n, 0.0);
    density_from_coeff(n, C.data(), nocc, 2.0, P.data());
    pre = 1;
  } else if (guess == 2) {
    std::vector<double> w(n), C((size_t)n * n);
    sym_eigen(n, core.h.data(), w.data(), C.data());
    P.assign((size_t)n * n, 0.0);
    density_from_coeff(n, C.data(), nocc, 2.0, P.data());
  } else {
    sad_density(m, P);
    if (guess == 3) {  // src/scf.rs:894-916
      int target = -1;
      double best = 0.0;
      for (int i = 0; i < m.nat; ++i)
        if (m.norb[i] >= 4 && m.el[i]->iop > best) { best = m.el[i]->iop; target = i; }
      if (target >= 0) {
        int off = m.off[target];
        double tilt = std::min(0.2, P[(size_t)(off + 3) * n + off + 3]);
        P[(size_t)(off + 3) * n + off + 3] -= tilt;
        P[(size_t)(off + 1) * n + off + 1] += tilt;
      }
    }
  }
  int accel = o.use_diis ? accel_in : 0;
  bool keep_d = accel == 2;
  size_t slot3 = (size_t)3 * n * n * 8;
  if (keep_d && !(o.scf_memory_mb == 0 || 2 * slot3 <= (size_t)o.scf_memory_mb * 1024 * 1024)) keep_d = false;
  if (accel == 2 && !keep_d) accel = 1;
  History h;
  h.depth = diis_depth(n, keep_d ? 3 : 2, o.scf_memory_mb);
  h.keep_d = keep_d;
  std::vector<double> F((size_t)n * n), FP((size_t)n * n), Pn((size_t)n * n), eps(n), C((size_t)n * n);
  double e_old = 0.0;
  st.converged = false;
  st.last_change = INFINITY;
  for (int it = 0; it < o.max_scf; ++it) {
    st.iterations = it + 1;
    build_fock(m, core, P.data(), F.data());
    double e_elec = 0.5 * (dot(P, core.h) + dot(P, F));
    std::vector<double> Fuse;
    if (accel == 0) {
      Fuse = F;
    } else {
      matmul(n, F.data(), P.data(), FP.data());
      std::vector<double> E(FP);
      for (int i = 0; i < n; ++i) {
        E[(size_t)i * n + i] = 0.0;
        for (int j = 0; j < i; ++j) {
          double up = FP[(size_t)j * n + i], lo = FP[(size_t)i * n + j];
          E[(size_t)i * n + j] = lo - up;
          E[(size_t)j * n + i] = up - lo;
        }
      }
      double en = std::sqrt(h.push(F, std::move(E), keep_d ? &P : nullptr));
      bool ok = (accel == 2 && en > o.adiis_switch) ? h.adiis(Fuse) : h.cdiis(Fuse);
      if (!ok) Fuse = h.F.back();
    }
    sym_eigen(n, Fuse.data(), eps.data(), C.data());
    density_from_coeff(n, C.data(), nocc, 2.0, Pn.data());
    double s = 0.0;
    for (size_t k = 0; k < Pn.size(); ++k) s += (Pn[k] - P[k]) * (Pn[k] - P[k]);
    double dp = std::sqrt(s / (double)std::max<size_t>(Pn.size(), 1));
    st.last_change = dp;
    double de = std::fabs(e_elec - e_old);
    st.eps = eps;
    st.C = C;
    P.swap(Pn);
    e_old = e_elec;
    if (it > 0 && de < o.e_tol && dp < o.p_tol) {
      st.converged = true;
      break;
    }
  }
  build_fock(m, core, P.data(), F.data());
  st.e_elec = 0.5 * (dot(P, core.h) + dot(P, F));
  st.P = P;
  st.iterations += pre;
}

struct UState {
  std::vector<double> Pa, Pb, ea, eb, Ca, Cb;
  double e_elec = 0, last_change = INFINITY;
  int iterations = 0;
  bool converged = false;
};

// The starting (P_a, P_b) pair, uhf_guess (src/scf.rs:937-980): core-Hamiltonian orbitals filled
// to the two depths, or the superposition (or symmetry-broken, src/scf.rs:894-916) density scaled
// by n_a / n and n_b / n.
static void uhf_guess(const Mol& m, const Core& core, int na, int nb, int guess, std::vector<double>& Pa,
                      std::vector<double>& Pb) {
  int n = m.nao;
  Pa.assign((size_t)n * n, 0.0);
  Pb.assign((size_t)n * n, 0.0);
  if (guess == 2) {
    std::vector<double> w(n), C((size_t)n * n);
    sym_eigen(n, core.h.data(), w.data(), C.data());
    density_from_coeff(n, C.data(), na, 1.0, Pa.data());
    density_from_coeff(n, C.data(), nb, 1.0, Pb.data());
    return;
  }
  std::vector<double> P;
  sad_density(m, P);
  if (guess == 3) {
    int target = -1;
    double best = 0.0;
    for (int i = 0; i < m.nat; ++i)
      if (m.norb[i] >= 4 && m.el[i]->iop > best) { best = m.el[i]->iop; target = i; }
    if (target >= 0) {
      int off = m.off[target];
      double tilt = std::min(0.2, P[(size_t)(off + 3) * n + off + 3]);
      P[(size_t)(off + 3) * n + off + 3] -= tilt;
      P[(size_t)(off + 1) * n + off + 1] += tilt;
    }
  }
  const double ne = (double)(na + nb);
  const double fa = ne > 0.0 ? na / ne : 0.5, fb = ne > 0.0 ? nb / ne : 0.5;
  for (size_t k = 0; k < P.size(); ++k) {
    Pa[k] = P[k] * fa;
    Pb[k] = P[k] * fb;
  }
}

// E = F P - P F for symmetric F and P (src/scf.rs:1039-1054), into E[off..off + n^2).
static void commutator_into(int n, const double* F, const double* P, double* E) {
  std::vector<double> FP((size_t)n * n);
  matmul(n, F, P, FP.data());
  for (int i = 0; i < n; ++i) {
    E[(size_t)i * n + i] = 0.0;
    for (int j = 0; j < i; ++j) {
      double up = FP[(size_t)j * n + i], lo = FP[(size_t)i * n + j];
      E[(size_t)i * n + j] = lo - up;
      E[(size_t)j * n + i] = up - lo;
    }
  }
}

// uhf_loop (src/scf.rs:1319-1496) for s/p elements, so without the d-block penalty: spin Fock
// matrices from (P_a + P_b, P_s), paired CDIIS on the stacked [E_a; E_b] error only when CDIIS is
// the accelerator, level shift on each spin's virtual space, aufbau densities of weight 1, the
// sum of the two RMS density changes, then damping; the energy is
// (P.H + P_a.F_a + P_b.F_b) / 2.
static void uhf_loop(const Mol& m, const Core& core, int na, int nb, const Options& o, int guess, int accel,
                     double shift, double damping, UState& st) {
  const int n = m.nao;
  const size_t nn = (size_t)n * n;
  std::vector<double> Pa, Pb;
  uhf_guess(m, core, na, nb, guess, Pa, Pb);
  const bool use_cdiis = o.use_diis && accel == 1;
  const size_t max_diis = (size_t)diis_depth(n, 4, o.scf_memory_mb);
  std::vector<std::vector<double>> hfa, hfb, herr;
  std::vector<double> Pt(nn), Fa(nn), Fb(nn), ea(n), eb(n), Ca(nn), Cb(nn), Pan(nn), Pbn(nn);
  double e_old = 0.0;
  st.converged = false;
  st.last_change = INFINITY;
  for (int it = 0; it < o.max_scf; ++it) {
    st.iterations = it + 1;
    for (size_t k = 0; k < nn; ++k) Pt[k] = Pa[k] + Pb[k];
    build_fock_spin(m, core, Pt.data(), Pa.data(), Fa.data());
    build_fock_spin(m, core, Pt.data(), Pb.data(), Fb.data());
    const double e_elec = 0.5 * (dot(Pt, core.h) + dot(Pa, Fa) + dot(Pb, Fb));
    std::vector<double> fa = Fa, fb = Fb;
    if (use_cdiis) {
      std::vector<double> err(2 * nn);
      commutator_into(n, Fa.data(), Pa.data(), err.data());
      commutator_into(n, Fb.data(), Pb.data(), err.data() + nn);
      hfa.push_back(Fa);
      hfb.push_back(Fb);
      herr.push_back(std::move(err));
      if (hfa.size() > max_diis) {
        hfa.erase(hfa.begin());
        hfb.erase(hfb.begin());
        herr.erase(herr.begin());
      }
      std::vector<std::vector<double>> gram(herr.size(), std::vector<double>(herr.size()));
      for (size_t i = 0; i < herr.size(); ++i)
        for (size_t j = 0; j < herr.size(); ++j) gram[i][j] = dot(herr[i], herr[j]);
      std::vector<double> c;
      if (diis_coeffs(gram, c)) {
        for (size_t k = 0; k < nn; ++k) {
          double sa = 0.0, sb = 0.0;
          for (size_t i = 0; i < c.size(); ++i) {
            sa += c[i] * hfa[i][k];
            sb += c[i] * hfb[i][k];
          }
          fa[k] = sa;
          fb[k] = sb;
        }
      }
    }
    if (shift != 0.0)  // level_shift (src/scf.rs:337-350)
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          const double id = i == j ? 1.0 : 0.0;
          fa[(size_t)i * n + j] += shift * (id - Pa[(size_t)i * n + j]);
          fb[(size_t)i * n + j] += shift * (id - Pb[(size_t)i * n + j]);
        }
    sym_eigen(n, fa.data(), ea.data(), Ca.data());
    sym_eigen(n, fb.data(), eb.data(), Cb.data());
    density_from_coeff(n, Ca.data(), na, 1.0, Pan.data());
    density_from_coeff(n, Cb.data(), nb, 1.0, Pbn.data());
    double sa = 0.0, sb = 0.0;
    for (size_t k = 0; k < nn; ++k) {
      sa += (Pan[k] - Pa[k]) * (Pan[k] - Pa[k]);
      sb += (Pbn[k] - Pb[k]) * (Pbn[k] - Pb[k]);
    }
    const double dp = std::sqrt(sa / (double)std::max<size_t>(nn, 1)) + std::sqrt(sb / (double)std::max<size_t>(nn, 1));
    st.last_change = dp;
    if (damping > 0.0)
      for (size_t k = 0; k < nn; ++k) {
        Pan[k] = (1.0 - damping) * Pan[k] + damping * Pa[k];
        Pbn[k] = (1.0 - damping) * Pbn[k] + damping * Pb[k];
      }
    const double de = std::fabs(e_elec - e_old);
    st.ea = ea;
    st.eb = eb;
    st.Ca = Ca;
    st.Cb = Cb;
    Pa.swap(Pan);
    Pb.swap(Pbn);
    e_old = e_elec;
    if (it > 0 && de < o.e_tol && dp < o.p_tol) {
      st.converged = true;
      break;
    }
  }
  for (size_t k = 0; k < nn; ++k) Pt[k] = Pa[k] + Pb[k];
  build_fock_spin(m, core, Pt.data(), Pa.data(), Fa.data());
  build_fock_spin(m, core, Pt.data(), Pb.data(), Fb.data());
  st.e_elec = 0.5 * (dot(Pt, core.h) + dot(Pa, Fa) + dot(Pb, Fb));
  st.Pa = Pa;
  st.Pb = Pb;
}

// Core-core repulsion summed over all atom pairs in index order (core_core_energy,
// src/repulsion.rs:13-27).
static double orc_core_core_impl(const Mol& m) {
  double e = 0.0;
  const double* xyz = m.xyz.data();
  for (int i = 0; i < m.nat; ++i)
    for (int j = i + 1; j < m.nat; ++j) {
      double d[3] = {xyz[3 * j] - xyz[3 * i], xyz[3 * j + 1] - xyz[3 * i + 1], xyz[3 * j + 2] - xyz[3 * i + 2]};
      e += pair_core_energy<double>(*m.el[i], *m.el[j], std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]));
    }
  return e;
}

}  // namespace orc

// ---------------------------------------------------------------------------------------
// C API used by tests/ and bench.py (through ctypes).
using namespace orc;

extern "C" {

int orc_element(int z, double* out /* 32 */) {
  const Table& T = table();
  auto it = T.el.find(z);
  if (it == T.el.end()) return E_MISSING_ELEMENT;
  const Elem& e = it->second;
  double v[32] = {(double)e.n_orb, e.dd, e.qq, e.rho0, e.rho1, e.rho2, e.po[1], e.po[2], e.po[3], e.po[7], e.po[9],
                  e.ddp[2], e.ddp[3], e.e_isol, e.eheat_ev, e.tore, e.mass, e.hsp};
  std::memcpy(out, v, sizeof(v));
  return OK;
}

// Integrals of the ordered pair (za first, the heavy atom of a mixed pair) at displacement
// d = R_b - R_a (Bohr): w (npack_a x npack_b row-major), e1b, e2a, overlap S (4x4 each).
int orc_pair_integrals(int za, int zb, const double* d, double* w, double* e1b, double* e2a, double* s) {
  const Table& T = table();
  auto ia = T.el.find(za), ib = T.el.find(zb);
  if (ia == T.el.end() || ib == T.el.end()) return E_MISSING_ELEMENT;
  double r = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  double x[3] = {d[0] / r, d[1] / r, d[2] / r};
  double dv[3] = {x[0] * r, x[1] * r, x[2] * r};
  PairInts<double> p;
  pair_two_electron<double>(ia->second, ib->second, dv, p);
  for (int k = 0; k < p.pa * p.pb; ++k) w[k] = p.w[k];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      e1b[a * 4 + b] = p.e1b[a][b];
      e2a[a * 4 + b] = p.e2a[a][b];
    }
  double o[3] = {0, 0, 0}, S_[4][4];
  if (!diatom_overlap(ia->second, o, ib->second, d, S_)) return E_UNSUPPORTED;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) s[a * 4 + b] = S_[a][b];
  return OK;
}

int orc_nao(int nat, const int* z) {
  const Table& T = table();
  int n = 0;
  for (int i = 0; i < nat; ++i) {
    auto it = T.el.find(z[i]);
    if (it == T.el.end()) return -1;
    n += it->second.n_orb;
  }
  return n;
}

int orc_hcore(int nat, const int* z, const double* xyz, double* h) {
  Mol m;
  int st = make_mol(nat, z, xyz, m);
  if (st) return st;
  Core c;
  if ((st = build_core(m, c))) return st;
  std::memcpy(h, c.h.data(), sizeof(double) * c.h.size());
  return OK;
}

int orc_fock(int nat, const int* z, const double* xyz, const double* P, double* F) {
  Mol m;
  int st = make_mol(nat, z, xyz, m);
  if (st) return st;
  Core c;
  if ((st = build_core(m, c))) return st;
  build_fock(m, c, P, F);
  return OK;
}

// Literal four-index two-centre contraction (src/fock.rs:510-559) plus the same one-centre
// block, as an independent check of the packed build.
int orc_fock_fourindex(int nat, const int* z, const double* xyz, const double* P, double* F) {
  Mol m;
  int st = make_mol(nat, z, xyz, m);
  if (st) return st;
  Core c;
  if ((st = build_core(m, c))) return st;
  int n = m.nao;
  std::memcpy(F, c.h.data(), sizeof(double) * (size_t)n * n);
  for (int ia = 0; ia < m.nat; ++ia) {
    const Elem& e = *m.el[ia];
    int o = m.off[ia], nn = m.norb[ia];
    for (int mu = 0; mu < nn; ++mu)
      for (int nu = 0; nu < nn; ++nu) {
        double acc = 0.0;
        for (int la = 0; la < nn; ++la)
          for (int si = 0; si < nn; ++si) {
            acc += P[(size_t)(o + la) * n + o + si] * oc_two_electron(mu, nu, la, si, e);
            acc -= 0.5 * P[(size_t)(o + la) * n + o + si] * oc_two_electron(mu, la, nu, si, e);
          }
        F[(size_t)(o + mu) * n + o + nu] += acc;
      }
  }
  for (const Pair& p : c.pairs) {
    const auto& t = p.t;
    int oa = m.off[p.a], ob = m.off[p.b], na = t.na, nb = t.nb;
    auto W = [&](int mu, int nu, int la, int si) { return t.w[pack(mu, nu) * t.pb + pack(la, si)]; };
    for (int mu = 0; mu < na; ++mu)
      for (int nu = 0; nu < na; ++nu) {
        double acc = 0.0;
        for (int la = 0; la < nb; ++la)
          for (int si = 0; si < nb; ++si) acc += P[(size_t)(ob + la) * n + ob + si] * W(mu, nu, la, si);
        F[(size_t)(oa + mu) * n + oa + nu] += acc;
      }
    for (int la = 0; la < nb; ++la)
      for (int si = 0; si < nb; ++si) {
        double acc = 0.0;
        for (int mu = 0; mu < na; ++mu)
          for (int nu = 0; nu < na; ++nu) acc += P[(size_t)(oa + mu) * n + oa + nu] * W(mu, nu, la, si);
        F[(size_t)(ob + la) * n + ob + si] += acc;
      }
    for (int mu = 0; mu < na; ++mu)
      for (int la = 0; la < nb; ++la) {
        double acc = 0.0;
        for (int nu = 0; nu < na; ++nu)
          for (int si = 0; si < nb; ++si) acc += 0.5 * P[(size_t)(oa + nu) * n + ob + si] * W(mu, nu, la, si);
        F[(size_t)(oa + mu) * n + ob + la] -= acc;
        F[(size_t)(ob + la) * n + oa + mu] = F[(size_t)(oa + mu) * n + ob + la];
      }
  }
  return OK;
}

// response_fock_rhf (src/fock.rs:444-485): G(dP) = J(dP) - K(dP / 2), no H_core; the one-centre
// blocks as in the Fock build, the two-centre terms of pair_contribution (:106-159) with each pair's
// switch sw (pair_sw[k] in core.pairs order, src/hamiltonian.rs:237-240; null: 1) on J_a, on d_a in
// J_b and on the exchange scale, a zero switch skipping the pair, then scatter_pair (:163-193).
int orc_response_fock(int nat, const int* z, const double* xyz, int nrhs, const double* dp, const double* pair_sw,
                      double* out) {
  Mol m;
  int st = make_mol(nat, z, xyz, m);
  if (st) return st;
  Core c;
  if ((st = build_core(m, c))) return st;
  const int n = m.nao;
  const size_t nn = (size_t)n * n;
  for (int r = 0; r < nrhs; ++r) {
    const double* P = dp + r * nn;
    double* F = out + r * nn;
    std::fill(F, F + nn, 0.0);
    for (int ia = 0; ia < m.nat; ++ia) {
      const Elem& e = *m.el[ia];
      const int o = m.off[ia], nb = m.norb[ia];
      for (int mu = 0; mu < nb; ++mu)
        for (int nu = 0; nu < nb; ++nu) {
          double acc = 0.0;
          for (int la = 0; la < nb; ++la)
            for (int si = 0; si < nb; ++si) {
              acc += P[(size_t)(o + la) * n + o + si] * oc_two_electron(mu, nu, la, si, e);
              acc -= 0.5 * P[(size_t)(o + la) * n + o + si] * oc_two_electron(mu, la, nu, si, e);
            }
          F[(size_t)(o + mu) * n + o + nu] += acc;
        }
    }
    for (size_t k = 0; k < c.pairs.size(); ++k) {
      const Pair& p = c.pairs[k];
      const double sw = pair_sw ? pair_sw[k] : 1.0;
      const auto& t = p.t;
      const int oa = m.off[p.a], ob = m.off[p.b], na = t.na, nb = t.nb, pa = t.pa, pb = t.pb;
      double ja[10] = {0}, jb[10] = {0}, kab[4][4] = {{0}};
      if (sw != 0.0) {
        double da[10], db[10];
        for (int mu = 0; mu < na; ++mu)
          for (int nu = 0; nu <= mu; ++nu)
            da[pack(mu, nu)] = (mu == nu ? 1.0 : 2.0) * P[(size_t)(oa + mu) * n + oa + nu];
        for (int mu = 0; mu < nb; ++mu)
          for (int nu = 0; nu <= mu; ++nu)
            db[pack(mu, nu)] = (mu == nu ? 1.0 : 2.0) * P[(size_t)(ob + mu) * n + ob + nu];
        for (int q = 0; q < pa; ++q) {
          double acc = 0.0;
          for (int s2 = 0; s2 < pb; ++s2) acc += db[s2] * t.w[q * pb + s2];
          ja[q] = sw * acc;
          const double d = sw * da[q];
          for (int s2 = 0; s2 < pb; ++s2) jb[s2] += d * t.w[q * pb + s2];
        }
        const double scale = sw * 0.5;
        for (int mu = 0; mu < na; ++mu)
          for (int nu = 0; nu < na; ++nu) {
            const double* row = &t.w[pack(mu, nu) * pb];
            for (int la = 0; la < nb; ++la) {
              double acc = 0.0;
              for (int si = 0; si < nb; ++si) acc += P[(size_t)(oa + nu) * n + ob + si] * row[pack(la, si)];
              kab[mu][la] -= scale * acc;
            }
          }
      }
      for (int mu = 0; mu < na; ++mu)
        for (int nu = 0; nu < na; ++nu) F[(size_t)(oa + mu) * n + oa + nu] += ja[pack(mu, nu)];
      for (int la = 0; la < nb; ++la)
        for (int si = 0; si < nb; ++si) F[(size_t)(ob + la) * n + ob + si] += jb[pack(la, si)];
      for (int mu = 0; mu < na; ++mu)
        for (int la = 0; la < nb; ++la) {
          const double v = F[(size_t)(oa + mu) * n + ob + la] + kab[mu][la];
          F[(size_t)(oa + mu) * n + ob + la] = v;
          F[(size_t)(ob + la) * n + oa + mu] = v;
        }
    }
  }
  return OK;
}

double orc_core_core(int nat, const int* z, const double* xyz) {
  Mol m;
  if (make_mol(nat, z, xyz, m)) return NAN;
  return orc_core_core_impl(m);
}

int orc_eigen(int n, const double* A, double* w, double* V) {
  sym_eigen(n, A, w, V);
  return OK;
}

// The state a solve hands back (ScfState, src/scf.rs:1300-1315 and :1482-1495): the total
// density, and for an unrestricted state the spin density and the beta spectrum too.
struct Scf {
  std::vector<double> P, S, ea, eb, Ca, Cb;  // S, eb, Cb empty for a restricted state
  double e_elec = 0.0, last_change = INFINITY;
  int iterations = 0, n_occ = 0, n_beta = 0;
  bool converged = false, unrestricted = false;
};

static Scf scf_of_rhf(const State& s, int nocc) {
  Scf r;
  r.P = s.P;
  r.ea = s.eps;
  r.Ca = s.C;
  r.e_elec = s.e_elec;
  r.last_change = s.last_change;
  r.iterations = s.iterations;
  r.converged = s.converged;
  r.n_occ = r.n_beta = nocc;
  return r;
}

static Scf scf_of_uhf(const UState& u, int na, int nb) {
  Scf r;
  r.P.resize(u.Pa.size());
  r.S.resize(u.Pa.size());
  for (size_t k = 0; k < u.Pa.size(); ++k) {
    r.P[k] = u.Pa[k] + u.Pb[k];
    r.S[k] = u.Pa[k] - u.Pb[k];
  }
  r.ea = u.ea;
  r.eb = u.eb;
  r.Ca = u.Ca;
  r.Cb = u.Cb;
  r.e_elec = u.e_elec;
  r.last_change = u.last_change;
  r.iterations = u.iterations;
  r.converged = u.converged;
  r.n_occ = na;
  r.n_beta = nb;
  r.unrestricted = true;
  return r;
}

// Frontier levels over both spin channels (frontier_gap, src/scf.rs:106-124; the reported
// HOMO / LUMO, :797-820): HOMO the higher occupied level, LUMO the lower virtual one.
static void frontier_of(const Scf& s, int n, bool& hh, double& h, bool& hl, double& l) {
  hh = hl = false;
  h = -INFINITY;
  l = INFINITY;
  auto take = [&](const std::vector<double>& e, int occ) {
    if (e.empty()) return;
    if (occ >= 1) { hh = true; h = std::max(h, e[occ - 1]); }
    if (occ < n) { hl = true; l = std::min(l, e[occ]); }
  };
  take(s.ea, s.n_occ);
  if (s.unrestricted) take(s.eb, s.n_beta);
}

static double frontier_gap(const Scf& s, int n) {
  bool hh, hl;
  double h, l;
  frontier_of(s, n, hh, h, hl, l);
  return hh && hl ? l - h : INFINITY;
}

// Largest element of P_a - P_b (spin_purity, src/scf.rs:80-86).
static double spin_purity(const Scf& s) {
  double m = 0.0;
  for (double v : s.S) m = std::max(m, std::fabs(v));
  return m;
}

// A spin-pure unrestricted state read as the restricted state it is (collapse_to_restricted,
// src/scf.rs:88-95): the alpha orbitals stand for both spins.
static void collapse(Scf& s) {
  s.S.clear();
  s.eb.clear();
  s.Cb.clear();
  s.n_beta = s.n_occ;
  s.unrestricted = false;
}

constexpr double STABILITY_GAP_EV = 6.0;   // src/scf.rs:74
constexpr double STABILITY_TOL_EV = 1.0e-6;  // src/scf.rs:79
constexpr double SPIN_PURE_TOL = 1.0e-8;   // src/scf.rs:100

// run_pm3 (src/scf.rs:450-848) for s/p elements.
//  - electron count and spin counts n_a = (N + M - 1) / 2, n_b = (N - M + 1) / 2 (:486-506);
//  - RHF or UHF by options.reference (:508-520): Auto picks UHF for n_a != n_b, RHF refuses an
//    open shell;
//  - the pair-cache ceiling (src/hamiltonian.rs:202-211, default 4096 MiB or
//    PM3_MAX_PAIR_CACHE_MB, src/hamiltonian.rs:44-49) -> E_RESOURCE;
//  - the first path; a level shift or damping "steers" it (:563-590): rhf_loop reads neither, so
//    for a restricted solve steering only switches the stability search off;
//  - the stability search (:599-715) when the frontier gap is below 6 eV (Auto) or always (Always),
//    never when steered.  Rungs (:647-673), guess x accelerator x shift x damping:
//      1 (primary, none)  2 (other, options)  3 (other, none)  4 (symmetry-broken, options)
//      5 (symmetry-broken, none)  6 (symmetry-broken, none, -2 eV, unrestricted)
//      7 (primary, CDIIS, damping 0.85, unrestricted)
//    with primary = Core if options.guess is Core else SAD.  A rung runs when its "unrestricted"
//    flag equals closed_shell (:676): a closed shell on the restricted path runs rungs 6 and 7
//    only, through uhf_loop, and keeps a converged one only when it stayed spin-pure (max|P_a -
//    P_b| < 1e-8), collapsed to the restricted state; an unrestricted solve runs all seven rungs
//    as ordinary unrestricted paths (rungs 6 and 7 with their flag false).  The lowest converged
//    electronic energy wins (1e-6 eV);
//  - core-core plus the field's nuclear term, HoF, Mulliken charges, dipole about the centre of
//    mass, HOMO / LUMO over both spin channels.
// P0 (restricted first path only) replaces the guess rule: see orc_run_from_density.
static int solve(int nat, const int* z, const double* xyz, double charge, int multiplicity, const Options& o,
                 const double* P0, Mol& m, Scf& best, Result& r) {
  r = Result{};
  if (multiplicity < 1) return E_INVALID;
  int st = make_mol(nat, z, xyz, m);
  if (st) return st;
  double n_elec = 0.0;
  for (int i = 0; i < nat; ++i) n_elec += m.el[i]->tore;
  n_elec -= charge;
  const long ne = std::lround(n_elec);
  if (std::fabs(n_elec - (double)ne) > 1.0e-6 || ne < 0) return E_INVALID;
  const long nu = multiplicity - 1;
  if (ne - nu < 0 || (ne - nu) % 2 != 0) return E_INVALID;
  const int na = (int)((ne + nu) / 2), nb = (int)((ne - nu) / 2);
  bool use_uhf = na != nb;
  if (o.reference == 2) use_uhf = true;
  if (o.reference == 1 && na != nb) return E_INVALID;
  const int n = m.nao;
  r.nao = n;
  long limit = o.integral_memory_mb;
  if (limit <= 0) {
    limit = 4096;
    if (const char* env = std::getenv("PM3_MAX_PAIR_CACHE_MB")) limit = std::strtol(env, nullptr, 10);
  }
  if (limit > 0) {
    const double need = pair_cache_bytes(m);
    if (need > (double)limit * 1048576.0) {
      r.required_mb = std::ceil(need / 1048576.0);
      r.limit_mb = (double)limit;
      return E_RESOURCE;
    }
  }
  Core core;
  double field_nuclear = 0.0;
  if ((st = build_core(m, core, o.has_field ? o.field : nullptr, &field_nuclear))) return st;
  bool has_he = false;
  for (int i = 0; i < nat; ++i) has_he |= z[i] == 2;
  auto run_scf = [&](double shift, double damp, int guess, int accel, bool force_u, const double* start) {
    Options oo = o;
    oo.use_diis = o.use_diis && accel != 0;
    if (use_uhf || force_u) {
      UState u;
      uhf_loop(m, core, na, nb, oo, guess, oo.use_diis ? accel : 0, shift, damp, u);
      return scf_of_uhf(u, na, nb);
    }
    State s;
    rhf_loop(m, core, na, oo, guess, (oo.use_diis && !has_he) ? accel : 0, s, start);
    return scf_of_rhf(s, na);
  };
  const bool steered = o.level_shift_ev != 0.0 || o.damping != 0.0;
  best = steered ? run_scf(o.level_shift_ev, o.damping, o.guess, o.accelerator, false, P0)
                 : run_scf(0.0, 0.0, o.guess, o.accelerator, false, P0);
  int paths = 1;
  const bool search = o.stability == 2 ? !steered : o.stability == 1 ? (!steered && frontier_gap(best, n) < STABILITY_GAP_EV)
                                                                     : false;
  const double first_ev = best.e_elec;
  if (search) {
    const int primary = o.guess == 2 ? 2 : 1, other = primary == 2 ? 1 : 2;
    const bool closed_shell = !use_uhf;
    const struct { int guess, accel; double shift, damp; bool unrestricted; } rungs[7] = {
        {primary, 0, 0.0, 0.0, false},        {other, o.accelerator, 0.0, 0.0, false}, {other, 0, 0.0, 0.0, false},
        {3, o.accelerator, 0.0, 0.0, false},  {3, 0, 0.0, 0.0, false},                 {3, 0, -2.0, 0.0, closed_shell},
        {primary, 1, 0.0, 0.85, closed_shell}};
    for (const auto& rg : rungs) {
      if (rg.unrestricted != closed_shell) continue;
      ++paths;
      Scf cand = run_scf(rg.shift, rg.damp, rg.guess, rg.accel, rg.unrestricted, nullptr);
      if (!cand.converged) continue;
      if (rg.unrestricted) {
        if (!(spin_purity(cand) < SPIN_PURE_TOL)) continue;
        collapse(cand);
      }
      if (!best.converged || cand.e_elec < best.e_elec - STABILITY_TOL_EV) best = cand;
    }
  }
  const double core_ev = orc_core_core_impl(m) + field_nuclear;
  const double total = best.e_elec + core_ev;
  double eis = 0.0, ehs = 0.0;
  for (int i = 0; i < nat; ++i) { eis += m.el[i]->e_isol; ehs += m.el[i]->eheat_ev; }
  r.total_ev = total;
  r.electronic_ev = best.e_elec;
  r.core_ev = core_ev;
  r.hof_kcal = (total - eis + ehs) * EV_TO_KCAL;
  r.iterations = best.iterations;
  r.converged = best.converged;
  r.last_change = best.last_change;
  r.scf_paths_tried = paths;
  r.scf_improvement_ev = std::max(first_ev - best.e_elec, 0.0);
  r.n_occ = best.n_occ;
  r.n_beta = best.n_beta;
  r.unrestricted = best.unrestricted;
  bool hh, hl;
  double h, l;
  frontier_of(best, n, hh, h, hl, l);
  r.has_homo = hh;
  r.has_lumo = hl;
  r.homo_ev = hh ? h : 0.0;
  r.lumo_ev = hl ? l : 0.0;
  return best.converged ? OK : E_NOT_CONVERGED;
}

// Dipole about the centre of mass (src/dipole.rs:64-140, src/scf.rs:782-791) and Mulliken charges
// from the total density P (src/scf.rs:762-772).
static void density_properties(const Mol& m, int nat, const double* xyz, const std::vector<double>& P, Result& r,
                               double* charges) {
  const int n = m.nao;
  double tm = 0.0, com[3] = {0, 0, 0};
  for (int i = 0; i < nat; ++i) {
    tm += m.el[i]->mass;
    for (int k = 0; k < 3; ++k) com[k] += xyz[3 * i + k] * m.el[i]->mass;
  }
  for (int k = 0; k < 3; ++k) com[k] = tm > 0 ? com[k] / tm : 0.0;
  double dip[3];
  for (int ax = 0; ax < 3; ++ax) {
    double nuc = 0.0, elec = 0.0;
    for (int i = 0; i < nat; ++i) {
      double disp = xyz[3 * i + ax] - com[ax];
      nuc += disp * m.el[i]->tore;
      int o2 = m.off[i];
      for (int mu = 0; mu < m.norb[i]; ++mu) elec += P[(size_t)(o2 + mu) * n + o2 + mu] * disp;
      if (m.norb[i] >= 4) {
        int p = o2 + ax + 1;
        elec += (P[(size_t)o2 * n + p] + P[(size_t)p * n + o2]) * m.el[i]->dd;
      }
    }
    dip[ax] = (nuc - elec) * AU_DIPOLE_TO_DEBYE;
  }
  for (int k = 0; k < 3; ++k) r.dipole[k] = dip[k];
  r.dipole_debye = std::sqrt(dip[0] * dip[0] + dip[1] * dip[1] + dip[2] * dip[2]);
  if (charges)
    for (int i = 0; i < nat; ++i) {
      double pop = 0.0;
      for (int mu = 0; mu < m.norb[i]; ++mu) pop += P[(size_t)(m.off[i] + mu) * n + m.off[i] + mu];
      charges[i] = m.el[i]->tore - pop;
    }
}

// solve plus the outputs.  Every output pointer may be null.  spin_density, mo_energies_beta and
// mo_coeff_beta are zero-filled for a restricted result.
static int run_out(int nat, const int* z, const double* xyz, double charge, int multiplicity, const Options* opt,
                   const double* P0, Result* res, double* density, double* spin_density, double* mo_energies,
                   double* mo_coeff, double* mo_energies_beta, double* mo_coeff_beta, double* charges) {
  Mol m;
  Scf best;
  Result r;
  const int st = solve(nat, z, xyz, charge, multiplicity, *opt, P0, m, best, r);
  if (st == OK || st == E_NOT_CONVERGED) {
    const int n = m.nao;
    const size_t nn = (size_t)n * n;
    density_properties(m, nat, xyz, best.P, r, charges);
    if (density) std::memcpy(density, best.P.data(), sizeof(double) * nn);
    if (spin_density) {
      if (best.unrestricted) std::memcpy(spin_density, best.S.data(), sizeof(double) * nn);
      else std::fill(spin_density, spin_density + nn, 0.0);
    }
    if (mo_energies) std::memcpy(mo_energies, best.ea.data(), sizeof(double) * n);
    if (mo_coeff) std::memcpy(mo_coeff, best.Ca.data(), sizeof(double) * nn);
    if (mo_energies_beta) {
      if (best.unrestricted) std::memcpy(mo_energies_beta, best.eb.data(), sizeof(double) * n);
      else std::fill(mo_energies_beta, mo_energies_beta + n, 0.0);
    }
    if (mo_coeff_beta) {
      if (best.unrestricted) std::memcpy(mo_coeff_beta, best.Cb.data(), sizeof(double) * nn);
      else std::fill(mo_coeff_beta, mo_coeff_beta + nn, 0.0);
    }
  }
  if (res) *res = r;
  return st;
}

// run_pm3 with a multiplicity and every output of Pm3Result (src/scf.rs:352-396).
int orc_run_pm3(int nat, const int* z, const double* xyz, double charge, int multiplicity, const Options* opt,
                Result* res, double* density, double* spin_density, double* mo_energies, double* mo_coeff,
                double* mo_energies_beta, double* mo_coeff_beta, double* charges) {
  return run_out(nat, z, xyz, charge, multiplicity, opt, nullptr, res, density, spin_density, mo_energies, mo_coeff,
                 mo_energies_beta, mo_coeff_beta, charges);
}

// run_pm3 of a singlet (multiplicity 1).
int orc_run(int nat, const int* z, const double* xyz, double charge, const Options* opt, Result* res, double* density,
            double* mo_energies, double* mo_coeff, double* charges) {
  return run_out(nat, z, xyz, charge, 1, opt, nullptr, res, density, nullptr, mo_energies, mo_coeff, nullptr, nullptr,
                 charges);
}

// run_pm3 with a multiplicity: total density, spin density and charges.
int orc_run_uhf(int nat, const int* z, const double* xyz, double charge, int multiplicity, const Options* opt,
                Result* res, double* density, double* spin_density, double* charges) {
  return run_out(nat, z, xyz, charge, multiplicity, opt, nullptr, res, density, spin_density, nullptr, nullptr,
                 nullptr, nullptr, charges);
}

// The same single point started from a caller-provided density P0 (n x n): P0 builds the first
// Fock matrix, whose aufbau density starts the unchanged SCF loop (one extra counted iteration),
// so only idempotent densities enter the DIIS history.  The stability rungs, when the gap rule
// fires, are those of the superposition guess.  This restates the learned-guess start.
int orc_run_from_density(int nat, const int* z, const double* xyz, double charge, const Options* opt, const double* P0,
                         Result* res, double* density) {
  if (!P0) return E_INVALID;
  Options o = *opt;
  if (o.guess == 2) o.guess = 1;
  return run_out(nat, z, xyz, charge, 1, &o, P0, res, density, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
}

// Hellmann-Feynman gradient at a fixed density P (eV/Bohr): src/gradient.rs:306-318 and
// :423-489 (electronic part, dual numbers) plus src/repulsion.rs:30-49 (core-core).  With a spin
// density S (may be null) the exchange term is the unrestricted one of
// electronic_gradient_fixed_density_spin (src/gradient.rs:523-599): P_a = (P + S) / 2,
// P_b = (P - S) / 2 and -(P_a P_a + P_b P_b) in place of -P P / 2.
int orc_fixed_density_gradient_spin(int nat, const int* z, const double* xyz, const double* P, const double* S,
                                    double* g);
int orc_fixed_density_gradient(int nat, const int* z, const double* xyz, const double* P, double* g) {
  return orc_fixed_density_gradient_spin(nat, z, xyz, P, nullptr, g);
}

int orc_fixed_density_gradient_spin(int nat, const int* z, const double* xyz, const double* P, const double* S,
                                    double* g) {
  Mol m;
  int st = make_mol(nat, z, xyz, m);
  if (st) return st;
  int n = m.nao;
  std::fill(g, g + 3 * nat, 0.0);
  for (int i = 0; i < nat; ++i)
    for (int j = i + 1; j < nat; ++j) {
      double d[3] = {xyz[3 * j] - xyz[3 * i], xyz[3 * j + 1] - xyz[3 * i + 1], xyz[3 * j + 2] - xyz[3 * i + 2]};
      double dist = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
      Dual e = pair_core_energy<Dual>(*m.el[i], *m.el[j], dmk(dist, 1, 0, 0));
      for (int k = 0; k < 3; ++k) {
        double c = d[k] / dist * (-e.d[0]);
        g[3 * i + k] += c;
        g[3 * j + k] -= c;
      }
    }
  for (int u = 0; u < nat; ++u)
    for (int v = u + 1; v < nat; ++v) {
      const Elem& eu = *m.el[u];
      const Elem& ev = *m.el[v];
      int a = u, b = v;
      if (!(eu.has_p() || !ev.has_p())) { a = v; b = u; }
      const Elem& ea = *m.el[a];
      const Elem& eb = *m.el[b];
      const double* pa = &xyz[3 * a];
      const double* pb = &xyz[3 * b];
      Dual dv[3] = {dmk(pb[0] - pa[0], 1, 0, 0), dmk(pb[1] - pa[1], 0, 1, 0), dmk(pb[2] - pa[2], 0, 0, 1)};
      PairInts<Dual> t;
      pair_two_electron<Dual>(ea, eb, dv, t);
      Dual S_[4][4];
      if (!diatom_overlap_dual(ea, pa, eb, pb, S_)) return E_UNSUPPORTED;
      int oa = m.off[a], ob = m.off[b], na = m.norb[a], nb = m.norb[b];
      double f[3] = {0, 0, 0};
      for (int i = 0; i < na; ++i) {
        double bi = beta_of(ea, i);
        for (int j = 0; j < nb; ++j) {
          double coef = P[(size_t)(oa + i) * n + ob + j] * (bi + beta_of(eb, j));
          for (int ax = 0; ax < 3; ++ax) f[ax] += coef * S_[i][j].d[ax];
        }
      }
      for (int i = 0; i < na; ++i)
        for (int j = 0; j < na; ++j) {
          double coef = P[(size_t)(oa + i) * n + oa + j];
          for (int ax = 0; ax < 3; ++ax) f[ax] += coef * t.e1b[i][j].d[ax];
        }
      for (int k = 0; k < nb; ++k)
        for (int l = 0; l < nb; ++l) {
          double coef = P[(size_t)(ob + k) * n + ob + l];
          for (int ax = 0; ax < 3; ++ax) f[ax] += coef * t.e2a[k][l].d[ax];
        }
      for (int mu = 0; mu < na; ++mu)
        for (int nu = 0; nu < na; ++nu)
          for (int la = 0; la < nb; ++la)
            for (int si = 0; si < nb; ++si) {
              const Dual& dw = t.w[pack(mu, nu) * t.pb + pack(la, si)];
              double coul = P[(size_t)(oa + mu) * n + oa + nu] * P[(size_t)(ob + la) * n + ob + si];
              double exch;
              if (S) {
                const size_t i1 = (size_t)(oa + mu) * n + ob + la, i2 = (size_t)(oa + nu) * n + ob + si;
                const double a1 = 0.5 * (P[i1] + S[i1]), a2 = 0.5 * (P[i2] + S[i2]);
                const double b1 = 0.5 * (P[i1] - S[i1]), b2 = 0.5 * (P[i2] - S[i2]);
                exch = -(a1 * a2 + b1 * b2);
              } else {
                exch = -0.5 * P[(size_t)(oa + mu) * n + ob + la] * P[(size_t)(oa + nu) * n + ob + si];
              }
              double coef = coul + exch;
              for (int ax = 0; ax < 3; ++ax) f[ax] += coef * dw.d[ax];
            }
      for (int ax = 0; ax < 3; ++ax) {
        g[3 * b + ax] += f[ax];
        g[3 * a + ax] -= f[ax];
      }
    }
  return OK;
}

// closed_form_gradient (src/gradient.rs:201-276): run_pm3, then the fixed-density gradient at the
// converged density (the spin-resolved one for an unrestricted result, :242-256), then the field's
// -q_A f (add_field_gradient, :296-301).
int orc_gradient_pm3(int nat, const int* z, const double* xyz, double charge, int multiplicity, const Options* opt,
                     Result* res, double* grad, double* density) {
  Mol m;
  int st = make_mol(nat, z, xyz, m);
  if (st) return st;
  const size_t nn = (size_t)m.nao * m.nao;
  std::vector<double> P(nn), S(nn), q(nat);
  Result r;
  st = run_out(nat, z, xyz, charge, multiplicity, opt, nullptr, &r, P.data(), S.data(), nullptr, nullptr, nullptr,
               nullptr, q.data());
  if (res) *res = r;
  if (st) return st;
  if (density) std::memcpy(density, P.data(), sizeof(double) * nn);
  st = orc_fixed_density_gradient_spin(nat, z, xyz, P.data(), r.unrestricted ? S.data() : nullptr, grad);
  if (st) return st;
  if (opt->has_field)
    for (int i = 0; i < nat; ++i)
      for (int ax = 0; ax < 3; ++ax) grad[3 * i + ax] -= opt->field[ax] * q[i];
  return OK;
}

int orc_gradient(int nat, const int* z, const double* xyz, double charge, const Options* opt, Result* res, double* grad,
                 double* density) {
  return orc_gradient_pm3(nat, z, xyz, charge, 1, opt, res, grad, density);
}

// A batch of independent single points spread over host threads (one solve per thread),
// the CPU-baseline harness.  Molecules are CSR-packed: atoms of molecule k are
// [atom_ptr[k], atom_ptr[k+1]).  Writes total energy, HoF, iterations and status per molecule.
int orc_run_batch(int nmol, const int* atom_ptr, const int* z, const double* xyz, const double* charge, const Options* opt,
                  int nthreads, double* total_ev, double* hof, int* iters, int* status) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int k = 0; k < nmol; ++k) {
    int a0 = atom_ptr[k], nat = atom_ptr[k + 1] - a0;
    Result r{};
    int s = orc_run(nat, z + a0, xyz + 3 * a0, charge ? charge[k] : 0.0, opt, &r, nullptr, nullptr, nullptr, nullptr);
    total_ev[k] = r.total_ev;
    hof[k] = r.hof_kcal;
    iters[k] = r.iterations;
    status[k] = s;
  }
  return OK;
}

// The same harness for SCF + closed_form_gradient (src/gradient.rs:201-276) per molecule.  grad
// holds 3 doubles per atom (eV/Bohr) in the CSR order of z.
int orc_gradient_batch(int nmol, const int* atom_ptr, const int* z, const double* xyz, const double* charge,
                       const Options* opt, int nthreads, double* total_ev, double* hof, int* iters, int* status,
                       double* grad) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
  for (int k = 0; k < nmol; ++k) {
    int a0 = atom_ptr[k], nat = atom_ptr[k + 1] - a0;
    Result r{};
    int s = orc_gradient(nat, z + a0, xyz + 3 * a0, charge ? charge[k] : 0.0, opt, &r, grad + 3 * a0, nullptr);
    total_ev[k] = r.total_ev;
    hof[k] = r.hof_kcal;
    iters[k] = r.iterations;
    status[k] = s;
  }
  return OK;
}

int orc_num_threads() {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

// ---------------------------------------------------------------------------------------
// Divide-and-conquer SCF (Yang-Lee) of a molecule: the partition (src/dc/partition.rs:184-275),
// the density pattern (src/dc/pattern.rs:62-171), the Pulay DIIS on the density
// (src/densitydiis.rs:34-119) and the SCF loop with one global chemical potential
// (src/dc/scf.rs:76-160, :352-690).  Matrices are dense nao x nao in the caller's AO order; every
// quantity the reference keeps on the pattern is held as a compact vector of the pattern's entries
// in its row-major order, so sums run in the reference's order.

struct DcOpts {  // DcOptions (src/dc/partition.rs:37-107), without long_range_cutoff
  double core_radius, buffer_radius;
  int min_core_atoms;
  double electron_tol, smearing_ev;
};

struct DcPart {
  std::vector<std::vector<int>> core, atoms, orbitals;
  std::vector<std::vector<char>> in_core;
  std::vector<int> core_of;
};

static inline double v3_norm2(const double* a, const double* b) {
  const double dx = a[0] - b[0], dy = a[1] - b[1], dz = a[2] - b[2];
  return dx * dx + dy * dy + dz * dz;
}

// Cores grown greedily from the outside in; buffers are every other atom within core + buffer
// radius of some core atom (src/dc/partition.rs:184-275).  Ties of the furthest-atom search go to
// the last index (Iterator::max_by), of the nearest-core search to the first (min_by).
static int dc_partition(const Mol& m, const DcOpts& o, DcPart& p) {
  const int nat = m.nat;
  if (nat == 0) return E_INVALID;
  if (o.core_radius <= 0.0 || o.buffer_radius < 0.0) return E_INVALID;
  const double* x = m.xyz.data();
  double c[3] = {0.0, 0.0, 0.0};
  for (int i = 0; i < nat; ++i)
    for (int k = 0; k < 3; ++k) c[k] += x[3 * i + k];
  for (int k = 0; k < 3; ++k) c[k] *= 1.0 / nat;
  const int NONE = -1;
  p.core_of.assign(nat, NONE);
  p.core.clear();
  int remaining = nat;
  while (remaining > 0) {
    int seed = -1;
    double best = 0.0;
    for (int a = 0; a < nat; ++a) {
      if (p.core_of[a] != NONE) continue;
      const double d = v3_norm2(x + 3 * a, c);
      if (seed < 0 || !(d < best)) { seed = a; best = d; }  // max_by keeps the later of equal elements
    }
    const int index = (int)p.core.size();
    std::vector<int> core;
    for (int a = 0; a < nat; ++a) {
      if (p.core_of[a] != NONE) continue;
      if (std::sqrt(v3_norm2(x + 3 * a, x + 3 * seed)) <= o.core_radius) {
        p.core_of[a] = index;
        core.push_back(a);
        --remaining;
      }
    }
    if ((int)core.size() < o.min_core_atoms && index > 0) {
      int nearest = 0;
      double dn = 0.0;
      for (int q = 0; q < index; ++q) {
        double dq = INFINITY;
        for (int i : p.core[q])
          for (int j : core) dq = std::min(dq, std::sqrt(v3_norm2(x + 3 * i, x + 3 * j)));
        if (q == 0 || dq < dn) { nearest = q; dn = dq; }
      }
      for (int a : core) p.core_of[a] = nearest;
      p.core[nearest].insert(p.core[nearest].end(), core.begin(), core.end());
      continue;
    }
    p.core.push_back(core);
  }
  const double reach = o.core_radius + o.buffer_radius;
  const int ns = (int)p.core.size();
  p.atoms.assign(ns, {});
  p.orbitals.assign(ns, {});
  p.in_core.assign(ns, {});
  for (int s = 0; s < ns; ++s) {
    const std::vector<int>& core = p.core[s];
    std::vector<int>& atoms = p.atoms[s];
    atoms = core;
    const int mine = p.core_of[core[0]];
    for (int a = 0; a < nat; ++a) {
      if (p.core_of[a] == mine) continue;
      bool close = false;
      for (int q : core) close = close || std::sqrt(v3_norm2(x + 3 * a, x + 3 * q)) <= reach;
      if (close) atoms.push_back(a);
    }
    for (int a : atoms) {
      const bool is_core = p.core_of[a] == mine;
      for (int k = 0; k < m.norb[a]; ++k) {
        p.orbitals[s].push_back(m.off[a] + k);
        p.in_core[s].push_back(is_core);
      }
    }
  }
  return OK;
}

static inline double dc_weight(bool mu_core, bool nu_core) {  // src/dc/partition.rs:171-178
  return mu_core && nu_core ? 1.0 : (mu_core || nu_core) ? 0.5 : 0.0;
}

// sum over subsystems of the Yang-Lee weights of every orbital pair (Partition::weight_sum)
static std::vector<double> dc_weight_sum(const DcPart& p, int nao) {
  std::vector<double> out((size_t)nao * nao, 0.0);
  for (size_t s = 0; s < p.orbitals.size(); ++s) {
    const std::vector<int>& o = p.orbitals[s];
    for (size_t i = 0; i < o.size(); ++i)
      for (size_t j = 0; j < o.size(); ++j)
        out[(size_t)o[i] * nao + o[j]] += dc_weight(p.in_core[s][i], p.in_core[s][j]);
  }
  return out;
}

// Pulay DIIS on a density against its residual P_out - P_in (src/densitydiis.rs:34-119).
struct DensityDiis {
  int depth = 8;
  std::vector<std::vector<double>> dens, res, gram;
  DensityDiis(size_t len, int memory_mb) {
    if (memory_mb == 0 || len == 0) {
      depth = 8;
    } else {
      const size_t per = 2 * len * sizeof(double);
      const size_t d = (size_t)memory_mb * 1024 * 1024 / std::max<size_t>(per, 1);
      depth = (int)std::min<size_t>(std::max<size_t>(d, 2), 8);
    }
  }
  static double vdot(const std::vector<double>& a, const std::vector<double>& b) {
    double s = 0.0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
  }
  void push(std::vector<double> d, std::vector<double> r) {
    std::vector<double> row;
    for (const auto& o : res) row.push_back(vdot(o, r));
    row.push_back(vdot(r, r));
    for (size_t i = 0; i < gram.size(); ++i) gram[i].push_back(row[i]);
    gram.push_back(row);
    dens.push_back(std::move(d));
    res.push_back(std::move(r));
    while ((int)dens.size() > depth) {
      dens.erase(dens.begin());
      res.erase(res.begin());
      gram.erase(gram.begin());
      for (auto& g : gram) g.erase(g.begin());
    }
  }
  // sum c_i P_i over the longest suffix of the history that solves with sum |c_i| <= 20
  bool extrapolate(std::vector<double>& out) const {
    const int n = (int)dens.size();
    for (int first = 0; first + 1 < n; ++first) {
      std::vector<std::vector<double>> block;
      for (int i = first; i < n; ++i) block.emplace_back(gram[i].begin() + first, gram[i].end());
      std::vector<double> c;
      if (!diis_coeffs(block, c)) continue;
      const int kept = n - first;
      double w = 0.0;
      for (int i = 0; i < kept; ++i) w += std::fabs(c[i]);
      if (!std::isfinite(w) || w > 20.0) continue;
      out.assign(dens[0].size(), 0.0);
      for (int i = 0; i < kept; ++i)
        for (size_t j = 0; j < out.size(); ++j) out[j] += c[i] * dens[first + i][j];
      return true;
    }
    return false;
  }
};

static inline double fermi_dirac(double e, double mu, double s) {  // src/dc/scf.rs:726-738
  if (s <= 0.0) return e <= mu ? 1.0 : 0.0;
  const double x = (e - mu) / s;
  if (x > 40.0) return 0.0;
  if (x < -40.0) return 1.0;
  return 1.0 / (1.0 + std::exp(x));
}

// The live entries of a partitioned density, row-major (DensityPattern::from_partition).
struct DcPattern {
  int nao = 0;
  std::vector<long long> at;  // dense index mu * nao + nu of every entry, ascending
  std::vector<int> slot;      // dense index -> entry, -1 outside
  void build(const DcPart& p, int n) {
    nao = n;
    slot.assign((size_t)n * n, -1);
    std::vector<char> live((size_t)n * n, 0);
    for (const auto& o : p.orbitals)
      for (int mu : o)
        for (int nu : o) live[(size_t)mu * n + nu] = 1;
    at.clear();
    for (long long e = 0; e < (long long)n * n; ++e)
      if (live[e]) {
        slot[e] = (int)at.size();
        at.push_back(e);
      }
  }
  size_t nnz() const { return at.size(); }
  void gather(const std::vector<double>& dense, std::vector<double>& v) const {
    v.resize(at.size());
    for (size_t i = 0; i < at.size(); ++i) v[i] = dense[at[i]];
  }
  void scatter(const std::vector<double>& v, std::vector<double>& dense) const {
    dense.assign((size_t)nao * nao, 0.0);
    for (size_t i = 0; i < at.size(); ++i) dense[at[i]] = v[i];
  }
};

// One spin channel (src/dc/scf.rs:604-690): every subsystem's Fock block diagonalised, the
// chemical potential bisected to `target` electrons, the density assembled with the Yang-Lee
// weights.  Returns mu; `out` is the assembled density on the pattern.
static double dc_solve_channel(const std::vector<double>& F, int nao, const DcPart& part, double target,
                               const DcOpts& dc, const DcPattern& pat, std::vector<double>& out) {
  const int ns = (int)part.orbitals.size();
  std::vector<std::vector<double>> w(ns), V(ns);
  for (int s = 0; s < ns; ++s) {
    const std::vector<int>& o = part.orbitals[s];
    const int n = (int)o.size();
    std::vector<double> loc((size_t)n * n);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) loc[(size_t)i * n + j] = F[(size_t)o[i] * nao + o[j]];
    w[s].resize(n);
    V[s].resize((size_t)n * n);
    sym_eigen(n, loc.data(), w[s].data(), V[s].data());
  }
  auto electrons_at = [&](double mu) {
    double total = 0.0;
    for (int s = 0; s < ns; ++s) {
      const int n = (int)w[s].size();
      for (int k = 0; k < n; ++k) {
        const double f = fermi_dirac(w[s][k], mu, dc.smearing_ev);
        if (f <= 0.0) continue;
        for (int i = 0; i < n; ++i)
          if (part.in_core[s][i]) {
            const double c = V[s][(size_t)i * n + k];
            total += f * c * c;
          }
      }
    }
    return total;
  };
  // bracket (src/dc/scf.rs:693-724)
  double lowest = INFINITY, highest = -INFINITY;
  for (int s = 0; s < ns; ++s)
    for (double e : w[s]) { lowest = std::min(lowest, e); highest = std::max(highest, e); }
  const double pad = 40.0 * std::max(dc.smearing_ev, 1.0);
  double low = lowest - pad, high = highest + pad;
  for (int wid = 0; electrons_at(low) > target && wid < 40; ++wid) low -= pad;
  for (int wid = 0; electrons_at(high) < target && wid < 40; ++wid) high += pad;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (low + high);
    if (electrons_at(mid) > target) high = mid;
    else low = mid;
    if (std::fabs(high - low) < dc.electron_tol) break;
  }
  const double mu = 0.5 * (low + high);
  out.assign(pat.nnz(), 0.0);
  for (int s = 0; s < ns; ++s) {
    const std::vector<int>& o = part.orbitals[s];
    const int n = (int)o.size();
    for (int k = 0; k < n; ++k) {
      const double f = fermi_dirac(w[s][k], mu, dc.smearing_ev);
      if (f <= 1.0e-14) continue;
      for (int i = 0; i < n; ++i) {
        const double ci = V[s][(size_t)i * n + k];
        if (ci == 0.0) continue;
        for (int j = 0; j < n; ++j) {
          const double wt = dc_weight(part.in_core[s][i], part.in_core[s][j]);
          if (wt == 0.0) continue;
          out[pat.slot[(size_t)o[i] * nao + o[j]]] += wt * f * ci * V[s][(size_t)j * n + k];
        }
      }
    }
  }
  return mu;
}

struct DcResultC {  // DcResult (src/dc/scf.rs:25-46) without the dense matrices
  double total_ev, electronic_ev, core_ev, hof_kcal, fermi_ev, fermi_beta_ev;
  int iterations, converged, unrestricted, dropped_pairs, largest_subsystem, n_subsystems;
};

// run_dc (src/dc/scf.rs:76-160) with the loop of :376-548.  dcopt = {core_radius, buffer_radius,
// min_core_atoms, electron_tol, smearing_ev}.  density / spin_density nao x nao, charges nat (any
// may be null).  E_NOT_CONVERGED leaves the outputs unwritten, as the reference returns an error.
int orc_run_dc(int nat, const int* z, const double* xyz, double charge, int multiplicity, const Options* opt,
               const double* dcopt, DcResultC* res, double* density, double* spin_density, double* charges) {
  const Options& o = *opt;
  DcOpts dc{dcopt[0], dcopt[1], (int)dcopt[2], dcopt[3], dcopt[4]};
  Mol m;
  int st = make_mol(nat, z, xyz, m);
  if (st) return st;
  // electron_counts (src/dc/scf.rs:357-398)
  const int mult = std::max(multiplicity, 1);
  double n_elec = 0.0;
  for (int i = 0; i < nat; ++i) n_elec += m.el[i]->tore;
  n_elec -= charge;
  const double n_unp = mult - 1;
  if (n_elec < n_unp || n_elec < 0.0) return E_INVALID;
  const double na = 0.5 * (n_elec + n_unp), nb = 0.5 * (n_elec - n_unp);
  bool uhf = na != nb;
  if (o.reference == 2) uhf = true;
  if (o.reference == 1) {
    if (na != nb) return E_INVALID;
    uhf = false;
  }
  long limit = o.integral_memory_mb;
  if (limit <= 0) {
    limit = 4096;
    if (const char* env = std::getenv("PM3_MAX_PAIR_CACHE_MB")) limit = std::strtol(env, nullptr, 10);
  }
  if (limit > 0 && pair_cache_bytes(m) > (double)limit * 1048576.0) return E_RESOURCE;
  Core core;
  double field_nuclear = 0.0;
  if ((st = build_core(m, core, o.has_field ? o.field : nullptr, &field_nuclear))) return st;
  DcPart part;
  if ((st = dc_partition(m, dc, part))) return st;
  const int n = m.nao;
  DcPattern pat;
  pat.build(part, n);
  const size_t nnz = pat.nnz();
  // loop_scf (src/dc/scf.rs:412-548)
  const double total_e = na + nb;
  auto share = [&](double e) { return total_e > 0.0 ? e / total_e : 0.5; };
  std::vector<double> atomic;
  sad_density(m, atomic);
  std::vector<double> a(nnz), b(nnz), tmp((size_t)n * n);
  for (size_t i = 0; i < nnz; ++i) {
    a[i] = atomic[pat.at[i]] * share(na);
    b[i] = atomic[pat.at[i]] * share(nb);
  }
  const double damping = o.damping > 0.0 ? std::min(std::max(o.damping, 0.0), 0.95) : 0.3;
  double last = INFINITY;
  bool converged = false;
  int iterations = 0;
  double mu_a = 0.0, mu_b = 0.0;
  DensityDiis hist(2 * nnz, o.scf_memory_mb);
  std::vector<double> ptot(nnz), fa, fb, newtot(nnz), Fa((size_t)n * n), Fb((size_t)n * n);
  std::vector<double> Pd, Pa, Pb, Fva, Fvb;
  for (int it = 1; it <= o.max_scf; ++it) {
    iterations = it;
    for (size_t i = 0; i < nnz; ++i) ptot[i] = a[i] + b[i];
    pat.scatter(ptot, Pd);
    pat.scatter(a, Pa);
    build_fock_spin(m, core, Pd.data(), Pa.data(), Fa.data());
    if (uhf) {
      pat.scatter(b, Pb);
      build_fock_spin(m, core, Pd.data(), Pb.data(), Fb.data());
    }
    mu_a = dc_solve_channel(Fa, n, part, na, dc, pat, fa);
    if (uhf) mu_b = dc_solve_channel(Fb, n, part, nb, dc, pat, fb);
    else mu_b = mu_a;
    const std::vector<double>& bfock = uhf ? Fb : Fa;
    const std::vector<double>& bfresh = uhf ? fb : fa;
    for (size_t i = 0; i < nnz; ++i) newtot[i] = fa[i] + bfresh[i];
    pat.gather(Fa, Fva);
    pat.gather(bfock, Fvb);
    double d1 = 0.0, d2 = 0.0, d3 = 0.0, sq = 0.0;
    for (size_t i = 0; i < nnz; ++i) d1 += newtot[i] * Fva[i];
    for (size_t i = 0; i < nnz; ++i) d2 += fa[i] * Fva[i];
    for (size_t i = 0; i < nnz; ++i) d3 += bfresh[i] * Fvb[i];
    for (size_t i = 0; i < nnz; ++i) sq += (newtot[i] - ptot[i]) * (newtot[i] - ptot[i]);
    const double electronic = 0.5 * (d1 + d2 + d3);
    const double change = std::sqrt(sq / (double)std::max((long long)n * n, 1LL));
    for (size_t i = 0; i < nnz; ++i) a[i] = damping * a[i] + (1.0 - damping) * fa[i];
    for (size_t i = 0; i < nnz; ++i) b[i] = damping * b[i] + (1.0 - damping) * bfresh[i];
    std::vector<double> cur(2 * nnz), res(2 * nnz);
    for (size_t i = 0; i < nnz; ++i) {
      cur[i] = a[i];
      cur[nnz + i] = b[i];
      res[i] = fa[i] - a[i];
      res[nnz + i] = bfresh[i] - b[i];
    }
    hist.push(cur, res);
    std::vector<double> blended;
    if (hist.extrapolate(blended)) {
      for (size_t i = 0; i < nnz; ++i) {
        a[i] = blended[i];
        b[i] = blended[nnz + i];
      }
    }
    if (std::fabs(electronic - last) < o.e_tol && change < o.p_tol) {
      converged = true;
      break;
    }
    last = electronic;
  }
  if (!converged) return E_NOT_CONVERGED;
  // the Fock matrices rebuilt at the converged density, the energy there (src/dc/scf.rs:554-572,
  // :110-122)
  std::vector<double> dens(nnz);
  for (size_t i = 0; i < nnz; ++i) dens[i] = a[i] + b[i];
  pat.scatter(dens, Pd);
  pat.scatter(a, Pa);
  build_fock_spin(m, core, Pd.data(), Pa.data(), Fa.data());
  if (uhf) {
    pat.scatter(b, Pb);
    build_fock_spin(m, core, Pd.data(), Pb.data(), Fb.data());
  } else {
    Fb = Fa;
  }
  std::vector<double> Hv;
  pat.gather(core.h, Hv);
  pat.gather(Fa, Fva);
  pat.gather(Fb, Fvb);
  double e1 = 0.0, e2 = 0.0, e3 = 0.0;
  for (size_t i = 0; i < nnz; ++i) e1 += dens[i] * Hv[i];
  for (size_t i = 0; i < nnz; ++i) e2 += a[i] * Fva[i];
  for (size_t i = 0; i < nnz; ++i) e3 += b[i] * Fvb[i];
  const double electronic = 0.5 * (e1 + e2 + e3);
  const double core_ev = orc_core_core_impl(m) + field_nuclear;
  double eis = 0.0, ehs = 0.0;
  for (int i = 0; i < nat; ++i) { eis += m.el[i]->e_isol; ehs += m.el[i]->eheat_ev; }
  std::vector<double> wsum = dc_weight_sum(part, n);
  int dropped = 0, largest = 0;
  for (double v : wsum) dropped += v < 0.5;
  for (const auto& ob : part.orbitals) largest = std::max(largest, (int)ob.size());
  if (res) {
    DcResultC r{};
    r.electronic_ev = electronic;
    r.core_ev = core_ev;
    r.total_ev = electronic + core_ev;
    r.hof_kcal = (r.total_ev - eis + ehs) * EV_TO_KCAL;
    r.fermi_ev = mu_a;
    r.fermi_beta_ev = mu_b;
    r.iterations = iterations;
    r.converged = 1;
    r.unrestricted = uhf;
    r.dropped_pairs = dropped;
    r.largest_subsystem = largest;
    r.n_subsystems = (int)part.orbitals.size();
    *res = r;
  }
  if (density) std::memcpy(density, Pd.data(), sizeof(double) * (size_t)n * n);
  if (spin_density) {
    std::vector<double> sd(nnz);
    for (size_t i = 0; i < nnz; ++i) sd[i] = a[i] - b[i];
    pat.scatter(sd, tmp);
    if (uhf) std::memcpy(spin_density, tmp.data(), sizeof(double) * (size_t)n * n);
    else std::fill(spin_density, spin_density + (size_t)n * n, 0.0);
  }
  if (charges)
    for (int i = 0; i < nat; ++i) {
      double pop = 0.0;
      for (int mu = 0; mu < m.norb[i]; ++mu) pop += Pd[(size_t)(m.off[i] + mu) * n + m.off[i] + mu];
      charges[i] = m.el[i]->tore - pop;
    }
  return OK;
}

// The partition alone: core_of[nat], and {subsystems, largest subsystem, dropped pairs}; with
// weight_sum (nao x nao, may be null) the summed Yang-Lee weights; with orbitals / in_core /
// sub_ptr (capacity cap entries, may be null) every subsystem's orbital list.
int orc_dc_partition(int nat, const int* z, const double* xyz, const double* dcopt, int* core_of, int* stats,
                     double* weight_sum, int cap, int* sub_ptr, int* orbitals, int* in_core) {
  DcOpts dc{dcopt[0], dcopt[1], (int)dcopt[2], dcopt[3], dcopt[4]};
  Mol m;
  int st = make_mol(nat, z, xyz, m);
  if (st) return st;
  DcPart part;
  if ((st = dc_partition(m, dc, part))) return st;
  for (int i = 0; i < nat; ++i) core_of[i] = part.core_of[i];
  std::vector<double> ws = dc_weight_sum(part, m.nao);
  int dropped = 0, largest = 0;
  for (double v : ws) dropped += v < 0.5;
  for (const auto& ob : part.orbitals) largest = std::max(largest, (int)ob.size());
  DcPattern pat;
  pat.build(part, m.nao);
  stats[0] = (int)part.orbitals.size();
  stats[1] = largest;
  stats[2] = dropped;
  stats[3] = (int)pat.nnz();
  if (weight_sum) std::memcpy(weight_sum, ws.data(), sizeof(double) * ws.size());
  if (sub_ptr) {
    int k = 0;
    sub_ptr[0] = 0;
    for (size_t s = 0; s < part.orbitals.size(); ++s) {
      for (size_t i = 0; i < part.orbitals[s].size(); ++i) {
        if (k >= cap) return E_INVALID;
        orbitals[k] = part.orbitals[s][i];
        in_core[k] = part.in_core[s][i];
        ++k;
      }
      sub_ptr[s + 1] = k;
    }
  }
  return OK;
}

// DensityDiis on nvec pushed (density, residual) pairs of length len: out = the extrapolation
// (returns 1) or untouched (returns 0); depth gets the history depth.
int orc_density_diis(int len, int memory_mb, int nvec, const double* dens, const double* res, double* out, int* depth) {
  DensityDiis d((size_t)len, memory_mb);
  for (int v = 0; v < nvec; ++v)
    d.push(std::vector<double>(dens + (size_t)v * len, dens + (size_t)(v + 1) * len),
           std::vector<double>(res + (size_t)v * len, res + (size_t)(v + 1) * len));
  if (depth) *depth = d.depth;
  std::vector<double> o;
  if (!d.extrapolate(o)) return 0;
  std::memcpy(out, o.data(), sizeof(double) * o.size());
  return 1;
}

}  // extern "C"
