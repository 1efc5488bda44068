// Host side of Pm3Parameters::standard() (src/params.rs:135-475): the embedded element rows plus
// the derived NDDO multipole quantities, packed into the constant-memory table the kernels read.
//
//   dd, qq (charge separations) ............ src/params.rs:584-597 (MOPAC calpar)
//   rho1, rho2 (exactly five secant steps) .. src/params.rs:605-649
//   po / ddp for main-group sp elements ..... src/params.rs:694-710
//   isolated-atom energy e_isol ............. src/params.rs:266-289, :527-552
#include <cmath>
#include <cstring>
#include <algorithm>

#include "pm3_host.hpp"

namespace pm3host {

namespace {

double row_pqn(int z) { return z <= 2 ? 1.0 : (z <= 10 ? 2.0 : (z <= 18 ? 3.0 : 0.0)); }

// Solve H_sp(au) = d/2 - 1/(2 sqrt(4 D1^2 + 1/d^2)) for d by five secant steps; rho1 = 1/(2d).
double secant_rho1(double hsp_ev, double D1) {
  const double target = std::max(hsp_ev, 1.0e-7) / pm3dev::EV;
  auto f = [D1](double d) { return 0.5 * d - 0.5 / std::sqrt(4.0 * D1 * D1 + 1.0 / (d * d)); };
  double x0 = std::pow(target / (D1 * D1), 1.0 / 3.0);
  double x1 = x0 + 0.04;
  for (int step = 0; step < 5; ++step) {
    const double f0 = f(x0), f1 = f(x1);
    if (std::fabs(f1 - f0) < 1.0e-25) break;
    const double x2 = x0 + (x1 - x0) * (target - f0) / (f1 - f0);
    x0 = x1;
    x1 = x2;
  }
  return 0.5 / x1;
}

// Solve H_pp(au) = q/4 - 1/(2 sqrt(4 D2^2 + 1/q^2)) + 1/(4 sqrt(8 D2^2 + 1/q^2)); rho2 = 1/(2q).
double secant_rho2(double hpp_ev, double D2) {
  const double target = hpp_ev / pm3dev::EV;
  auto f = [D2](double q) {
    return 0.25 * q - 0.5 / std::sqrt(4.0 * D2 * D2 + 1.0 / (q * q)) + 0.25 / std::sqrt(8.0 * D2 * D2 + 1.0 / (q * q));
  };
  const double d4 = D2 * D2 * D2 * D2;
  double x0 = std::pow(16.0 * target / (48.0 * d4), 0.2);
  double x1 = x0 + 0.04;
  for (int step = 0; step < 5; ++step) {
    const double f0 = f(x0), f1 = f(x1);
    if (std::fabs(f1 - f0) < 1.0e-25) break;
    const double x2 = x0 + (x1 - x0) * (target - f0) / (f1 - f0);
    x0 = x1;
    x1 = x2;
  }
  return 0.5 / x1;
}

struct Raw {
  int z;
  double uss, upp, zs, zp, bs, bp, gss, gsp, gpp, gp2, hsp, alp;
  double g1k, g1l, g1m, g2k, g2l, g2m, ios, iop, ns, np, eheat, mass, tore;
};

const Raw kRaw[] = {
#define PM3_ELEMENT(...) {__VA_ARGS__},
#define PM3_PAIR(...)
#include "pm3_table.inc"
#undef PM3_ELEMENT
#undef PM3_PAIR
};

struct RawPair { int zi, zj; double alpb, xfac; };
const RawPair kPairs[] = {
#define PM3_ELEMENT(...)
#define PM3_PAIR(ZI, ZJ, A, X) {ZI, ZJ, A, X},
#include "pm3_table.inc"
#undef PM3_ELEMENT
#undef PM3_PAIR
};

}  // namespace

bool build_params(Params& P) {
  std::memset(&P.table, 0, sizeof(P.table));
  std::fill(P.slot_of_z, P.slot_of_z + 128, -1);
  P.nslots = 0;
  for (const Raw& r : kRaw) {
    const int s = P.nslots++;
    if (s >= pm3dev::MAX_SLOTS) return false;
    P.slot_of_z[r.z] = s;
    pm3dev::ElemConst& e = P.table.el[s];
    ElemHost& h = P.host[s];
    e.z = r.z;
    e.n_s = (int)r.ns;
    e.uss = r.uss; e.upp = r.upp; e.betas = r.bs; e.betap = r.bp;
    e.gss = r.gss; e.gsp = r.gsp; e.gpp = r.gpp; e.gp2 = r.gp2;
    e.hsp = std::max(r.hsp, 1.0e-7);
    e.alp = r.alp;
    e.zs = r.zs; e.zp = r.zp; e.tore = r.tore; e.ios = r.ios; e.iop = r.iop; e.mass = r.mass;
    e.norb = r.zp > 0.0 ? 4 : 1;
    e.ngauss = 0;
    const double gk[2] = {r.g1k, r.g2k}, gl[2] = {r.g1l, r.g2l}, gm[2] = {r.g1m, r.g2m};
    for (int k = 0; k < 2; ++k)
      if (gk[k] != 0.0 || gl[k] != 0.0) {
        e.gk[e.ngauss] = gk[k]; e.gl[e.ngauss] = gl[k]; e.gm[e.ngauss] = gm[k]; ++e.ngauss;
      }
    e.rho0 = e.gss > 0.0 ? 0.5 * pm3dev::EV / e.gss : 0.0;
    if (e.norb == 4) {
      const double zp = std::max(e.zp, 0.3);
      const double n = row_pqn(e.z);
      e.dd = (2.0 * n + 1.0) * std::pow(4.0 * e.zs * zp, n + 0.5) / std::pow(e.zs + zp, 2.0 * n + 2.0) / std::sqrt(3.0);
      e.qq = std::sqrt((4.0 * n * n + 6.0 * n + 2.0) / 20.0) / zp;
      e.rho1 = secant_rho1(e.hsp, e.dd);
      e.rho2 = secant_rho2(std::max(0.5 * (e.gpp - e.gp2), 0.1), e.qq);
      e.po1 = e.rho0; e.po2 = e.rho1; e.po3 = e.rho2; e.po9 = e.rho0;
      e.ddp2 = e.dd; e.ddp3 = e.qq * std::sqrt(2.0);
    } else {
      e.dd = e.qq = e.rho1 = e.rho2 = 0.0;
      e.po1 = e.rho0; e.po2 = e.po3 = 0.0; e.po9 = e.rho0;
      e.ddp2 = e.ddp3 = 0.0;
    }
    // Isolated-atom electronic energy (average of configuration over the MOPAC shell
    // occupancies); hydrogen is U_ss itself.
    if (r.z == 1) {
      h.e_isol = e.uss;
    } else {
      const long kp = (long)r.iop;
      const long lp = std::min(kp, 6 - kp);
      const double c_gss = std::max(r.ios - 1.0, 0.0);
      const double c_gsp = r.ios * r.iop;
      const double c_hsp = -r.iop * r.ios * 0.5;
      const double c_gp2 = (double)((kp * (kp - 1)) / 2) + (double)(lp * (lp - 1)) / 4.0;
      const double c_gpp = -(double)(lp * (lp - 1)) / 4.0;
      h.e_isol = e.uss * r.ios + e.upp * r.iop + e.gss * c_gss + e.gpp * c_gpp + e.gsp * c_gsp + e.gp2 * c_gp2 + e.hsp * c_hsp;
    }
    h.eheat_ev = r.eheat / EV_TO_KCAL;
    h.mass = r.mass;
  }
  for (const RawPair& p : kPairs) {
    const int a = P.slot_of_z[p.zi], b = P.slot_of_z[p.zj];
    if (a < 0 || b < 0) continue;
    P.table.alpb[a * pm3dev::MAX_SLOTS + b] = P.table.alpb[b * pm3dev::MAX_SLOTS + a] = p.alpb;
    P.table.xfac[a * pm3dev::MAX_SLOTS + b] = P.table.xfac[b * pm3dev::MAX_SLOTS + a] = p.xfac;
  }
  return true;
}

}  // namespace pm3host
