// LEARNED-GUESS ORACLE -- test infrastructure only.
//
// A plain CPU restatement of the learned initial density that the product computes on the GPU:
// the directed pair list, the network input features, the three-layer network (forward and
// backward), the equivariant assembly of density blocks and the training loss with its output
// gradient.  The reference has no learned guess, so this restates the engine's own
// specification (the feature, network and assembly definitions documented with the product
// source); it is loaded only by tests/ through ctypes and never by the product.
//
// Pins (tests/test_guess_oracle.py): the features against an independent numpy evaluation of
// the definition, the network gradient (exact mode) and the loss gradient against central
// finite differences, rotation equivariance and symmetry of the assembled density, and the
// superposition density at zero outputs.  Rounded mode mirrors the device's bf16 rounding points
// (features, weights, hidden activations and back-propagated errors) with double accumulation.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace gorc {

constexpr int NF = 96, NH = 128, NO = 16;
constexpr int OFF_W1 = 0, OFF_W2 = NF * NH, OFF_B2 = OFF_W2 + NH * NH, OFF_W3 = OFF_B2 + NH, OFF_B3 = OFF_W3 + NH * NO;
constexpr int NPARAM = OFF_B3 + NO;
constexpr double BOHR_A = 0.529177210903;
constexpr double CUT_A = 4.0;
constexpr double PI = 3.14159265358979323846;

static int eidx(int z) { return z == 1 ? 0 : (z == 6 ? 1 : (z == 8 ? 2 : -1)); }

static uint16_t bf16_bits(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}
static float bf16_val(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
static double rnd(double v, bool on) { return on ? (double)bf16_val(bf16_bits((float)v)) : v; }

static double dist_bohr(const double* xyz, int a, int b) {
  const double dx = xyz[3 * b] - xyz[3 * a], dy = xyz[3 * b + 1] - xyz[3 * a + 1], dz = xyz[3 * b + 2] - xyz[3 * a + 2];
  return std::sqrt(dx * dx + dy * dy + dz * dz);
}

static void unit(const double* xyz, int a, int b, double* u) {
  const double r = dist_bohr(xyz, a, b);
  for (int k = 0; k < 3; ++k) u[k] = (xyz[3 * b + k] - xyz[3 * a + k]) / r;
}

// Directed pairs (a, b), a != b, r_ab < 4 Angstrom, ordered by a then b (input order).
static void pairs(int nat, const double* xyz, std::vector<int>& src, std::vector<int>& dst) {
  src.clear();
  dst.clear();
  for (int a = 0; a < nat; ++a)
    for (int b = 0; b < nat; ++b)
      if (a != b && dist_bohr(xyz, a, b) * BOHR_A < CUT_A) {
        src.push_back(a);
        dst.push_back(b);
      }
}

static void env_side(int nat, const int* z, const double* xyz, int s, int skip, const double* u, double* f, int off,
                     int aoff) {
  for (int c = 0; c < nat; ++c) {
    if (c == s || c == skip) continue;
    const double rb = dist_bohr(xyz, s, c), r = rb * BOHR_A;
    if (!(r < CUT_A)) continue;
    const double fc = 0.5 * (std::cos(PI * r / CUT_A) + 1.0);
    double cs = 0.0;
    for (int k = 0; k < 3; ++k) cs += u[k] * (xyz[3 * c + k] - xyz[3 * s + k]);
    cs /= rb;
    f[aoff] += fc * cs;
    f[aoff + 1] += fc * cs * cs;
    const int e = eidx(z[c]);
    for (int m = 0; m < 8; ++m) {
      const double t = (r - (1.0 + 3.0 * m / 7.0)) / 0.5;
      f[off + 8 * e + m] += fc * std::exp(-t * t);
    }
  }
}

static void features(int nat, const int* z, const double* xyz, int a, int b, double* f) {
  for (int k = 0; k < NF; ++k) f[k] = 0.0;
  double u[3];
  unit(xyz, a, b, u);
  const double r = dist_bohr(xyz, a, b) * BOHR_A;
  f[3 * eidx(z[a]) + eidx(z[b])] = 1.0;
  for (int k = 0; k < 32; ++k) {
    const double t = (r - (0.5 + 3.5 * k / 31.0)) / 0.125;
    f[9 + k] = std::exp(-t * t);
  }
  env_side(nat, z, xyz, a, b, u, f, 41, 89);
  const double v[3] = {-u[0], -u[1], -u[2]};
  env_side(nat, z, xyz, b, a, v, f, 65, 91);
  f[95] = 1.0;
}

static double silu(double x) { return x / (1.0 + std::exp(-x)); }
static double dsilu(double x) {
  const double s = 1.0 / (1.0 + std::exp(-x));
  return s * (1.0 + x * (1.0 - s));
}

struct Net {
  std::vector<double> w;  // parameters as used (bf16-rounded weights in rounded mode)
  bool r;
  Net(const float* p, bool rounded) : w(NPARAM), r(rounded) {
    for (int i = 0; i < NPARAM; ++i) {
      const bool is_bias = i >= OFF_B2 && i < OFF_W3 ? true : i >= OFF_B3;
      w[i] = is_bias ? (double)p[i] : rnd((double)p[i], rounded);
    }
  }
  double W1(int i, int h) const { return w[OFF_W1 + i * NH + h]; }
  double W2(int i, int h) const { return w[OFF_W2 + i * NH + h]; }
  double W3(int i, int o) const { return w[OFF_W3 + i * NO + o]; }
};

struct Act {
  double x[NF], z1[NH], h1[NH], z2[NH], h2[NH], y[NO];
};

static void forward(const Net& N, const uint16_t* xb, const double* xd, Act& A) {
  for (int i = 0; i < NF; ++i) A.x[i] = xb ? (double)bf16_val(xb[i]) : xd[i];
  for (int h = 0; h < NH; ++h) {
    double s = 0.0;
    for (int i = 0; i < NF; ++i) s += A.x[i] * N.W1(i, h);
    A.z1[h] = s;
    A.h1[h] = rnd(silu(s), N.r);
  }
  for (int h = 0; h < NH; ++h) {
    double s = N.w[OFF_B2 + h];
    for (int i = 0; i < NH; ++i) s += A.h1[i] * N.W2(i, h);
    A.z2[h] = s;
    A.h2[h] = rnd(silu(s), N.r);
  }
  for (int o = 0; o < NO; ++o) {
    double s = N.w[OFF_B3 + o];
    for (int i = 0; i < NH; ++i) s += A.h2[i] * N.W3(i, o);
    A.y[o] = s;
  }
}

static void backward(const Net& N, const Act& A, const double* dy_in, double* g) {
  double dy[NO], dz2[NH], dz1[NH];
  for (int o = 0; o < NO; ++o) dy[o] = rnd(dy_in[o], N.r);
  for (int h = 0; h < NH; ++h) {
    double s = 0.0;
    for (int o = 0; o < NO; ++o) s += N.W3(h, o) * dy[o];
    dz2[h] = rnd(s * dsilu(A.z2[h]), N.r);
  }
  for (int h = 0; h < NH; ++h) {
    double s = 0.0;
    for (int k = 0; k < NH; ++k) s += N.W2(h, k) * dz2[k];
    dz1[h] = rnd(s * dsilu(A.z1[h]), N.r);
  }
  for (int i = 0; i < NF; ++i)
    for (int h = 0; h < NH; ++h) g[OFF_W1 + i * NH + h] += A.x[i] * dz1[h];
  for (int i = 0; i < NH; ++i)
    for (int h = 0; h < NH; ++h) g[OFF_W2 + i * NH + h] += A.h1[i] * dz2[h];
  for (int h = 0; h < NH; ++h) g[OFF_B2 + h] += dz2[h];
  for (int i = 0; i < NH; ++i)
    for (int o = 0; o < NO; ++o) g[OFF_W3 + i * NO + o] += A.h2[i] * dy[o];
  for (int o = 0; o < NO; ++o) g[OFF_B3 + o] += dy[o];
}

static int norb(int z) { return z <= 2 ? 1 : 4; }

// Assembled density (n x n, input AO order) from per-directed-pair outputs y (16 each).
static void density(int nat, const int* z, const double* xyz, const std::vector<int>& src, const std::vector<int>& dst,
                    const float* y, double* P, int n) {
  std::vector<int> off(nat + 1, 0);
  for (int a = 0; a < nat; ++a) off[a + 1] = off[a] + norb(z[a]);
  std::memset(P, 0, sizeof(double) * (size_t)n * n);
  for (int a = 0; a < nat; ++a) {
    const int e = eidx(z[a]);
    P[(size_t)off[a] * n + off[a]] = e == 0 ? 1.0 : 2.0;
    if (norb(z[a]) == 4)
      for (int k = 1; k < 4; ++k) P[(size_t)(off[a] + k) * n + off[a] + k] = e == 1 ? 2.0 / 3.0 : 4.0 / 3.0;
  }
  const int np = (int)src.size();
  for (int j = 0; j < np; ++j) {
    const int a = src[j], b = dst[j], na = norb(z[a]), nb = norb(z[b]);
    const float* yj = y + (size_t)NO * j;
    double u[3];
    unit(xyz, a, b, u);
    // message to a's on-site block
    for (int m = 0; m < na; ++m)
      for (int l = 0; l < na; ++l) {
        double v;
        if (m == 0 && l == 0) v = yj[5];
        else if (m == 0) v = yj[6] * u[l - 1];
        else if (l == 0) v = yj[6] * u[m - 1];
        else v = yj[7] * u[m - 1] * u[l - 1] + yj[8] * ((m == l ? 1.0 : 0.0) - u[m - 1] * u[l - 1]);
        P[(size_t)(off[a] + m) * n + off[a] + l] += v;
      }
    // half of B(a->b) into (a, b) and its transpose into (b, a)
    for (int m = 0; m < na; ++m)
      for (int l = 0; l < nb; ++l) {
        double v;
        if (m == 0 && l == 0) v = yj[0];
        else if (m == 0) v = yj[1] * u[l - 1];
        else if (l == 0) v = yj[2] * u[m - 1];
        else v = yj[3] * u[m - 1] * u[l - 1] + yj[4] * ((m == l ? 1.0 : 0.0) - u[m - 1] * u[l - 1]);
        P[(size_t)(off[a] + m) * n + off[b] + l] += 0.5 * v;
        P[(size_t)(off[b] + l) * n + off[a] + m] += 0.5 * v;
      }
  }
}

}  // namespace gorc

using namespace gorc;

extern "C" {

int orc_guess_nparam() { return NPARAM; }

// Directed pairs; returns their count, filling src/dst when non-null (capacity nat*(nat-1)).
int orc_guess_pairs(int nat, const int* z, const double* xyz, int* src, int* dst) {
  (void)z;
  std::vector<int> s, d;
  pairs(nat, xyz, s, d);
  if (src) std::memcpy(src, s.data(), sizeof(int) * s.size());
  if (dst) std::memcpy(dst, d.data(), sizeof(int) * d.size());
  return (int)s.size();
}

// Features of every directed pair: f (double, before rounding) and x (bf16 bits), NF per pair.
int orc_guess_features(int nat, const int* z, const double* xyz, double* f, uint16_t* x) {
  for (int a = 0; a < nat; ++a)
    if (eidx(z[a]) < 0) return 4;
  std::vector<int> s, d;
  pairs(nat, xyz, s, d);
  std::vector<double> tmp(NF);
  for (size_t j = 0; j < s.size(); ++j) {
    features(nat, z, xyz, s[j], d[j], tmp.data());
    for (int k = 0; k < NF; ++k) {
      if (f) f[j * NF + k] = tmp[k];
      if (x) x[j * NF + k] = bf16_bits((float)tmp[k]);
    }
  }
  return 0;
}

// Network outputs y (n x NO) for inputs x (bf16 bits) or xd (double, when x is null).
// rounded != 0 mirrors the device's bf16 rounding points.
int orc_guess_forward(const float* params, int n, const uint16_t* x, const double* xd, int rounded, double* y) {
  Net N(params, rounded != 0);
  Act A;
  for (int p = 0; p < n; ++p) {
    forward(N, x ? x + (size_t)p * NF : nullptr, xd ? xd + (size_t)p * NF : nullptr, A);
    for (int o = 0; o < NO; ++o) y[(size_t)p * NO + o] = A.y[o];
  }
  return 0;
}

// Parameter gradient of sum_p dy_p . y_p (NPARAM doubles, accumulated from zero).
int orc_guess_backward(const float* params, int n, const uint16_t* x, const double* xd, const double* dy, int rounded,
                       double* grad) {
  Net N(params, rounded != 0);
  Act A;
  std::memset(grad, 0, sizeof(double) * NPARAM);
  for (int p = 0; p < n; ++p) {
    forward(N, x ? x + (size_t)p * NF : nullptr, xd ? xd + (size_t)p * NF : nullptr, A);
    backward(N, A, dy + (size_t)p * NO, grad);
  }
  return 0;
}

// Assembled guess density (n x n, input AO order) from outputs y (directed pairs in
// orc_guess_pairs order, NO floats each).
int orc_guess_density(int nat, const int* z, const double* xyz, const float* y, double* P) {
  int n = 0;
  for (int a = 0; a < nat; ++a) {
    if (eidx(z[a]) < 0) return 4;
    n += norb(z[a]);
  }
  std::vector<int> s, d;
  pairs(nat, xyz, s, d);
  density(nat, z, xyz, s, d, y, P, n);
  return 0;
}

// Training loss sum_a |P_aa - T_aa|^2 + sum_(a,b) |P_ab - T_ab|^2 over on-site blocks and
// directed pairs inside the cutoff, for target density T (n x n), and its gradient with
// respect to the outputs y (dy, NO per directed pair, unused outputs zero).
double orc_guess_loss(int nat, const int* z, const double* xyz, const float* y, const double* T, double* dy) {
  int n = 0;
  for (int a = 0; a < nat; ++a) n += norb(z[a]);
  std::vector<int> s, d;
  pairs(nat, xyz, s, d);
  std::vector<double> P((size_t)n * n);
  density(nat, z, xyz, s, d, y, P.data(), n);
  std::vector<int> off(nat + 1, 0);
  for (int a = 0; a < nat; ++a) off[a + 1] = off[a] + norb(z[a]);
  auto R = [&](int a, int b, int m, int l) {
    const size_t i = (size_t)(off[a] + m) * n + off[b] + l;
    return P[i] - T[i];
  };
  double loss = 0.0;
  for (int a = 0; a < nat; ++a)
    for (int m = 0; m < norb(z[a]); ++m)
      for (int l = 0; l < norb(z[a]); ++l) loss += R(a, a, m, l) * R(a, a, m, l);
  const int np = (int)s.size();
  for (int j = 0; j < np; ++j) {
    const int a = s[j], b = d[j], na = norb(z[a]), nb = norb(z[b]);
    for (int m = 0; m < na; ++m)
      for (int l = 0; l < nb; ++l) loss += R(a, b, m, l) * R(a, b, m, l);
    if (!dy) continue;
    double u[3];
    unit(xyz, a, b, u);
    double* g = dy + (size_t)NO * j;
    for (int o = 0; o < NO; ++o) g[o] = 0.0;
    for (int m = 0; m < na; ++m)
      for (int l = 0; l < nb; ++l) {
        const double r2 = 2.0 * R(a, b, m, l);
        if (m == 0 && l == 0) g[0] += r2;
        else if (m == 0) g[1] += r2 * u[l - 1];
        else if (l == 0) g[2] += r2 * u[m - 1];
        else {
          const double uu = u[m - 1] * u[l - 1];
          g[3] += r2 * uu;
          g[4] += r2 * ((m == l ? 1.0 : 0.0) - uu);
        }
      }
    for (int m = 0; m < na; ++m)
      for (int l = 0; l < na; ++l) {
        const double r2 = 2.0 * R(a, a, m, l);
        if (m == 0 && l == 0) g[5] += r2;
        else if (m == 0) g[6] += r2 * u[l - 1];
        else if (l == 0) g[6] += r2 * u[m - 1];
        else {
          const double uu = u[m - 1] * u[l - 1];
          g[7] += r2 * uu;
          g[8] += r2 * ((m == l ? 1.0 : 0.0) - uu);
        }
      }
  }
  return loss;
}

}  // extern "C"
