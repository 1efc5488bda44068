// The Cartesian Hessian by central differences of the analytic gradient (numerical_hessian,
// src/hessian.rs:268-322): column j is (g(x + h e_j) - g(x - h e_j)) / 2h, and the matrix is
// symmetrised by averaging.  The reference runs the 6N displaced gradients one after another on a
// thread pool; here all of them are one batch, solved in one persistent launch with their
// gradients, under the reference's tightened SCF options (tighten_scf_for_hessian, :169-175).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pm3_gpu.h"

extern "C" {

pm3_status pm3_numerical_hessian(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts_in,
                                 double step, double* hessian) {
  if (!mol || !params || !hessian || mol->n_atoms < 1 || !mol->z || !mol->xyz_bohr) return PM3_INVALID_INPUT;
  if (!(step > 0.0) || !std::isfinite(step)) return PM3_INVALID_INPUT;
  pm3_options opts;
  pm3_options_default(&opts);
  if (opts_in) opts = *opts_in;
  opts.e_tol = std::min(opts.e_tol, 1.0e-12);
  opts.p_tol = std::min(opts.p_tol, 1.0e-11);
  opts.max_scf = std::max(opts.max_scf, 500);
  const int nat = mol->n_atoms, ndof = 3 * nat, nmol = 2 * ndof;
  std::vector<int32_t> ptr(nmol + 1), z((size_t)nmol * nat), mult(nmol, mol->multiplicity);
  std::vector<double> xyz((size_t)nmol * ndof), q(nmol, mol->charge);
  for (int k = 0; k < nmol; ++k) {
    ptr[k] = k * nat;
    std::memcpy(z.data() + (size_t)k * nat, mol->z, sizeof(int32_t) * nat);
    double* x = xyz.data() + (size_t)k * ndof;
    std::memcpy(x, mol->xyz_bohr, sizeof(double) * ndof);
    x[k / 2] += (k % 2 == 0 ? step : -step);  // molecule 2j: +h along coordinate j, 2j + 1: -h
  }
  ptr[nmol] = nmol * nat;
  std::vector<pm3_status> st(nmol);
  std::vector<pm3_result> res(nmol);
  std::vector<double> g((size_t)nmol * ndof);
  const pm3_status s = pm3_run_batch(nmol, ptr.data(), z.data(), xyz.data(), q.data(), mult.data(), params, &opts, 1,
                                     st.data(), res.data(), g.data());
  if (s != PM3_OK) return s;
  for (int k = 0; k < nmol; ++k)
    if (st[k] != PM3_OK) return st[k];
  for (int j = 0; j < ndof; ++j) {
    const double* gp = g.data() + (size_t)(2 * j) * ndof;
    const double* gm = g.data() + (size_t)(2 * j + 1) * ndof;
    for (int i = 0; i < ndof; ++i) hessian[(size_t)i * ndof + j] = (gp[i] - gm[i]) / (2.0 * step);
  }
  for (int i = 0; i < ndof; ++i)
    for (int j = i + 1; j < ndof; ++j) {
      const double v = 0.5 * (hessian[(size_t)i * ndof + j] + hessian[(size_t)j * ndof + i]);
      hessian[(size_t)i * ndof + j] = v;
      hessian[(size_t)j * ndof + i] = v;
    }
  return PM3_OK;
}

}  // extern "C"
