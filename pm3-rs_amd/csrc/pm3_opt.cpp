// Batched L-BFGS geometry optimisation on the PM3 surface: the optimiser of src/optimizer.rs
// (drive, :77-221, with optimize, :233-262) run for many molecules at once over the batched
// engine.  Each round uploads every running molecule's current point into one device-resident
// batch (pm3_batch_set_positions), solves SCF plus analytic gradient for all of them in one
// launch, and advances each molecule's own state machine: the first evaluation, then line-search
// trials.  Per molecule the steps are the reference's: the two-loop recursion over at most
// `history` pairs with the s.y / y.y initial scaling (0.1 / max|g| before any pair), the uphill
// guard, the backtracking Armijo search (c1 = 1e-4, halving down to 1e-8) and the curvature
// filter s.y > 1e-10.  The reference asks a trial point for its energy alone and, once the
// trial is accepted, evaluates the same point again for the gradient; both are the same SCF,
// so here the trial's solve carries the gradient and an accepted point costs no second solve.
// When fewer than half of a batch's molecules are still running, the batch is rebuilt from the
// running ones, so the last few slow optimisations do not pay for the finished ones.  With
// warm_start each round's SCF starts from the molecule's density at its previous point, carried
// into a rebuilt batch too (pm3_batch_carry_densities).
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pm3_gpu.h"

namespace {

struct OptState {
  int k = 0;    // caller index
  int a0 = 0;   // first atom in the caller's arrays
  int nat = 0;
  std::vector<double> x, g, d, xt;  // current point, its gradient, direction, point being solved
  std::vector<std::vector<double>> s, y;
  std::vector<double> rho;
  double energy = 0.0, heat = 0.0, maxg = 0.0, step = 1.0, gdd = 0.0;
  int iter = 0;         // index of the reference's outer loop
  bool trial = false;   // false: this round solves the first point; true: a line-search trial
  bool done = false;
  bool converged = false;
};

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

void axpy(std::vector<double>& y, double a, const std::vector<double>& x) {
  for (size_t i = 0; i < y.size(); ++i) y[i] += a * x[i];
}

double max_abs(const std::vector<double>& v) {
  double m = 0.0;
  for (double x : v) m = std::fmax(m, std::fabs(x));
  return m;
}

// Top of the reference's outer loop (src/optimizer.rs:111-158): stop, or the next direction
// and the first trial point of its line search.
void loop_top(OptState& o, const pm3_opt_options& opt, pm3_opt_result& r) {
  if (o.iter >= opt.max_iter) {
    o.done = true;
    return;
  }
  r.iterations = o.iter + 1;
  if (o.converged) {
    o.done = true;
    return;
  }
  std::vector<double> q = o.g;
  const int m = (int)o.s.size();
  std::vector<double> alpha(m, 0.0);
  for (int i = m - 1; i >= 0; --i) {
    const double a = o.rho[i] * dot(o.s[i], q);
    alpha[i] = a;
    axpy(q, -a, o.y[i]);
  }
  double gamma;
  if (m > 0) {
    const double sy = dot(o.s[m - 1], o.y[m - 1]);
    const double yy = dot(o.y[m - 1], o.y[m - 1]);
    gamma = yy > 0.0 ? sy / yy : 1.0;
  } else {
    gamma = 0.1 / std::fmax(o.maxg, 1.0e-6);
  }
  for (double& v : q) v *= gamma;
  for (int i = 0; i < m; ++i) {
    const double beta = o.rho[i] * dot(o.y[i], q);
    axpy(q, alpha[i] - beta, o.s[i]);
  }
  o.d.resize(q.size());
  for (size_t i = 0; i < q.size(); ++i) o.d[i] = -q[i];
  if (dot(o.d, o.g) > 0.0)
    for (size_t i = 0; i < q.size(); ++i) o.d[i] = -o.g[i];
  o.gdd = dot(o.g, o.d);
  o.step = 1.0;
  o.xt = o.x;
  axpy(o.xt, o.step, o.d);
  o.trial = true;
}

// Next trial of the line search after a rejected one (src/optimizer.rs:161-181); false once the
// step has fallen below 1e-8 (the optimisation stops at the current point).
bool next_trial(OptState& o) {
  o.step *= 0.5;
  if (o.step < 1.0e-8) return false;
  o.xt = o.x;
  axpy(o.xt, o.step, o.d);
  return true;
}

// One OptStep of the trajectory (src/optimizer.rs:121-126,207-212): the point just accepted.
void record(const OptState& o, pm3_opt_result& r, size_t cap, pm3_opt_step* steps, double* traj_xyz) {
  const size_t j = (size_t)r.trajectory_len;
  if (j >= cap) return;
  if (steps) steps[(size_t)o.k * cap + j] = pm3_opt_step{o.energy, o.heat, o.maxg};
  if (traj_xyz) std::memcpy(traj_xyz + 3 * (cap * (size_t)o.a0 + j * (size_t)o.nat), o.x.data(), sizeof(double) * o.x.size());
  r.trajectory_len += 1;
}

bool coincident(const std::vector<double>& x) {
  const size_t n = x.size() / 3;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < i; ++j) {
      const double dx = x[3 * i] - x[3 * j], dy = x[3 * i + 1] - x[3 * j + 1], dz = x[3 * i + 2] - x[3 * j + 2];
      if (!(dx * dx + dy * dy + dz * dz > 1.0e-16)) return true;
    }
  return false;
}

}  // namespace

extern "C" {

void pm3_opt_options_default(pm3_opt_options* out) {
  if (!out) return;
  out->max_iter = 200;
  out->gtol = 1.0e-3;
  out->grad_step = 5.0e-4;
  out->history = 8;
  out->warm_start = 0;
}

pm3_status pm3_optimize_batch(int32_t n_mol, const int32_t* atom_ptr, const int32_t* z, const double* xyz_bohr,
                              const double* charge, const int32_t* multiplicity, const pm3_params* params,
                              const pm3_options* opts, const pm3_opt_options* opt_in, pm3_opt_result* results,
                              double* xyz_out, pm3_opt_step* steps, double* traj_xyz) {
  if (n_mol < 0 || !params || (n_mol > 0 && (!atom_ptr || !z || !xyz_bohr || !results || !xyz_out)))
    return PM3_INVALID_INPUT;
  if (n_mol > 0 && atom_ptr[0] != 0) return PM3_INVALID_INPUT;
  for (int k = 0; k < n_mol; ++k)
    if (atom_ptr[k + 1] < atom_ptr[k]) return PM3_INVALID_INPUT;
  pm3_opt_options opt;
  pm3_opt_options_default(&opt);
  if (opt_in) opt = *opt_in;
  if (opt.max_iter < 0 || opt.history < 0 || !(opt.gtol >= 0.0)) return PM3_INVALID_INPUT;

  // trajectory slots per molecule; the number of accepted points never exceeds max_iter
  const size_t cap = (size_t)opt.max_iter + 1;
  std::vector<OptState> st(n_mol);
  for (int k = 0; k < n_mol; ++k) {
    OptState& o = st[k];
    o.k = k;
    o.a0 = atom_ptr[k];
    o.nat = atom_ptr[k + 1] - o.a0;
    o.x.assign(xyz_bohr + 3 * (size_t)o.a0, xyz_bohr + 3 * (size_t)(o.a0 + o.nat));
    o.xt = o.x;
    pm3_opt_result& r = results[k];
    std::memset(&r, 0, sizeof(r));
  }

  pm3_batch* batch = nullptr;
  std::vector<int> members;  // caller indices of the batch's molecules, in batch order
  std::vector<int32_t> bptr, bz, bm;
  std::vector<double> bx, bq, bgrad;
  std::vector<pm3_status> bst;
  std::vector<pm3_result> bres;
  pm3_status fatal = PM3_OK;

  for (;;) {
    std::vector<int> run;
    for (int k = 0; k < n_mol; ++k)
      if (!st[k].done) run.push_back(k);
    if (run.empty()) break;
    // a trial point with coincident atoms cannot be solved: it is a rejected trial
    bool again = false;
    for (int k : run) {
      OptState& o = st[k];
      while (!o.done && o.trial && coincident(o.xt))
        if (!next_trial(o)) o.done = true;
      again = again || o.done;
    }
    if (again) continue;
    if (!batch || 2 * run.size() <= members.size()) {
      pm3_batch* old = batch;
      const std::vector<int> old_members = members;
      batch = nullptr;
      members = run;
      bptr.assign(1, 0);
      bz.clear();
      bx.clear();
      bq.clear();
      bm.clear();
      for (int k : members) {
        const OptState& o = st[k];
        bptr.push_back(bptr.back() + o.nat);
        bz.insert(bz.end(), z + o.a0, z + o.a0 + o.nat);
        bx.insert(bx.end(), o.xt.begin(), o.xt.end());
        bq.push_back(charge ? charge[k] : 0.0);
        bm.push_back(multiplicity ? multiplicity[k] : 1);
      }
      const pm3_status s = pm3_batch_create((int32_t)members.size(), bptr.data(), bz.data(), bx.data(), bq.data(),
                                            bm.data(), params, opts,
                                            PM3_BATCH_WANT_GRADIENT | PM3_BATCH_KEEP_BATCHED |
                                                (opt.warm_start ? PM3_BATCH_WARM_START : 0),
                                            &batch);
      pm3_status carried = PM3_OK;
      if (s == PM3_OK && old && opt.warm_start) {
        // the running molecules keep their densities across the rebuild
        std::vector<int32_t> pos(n_mol, -1), src(members.size());
        for (size_t p = 0; p < old_members.size(); ++p) pos[old_members[p]] = (int32_t)p;
        for (size_t p = 0; p < members.size(); ++p) src[p] = pos[members[p]];
        carried = pm3_batch_carry_densities(batch, old, src.data());
      }
      if (old) pm3_batch_destroy(old);
      if (s != PM3_OK || carried != PM3_OK) {
        fatal = s != PM3_OK ? s : carried;
        break;
      }
      bgrad.assign(bx.size(), 0.0);
      bst.assign(members.size(), PM3_OK);
      bres.assign(members.size(), pm3_result{});
    } else {
      // finished members keep their last point; their solve is discarded
      for (size_t p = 0; p < members.size(); ++p) {
        const OptState& o = st[members[p]];
        const std::vector<double>& src = o.done ? o.x : o.xt;
        std::memcpy(bx.data() + 3 * (size_t)bptr[p], src.data(), sizeof(double) * src.size());
      }
      const pm3_status s = pm3_batch_set_positions(batch, bx.data());
      if (s != PM3_OK) {
        fatal = s;
        break;
      }
    }
    pm3_status s = pm3_batch_solve(batch);
    if (s == PM3_OK) s = pm3_batch_results(batch, bst.data(), bres.data(), nullptr, nullptr);
    if (s == PM3_OK) s = pm3_batch_gradient(batch, bgrad.data());
    if (s != PM3_OK) {
      fatal = s;
      break;
    }
    for (size_t p = 0; p < members.size(); ++p) {
      OptState& o = st[members[p]];
      if (o.done) continue;
      pm3_opt_result& r = results[o.k];
      r.evaluations += 1;
      const pm3_result& e = bres[p];
      const bool ok = bst[p] == PM3_OK && e.converged;
      std::vector<double> gn(bgrad.begin() + 3 * (size_t)bptr[p], bgrad.begin() + 3 * (size_t)(bptr[p] + o.nat));
      if (!o.trial) {
        // the first evaluation (src/optimizer.rs:88-107): its failure is the molecule's error
        if (!ok) {
          r.status = bst[p] != PM3_OK ? bst[p] : PM3_SCF_NOT_CONVERGED;
          o.done = true;
          continue;
        }
        o.energy = e.total_ev;
        o.heat = e.heat_of_formation_kcal;
        o.g = gn;
        o.maxg = max_abs(gn);
        r.first_energy_ev = o.energy;
        o.converged = o.maxg < opt.gtol;
        o.iter = 0;
        record(o, r, cap, steps, traj_xyz);
        loop_top(o, opt, r);
        continue;
      }
      if (ok && e.total_ev <= o.energy + 1.0e-4 * o.step * o.gdd) {
        // accepted: the L-BFGS memory update (src/optimizer.rs:183-197), then the next iteration
        std::vector<double> sv(o.x.size()), yv(o.x.size());
        for (size_t i = 0; i < sv.size(); ++i) {
          sv[i] = o.xt[i] - o.x[i];
          yv[i] = gn[i] - o.g[i];
        }
        const double sy = dot(sv, yv);
        if (sy > 1.0e-10) {
          o.s.push_back(std::move(sv));
          o.y.push_back(std::move(yv));
          o.rho.push_back(1.0 / sy);
          if ((int)o.s.size() > opt.history) {
            o.s.erase(o.s.begin());
            o.y.erase(o.y.begin());
            o.rho.erase(o.rho.begin());
          }
        }
        o.x = o.xt;
        o.g = gn;
        o.energy = e.total_ev;
        o.heat = e.heat_of_formation_kcal;
        o.maxg = max_abs(gn);
        o.converged = o.maxg < opt.gtol;
        o.iter += 1;
        record(o, r, cap, steps, traj_xyz);
        loop_top(o, opt, r);
      } else if (!next_trial(o)) {
        o.done = true;  // no progress possible: stop at the current point
      }
    }
  }
  if (batch) pm3_batch_destroy(batch);
  for (int k = 0; k < n_mol; ++k) {
    const OptState& o = st[k];
    pm3_opt_result& r = results[k];
    if (fatal != PM3_OK && r.status == PM3_OK) r.status = fatal;
    r.converged = o.converged ? 1 : 0;
    r.energy_ev = o.energy;
    r.heat_of_formation_kcal = o.heat;
    r.max_gradient = o.maxg;
    std::memcpy(xyz_out + 3 * (size_t)o.a0, o.x.data(), sizeof(double) * o.x.size());
  }
  return fatal;
}

pm3_status pm3_optimize(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                        const pm3_opt_options* opt, pm3_opt_result* result, double* xyz_out, pm3_opt_step* steps,
                        double* traj_xyz) {
  if (!mol || !params || !result || !xyz_out || mol->n_atoms < 1 || !mol->z || !mol->xyz_bohr)
    return PM3_INVALID_INPUT;
  if (mol->multiplicity < 1) return PM3_INVALID_INPUT;
  const int32_t ptr[2] = {0, mol->n_atoms};
  const double q = mol->charge;
  const int32_t mult = mol->multiplicity;
  const pm3_status s =
      pm3_optimize_batch(1, ptr, mol->z, mol->xyz_bohr, &q, &mult, params, opts, opt, result, xyz_out, steps, traj_xyz);
  return s != PM3_OK ? s : (pm3_status)result->status;
}

// optimize_dc (src/optimizer.rs:277-310): the same optimiser for one molecule over the
// divide-and-conquer gradient.  Every point is one pm3_dc_gradient call (run_dc, then the gradient
// at the partitioned density); a trial's call carries its gradient, so an accepted point costs no
// second solve (the reference asks a trial for run_dc's energy alone and evaluates the accepted
// point again: the same SCF).  A trial that fails in any way is a rejected step, as the reference's
// energy_only maps an error to None; a failure of the first evaluation is the call's error.
pm3_status pm3_optimize_dc(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                           const pm3_dc_options* dc, const pm3_opt_options* opt_in, pm3_opt_result* result,
                           pm3_dc_result* scf, double* xyz_out, pm3_opt_step* steps, double* traj_xyz) {
  if (!mol || !params || !result || !xyz_out || mol->n_atoms < 1 || !mol->z || !mol->xyz_bohr)
    return PM3_INVALID_INPUT;
  if (mol->multiplicity < 1) return PM3_INVALID_INPUT;
  pm3_opt_options opt;
  pm3_opt_options_default(&opt);
  if (opt_in) opt = *opt_in;
  if (opt.max_iter < 0 || opt.history < 0 || !(opt.gtol >= 0.0)) return PM3_INVALID_INPUT;
  const size_t cap = (size_t)opt.max_iter + 1;
  std::memset(result, 0, sizeof(*result));
  OptState o;
  o.nat = mol->n_atoms;
  o.x.assign(mol->xyz_bohr, mol->xyz_bohr + 3 * (size_t)o.nat);
  o.xt = o.x;
  auto eval = [&](const std::vector<double>& x, pm3_dc_result& r, std::vector<double>& g) {
    pm3_molecule m = *mol;
    m.xyz_bohr = x.data();
    g.assign(x.size(), 0.0);
    return pm3_dc_gradient(&m, params, opts, dc, &r, g.data(), nullptr, nullptr, nullptr);
  };
  pm3_dc_result cur{};
  std::vector<double> g;
  pm3_status s = eval(o.x, cur, g);
  result->evaluations = 1;
  if (s != PM3_OK) {
    result->status = s;
    if (scf) *scf = cur;
    std::memcpy(xyz_out, o.x.data(), sizeof(double) * o.x.size());
    return s;
  }
  o.energy = cur.total_ev;
  o.heat = cur.heat_of_formation_kcal;
  o.g = g;
  o.maxg = max_abs(g);
  result->first_energy_ev = o.energy;
  o.converged = o.maxg < opt.gtol;
  record(o, *result, cap, steps, traj_xyz);
  loop_top(o, opt, *result);
  while (!o.done) {
    if (coincident(o.xt)) {  // a trial point with coincident atoms cannot be solved: rejected
      if (!next_trial(o)) o.done = true;
      continue;
    }
    pm3_dc_result tr{};
    std::vector<double> gn;
    s = eval(o.xt, tr, gn);
    result->evaluations += 1;
    if (s == PM3_OK && tr.total_ev <= o.energy + 1.0e-4 * o.step * o.gdd) {
      // accepted: the L-BFGS memory update (src/optimizer.rs:183-197), then the next iteration
      std::vector<double> sv(o.x.size()), yv(o.x.size());
      for (size_t i = 0; i < sv.size(); ++i) {
        sv[i] = o.xt[i] - o.x[i];
        yv[i] = gn[i] - o.g[i];
      }
      const double sy = dot(sv, yv);
      if (sy > 1.0e-10) {
        o.s.push_back(std::move(sv));
        o.y.push_back(std::move(yv));
        o.rho.push_back(1.0 / sy);
        if ((int)o.s.size() > opt.history) {
          o.s.erase(o.s.begin());
          o.y.erase(o.y.begin());
          o.rho.erase(o.rho.begin());
        }
      }
      o.x = o.xt;
      o.g = gn;
      cur = tr;
      o.energy = tr.total_ev;
      o.heat = tr.heat_of_formation_kcal;
      o.maxg = max_abs(gn);
      o.converged = o.maxg < opt.gtol;
      o.iter += 1;
      record(o, *result, cap, steps, traj_xyz);
      loop_top(o, opt, *result);
    } else if (!next_trial(o)) {
      o.done = true;  // no progress possible: stop at the current point
    }
  }
  result->converged = o.converged ? 1 : 0;
  result->energy_ev = o.energy;
  result->heat_of_formation_kcal = o.heat;
  result->max_gradient = o.maxg;
  if (scf) *scf = cur;
  std::memcpy(xyz_out, o.x.data(), sizeof(double) * o.x.size());
  return PM3_OK;
}

}  // extern "C"
