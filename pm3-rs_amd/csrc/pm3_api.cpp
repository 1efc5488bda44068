// C ABI of libpm3gpu.so (declared in include/pm3_gpu.h).  Thin layer: validation and the error
// surface of the reference (src/error.rs:12-47), option defaults (src/scf.rs:259-287), heat of
// formation bookkeeping (src/scf.rs:745-752); all numerical work is on the device.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pm3_gpu.h"
#include "../../include/pm3_guess.h"
#include "pm3_guess.hpp"
#include "pm3_host.hpp"

using namespace pm3host;

struct pm3_params {
  Params p;
};

struct pm3_guess_model {
  GuessModel* m = nullptr;
};

struct pm3_guess_data {
  GuessData* d = nullptr;
};

struct pm3_batch {
  Batch* b = nullptr;
  std::vector<int32_t> z;
  std::vector<double> xyz;
  std::vector<MolIn> mols;
  const pm3_params* params = nullptr;
  std::vector<MolResult> results;
  bool fetched = false;
  bool want_density = false;  // created with PM3_BATCH_WANT_DENSITY
};

namespace {

thread_local std::string g_last_error;

pm3_status fail(pm3_status s, const std::string& msg) {
  g_last_error = msg;
  return s;
}

const char* status_text(int s) {
  switch (s) {
    case PM3_OK: return "ok";
    case PM3_INVALID_INPUT: return "invalid input";
    case PM3_MISSING_ELEMENT: return "missing element";
    case PM3_SCF_NOT_CONVERGED: return "SCF not converged";
    case PM3_UNSUPPORTED: return "unsupported";
    case PM3_RESOURCE_LIMIT: return "resource limit";
    default: return "device error";
  }
}

bool engine_options(const pm3_options* o, EngineOptions& e, std::string& why) {
  pm3_options d;
  pm3_options_default(&d);
  if (!o) o = &d;
  if (o->max_scf < 1) { why = "max_scf must be >= 1"; return false; }
  if (!(o->e_tol >= 0.0) || !(o->p_tol >= 0.0)) { why = "tolerances must be >= 0"; return false; }
  if (o->accelerator < 0 || o->accelerator > 2) { why = "unknown accelerator"; return false; }
  if (o->guess < 0 || o->guess > 4) { why = "unknown guess"; return false; }
  if (o->stability < 0 || o->stability > 2) { why = "unknown stability mode"; return false; }
  if (o->reference < 0 || o->reference > 2) { why = "unknown reference"; return false; }
  if (!(o->damping >= 0.0 && o->damping < 1.0)) { why = "damping must be in [0, 1)"; return false; }
  if (!std::isfinite(o->level_shift_ev)) { why = "level shift must be finite"; return false; }
  if (o->has_field && !(std::isfinite(o->field[0]) && std::isfinite(o->field[1]) && std::isfinite(o->field[2]))) {
    why = "field must be finite";
    return false;
  }
  if (o->integral_memory_mb < 0) { why = "integral_memory_mb must be >= 0"; return false; }
  e.max_scf = o->max_scf;
  e.e_tol = o->e_tol;
  e.p_tol = o->p_tol;
  e.accel = o->use_diis ? o->accelerator : 0;
  e.adiis_switch = o->adiis_switch;
  e.guess = o->guess == PM3_GUESS_AUTO ? 1 : o->guess;
  e.guess_model = nullptr;
  if (o->guess == PM3_GUESS_LEARNED) {
    if (!o->guess_model || !o->guess_model->m) { why = "PM3_GUESS_LEARNED needs options.guess_model"; return false; }
    e.guess_model = o->guess_model->m;
  }
  e.stability = o->stability;
  e.scf_memory_mb = o->scf_memory_mb;
  e.use_diis = o->use_diis ? 1 : 0;
  e.level_shift = o->level_shift_ev;
  e.damping = o->damping;
  e.reference = o->reference;
  e.has_field = o->has_field ? 1 : 0;
  for (int x = 0; x < 3; ++x) e.field[x] = o->has_field ? o->field[x] : 0.0;
  e.integral_memory_mb = o->integral_memory_mb;
  return true;
}

void fill_result(const MolResult& r, pm3_result* out) {
  std::memset(out, 0, sizeof(*out));
  out->electronic_ev = r.e_elec;
  out->core_ev = r.e_core;
  out->total_ev = r.e_elec + r.e_core;
  out->heat_of_formation_kcal = r.hof_kcal;
  for (int k = 0; k < 3; ++k) out->dipole_debye[k] = r.dipole[k];
  out->dipole_magnitude = std::sqrt(r.dipole[0] * r.dipole[0] + r.dipole[1] * r.dipole[1] + r.dipole[2] * r.dipole[2]);
  out->homo_ev = r.homo;
  out->lumo_ev = r.lumo;
  out->has_homo = r.nocc >= 1;
  out->has_lumo = r.nocc < r.nao;
  out->last_change = r.last_change;
  out->scf_improvement_ev = r.improvement;
  out->iterations = r.iterations;
  out->converged = r.converged;
  out->scf_paths_tried = r.paths;
  out->n_occ = r.nocc;
  out->nao = r.nao;
  out->guess_used = r.guess_used;
  out->unrestricted = r.unrestricted;
  out->n_beta = r.nbeta;
  out->missing_z = r.missing_z;
  out->required_mb = r.required_mb;
  out->limit_mb = r.limit_mb;
  if (r.unrestricted) {
    out->has_homo = r.nocc >= 1 || r.nbeta >= 1;
    out->has_lumo = r.nocc < r.nao || r.nbeta < r.nao;
  }
}

pm3_status check_molecule(const pm3_molecule* mol) {
  if (!mol || mol->n_atoms < 1 || !mol->z || !mol->xyz_bohr) return fail(PM3_INVALID_INPUT, "empty molecule");
  if (mol->multiplicity < 1) return fail(PM3_INVALID_INPUT, "multiplicity must be >= 1");
  for (int i = 0; i < 3 * mol->n_atoms; ++i)
    if (!std::isfinite(mol->xyz_bohr[i])) return fail(PM3_INVALID_INPUT, "non-finite coordinate");
  return PM3_OK;
}

// The error text of a molecule's status, with the details Pm3Error carries (src/error.rs:12-47).
pm3_status molecule_status(int st, const MolResult& r) {
  if (st == PM3_OK) return PM3_OK;
  if (st == PM3_SCF_NOT_CONVERGED)
    return fail(PM3_SCF_NOT_CONVERGED, "SCF did not converge in " + std::to_string(r.iterations) +
                                           " iterations (last RMS density change " + std::to_string(r.last_change) + ")");
  if (st == PM3_MISSING_ELEMENT)
    return fail(PM3_MISSING_ELEMENT, "missing PM3 parameter block for Z=" + std::to_string(r.missing_z));
  if (st == PM3_UNSUPPORTED && r.missing_z > 0)
    return fail(PM3_UNSUPPORTED, "this engine covers the s/p elements H, Li-F, Na-Cl; Z=" + std::to_string(r.missing_z) +
                                     " is not among them");
  if (st == PM3_UNSUPPORTED)
    return fail(PM3_UNSUPPORTED, "unrestricted molecules above the single-workgroup size (about 190 AOs) are not handled");
  if (st == PM3_RESOURCE_LIMIT && r.limit_mb > 0.0)
    return fail(PM3_RESOURCE_LIMIT, "PM3 two-electron pair cache needs " + std::to_string((long long)r.required_mb) +
                                        " MiB, above the limit of " + std::to_string((long long)r.limit_mb) + " MiB");
  return fail((pm3_status)st, status_text(st));
}

// Every atom's electron count and the multiplicity's parity, and RHF on an open shell, are input
// errors the engine also reports per molecule; this gives run_pm3's message for the RHF case.
pm3_status reference_check(const pm3_molecule* mol, const pm3_params* params, const EngineOptions& eo) {
  if (eo.reference != 1 || mol->multiplicity == 1) return PM3_OK;
  double ne = -mol->charge;
  for (int i = 0; i < mol->n_atoms; ++i) {
    const int z = mol->z[i];
    if (z < 0 || z >= 128 || params->p.slot_of_z[z] < 0) return PM3_OK;  // reported as a missing element
    ne += params->p.table.el[params->p.slot_of_z[z]].tore;
  }
  const long n = std::lround(ne), nu = mol->multiplicity - 1;
  if (n - nu >= 0 && (n - nu) % 2 == 0)
    return fail(PM3_INVALID_INPUT, "RHF requested for an open-shell system (n_alpha=" + std::to_string((n + nu) / 2) +
                                       " != n_beta=" + std::to_string((n - nu) / 2) + "); use UHF or Auto");
  return PM3_OK;
}

}  // namespace

extern "C" {

const pm3_params* pm3_params_standard(void) {
  static pm3_params* g = nullptr;
  static std::once_flag once;
  std::call_once(once, [] {
    pm3_params* p = new pm3_params();
    if (build_params(p->p)) g = p;
    else delete p;
  });
  if (!g) g_last_error = "failed to build the PM3 parameter table";
  return g;
}

void pm3_options_default(pm3_options* out) {
  out->max_scf = 200;
  out->e_tol = 1.0e-8;
  out->p_tol = 1.0e-7;
  out->use_diis = 1;
  out->accelerator = PM3_ACCEL_ADIIS_CDIIS;
  out->adiis_switch = 0.1;
  out->guess = PM3_GUESS_AUTO;
  out->stability = PM3_STABILITY_AUTO;
  out->scf_memory_mb = 512;
  out->guess_model = nullptr;
  out->level_shift_ev = 0.0;
  out->damping = 0.0;
  out->reference = PM3_REFERENCE_AUTO;
  out->has_field = 0;
  out->field[0] = out->field[1] = out->field[2] = 0.0;
  out->integral_memory_mb = 0;
}

int32_t pm3_query_nao(const pm3_molecule* mol, const pm3_params* params) {
  if (!mol || !params) return -1;
  int n = 0;
  for (int i = 0; i < mol->n_atoms; ++i) {
    const int z = mol->z[i];
    if (z < 0 || z >= 128 || params->p.slot_of_z[z] < 0) return -1;
    n += params->p.table.el[params->p.slot_of_z[z]].norb;
  }
  return n;
}

double pm3_element_mass(const pm3_params* params, int32_t z) {
  if (!params || z < 0 || z >= 128 || params->p.slot_of_z[z] < 0) return std::nan("");
  return params->p.table.el[params->p.slot_of_z[z]].mass;
}

pm3_status pm3_run(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts, pm3_result* result,
                   double* density, double* mo_energies, double* mo_coeff, double* charges, double* spin_density,
                   double* mo_energies_beta, double* mo_coeff_beta) {
  g_last_error.clear();
  if (!params) return fail(PM3_INVALID_INPUT, "null parameter set");
  pm3_status s = check_molecule(mol);
  if (s != PM3_OK) return s;
  EngineOptions eo;
  std::string why;
  if (!engine_options(opts, eo, why)) return fail(PM3_INVALID_INPUT, why);
  if ((s = reference_check(mol, params, eo)) != PM3_OK) return s;
  std::vector<MolIn> mols{MolIn{mol->n_atoms, mol->z, mol->xyz_bohr, mol->charge, mol->multiplicity}};
  RunRequest req;
  req.solo_large = true;
  req.want_density = density != nullptr || spin_density != nullptr;
  req.want_orbitals = mo_energies != nullptr || mo_coeff != nullptr || mo_energies_beta != nullptr ||
                      mo_coeff_beta != nullptr;
  std::vector<MolResult> out;
  std::string err;
  if (!engine_run(params->p, mols, eo, req, out, err)) return fail(PM3_DEVICE_ERROR, err);
  const MolResult& r = out[0];
  if (result) fill_result(r, result);
  if (r.status == PM3_OK || r.status == PM3_SCF_NOT_CONVERGED) {
    const size_t n2 = (size_t)r.nao * r.nao, n = (size_t)r.nao;
    auto put = [](double* dst, const std::vector<double>& v, size_t count) {
      if (!dst) return;
      if (v.size() == count) std::memcpy(dst, v.data(), count * sizeof(double));
      else std::memset(dst, 0, count * sizeof(double));
    };
    put(density, r.density, n2);
    put(mo_coeff, r.mo_coeff, n2);
    put(mo_energies, r.mo_energies, n);
    put(charges, r.charges, (size_t)mol->n_atoms);
    put(spin_density, r.spin_density, n2);
    put(mo_energies_beta, r.mo_energies_beta, n);
    put(mo_coeff_beta, r.mo_coeff_beta, n2);
  }
  return molecule_status(r.status, r);
}

pm3_status pm3_run_noalloc(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                           pm3_result* result, double* density, double* charges) {
  static_assert(PM3_NOALLOC_MAX_ATOMS == NOALLOC_MAX_ATOMS && PM3_NOALLOC_MAX_NAO == NOALLOC_MAX_NAO,
                "header and engine limits of the allocation-free entry point");
  g_last_error.clear();
  if (!params) return fail(PM3_INVALID_INPUT, "null parameter set");
  pm3_status s = check_molecule(mol);
  if (s != PM3_OK) return s;
  EngineOptions eo;
  std::string why;
  if (!engine_options(opts, eo, why)) return fail(PM3_INVALID_INPUT, why);
  if ((s = reference_check(mol, params, eo)) != PM3_OK) return s;
  const int nao = pm3_query_nao(mol, params);
  if (nao >= 0 && (mol->n_atoms > NOALLOC_MAX_ATOMS || nao > NOALLOC_MAX_NAO))
    return fail(PM3_RESOURCE_LIMIT, "the allocation-free entry point takes at most " + std::to_string(NOALLOC_MAX_ATOMS) +
                                        " atoms and " + std::to_string(NOALLOC_MAX_NAO) + " AOs (this molecule: " +
                                        std::to_string(mol->n_atoms) + " atoms, " + std::to_string(nao) + " AOs)");
  MolResult r;
  std::string err;
  const int code = engine_run_noalloc(params->p, MolIn{mol->n_atoms, mol->z, mol->xyz_bohr, mol->charge,
                                                       mol->multiplicity}, eo, density != nullptr, r, err);
  if (code != 0) return fail((pm3_status)code, err);
  if (result) fill_result(r, result);
  if (r.status == PM3_OK || r.status == PM3_SCF_NOT_CONVERGED) {
    const size_t n2 = (size_t)r.nao * r.nao;
    if (density && r.density.size() == n2) std::memcpy(density, r.density.data(), n2 * sizeof(double));
    if (charges && (int)r.charges.size() == mol->n_atoms) std::memcpy(charges, r.charges.data(), mol->n_atoms * sizeof(double));
  }
  return molecule_status(r.status, r);
}

pm3_status pm3_batch_create(int32_t n_mol, const int32_t* atom_ptr, const int32_t* z, const double* xyz_bohr,
                            const double* charge, const int32_t* multiplicity, const pm3_params* params,
                            const pm3_options* opts, int32_t flags, pm3_batch** out) {
  g_last_error.clear();
  if (!out) return fail(PM3_INVALID_INPUT, "null batch handle");
  *out = nullptr;
  if (!params || n_mol < 0 || (n_mol > 0 && (!atom_ptr || !z || !xyz_bohr))) return fail(PM3_INVALID_INPUT, "bad batch arguments");
  EngineOptions eo;
  std::string why;
  if (!engine_options(opts, eo, why)) return fail(PM3_INVALID_INPUT, why);
  pm3_batch* h = new pm3_batch();
  const int na = n_mol > 0 ? atom_ptr[n_mol] : 0;
  if (n_mol > 0 && atom_ptr[0] != 0) { delete h; return fail(PM3_INVALID_INPUT, "atom_ptr[0] must be 0"); }
  for (int k = 0; k < n_mol; ++k)
    if (atom_ptr[k + 1] < atom_ptr[k]) { delete h; return fail(PM3_INVALID_INPUT, "atom_ptr must be non-decreasing"); }
  h->z.assign(z, z + na);
  h->xyz.assign(xyz_bohr, xyz_bohr + 3 * (size_t)na);
  h->params = params;
  for (int k = 0; k < n_mol; ++k) {
    const int a0 = atom_ptr[k];
    h->mols.push_back(MolIn{atom_ptr[k + 1] - a0, h->z.data() + a0, h->xyz.data() + 3 * (size_t)a0,
                            charge ? charge[k] : 0.0, multiplicity ? multiplicity[k] : 1});
  }
  RunRequest req;
  req.solo_large = (flags & (PM3_BATCH_KEEP_BATCHED | PM3_BATCH_WARM_START)) == 0;
  req.want_density = (flags & PM3_BATCH_WANT_DENSITY) != 0;
  req.want_gradient = (flags & PM3_BATCH_WANT_GRADIENT) != 0;
  req.force_large = (flags & PM3_BATCH_FORCE_LARGE) != 0;
  req.warm_start = (flags & PM3_BATCH_WARM_START) != 0;
  h->want_density = req.want_density;
  std::string err;
  if (!batch_create(params->p, h->mols, eo, req, h->b, err)) {
    batch_destroy(h->b);
    delete h;
    return fail(PM3_DEVICE_ERROR, err);
  }
  *out = h;
  return PM3_OK;
}

pm3_status pm3_batch_solve(pm3_batch* h) {
  g_last_error.clear();
  if (!h) return fail(PM3_INVALID_INPUT, "null batch");
  std::string err;
  if (!batch_solve(h->b, err)) return fail(PM3_DEVICE_ERROR, err);
  h->fetched = false;
  return PM3_OK;
}

pm3_status pm3_batch_results(pm3_batch* h, pm3_status* status, pm3_result* results, double* charges, double* density) {
  g_last_error.clear();
  if (!h) return fail(PM3_INVALID_INPUT, "null batch");
  if (density && !h->want_density) return fail(PM3_INVALID_INPUT, "densities need PM3_BATCH_WANT_DENSITY at creation");
  std::string err;
  if (!h->fetched) {
    if (!batch_fetch(h->b, h->params->p, h->mols, h->results, err)) return fail(PM3_DEVICE_ERROR, err);
    h->fetched = true;
  }
  size_t aoff = 0, moff = 0;
  for (size_t k = 0; k < h->mols.size(); ++k) {
    const MolResult& r = h->results[k];
    if (status) status[k] = (pm3_status)r.status;
    if (results) fill_result(r, &results[k]);
    if (charges && (int)r.charges.size() == h->mols[k].n_atoms)
      std::memcpy(charges + aoff, r.charges.data(), r.charges.size() * sizeof(double));
    const size_t n2 = (size_t)r.nao * r.nao;
    if (density && r.density.size() == n2 && n2 > 0) std::memcpy(density + moff, r.density.data(), n2 * sizeof(double));
    aoff += h->mols[k].n_atoms;
    moff += n2;
  }
  return PM3_OK;
}

pm3_status pm3_batch_spin_densities(pm3_batch* h, double* spin_density) {
  g_last_error.clear();
  if (!h || !spin_density) return fail(PM3_INVALID_INPUT, "null argument");
  if (!h->want_density) return fail(PM3_INVALID_INPUT, "spin densities need PM3_BATCH_WANT_DENSITY at creation");
  pm3_status s = pm3_batch_results(h, nullptr, nullptr, nullptr, nullptr);
  if (s != PM3_OK) return s;
  size_t moff = 0;
  for (size_t k = 0; k < h->mols.size(); ++k) {
    const MolResult& r = h->results[k];
    const size_t n2 = (size_t)r.nao * r.nao;
    if (r.spin_density.size() == n2 && n2 > 0) std::memcpy(spin_density + moff, r.spin_density.data(), n2 * sizeof(double));
    else if (n2 > 0) std::memset(spin_density + moff, 0, n2 * sizeof(double));
    moff += n2;
  }
  return PM3_OK;
}

double pm3_batch_last_kernel_ms(const pm3_batch* h) { return h ? batch_kernel_ms(h->b) : 0.0; }

pm3_status pm3_batch_large_timing(const pm3_batch* h, double* product_ms, int64_t* products, int64_t* product_flop,
                                  double* fock_ms, int64_t* fock_builds, int64_t* lanczos_steps, int64_t* lanczos_runs) {
  g_last_error.clear();
  if (!h) return fail(PM3_INVALID_INPUT, "null batch");
  double pms = 0.0, fms = 0.0;
  long long np = 0, nf = 0, fl = 0, ls = 0, lr = 0;
  batch_large_timing(h->b, pms, np, fl, fms, nf, ls, lr);
  if (lanczos_steps) *lanczos_steps = ls;
  if (lanczos_runs) *lanczos_runs = lr;
  if (product_ms) *product_ms = pms;
  if (products) *products = np;
  if (product_flop) *product_flop = fl;
  if (fock_ms) *fock_ms = fms;
  if (fock_builds) *fock_builds = nf;
  return PM3_OK;
}

double pm3_batch_last_pair_ms(const pm3_batch* h) { return h ? batch_pair_ms(h->b) : 0.0; }

double pm3_batch_last_guess_ms(const pm3_batch* h) { return h ? batch_guess_ms(h->b) : 0.0; }

pm3_status pm3_batch_guess_timing(const pm3_batch* h, double* features_ms, double* network_ms, int64_t* directed_pairs) {
  if (!h) return fail(PM3_INVALID_INPUT, "null batch");
  double f = 0.0, n = 0.0;
  long long p = 0;
  batch_guess_timing(h->b, f, n, p);
  if (features_ms) *features_ms = f;
  if (network_ms) *network_ms = n;
  if (directed_pairs) *directed_pairs = p;
  return PM3_OK;
}

pm3_status pm3_batch_phase_ticks(const pm3_batch* h, uint64_t* out) {
  if (!h || !out) return fail(PM3_INVALID_INPUT, "null argument");
  std::string err;
  unsigned long long t[16];
  if (!batch_phase_ticks(h->b, t, err)) return fail(PM3_DEVICE_ERROR, err);
  for (int k = 0; k < 16; ++k) out[k] = t[k];
  return PM3_OK;
}

pm3_status pm3_batch_gradient(pm3_batch* h, double* gradient) {
  g_last_error.clear();
  if (!h || !gradient) return fail(PM3_INVALID_INPUT, "null argument");
  pm3_status s = pm3_batch_results(h, nullptr, nullptr, nullptr, nullptr);
  if (s != PM3_OK) return s;
  size_t off = 0;
  for (size_t k = 0; k < h->mols.size(); ++k) {
    const MolResult& r = h->results[k];
    const size_t na = (size_t)h->mols[k].n_atoms;
    if (r.gradient.size() == 3 * na) std::memcpy(gradient + 3 * off, r.gradient.data(), 3 * na * sizeof(double));
    else std::memset(gradient + 3 * off, 0, 3 * na * sizeof(double));
    off += na;
  }
  return PM3_OK;
}

pm3_status pm3_batch_set_positions(pm3_batch* h, const double* xyz_bohr) {
  g_last_error.clear();
  if (!h || !h->b) return fail(PM3_INVALID_INPUT, "null batch");
  if (!xyz_bohr && !h->xyz.empty()) return fail(PM3_INVALID_INPUT, "null positions");
  // the host copy (read by pm3_batch_results) changes only once the device has the new values
  std::vector<double> keep(h->xyz);
  std::copy(xyz_bohr, xyz_bohr + h->xyz.size(), h->xyz.begin());
  std::string why;
  const int s = batch_set_positions(h->b, h->mols, why);
  if (s != 0) {
    std::copy(keep.begin(), keep.end(), h->xyz.begin());
    return fail((pm3_status)s, why);
  }
  h->fetched = false;
  return PM3_OK;
}

pm3_status pm3_batch_carry_densities(pm3_batch* dst, const pm3_batch* src, const int32_t* src_index) {
  g_last_error.clear();
  if (!dst || !dst->b || !src || !src->b || (!src_index && !dst->mols.empty())) return fail(PM3_INVALID_INPUT, "null argument");
  std::vector<int> idx(src_index, src_index + dst->mols.size());
  for (size_t k = 0; k < idx.size(); ++k) {
    const int j = idx[k];
    if (j < 0 || j >= (int)src->mols.size() || src->mols[j].n_atoms != dst->mols[k].n_atoms ||
        !std::equal(dst->mols[k].z, dst->mols[k].z + dst->mols[k].n_atoms, src->mols[j].z))
      return fail(PM3_INVALID_INPUT, "molecule " + std::to_string(k) + " is not its source's molecule");
  }
  std::string why;
  const int s = batch_carry_densities(dst->b, src->b, idx, why);
  return s == 0 ? PM3_OK : fail((pm3_status)s, why);
}

void pm3_batch_destroy(pm3_batch* h) {
  if (!h) return;
  batch_destroy(h->b);
  delete h;
}

pm3_status pm3_run_batch(int32_t n_mol, const int32_t* atom_ptr, const int32_t* z, const double* xyz_bohr,
                         const double* charge, const int32_t* multiplicity, const pm3_params* params,
                         const pm3_options* opts, int32_t want_gradient, pm3_status* status, pm3_result* results,
                         double* gradient) {
  if (want_gradient && !gradient) return fail(PM3_INVALID_INPUT, "gradient buffer required");
  pm3_batch* h = nullptr;
  pm3_status s = pm3_batch_create(n_mol, atom_ptr, z, xyz_bohr, charge, multiplicity, params, opts,
                                  want_gradient ? PM3_BATCH_WANT_GRADIENT : 0, &h);
  if (s != PM3_OK) return s;
  s = pm3_batch_solve(h);
  if (s == PM3_OK) s = pm3_batch_results(h, status, results, nullptr, nullptr);
  if (s == PM3_OK && want_gradient) s = pm3_batch_gradient(h, gradient);
  pm3_batch_destroy(h);
  return s;
}

pm3_status pm3_gradient(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                        pm3_result* result, double* gradient, double* density, double* charges) {
  g_last_error.clear();
  if (!params || !gradient) return fail(PM3_INVALID_INPUT, "null argument");
  pm3_status s = check_molecule(mol);
  if (s != PM3_OK) return s;
  EngineOptions eo;
  std::string why;
  if (!engine_options(opts, eo, why)) return fail(PM3_INVALID_INPUT, why);
  if ((s = reference_check(mol, params, eo)) != PM3_OK) return s;
  std::vector<MolIn> mols{MolIn{mol->n_atoms, mol->z, mol->xyz_bohr, mol->charge, mol->multiplicity}};
  RunRequest req;
  req.solo_large = true;
  req.want_density = true;
  req.want_gradient = true;
  std::vector<MolResult> out;
  std::string err;
  if (!engine_run(params->p, mols, eo, req, out, err)) return fail(PM3_DEVICE_ERROR, err);
  const MolResult& r = out[0];
  if (result) fill_result(r, result);
  if (r.status == PM3_OK) {
    std::memcpy(gradient, r.gradient.data(), 3 * (size_t)mol->n_atoms * sizeof(double));
    if (density) std::memcpy(density, r.density.data(), (size_t)r.nao * r.nao * sizeof(double));
    if (charges) std::memcpy(charges, r.charges.data(), (size_t)mol->n_atoms * sizeof(double));
  }
  return molecule_status(r.status, r);
}

pm3_status pm3_response_fock(const pm3_molecule* mol, const pm3_params* params, int32_t n_rhs, const double* dp,
                             const double* pair_sw, double* out) {
  g_last_error.clear();
  if (!params || n_rhs < 0 || (n_rhs > 0 && (!dp || !out))) return fail(PM3_INVALID_INPUT, "null argument");
  pm3_status s = check_molecule(mol);
  if (s != PM3_OK) return s;
  if (n_rhs == 0) return PM3_OK;
  if (pair_sw) {
    const long long np = (long long)mol->n_atoms * (mol->n_atoms - 1) / 2;
    for (long long p = 0; p < np; ++p)
      if (!std::isfinite(pair_sw[p])) return fail(PM3_INVALID_INPUT, "pair_sw must be finite");
  }
  std::string err;
  const int st = response_fock_run(params->p, MolIn{mol->n_atoms, mol->z, mol->xyz_bohr, mol->charge, mol->multiplicity},
                                   n_rhs, dp, pair_sw, out, err);
  return st == 0 ? PM3_OK : fail((pm3_status)st, err);
}

void pm3_dc_options_default(pm3_dc_options* out) {
  const DcOptionsHost d;
  out->core_radius = d.core_radius;
  out->buffer_radius = d.buffer_radius;
  out->min_core_atoms = d.min_core_atoms;
  out->electron_tol = d.electron_tol;
  out->smearing_ev = d.smearing_ev;
}

static pm3_status run_dc_impl(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                              const pm3_dc_options* dc, pm3_dc_result* result, double* gradient, double* density,
                              double* spin_density, double* charges);

pm3_status pm3_run_dc(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                      const pm3_dc_options* dc, pm3_dc_result* result, double* density, double* spin_density,
                      double* charges) {
  return run_dc_impl(mol, params, opts, dc, result, nullptr, density, spin_density, charges);
}

pm3_status pm3_dc_gradient(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                           const pm3_dc_options* dc, pm3_dc_result* result, double* gradient, double* density,
                           double* spin_density, double* charges) {
  if (!gradient) {
    g_last_error = "null gradient array";
    return PM3_INVALID_INPUT;
  }
  return run_dc_impl(mol, params, opts, dc, result, gradient, density, spin_density, charges);
}

static pm3_status run_dc_impl(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                              const pm3_dc_options* dc, pm3_dc_result* result, double* gradient, double* density,
                              double* spin_density, double* charges) {
  g_last_error.clear();
  if (!params) return fail(PM3_INVALID_INPUT, "null parameter set");
  pm3_status s = check_molecule(mol);
  if (s != PM3_OK) return s;
  EngineOptions eo;
  std::string why;
  if (!engine_options(opts, eo, why)) return fail(PM3_INVALID_INPUT, why);
  DcOptionsHost d;
  if (dc) {
    d.core_radius = dc->core_radius;
    d.buffer_radius = dc->buffer_radius;
    d.min_core_atoms = dc->min_core_atoms;
    d.electron_tol = dc->electron_tol;
    d.smearing_ev = dc->smearing_ev;
  }
  if (!std::isfinite(d.core_radius) || !std::isfinite(d.buffer_radius) || !std::isfinite(d.electron_tol) ||
      !std::isfinite(d.smearing_ev) || d.electron_tol <= 0.0 || d.smearing_ev < 0.0)
    return fail(PM3_INVALID_INPUT, "divide-and-conquer options must be finite, electron_tol > 0, smearing_ev >= 0");
  DcResultHost r;
  std::string err;
  const int st = dc_run(params->p, MolIn{mol->n_atoms, mol->z, mol->xyz_bohr, mol->charge, mol->multiplicity}, eo, d,
                        r, err, gradient != nullptr);
  if (result && (st == 0 || st == PM3_SCF_NOT_CONVERGED)) {
    result->total_ev = r.total_ev;
    result->electronic_ev = r.electronic_ev;
    result->core_ev = r.core_ev;
    result->heat_of_formation_kcal = r.hof_kcal;
    result->fermi_ev = r.fermi_ev;
    result->fermi_beta_ev = r.fermi_beta_ev;
    result->iterations = r.iterations;
    result->converged = r.converged;
    result->unrestricted = r.unrestricted;
    result->dropped_pairs = r.dropped_pairs;
    result->largest_subsystem = r.largest_subsystem;
    result->n_subsystems = r.n_subsystems;
    result->nao = r.nao;
  }
  if (st == PM3_SCF_NOT_CONVERGED)
    return fail(PM3_SCF_NOT_CONVERGED, "divide-and-conquer SCF did not converge in " + std::to_string(r.iterations) +
                                           " iterations");
  if (st == PM3_MISSING_ELEMENT)
    return fail(PM3_MISSING_ELEMENT, "missing PM3 parameter block for Z=" + std::to_string(r.missing_z));
  if (st == PM3_UNSUPPORTED)
    return fail(PM3_UNSUPPORTED, "this engine covers the s/p elements H, Li-F, Na-Cl; Z=" + std::to_string(r.missing_z) +
                                     " is not among them");
  if (st == PM3_RESOURCE_LIMIT)
    return fail(PM3_RESOURCE_LIMIT, "PM3 two-electron pair cache needs " + std::to_string((long long)r.required_mb) +
                                        " MiB, above the limit of " + std::to_string((long long)r.limit_mb) + " MiB");
  if (st != 0) return fail((pm3_status)st, err.empty() ? std::string("divide-and-conquer failed") : err);
  const size_t n2 = (size_t)r.nao * r.nao;
  if (density) std::memcpy(density, r.density.data(), n2 * sizeof(double));
  if (spin_density) std::memcpy(spin_density, r.spin_density.data(), n2 * sizeof(double));
  if (charges) std::memcpy(charges, r.charges.data(), (size_t)mol->n_atoms * sizeof(double));
  if (gradient) std::memcpy(gradient, r.gradient.data(), 3 * (size_t)mol->n_atoms * sizeof(double));
  return PM3_OK;
}

// ------------------------------------------------------------------------------------------
// Tuning table and autotuner (no reference counterpart: the reference has no device kernels).

static void to_c(const Tuning& t, pm3_tuning* c) {
  c->pair_block = t.pair_block;
  c->pair_blocks_per_cu = t.pair_blocks_per_cu;
  c->sp2_tile16_n = t.sp2_tile16_n;
  c->sp2_lds = t.sp2_lds;
  c->sp2_tight_n = t.sp2_tight_n;
  c->large_syrk = t.large_syrk;
  c->large_split = t.large_split;
  c->wg128_nao = t.wg128_nao;
  c->wg256_nao = t.wg256_nao;
  c->solo_large_nao = t.solo_large_nao;
}

static Tuning from_c(const pm3_tuning* c) {
  Tuning t;
  t.pair_block = c->pair_block;
  t.pair_blocks_per_cu = c->pair_blocks_per_cu;
  t.sp2_tile16_n = c->sp2_tile16_n;
  t.sp2_lds = c->sp2_lds;
  t.sp2_tight_n = c->sp2_tight_n;
  t.large_syrk = c->large_syrk;
  t.large_split = c->large_split;
  t.wg128_nao = c->wg128_nao;
  t.wg256_nao = c->wg256_nao;
  t.solo_large_nao = c->solo_large_nao;
  return t;
}

void pm3_tuning_default(pm3_tuning* t) {
  if (t) to_c(Tuning(), t);
}

pm3_status pm3_tuning_get(pm3_tuning* t) {
  g_last_error.clear();
  if (!t) return fail(PM3_INVALID_INPUT, "null tuning");
  to_c(tuning_table(), t);
  return PM3_OK;
}

pm3_status pm3_tuning_set(const pm3_tuning* t) {
  g_last_error.clear();
  if (!t) return fail(PM3_INVALID_INPUT, "null tuning");
  const Tuning v = from_c(t);
  std::string why;
  if (!tuning_valid(v, why)) return fail(PM3_INVALID_INPUT, why);
  tuning_table() = v;
  return PM3_OK;
}

pm3_status pm3_tuning_load(const char* path) {
  g_last_error.clear();
  if (!path) return fail(PM3_INVALID_INPUT, "null path");
  Tuning t;
  std::string err;
  if (!tuning_load(path, t, err)) return fail(PM3_INVALID_INPUT, err);
  tuning_table() = t;
  return PM3_OK;
}

pm3_status pm3_tuning_save(const char* path) {
  g_last_error.clear();
  if (!path) return fail(PM3_INVALID_INPUT, "null path");
  std::string err;
  if (!tuning_save(path, tuning_table(), err)) return fail(PM3_INVALID_INPUT, err);
  return PM3_OK;
}

// Coordinate descent over the table's entries on the caller's workload: each candidate table is
// used for a fresh batch (the table is read at creation), solved once untimed and `reps` times
// timed (device time of pair integrals + learned guess + SCF/gradient, HIP events; the median
// counts).  A candidate is kept when it is faster by more than max(0.2 %, 0.02 ms) and every
// molecule's status, and heat of formation to 1e-5 kcal/mol, equal those of the starting table.
// PM3_TUNE_VERBOSE=1 prints each candidate's time to stderr.
pm3_status pm3_autotune(int32_t n_mol, const int32_t* atom_ptr, const int32_t* z, const double* xyz_bohr,
                        const double* charge, const pm3_params* params, const pm3_options* opts, int32_t flags,
                        int32_t reps, pm3_tuning* best_out, pm3_tune_report* report) {
  g_last_error.clear();
  if (!params || n_mol <= 0 || !atom_ptr || !z || !xyz_bohr) return fail(PM3_INVALID_INPUT, "bad autotune arguments");
  if (atom_ptr[0] != 0) return fail(PM3_INVALID_INPUT, "atom_ptr[0] must be 0");
  for (int k = 0; k < n_mol; ++k)
    if (atom_ptr[k + 1] < atom_ptr[k]) return fail(PM3_INVALID_INPUT, "atom_ptr must be non-decreasing");
  EngineOptions eo;
  std::string why;
  if (!engine_options(opts, eo, why)) return fail(PM3_INVALID_INPUT, why);
  std::vector<MolIn> mols;
  for (int k = 0; k < n_mol; ++k)
    mols.push_back(MolIn{atom_ptr[k + 1] - atom_ptr[k], z + atom_ptr[k], xyz_bohr + 3 * (size_t)atom_ptr[k],
                         charge ? charge[k] : 0.0});
  RunRequest req;
  req.solo_large = (flags & (PM3_BATCH_KEEP_BATCHED | PM3_BATCH_WARM_START)) == 0;
  req.want_density = (flags & PM3_BATCH_WANT_DENSITY) != 0;
  req.want_gradient = (flags & PM3_BATCH_WANT_GRADIENT) != 0;
  req.force_large = (flags & PM3_BATCH_FORCE_LARGE) != 0;
  reps = reps < 1 ? 1 : reps;
  const Tuning start = tuning_table();
  int nlarge = 0, nsmall = 0;
  auto measure = [&](const Tuning& t, double& ms, std::vector<MolResult>& res, std::string& err) -> bool {
    tuning_table() = t;
    Batch* b = nullptr;
    bool ok = batch_create(params->p, mols, eo, req, b, err) && batch_solve(b, err);
    std::vector<double> ts;
    for (int r = 0; ok && r < reps; ++r) {
      ok = batch_solve(b, err);
      ts.push_back(batch_pair_ms(b) + batch_guess_ms(b) + batch_kernel_ms(b));
    }
    if (ok) ok = batch_fetch(b, params->p, mols, res, err);
    if (ok) {
      nlarge = batch_large_count(b);
      nsmall = batch_small_count(b);
    }
    batch_destroy(b);
    if (!ok) return false;
    std::sort(ts.begin(), ts.end());
    ms = ts[ts.size() / 2];
    return true;
  };
  std::string err;
  std::vector<MolResult> ref;
  double t0 = 0.0;
  if (!measure(start, t0, ref, err)) {
    tuning_table() = start;
    return fail(PM3_DEVICE_ERROR, err);
  }
  // entries of the persistent kernel (small) and of the host-driven path (large) are tried only
  // when the batch has molecules there
  struct Axis {
    int Tuning::*m;
    std::vector<int> values;
    int needs;  // 0 any batch, 1 molecules on the persistent kernel, 2 on the host-driven path
  };
  const Axis axes[] = {
      {&Tuning::pair_block, {64, 128, 256}, 0},
      {&Tuning::pair_blocks_per_cu, {2, 4, 8, 16, 32}, 0},
      {&Tuning::sp2_tile16_n, {0, 32, 48, 64, 96, 128, 160, 200}, 1},
      {&Tuning::sp2_lds, {0, 1, 2, 3, 4, 5, 6}, 1},
      {&Tuning::sp2_tight_n, {24, 32, 48, 64, 96, 1 << 20}, 1},
      {&Tuning::large_syrk, {0, 1}, 2},
      {&Tuning::large_split, {0, 1, 2, 3, 4}, 2},
      {&Tuning::wg128_nao, {0, 32, 48, 64, 80, 96}, 1},
      {&Tuning::wg256_nao, {0, 64, 80, 96, 112, 128}, 1},
      {&Tuning::solo_large_nao, {96, 128, 160, 192, 1 << 20}, 0},
  };
  Tuning best = start;
  double tbest = t0;
  int candidates = 1, rejected = 0;
  const char* vb = std::getenv("PM3_TUNE_VERBOSE");
  const bool verbose = vb && vb[0] == '1';
  if (verbose) std::fprintf(stderr, "pm3_autotune: starting table %.3f ms\n", t0);
  for (const Axis& ax : axes) {
    if ((ax.needs == 1 && nsmall == 0) || (ax.needs == 2 && nlarge == 0)) continue;
    const int base = best.*(ax.m);
    for (int v : ax.values) {
      if (v == base) continue;
      Tuning c = best;
      c.*(ax.m) = v;
      double tc = 0.0;
      std::vector<MolResult> res;
      if (!measure(c, tc, res, err)) {
        tuning_table() = start;
        return fail(PM3_DEVICE_ERROR, err);
      }
      ++candidates;
      bool same = res.size() == ref.size();
      for (size_t k = 0; same && k < res.size(); ++k) {
        if (res[k].status != ref[k].status) same = false;
        else if (res[k].status == 0 && std::fabs(res[k].hof_kcal - ref[k].hof_kcal) > 1.0e-5) same = false;
      }
      if (verbose) {
        std::string name = "?";
        if (ax.m == &Tuning::pair_block) name = "pair_block";
        else if (ax.m == &Tuning::pair_blocks_per_cu) name = "pair_blocks_per_cu";
        else if (ax.m == &Tuning::sp2_tile16_n) name = "sp2_tile16_n";
        else if (ax.m == &Tuning::sp2_lds) name = "sp2_lds";
        else if (ax.m == &Tuning::sp2_tight_n) name = "sp2_tight_n";
        else if (ax.m == &Tuning::large_syrk) name = "large_syrk";
        else if (ax.m == &Tuning::large_split) name = "large_split";
        else if (ax.m == &Tuning::wg128_nao) name = "wg128_nao";
        else if (ax.m == &Tuning::wg256_nao) name = "wg256_nao";
        else if (ax.m == &Tuning::solo_large_nao) name = "solo_large_nao";
        std::fprintf(stderr, "pm3_autotune: %s %d: %.3f ms%s\n", name.c_str(), v, tc, same ? "" : " (answers differ)");
      }
      if (!same) {
        ++rejected;
        continue;
      }
      if (tc < tbest - std::max(0.002 * tbest, 0.02)) {
        best = c;
        tbest = tc;
      }
    }
  }
  tuning_table() = best;
  if (best_out) to_c(best, best_out);
  if (report) {
    report->default_ms = t0;
    report->best_ms = tbest;
    report->candidates = candidates;
    report->rejected = rejected;
  }
  return PM3_OK;
}

const char* pm3_last_error_message(void) { return g_last_error.c_str(); }

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Kernel-level stages (include/pm3_gpu_stages.h).
#include "../../include/pm3_gpu_stages.h"

namespace {

pm3_status run_stage(const pm3_molecule* mol, int mode, const std::vector<double>& in, std::vector<double>& out) {
  const pm3_params* params = pm3_params_standard();
  if (!params) return fail(PM3_DEVICE_ERROR, "parameter table unavailable");
  MolIn m{mol->n_atoms, mol->z, mol->xyz_bohr, mol->charge};
  int st = 0;
  std::string err;
  if (!stage_run(params->p, m, mode, in, out, st, err)) return fail(PM3_DEVICE_ERROR, err);
  if (st != 0) return fail((pm3_status)st, status_text(st));
  return PM3_OK;
}

// A chain of n hydrogens 3 Bohr apart: a carrier whose AO count is n for the generic stages.
pm3_status generic_stage(int n, int mode, const std::vector<double>& in, std::vector<double>& out) {
  if (n < 1 || n > 190) return fail(PM3_INVALID_INPUT, "n must be in [1, 190]");
  std::vector<int32_t> z(n, 1);
  std::vector<double> xyz(3 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i) xyz[3 * i] = 3.0 * i;
  pm3_molecule mol{n, z.data(), xyz.data(), (double)(n % 2), 1};
  return run_stage(&mol, mode, in, out);
}

}  // namespace

extern "C" {

pm3_status pm3_stage_hcore(const pm3_molecule* mol, const pm3_params* params, double* h_core, double* core_ev) {
  g_last_error.clear();
  if (!params) return fail(PM3_INVALID_INPUT, "null parameter set");
  pm3_status s = check_molecule(mol);
  if (s != PM3_OK) return s;
  std::vector<double> out;
  s = run_stage(mol, 0, {}, out);
  if (s != PM3_OK) return s;
  const size_t n2 = out.size() - 1;
  if (h_core) std::memcpy(h_core, out.data(), n2 * sizeof(double));
  if (core_ev) *core_ev = out[n2];
  return PM3_OK;
}

pm3_status pm3_stage_fock(const pm3_molecule* mol, const pm3_params* params, const double* density, double* fock) {
  g_last_error.clear();
  if (!params || !density || !fock) return fail(PM3_INVALID_INPUT, "null argument");
  pm3_status s = check_molecule(mol);
  if (s != PM3_OK) return s;
  const int n = pm3_query_nao(mol, params);
  if (n < 0) return fail(PM3_MISSING_ELEMENT, "missing element");
  std::vector<double> in(density, density + (size_t)n * n), out;
  s = run_stage(mol, 1, in, out);
  if (s != PM3_OK) return s;
  std::memcpy(fock, out.data(), (size_t)n * n * sizeof(double));
  return PM3_OK;
}

pm3_status pm3_stage_eigen(int32_t n, const double* a, double* w, double* v) {
  g_last_error.clear();
  if (!a || !w || !v) return fail(PM3_INVALID_INPUT, "null argument");
  std::vector<double> in(a, a + (size_t)n * n), out;
  pm3_status s = generic_stage(n, 2, in, out);
  if (s != PM3_OK) return s;
  std::memcpy(w, out.data(), n * sizeof(double));
  std::memcpy(v, out.data() + n, (size_t)n * n * sizeof(double));
  return PM3_OK;
}

pm3_status pm3_stage_eigen_warm(int32_t n, const double* a, const double* c, double* w, double* v) {
  g_last_error.clear();
  if (!a || !c || !w || !v) return fail(PM3_INVALID_INPUT, "null argument");
  std::vector<double> in(a, a + (size_t)n * n), out;
  in.insert(in.end(), c, c + (size_t)n * n);
  pm3_status s = generic_stage(n, 4, in, out);
  if (s != PM3_OK) return s;
  std::memcpy(w, out.data(), n * sizeof(double));
  std::memcpy(v, out.data() + n, (size_t)n * n * sizeof(double));
  return PM3_OK;
}

pm3_status pm3_stage_purify(const pm3_molecule* mol, const pm3_params* params, const double* fock, double* density,
                            int32_t* products, double* homo, double* lumo) {
  g_last_error.clear();
  if (!params || !fock || !density) return fail(PM3_INVALID_INPUT, "null argument");
  pm3_status s = check_molecule(mol);
  if (s != PM3_OK) return s;
  const int n = pm3_query_nao(mol, params);
  if (n < 0) return fail(PM3_MISSING_ELEMENT, "missing element");
  std::vector<double> in(fock, fock + (size_t)n * n), out;
  s = run_stage(mol, 6, in, out);
  if (s != PM3_OK) return s;
  std::memcpy(density, out.data(), (size_t)n * n * sizeof(double));
  if (products) *products = (int32_t)out[(size_t)n * n];
  if (homo) *homo = out[(size_t)n * n + 1];
  if (lumo) *lumo = out[(size_t)n * n + 2];
  return PM3_OK;
}

pm3_status pm3_stage_gemm(int32_t n, const double* a, const double* b, double* c) {
  g_last_error.clear();
  if (!a || !b || !c) return fail(PM3_INVALID_INPUT, "null argument");
  std::vector<double> in(a, a + (size_t)n * n), out;
  in.insert(in.end(), b, b + (size_t)n * n);
  pm3_status s = generic_stage(n, 3, in, out);
  if (s != PM3_OK) return s;
  std::memcpy(c, out.data(), (size_t)n * n * sizeof(double));
  return PM3_OK;
}

pm3_status pm3_stage_sym_step(int32_t n, int32_t square, int32_t slab, const double* x, double* y, double* trace,
                              double* norm2, int32_t reps, int32_t workgroups, double* ms_per_launch) {
  g_last_error.clear();
  if (!x || !y) return fail(PM3_INVALID_INPUT, "null argument");
  std::string err;
  const int r = sym_step_stage(n, square, slab, x, y, trace, norm2, reps, workgroups, ms_per_launch, err);
  if (r == 2) return fail(PM3_INVALID_INPUT, err);
  if (r != 0) return fail(PM3_DEVICE_ERROR, err);
  return PM3_OK;
}

}  // extern "C"

// ------------------------------------------------------------------------------------------
// Learned initial density (include/pm3_guess.h).
static_assert(PM3_GUESS_NPARAM == pm3guess::NPARAM && PM3_GUESS_NF == pm3guess::NF && PM3_GUESS_NH == pm3guess::NH &&
                  PM3_GUESS_NO == pm3guess::NO,
              "pm3_guess.h sizes must match the network");

extern "C" {

pm3_status pm3_guess_model_create(uint64_t seed, pm3_guess_model** out) {
  g_last_error.clear();
  if (!out) return fail(PM3_INVALID_INPUT, "null handle");
  *out = nullptr;
  std::string err;
  GuessModel* m = nullptr;
  if (!guess_model_create(seed, m, err)) return fail(PM3_DEVICE_ERROR, err);
  *out = new pm3_guess_model{m};
  return PM3_OK;
}

pm3_status pm3_guess_model_load(const char* path, pm3_guess_model** out) {
  g_last_error.clear();
  if (!out || !path) return fail(PM3_INVALID_INPUT, "null argument");
  *out = nullptr;
  std::string err;
  GuessModel* m = nullptr;
  if (!guess_model_load(path, m, err)) return fail(m ? PM3_DEVICE_ERROR : PM3_INVALID_INPUT, err);
  *out = new pm3_guess_model{m};
  return PM3_OK;
}

pm3_status pm3_guess_model_save(const pm3_guess_model* model, const char* path) {
  g_last_error.clear();
  if (!model || !model->m || !path) return fail(PM3_INVALID_INPUT, "null argument");
  std::string err;
  if (!guess_model_save(model->m, path, err)) return fail(PM3_INVALID_INPUT, err);
  return PM3_OK;
}

pm3_status pm3_guess_model_get_params(const pm3_guess_model* model, float* out) {
  g_last_error.clear();
  if (!model || !model->m || !out) return fail(PM3_INVALID_INPUT, "null argument");
  std::string err;
  if (!guess_model_params(model->m, out, err)) return fail(PM3_DEVICE_ERROR, err);
  return PM3_OK;
}

pm3_status pm3_guess_model_set_params(pm3_guess_model* model, const float* params) {
  g_last_error.clear();
  if (!model || !model->m || !params) return fail(PM3_INVALID_INPUT, "null argument");
  for (int i = 0; i < PM3_GUESS_NPARAM; ++i)
    if (!std::isfinite(params[i])) return fail(PM3_INVALID_INPUT, "non-finite parameter");
  std::string err;
  if (!guess_model_set_params(model->m, params, err)) return fail(PM3_DEVICE_ERROR, err);
  return PM3_OK;
}

int32_t pm3_guess_model_steps(const pm3_guess_model* model) { return model ? guess_model_steps(model->m) : 0; }

void pm3_guess_model_destroy(pm3_guess_model* model) {
  if (!model) return;
  guess_model_destroy(model->m);
  delete model;
}

pm3_status pm3_guess_data_create(pm3_guess_data** out) {
  g_last_error.clear();
  if (!out) return fail(PM3_INVALID_INPUT, "null handle");
  std::string err;
  GuessData* d = nullptr;
  if (!guess_data_create(d, err)) return fail(PM3_DEVICE_ERROR, err);
  *out = new pm3_guess_data{d};
  return PM3_OK;
}

pm3_status pm3_guess_data_add(pm3_guess_data* data, pm3_batch* solved, int64_t* n_added) {
  g_last_error.clear();
  if (!data || !solved) return fail(PM3_INVALID_INPUT, "null argument");
  long long added = 0;
  std::string err;
  if (!guess_data_add(data->d, solved->b, added, err)) return fail(PM3_INVALID_INPUT, err);
  if (n_added) *n_added = added;
  return PM3_OK;
}

pm3_status pm3_guess_data_stats(const pm3_guess_data* data, int64_t* n_mol, int64_t* n_atoms, int64_t* n_pairs) {
  if (!data) return fail(PM3_INVALID_INPUT, "null argument");
  long long a = 0, b = 0, c = 0;
  guess_data_stats(data->d, a, b, c);
  if (n_mol) *n_mol = a;
  if (n_atoms) *n_atoms = b;
  if (n_pairs) *n_pairs = c;
  return PM3_OK;
}

void pm3_guess_data_destroy(pm3_guess_data* data) {
  if (!data) return;
  guess_data_destroy(data->d);
  delete data;
}

void pm3_guess_train_options_default(pm3_guess_train_options* out) {
  out->steps = 1000;
  out->batch_molecules = 1024;
  out->lr = 1.0e-3f;
  out->beta1 = 0.9f;
  out->beta2 = 0.999f;
  out->eps = 1.0e-8f;
  out->seed = 1;
  out->lr_final = 1.0e-5f;
  out->decay_steps = 0;
}

pm3_status pm3_guess_train(pm3_guess_model* model, pm3_guess_data* data, const pm3_guess_train_options* opts,
                           float* losses, double* seconds) {
  g_last_error.clear();
  if (!model || !model->m || !data || !opts) return fail(PM3_INVALID_INPUT, "null argument");
  if (opts->steps < 1 || opts->batch_molecules < 1 || !(opts->lr > 0.0f)) return fail(PM3_INVALID_INPUT, "bad training options");
  GuessTrainOptions o;
  o.steps = opts->steps;
  o.batch_molecules = opts->batch_molecules;
  o.lr = opts->lr;
  o.beta1 = opts->beta1;
  o.beta2 = opts->beta2;
  o.eps = opts->eps;
  o.seed = opts->seed;
  o.lr_final = opts->lr_final;
  o.decay_steps = opts->decay_steps;
  if (o.decay_steps < 0 || !(o.lr_final >= 0.0f)) return fail(PM3_INVALID_INPUT, "bad learning-rate schedule");
  std::string err;
  if (!guess_train(model->m, data->d, o, losses, seconds, err)) return fail(PM3_DEVICE_ERROR, err);
  return PM3_OK;
}

pm3_status pm3_guess_eval(const pm3_guess_model* model, pm3_guess_data* data, double* loss) {
  g_last_error.clear();
  if (!model || !model->m || !data || !loss) return fail(PM3_INVALID_INPUT, "null argument");
  std::string err;
  if (!guess_eval(model->m, data->d, *loss, err)) return fail(PM3_DEVICE_ERROR, err);
  return PM3_OK;
}

pm3_status pm3_guess_grad(const pm3_guess_model* model, pm3_guess_data* data, float* grad, double* loss) {
  g_last_error.clear();
  if (!model || !model->m || !data || !grad || !loss) return fail(PM3_INVALID_INPUT, "null argument");
  std::string err;
  if (!guess_grad(model->m, data->d, grad, *loss, err)) return fail(PM3_DEVICE_ERROR, err);
  return PM3_OK;
}

pm3_status pm3_guess_mlp(const pm3_guess_model* model, int32_t n, const uint16_t* x, float* y, const float* dy,
                         float* grad) {
  g_last_error.clear();
  if (!model || !model->m || !x || !y || n < 1 || (dy && !grad)) return fail(PM3_INVALID_INPUT, "bad arguments");
  std::string err;
  if (!guess_mlp(model->m, n, x, y, dy, grad, err)) return fail(PM3_DEVICE_ERROR, err);
  return PM3_OK;
}

pm3_status pm3_guess_features(const pm3_molecule* mol, const pm3_params* params, int32_t* n_pairs, int32_t* src,
                              int32_t* dst, uint16_t* x) {
  g_last_error.clear();
  if (!params || !n_pairs) return fail(PM3_INVALID_INPUT, "null argument");
  pm3_status s = check_molecule(mol);
  if (s != PM3_OK) return s;
  for (int i = 0; i < mol->n_atoms; ++i)
    if (pm3guess::elem_index(mol->z[i]) < 0) return fail(PM3_UNSUPPORTED, "the learned guess covers H, C and O only");
  std::vector<MolIn> mols{MolIn{mol->n_atoms, mol->z, mol->xyz_bohr, mol->charge}};
  EngineOptions eo;
  std::string why, err;
  if (!engine_options(nullptr, eo, why)) return fail(PM3_INVALID_INPUT, why);
  Batch* b = nullptr;
  std::vector<int> gs, gd, cm;
  std::vector<uint16_t> gx;
  const bool ok = batch_create(params->p, mols, eo, RunRequest(), b, err) && guess_batch_features(b, gs, gd, gx, cm, err);
  batch_destroy(b);
  if (!ok) return fail(PM3_DEVICE_ERROR, err);
  const int32_t cap = *n_pairs;
  if (src || dst || x) {
    if (!src || !dst || !x || cap < (int32_t)gs.size()) {
      *n_pairs = (int32_t)gs.size();
      return fail(PM3_INVALID_INPUT, "pair buffers missing or too small");
    }
    for (size_t j = 0; j < gs.size(); ++j) {
      src[j] = cm[gs[j]];
      dst[j] = cm[gd[j]];
    }
    if (!gx.empty()) std::memcpy(x, gx.data(), gx.size() * sizeof(uint16_t));
  }
  *n_pairs = (int32_t)gs.size();
  return PM3_OK;
}

pm3_status pm3_guess_density(const pm3_guess_model* model, const pm3_molecule* mol, const pm3_params* params,
                             double* density, double* c2l, double* c2s) {
  g_last_error.clear();
  if (!model || !model->m || !params || !density) return fail(PM3_INVALID_INPUT, "null argument");
  pm3_status s = check_molecule(mol);
  if (s != PM3_OK) return s;
  MolIn m{mol->n_atoms, mol->z, mol->xyz_bohr, mol->charge};
  std::vector<double> out;
  int st = 0;
  std::string err;
  if (!stage_run(params->p, m, 5, {}, out, st, err, model->m)) return fail(PM3_DEVICE_ERROR, err);
  if (st != 0) return fail((pm3_status)st, st == 4 ? "the learned guess covers H, C and O only" : status_text(st));
  const size_t n2 = out.size() - 2;
  std::memcpy(density, out.data(), n2 * sizeof(double));
  if (c2l) *c2l = out[n2];
  if (c2s) *c2s = out[n2 + 1];
  return PM3_OK;
}

}  // extern "C"
