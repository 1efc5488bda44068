// Host-side declarations shared by the C ABI (pm3_api.cpp) and the engine (pm3_engine.hip).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "nddo_device.hpp"

namespace pm3host {

constexpr double EV_TO_KCAL = 23.06054783061903;     // src/constants.rs:27
constexpr double AU_DIPOLE_TO_DEBYE = 2.541746473;   // src/constants.rs:56

struct ElemHost {
  double e_isol = 0.0, eheat_ev = 0.0, mass = 0.0;
};

struct Params {
  pm3dev::TableConst table;
  ElemHost host[pm3dev::MAX_SLOTS];
  int slot_of_z[128];
  int nslots = 0;
};

bool build_params(Params& P);

// Options as the engine consumes them (mirrors pm3_options).
struct EngineOptions {
  int max_scf;
  double e_tol, p_tol;
  int accel;        // 0 none, 1 CDIIS, 2 A-DIIS -> CDIIS (already combined with use_diis)
  double adiis_switch;
  int guess;        // 1 SAD, 2 core, 3 symmetry-broken, 4 learned (superposition fallback)
  int stability;    // 0 off, 1 auto, 2 always
  int scf_memory_mb;
  const struct GuessModel* guess_model = nullptr;  // network of the learned guess (guess 4)
  int use_diis = 1;          // Pm3Options::use_diis (accel already folds it in; the CDIIS rung reads it)
  double level_shift = 0.0;  // Pm3Options::level_shift_ev (eV)
  double damping = 0.0;      // Pm3Options::damping
  int reference = 0;         // 0 auto, 1 RHF, 2 UHF (Pm3Options::reference)
  int has_field = 0;         // Pm3Options::field is Some
  double field[3] = {0.0, 0.0, 0.0};  // eV per Bohr per elementary charge
  int integral_memory_mb = 0;         // pair-cache ceiling (MiB); 0 the process default
};

// One molecule as handed to the engine, atoms in caller order.
struct MolIn {
  int n_atoms;
  const int32_t* z;
  const double* xyz;
  double charge;
  int multiplicity = 1;
};

// Per-molecule engine output.  Matrices and vectors are in the caller's AO / atom order.
struct MolResult {
  int status = 0;           // pm3_status
  double e_elec = 0, e_core = 0, hof_kcal = 0, last_change = 0, improvement = 0;
  double homo = 0, lumo = 0;
  double dipole[3] = {0, 0, 0};  // Debye
  int iterations = 0, converged = 0, paths = 0, nocc = 0, nao = 0;
  int guess_used = 0;  // 0 superposition, 1 learned, 2 learned rejected by the commutator test
  int unrestricted = 0, nbeta = 0;
  int missing_z = 0;                          // status 2 / 4: the atomic number refused
  double required_mb = 0.0, limit_mb = 0.0;   // status 5: pair-cache size and ceiling (MiB)
  std::vector<double> charges, density, mo_energies, mo_coeff, gradient;
  std::vector<double> spin_density, mo_energies_beta, mo_coeff_beta;  // unrestricted results only
};

struct RunRequest {
  bool want_density = false;
  bool want_orbitals = false;
  bool want_gradient = false;
  bool force_large = false;  // solve every molecule on the host-driven multi-kernel path
  bool warm_start = false;   // solves after the first start from each molecule's last density
  bool solo_large = false;   // a batch of one molecule above Tuning::solo_large_nao AOs takes the
                             // multi-kernel path (the caller-facing entry points set it)
  bool dc = false;           // the divide-and-conquer driver's batch: open shells on the multi-kernel
                             // path too (it builds spin Fock matrices there)
};

// DcOptions (src/dc/partition.rs:40-97) without long_range_cutoff.
struct DcOptionsHost {
  double core_radius = 6.0, buffer_radius = 9.0;  // Bohr
  int min_core_atoms = 1;
  double electron_tol = 1.0e-9, smearing_ev = 0.1;
};

// DcResult (src/dc/scf.rs:25-46); matrices in the caller's AO order.
struct DcResultHost {
  double total_ev = 0, electronic_ev = 0, core_ev = 0, hof_kcal = 0, fermi_ev = 0, fermi_beta_ev = 0;
  int iterations = 0, converged = 0, unrestricted = 0, dropped_pairs = 0, largest_subsystem = 0, n_subsystems = 0;
  int nao = 0, missing_z = 0;
  double required_mb = 0.0, limit_mb = 0.0;
  std::vector<double> density, spin_density, charges;
  std::vector<double> gradient;  // dc_gradient, eV/Bohr, n_atoms x 3 (when asked for)
};

// Launch shapes and tile choices.  They change speed only: results agree to rounding whatever
// the values (pm3_autotune checks that on the workload it tunes).  The process-wide table is read
// by batch_create; environment knobs (PM3_SP2_LDS, PM3_SP2_TIGHT_N) still override its entries.
struct Tuning {
  int pair_block = 256;         // threads per block of the pair and gradient-pair kernels: 64, 128 or 256
  int pair_blocks_per_cu = 8;   // grid of those grid-stride kernels, blocks per CU: 1 to 64
  int sp2_tile16_n = 128;       // purification products on 16 x 16 wave tiles up to this size (AOs)
  int sp2_lds = 6;              // purification operand placements: 0 HBM, 1 + both in LDS, 2 + X in LDS,
                                // 3 + X in LDS overwritten in place by each product,
                                // 4 + the leading rows of X in LDS where the whole of X does not fit,
                                // 5 + X's upper 16 x 16 blocks (tiled, swizzled) in LDS, updated in place,
                                // 6 + that tiled image split between LDS and HBM where it does not fit
  int sp2_tight_n = 48;         // purification bounds from Lanczos from this size (AOs) on
  int large_syrk = 0;           // host-driven path: purification products by dsyrk (1) or dgemm (0)
  int large_split = 1;          // host-driven path, matrices above 1024 AOs: 1 purification products by the
                                // lower-triangle 64 x 64 tile kernel (update fused into its store);
                                // q in 2..4 only the q x q block-lower part of each product by q batched
                                // rocBLAS dgemm rounds when q divides n; 0 one rocBLAS dgemm (at 3072 AOs:
                                // 0.59 ms per step against 0.70 + 0.04 for q = 2 and 0.87 + 0.04 for 0)
  int wg128_nao = 80;           // molecules up to this size (AOs) whose LDS needs fit 36 KiB run on
                                // 128-thread workgroups, four per CU
  int wg256_nao = 96;           // then those up to this size fitting 76 KiB on 256-thread workgroups, two
                                // per CU; the rest on 512-thread workgroups, one per CU
  int solo_large_nao = 128;     // a batch of a single restricted molecule above this size (AOs) runs on
                                // the multi-kernel path, which spreads each step over the whole GPU
};
// The table batch_create reads; on first use it is loaded from the file PM3_TUNING_FILE names,
// if set.  Set or loaded tables apply to batches created afterwards (not thread-safe against a
// concurrent batch_create).
Tuning& tuning_table();
bool tuning_valid(const Tuning& t, std::string& why);
bool tuning_load(const char* path, Tuning& t, std::string& err);
bool tuning_save(const char* path, const Tuning& t, std::string& err);

// A batch whose inputs and scratch live on the device; batch_solve is the launch alone.
struct Batch;
bool batch_create(const Params& P, const std::vector<MolIn>& mols, const EngineOptions& opt, const RunRequest& req,
                  Batch*& out, std::string& err);
bool batch_solve(Batch* b, std::string& err);
bool batch_fetch(Batch* b, const Params& P, const std::vector<MolIn>& mols, std::vector<MolResult>& out,
                 std::string& err);
void batch_destroy(Batch* b);
// New positions for the batch's molecules (the same atoms, in the caller's order, as mols at
// creation); the next solve's kernels read them.  Returns 0, or a pm3_status code with err set:
// 1 bad positions (not finite, coincident atoms, wrong counts), 4 a learned-guess batch, 6 a
// device failure.
int batch_set_positions(Batch* b, const std::vector<MolIn>& mols, std::string& err);
// Each molecule's last density and result record of src (molecule src_of[k] for dst's k, the
// same atoms) into dst, for dst's next warm-started solve.  Returns 0 or a pm3_status code.
int batch_carry_densities(Batch* dst, const Batch* src, const std::vector<int>& src_of, std::string& err);
double batch_kernel_ms(const Batch* b);  // device time of the last solve's SCF launch (HIP events)
double batch_pair_ms(const Batch* b);    // device time of the last solve's pair-integral launch
double batch_guess_ms(const Batch* b);   // device time of the last solve's learned-guess launches
int batch_large_count(const Batch* b);  // molecules of the batch on the host-driven multi-kernel path
int batch_small_count(const Batch* b);  // molecules of the batch on the persistent one-workgroup kernel
// Host-driven path, last solve: device time, count and FLOP of the purification products
// (HIP events around each rocBLAS call), device time and count of the Fock builds, and the
// frontier-level Lanczos steps and runs.
void batch_large_timing(const Batch* b, double& product_ms, long long& products, long long& product_flop,
                        double& fock_ms, long long& fock_builds, long long& lanczos_steps, long long& lanczos_runs);
// The same split into the feature kernel and the network kernel, and the batch's directed pairs.
void batch_guess_timing(const Batch* b, double& features_ms, double& network_ms, long long& directed_pairs);
// Per-phase wall-clock ticks (10 ns) summed over workgroups, when PM3_PHASE_TIMERS=1 at creation:
// 0 H_core finish, 1 Fock, 2 energy+DIIS, 3 warm-start products, 4 Jacobi, 5 sort+C V, 6 density,
// 7 other, 8 DIIS coefficient solve, 9 DIIS extrapolation; [10] frontier Lanczos steps, [11] frontier
// Lanczos runs, [13] Jacobi sweeps or purification products, [14] eigensolves or purifications,
// [15] Jacobi rounds or purification fallbacks (16 slots).
bool batch_phase_ticks(const Batch* b, unsigned long long* out, std::string& err);

// One step of the large path's purification tile kernel on a host n x n matrix: y = x x (square)
// or 2 x - x x, tr y and |y|_F^2; reps more launches timed (ms per launch); workgroups > 0 caps
// the grid.  Returns 0, 2 (bad argument) or 6 (device error), with err set.
int sym_step_stage(int n, int square, int slab, const double* x, double* y, double* trace, double* norm2, int reps,
                   int workgroups, double* ms, std::string& err);

// One stage of the pipeline on a single molecule (kernel-level parity hooks); see stage_kernel.
bool stage_run(const Params& P, const MolIn& mol, int mode, const std::vector<double>& in, std::vector<double>& out,
               int& status, std::string& err, const struct GuessModel* model = nullptr);

// response_fock_rhf (src/fock.rs:444-485) of one molecule for nrhs densities dp[nrhs][n][n]
// (caller's AO order): out[r] = J(dp_r) - K(dp_r / 2), no H_core, two-centre terms weighted by
// pair_sw (one weight per atom pair in the reference's order, u < v row by row; null: all 1).
// Returns 0, or a pm3_status code with err set.
int response_fock_run(const Params& P, const MolIn& mol, int nrhs, const double* dp, const double* pair_sw,
                      double* out, std::string& err);

// create + solve + fetch.  Returns false (and sets err) only for a device/launch failure;
// per-molecule failures are in out[k].status.
bool engine_run(const Params& P, const std::vector<MolIn>& mols, const EngineOptions& opt, const RunRequest& req,
                std::vector<MolResult>& out, std::string& err);

// The same single point with no device allocation: every buffer is carved from the library's
// static per-GPU scratch block (NOALLOC_BYTES) and the work runs on that GPU's own stream, both
// set up once per device.  Returns 0, or a pm3_status code with err set (5 when the molecule's
// buffers do not fit the block, 6 for a device failure).  Calls on one GPU share the block: they
// must not overlap.
// run_dc (src/dc/scf.rs:74-160) of one molecule on the device, with want_gradient also dc_gradient
// (src/dc/derivatives.rs:55-121).  Returns 0, or a pm3_status code with err set (1 invalid input,
// 2 missing element, 3 not converged: the scalars are then set and the matrices empty, 4
// unsupported element, 5 resource limit, 6 device failure).
int dc_run(const Params& P, const MolIn& mol, const EngineOptions& opt, const DcOptionsHost& dc, DcResultHost& out,
           std::string& err, bool want_gradient = false);

constexpr int NOALLOC_MAX_ATOMS = 128;
constexpr int NOALLOC_MAX_NAO = 128;
int engine_run_noalloc(const Params& P, const MolIn& mol, const EngineOptions& opt, bool want_density, MolResult& out,
                       std::string& err);

}  // namespace pm3host
