/* C ABI of the MI355X PM3 engine (libpm3gpu.so).
 *
 * The entry points replace the reference's molecular RHF boundary:
 *   pm3_run        <- run_pm3(&Molecule, &Pm3Parameters, &Pm3Options) -> Result<Pm3Result>
 *                     (src/scf.rs:450-454, re-exported at src/lib.rs:108)
 *   pm3_gradient   <- closed_form_gradient(&Molecule, &Pm3Parameters, &Pm3Options)
 *                     -> Result<GradientResult> (src/gradient.rs:201-205, src/lib.rs:72)
 *   pm3_params_standard <- Pm3Parameters::standard() (src/params.rs:135-137)
 *   pm3_run_batch  -- many independent molecules in one persistent launch (no reference
 *                     counterpart: the reference solves one molecule per call).
 *   pm3_optimize   <- optimize(&Molecule, &Pm3Parameters, &Pm3Options, &OptOptions)
 *                     -> Result<OptResult> (src/optimizer.rs:233-262)
 *   pm3_optimize_batch -- the same L-BFGS for many molecules, one batched solve per round.
 *   pm3_optimize_dc <- optimize_dc(&Molecule, &Pm3Parameters, &Pm3Options, &DcOptions, &OptOptions)
 *                     -> Result<DcOptResult> (src/optimizer.rs:277-320)
 *
 * Units follow the reference (src/system.rs:21-26): positions in Bohr, energies in eV, heat
 * of formation in kcal/mol, gradient in eV/Bohr, dipole in Debye.  Matrices are n x n
 * row-major doubles (n = pm3_query_nao), eigenvectors as columns, eigenvalues ascending
 * (src/linalg.rs:217-263).  All buffers are caller-allocated host memory.
 *
 * Errors: every call returns a pm3_status whose values map one-to-one onto the reference's
 * Pm3Error variants (src/error.rs:12-47); pm3_last_error_message() gives the detail text of the
 * calling thread's last failure.  Non-convergence is an error (PM3_SCF_NOT_CONVERGED), never a
 * result with converged = 0 (src/scf.rs:754-760); the result struct still carries the
 * iteration count and last density change.
 */
#ifndef PM3_GPU_H
#define PM3_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  PM3_OK = 0,
  PM3_INVALID_INPUT = 1,       /* Pm3Error::InvalidInput */
  PM3_MISSING_ELEMENT = 2,     /* Pm3Error::MissingElement(z) */
  PM3_SCF_NOT_CONVERGED = 3,   /* Pm3Error::ScfNotConverged{iterations, error} */
  PM3_UNSUPPORTED = 4,         /* d elements, He/Ne, n >= 4 overlaps: Pm3Error::UnsupportedDOrbitals(z) and
                                  the elements this engine has no closed-form overlap for */
  PM3_RESOURCE_LIMIT = 5,      /* Pm3Error::ResourceLimit{operation, required_mb, limit_mb} */
  PM3_DEVICE_ERROR = 6         /* HIP runtime failure (no reference counterpart) */
} pm3_status;

/* Reference (src/scf.rs:120-129): AUTO picks RHF for n_alpha == n_beta and UHF otherwise, RHF refuses an
 * open shell (PM3_INVALID_INPUT), UHF takes the unrestricted path even for a singlet. */
typedef enum { PM3_REFERENCE_AUTO = 0, PM3_REFERENCE_RHF = 1, PM3_REFERENCE_UHF = 2 } pm3_reference;

typedef enum { PM3_ACCEL_NONE = 0, PM3_ACCEL_CDIIS = 1, PM3_ACCEL_ADIIS_CDIIS = 2 } pm3_accelerator;
/* ScfGuess (src/scf.rs:133-158) plus the learned initial guess of this engine. */
typedef enum { PM3_GUESS_AUTO = 0, PM3_GUESS_SAD = 1, PM3_GUESS_CORE = 2, PM3_GUESS_SYMMETRY_BROKEN = 3,
               PM3_GUESS_LEARNED = 4 } pm3_guess;
typedef enum { PM3_STABILITY_OFF = 0, PM3_STABILITY_AUTO = 1, PM3_STABILITY_ALWAYS = 2 } pm3_stability;

/* Molecule (src/system.rs:28-40): z[i] and xyz_bohr[3*i..3*i+2] per atom. */
typedef struct {
  int32_t n_atoms;
  const int32_t* z;
  const double* xyz_bohr;
  double charge;
  int32_t multiplicity;
} pm3_molecule;

struct pm3_guess_model; /* trained network of the learned guess, see pm3_guess.h */

/* The fields of Pm3Options (src/scf.rs:160-287) that the molecular single point reads, plus the
 * network that PM3_GUESS_LEARNED uses.  The fields that only the reference's Hessian, periodic,
 * correction-variant and d-element paths read (variant, mmok, d_penalty_*, hessian_*, cphf_max_iter)
 * have no counterpart here; charge and multiplicity live in pm3_molecule. */
typedef struct {
  int32_t max_scf;        /* default 200 */
  double e_tol;           /* default 1e-8 eV */
  double p_tol;           /* default 1e-7 (RMS density change) */
  int32_t use_diis;       /* default 1 */
  int32_t accelerator;    /* pm3_accelerator, default ADIIS_CDIIS */
  double adiis_switch;    /* default 0.1 */
  int32_t guess;          /* pm3_guess, default AUTO (= SAD) */
  int32_t stability;      /* pm3_stability, default AUTO */
  int32_t scf_memory_mb;  /* default 512: sets the DIIS depth (src/scf.rs:309-327) */
  const struct pm3_guess_model* guess_model; /* required by PM3_GUESS_LEARNED, default NULL */
  double level_shift_ev;  /* default 0.  With damping non-zero it "steers" the first path and switches the
                             stability search off (src/scf.rs:563-607); the unrestricted loop applies
                             F + shift (I - P_spin) and damping, the restricted loop reads neither */
  double damping;         /* default 0, in [0, 1) */
  int32_t reference;      /* pm3_reference, default AUTO */
  int32_t has_field;      /* default 0; 1: the uniform field below is on (Pm3Options::field = Some) */
  double field[3];        /* eV per Bohr per elementary charge, referenced to the coordinate origin */
  int32_t integral_memory_mb; /* pair-cache ceiling in MiB (src/hamiltonian.rs:202-211); default 0 = the
                             process default, PM3_MAX_PAIR_CACHE_MB or 4096 (a value 0 there disables it) */
} pm3_options;

/* Scalars of Pm3Result (src/scf.rs:352-396), plus the detail of a failure. */
typedef struct {
  double total_ev, electronic_ev, core_ev, heat_of_formation_kcal;
  double dipole_debye[3], dipole_magnitude;
  double homo_ev, lumo_ev;     /* over both spin channels for an unrestricted result (src/scf.rs:797-820) */
  double last_change;          /* RMS density change of the last iteration (the sum over spins for UHF) */
  double scf_improvement_ev;
  int32_t iterations, converged, scf_paths_tried, n_occ, nao, has_homo, has_lumo;
  int32_t guess_used;          /* 0 superposition / other rule, 1 learned, 2 learned rejected
                                  (first-iteration commutator norm above the superposition's) */
  int32_t unrestricted;        /* Pm3Result::unrestricted */
  int32_t n_beta;              /* Pm3Result::n_beta (= n_occ for a restricted result) */
  int32_t missing_z;           /* PM3_MISSING_ELEMENT / PM3_UNSUPPORTED: the atomic number refused */
  double required_mb, limit_mb; /* PM3_RESOURCE_LIMIT: the pair cache's size and the ceiling, MiB */
} pm3_result;

typedef struct pm3_params pm3_params; /* opaque, immutable once built */

/* Pm3Parameters::standard(): the embedded PM3 table for the supported s/p elements
 * (H, Li, Be, B, C, N, O, F, Na, Mg, Al, Si, P, S, Cl).  Returns NULL on failure. */
const pm3_params* pm3_params_standard(void);

/* Fills *out with the reference Pm3Options::default() values (src/scf.rs:259-287). */
void pm3_options_default(pm3_options* out);

/* Number of AOs of the molecule (Basis::build, src/basis.rs:39-62); -1 for an unknown element. */
int32_t pm3_query_nao(const pm3_molecule* mol, const pm3_params* params);
/* Atomic mass (amu) of element z in the parameter set (Pm3Element::mass, src/params.rs); NaN for
 * an element the set lacks. */
double pm3_element_mass(const pm3_params* params, int32_t z);

/* run_pm3 (src/scf.rs:450-848).  Optional outputs may be NULL: density[n*n] (total), mo_energies[n],
 * mo_coeff[n*n] (alpha for an unrestricted result), charges[n_atoms], and for an unrestricted result
 * spin_density[n*n] (P_alpha - P_beta), mo_energies_beta[n], mo_coeff_beta[n*n] (zero-filled for a
 * restricted one). */
pm3_status pm3_run(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts, pm3_result* result,
                   double* density, double* mo_energies, double* mo_coeff, double* charges, double* spin_density,
                   double* mo_energies_beta, double* mo_coeff_beta);

/* pm3_run without any device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (code under HIP graph capture, serving loops that forbid hipMalloc).
 * Every device buffer of the call -- staged inputs, integrals, SCF workspace, outputs -- is carved
 * from one static block of PM3_NOALLOC_SCRATCH_MB MiB per GPU, and the work runs on one stream per
 * GPU, both set up at the first call on that device; no hipMalloc, hipFree or allocating library
 * routine is called.  Molecules up to PM3_NOALLOC_MAX_ATOMS atoms and PM3_NOALLOC_MAX_NAO AOs;
 * larger ones get PM3_RESOURCE_LIMIT before any device work.  Outputs as pm3_run (density[n*n] and
 * charges[n_atoms] may be NULL; densities are purified, so no orbitals).  Calls on one GPU share the
 * block and the stream and must not overlap: concurrent calls on the same GPU race. */
enum { PM3_NOALLOC_MAX_ATOMS = 128, PM3_NOALLOC_MAX_NAO = 128, PM3_NOALLOC_SCRATCH_MB = 32 };
pm3_status pm3_run_noalloc(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                           pm3_result* result, double* density, double* charges);

/* closed_form_gradient (src/gradient.rs:201-276): SCF plus the Hellmann-Feynman gradient at the
 * converged density.  gradient[3*n_atoms] in eV/Bohr; density[nao*nao] and the Mulliken charges
 * of GradientResult::scf, charges[n_atoms], may be NULL. */
pm3_status pm3_gradient(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                        pm3_result* result, double* gradient, double* density, double* charges);

/* response_fock_rhf (src/fock.rs:444-485), the CPHF response operator of the analytic Hessian
 * (src/hessian.rs:1025-1027), for n_rhs densities at once: out[r] = J(dP_r) - K(dP_r / 2) with no
 * H_core, for symmetric dp[r] (n_rhs x nao x nao, row-major, the molecule's AO order).  pair_sw
 * (may be NULL: every pair 1) weights each atom pair's two-centre terms, one finite value per pair
 * in the reference's order ((u, v), u < v, row by row over the atoms); a zero drops the pair.
 * The molecule must be one run_pm3 accepts (charge and multiplicity); its electrons are not
 * otherwise used.  out: n_rhs x nao x nao. */
pm3_status pm3_response_fock(const pm3_molecule* mol, const pm3_params* params, int32_t n_rhs, const double* dp,
                             const double* pair_sw, double* out);

/* numerical_hessian (src/hessian.rs:268-322): the 3N x 3N Cartesian Hessian (eV/Bohr^2, row-major)
 * by central differences of the analytic gradient with displacement `step` (Bohr), symmetrised by
 * averaging, under the options tightened as tighten_scf_for_hessian does (:169-175: e_tol <= 1e-12,
 * p_tol <= 1e-11, max_scf >= 500).  All 6N displaced gradients are one batched solve.  Returns the
 * first displaced solve's failure, if any. */
pm3_status pm3_numerical_hessian(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                                 double step, double* hessian);

/* DcOptions (src/dc/partition.rs:40-97): how divide-and-conquer cuts the molecule.  The screened far
 * field (long_range_cutoff) has no counterpart here: every pair keeps its exact two-electron table. */
typedef struct {
  double core_radius;      /* Bohr, default 6.0: target radius of a core region */
  double buffer_radius;    /* Bohr, default 9.0: how far beyond a core its buffer reaches */
  int32_t min_core_atoms;  /* default 1: a smaller core joins the nearest existing one */
  double electron_tol;     /* default 1e-9: bracket width at which the chemical potential is taken */
  double smearing_ev;      /* default 0.1: Fermi-Dirac broadening of the subsystem occupations (eV) */
} pm3_dc_options;

/* Scalars of DcResult (src/dc/scf.rs:25-46). */
typedef struct {
  double total_ev, electronic_ev, core_ev, heat_of_formation_kcal;
  double fermi_ev, fermi_beta_ev;  /* global chemical potential of each spin channel (eV) */
  int32_t iterations, converged, unrestricted;
  int32_t dropped_pairs, largest_subsystem, n_subsystems, nao;
} pm3_dc_result;

/* Fills *out with DcOptions::default() (src/dc/partition.rs:85-97). */
void pm3_dc_options_default(pm3_dc_options* out);

/* run_dc (src/dc/scf.rs:74-160): divide-and-conquer SCF of one molecule on the device -- every
 * subsystem's Fock block diagonalised in one batched eigensolve per iteration, one global chemical
 * potential, the Yang-Lee assembly and Pulay DIIS on the density (src/densitydiis.rs:34-119).  The
 * options read are max_scf, e_tol, p_tol, scf_memory_mb (the density DIIS depth), damping (0: 0.3),
 * reference, has_field / field and integral_memory_mb.  Optional outputs: density[n*n] (total),
 * spin_density[n*n] (P_alpha - P_beta, zero-filled for a restricted run), charges[n_atoms].
 * PM3_SCF_NOT_CONVERGED fills *result and leaves the arrays unwritten, as the reference returns an
 * error there. */
pm3_status pm3_run_dc(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                      const pm3_dc_options* dc, pm3_dc_result* result, double* density, double* spin_density,
                      double* charges);

/* dc_gradient (src/dc/derivatives.rs:55-121): run_dc, then the fixed-density analytic gradient at the
 * partitioned density (eV/Bohr, gradient[3*n_atoms]; the spin form for an unrestricted result),
 * minus f q_A in a field.  The other outputs as pm3_run_dc. */
pm3_status pm3_dc_gradient(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                           const pm3_dc_options* dc, pm3_dc_result* result, double* gradient, double* density,
                           double* spin_density, double* charges);

/* Batched solve of many independent molecules in one persistent launch.  Atoms of molecule k are
 * [atom_ptr[k], atom_ptr[k+1]) of z / xyz_bohr; charge may be NULL (all neutral), multiplicity may
 * be NULL (all singlets).  Per molecule: status[k] (pm3_status), results[k].  want_gradient != 0
 * also fills gradient[3 * atom_ptr[k] ...] (may be NULL otherwise).  Returns PM3_OK when the launch
 * itself succeeded; per-molecule failures are reported in status[]. */
pm3_status pm3_run_batch(int32_t n_mol, const int32_t* atom_ptr, const int32_t* z, const double* xyz_bohr,
                         const double* charge, const int32_t* multiplicity, const pm3_params* params,
                         const pm3_options* opts, int32_t want_gradient, pm3_status* status, pm3_result* results,
                         double* gradient);

/* Device-resident batch: pm3_batch_create uploads the molecules and allocates the workgroup
 * scratch once; pm3_batch_solve is the persistent launch alone (inputs already in HBM);
 * pm3_batch_results copies back per-molecule status, scalars and optionally the Mulliken charges
 * (charges[atom_ptr[k] + i]) and densities (concatenated nao_k^2 blocks; only for a batch
 * created with PM3_BATCH_WANT_DENSITY, else PM3_INVALID_INPUT).  No reference
 * counterpart: this is the batched entry the reference lacks; pm3_run_batch is
 * create + solve + results + destroy. */
typedef struct pm3_batch pm3_batch;
/* PM3_BATCH_FORCE_LARGE solves every molecule on the multi-kernel path that molecules too large
 * for one workgroup (more than about 190 AOs) always take: a host-driven SCF whose steps are
 * grid-wide kernels, rocBLAS products and a rocSOLVER eigensolve. */
/* PM3_BATCH_WARM_START: every solve after the first starts each molecule's first SCF path from
 * that molecule's converged density of the previous solve, kept in device memory, in place of
 * the options' guess (the stability rungs keep theirs; a molecule whose previous solve did not
 * converge starts from the guess).  Meant for a batch whose geometries move
 * a little between solves (pm3_batch_set_positions); the fixed point is the same, the SCF path
 * shorter.  Molecules on the multi-kernel path keep the options' guess. */
/* A batch of one restricted molecule above the tuning entry solo_large_nao AOs (and so a pm3_run or
 * pm3_gradient call of one) takes the multi-kernel path, unless it is a warm-start batch or asks
 * for PM3_BATCH_KEEP_BATCHED: one workgroup would leave the rest of the GPU idle. */
enum { PM3_BATCH_WANT_DENSITY = 1, PM3_BATCH_WANT_GRADIENT = 2, PM3_BATCH_FORCE_LARGE = 4, PM3_BATCH_WARM_START = 8,
       PM3_BATCH_KEEP_BATCHED = 16 };
pm3_status pm3_batch_create(int32_t n_mol, const int32_t* atom_ptr, const int32_t* z, const double* xyz_bohr,
                            const double* charge, const int32_t* multiplicity, const pm3_params* params,
                            const pm3_options* opts, int32_t flags, pm3_batch** out);
pm3_status pm3_batch_solve(pm3_batch* batch);
pm3_status pm3_batch_results(pm3_batch* batch, pm3_status* status, pm3_result* results, double* charges,
                             double* density);
/* Spin densities P_alpha - P_beta of the last solve (a PM3_BATCH_WANT_DENSITY batch), in the layout of
 * pm3_batch_results' densities; zeros for restricted results. */
pm3_status pm3_batch_spin_densities(pm3_batch* batch, double* spin_density);
void pm3_batch_destroy(pm3_batch* batch);
/* New positions (Bohr, the creation layout xyz_bohr[3 * (atom_ptr[k] + i) + c], the same atoms)
 * for every molecule of the batch; the next pm3_batch_solve computes integrals, SCF and gradients
 * there.  A batch using PM3_GUESS_LEARNED rebuilds its 4 Angstrom neighbour lists for the new
 * positions (growing their device buffers when the pair count outgrows them), so the next solve's
 * learned guess is the one a batch created at these positions would make.  PM3_INVALID_INPUT for
 * positions that are not finite or put two atoms of a molecule together; the batch is unchanged
 * on such a failure.  No reference counterpart: a geometry change in the reference is a new
 * run_pm3 call (src/optimizer.rs:344-347). */
pm3_status pm3_batch_set_positions(pm3_batch* batch, const double* xyz_bohr);
/* For a PM3_BATCH_WARM_START batch dst: take each molecule's last density and result record
 * from the solved batch src, where src_index[k] is dst molecule k's index in src (the same atoms
 * in the same order), so that dst's first solve is warm-started too.  PM3_INVALID_INPUT when a
 * molecule is not its source's.  No reference counterpart: it lets a caller rebuild a smaller
 * batch of moving molecules without losing their densities. */
pm3_status pm3_batch_carry_densities(pm3_batch* dst, const pm3_batch* src, const int32_t* src_index);
/* Analytic gradients of the last solve (batch created with PM3_BATCH_WANT_GRADIENT), eV/Bohr,
 * gradient[3 * (atom_ptr[k] + i) + c]; the same per-molecule gradient pm3_gradient returns. */
pm3_status pm3_batch_gradient(pm3_batch* batch, double* gradient);
/* Device time (ms, HIP events on the launch stream) of the last pm3_batch_solve's kernel. */
double pm3_batch_last_kernel_ms(const pm3_batch* batch);
/* Device time (ms) of the last pm3_batch_solve's pair-integral kernel. */
double pm3_batch_last_pair_ms(const pm3_batch* batch);
/* Device time (ms) of the last pm3_batch_solve's learned-guess kernels (features + network). */
double pm3_batch_last_guess_ms(const pm3_batch* batch);
/* The same split into the feature kernel and the network kernel, plus the batch's number of
 * directed atom pairs inside the 4 Angstrom cutoff (the network's rows). */
pm3_status pm3_batch_guess_timing(const pm3_batch* batch, double* features_ms, double* network_ms,
                                  int64_t* directed_pairs);
/* Per-phase wall-clock ticks (10 ns) of the last solve summed over workgroups, collected only when
 * the environment has PM3_PHASE_TIMERS=1 at pm3_batch_create (zeros otherwise): out[0..9] H_core
 * finish, Fock, energy + DIIS history and Gram sums, dense products (F P, warm-start transform),
 * eigensolve or purification, sort + C V or frontier levels, density change, bookkeeping, DIIS
 * coefficient solve, DIIS extrapolation; out[10] frontier-level Lanczos steps, out[11] frontier-level
 * Lanczos runs (purified route); out[12] unused; out[13] Jacobi sweeps or
 * purification products, out[14] eigensolves or purifications, out[15] Jacobi rounds or
 * purification fallbacks. */
pm3_status pm3_batch_phase_ticks(const pm3_batch* batch, uint64_t* out16);
/* Molecules on the multi-kernel path, last solve: device time (ms, HIP events around each rocBLAS
 * call), count and FLOP issued (2 n^3 per dgemm, n^3 per dsyrk, 2 b^3 per b x b block product of
 * the block-lower form: 1.5 n^3 at q = 2) of the purification products, device
 * time and count of the Fock builds, and the Lanczos steps and runs of the frontier levels.
 * Zeros when the batch has no such molecule; any output pointer may be null. */
pm3_status pm3_batch_large_timing(const pm3_batch* batch, double* product_ms, int64_t* products,
                                  int64_t* product_flop, double* fock_ms, int64_t* fock_builds,
                                  int64_t* lanczos_steps, int64_t* lanczos_runs);

/* Launch shapes and tile choices of the kernels, and an autotuner for them.  No reference
 * counterpart (the reference has no device kernels).  The entries change speed only: results
 * agree to rounding whatever the values, and pm3_autotune checks that on the workload it tunes.
 * A process-wide table is read by every pm3_batch_create / pm3_run / pm3_gradient; it starts at
 * the defaults, or at the file the environment variable PM3_TUNING_FILE names.  Setting or
 * loading it is not thread-safe against concurrent batch creation. */
typedef struct {
  int32_t pair_block;         /* threads per block of the pair-integral and gradient-pair kernels: 64, 128, 256 */
  int32_t pair_blocks_per_cu; /* their grid (grid-stride loops), blocks per CU: 1..64 */
  int32_t sp2_tile16_n;       /* purification products on 16x16 wave tiles up to this many AOs, 32x32 above */
  int32_t sp2_lds;            /* purification operand placements: 0 HBM, 1 + both in LDS, 2 + X in LDS,
                                 3 + X in LDS overwritten in place by each product,
                                 4 + the leading rows of X in LDS where the whole of X does not fit,
                                 5 + X's upper 16x16 blocks (tiled, swizzled) in LDS, updated in place,
                                 6 + that tiled image split between LDS and HBM where it does not fit */
  int32_t sp2_tight_n;        /* purification spectral bounds from Lanczos from this many AOs on */
  int32_t large_syrk;         /* multi-kernel path: purification products by dsyrk (1) or dgemm (0) */
  int32_t large_split;        /* multi-kernel path above 1024 AOs: 1 purification steps by the library's
                                 lower-triangle tile kernel (product and update in one launch); q in
                                 2..4 only the q x q block-lower part of each product by q batched
                                 rocBLAS dgemm rounds when q divides the AO count; 0 one dgemm */
  int32_t wg128_nao;          /* SCF workgroup width: molecules up to this many AOs (and LDS needs within
                                 36 KiB) on 128-thread workgroups, four per CU */
  int32_t wg256_nao;          /* then up to this many AOs (within 76 KiB) on 256-thread workgroups, two per
                                 CU; the rest on 512-thread workgroups, one per CU */
  int32_t solo_large_nao;     /* a batch (or pm3_run / pm3_gradient call) of one restricted molecule above
                                 this many AOs takes the multi-kernel path */
} pm3_tuning;
void pm3_tuning_default(pm3_tuning* t);
pm3_status pm3_tuning_get(pm3_tuning* t);
pm3_status pm3_tuning_set(const pm3_tuning* t); /* PM3_INVALID_INPUT for an entry out of range */
/* Text file, one "key value" line per entry (keys as the field names), '#' comments. */
pm3_status pm3_tuning_load(const char* path);
pm3_status pm3_tuning_save(const char* path);
typedef struct {
  double default_ms;  /* device time per solve with the starting table (median of reps) */
  double best_ms;     /* the same with the returned table */
  int32_t candidates; /* tables timed, the starting one included */
  int32_t rejected;   /* candidates whose statuses or heats of formation (1e-5 kcal/mol) differed */
} pm3_tune_report;
/* Coordinate descent over the entries on this workload (arguments as pm3_batch_create; one
 * batch per candidate, one untimed and `reps` timed solves each).  On success the process-wide
 * table is the best one found and is returned in *best; on failure it is left as it was. */
pm3_status pm3_autotune(int32_t n_mol, const int32_t* atom_ptr, const int32_t* z, const double* xyz_bohr,
                        const double* charge, const pm3_params* params, const pm3_options* opts, int32_t flags,
                        int32_t reps, pm3_tuning* best, pm3_tune_report* report);

/* Detail text of the calling thread's last error ("" if none). */
const char* pm3_last_error_message(void);

/* OptOptions (src/optimizer.rs:15-35).  grad_step is the reference's finite-difference step; the
 * gradient here is analytic, so it is accepted and not used.  warm_start has no reference
 * counterpart: 1 starts each round's SCF from the molecule's density of its previous point
 * (PM3_BATCH_WARM_START) instead of the options' guess, as the reference does every time. */
typedef struct {
  int32_t max_iter;   /* default 200 */
  double gtol;        /* default 1e-3 eV/Bohr, on the largest gradient component */
  double grad_step;   /* default 5e-4 Bohr */
  int32_t history;    /* default 8 L-BFGS pairs */
  int32_t warm_start; /* default 0 */
} pm3_opt_options;
void pm3_opt_options_default(pm3_opt_options* out);

/* OptResult (src/optimizer.rs:45-52): its flags and counts, and the scalars of the first and last
 * trajectory steps (OptStep, :37-43).  evaluations counts the SCF + gradient solves of the
 * molecule: the first point and every line-search trial (an accepted trial is not solved again). */
typedef struct {
  int32_t status; /* PM3_OK, or the failure of the first evaluation */
  int32_t converged, iterations, evaluations;
  int32_t trajectory_len; /* accepted points, the starting one included (<= max_iter + 1) */
  double energy_ev, heat_of_formation_kcal, max_gradient; /* at the final geometry */
  double first_energy_ev;                                 /* at the starting geometry */
} pm3_opt_result;

/* OptStep (src/optimizer.rs:37-43) without its positions, which go to a separate array. */
typedef struct {
  double energy_ev, heat_of_formation_kcal, max_gradient;
} pm3_opt_step;

/* optimize (src/optimizer.rs:233-262) for many molecules at once: L-BFGS with the
 * backtracking Armijo search of drive (:77-221) per molecule, every round one batched SCF +
 * gradient solve over all running molecules.  Molecules in the atom_ptr / z / xyz_bohr / charge /
 * multiplicity layout of pm3_run_batch; results[k] per molecule, final positions xyz_out[3 * atom_ptr[n_mol]]
 * (Bohr).  Returns the first device or launch failure, else PM3_OK (per-molecule failures are in
 * results[k].status).  With PM3_GUESS_LEARNED a solve starts from the learned guess at the
 * molecule's current positions, unless warm_start gives it the molecule's last converged density.
 *
 * The trajectory (OptResult::trajectory, drive :121-126,207-212: the starting point and every
 * accepted point) is written when the optional arrays are given, with room for T = max_iter + 1
 * steps per molecule: steps[k * T + j] and the positions of step j of molecule k at
 * traj_xyz[3 * (T * atom_ptr[k] + j * n_k)] (Bohr, n_k atoms).  results[k].trajectory_len is
 * the number of steps written. */
pm3_status pm3_optimize_batch(int32_t n_mol, const int32_t* atom_ptr, const int32_t* z, const double* xyz_bohr,
                              const double* charge, const int32_t* multiplicity, const pm3_params* params,
                              const pm3_options* opts, const pm3_opt_options* opt, pm3_opt_result* results,
                              double* xyz_out, pm3_opt_step* steps, double* traj_xyz);
/* optimize for one molecule (pm3_optimize_batch of one); xyz_out[3 * n_atoms], and optionally
 * steps[max_iter + 1] and traj_xyz[3 * n_atoms * (max_iter + 1)]. */
pm3_status pm3_optimize(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                        const pm3_opt_options* opt, pm3_opt_result* result, double* xyz_out, pm3_opt_step* steps,
                        double* traj_xyz);

/* optimize_dc (src/optimizer.rs:277-310) -> Result<DcOptResult> (:313-320): the optimiser of
 * pm3_optimize over the divide-and-conquer gradient (pm3_dc_gradient at every point, a failed trial
 * a rejected step).  *result as pm3_optimize's; *scf (may be NULL) the DcResult scalars of the last
 * accepted point; xyz_out, steps and traj_xyz as pm3_optimize's.  Returns the first evaluation's
 * error, else PM3_OK (not converged within max_iter is result->converged = 0, as the reference). */
pm3_status pm3_optimize_dc(const pm3_molecule* mol, const pm3_params* params, const pm3_options* opts,
                           const pm3_dc_options* dc, const pm3_opt_options* opt, pm3_opt_result* result,
                           pm3_dc_result* scf, double* xyz_out, pm3_opt_step* steps, double* traj_xyz);

#ifdef __cplusplus
}
#endif
#endif /* PM3_GPU_H */
