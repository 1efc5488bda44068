/* Kernel-level entry points of libpm3gpu.so: one stage of the device pipeline at a time, so each
 * kernel can be checked against the CPU oracle on its own.  Same conventions as pm3_gpu.h
 * (Bohr, eV, n x n row-major in the caller's AO order unless stated).  Each runs the same device
 * code the SCF kernel runs, on one workgroup. */
#ifndef PM3_GPU_STAGES_H
#define PM3_GPU_STAGES_H

#include "pm3_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

/* H_core and the core-core energy: replaces build_core_limited's h_core
 * (src/hamiltonian.rs:191-344) and core_core_energy (src/repulsion.rs:13-27). */
pm3_status pm3_stage_hcore(const pm3_molecule* mol, const pm3_params* params, double* h_core, double* core_ev);

/* F = H_core + G(P) for a given density: replaces build_fock (src/fock.rs:488-500). */
pm3_status pm3_stage_fock(const pm3_molecule* mol, const pm3_params* params, const double* density, double* fock);

/* Eigenpairs of a symmetric n x n matrix, ascending, eigenvectors as columns: replaces
 * symmetric_eigen (src/linalg.rs:221-263).  n <= 190. */
pm3_status pm3_stage_eigen(int32_t n, const double* a, double* w, double* v);

/* Same, warm-started in an orthonormal basis c (columns): the form the SCF loop uses. */
pm3_status pm3_stage_eigen_warm(int32_t n, const double* a, const double* c, double* w, double* v);

/* C = A B for n x n matrices on the FP64 matrix cores: replaces Matrix::matmul
 * (src/linalg.rs:65-122).  n <= 190. */
pm3_status pm3_stage_gemm(int32_t n, const double* a, const double* b, double* c);

/* Aufbau density 2 C_occ C_occ^T of a symmetric Fock matrix (caller's AO order) without its
 * eigenvectors, by trace-correcting purification, n_occ from the molecule's electron count, plus
 * the HOMO and LUMO by Lanczos: what the batched SCF does in place of the eigensolve and
 * density_from_coeff when no orbitals are requested; replaces the pair symmetric_eigen
 * (src/scf.rs:1264) + density_from_coeff (src/scf.rs:982-994).  products = number of
 * purification products, -1 when the frontier gap is closed (no density returned then). */
pm3_status pm3_stage_purify(const pm3_molecule* mol, const pm3_params* params, const double* fock, double* density,
                            int32_t* products, double* homo, double* lumo);

/* One step of the multi-kernel path's purification on the FP64 matrix cores, for a symmetric
 * n x n matrix x (1 <= n <= 16384): y = x x (square != 0) or 2 x - x x, computed from the 64 x 64
 * tiles of the lower triangle and mirrored, with trace = tr y and norm2 = |y|_F^2 = tr y^2 summed
 * from the kernel's partials; slab = 8, 16 or 32 columns of the operand panels per LDS stage.  With
 * reps > 0 the same step runs reps more times and ms_per_launch is their mean device time.
 * workgroups = 0 launches one workgroup per tile; a positive count caps the grid (a multiple of 8,
 * at least 8) so that each workgroup takes several tiles, as beyond about 10 000 AOs.  One
 * product of sp2_density's loop in the large path (the reference obtains the density from
 * symmetric_eigen + density_from_coeff instead, src/scf.rs:1264, src/scf.rs:982-994). */
pm3_status pm3_stage_sym_step(int32_t n, int32_t square, int32_t slab, const double* x, double* y, double* trace,
                              double* norm2, int32_t reps, int32_t workgroups, double* ms_per_launch);

#ifdef __cplusplus
}
#endif
#endif /* PM3_GPU_STAGES_H */
