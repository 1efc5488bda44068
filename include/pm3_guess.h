/* Learned initial density of the MI355X PM3 engine (libpm3gpu.so).  No reference counterpart:
 * the reference starts every SCF from the superposition (SAD) density (src/scf.rs:855-881).
 *
 * A per-atom-pair network (three layers, hidden width 128, SiLU, bf16 weights with fp32
 * accumulation on the matrix cores) predicts density blocks for every directed pair of atoms
 * closer than 4 Angstrom; they are assembled into a rotation-equivariant guess density.  The SCF
 * keeps it only when its first-iteration commutator norm |F P - P F| is not larger than the
 * superposition density's; the SCF itself (and so the converged answer) is unchanged.  The
 * network covers molecules made of H, C and O; others use the superposition guess.
 *
 * Select it with pm3_options.guess = PM3_GUESS_LEARNED and pm3_options.guess_model.
 * Training runs on the GPU from densities this engine converged (pm3_guess_data_add).
 * Handles are independent; calls that share one handle must not overlap.
 */
#ifndef PM3_GUESS_H
#define PM3_GUESS_H

#include "pm3_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { PM3_GUESS_NF = 96, PM3_GUESS_NH = 128, PM3_GUESS_NO = 16, PM3_GUESS_NPARAM = 30864 };

typedef struct pm3_guess_model pm3_guess_model;
typedef struct pm3_guess_data pm3_guess_data;

/* Network with seeded random weights (outputs near zero, so the guess starts near SAD). */
pm3_status pm3_guess_model_create(uint64_t seed, pm3_guess_model** out);
/* Load / save the fp32 master parameters (little-endian file written by this library). */
pm3_status pm3_guess_model_load(const char* path, pm3_guess_model** out);
pm3_status pm3_guess_model_save(const pm3_guess_model* model, const char* path);
/* Parameters as fp32 [W1 (96x128, row 95 = first bias) | W2 (128x128) | b2 | W3 (128x16) | b3],
 * row-major [in][out]; set_params also resets the optimiser state. */
pm3_status pm3_guess_model_get_params(const pm3_guess_model* model, float* out);
pm3_status pm3_guess_model_set_params(pm3_guess_model* model, const float* params);
int32_t pm3_guess_model_steps(const pm3_guess_model* model); /* optimiser steps taken */
void pm3_guess_model_destroy(pm3_guess_model* model);

/* Training set: converged densities of solved batches (created with PM3_BATCH_WANT_DENSITY),
 * kept in device memory.  Molecules that did not converge, contain elements other than H, C and
 * O, or have more than 256 atoms are skipped; *n_added counts the rest. */
pm3_status pm3_guess_data_create(pm3_guess_data** out);
pm3_status pm3_guess_data_add(pm3_guess_data* data, pm3_batch* solved, int64_t* n_added);
pm3_status pm3_guess_data_stats(const pm3_guess_data* data, int64_t* n_mol, int64_t* n_atoms, int64_t* n_pairs);
void pm3_guess_data_destroy(pm3_guess_data* data);

/* Adam on the per-atom mean squared error of the assembled on-site and pair blocks.  Each step
 * draws batch_molecules molecules without replacement (reshuffled per epoch, seeded).
 * losses[steps] receives the loss before each step (may be NULL); seconds the wall time.
 * With decay_steps > 0 the step size follows a cosine from lr to lr_final over the model's
 * optimiser steps 0..decay_steps (counted across calls), and stays at lr_final after. */
typedef struct {
  int32_t steps;
  int32_t batch_molecules; /* default 1024 */
  float lr, beta1, beta2, eps; /* defaults 1e-3, 0.9, 0.999, 1e-8 */
  uint64_t seed;
  float lr_final;          /* default 1e-5 */
  int32_t decay_steps;     /* default 0: constant lr */
} pm3_guess_train_options;
void pm3_guess_train_options_default(pm3_guess_train_options* out);
pm3_status pm3_guess_train(pm3_guess_model* model, pm3_guess_data* data, const pm3_guess_train_options* opts,
                           float* losses, double* seconds);
/* The same loss over the whole set (no update). */
pm3_status pm3_guess_eval(const pm3_guess_model* model, pm3_guess_data* data, double* loss);

/* Gradient (grad[PM3_GUESS_NPARAM]) and loss of the whole set as a single step, no update. */
pm3_status pm3_guess_grad(const pm3_guess_model* model, pm3_guess_data* data, float* grad, double* loss);

/* Kernel-level hooks for checks against the CPU restatement.
 * pm3_guess_mlp: outputs y[n*16] of inputs x[n*96] (bf16 bits); when dy is not NULL also the
 * gradient of sum_p dy_p . y_p with respect to the parameters (grad[PM3_GUESS_NPARAM]).
 * pm3_guess_features: directed pairs (caller atom indices) and their inputs; call with NULL
 * arrays to get *n_pairs first.
 * pm3_guess_density: the assembled learned density (n x n, caller AO order) and the squared
 * first-iteration commutator norms of it and of the superposition density. */
pm3_status pm3_guess_mlp(const pm3_guess_model* model, int32_t n, const uint16_t* x, float* y, const float* dy,
                         float* grad);
pm3_status pm3_guess_features(const pm3_molecule* mol, const pm3_params* params, int32_t* n_pairs, int32_t* src,
                              int32_t* dst, uint16_t* x);
pm3_status pm3_guess_density(const pm3_guess_model* model, const pm3_molecule* mol, const pm3_params* params,
                             double* density, double* commutator2_learned, double* commutator2_sad);

#ifdef __cplusplus
}
#endif
#endif /* PM3_GUESS_H */
