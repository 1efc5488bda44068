// Host-side interface of the learned initial density: the network (GuessModel), the training
// set built from solved batches (GuessData), and the view of a batch that the guess kernels read.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "guess_device.hpp"

namespace pm3host {

struct Batch;

// Device pointers and host bookkeeping of a batch's directed pairs (see batch_create).  Atom
// indices are global and in the engine's heavy-first order; pairs of a source atom are
// contiguous and ordered by the caller's index of the destination.
struct BatchGuessView {
  int nmol = 0;
  std::vector<int> atom0, nat, nao, ngp, eligible, solved;  // per caller molecule
  std::vector<long long> gp0, out_mat;
  std::vector<int> caller_atom;  // per global atom: the caller's index within its molecule
  long long ngp_total = 0, atoms_total = 0;
  const double *ax = nullptr, *ay = nullptr, *az = nullptr;
  const int8_t* elem = nullptr;
  const int *aoff = nullptr, *gbeg = nullptr, *gcnt = nullptr, *gsrc = nullptr, *gdst = nullptr, *gpart = nullptr;
  const double* out_P = nullptr;  // converged densities (n x n per molecule), null without density output
  uint16_t* X = nullptr;          // features of every directed pair (inference)
  float* Y = nullptr;             // network outputs of every directed pair (inference)
};

bool batch_guess_view(Batch* b, BatchGuessView& v, std::string& err);  // statuses need a solved batch

struct GuessModel;
struct GuessData;

bool guess_model_create(uint64_t seed, GuessModel*& out, std::string& err);
bool guess_model_load(const char* path, GuessModel*& out, std::string& err);
bool guess_model_save(const GuessModel* m, const char* path, std::string& err);
bool guess_model_params(const GuessModel* m, float* out, std::string& err);
bool guess_model_set_params(GuessModel* m, const float* in, std::string& err);
int guess_model_steps(const GuessModel* m);
void guess_model_destroy(GuessModel* m);

// Features and network outputs of every directed pair of a batch (into v.X / v.Y); `between`, when
// given, is recorded on the stream after the feature kernel.
bool guess_infer(const GuessModel* m, const BatchGuessView& v, std::string& err, hipStream_t stream,
                 hipEvent_t between = nullptr);

bool guess_data_create(GuessData*& out, std::string& err);
bool guess_data_add(GuessData* d, Batch* solved, long long& added, std::string& err);
void guess_data_stats(const GuessData* d, long long& n_mol, long long& n_atoms, long long& n_pairs);
void guess_data_destroy(GuessData* d);

struct GuessTrainOptions {
  int steps = 1000;
  int batch_molecules = 1024;
  float lr = 1.0e-3f, beta1 = 0.9f, beta2 = 0.999f, eps = 1.0e-8f;
  uint64_t seed = 1;
  float lr_final = 1.0e-5f;
  int decay_steps = 0;  // cosine from lr to lr_final over the model's steps [0, decay_steps)
};
bool guess_train(GuessModel* m, GuessData* d, const GuessTrainOptions& o, float* losses, double* seconds,
                 std::string& err);
bool guess_eval(const GuessModel* m, GuessData* d, double& loss, std::string& err);
// Gradient and loss of the whole set as a single step (no update).
bool guess_grad(const GuessModel* m, GuessData* d, float* grad, double& loss, std::string& err);

// Kernel-level hooks: the network on given inputs (x bf16, n x NF) -> y (n x NO); with dy, also
// the gradient of sum_p dy_p . y_p with respect to the parameters (NPARAM floats).
bool guess_mlp(const GuessModel* m, int n, const uint16_t* x, float* y, const float* dy, float* grad,
               std::string& err);
// Features of one batch's directed pairs (host copies), after batch_create.
// src/dst are global engine atom indices; caller[a] is the caller's index of global atom a within
// its molecule.
bool guess_batch_features(Batch* b, std::vector<int>& src, std::vector<int>& dst, std::vector<uint16_t>& x,
                          std::vector<int>& caller, std::string& err);

}  // namespace pm3host
