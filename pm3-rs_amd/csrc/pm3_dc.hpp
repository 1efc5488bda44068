// Divide-and-conquer SCF of one molecule on the device (run_dc, src/dc/scf.rs:74-160, loop
// src/dc/scf.rs:412-548, subsystem solve :604-690), included by pm3_engine.hip's main build.
//
// The molecule goes through the batch machinery on the host-driven path, which supplies its pair
// tables, H_core and the grid-wide Fock kernels.  Every density and Fock matrix is a dense np x np
// matrix in the engine's AO order that is zero outside the partition's pattern (a density only
// ever receives pattern entries, and only pattern entries of a Fock matrix are read), so the
// dense Fock build equals the reference's pattern-restricted one (build_fock_spin_sparse,
// src/fock.rs:359-442) on every entry that is used.  Per iteration:
//   Fock build(s) -> every subsystem's Fock block gathered into its size bucket's padded batch ->
//   batched symmetric eigensolve per bucket (rocSOLVER, strided batched) -> core weights of each
//   eigenvector -> the global chemical potential bisected by one workgroup -> occupied blocks
//   V f V^T as one batched product per bucket -> Yang-Lee weighted assembly, subsystem by
//   subsystem in the partition's order -> energy, change, damping and the density DIIS on the
//   pattern entries (src/densitydiis.rs:34-119).
// Buckets: subsystems are grouped by their size rounded up to BUCKET orbitals (capped at the
// largest subsystem), and neighbouring groups merged where the latency of another batched call
// costs more than the padding it saves (LATENCY), so the eigensolves and products, cubic in the
// padded size, run at a fraction of the volume of one batch padded to the largest subsystem.
// Padding: subsystem s of n_s orbitals sits in an nb x nb block (nb its bucket's size) whose extra
// diagonal carries a value above every eigenvalue of the Fock matrix (its Gershgorin bound + 1 eV),
// so the first n_s eigenpairs (ascending) are the subsystem's own and the rest are unit vectors on
// the padding.

namespace dc {

struct Part {
  std::vector<std::vector<int>> core, atoms, orb;  // orb: caller AO indices, core atoms first
  std::vector<std::vector<char>> in_core;
  std::vector<int> core_of;
};

// partition (src/dc/partition.rs:175-275): cores grown from the unassigned atom furthest from the
// centroid (the last such atom on ties), each claiming every free atom within core_radius of its
// seed; a core smaller than min_core_atoms joins the existing core nearest to it (the first on
// ties); a subsystem is its core plus every other atom within core_radius + buffer_radius of one
// of its core atoms.  Positions in Bohr, atoms and AOs in the caller's order.
static bool partition(int nat, const double* x, const std::vector<int>& norb, const std::vector<int>& off,
                      const DcOptionsHost& o, Part& p, std::string& err) {
  if (nat == 0) {
    err = "cannot partition an empty molecule";
    return false;
  }
  if (!(o.core_radius > 0.0) || !(o.buffer_radius >= 0.0)) {
    err = "divide-and-conquer needs a positive core radius and a non-negative buffer";
    return false;
  }
  auto d2 = [&](int a, const double* c) {
    const double dx = x[3 * a] - c[0], dy = x[3 * a + 1] - c[1], dz = x[3 * a + 2] - c[2];
    return dx * dx + dy * dy + dz * dz;
  };
  auto dist = [&](int a, int b) { return std::sqrt(d2(a, x + 3 * b)); };
  double cen[3] = {0.0, 0.0, 0.0};
  for (int a = 0; a < nat; ++a)
    for (int k = 0; k < 3; ++k) cen[k] += x[3 * a + k];
  for (int k = 0; k < 3; ++k) cen[k] *= 1.0 / (double)nat;
  const int none = -1;
  p.core_of.assign(nat, none);
  p.core.clear();
  int left = nat;
  while (left > 0) {
    int seed = -1;
    double far = 0.0;
    for (int a = 0; a < nat; ++a) {
      if (p.core_of[a] != none) continue;
      const double v = d2(a, cen);
      if (seed < 0 || v >= far) {  // the last of equal maxima
        seed = a;
        far = v;
      }
    }
    const int idx = (int)p.core.size();
    std::vector<int> mine;
    for (int a = 0; a < nat; ++a)
      if (p.core_of[a] == none && dist(a, seed) <= o.core_radius) {
        p.core_of[a] = idx;
        mine.push_back(a);
        --left;
      }
    if ((int)mine.size() < o.min_core_atoms && idx > 0) {
      int near = 0;
      double best = 0.0;
      for (int q = 0; q < idx; ++q) {
        double dq = INFINITY;
        for (int i : p.core[q])
          for (int j : mine) dq = std::min(dq, dist(i, j));
        if (q == 0 || dq < best) {  // the first of equal minima
          near = q;
          best = dq;
        }
      }
      for (int a : mine) p.core_of[a] = near;
      p.core[near].insert(p.core[near].end(), mine.begin(), mine.end());
      continue;
    }
    p.core.push_back(mine);
  }
  const double reach = o.core_radius + o.buffer_radius;
  const size_t ns = p.core.size();
  p.atoms.assign(ns, {});
  p.orb.assign(ns, {});
  p.in_core.assign(ns, {});
  for (size_t s = 0; s < ns; ++s) {
    const int id = p.core_of[p.core[s][0]];
    std::vector<int>& at = p.atoms[s];
    at = p.core[s];
    for (int a = 0; a < nat; ++a) {
      if (p.core_of[a] == id) continue;
      bool close = false;
      for (int c : p.core[s]) close = close || dist(a, c) <= reach;
      if (close) at.push_back(a);
    }
    for (int a : at)
      for (int k = 0; k < norb[a]; ++k) {
        p.orb[s].push_back(off[a] + k);
        p.in_core[s].push_back(p.core_of[a] == id);
      }
  }
  return true;
}

__host__ __device__ inline double fermi_dirac(double e, double mu, double s) {  // src/dc/scf.rs:726-738
  if (s <= 0.0) return e <= mu ? 1.0 : 0.0;
  const double x = (e - mu) / s;
  if (x > 40.0) return 0.0;
  if (x < -40.0) return 1.0;
  return 1.0 / (1.0 + exp(x));
}

constexpr int THREADS = 256;
constexpr int BLOCKS = 1024;
constexpr int MU_THREADS = 1024;
constexpr int BUCKET = 64;  // subsystem sizes are rounded up to a multiple of this for the batches
// Cost of one batched eigensolve and product in units of one padded block's nb^3: LATENCY nb +
// count nb^3.  The latency term (the solver's n sequential reduction steps, whatever the batch)
// is fitted to single-batch solves of 64- and 512-water clusters (51 x 319 and 348 x 587 AOs).
constexpr double LATENCY = 1.0e7;

__device__ inline int gt() { return blockIdx.x * blockDim.x + threadIdx.x; }
__device__ inline int gs() { return gridDim.x * blockDim.x; }

__global__ __launch_bounds__(THREADS) void add_kernel(const double* a, const double* b, double* out, long long m) {
  for (long long e = gt(); e < m; e += gs()) out[e] = a[e] + b[e];
}
__global__ __launch_bounds__(THREADS) void sub_kernel(const double* a, const double* b, double* out, long long m) {
  for (long long e = gt(); e < m; e += gs()) out[e] = a[e] - b[e];
}
// c <- d c + (1 - d) f (damp_sparse, src/dc/pattern.rs:131-136)
__global__ __launch_bounds__(THREADS) void damp_kernel(double* c, const double* f, long long m, double d) {
  for (long long e = gt(); e < m; e += gs()) c[e] = d * c[e] + (1.0 - d) * f[e];
}
__global__ __launch_bounds__(THREADS) void scale_kernel(const double* a, double* out, long long m, double s) {
  for (long long e = gt(); e < m; e += gs()) out[e] = a[e] * s;
}
// Zeros everywhere but on the pattern entries (the starting density read onto the pattern).
__global__ __launch_bounds__(THREADS) void keep_pattern_kernel(const double* src, const long long* at, long long nnz,
                                                               double* dst) {
  for (long long e = gt(); e < nnz; e += gs()) dst[at[e]] = src[at[e]];
}

// Every subsystem of one bucket into its padded nb x nb slot of the bucket's batch (column-major;
// the block is symmetric).  orb and nsz point at the bucket's first subsystem; the orbital table
// has ostride entries per subsystem.
__global__ __launch_bounds__(THREADS) void gather_kernel(const double* F, int np, const int* orb, const int* nsz,
                                                         int nsub, int nb, int ostride, double pad, double* Fs) {
  const long long per = (long long)nb * nb, tot = per * nsub;
  for (long long e = gt(); e < tot; e += gs()) {
    const int s = (int)(e / per), r = (int)(e % per), i = r % nb, j = r / nb, n = nsz[s];
    double v = 0.0;
    if (i < n && j < n) v = F[(long long)orb[(long long)s * ostride + i] * np + orb[(long long)s * ostride + j]];
    else if (i == j) v = pad;
    Fs[e] = v;
  }
}

// q[s][k] = sum over the core rows i of V[s](i, k)^2 (the eigenvector's electron count at full
// occupation), k < n_s; the padding's columns get 0.  One bucket: V holds its nb x nb blocks,
// core and q have ostride entries per subsystem.
__global__ __launch_bounds__(THREADS) void core_weight_kernel(const double* V, const int* nsz, const char* core,
                                                              int nsub, int nb, int ostride, double* q) {
  const long long tot = (long long)nsub * nb;
  for (long long e = gt(); e < tot; e += gs()) {
    const int s = (int)(e / nb), k = (int)(e % nb), n = nsz[s];
    double t = 0.0;
    if (k < n) {
      const double* col = V + (long long)s * nb * nb + (long long)k * nb;
      for (int i = 0; i < n; ++i)
        if (core[(long long)s * ostride + i]) t += col[i] * col[i];
    }
    q[(long long)s * ostride + k] = t;
  }
}

// The chemical potential that puts `target` electrons in the subsystems' core diagonals: the
// bracket of src/dc/scf.rs:693-724 (lowest eigenvalue - pad, highest + pad, widened at most 40
// times each way) and the bisection of :645-657 (200 halvings at most, stopping once the bracket is
// below electron_tol).  One workgroup; electron counts summed in a fixed order.  out[0] = mu.
__global__ __launch_bounds__(MU_THREADS) void mu_kernel(const double* w, const double* q, const int* nsz, int nsub,
                                                        int nmax, double target, double smear, double etol,
                                                        double* out) {
  __shared__ double part[MU_THREADS / 64];
  __shared__ double bc;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const long long tot = (long long)nsub * nmax;
  auto total = [&](double v, bool mx) {  // sum (mx: max) over the workgroup, same value everywhere
    for (int o = 32; o > 0; o >>= 1) {
      const double u = __shfl_xor(v, o, 64);
      v = mx ? fmax(v, u) : v + u;
    }
    __syncthreads();
    if (lane == 0) part[wv] = v;
    __syncthreads();
    double t = part[0];
    for (int x = 1; x < MU_THREADS / 64; ++x) t = mx ? fmax(t, part[x]) : t + part[x];
    return t;
  };
  auto electrons = [&](double mu) {
    double t = 0.0;
    for (long long e = tid; e < tot; e += MU_THREADS) {
      const int s = (int)(e / nmax), k = (int)(e % nmax);
      if (k >= nsz[s]) continue;
      const double f = fermi_dirac(w[e], mu, smear);
      if (f <= 0.0) continue;
      t += f * q[e];
    }
    return total(t, false);
  };
  double lo = INFINITY, hi = -INFINITY;
  for (long long e = tid; e < tot; e += MU_THREADS) {
    const int s = (int)(e / nmax), k = (int)(e % nmax);
    if (k >= nsz[s]) continue;
    lo = fmin(lo, w[e]);
    hi = fmax(hi, w[e]);
  }
  lo = -total(-lo, true);
  hi = total(hi, true);
  const double pad = 40.0 * fmax(smear, 1.0);
  double low = lo - pad, high = hi + pad;
  for (int wd = 0; wd < 40 && electrons(low) > target; ++wd) low -= pad;
  for (int wd = 0; wd < 40 && electrons(high) < target; ++wd) high += pad;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (low + high);
    if (electrons(mid) > target) high = mid;
    else low = mid;
    if (fabs(high - low) < etol) break;
  }
  if (tid == 0) out[0] = 0.5 * (low + high);
  (void)bc;
}

// Y[s](:, k) = f_k V[s](:, k) with f_k the Fermi-Dirac occupation at mu (0 where the reference
// skips f <= 1e-14, and for the padding).  One bucket: V and Y hold its nb x nb blocks, w has
// ostride entries per subsystem.
__global__ __launch_bounds__(THREADS) void occupy_kernel(const double* V, const double* w, const int* nsz, int nsub,
                                                         int nb, int ostride, const double* mu, double smear, double* Y) {
  const long long per = (long long)nb * nb, tot = per * nsub;
  const double m = mu[0];
  for (long long e = gt(); e < tot; e += gs()) {
    const int s = (int)(e / per), r = (int)(e % per), k = r / nb;
    double f = 0.0;
    if (k < nsz[s]) {
      f = fermi_dirac(w[(long long)s * ostride + k], m, smear);
      if (f <= 1.0e-14) f = 0.0;
    }
    Y[e] = f * V[e];
  }
}

// Subsystem s's block of V f V^T added to the dense density with the Yang-Lee weights (1 core-core,
// 1/2 core-buffer, 0 buffer-buffer, src/dc/partition.rs:161-169); one launch per subsystem, in order.
__global__ __launch_bounds__(THREADS) void assemble_kernel(const double* C, int ld, int n, const int* orb,
                                                           const char* core, double* P, int np) {
  const long long tot = (long long)n * n;
  for (long long e = gt(); e < tot; e += gs()) {
    const int i = (int)(e % n), j = (int)(e / n);
    const bool ci = core[i], cj = core[j];
    const double wt = ci && cj ? 1.0 : (ci || cj) ? 0.5 : 0.0;
    if (wt == 0.0) continue;
    P[(long long)orb[i] * np + orb[j]] += wt * C[(long long)j * ld + i];
  }
}

// Pattern entries of (a, b) into cur and of (fa - a, fb - b) into res (alpha first,
// gather_pair_sparse).
__global__ __launch_bounds__(THREADS) void pack_kernel(const double* a, const double* b, const double* fa,
                                                       const double* fb, const long long* at, long long nnz,
                                                       double* cur, double* res) {
  for (long long e = gt(); e < nnz; e += gs()) {
    const long long x = at[e];
    cur[e] = a[x];
    cur[nnz + e] = b[x];
    res[e] = fa[x] - a[x];
    res[nnz + e] = fb[x] - b[x];
  }
}

struct Mix {
  const double* v[8];
  double c[8];
  int k;
};
// sum_i c_i dens_i over the kept slots, written back onto the pattern of a and b
// (scatter_pair_sparse).
__global__ __launch_bounds__(THREADS) void extrapolate_kernel(Mix m, const long long* at, long long nnz, double* a,
                                                              double* b) {
  for (long long e = gt(); e < 2 * nnz; e += gs()) {
    double s = 0.0;
    for (int i = 0; i < m.k; ++i) s += m.c[i] * m.v[i][e];
    if (e < nnz) a[at[e]] = s;
    else b[at[e - nnz]] = s;
  }
}

}  // namespace dc

// The spin Fock build F = H + J(P) - K(X) on the grid (build_fock_spin, src/fock.rs:285-332).
__global__ __launch_bounds__(LG_THREADS) void lg_fock_pairs_x_kernel(KArgs A, MolDesc md, const double* P,
                                                                     const double* X, double xf, double* F,
                                                                     double* slot) {
  const Mol M = mol_of(A, md, slot);
  fock_pairs(A, M, M.H, P, X, xf, F, gtid(), gstride());
}
__global__ __launch_bounds__(LG_THREADS) void lg_fock_diag_x_kernel(KArgs A, MolDesc md, const double* P,
                                                                    const double* X, double xf, double* F,
                                                                    const double* J) {
  const Mol M = mol_of(A, md, nullptr);
  fock_diag(A, M, M.H, P, X, xf, F, J, gtid(), gstride());
}

static bool dc_fock(Batch& b, const MolDesc& d, const double* P, const double* X, double xf, double* F,
                    std::string& err) {
  hipLaunchKernelGGL(lg_fock_pairs_x_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, b.args, d, P, X, xf, F,
                     b.lg.slot);
  hipLaunchKernelGGL(lg_fock_jsum_wide_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, b.args, d, P, b.lg.J,
                     b.lg.slot);
  hipLaunchKernelGGL(lg_fock_diag_x_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, b.args, d, P, X, xf, F,
                     (const double*)b.lg.J);
  HIP_OK(hipGetLastError());
  return true;
}

// Status-returning checks for dc_run's own statements (its lambdas return bool and use HIP_OK):
// a HIP or rocBLAS failure there is status 6 with the message in err.
#define DC_HIP(x)                                           \
  do {                                                      \
    hipError_t e_ = (x);                                    \
    if (e_ != hipSuccess) {                                 \
      err = std::string(#x) + ": " + hipGetErrorString(e_); \
      return 6;                                             \
    }                                                       \
  } while (0)
#define DC_BLAS(x)                                                 \
  do {                                                             \
    rocblas_status s_ = (x);                                       \
    if (s_ != rocblas_status_success) {                            \
      err = std::string(#x) + ": " + rocblas_status_to_string(s_); \
      return 6;                                                    \
    }                                                              \
  } while (0)

int dc_run(const Params& P, const MolIn& mol, const EngineOptions& opt, const DcOptionsHost& dco, DcResultHost& out,
           std::string& err, bool want_gradient) {
  out = DcResultHost();
  // electron counts (src/dc/scf.rs:357-398)
  double ne = -mol.charge;
  std::vector<int> norb(mol.n_atoms), off(mol.n_atoms + 1, 0);
  for (int i = 0; i < mol.n_atoms; ++i) {
    const int z = mol.z[i];
    if (z < 0 || z >= 128 || P.slot_of_z[z] < 0) {
      out.missing_z = z;
      return 2;
    }
    const ElemConst& e = P.table.el[P.slot_of_z[z]];
    ne += e.tore;
    norb[i] = e.norb;
    off[i + 1] = off[i] + e.norb;
  }
  const int mult = std::max(mol.multiplicity, 1);
  const double nunp = mult - 1;
  if (ne < nunp || ne < 0.0) {
    err = "electron count " + std::to_string(ne) + " is incompatible with multiplicity " + std::to_string(mult);
    return 1;
  }
  const double na = 0.5 * (ne + nunp), nb = 0.5 * (ne - nunp);
  bool uhf = na != nb;
  if (opt.reference == 2) uhf = true;
  if (opt.reference == 1) {
    if (na != nb) {
      err = "RHF requested for an open-shell system";
      return 1;
    }
    uhf = false;
  }
  dc::Part part;
  if (!dc::partition(mol.n_atoms, mol.xyz, norb, off, dco, part, err)) return 1;
  // the molecule on the host-driven path: pair tables, H_core, the grid-wide Fock kernels
  RunRequest req;
  req.force_large = true;
  req.dc = true;
  req.want_gradient = want_gradient;
  Batch* bp = nullptr;
  if (!batch_create(P, {mol}, opt, req, bp, err)) return 6;
  std::unique_ptr<Batch, void (*)(Batch*)> hold(bp, batch_destroy);
  Batch& b = *bp;
  if (b.pre_status[0] != 0) {
    out.missing_z = b.pre_z[0];
    out.required_mb = b.pre_req[0];
    out.limit_mb = b.pre_lim[0];
    return b.pre_status[0];
  }
  const MolDesc& d = b.desc[0];
  const int n = d.nao, np = d.np;
  const long long m2 = (long long)np * np;
  std::vector<int> eng_of(n);  // caller AO -> engine AO
  for (int e = 0; e < n; ++e) eng_of[b.ao_map[0][e]] = e;
  // pattern (DensityPattern::from_partition, src/dc/pattern.rs:62-95): the orbital pairs some
  // subsystem holds, in the caller's row-major order, as engine dense indices
  const int nsub = (int)part.orb.size();
  int nmax = 0;
  for (const auto& o : part.orb) nmax = std::max(nmax, (int)o.size());
  std::vector<char> live((size_t)n * n, 0);
  for (const auto& o : part.orb)
    for (int mu : o)
      for (int nu : o) live[(size_t)mu * n + nu] = 1;
  std::vector<long long> at;
  for (int mu = 0; mu < n; ++mu)
    for (int nu = 0; nu < n; ++nu)
      if (live[(size_t)mu * n + nu]) at.push_back((long long)eng_of[mu] * np + eng_of[nu]);
  const long long nnz = (long long)at.size();
  {
    std::vector<double> ws((size_t)n * n, 0.0);
    for (int s = 0; s < nsub; ++s) {
      const auto& o = part.orb[s];
      for (size_t i = 0; i < o.size(); ++i)
        for (size_t j = 0; j < o.size(); ++j) {
          const bool ci = part.in_core[s][i], cj = part.in_core[s][j];
          ws[(size_t)o[i] * n + o[j]] += ci && cj ? 1.0 : (ci || cj) ? 0.5 : 0.0;
        }
    }
    for (double v : ws) out.dropped_pairs += v < 0.5;
  }
  out.largest_subsystem = nmax;
  out.n_subsystems = nsub;
  out.nao = n;
  out.unrestricted = uhf ? 1 : 0;
  // size buckets: each subsystem's size rounded up to BUCKET (at most nmax), then consecutive
  // rounded sizes merged where one batch at the larger size costs less than two calls
  // (dc::LATENCY; least total cost over every grouping, by dynamic programming); nb is the
  // bucket's padded size.  The batch order sorts the subsystems by nb (stably), so each bucket is
  // one contiguous run.
  std::vector<int> bsize(nsub), order(nsub), slot_of(nsub);
  for (int s = 0; s < nsub; ++s)
    bsize[s] = std::min(nmax, ((int)part.orb[s].size() + dc::BUCKET - 1) / dc::BUCKET * dc::BUCKET);
  {
    std::vector<int> sizes(bsize);
    std::sort(sizes.begin(), sizes.end());
    std::vector<int> cand, cnt;  // distinct rounded sizes and their subsystem counts
    for (int v : sizes) {
      if (cand.empty() || cand.back() != v) {
        cand.push_back(v);
        cnt.push_back(0);
      }
      ++cnt.back();
    }
    const int m = (int)cand.size();
    std::vector<double> best(m + 1, 0.0);
    std::vector<int> from(m + 1, 0);
    for (int j = 1; j <= m; ++j) {  // best[j]: the first j sizes grouped, the last group ending at j - 1
      const double nb = cand[j - 1];
      best[j] = INFINITY;
      long long c = 0;
      for (int i = j; i >= 1; --i) {
        c += cnt[i - 1];
        const double v = best[i - 1] + dc::LATENCY * nb + (double)c * nb * nb * nb;
        if (v < best[j]) {
          best[j] = v;
          from[j] = i - 1;
        }
      }
    }
    std::vector<int> to(m);  // rounded size -> its group's padded size
    for (int j = m; j > 0; j = from[j])
      for (int i = from[j]; i < j; ++i) to[i] = cand[j - 1];
    for (int s = 0; s < nsub; ++s)
      bsize[s] = to[std::lower_bound(cand.begin(), cand.end(), bsize[s]) - cand.begin()];
  }
  for (int s = 0; s < nsub; ++s) order[s] = s;
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return bsize[a] < bsize[c]; });
  struct Bucket {
    int nb, first, count;
    long long off;  // offset of its first nb x nb block in the batch buffers
  };
  std::vector<Bucket> buckets;
  std::vector<long long> boff(nsub);  // batch offset of each slot's block
  long long per_total = 0;
  for (int t = 0; t < nsub; ++t) {
    const int s = order[t], nb = bsize[s];
    slot_of[s] = t;
    if (buckets.empty() || buckets.back().nb != nb) buckets.push_back({nb, t, 0, per_total});
    ++buckets.back().count;
    boff[t] = per_total;
    per_total += (long long)nb * nb;
  }
  // subsystem tables in batch order (engine AOs, nmax entries per subsystem)
  std::vector<int> horb((size_t)nsub * nmax, 0), hsz(nsub);
  std::vector<char> hcore((size_t)nsub * nmax, 0);
  for (int t = 0; t < nsub; ++t) {
    const int s = order[t];
    hsz[t] = (int)part.orb[s].size();
    for (int i = 0; i < hsz[t]; ++i) {
      horb[(size_t)t * nmax + i] = eng_of[part.orb[s][i]];
      hcore[(size_t)t * nmax + i] = part.in_core[s][i];
    }
  }
  // device buffers
  double *Pa, *Pb, *Pt, *Fa, *Fb, *fa, *fb, *nt, *Fs, *Ws, *Ew, *Qw, *Ys, *Cs, *mu;
  int *dorb, *dsz, *dinfo;
  char* dcore;
  long long* dat;
  const int depth_mb = opt.scf_memory_mb;
  int depth = 8;
  if (depth_mb != 0 && nnz > 0) {
    const long long slot_bytes = 2LL * (2 * nnz) * 8;
    depth = (int)std::min<long long>(8, std::max<long long>(2, (long long)depth_mb * 1024 * 1024 / slot_bytes));
  }
  double *Hd, *Hr;
  if (!b.alloc(Pa, m2, err) || !b.alloc(Pb, m2, err) || !b.alloc(Pt, m2, err) || !b.alloc(Fa, m2, err) ||
      !b.alloc(Fb, m2, err) || !b.alloc(fa, m2, err) || !b.alloc(fb, m2, err) || !b.alloc(nt, m2, err) ||
      !b.alloc(Fs, per_total, err) || !b.alloc(Ws, (long long)nmax * nsub, err) ||
      !b.alloc(Ew, (long long)nmax * nsub, err) || !b.alloc(Qw, (long long)nmax * nsub, err) ||
      !b.alloc(Ys, per_total, err) || !b.alloc(Cs, per_total, err) || !b.alloc(mu, 2, err) ||
      !b.alloc(dorb, (long long)nsub * nmax, err) || !b.alloc(dsz, nsub, err) || !b.alloc(dinfo, nsub, err) ||
      !b.alloc(dcore, (long long)nsub * nmax, err) || !b.alloc(dat, std::max(1LL, nnz), err) ||
      !b.alloc(Hd, 2 * nnz * depth, err) || !b.alloc(Hr, 2 * nnz * depth, err))
    return 6;
  DC_HIP(scopy(b.stream, dorb, horb.data(), sizeof(int) * horb.size(), hipMemcpyHostToDevice));
  DC_HIP(scopy(b.stream, dsz, hsz.data(), sizeof(int) * hsz.size(), hipMemcpyHostToDevice));
  DC_HIP(scopy(b.stream, dcore, hcore.data(), hcore.size(), hipMemcpyHostToDevice));
  DC_HIP(scopy(b.stream, dat, at.data(), sizeof(long long) * at.size(), hipMemcpyHostToDevice));
  // pair tables and H_core
  DC_HIP(hipMemsetAsync(b.args.queue, 0, 3 * sizeof(int), b.stream));
  hipLaunchKernelGGL(pair_kernel, dim3(b.pair_blocks), dim3(b.pair_block), 0, b.stream, b.args);
  DC_HIP(hipGetLastError());
  hipLaunchKernelGGL(lg_core_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, b.args, d, b.lg.red);
  DC_HIP(hipGetLastError());
  double ecore = 0.0;
  if (!large_sum1(b, ecore, err)) return 6;
  const dim3 G(dc::BLOCKS), T(dc::THREADS);
  // starting density: the superposition of atomic densities, each channel its electron share,
  // read onto the pattern (src/dc/scf.rs:431-433)
  const double tot_e = na + nb;
  auto share = [&](double e) { return tot_e > 0.0 ? e / tot_e : 0.5; };
  DC_HIP(hipMemsetAsync(nt, 0, sizeof(double) * m2, b.stream));
  hipLaunchKernelGGL(lg_sad_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, b.args, d, nt, -1);
  DC_HIP(hipMemsetAsync(Pa, 0, sizeof(double) * m2, b.stream));
  DC_HIP(hipMemsetAsync(Pb, 0, sizeof(double) * m2, b.stream));
  hipLaunchKernelGGL(dc::scale_kernel, G, T, 0, b.stream, (const double*)nt, Fa, m2, share(na));
  hipLaunchKernelGGL(dc::keep_pattern_kernel, G, T, 0, b.stream, (const double*)Fa, (const long long*)dat, nnz, Pa);
  hipLaunchKernelGGL(dc::scale_kernel, G, T, 0, b.stream, (const double*)nt, Fa, m2, share(nb));
  hipLaunchKernelGGL(dc::keep_pattern_kernel, G, T, 0, b.stream, (const double*)Fa, (const long long*)dat, nnz, Pb);
  DC_HIP(hipGetLastError());
  const double damping = opt.damping > 0.0 ? std::min(std::max(opt.damping, 0.0), 0.95) : 0.3;
  // Fock matrices of one channel: F = H + J(P_total) - K(P_spin); a restricted run builds the
  // alpha matrix only (the beta channel reads it)
  auto fock = [&](double* F, const double* Pspin) { return dc_fock(b, d, Pt, Pspin, 1.0, F, err); };
  // one stream and solver handle per bucket after the first (default priority, non-blocking)
  if (!b.dc_fork) DC_HIP(hipEventCreateWithFlags(&b.dc_fork, hipEventDisableTiming));
  while (b.dc_lanes.size() + 1 < buckets.size()) {
    b.dc_lanes.emplace_back();
    Batch::DcLane& l = b.dc_lanes.back();
    DC_HIP(hipStreamCreateWithFlags(&l.s, hipStreamNonBlocking));
    DC_HIP(hipEventCreateWithFlags(&l.done, hipEventDisableTiming));
    DC_BLAS(rocblas_create_handle(&l.h));
    DC_BLAS(rocblas_set_stream(l.h, l.s));
  }
  const int nlanes = (int)buckets.size() - 1;
  auto lane_stream = [&](int k) { return k == 0 ? b.stream : b.dc_lanes[k - 1].s; };
  auto lane_blas = [&](int k) { return k == 0 ? b.blas : b.dc_lanes[k - 1].h; };
  auto fork = [&]() -> bool {  // the lanes start after the work queued so far on the batch stream
    HIP_OK(hipEventRecord(b.dc_fork, b.stream));
    for (int k = 0; k < nlanes; ++k) HIP_OK(hipStreamWaitEvent(b.dc_lanes[k].s, b.dc_fork, 0));
    return true;
  };
  auto join = [&]() -> bool {  // the batch stream continues after every lane's work
    for (int k = 0; k < nlanes; ++k) HIP_OK(hipEventRecord(b.dc_lanes[k].done, b.dc_lanes[k].s));
    for (int k = 0; k < nlanes; ++k) HIP_OK(hipStreamWaitEvent(b.stream, b.dc_lanes[k].done, 0));
    return true;
  };
  // one spin channel: subsystem eigenpairs, the chemical potential, the assembled density
  auto solve_channel = [&](const double* F, double target, double* dst, double& mu_out) -> bool {
    hipLaunchKernelGGL(lg_gershgorin_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, F, n, np, b.lg.red);
    HIP_OK(hipGetLastError());
    HIP_OK(scopy(b.stream, b.h_red.data(), b.lg.red, sizeof(double) * (size_t)LG_BLOCKS * LG_MAXDOT,
                 hipMemcpyDeviceToHost));
    double ghi = -1.0e300;
    for (int blk = 0; blk < LG_BLOCKS; ++blk) ghi = std::max(ghi, b.h_red[(size_t)blk * LG_MAXDOT + 1]);
    if (!fork()) return false;
    for (int bk = 0; bk < (int)buckets.size(); ++bk) {
      const Bucket& k = buckets[bk];
      const long long kper = (long long)k.nb * k.nb;
      hipLaunchKernelGGL(dc::gather_kernel, G, T, 0, lane_stream(bk), F, np,
                         (const int*)(dorb + (long long)k.first * nmax), (const int*)(dsz + k.first), k.count, k.nb,
                         nmax, ghi + 1.0, Fs + k.off);
      HIP_OK(hipGetLastError());
      BLAS_OK(rocsolver_dsyevd_strided_batched(lane_blas(bk), rocblas_evect_original, rocblas_fill_upper, k.nb,
                                               Fs + k.off, k.nb, kper, Ws + (long long)k.first * nmax, nmax,
                                               Ew + (long long)k.first * nmax, nmax, dinfo + k.first, k.count));
      hipLaunchKernelGGL(dc::core_weight_kernel, G, T, 0, lane_stream(bk), (const double*)(Fs + k.off),
                         (const int*)(dsz + k.first), (const char*)(dcore + (long long)k.first * nmax), k.count, k.nb,
                         nmax, Qw + (long long)k.first * nmax);
      HIP_OK(hipGetLastError());
    }
    if (!join()) return false;
    std::vector<int> info(nsub);
    HIP_OK(scopy(b.stream, info.data(), dinfo, sizeof(int) * nsub, hipMemcpyDeviceToHost));
    for (int t = 0; t < nsub; ++t)
      if (info[t] != 0) {
        err = "rocsolver_dsyevd_strided_batched did not converge on subsystem " + std::to_string(order[t]);
        return false;
      }
    hipLaunchKernelGGL(dc::mu_kernel, dim3(1), dim3(dc::MU_THREADS), 0, b.stream, (const double*)Ws, (const double*)Qw,
                       (const int*)dsz, nsub, nmax, target, dco.smearing_ev, dco.electron_tol, mu);
    HIP_OK(hipGetLastError());
    // C_s = Y_s V_s^T (column-major nb x nb per subsystem, the k sum over every column)
    const double one = 1.0, zero = 0.0;
    if (!fork()) return false;
    for (int bk = 0; bk < (int)buckets.size(); ++bk) {
      const Bucket& k = buckets[bk];
      const long long kper = (long long)k.nb * k.nb;
      hipLaunchKernelGGL(dc::occupy_kernel, G, T, 0, lane_stream(bk), (const double*)(Fs + k.off),
                         (const double*)(Ws + (long long)k.first * nmax), (const int*)(dsz + k.first), k.count, k.nb,
                         nmax, (const double*)mu, dco.smearing_ev, Ys + k.off);
      HIP_OK(hipGetLastError());
      BLAS_OK(rocblas_dgemm_strided_batched(lane_blas(bk), rocblas_operation_none, rocblas_operation_transpose, k.nb,
                                            k.nb, k.nb, &one, Ys + k.off, k.nb, kper, Fs + k.off, k.nb, kper, &zero,
                                            Cs + k.off, k.nb, kper, k.count));
    }
    if (!join()) return false;
    HIP_OK(hipMemsetAsync(dst, 0, sizeof(double) * m2, b.stream));
    for (int s = 0; s < nsub; ++s) {  // the partition's order
      const int t = slot_of[s];
      const long long nn = (long long)hsz[t] * hsz[t];
      const int blocks = (int)std::min<long long>(dc::BLOCKS, (nn + dc::THREADS - 1) / dc::THREADS);
      hipLaunchKernelGGL(dc::assemble_kernel, dim3(std::max(1, blocks)), T, 0, b.stream, (const double*)(Cs + boff[t]),
                         bsize[s], hsz[t], (const int*)(dorb + (long long)t * nmax),
                         (const char*)(dcore + (long long)t * nmax), dst, np);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(scopy(b.stream, &mu_out, mu, sizeof(double), hipMemcpyDeviceToHost));
    return true;
  };
  auto dots = [&](std::initializer_list<std::pair<const double*, const double*>> xy, double* res) {
    Dots D{};
    D.k = 0;
    for (const auto& p : xy) {
      D.x[D.k] = p.first;
      D.y[D.k] = p.second;
      ++D.k;
    }
    return large_dots(b, D, m2, res, err);
  };
  std::vector<std::vector<double>> gram;  // <r_i, r_j> of the kept history
  std::vector<int> slots;                 // history slot of each kept entry, oldest first
  double last = INFINITY, mu_a = 0.0, mu_b = 0.0;
  bool converged = false;
  int iterations = 0;
  for (int it = 1; it <= opt.max_scf; ++it) {
    iterations = it;
    hipLaunchKernelGGL(dc::add_kernel, G, T, 0, b.stream, (const double*)Pa, (const double*)Pb, Pt, m2);
    HIP_OK(hipGetLastError());
    if (!fock(Fa, Pa)) return 6;
    if (uhf && !fock(Fb, Pb)) return 6;
    if (!solve_channel(Fa, na, fa, mu_a)) return 6;
    if (uhf) {
      if (!solve_channel(Fb, nb, fb, mu_b)) return 6;
    } else {
      mu_b = mu_a;
    }
    const double* Fbv = uhf ? Fb : Fa;
    const double* fbv = uhf ? fb : fa;
    hipLaunchKernelGGL(dc::add_kernel, G, T, 0, b.stream, (const double*)fa, fbv, nt, m2);
    HIP_OK(hipGetLastError());
    double e3[3];
    if (!dots({{nt, Fa}, {fa, Fa}, {fbv, Fbv}}, e3)) return 6;
    const double electronic = 0.5 * (e3[0] + e3[1] + e3[2]);
    hipLaunchKernelGGL(lg_diff2_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, (const double*)nt,
                       (const double*)Pt, m2, b.lg.red);
    HIP_OK(hipGetLastError());
    double sq = 0.0;
    if (!large_sum1(b, sq, err)) return 6;
    const double change = std::sqrt(sq / (double)std::max(n * n, 1));
    hipLaunchKernelGGL(dc::damp_kernel, G, T, 0, b.stream, Pa, (const double*)fa, m2, damping);
    hipLaunchKernelGGL(dc::damp_kernel, G, T, 0, b.stream, Pb, fbv, m2, damping);
    HIP_OK(hipGetLastError());
    // the density DIIS (push, then the longest suffix that solves with sum |c| <= 20)
    int ns = (int)slots.size();
    if (ns == depth) {
      const int s0 = slots[0];
      slots.erase(slots.begin());
      gram.erase(gram.begin());
      for (auto& r : gram) r.erase(r.begin());
      slots.push_back(s0);
    } else {
      slots.push_back(ns);
    }
    ns = (int)slots.size();
    const int sn = slots.back();
    double* cur = Hd + 2 * nnz * (long long)sn;
    double* res = Hr + 2 * nnz * (long long)sn;
    hipLaunchKernelGGL(dc::pack_kernel, G, T, 0, b.stream, (const double*)Pa, (const double*)Pb, (const double*)fa,
                       fbv, (const long long*)dat, nnz, cur, res);
    HIP_OK(hipGetLastError());
    {
      Dots D{};
      D.k = ns;
      for (int i = 0; i < ns; ++i) {
        D.x[i] = Hr + 2 * nnz * (long long)slots[i];
        D.y[i] = res;
      }
      std::vector<double> row(ns);
      if (!large_dots(b, D, 2 * nnz, row.data(), err)) return 6;
      for (int i = 0; i + 1 < ns; ++i) gram[i].push_back(row[i]);
      gram.push_back(row);
    }
    for (int first = 0; first + 1 < ns; ++first) {
      Diis Dm{};
      Dm.len = ns - first;
      for (int i = first; i < ns; ++i)
        for (int j = first; j < ns; ++j) Dm.b[i - first][j - first] = gram[i][j];
      double c[MAXD];
      if (!cdiis_coeffs(Dm, c)) continue;
      double wsum = 0.0;
      for (int i = 0; i < Dm.len; ++i) wsum += std::fabs(c[i]);
      if (!std::isfinite(wsum) || wsum > 20.0) continue;
      dc::Mix mx{};
      mx.k = Dm.len;
      for (int i = 0; i < Dm.len; ++i) {
        mx.v[i] = Hd + 2 * nnz * (long long)slots[first + i];
        mx.c[i] = c[i];
      }
      hipLaunchKernelGGL(dc::extrapolate_kernel, G, T, 0, b.stream, mx, (const long long*)dat, nnz, Pa, Pb);
      HIP_OK(hipGetLastError());
      break;
    }
    if (std::fabs(electronic - last) < opt.e_tol && change < opt.p_tol) {
      converged = true;
      break;
    }
    last = electronic;
  }
  out.iterations = iterations;
  out.converged = converged ? 1 : 0;
  if (!converged) return 3;
  // the Fock matrices at the converged density and the energy there (src/dc/scf.rs:554-572, :110-122)
  hipLaunchKernelGGL(dc::add_kernel, G, T, 0, b.stream, (const double*)Pa, (const double*)Pb, Pt, m2);
  DC_HIP(hipGetLastError());
  if (!fock(Fa, Pa)) return 6;
  if (uhf && !fock(Fb, Pb)) return 6;
  const Mol M = mol_of(b.args, d, nullptr);
  double e3[3];
  if (!dots({{Pt, M.H}, {Pa, Fa}, {Pb, uhf ? Fb : Fa}}, e3)) return 6;
  out.electronic_ev = 0.5 * (e3[0] + e3[1] + e3[2]);
  out.core_ev = ecore;
  out.total_ev = out.electronic_ev + ecore;
  double eis = 0.0, ehs = 0.0;
  for (int i = 0; i < mol.n_atoms; ++i) {
    const int sl = P.slot_of_z[mol.z[i]];
    eis += P.host[sl].e_isol;
    ehs += P.host[sl].eheat_ev;
  }
  out.hof_kcal = (out.total_ev - eis + ehs) * EV_TO_KCAL;
  out.fermi_ev = mu_a;
  out.fermi_beta_ev = mu_b;
  // densities back in the caller's AO order; charges from the total diagonal
  std::vector<double> ha(m2), hb(m2);
  DC_HIP(scopy(b.stream, ha.data(), Pa, sizeof(double) * m2, hipMemcpyDeviceToHost));
  DC_HIP(scopy(b.stream, hb.data(), Pb, sizeof(double) * m2, hipMemcpyDeviceToHost));
  out.density.assign((size_t)n * n, 0.0);
  out.spin_density.assign((size_t)n * n, 0.0);
  for (int mu2 = 0; mu2 < n; ++mu2)
    for (int nu = 0; nu < n; ++nu) {
      const size_t e = (size_t)eng_of[mu2] * np + eng_of[nu];
      out.density[(size_t)mu2 * n + nu] = ha[e] + hb[e];
      if (uhf) out.spin_density[(size_t)mu2 * n + nu] = ha[e] - hb[e];
    }
  out.charges.assign(mol.n_atoms, 0.0);
  for (int i = 0; i < mol.n_atoms; ++i) {
    double pop = 0.0;
    for (int k = 0; k < norb[i]; ++k) pop += out.density[(size_t)(off[i] + k) * n + off[i] + k];
    out.charges[i] = P.table.el[P.slot_of_z[mol.z[i]]].tore - pop;
  }
  if (want_gradient) {
    // dc_gradient (src/dc/derivatives.rs:55-121): the fixed-density gradient at the partitioned
    // density (the spin form for an unrestricted result) by the batched path's pair-gradient
    // kernels, which read the density (and spin density) from the output slots, minus f q_A in a
    // field
    hipLaunchKernelGGL(lg_pack_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, (const double*)Pt,
                       b.args.out_P + d.out_mat, n, np);
    if (uhf) {
      hipLaunchKernelGGL(dc::sub_kernel, G, T, 0, b.stream, (const double*)Pa, (const double*)Pb, nt, m2);
      hipLaunchKernelGGL(lg_pack_kernel, dim3(LG_BLOCKS), dim3(LG_THREADS), 0, b.stream, (const double*)nt,
                         b.args.out_S + d.out_mat, n, np);
    }
    hipLaunchKernelGGL(lg_epilogue_kernel, dim3(1), dim3(NT), 0, b.stream, b.args, d, (const double*)Pt, b.lg.dip);
    hipLaunchKernelGGL(grad_pair_kernel, dim3(b.pair_blocks), dim3(b.pair_block), 0, b.stream, b.args);
    hipLaunchKernelGGL(grad_gather_wide_kernel, dim3(1, (unsigned)std::min(256, (d.nat + 3) / 4)), dim3(256), 0,
                       b.stream, b.args, 0);
    HIP_OK(hipGetLastError());
    std::vector<double> hg(3 * (size_t)d.nat);
    HIP_OK(scopy(b.stream, hg.data(), b.args.out_g + 3LL * d.atom0, sizeof(double) * hg.size(), hipMemcpyDeviceToHost));
    out.gradient.assign(3 * (size_t)mol.n_atoms, 0.0);
    for (int t = 0; t < d.nat; ++t)
      for (int c = 0; c < 3; ++c) out.gradient[3 * (size_t)b.atom_map[0][t] + c] = hg[3 * (size_t)t + c];
  }
  return 0;
}
