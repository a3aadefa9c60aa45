// vox8 — solver-side interfaces (batched PCG, LOBPCG) used by the C-ABI layer.
#pragma once
#include <functional>

#include "vox8_kernels.h"

namespace vox8 {

// Operator bundle a solver runs on: y = A x for `nrhs` column-major vectors.
template <typename T>
struct PcgOps {
  const Grid* g = nullptr;
  const uint8_t* fixed = nullptr;   // per-node fixed-DOF bitmask
  const T* dinv = nullptr;          // Jacobi inverse diagonal (0 at fixed DOFs)
  std::function<int(const T* x, int64_t ldx, T* y, int64_t ldy, int nrhs, hipStream_t st,
                    DotFuse fuse)>
      apply;
};

template <typename T>
struct PcgWork {
  T* r = nullptr;
  T* p = nullptr;
  T* q = nullptr;
  T* x = nullptr;                 // only used as the correction of the mixed solve
  double* rhs64 = nullptr;        // FP64 right-hand side of the inner solve (+nrhs scales)
  int* skip = nullptr;            // device [ncols]: column finished (multigrid skips it)
  PcgScalars* scal = nullptr;     // device [ncols]
  PcgScalars* host_scal = nullptr;  // pinned [ncols]
  double* host_scale = nullptr;     // pinned [ncols]
  double* partials = nullptr;
  unsigned int* counters = nullptr;
  int ncols = 0;
  int64_t n = 0;
};

// Preconditioner z = M r for nrhs column-major vectors (ld = n_dofs); r and z
// may alias.  skip (device, per column, may be null): non-zero columns are
// left untouched.
template <typename T>
using Precond = std::function<int(const T* r, T* z, int nrhs, hipStream_t st, const int* skip)>;

// The LOBPCG residual formed with the K and M products of X: W = D^-1 r
// (dinv null: W = r), r = K x_j - lam_j M x_j, zero for the columns flagged
// in conv; |r_j|^2 into res[j].
struct ResidualFuse {
  const double* lam;
  const int* conv;
  const double* dinv;
  double* w;               // nrhs columns, ld n_dofs
  float* w32;              // instead of w (may be null): W in FP32, ld n_dofs
  double* partials;        // partials_cap doubles
  int64_t partials_cap;
  double* res;
};

// K and M products of nrhs columns; skip as above (those outputs untouched).
// rf (may be null): also form that residual if the product can (*fused is
// then true; otherwise the caller forms it).
using ApplyKM = std::function<int(const double* x, double* yK, double* yM, int nrhs, hipStream_t st,
                                  const int* skip, const ResidualFuse* rf, bool* fused)>;

struct PcgResult {
  int iters;      // CG iterations (summed over inner solves for the mixed path)
  int info;       // 0 converged, >0 not converged (scipy convention)
  double relres;  // final ||r|| / ||b|| (FP64 residual)
  int outer;      // refinement steps (mixed path), 1 for the plain path
};

// Batched PCG; M == nullptr: Jacobi (A.dinv) fused into the vector kernels.
// M32 (with M, may be null): where M reads its input in FP32 when called
// with r == nullptr; the residual update writes it there as well.
int pcg_fp64(const PcgOps<double>& A, const double* F, int64_t ldf, double* X, bool x0_given,
             int nrhs, double rtol, int maxiter, PcgWork<double>& w, hipStream_t st,
             PcgResult* res, const Precond<double>* M = nullptr, float* M32 = nullptr);
int convert_to_float(const double* a, float* b, int64_t n, hipStream_t st);

// ---- LOBPCG ----
struct LobpcgResult {
  int iters;
  int info;            // 0 converged, >0 not converged (iterations done), <0 failure
  double max_res;      // final max absolute residual ||K x - lam M x||
};

struct LobpcgWork {
  double* S = nullptr;    // [3k][n]  X | W | P
  double* AS = nullptr;   // [3k][n]
  double* BS = nullptr;   // [3k][n]
  double* G = nullptr;    // device: GA (m*m) | GB (m*m) | C (m*k) | lam (k) | status
  double* gram_partials = nullptr;
  double* res = nullptr;      // device residual norms^2 [k]
  double* host = nullptr;     // pinned mirror of lam|res|status
  double* res_partials = nullptr;
  int64_t res_partials_cap = 0;   // doubles in res_partials
  unsigned int* counters = nullptr;
  int* conv = nullptr;        // device: sticky convergence flags [12] (soft locking)
  double* guard = nullptr;    // last solve's guard vectors [gmax][n] (warm start of the next)
  int guard_cols = 0;         // columns held in `guard` (0 = none)
  int guard_k = 0;            // wanted pairs of the solve that left them
  double ritz[12] = {};       // Ritz values of the last solve's block (wanted + guard), ascending
  int ritz_n = 0;
  int kmax = 0;
  int64_t n = 0;
  int gram_chunks = 0;
};

constexpr int LOBPCG_GUARD_MAX = 2;

// k wanted pairs, iterated with `guard` extra columns (k + guard <= 12); only
// the k lowest are tested for convergence and returned.
// M32 (with M, may be null): where M reads its input in FP32 when called
// with r == nullptr; the residual is then written there directly.
int lobpcg_fp64(const PcgOps<double>& A, const ApplyKM& applyKM,
                int k, int guard, double tol, int maxiter, const double* X0, double* Xout,
                double* lam_out, LobpcgWork& w, hipStream_t st, LobpcgResult* res, uint64_t seed,
                const Precond<double>* M = nullptr, float* M32 = nullptr);

}  // namespace vox8
