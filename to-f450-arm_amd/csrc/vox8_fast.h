// vox8 — structured (Hadamard-basis) stencil: coefficient block and launcher.
#pragma once
#include "vox8_kernels.h"

namespace vox8 {

// PAT_ORTHO: orthotropic non-zero pattern (45 entries); PAT_BLOCK: general
// blocks; PAT_CUBE: the orthotropic pattern with the equalities of an isotropic
// material on a cubic voxel (8 distinct stiffness and 4 distinct mass values).
enum Pattern : int { PAT_DENSE = -1, PAT_ORTHO = 0, PAT_BLOCK = 1, PAT_CUBE = 2 };

// Element operator in the monomial basis, scaled by 1/64 (see vox8_stencil_fast.hip):
// lin: {xi, eta, zeta} x {x,y,z}, bil: {xi eta, eta zeta, xi zeta} x {x,y,z},
// tri: {xi eta zeta} x {x,y,z}; md: mass diagonal per monomial (same for x,y,z).
template <typename T>
struct FastCoef {
  T lin[9][9];
  T bil[9][9];
  T tri[3][3];
  T md[8];
};

// Index pairs (row, col) of the orthotropic non-zero pattern.
constexpr int ORTHO_LIN[21][2] = {{0, 0}, {0, 4}, {0, 8}, {1, 1}, {1, 3}, {2, 2}, {2, 6},
                                  {3, 1}, {3, 3}, {4, 0}, {4, 4}, {4, 8}, {5, 5}, {5, 7},
                                  {6, 2}, {6, 6}, {7, 5}, {7, 7}, {8, 0}, {8, 4}, {8, 8}};
constexpr int ORTHO_BIL[21][2] = {{0, 0}, {0, 5}, {1, 1}, {1, 8}, {2, 2}, {2, 3}, {2, 7},
                                  {3, 2}, {3, 3}, {3, 7}, {4, 4}, {4, 6}, {5, 0}, {5, 5},
                                  {6, 4}, {6, 6}, {7, 2}, {7, 3}, {7, 7}, {8, 1}, {8, 8}};

// Classes of the 21 lin / 21 bil / 3 tri orthotropic entries (ORTHO_LIN,
// ORTHO_BIL order, then tri diagonal) under cubic isotropy.  The structured
// kernel reads the class values from lin[0][0], lin[0][4], lin[1][1],
// bil[0][0], bil[0][5], bil[2][2], bil[2][3], tri[0][0].
constexpr int CUBE_CLASS[45] = {0, 1, 1, 2, 2, 2, 2, 2, 2, 1, 0, 1, 2, 2, 2, 2, 2, 2, 1, 1, 0,
                                3, 4, 3, 4, 5, 6, 6, 6, 5, 6, 3, 4, 4, 3, 4, 3, 6, 6, 5, 4, 3,
                                7, 7, 7};

// Classify Ke/Me (24x24 row-major, reference local node order) and derive the
// coefficients.  Returns PAT_CUBE, PAT_ORTHO, PAT_BLOCK or PAT_DENSE.
int fast_coefficients(const double* Ke, const double* Me, FastCoef<double>* c64,
                      FastCoef<float>* c32);

// EPI_CHEB0: the first Chebyshev step from a zero iterate, x = 0, r = b:
// the input vector is b and the direction d = c0 D^-1 b is formed on load.
// EPI_X0 / EPI_X: Chebyshev steps in the three-term form on the iterates
// alone (no residual or direction vectors),
//   x_{k+1} = x_k + c1 (x_k - x_{k-1}) + c2 D^-1 (b - A x_k),
// with the input vector x_k.  EPI_X0 is the step from x_1 = c0 D^-1 b with
// x_0 = 0: the input is b and x_1 is formed on load.  EPI_X takes x_{k-1}
// from `xprev` (prev 1), as c0 D^-1 b (prev 2: the step after EPI_X0), or
// drops the c1 term (prev 0: the first step from a non-zero iterate).
// EPI_LOBRES (FP64 K and M products, LOBPCG): besides K x and M x, the
// residual r = K x - lam_col M x of each column: W = D^-1 r (r without dinv),
// zero for the columns flagged in conv, and per-workgroup partial sums of
// |r|^2 (partials[col * tiles + tile]).
enum EpiMode : int { EPI_PLAIN = 0, EPI_RESID = 1, EPI_CHEB = 2, EPI_CHEB0 = 3, EPI_X0 = 4,
                     EPI_X = 5, EPI_LOBRES = 6 };

// Node-local epilogue of a K product (multigrid smoothing), vectors with ld.
template <typename T>
struct Epilogue {
  int mode = EPI_PLAIN;
  const T* b = nullptr;      // RESID, X: right-hand side
  T* r = nullptr;            // RESID: out; CHEB: in/out residual
  T* xs = nullptr;           // CHEB: iterate (+= d); X0, X: out, x_{k+1}
  T* dnew = nullptr;         // out: next direction (RESID: may be null)
  const T* dinv = nullptr;   // Jacobi inverse diagonal (one vector, no ld)
  T c0 = 0, c1 = 0, c2 = 0;
  int64_t ld = 0;
  const int* skip = nullptr;  // per column: non-zero = leave this column untouched
  const T* xprev = nullptr;  // EPI_X, prev 1: x_{k-1}
  int prev = 0;              // EPI_X: where x_{k-1} comes from (see EpiMode)
  double* out64 = nullptr;   // EPI_X (FP32 structured stencil): x_{k+1} stored as FP64 here
                             // instead of into xs
  const double* lam = nullptr;   // EPI_LOBRES: per-column eigenvalue estimate
  const int* conv = nullptr;     //   per-column lock flags (W written as zero), may be null
  double* w = nullptr;           //   W (ld), the residual's output
  float* w32 = nullptr;          //   instead of w: W in FP32 (ld), a multigrid cycle's input
  double* partials = nullptr;    //   |r|^2 per (column, workgroup)
  double* res = nullptr;         //   |r|^2 per column (summed in a fixed order after the launch)
  int64_t partials_cap = 0;      //   doubles in partials
};


// Element-tile rows (8 or 12) and z chunk (node planes per workgroup).
struct StencilShape {
  int ty = 8, zc = 1;
};

// Autotune the FP32 launch shape for a grid and column count: times a
// Chebyshev smoothing step (three-term form: x_k in d, x_{k-1} in r, b in
// xs, x_{k+1} written to dnew; ncols columns, leading dimension ld; pass work
// vectors) of each candidate shape and keeps the choice for the process:
// every FP32 launch on such a grid uses it.  Synchronises the stream; not for
// use during graph capture.  VOX8_FP32_SHAPE="rows/chunk" fixes the shape
// instead of timing (profiler runs).
int stencil_tune_fp32(const Grid& g, int pattern, const FastCoef<float>& C, const float* fK,
                      const uint8_t* fixed, const float* dinv, float* d, float* r, float* xs,
                      float* dnew, int64_t ld, int ncols, hipStream_t st);
bool stencil_tuned_fp32(const Grid& g, int ncols);
// Most workgroups one column of a structured-stencil launch can have (any
// tile rows and z chunk): the size of per-workgroup partial sums.
int64_t stencil_max_blocks(const Grid& g);

template <typename T>
int launch_apply_fast(const Grid& g, int pattern, int op, const FastCoef<T>& C, const T* fK,
                      const T* fM, const uint8_t* fixed, const T* x, int64_t ldx, T* yK, T* yM,
                      int64_t ldy, int nrhs, hipStream_t st, DotFuse fuse = DotFuse(),
                      Epilogue<T> ep = Epilogue<T>());

}  // namespace vox8
