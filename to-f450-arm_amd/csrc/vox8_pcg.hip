// vox8 — batched Jacobi-preconditioned CG on the matrix-free operator.
//
// Replaces scipy.sparse.linalg.cg as called by the reference
// (src/fem.py:543-547; GPU variant src/fem_gpu.py:61-81).  Semantics follow
// SciPy's cg (scipy/sparse/linalg/_isolve/iterative.py, cg):
//   atol = rtol*||b||; before each iteration stop if ||r|| < atol;
//   z = D^-1 r; rho = r.z; p = z + (rho/rho_prev) p; q = A p; alpha = rho/(p.q);
//   x += alpha p; r -= alpha q.  Not converged after maxiter -> info = maxiter.
// Each right-hand side is an independent CG (identical to the reference's
// one-solve-per-load-case loop, src/topopt.py:287-290), but all columns share
// every launch: the element factors and stencil are read once per iteration
// for all load cases.  All inner products and norms accumulate in FP64.
#include "vox8_kernels.h"
#include "vox8_pcg.h"

namespace vox8 {

// out[j] = mask(in[j]) (zero at fixed DOFs), converting precision.
template <typename Tin, typename Tout>
__global__ void k_mask_copy(int64_t n_dofs, const uint8_t* __restrict__ fixed,
                            const Tin* __restrict__ in, int64_t ldi, Tout* __restrict__ out,
                            int64_t ldo) {
  const int col = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_dofs;
       i += (int64_t)gridDim.x * blockDim.x) {
    const bool fx = (fixed[i / 3] >> (i % 3)) & 1;
    out[col * ldo + i] = fx ? Tout(0) : (in ? (Tout)in[col * ldi + i] : Tout(0));
  }
}

// r = mask(F) - (have_q ? q : 0); reduce ||F_f||^2, ||r||^2, r.D^-1 r and set
// up the column state.  F is FP64 (the caller's load), r is T.
template <typename T>
__global__ void __launch_bounds__(256)
k_pcg_init(int64_t n_dofs, const uint8_t* __restrict__ fixed, const double* __restrict__ F,
           int64_t ldf, const T* __restrict__ q, int64_t ldq, const T* __restrict__ dinv,
           T* __restrict__ r, int64_t ldr, double rtol, PcgScalars* __restrict__ scal,
           double* __restrict__ partials, unsigned int* __restrict__ counters) {
  __shared__ double sh[8];
  const int col = blockIdx.y;
  double bb = 0.0, rr = 0.0, rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_dofs;
       i += (int64_t)gridDim.x * blockDim.x) {
    const bool fx = (fixed[i / 3] >> (i % 3)) & 1;
    const double f = fx ? 0.0 : F[col * ldf + i];
    const T ri = fx ? T(0) : (T)(f - (q ? (double)q[col * ldq + i] : 0.0));
    r[col * ldr + i] = ri;
    bb += f * f;
    rr += (double)ri * (double)ri;
    rz += (double)ri * (double)dinv[i] * (double)ri;
  }
  double v[3] = {block_sum(bb, sh), block_sum(rr, sh), block_sum(rz, sh)}, tot[3];
  if (grid_reduce_last<3>(v, partials + (int64_t)col * gridDim.x * 3, counters + col, gridDim.x,
                          blockIdx.x, tot, sh)) {
    if (threadIdx.x == 0) {
      PcgScalars s;
      s.bnorm = sqrt(tot[0]);
      s.atol2 = rtol * rtol * tot[0];
      s.rr = tot[1];
      s.rz = tot[2];
      s.rz_prev = 0.0;
      s.pq = 0.0;
      s.iters = 0;
      s.done = (tot[0] == 0.0) ? 1 : 0;
      s.pad0 = s.pad1 = 0;
      scal[col] = s;
    }
  }
}

// Start of an iteration: stop test, then p = D^-1 r + beta p.
template <typename T>
__global__ void k_pcg_dir(int64_t n_dofs, const T* __restrict__ dinv, const T* __restrict__ r,
                          T* __restrict__ p, int64_t ld, PcgScalars* __restrict__ scal,
                          int maxiter) {
  const int col = blockIdx.y;
  const PcgScalars s = scal[col];
  if (s.done || s.iters >= maxiter) return;
  if (s.rr < s.atol2) {
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[col].done = 1;
    return;
  }
  const double beta = s.iters > 0 ? s.rz / s.rz_prev : 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_dofs;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double z = (double)dinv[i] * (double)r[col * ld + i];
    p[col * ld + i] = (T)(s.iters > 0 ? z + beta * (double)p[col * ld + i] : z);
  }
}

// End of an iteration: alpha = rho/(p.q); x += alpha p; r -= alpha q; new rho, ||r||^2.
template <typename T>
__global__ void __launch_bounds__(256)
k_pcg_step(int64_t n_dofs, const T* __restrict__ dinv, T* __restrict__ x,
           T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ q, int64_t ld,
           PcgScalars* __restrict__ scal, int maxiter, double* __restrict__ partials,
           unsigned int* __restrict__ counters) {
  __shared__ double sh[8];
  const int col = blockIdx.y;
  const PcgScalars s = scal[col];
  if (s.done || s.iters >= maxiter) return;
  const double alpha = s.rz / s.pq;
  double rr = 0.0, rz = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_dofs;
       i += (int64_t)gridDim.x * blockDim.x) {
    const int64_t j = col * ld + i;
    x[j] = (T)((double)x[j] + alpha * (double)p[j]);
    const T rn = (T)((double)r[j] - alpha * (double)q[j]);
    r[j] = rn;
    rr += (double)rn * (double)rn;
    rz += (double)rn * (double)dinv[i] * (double)rn;
  }
  double v[2] = {block_sum(rr, sh), block_sum(rz, sh)}, tot[2];
  if (grid_reduce_last<2>(v, partials + (int64_t)col * gridDim.x * 2, counters + col, gridDim.x,
                          blockIdx.x, tot, sh)) {
    if (threadIdx.x == 0) {
      scal[col].rr = tot[0];
      scal[col].rz_prev = s.rz;
      scal[col].rz = tot[1];
      scal[col].iters = s.iters + 1;
    }
  }
}

// x64 += mask(d) (iterative-refinement correction), per column.
template <typename T>
__global__ void k_add_correction(int64_t n_dofs, const T* __restrict__ d, int64_t ldd,
                                 double* __restrict__ x, int64_t ldx) {
  const int col = blockIdx.y;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_dofs;
       i += (int64_t)gridDim.x * blockDim.x)
    x[col * ldx + i] += (double)d[col * ldd + i];
}

__global__ void k_to_float(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// ---------------------------------------------------------------------------
static inline unsigned vec_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (unsigned)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

int pcg_reduce_capacity_blocks() { return 1024; }

template <typename T>
static int run_cg(const PcgOps<T>& A, const double* F, int64_t ldf, T* x, const T* x0_or_null,
                  bool x0_given, int nrhs, double rtol, int maxiter, PcgWork<T>& w,
                  hipStream_t st, int chunk0) {
  const Grid& g = *A.g;
  const int64_t n = g.n_dofs;
  const unsigned nb = vec_blocks(n);
  const dim3 grid(nb, (unsigned)nrhs), blk(256);
  // x = mask(x0) or 0
  hipLaunchKernelGGL((k_mask_copy<T, T>), grid, blk, 0, st, n, A.fixed,
                     x0_given ? x0_or_null : (const T*)nullptr, n, x, n);
  VOX8_HIP(hipGetLastError());
  if (x0_given) VOX8_TRY(A.apply(x, n, w.q, n, nrhs, st, DotFuse()));
  hipLaunchKernelGGL((k_pcg_init<T>), grid, blk, 0, st, n, A.fixed, F, ldf,
                     x0_given ? (const T*)w.q : (const T*)nullptr, n, A.dinv, w.r, n, rtol,
                     w.scal, w.partials, w.counters);
  VOX8_HIP(hipGetLastError());
  DotFuse fuse;
  fuse.scal = w.scal;
  fuse.partials = w.partials;
  fuse.counters = w.counters;
  int launched = 0, chunk = chunk0;
  while (launched < maxiter) {
    const int todo = chunk < maxiter - launched ? chunk : maxiter - launched;
    for (int it = 0; it < todo; ++it) {
      hipLaunchKernelGGL((k_pcg_dir<T>), grid, blk, 0, st, n, A.dinv, w.r, w.p, n, w.scal, maxiter);
      VOX8_TRY(A.apply(w.p, n, w.q, n, nrhs, st, fuse));
      hipLaunchKernelGGL((k_pcg_step<T>), grid, blk, 0, st, n, A.dinv, x, w.r, w.p, w.q, n,
                         w.scal, maxiter, w.partials, w.counters);
    }
    VOX8_HIP(hipGetLastError());
    launched += todo;
    VOX8_HIP(hipMemcpyAsync(w.host_scal, w.scal, sizeof(PcgScalars) * nrhs, hipMemcpyDeviceToHost, st));
    VOX8_HIP(hipStreamSynchronize(st));
    bool all = true;
    for (int j = 0; j < nrhs; ++j)
      if (!w.host_scal[j].done && w.host_scal[j].iters < maxiter) all = false;
    if (all) break;
    chunk = chunk * 2 < 256 ? chunk * 2 : 256;
  }
  // As in SciPy, a column that reaches maxiter is reported unconverged even if
  // its last step met the tolerance (the stop test runs before an iteration).
  return OK;
}

int pcg_fp64(const PcgOps<double>& A, const double* F, int64_t ldf, double* X, bool x0_given,
             int nrhs, double rtol, int maxiter, PcgWork<double>& w, hipStream_t st,
             PcgResult* res) {
  VOX8_TRY(run_cg<double>(A, F, ldf, X, X, x0_given, nrhs, rtol, maxiter, w, st, 8));
  for (int j = 0; j < nrhs; ++j) {
    const PcgScalars& s = w.host_scal[j];
    res[j].iters = s.iters;
    res[j].info = s.done ? 0 : maxiter;
    res[j].relres = s.bnorm > 0 ? sqrt(s.rr) / s.bnorm : 0.0;
    res[j].outer = 1;
  }
  return OK;
}

// Mixed precision: FP64 outer residual r = F - K64 x, FP32 inner PCG on the
// correction, x += d, until ||r|| < rtol ||F_f|| in FP64 (classical iterative
// refinement; the converged answer is the FP64 one).
int pcg_mixed(const PcgOps<double>& A64, const PcgOps<float>& A32, const double* F, int64_t ldf,
              double* X, bool x0_given, int nrhs, double rtol, int maxiter, double inner_rtol,
              PcgWork<double>& w64, PcgWork<float>& w32, hipStream_t st, PcgResult* res) {
  const Grid& g = *A64.g;
  const int64_t n = g.n_dofs;
  const unsigned nb = vec_blocks(n);
  const dim3 grid(nb, (unsigned)nrhs), blk(256);
  hipLaunchKernelGGL((k_mask_copy<double, double>), grid, blk, 0, st, n, A64.fixed,
                     x0_given ? (const double*)X : (const double*)nullptr, n, X, n);
  VOX8_HIP(hipGetLastError());
  for (int j = 0; j < nrhs; ++j) { res[j].iters = 0; res[j].outer = 0; res[j].info = maxiter; }
  std::vector<int> done(nrhs, 0);
  int total_inner = 0;
  for (int outer = 0; outer < 64; ++outer) {
    // r64 = F - K64 x  (and FP64 norms)
    VOX8_TRY(A64.apply(X, n, w64.q, n, nrhs, st, DotFuse()));
    hipLaunchKernelGGL((k_pcg_init<double>), grid, blk, 0, st, n, A64.fixed, F, ldf,
                       (const double*)w64.q, n, A64.dinv, w64.r, n, rtol, w64.scal,
                       w64.partials, w64.counters);
    VOX8_HIP(hipGetLastError());
    VOX8_HIP(hipMemcpyAsync(w64.host_scal, w64.scal, sizeof(PcgScalars) * nrhs,
                            hipMemcpyDeviceToHost, st));
    VOX8_HIP(hipStreamSynchronize(st));
    bool all = true;
    for (int j = 0; j < nrhs; ++j) {
      const PcgScalars& s = w64.host_scal[j];
      res[j].relres = s.bnorm > 0 ? sqrt(s.rr) / s.bnorm : 0.0;
      if (s.bnorm == 0.0 || s.rr < s.atol2) {
        if (!done[j]) { done[j] = 1; res[j].info = 0; res[j].outer = outer; }
      } else {
        all = false;
      }
    }
    if (all || total_inner >= maxiter) break;
    // inner: solve K d = r in FP32 (columns already converged have zero weight:
    // their rhs is replaced by zeros so the inner CG finishes immediately)
    for (int j = 0; j < nrhs; ++j) w64.host_scale[j] = done[j] ? 0.0 : 1.0;
    VOX8_TRY(scale_columns_to_float(w64.r, n, w32.rhs64, w64.host_scale, nrhs, n, st));
    VOX8_TRY(run_cg<float>(A32, w32.rhs64, n, w32.x, nullptr, false, nrhs, inner_rtol,
                           maxiter - total_inner, w32, st, 16));
    int mx = 0;
    for (int j = 0; j < nrhs; ++j) {
      if (!done[j]) res[j].iters += w32.host_scal[j].iters;
      mx = w32.host_scal[j].iters > mx ? w32.host_scal[j].iters : mx;
    }
    total_inner += mx > 0 ? mx : 1;
    hipLaunchKernelGGL((k_add_correction<float>), grid, blk, 0, st, n, (const float*)w32.x, n, X, n);
    VOX8_HIP(hipGetLastError());
  }
  return OK;
}

__global__ void k_scale_cols(int64_t n, const double* __restrict__ a, int64_t lda,
                             double* __restrict__ b, int64_t ldb, const double* __restrict__ s) {
  const int col = blockIdx.y;
  const double sc = s[col];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[col * ldb + i] = sc * a[col * lda + i];
}

int scale_columns_to_float(const double* a, int64_t lda, double* b, const double* host_scale,
                           int nrhs, int64_t n, hipStream_t st) {
  // host_scale lives in pinned memory; copy to the device slot after it
  double* dscale = b + (int64_t)nrhs * n;
  VOX8_HIP(hipMemcpyAsync(dscale, host_scale, sizeof(double) * nrhs, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_scale_cols, dim3(vec_blocks(n), nrhs), dim3(256), 0, st, n, a, lda, b, n,
                     (const double*)dscale);
  VOX8_HIP(hipGetLastError());
  return OK;
}

int convert_to_float(const double* a, float* b, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_float, dim3(vec_blocks(n)), dim3(256), 0, st, n, a, b);
  VOX8_HIP(hipGetLastError());
  return OK;
}

}  // namespace vox8
