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
#include <algorithm>
#include <cmath>
#include <vector>

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

// scipy cg returns b (= 0) when ||b|| = 0, whatever x0 was.
template <typename T>
__global__ void k_pcg_zero_null(int64_t n_dofs, T* __restrict__ x, int64_t ldx,
                                const PcgScalars* __restrict__ scal) {
  const int col = blockIdx.y;
  if (scal[col].bnorm != 0.0) return;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n_dofs;
       i += (int64_t)gridDim.x * blockDim.x)
    x[col * ldx + i] = T(0);
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
// The grid-stride loop is unrolled 4x (16 independent loads in flight per
// thread); each thread's partial sums keep a fixed order.
template <typename T, bool RZ>
__global__ void __launch_bounds__(256)
k_pcg_step(int64_t n_dofs, const T* __restrict__ dinv, T* __restrict__ x,
           T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ q, int64_t ld,
           PcgScalars* __restrict__ scal, int maxiter, double* __restrict__ partials,
           unsigned int* __restrict__ counters, float* __restrict__ r32) {
  __shared__ double sh[8];
  const int col = blockIdx.y;
  const PcgScalars s = scal[col];
  if (s.done || s.iters >= maxiter) return;
  const double alpha = s.rz / s.pq;
  double rr = 0.0, rz = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  x += col * ld; r += col * ld; p += col * ld; q += col * ld;
  if (r32) r32 += col * ld;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n_dofs; i += 4 * stride) {
    T xv[4], pv[4], rv[4], qv[4], dv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t j = i + u * stride;
      xv[u] = x[j]; pv[u] = p[j]; rv[u] = r[j]; qv[u] = q[j];
      if (RZ) dv[u] = dinv[j];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t j = i + u * stride;
      x[j] = (T)((double)xv[u] + alpha * (double)pv[u]);
      const T rn = (T)((double)rv[u] - alpha * (double)qv[u]);
      r[j] = rn;
      if (r32) r32[j] = (float)rn;
      rr += (double)rn * (double)rn;
      if (RZ) rz += (double)rn * (double)dv[u] * (double)rn;
    }
  }
  for (; i < n_dofs; i += stride) {
    x[i] = (T)((double)x[i] + alpha * (double)p[i]);
    const T rn = (T)((double)r[i] - alpha * (double)q[i]);
    r[i] = rn;
    if (r32) r32[i] = (float)rn;
    rr += (double)rn * (double)rn;
    if (RZ) rz += (double)rn * (double)dinv[i] * (double)rn;
  }
  const double v[2] = {block_sum(rr, sh), RZ ? block_sum(rz, sh) : 0.0};
  write_partials<2>(v, partials + (int64_t)col * gridDim.x * 2, blockIdx.x);
}

// After k_pcg_step: ||r||^2 (and r.D^-1 r) from the partials, iteration count.
template <bool RZ>
__global__ void __launch_bounds__(256)
k_pcg_step_finish(PcgScalars* __restrict__ scal, int maxiter, const double* __restrict__ partials,
                  int nb) {
  __shared__ double sh[8];
  const int col = blockIdx.y;
  const PcgScalars s = scal[col];
  if (s.done || s.iters >= maxiter) return;
  double tot[2];
  sum_partials<2>(partials + (int64_t)col * nb * 2, nb, tot, sh);
  if (threadIdx.x == 0) {
    scal[col].rr = tot[0];
    if (RZ) {
      scal[col].rz_prev = s.rz;
      scal[col].rz = tot[1];
    }
    scal[col].iters = s.iters + 1;
  }
}

// Preconditioned variant (explicit z = M r): stop test, then rho = r.z.
// skip[col] tells the preconditioner which columns need no more work.
__global__ void k_pcg_check(PcgScalars* __restrict__ scal, int maxiter, int* __restrict__ skip) {
  const int col = blockIdx.y;
  PcgScalars& s = scal[col];
  if (!s.done && s.iters < maxiter && s.rr < s.atol2) s.done = 1;
  skip[col] = s.done || s.iters >= maxiter;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_pcg_rz(int64_t n_dofs, const T* __restrict__ r, const T* __restrict__ z, int64_t ld,
         PcgScalars* __restrict__ scal, int maxiter, double* __restrict__ partials,
         unsigned int* __restrict__ counters) {
  __shared__ double sh[8];
  const int col = blockIdx.y;
  const PcgScalars s = scal[col];
  if (s.done || s.iters >= maxiter) return;
  double rz = 0.0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  r += col * ld; z += col * ld;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n_dofs; i += 4 * stride) {
    T rv[4], zv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { rv[u] = r[i + u * stride]; zv[u] = z[i + u * stride]; }
#pragma unroll
    for (int u = 0; u < 4; ++u) rz += (double)rv[u] * (double)zv[u];
  }
  for (; i < n_dofs; i += stride) rz += (double)r[i] * (double)z[i];
  const double v = block_sum(rz, sh);
  write_partials<1>(&v, partials + (int64_t)col * gridDim.x, blockIdx.x);
}

// After k_pcg_rz: rho = r.z from the partials.
__global__ void __launch_bounds__(256)
k_pcg_rz_finish(PcgScalars* __restrict__ scal, int maxiter, const double* __restrict__ partials,
                int nb) {
  __shared__ double sh[8];
  const int col = blockIdx.y;
  const PcgScalars s = scal[col];
  if (s.done || s.iters >= maxiter) return;
  double tot;
  sum_partials<1>(partials + (int64_t)col * nb, nb, &tot, sh);
  if (threadIdx.x == 0) {
    scal[col].rz_prev = s.rz;
    scal[col].rz = tot;
  }
}

// After a fused K p product: p.q from the partials.
__global__ void __launch_bounds__(256)
k_pq_finish(PcgScalars* __restrict__ scal, const double* __restrict__ partials, int nb) {
  __shared__ double sh[8];
  const int col = blockIdx.y;
  if (scal[col].done) return;
  double tot;
  sum_partials<1>(partials + (int64_t)col * nb, nb, &tot, sh);
  if (threadIdx.x == 0) scal[col].pq = tot;
}

// Column sums of per-workgroup partials (partials[col * nb + b]), fixed order.
__global__ void __launch_bounds__(256)
k_partials_finish(const double* __restrict__ partials, int nb, double* __restrict__ out) {
  __shared__ double sh[8];
  const int col = blockIdx.y;
  double tot;
  sum_partials<1>(partials + (int64_t)col * nb, nb, &tot, sh);
  if (threadIdx.x == 0) out[col] = tot;
}

int launch_partials_finish(const double* partials, int nb, int ncols, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_partials_finish, dim3(1, ncols), dim3(256), 0, st, partials, nb, out);
  VOX8_HIP(hipGetLastError());
  return OK;
}

int launch_pq_finish(const DotFuse& fuse, int nb, int ncols, hipStream_t st) {
  hipLaunchKernelGGL(k_pq_finish, dim3(1, ncols), dim3(256), 0, st, fuse.scal,
                     (const double*)fuse.partials, nb);
  VOX8_HIP(hipGetLastError());
  return OK;
}

template <typename T>
__global__ void __launch_bounds__(256)
k_pcg_dir_z(int64_t n_dofs, const T* __restrict__ z, T* __restrict__ p, int64_t ld,
            const PcgScalars* __restrict__ scal, int maxiter) {
  const int col = blockIdx.y;
  const PcgScalars s = scal[col];
  if (s.done || s.iters >= maxiter) return;
  const double beta = s.iters > 0 ? s.rz / s.rz_prev : 0.0;
  const bool first = s.iters == 0;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  z += col * ld; p += col * ld;
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i + 3 * stride < n_dofs; i += 4 * stride) {
    T zv[4], pv[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { zv[u] = z[i + u * stride]; pv[u] = first ? T(0) : p[i + u * stride]; }
#pragma unroll
    for (int u = 0; u < 4; ++u) p[i + u * stride] = (T)((double)zv[u] + beta * (double)pv[u]);
  }
  for (; i < n_dofs; i += stride) p[i] = (T)(first ? (double)z[i] : (double)z[i] + beta * (double)p[i]);
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
                  hipStream_t st, int chunk0, const Precond<T>* M, float* M32 = nullptr) {
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
  hipLaunchKernelGGL((k_pcg_zero_null<T>), grid, blk, 0, st, n, x, n, (const PcgScalars*)w.scal);
  VOX8_HIP(hipGetLastError());
  DotFuse fuse;
  fuse.scal = w.scal;
  fuse.partials = w.partials;
  fuse.counters = w.counters;
  // Iterations are launched in chunks between host reads of the column
  // states.  Iterations launched past a column's convergence are no-ops on
  // the device, but each still costs its launches (a multigrid cycle is ~150
  // of them), so after the first reads the next chunk is the remaining
  // iteration count extrapolated from the slowest column's observed
  // residual reduction per iteration (plus one), doubling only while no
  // rate is known yet.
  int launched = 0, chunk = chunk0;
  bool r32_ok = false;
  std::vector<double> rr_seen(nrhs, -1.0);
  std::vector<int> it_seen(nrhs, 0);
  while (launched < maxiter) {
    const int todo = chunk < maxiter - launched ? chunk : maxiter - launched;
    for (int it = 0; it < todo; ++it) {
      if (!M) {
        hipLaunchKernelGGL((k_pcg_dir<T>), grid, blk, 0, st, n, A.dinv, w.r, w.p, n, w.scal, maxiter);
        VOX8_TRY(A.apply(w.p, n, w.q, n, nrhs, st, fuse));
        hipLaunchKernelGGL((k_pcg_step<T, true>), grid, blk, 0, st, n, A.dinv, x, w.r, w.p, w.q, n,
                           w.scal, maxiter, w.partials, w.counters, (float*)nullptr);
        hipLaunchKernelGGL(k_pcg_step_finish<true>, dim3(1, nrhs), blk, 0, st, w.scal, maxiter,
                           (const double*)w.partials, (int)nb);
      } else {
        hipLaunchKernelGGL(k_pcg_check, dim3(1, nrhs), dim3(1), 0, st, w.scal, maxiter, w.skip);
        // after the first step the step kernel has written the residual in
        // FP32 where the preconditioner reads it (M32), so M skips its conversion
        VOX8_TRY((*M)(r32_ok ? nullptr : w.r, w.x, nrhs, st, w.skip));
        hipLaunchKernelGGL((k_pcg_rz<T>), grid, blk, 0, st, n, (const T*)w.r, (const T*)w.x, n,
                           w.scal, maxiter, w.partials, w.counters);
        hipLaunchKernelGGL(k_pcg_rz_finish, dim3(1, nrhs), blk, 0, st, w.scal, maxiter,
                           (const double*)w.partials, (int)nb);
        hipLaunchKernelGGL((k_pcg_dir_z<T>), grid, blk, 0, st, n, (const T*)w.x, w.p, n,
                           (const PcgScalars*)w.scal, maxiter);
        VOX8_TRY(A.apply(w.p, n, w.q, n, nrhs, st, fuse));
        hipLaunchKernelGGL((k_pcg_step<T, false>), grid, blk, 0, st, n, A.dinv, x, w.r, w.p, w.q, n,
                           w.scal, maxiter, w.partials, w.counters, M32);
        r32_ok = M32 != nullptr;
        hipLaunchKernelGGL(k_pcg_step_finish<false>, dim3(1, nrhs), blk, 0, st, w.scal, maxiter,
                           (const double*)w.partials, (int)nb);
      }
    }
    VOX8_HIP(hipGetLastError());
    launched += todo;
    VOX8_HIP(hipMemcpyAsync(w.host_scal, w.scal, sizeof(PcgScalars) * nrhs, hipMemcpyDeviceToHost, st));
    VOX8_HIP(hipStreamSynchronize(st));
    bool all = true;
    double need = 0.0;              // iterations the slowest running column still needs
    bool known = true;
    for (int j = 0; j < nrhs; ++j) {
      const PcgScalars& sc = w.host_scal[j];
      if (sc.done || sc.iters >= maxiter) continue;
      all = false;
      const double rr = sc.rr;
      if (rr_seen[j] > 0.0 && rr > 0.0 && sc.iters > it_seen[j] && rr < rr_seen[j] && sc.atol2 > 0.0) {
        const double per_it = std::log(rr / rr_seen[j]) / (sc.iters - it_seen[j]);   // < 0
        need = std::max(need, std::log(sc.atol2 / rr) / per_it);
      } else {
        known = false;
      }
      rr_seen[j] = rr;
      it_seen[j] = sc.iters;
    }
    if (all) break;
    if (known)
      chunk = need < 255.0 ? std::max(1, (int)std::ceil(need) + 1) : 256;
    else
      chunk = chunk * 2 < 256 ? chunk * 2 : 256;
  }
  // As in SciPy, a column that reaches maxiter is reported unconverged even if
  // its last step met the tolerance (the stop test runs before an iteration).
  return OK;
}

int pcg_fp64(const PcgOps<double>& A, const double* F, int64_t ldf, double* X, bool x0_given,
             int nrhs, double rtol, int maxiter, PcgWork<double>& w, hipStream_t st,
             PcgResult* res, const Precond<double>* M, float* M32) {
  VOX8_TRY(run_cg<double>(A, F, ldf, X, X, x0_given, nrhs, rtol, maxiter, w, st, M ? 2 : 8, M,
                          M ? M32 : nullptr));
  for (int j = 0; j < nrhs; ++j) {
    const PcgScalars& s = w.host_scal[j];
    res[j].iters = s.iters;
    res[j].info = s.done ? 0 : maxiter;
    res[j].relres = s.bnorm > 0 ? sqrt(s.rr) / s.bnorm : 0.0;
    res[j].outer = 1;
  }
  return OK;
}

int convert_to_float(const double* a, float* b, int64_t n, hipStream_t st) {
  hipLaunchKernelGGL(k_to_float, dim3(vec_blocks(n)), dim3(256), 0, st, n, a, b);
  VOX8_HIP(hipGetLastError());
  return OK;
}

}  // namespace vox8
