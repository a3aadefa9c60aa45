// vox8 — general (dense 24x24) matrix-free K.u / M.u, element factors, Jacobi diagonal.
//
// Replaces the assembled-CSR products of the reference:
//   assemble_K / assemble_M + CSR SpMV      src/fem.py:311-325, 362-398
//   K_ff.diagonal() Jacobi                   src/fem.py:441-443, 543-544
// This node-centric gather kernel works for ANY element matrix pair (Ke, Me),
// e.g. the anisotropic Ke a caller writes into FEM3D.Ke_solid. The fast
// structured kernel (vox8_stencil_fast.hip) is used when Ke has the box-element
// block structure; this one is the general path and the fast path's test oracle.
#include "vox8_kernels.h"

namespace vox8 {

// f_K = e_min + (1-e_min) rho^p ; f_M = Pedersen factor (src/fem.py:320, 334-349).
// Inactive elements get 0 for both, so they drop out of every product.
// changed (may be null): set to 1 when any factor differs from the value it
// replaces (bitwise), so a context can tell a repeated upload of the same
// density from a new one.
__global__ void k_element_factors(int64_t n_elem, const double* __restrict__ rho,
                                  const uint8_t* __restrict__ active, double p, double emin,
                                  double mthr, double mq, double* __restrict__ fK,
                                  double* __restrict__ fM, float* __restrict__ fK32,
                                  float* __restrict__ fM32, unsigned int* __restrict__ changed) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  double r = rho[e];
  double k = 0.0, m = 0.0;
  if (active[e]) {
    k = emin + (1.0 - emin) * pow(r, p);
    m = (r >= mthr) ? r : mthr * pow(fmax(r, 0.0) / mthr, mq);
  }
  if (changed && (__double_as_longlong(fK[e]) != __double_as_longlong(k) ||
                  __double_as_longlong(fM[e]) != __double_as_longlong(m)))
    *changed = 1u;                          // every writer stores the same value
  fK[e] = k;
  fM[e] = m;
  fK32[e] = (float)k;
  fM32[e] = (float)m;
}

// y = A x on free DOFs (A = sum_e f_e Ae), x read as 0 at fixed DOFs, y = 0 at
// fixed DOFs.  One thread per node, gathering its <= 8 elements.
//
// With `fuse.scal` set (PCG use) a column whose solve has finished is skipped and
// the dot product x.yK of the column is reduced into fuse.scal[col].pq.
template <typename T, bool DO_K, bool DO_M>
__global__ void __launch_bounds__(256)
k_apply_dense(Grid g, const T* __restrict__ fK, const T* __restrict__ fM,
              const T* __restrict__ Ke, const T* __restrict__ Me,
              const uint8_t* __restrict__ fixed, const T* __restrict__ x, int64_t ldx,
              T* __restrict__ yK, T* __restrict__ yM, int64_t ldy, DotFuse fuse) {
  __shared__ double sh[8];
  const int col = blockIdx.y;
  if (fuse.scal && fuse.scal[col].done) return;
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = n < g.n_nodes;
  if (!live) n = 0;
  x += col * ldx;
  const int ix = (int)(n % g.NX);
  const int iy = (int)((n / g.NX) % g.NY);
  const int iz = (int)(n / ((int64_t)g.NX * g.NY));
  T aK[3] = {0, 0, 0}, aM[3] = {0, 0, 0};
  const uint8_t myfix = fixed[n];
  for (int s = 0; s < 8; ++s) {
    const int sx = s & 1, sy = (s >> 1) & 1, sz = s >> 2;
    const int ex = ix - 1 + sx, ey = iy - 1 + sy, ez = iz - 1 + sz;
    if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
    const int64_t e = ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez);
    const T fk = DO_K ? fK[e] : T(0);
    const T fm = DO_M ? fM[e] : T(0);
    if (fk == T(0) && fm == T(0)) continue;
    const int a = local_idx(1 - sx, 1 - sy, 1 - sz);
    T tK[3] = {0, 0, 0}, tM[3] = {0, 0, 0};
    for (int b = 0; b < 8; ++b) {
      const int64_t nb = (ex + local_off(b, 0))
                       + (int64_t)g.NX * ((ey + local_off(b, 1)) + (int64_t)g.NY * (ez + local_off(b, 2)));
      const uint8_t fb = fixed[nb];
      T xb[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) xb[d] = (fb >> d) & 1 ? T(0) : x[3 * nb + d];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int idx = (3 * a + c) * 24 + 3 * b + d;
          if (DO_K) tK[c] += Ke[idx] * xb[d];
          if (DO_M) tM[c] += Me[idx] * xb[d];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (DO_K) aK[c] += fk * tK[c];
      if (DO_M) aM[c] += fm * tM[c];
    }
  }
  double dot = 0.0;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const bool fx = (myfix >> c) & 1;
    const T vk = fx ? T(0) : aK[c];
    if (live) {
      if (DO_K) yK[col * ldy + 3 * n + c] = vk;
      if (DO_M) yM[col * ldy + 3 * n + c] = fx ? T(0) : aM[c];
    }
    if (DO_K && fuse.scal && live) dot += (double)vk * (double)(fx ? T(0) : x[3 * n + c]);
  }
  if (fuse.scal) {
    const double v = block_sum(dot, sh);
    write_partials<1>(&v, fuse.partials + (int64_t)col * gridDim.x, blockIdx.x);
  }
}

// Lumped tip mass on the translational DOFs of the listed nodes (src/fem.py:383-398):
// y[3n+c] += m * x[3n+c] (fixed DOFs stay zero).
template <typename T>
__global__ void k_tip_mass(const int64_t* __restrict__ nodes, int n_tip, T m,
                           const uint8_t* __restrict__ fixed, const T* __restrict__ x,
                           int64_t ldx, T* __restrict__ y, int64_t ldy) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_tip) return;
  const int col = blockIdx.y;
  const int64_t n = nodes[i];
  const uint8_t f = fixed[n];
  for (int c = 0; c < 3; ++c)
    if (!((f >> c) & 1)) y[col * ldy + 3 * n + c] += m * x[col * ldx + 3 * n + c];
}

// Jacobi inverse diagonal of A_ff: dinv = 1/diag (guard as src/fem.py:442 / :544),
// 0 at fixed DOFs so a preconditioned vector stays zero there.
__global__ void k_jacobi(Grid g, const double* __restrict__ f, const double* __restrict__ Ae,
                         const uint8_t* __restrict__ fixed, int positive_guard,
                         const int64_t* __restrict__ tip_nodes, int n_tip, double tip_m,
                         double* __restrict__ diag_out, double* __restrict__ dinv) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes) return;
  const int ix = (int)(n % g.NX);
  const int iy = (int)((n / g.NX) % g.NY);
  const int iz = (int)(n / ((int64_t)g.NX * g.NY));
  double d[3] = {0, 0, 0};
  for (int s = 0; s < 8; ++s) {
    const int sx = s & 1, sy = (s >> 1) & 1, sz = s >> 2;
    const int ex = ix - 1 + sx, ey = iy - 1 + sy, ez = iz - 1 + sz;
    if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
    const double fe = f[ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez)];
    const int a = local_idx(1 - sx, 1 - sy, 1 - sz);
    for (int c = 0; c < 3; ++c) d[c] += fe * Ae[(3 * a + c) * 25];
  }
  for (int i = 0; i < n_tip; ++i)      // only used for M with a lump (tiny list)
    if (tip_nodes[i] == n) { d[0] += tip_m; d[1] += tip_m; d[2] += tip_m; }
  const uint8_t fx = fixed[n];
  for (int c = 0; c < 3; ++c) {
    const bool isfix = (fx >> c) & 1;
    const double v = d[c];
    if (diag_out) diag_out[3 * n + c] = isfix ? 0.0 : v;
    if (dinv) {
      const bool ok = positive_guard ? (v > 0.0) : (v != 0.0);
      dinv[3 * n + c] = isfix ? 0.0 : (ok ? 1.0 / v : 1.0);
    }
  }
}

// ---------------------------------------------------------------------------
int launch_element_factors(const Grid& g, const double* rho, const uint8_t* active, double p,
                           double emin, double mthr, double mq, double* fK, double* fM,
                           float* fK32, float* fM32, hipStream_t st, unsigned int* changed) {
  const int B = 256;
  hipLaunchKernelGGL(k_element_factors, dim3((unsigned)((g.n_elem + B - 1) / B)), dim3(B), 0, st,
                     g.n_elem, rho, active, p, emin, mthr, mq, fK, fM, fK32, fM32, changed);
  VOX8_HIP(hipGetLastError());
  return OK;
}

template <typename T>
int launch_apply_dense(const Grid& g, int op, const T* fK, const T* fM, const T* Ke,
                       const T* Me, const uint8_t* fixed, const T* x, int64_t ldx, T* yK,
                       T* yM, int64_t ldy, int nrhs, hipStream_t st, DotFuse fuse) {
  const int B = 256;
  dim3 grid((unsigned)((g.n_nodes + B - 1) / B), (unsigned)nrhs);
  if (op == OP_K)
    hipLaunchKernelGGL((k_apply_dense<T, true, false>), grid, dim3(B), 0, st, g, fK, fM, Ke, Me,
                       fixed, x, ldx, yK, yM, ldy, fuse);
  else if (op == OP_M)
    hipLaunchKernelGGL((k_apply_dense<T, false, true>), grid, dim3(B), 0, st, g, fK, fM, Ke, Me,
                       fixed, x, ldx, yK, yM, ldy, fuse);
  else
    hipLaunchKernelGGL((k_apply_dense<T, true, true>), grid, dim3(B), 0, st, g, fK, fM, Ke, Me,
                       fixed, x, ldx, yK, yM, ldy, fuse);
  VOX8_HIP(hipGetLastError());
  if (fuse.scal) VOX8_TRY(launch_pq_finish(fuse, (int)grid.x, nrhs, st));
  return OK;
}
template int launch_apply_dense<double>(const Grid&, int, const double*, const double*,
                                        const double*, const double*, const uint8_t*,
                                        const double*, int64_t, double*, double*, int64_t, int,
                                        hipStream_t, DotFuse);
template int launch_apply_dense<float>(const Grid&, int, const float*, const float*,
                                       const float*, const float*, const uint8_t*, const float*,
                                       int64_t, float*, float*, int64_t, int, hipStream_t,
                                       DotFuse);

template <typename T>
int launch_tip_mass(const int64_t* nodes, int n_tip, double m, const uint8_t* fixed, const T* x,
                    int64_t ldx, T* y, int64_t ldy, int nrhs, hipStream_t st) {
  if (n_tip <= 0 || m == 0.0) return OK;
  const int B = 128;
  hipLaunchKernelGGL(k_tip_mass<T>, dim3((n_tip + B - 1) / B, nrhs), dim3(B), 0, st, nodes, n_tip,
                     (T)m, fixed, x, ldx, y, ldy);
  VOX8_HIP(hipGetLastError());
  return OK;
}
template int launch_tip_mass<double>(const int64_t*, int, double, const uint8_t*, const double*,
                                     int64_t, double*, int64_t, int, hipStream_t);
template int launch_tip_mass<float>(const int64_t*, int, double, const uint8_t*, const float*,
                                    int64_t, float*, int64_t, int, hipStream_t);

int launch_jacobi(const Grid& g, const double* f, const double* Ae, const uint8_t* fixed,
                  int positive_guard, const int64_t* tip_nodes, int n_tip, double tip_m,
                  double* diag_out, double* dinv, hipStream_t st) {
  const int B = 256;
  hipLaunchKernelGGL(k_jacobi, dim3((unsigned)((g.n_nodes + B - 1) / B)), dim3(B), 0, st, g, f, Ae,
                     fixed, positive_guard, tip_nodes, n_tip, tip_m, diag_out, dinv);
  VOX8_HIP(hipGetLastError());
  return OK;
}

}  // namespace vox8
