// vox8 — SIMP-loop kernels: masked cone filter, element quadratic forms and the
// compliance / eigenvalue sensitivities, OC bisection, small reductions.
//
// Reference counterparts:
//   filter_field (scipy.ndimage.convolve, constant mode)     src/topopt.py:44-58
//   elem_compliance + dc = -p(1-e_min) rho^(p-1) ce          src/fem.py:605-615, src/topopt.py:292-306
//   eigenvalue_sensitivity                                    src/fem.py:472-526
//   _oc_update (bisection on the multiplier)                  src/topopt.py:190-214
//   _select_mode_by_mac, change / mass bookkeeping            src/topopt.py:178-188, 383-397
// All grid arrays are element-ordered e = ex + nx*(ey + ny*ez).
#include "vox8_kernels.h"
#include "vox8_topopt.h"

namespace vox8 {

static unsigned nblk(int64_t n, int cap = 1024) {
  int64_t b = (n + 255) / 256;
  return (unsigned)(b < cap ? (b < 1 ? 1 : b) : cap);
}

// Dynamic-count variant of the last-block grid reduction.
__device__ bool grid_reduce_last_dyn(const double* vals /*smem, nv*/, int nv,
                                     double* partials, unsigned int* counter, double* out,
                                     double* sh) {
  __shared__ int s_last;
  const int nb = gridDim.x, bid = blockIdx.x;
  if (threadIdx.x == 0) {
    for (int v = 0; v < nv; ++v)
      __hip_atomic_store(&partials[(int64_t)bid * nv + v], vals[v], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    s_last = atomicAdd(counter, 1u) == (unsigned)(nb - 1);
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
  for (int v = 0; v < nv; ++v) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x)
      t += __hip_atomic_load(&partials[(int64_t)b * nv + v], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    t = block_sum(t, sh);
    if (threadIdx.x == 0) out[v] = t;
  }
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

// ---- masked cone filter ------------------------------------------------------
// mode FILTER_FWD:  out[e] = mask[e] ? sum_t w_t in[e+t] mask[e+t] / den[e] : 0
// mode FILTER_ADJ:  out[e] = mask[e] ? sum_t w_t in[e+t] mask[e+t] / den[e+t] : 0
//                   (the transpose of FILTER_FWD: the tap set is symmetric)
// mode FILTER_DEN:  out[e] = max(sum_t w_t mask[e+t], 1e-12)  (the denominator)
__global__ void __launch_bounds__(256)
k_filter(int nx, int ny, int nz, int ntap, const int* __restrict__ toff,
         const double* __restrict__ tw, const uint8_t* __restrict__ mask,
         const double* __restrict__ den, const double* __restrict__ in, double* __restrict__ out,
         int mode, const double* __restrict__ gate) {
  if (gate && !(*gate > 0.0)) return;
  const int64_t n = (int64_t)nx * ny * nz;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int ex = (int)(e % nx), ey = (int)((e / nx) % ny), ez = (int)(e / ((int64_t)nx * ny));
    if (mode != FILTER_DEN && !mask[e]) { out[e] = 0.0; continue; }
    double s = 0.0;
    for (int t = 0; t < ntap; ++t) {
      const int dx = toff[3 * t], dy = toff[3 * t + 1], dz = toff[3 * t + 2];
      const int x = ex + dx, y = ey + dy, z = ez + dz;
      if (x < 0 || y < 0 || z < 0 || x >= nx || y >= ny || z >= nz) continue;
      const int64_t f = x + (int64_t)nx * (y + (int64_t)ny * z);
      if (!mask[f]) continue;
      const double v = mode == FILTER_DEN ? 1.0 : mode == FILTER_ADJ ? in[f] / den[f] : in[f];
      s += tw[t] * v;
    }
    out[e] = mode == FILTER_FWD ? s / den[e] : mode == FILTER_ADJ ? s : fmax(s, 1e-12);
  }
}

int launch_filter(int nx, int ny, int nz, int ntap, const int* toff, const double* tw,
                  const uint8_t* mask, const double* den, const double* in, double* out,
                  int mode, hipStream_t st, const double* gate) {
  hipLaunchKernelGGL(k_filter, dim3(nblk((int64_t)nx * ny * nz, 4096)), dim3(256), 0, st, nx, ny, nz,
                     ntap, toff, tw, mask, den, in, out, mode, gate);
  VOX8_HIP(hipGetLastError());
  return OK;
}

// ---- element quadratic forms + compliance / eigen sums -------------------------
// For every element (factor fK != 0, i.e. active):
//   ce[e]  = sum_j w_j U_j,e^T Ke U_j,e        (U: nrhs column-major vectors)
//   kphi[e]= phi_e^T Ke phi_e, mphi[e] = phi_e^T Me phi_e   (if phi given)
// Reduced: comp[j] = sum_e fK_e U_j,e^T Ke U_j,e  and  vmv = sum_e fM_e mphi[e].
__global__ void __launch_bounds__(256)
k_quadforms(Grid g, const double* __restrict__ fK, const double* __restrict__ fM,
            const uint8_t* __restrict__ active, const double* __restrict__ Ke,
            const double* __restrict__ Me, const double* __restrict__ U, int64_t ldu, int nrhs,
            const double* __restrict__ wts, const double* __restrict__ phi,
            double* __restrict__ ce, double* __restrict__ kphi, double* __restrict__ mphi,
            double* __restrict__ partials, unsigned int* __restrict__ counter,
            double* __restrict__ sums, const int64_t* __restrict__ tip_nodes, int n_tip,
            double tip_m, const double* __restrict__ gate) {
  if (gate && !(*gate > 0.0)) return;     // uniform over the grid: the counter is untouched
  __shared__ double sh[8], sv[20];
  double acc[17];
  const int nv = nrhs + 1;
  for (int v = 0; v < 17; ++v) acc[v] = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < g.n_elem;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (!active[e]) {
      if (ce) ce[e] = 0.0;
      if (phi) { kphi[e] = 0.0; mphi[e] = 0.0; }
      continue;
    }
    const int ex = (int)(e % g.nx), ey = (int)((e / g.nx) % g.ny),
              ez = (int)(e / ((int64_t)g.nx * g.ny));
    int64_t dof[8];
    for (int b = 0; b < 8; ++b)
      dof[b] = 3 * ((ex + local_off(b, 0)) +
                    (int64_t)g.NX * ((ey + local_off(b, 1)) + (int64_t)g.NY * (ez + local_off(b, 2))));
    double cw = 0.0;
    for (int j = 0; j < nrhs; ++j) {
      double u[24];
      for (int b = 0; b < 8; ++b)
        for (int c = 0; c < 3; ++c) u[3 * b + c] = U[j * ldu + dof[b] + c];
      double q = 0.0;
      for (int r = 0; r < 24; ++r) {
        double t = 0.0;
        for (int s = 0; s < 24; ++s) t += Ke[r * 24 + s] * u[s];
        q += u[r] * t;
      }
      cw += wts[j] * q;
      acc[j] += fK[e] * q;
    }
    if (ce) ce[e] = cw;
    if (phi) {
      double u[24];
      for (int b = 0; b < 8; ++b)
        for (int c = 0; c < 3; ++c) u[3 * b + c] = phi[dof[b] + c];
      double qk = 0.0, qm = 0.0;
      for (int r = 0; r < 24; ++r) {
        double tk = 0.0, tm = 0.0;
        for (int s = 0; s < 24; ++s) {
          tk += Ke[r * 24 + s] * u[s];
          tm += Me[r * 24 + s] * u[s];
        }
        qk += u[r] * tk;
        qm += u[r] * tm;
      }
      kphi[e] = qk;
      mphi[e] = qm;
      acc[nrhs] += fM[e] * qm;
    }
  }
  for (int v = 0; v < nv; ++v) {
    const double t = block_sum(acc[v], sh);
    if (threadIdx.x == 0) sv[v] = t;
  }
  __syncthreads();
  if (grid_reduce_last_dyn(sv, nv, partials, counter, sums, sh) && phi && threadIdx.x == 0) {
    // v^T M v includes the rho-independent tip lump (src/fem.py:511-520)
    double t = 0.0;
    for (int i = 0; i < n_tip; ++i)
      for (int c = 0; c < 3; ++c) t += phi[3 * tip_nodes[i] + c] * phi[3 * tip_nodes[i] + c];
    sums[nrhs] += tip_m * t;
  }
}

// dc[e] = -p (1-e_min) rho^(p-1) ce[e]           (src/topopt.py:303)
// dlam[e] = ((1-e_min) p max(rho,1e-9)^(p-1) kphi - lam m'(rho) mphi) / vMv
//                                                 (src/fem.py:492-526)
__global__ void k_sens_finish(int64_t n, const uint8_t* __restrict__ active,
                              const double* __restrict__ rho, double p, double emin,
                              double mthr, double mq, const double* __restrict__ ce,
                              double* __restrict__ dc, const double* __restrict__ kphi,
                              const double* __restrict__ mphi, double lam,
                              const double* __restrict__ vmv_parts, double* __restrict__ dlam,
                              const double* __restrict__ lam_dev, const double* __restrict__ gate) {
  if (gate && !(*gate > 0.0)) return;
  if (lam_dev) lam = *lam_dev;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (!active[e]) {
      if (dc) dc[e] = 0.0;
      if (dlam) dlam[e] = 0.0;
      continue;
    }
    const double r = rho[e];
    if (dc) dc[e] = -p * (1.0 - emin) * pow(r, p - 1.0) * ce[e];
    if (dlam) {
      double vmv = *vmv_parts;
      if (vmv <= 0.0) vmv = 1.0;
      const double dk = (1.0 - emin) * p * pow(fmax(r, 1e-9), p - 1.0);
      const double dm = r >= mthr ? 1.0 : mq * pow(fmax(r, 0.0) / mthr, mq - 1.0);
      dlam[e] = (dk * kphi[e] - lam * dm * mphi[e]) / vmv;
    }
  }
}

// The same forms for a box element whose Ke / Me the structured stencil
// handles (pattern != dense): with m = H^T v the butterfly coefficients of the
// element's corner values v (per component, monomial order of the stencil),
// v^T Ke v = m^T Kt m over the three diagonal blocks of Kt (lin 9x9, bil 9x9,
// tri 3x3 in C, scaled so that no other factor enters) and v^T Me v =
// sum_k md[k] sum_c m_c[k]^2: about 240 flops per element and vector instead
// of 1152.  Block sums go to partials[block][nv] and k_quadforms_finish adds
// them in a fixed order (no grid-wide fence, bitwise reproducible).
__device__ __forceinline__ void qf_mono(const double* __restrict__ v, int64_t ld,
                                        const int64_t dof[8], double m[3][8]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double w[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) w[b] = v[dof[b] + c];
    // corners in the reference's local order -> x + 2y + 4z order of the butterflies
    const double x0 = w[0], x1 = w[1], x2 = w[3], x3 = w[2], x4 = w[4], x5 = w[5], x6 = w[7], x7 = w[6];
    const double s0 = x1 + x0, d0 = x1 - x0, s1 = x3 + x2, d1 = x3 - x2;
    const double s2 = x5 + x4, d2 = x5 - x4, s3 = x7 + x6, d3 = x7 - x6;
    const double ss0 = s1 + s0, sd0 = s1 - s0, ds0 = d1 + d0, dd0 = d1 - d0;
    const double ss1 = s3 + s2, sd1 = s3 - s2, ds1 = d3 + d2, dd1 = d3 - d2;
    m[c][0] = ss1 + ss0; m[c][1] = ds1 + ds0; m[c][2] = sd1 + sd0; m[c][3] = ss1 - ss0;
    m[c][4] = dd1 + dd0; m[c][5] = sd1 - sd0; m[c][6] = ds1 - ds0; m[c][7] = dd1 - dd0;
  }
  (void)ld;
}

__device__ __forceinline__ double qf_stiff(const FastCoef<double>& C, const double m[3][8]) {
  double al[9], ab[9], at[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    al[c] = m[c][1]; al[3 + c] = m[c][2]; al[6 + c] = m[c][3];
    ab[c] = m[c][4]; ab[3 + c] = m[c][5]; ab[6 + c] = m[c][6];
    at[c] = m[c][7];
  }
  double q = 0.0;
#pragma unroll
  for (int r = 0; r < 9; ++r) {
    double tl = 0.0, tb = 0.0;
#pragma unroll
    for (int s = 0; s < 9; ++s) { tl += C.lin[r][s] * al[s]; tb += C.bil[r][s] * ab[s]; }
    q += al[r] * tl + ab[r] * tb;
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) q += at[r] * (C.tri[r][0] * at[0] + C.tri[r][1] * at[1] + C.tri[r][2] * at[2]);
  return q;
}

__global__ void __launch_bounds__(256)
k_quadforms_mono(Grid g, FastCoef<double> C, const double* __restrict__ fK,
                 const double* __restrict__ fM, const uint8_t* __restrict__ active,
                 const double* __restrict__ U, int64_t ldu, int nrhs, const double* __restrict__ wts,
                 const double* __restrict__ phi, double* __restrict__ ce, double* __restrict__ kphi,
                 double* __restrict__ mphi, double* __restrict__ partials,
                 const double* __restrict__ gate) {
  if (gate && !(*gate > 0.0)) return;
  __shared__ double sh[8];
  double acc[17];
#pragma unroll
  for (int v = 0; v < 17; ++v) acc[v] = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < g.n_elem;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (!active[e]) {
      if (ce) ce[e] = 0.0;
      if (phi) { kphi[e] = 0.0; mphi[e] = 0.0; }
      continue;
    }
    const int ex = (int)(e % g.nx), ey = (int)((e / g.nx) % g.ny),
              ez = (int)(e / ((int64_t)g.nx * g.ny));
    int64_t dof[8];
#pragma unroll
    for (int b = 0; b < 8; ++b)
      dof[b] = 3 * ((ex + local_off(b, 0)) +
                    (int64_t)g.NX * ((ey + local_off(b, 1)) + (int64_t)g.NY * (ez + local_off(b, 2))));
    double cw = 0.0;
    const double fk = fK[e];
#pragma unroll 1
    for (int j = 0; j < nrhs; ++j) {
      double m[3][8];
      qf_mono(U + j * ldu, ldu, dof, m);
      const double q = qf_stiff(C, m);
      cw += wts[j] * q;
#pragma unroll
      for (int v = 0; v < 16; ++v)
        if (v == j) acc[v] += fk * q;            // unrolled: acc stays in registers
    }
    if (ce) ce[e] = cw;
    if (phi) {
      double m[3][8];
      qf_mono(phi, 0, dof, m);
      const double qk = qf_stiff(C, m);
      double qm = 0.0;
#pragma unroll
      for (int k = 0; k < 8; ++k) qm += C.md[k] * (m[0][k] * m[0][k] + m[1][k] * m[1][k] + m[2][k] * m[2][k]);
      kphi[e] = qk;
      mphi[e] = qm;
      acc[16] += fM[e] * qm;
    }
  }
  // block sums of the nrhs compliance values and of phi^T M phi (slot 16)
#pragma unroll
  for (int v = 0; v < 17; ++v) {
    const double t = block_sum(acc[v], sh);
    if (threadIdx.x == 0) partials[(int64_t)blockIdx.x * 17 + v] = t;
  }
}

// sums[j] (j < nrhs) and sums[nrhs] = phi^T M phi + the tip lump, from the
// block partials in a fixed order; one block.
__global__ void __launch_bounds__(256)
k_quadforms_finish(int nb, int nrhs, bool have_phi, const double* __restrict__ partials,
                   const double* __restrict__ phi, const int64_t* __restrict__ tip_nodes, int n_tip,
                   double tip_m, double* __restrict__ sums, const double* __restrict__ gate) {
  if (gate && !(*gate > 0.0)) return;
  __shared__ double sh[8];
  for (int v = 0; v <= nrhs; ++v) {
    if (v == nrhs && !have_phi) break;
    const int slot = v < nrhs ? v : 16;
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) t += partials[(int64_t)b * 17 + slot];
    t = block_sum(t, sh);
    if (threadIdx.x == 0) {
      if (v == nrhs) {
        // v^T M v includes the rho-independent tip lump (src/fem.py:511-520)
        double tl = 0.0;
        for (int i = 0; i < n_tip; ++i)
          for (int c = 0; c < 3; ++c) tl += phi[3 * tip_nodes[i] + c] * phi[3 * tip_nodes[i] + c];
        t += tip_m * tl;
      }
      sums[v] = t;
    }
  }
}

int launch_quadforms(const Grid& g, const double* fK, const double* fM, const uint8_t* active,
                     const double* Ke, const double* Me, const double* U, int64_t ldu, int nrhs,
                     const double* wts, const double* phi, double* ce, double* kphi,
                     double* mphi, double* partials, unsigned int* counter, double* sums,
                     const int64_t* tip_nodes, int n_tip, double tip_m, hipStream_t st,
                     const double* gate, const FastCoef<double>* mono) {
  VOX8_CHECK(nrhs >= 0 && nrhs <= 16, "quadforms: at most 16 load cases");
  if (mono) {
    const unsigned nb = nblk(g.n_elem, 2048);    // partials: nb x 17 doubles
    hipLaunchKernelGGL(k_quadforms_mono, dim3(nb), dim3(256), 0, st, g, *mono, fK, fM, active, U, ldu,
                       nrhs, wts, phi, ce, kphi, mphi, partials, gate);
    hipLaunchKernelGGL(k_quadforms_finish, dim3(1), dim3(256), 0, st, (int)nb, nrhs, phi != nullptr,
                       (const double*)partials, phi, tip_nodes, n_tip, tip_m, sums, gate);
  } else {
    hipLaunchKernelGGL(k_quadforms, dim3(nblk(g.n_elem)), dim3(256), 0, st, g, fK, fM, active, Ke,
                       Me, U, ldu, nrhs, wts, phi, ce, kphi, mphi, partials, counter, sums,
                       tip_nodes, n_tip, tip_m, gate);
  }
  VOX8_HIP(hipGetLastError());
  return OK;
}

int launch_sens_finish(int64_t n, const uint8_t* active, const double* rho, double p, double emin,
                       double mthr, double mq, const double* ce, double* dc, const double* kphi,
                       const double* mphi, double lam, const double* vmv, double* dlam,
                       hipStream_t st, const double* lam_dev, const double* gate) {
  hipLaunchKernelGGL(k_sens_finish, dim3(nblk(n, 4096)), dim3(256), 0, st, n, active, rho, p, emin,
                     mthr, mq, ce, dc, kphi, mphi, lam, vmv, dlam, lam_dev, gate);
  VOX8_HIP(hipGetLastError());
  return OK;
}

// ---- OC bisection ---------------------------------------------------------------
// state: [0] l1, [1] l2, [2] lmid of the last step, [3] done flag, [4] steps
__device__ __forceinline__ double oc_x(double x, double s, double lmid, double Ve, double move) {
  const double t = x * sqrt(s / (lmid * Ve));
  return fmax(0.001, fmax(x - move, fmin(1.0, fmin(x + move, t))));
}

__global__ void __launch_bounds__(256)
k_oc_step(int64_t n, const uint8_t* __restrict__ design, const double* __restrict__ x,
          const double* __restrict__ dc, double Ve, double target, double move,
          double* __restrict__ state, double* __restrict__ partials,
          unsigned int* __restrict__ counter) {
  __shared__ double sh[8], sv[1];
  const double l1 = state[0], l2 = state[1];
  if (state[3] != 0.0) return;
  if ((l2 - l1) / (l1 + l2) <= 1e-4) {
    if (blockIdx.x == 0 && threadIdx.x == 0) state[3] = 1.0;
    return;
  }
  const double lmid = 0.5 * (l1 + l2);
  double v = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    if (design[e]) v += oc_x(x[e], fabs(dc[e]), lmid, Ve, move);
  const double t = block_sum(v, sh);
  if (threadIdx.x == 0) sv[0] = t;
  __syncthreads();
  double tot;
  if (grid_reduce_last_dyn(sv, 1, partials, counter, &tot, sh)) {
    if (threadIdx.x == 0) {
      if (tot * Ve > target) state[0] = lmid; else state[1] = lmid;
      state[2] = lmid;
      state[4] += 1.0;
    }
  }
}

__global__ void k_oc_apply(int64_t n, const uint8_t* __restrict__ design,
                           const double* __restrict__ x, const double* __restrict__ dc, double Ve,
                           double move, const double* __restrict__ state,
                           double* __restrict__ xnew) {
  const double lmid = state[2];
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    xnew[e] = design[e] ? oc_x(x[e], fabs(dc[e]), lmid, Ve, move) : x[e];
}

__global__ void k_oc_init(double* state) {
  state[0] = 1e-9; state[1] = 1e9; state[2] = 0.5 * (1e-9 + 1e9); state[3] = 0.0; state[4] = 0.0;
}

int launch_oc(int64_t n, const uint8_t* design, const double* x, const double* dc, double Ve,
              double target, double move, double* state, double* partials, unsigned int* counter,
              double* xnew, int max_steps, hipStream_t st) {
  hipLaunchKernelGGL(k_oc_init, dim3(1), dim3(1), 0, st, state);
  const unsigned nb = nblk(n, 512);
  for (int s = 0; s < max_steps; ++s)
    hipLaunchKernelGGL(k_oc_step, dim3(nb), dim3(256), 0, st, n, design, x, dc, Ve, target, move,
                       state, partials, counter);
  hipLaunchKernelGGL(k_oc_apply, dim3(nblk(n, 4096)), dim3(256), 0, st, n, design, x, dc, Ve, move,
                     (const double*)state, xnew);
  VOX8_HIP(hipGetLastError());
  return OK;
}

// ---- small reductions -------------------------------------------------------------
// out[0] = sum_mask a, out[1] = sum_mask |a|, out[2] = max_mask |a - b| (b may be null),
// out[3] = count(mask)
__global__ void __launch_bounds__(256)
k_stats(int64_t n, const uint8_t* __restrict__ mask, const double* __restrict__ a,
        const double* __restrict__ b, double* __restrict__ partials, unsigned int* __restrict__ counter,
        double* __restrict__ out, const double* __restrict__ gate) {
  if (gate && !(*gate > 0.0)) return;     // uniform over the grid: the counter is untouched
  __shared__ double sh[8], sv[4], smax[4];
  double s = 0.0, sa = 0.0, mx = 0.0, cnt = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x) {
    if (mask && !mask[e]) continue;
    const double v = a[e];
    s += v;
    sa += fabs(v);
    cnt += 1.0;
    if (b) mx = fmax(mx, fabs(v - b[e]));
  }
  // max via wave/block reduction (max is order independent)
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) smax[w] = mx;
  const double t0 = block_sum(s, sh);
  const double t1 = block_sum(sa, sh);
  const double t3 = block_sum(cnt, sh);
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) m = fmax(m, smax[i]);
    sv[0] = t0; sv[1] = t1; sv[2] = m; sv[3] = t3;
  }
  __syncthreads();
  // combine: sums by the deterministic reduction; the max through a separate slot
  __shared__ int s_last;
  const int nb = gridDim.x;
  if (threadIdx.x == 0) {
    for (int v = 0; v < 4; ++v)
      __hip_atomic_store(&partials[(int64_t)blockIdx.x * 4 + v], sv[v], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    s_last = atomicAdd(counter, 1u) == (unsigned)(nb - 1);
  }
  __syncthreads();
  if (!s_last) return;
  __threadfence();
  double r[4] = {0, 0, 0, 0};
  for (int bb = threadIdx.x; bb < nb; bb += blockDim.x) {
    r[0] += __hip_atomic_load(&partials[(int64_t)bb * 4 + 0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    r[1] += __hip_atomic_load(&partials[(int64_t)bb * 4 + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    r[2] = fmax(r[2], __hip_atomic_load(&partials[(int64_t)bb * 4 + 2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
    r[3] += __hip_atomic_load(&partials[(int64_t)bb * 4 + 3], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  const double f0 = block_sum(r[0], sh), f1 = block_sum(r[1], sh), f3 = block_sum(r[3], sh);
  double m = r[2];
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_xor(m, off, 64));
  __syncthreads();
  if (lane == 0) smax[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) mm = fmax(mm, smax[i]);
    out[0] = f0; out[1] = f1; out[2] = mm; out[3] = f3;
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

int launch_stats(int64_t n, const uint8_t* mask, const double* a, const double* b, double* partials,
                 unsigned int* counter, double* out, hipStream_t st, const double* gate) {
  hipLaunchKernelGGL(k_stats, dim3(nblk(n, 512)), dim3(256), 0, st, n, mask, a, b, partials, counter,
                     out, gate);
  VOX8_HIP(hipGetLastError());
  return OK;
}

// out[j] = X_j . v, out[k+j] = X_j . X_j, out[2k] = v . v   (X: k column-major vectors)
__global__ void __launch_bounds__(256)
k_block_dots(int64_t n, int k, const double* __restrict__ X, int64_t ldx,
             const double* __restrict__ v, double* __restrict__ partials,
             unsigned int* __restrict__ counter, double* __restrict__ out,
             const double* const* __restrict__ Xind) {
  __shared__ double sh[8], sv[33];
  if (Xind) X = *Xind;
  double acc[33];
  const int nv = 2 * k + 1;
  for (int i = 0; i < 33; ++i) acc[i] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double b = v[i];
    for (int j = 0; j < k; ++j) {
      const double a = X[j * ldx + i];
      acc[j] += a * b;
      acc[k + j] += a * a;
    }
    acc[2 * k] += b * b;
  }
  for (int j = 0; j < nv; ++j) {
    const double t = block_sum(acc[j], sh);
    if (threadIdx.x == 0) sv[j] = t;
  }
  __syncthreads();
  grid_reduce_last_dyn(sv, nv, partials, counter, out, sh);
}

int launch_block_dots(int64_t n, int k, const double* X, int64_t ldx, const double* v,
                      double* partials, unsigned int* counter, double* out, hipStream_t st,
                      const double* const* Xind) {
  VOX8_CHECK(k >= 1 && k <= 16, "block_dots: 1..16 vectors");
  hipLaunchKernelGGL(k_block_dots, dim3(nblk(n, 512)), dim3(256), 0, st, n, k, X, ldx, v, partials,
                     counter, out, Xind);
  VOX8_HIP(hipGetLastError());
  return OK;
}

// y = a*x + b*y on the mask (outside: y unchanged)
__global__ void k_axpby(int64_t n, const uint8_t* __restrict__ mask, double a,
                        const double* __restrict__ x, double b, double* __restrict__ y) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    if (!mask || mask[e]) y[e] = a * x[e] + b * y[e];
}

int launch_axpby(int64_t n, const uint8_t* mask, double a, const double* x, double b, double* y,
                 hipStream_t st) {
  hipLaunchKernelGGL(k_axpby, dim3(nblk(n, 4096)), dim3(256), 0, st, n, mask, a, x, b, y);
  VOX8_HIP(hipGetLastError());
  return OK;
}

// ---- device-resident loop tail -------------------------------------------------
// state layout (doubles): [0] tracked mode (sorted position), [1] its eigenvalue,
// [2] violation, [3..6] stats of the filtered compliance sensitivity on the
// design, [7..10] stats of the filtered eigenvalue sensitivity, [11..14] stats
// of the change, [15..18] stats of the new design, [19] penalty coefficient.

// Mode tracking by MAC against the previous tracked mode (src/topopt.py:178-188,
// 327-339): dots = {X_j.v, X_j.X_j, v.v} of the eigenvalue-sorted modes; the
// first maximum wins (numpy argmax); no previous mode -> the lowest.
__global__ void k_tail_select(TailSelect s, const TailPar* __restrict__ par,
                              const double* __restrict__ dots, double* __restrict__ state) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int best = 0;
  if (s.have_prev && s.k > 0) {
    double bm = 0.0;
    for (int j = 0; j < s.k; ++j) {
      const double num = dots[j];
      const double mac = num * num / fmax(dots[s.k + j] * dots[2 * s.k], 1e-24);
      if (j == 0 || mac > bm) { bm = mac; best = j; }
    }
  }
  const double lam1 = s.k > 0 ? par->lam[best] : 0.0;
  for (int i = 3; i < TAIL_STATE; ++i) state[i] = 0.0;   // gated stages leave zeros
  state[0] = (double)best;
  state[1] = lam1;
  state[2] = s.k > 0 ? fmax(0.0, 1.0 - lam1 / s.w_target) : 0.0;
}

// v = X[:, state[0]] (the tracked mode becomes the next step's reference)
__global__ void k_gather_col(int64_t n, const TailPar* __restrict__ par, int64_t ldx,
                             const double* __restrict__ state, double* __restrict__ v) {
  const double* __restrict__ X = par->X;
  const int64_t col = (int64_t)state[0];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    v[i] = X[col * ldx + i];
}

// Augmented-Lagrangian frequency penalty (src/topopt.py:339-361): when the
// violation is positive, y += a x with a = -mu viol (s_c / max(s_m, 1e-30)),
// s_c and s_m the mean |.| of the filtered compliance and eigenvalue
// sensitivities over the design (both divided by the design count).
__global__ void k_tail_penalty(int64_t n, const TailPar* __restrict__ par, double* __restrict__ state,
                               const double* __restrict__ x, double b, double* __restrict__ y) {
  const double mu = par->mu;
  const double viol = state[2];
  if (!(viol > 0.0)) return;
  const double cnt = state[6];
  const double s_c = state[4] / cnt, s_m = state[8] / cnt;
  const double a = -mu * viol * (s_c / fmax(s_m, 1e-30));
  if (blockIdx.x == 0 && threadIdx.x == 0) state[19] = a;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    y[e] = a * x[e] + b * y[e];
}

int launch_tail_select(const TailSelect& s, const TailPar* par, const double* dots, double* state,
                       hipStream_t st) {
  hipLaunchKernelGGL(k_tail_select, dim3(1), dim3(64), 0, st, s, par, dots, state);
  VOX8_HIP(hipGetLastError());
  return OK;
}

int launch_gather_col(int64_t n, const TailPar* par, int64_t ldx, const double* state, double* v,
                      hipStream_t st) {
  hipLaunchKernelGGL(k_gather_col, dim3(nblk(n, 4096)), dim3(256), 0, st, n, par, ldx, state, v);
  VOX8_HIP(hipGetLastError());
  return OK;
}

int launch_tail_penalty(int64_t n, const TailPar* par, double* state, const double* x, double* y,
                        hipStream_t st) {
  hipLaunchKernelGGL(k_tail_penalty, dim3(nblk(n, 4096)), dim3(256), 0, st, n, par, state, x, 1.0, y);
  VOX8_HIP(hipGetLastError());
  return OK;
}

// Composite physical density (src/topopt.py:273-281):
// out = design ? filtered : (in_arm ? 1 : 0)
__global__ void k_compose(int64_t n, const uint8_t* __restrict__ design,
                          const uint8_t* __restrict__ in_arm, const double* __restrict__ filt,
                          double* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = design[e] ? filt[e] : (in_arm[e] ? 1.0 : 0.0);
}

int launch_compose(int64_t n, const uint8_t* design, const uint8_t* in_arm, const double* filt,
                   double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_compose, dim3(nblk(n, 4096)), dim3(256), 0, st, n, design, in_arm, filt, out);
  VOX8_HIP(hipGetLastError());
  return OK;
}

}  // namespace vox8

// ---- modal effective mass (sims/modal_participation.py:37-57) ------------------
namespace vox8 {

// r = unit rigid translation in direction dir on the free DOFs (zero at fixed DOFs)
__global__ void k_unit_translation(int64_t n_nodes, int dir, const uint8_t* __restrict__ fixed,
                                   double* __restrict__ r) {
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < n_nodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t f = fixed[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) r[3 * n + c] = (c == dir && !((f >> c) & 1)) ? 1.0 : 0.0;
  }
}

// Block partials, column j = blockIdx.y of MX (ld = 3 n_nodes, zero at fixed
// DOFs): L_j = sum of its direction-dir entries (= X_j^T M r, M symmetric,
// r the free-DOF translation) and, for j < k, g_j = X_j . (M X_j).  Column k
// holds M r, whose L is r^T M r.
__global__ void __launch_bounds__(256)
k_em_partials(int64_t n_nodes, int dir, int k, const double* __restrict__ X, int64_t ldx,
              const double* __restrict__ MX, double* __restrict__ partials) {
  __shared__ double sh[8];
  const int col = blockIdx.y;
  const double* y = MX + (int64_t)col * 3 * n_nodes;
  const double* x = col < k ? X + (int64_t)col * ldx : nullptr;
  double s1 = 0.0, s2 = 0.0;
  for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < n_nodes;
       n += (int64_t)gridDim.x * blockDim.x) {
    const double y0 = y[3 * n], y1 = y[3 * n + 1], y2 = y[3 * n + 2];
    s1 += dir == 0 ? y0 : (dir == 1 ? y1 : y2);
    if (x) s2 += x[3 * n] * y0 + x[3 * n + 1] * y1 + x[3 * n + 2] * y2;
  }
  double t[2];
  t[0] = block_sum(s1, sh);
  t[1] = block_sum(s2, sh);
  write_partials<2>(t, partials + (int64_t)col * gridDim.x * 2, blockIdx.x);
}

// out[j] = L_j^2 / max(g_j, 1e-30) / (r^T M r), out[k + j] = L_j,
// out[2k + j] = g_j, out[3k] = r^T M r; fixed-order sums (one block).
__global__ void __launch_bounds__(256)
k_em_finish(int nb, int k, const double* __restrict__ partials, double* __restrict__ out) {
  __shared__ double sh[8];
  double tot[2];
  sum_partials<2>(partials + (int64_t)k * nb * 2, nb, tot, sh);
  const double total = tot[0];
  for (int j = 0; j < k; ++j) {
    double t[2];
    sum_partials<2>(partials + (int64_t)j * nb * 2, nb, t, sh);
    if (threadIdx.x == 0) {
      out[j] = total > 0.0 ? t[0] * t[0] / fmax(t[1], 1e-30) / total : 0.0;
      out[k + j] = t[0];
      out[2 * k + j] = t[1];
    }
  }
  if (threadIdx.x == 0) out[3 * k] = total;
}

int launch_unit_translation(int64_t n_nodes, int dir, const uint8_t* fixed, double* r,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_unit_translation, dim3(nblk(n_nodes, 4096)), dim3(256), 0, st, n_nodes, dir,
                     fixed, r);
  VOX8_HIP(hipGetLastError());
  return OK;
}

int launch_effective_mass(int64_t n_nodes, int dir, int k, const double* X, int64_t ldx,
                          const double* MX, double* partials, int max_doubles, double* out,
                          hipStream_t st) {
  VOX8_CHECK(k >= 1 && k <= 16 && dir >= 0 && dir <= 2, "effective mass: 1..16 modes, direction 0..2");
  int nb = (int)nblk(n_nodes, 512);
  if ((int64_t)nb * 2 * (k + 1) > max_doubles) nb = max_doubles / (2 * (k + 1));
  VOX8_CHECK(nb >= 1, "effective mass: reduction workspace too small");
  hipLaunchKernelGGL(k_em_partials, dim3(nb, k + 1), dim3(256), 0, st, n_nodes, dir, k, X, ldx, MX,
                     partials);
  hipLaunchKernelGGL(k_em_finish, dim3(1), dim3(256), 0, st, nb, k, (const double*)partials, out);
  VOX8_HIP(hipGetLastError());
  return OK;
}

}  // namespace vox8
