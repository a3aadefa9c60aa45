// vox8 — element stress recovery and stress invariants (post-processing of a
// static solve, e.g. the fatigue evaluation of a design).
//
// Replaces FEM3D.compute_stress, von_mises and principal_stress_max
// (src/fem.py:564-602).  The reference gathers every element's 24 DOFs on the
// host and loops over elements for the principal stress (numpy eigvalsh per
// element); here one thread per element does the gather, the 6x24 and 6x6
// products and both invariants in FP64.
#include "vox8_kernels.h"
#include "vox8_post.h"

namespace vox8 {

// stress_e = rho_e D (B u_e) at the element centre, element-grid order; 0 on
// inactive elements.  u_e is in the reference's local node order
// (src/mesh3d.py:140-146), the order of B's columns.
__global__ void __launch_bounds__(256)
k_stress(Grid g, StressCoef sc, const double* __restrict__ u, const double* __restrict__ rho,
         const uint8_t* __restrict__ active, double* __restrict__ out) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < g.n_elem;
       e += (int64_t)gridDim.x * blockDim.x) {
    double s[6] = {0, 0, 0, 0, 0, 0};
    if (active[e]) {
      const int64_t ex = e % g.nx, ey = (e / g.nx) % g.ny, ez = e / ((int64_t)g.nx * g.ny);
      double ue[24];
#pragma unroll
      for (int a = 0; a < 8; ++a) {
        const int64_t n = (ex + local_off(a, 0)) +
                          (int64_t)g.NX * ((ey + local_off(a, 1)) + (int64_t)g.NY * (ez + local_off(a, 2)));
#pragma unroll
        for (int c = 0; c < 3; ++c) ue[3 * a + c] = u[3 * n + c];
      }
      double eps[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 24; ++k) t += sc.B[24 * i + k] * ue[k];
        eps[i] = t;
      }
      const double r = rho[e];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        double t = 0.0;
#pragma unroll
        for (int k = 0; k < 6; ++k) t += sc.D[6 * i + k] * eps[k];
        s[i] = r * t;
      }
    }
#pragma unroll
    for (int i = 0; i < 6; ++i) out[6 * e + i] = s[i];
  }
}

// Largest eigenvalue of the symmetric 3x3 [[a, d, f], [d, b, e], [f, e, c]]
// by the trigonometric solution of the characteristic cubic (Smith, 1961):
// with q = tr/3, p = sqrt(tr((A - qI)^2)/6) and B = (A - qI)/p, the
// eigenvalues are q + 2p cos(phi + 2 pi k/3), phi = acos(det(B)/2)/3; k = 0
// gives the largest.
__device__ double max_eig_sym3(double a, double b, double c, double d, double e, double f) {
  const double q = (a + b + c) / 3.0;
  const double aq = a - q, bq = b - q, cq = c - q;
  const double p2 = aq * aq + bq * bq + cq * cq + 2.0 * (d * d + e * e + f * f);
  if (!(p2 > 0.0)) return q;                       // multiple of the identity
  const double p = sqrt(p2 / 6.0), ip = 1.0 / p;
  const double ba = aq * ip, bb = bq * ip, bc = cq * ip, bd = d * ip, be = e * ip, bf = f * ip;
  const double det = ba * (bb * bc - be * be) - bd * (bd * bc - be * bf) + bf * (bd * be - bb * bf);
  const double r = fmin(1.0, fmax(-1.0, 0.5 * det));
  return q + 2.0 * p * cos(acos(r) / 3.0);
}

// von Mises and largest principal stress of n Voigt rows [xx, yy, zz, xy, yz, xz].
__global__ void __launch_bounds__(256)
k_stress_inv(int64_t n, const double* __restrict__ s, double* __restrict__ vm,
             double* __restrict__ pmax) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double xx = s[6 * i], yy = s[6 * i + 1], zz = s[6 * i + 2];
    const double xy = s[6 * i + 3], yz = s[6 * i + 4], xz = s[6 * i + 5];
    if (vm) {
      const double v2 = 0.5 * ((xx - yy) * (xx - yy) + (yy - zz) * (yy - zz) + (zz - xx) * (zz - xx)) +
                        3.0 * (xy * xy + yz * yz + xz * xz);
      vm[i] = sqrt(fmax(v2, 0.0));
    }
    if (pmax) pmax[i] = max_eig_sym3(xx, yy, zz, xy, yz, xz);
  }
}

static unsigned blocks_for(int64_t n) {
  const int64_t b = (n + 255) / 256;
  return (unsigned)(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

int launch_stress(const Grid& g, const StressCoef& sc, const double* u, const double* rho,
                  const uint8_t* active, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_stress, dim3(blocks_for(g.n_elem)), dim3(256), 0, st, g, sc, u, rho, active, out);
  VOX8_HIP(hipGetLastError());
  return OK;
}

int launch_stress_invariants(int64_t n, const double* s, double* vm, double* pmax, hipStream_t st) {
  if (n <= 0) return OK;
  hipLaunchKernelGGL(k_stress_inv, dim3(blocks_for(n)), dim3(256), 0, st, n, s, vm, pmax);
  VOX8_HIP(hipGetLastError());
  return OK;
}

}  // namespace vox8
