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

// PA12 SLS S-N curve at R = -1 (src/fatigue.py:20-21): life table in log-log
// space, stress abscissae increasing (the repeated 12 MPa plateau value is
// separated by 1e-12, as the reference's interpolation table does).
struct SnTable {
  double xs[7], ys[7];   // log10(MPa), log10(cycles), ascending stress
  double endurance_mpa, overload_mpa, fos_ref_mpa;
};

// life of an equivalent alternating stress (src/fatigue.py:24-41): linear
// interpolation of log10(life) in log10(stress), clamped at the table ends;
// below 0.99x the endurance stress 1e9 cycles, above 1.5x the top stress 1.
__device__ double sn_life_dev(const SnTable& t, double s_pa) {
  const double s = s_pa / 1e6;
  const double ls = log10(s > 0.0 ? s : 1e-10);
  double ly;
  if (ls <= t.xs[0]) {
    ly = t.ys[0];
  } else if (ls >= t.xs[6]) {
    ly = t.ys[6];
  } else {
    int k = 0;
    while (k < 5 && ls >= t.xs[k + 1]) ++k;
    ly = t.ys[k] + (ls - t.xs[k]) * (t.ys[k + 1] - t.ys[k]) / (t.xs[k + 1] - t.xs[k]);
  }
  double life = pow(10.0, ly);
  if (s < t.endurance_mpa * 0.99) life = 1e9;
  if (s > t.overload_mpa * 1.5) life = 1.0;
  return life;
}

// Goodman / S-N field of two load-case stress fields (src/fatigue.py:44-121).
__global__ void __launch_bounds__(256)
k_fatigue(int64_t n, SnTable t, const double* __restrict__ s1, const double* __restrict__ s2,
          double uts, double* __restrict__ mean, double* __restrict__ alt, double* __restrict__ eq,
          double* __restrict__ life, double* __restrict__ damage, double* __restrict__ fos) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double a = s1[i], b = s2[i];
    const double peak = fmax(a, b), valley = fmin(a, b);
    const double m = 0.5 * (peak + valley), al = 0.5 * (peak - valley);
    const double e = al / fmax(1.0 - fmax(m, 0.0) / uts, 1e-6);
    const double L = sn_life_dev(t, e);
    if (mean) mean[i] = m;
    if (alt) alt[i] = al;
    if (eq) eq[i] = e;
    if (life) life[i] = L;
    if (damage) damage[i] = 1.0 / fmax(L, 1.0);
    if (fos) fos[i] = t.fos_ref_mpa * 1e6 / fmax(e, 1e-6);
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

int launch_fatigue(int64_t n, const double* s1, const double* s2, double uts, const double* h_xs,
                   const double* h_ys, double endurance_mpa, double overload_mpa, double fos_ref_mpa,
                   double* mean, double* alt, double* eq, double* life, double* damage, double* fos,
                   hipStream_t st) {
  if (n <= 0) return OK;
  SnTable t;
  for (int i = 0; i < 7; ++i) { t.xs[i] = h_xs[i]; t.ys[i] = h_ys[i]; }
  t.endurance_mpa = endurance_mpa;
  t.overload_mpa = overload_mpa;
  t.fos_ref_mpa = fos_ref_mpa;
  hipLaunchKernelGGL(k_fatigue, dim3(blocks_for(n)), dim3(256), 0, st, n, t, s1, s2, uts, mean, alt,
                     eq, life, damage, fos);
  VOX8_HIP(hipGetLastError());
  return OK;
}

}  // namespace vox8
