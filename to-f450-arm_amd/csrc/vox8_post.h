// vox8 — post-processing kernels (element stresses, stress invariants).
#pragma once
#include "vox8_common.h"

namespace vox8 {

// Centre strain-displacement matrix (6x24, row-major, columns in the
// element's local DOF order) and constitutive matrix (6x6), by value.
struct StressCoef {
  double B[6 * 24];
  double D[36];
};

int launch_stress(const Grid& g, const StressCoef& sc, const double* u, const double* rho,
                  const uint8_t* active, double* out, hipStream_t st);
int launch_stress_invariants(int64_t n, const double* s, double* vm, double* pmax, hipStream_t st);
// Goodman / S-N fields; h_xs / h_ys the 7-point log10 table (stress ascending).
int launch_fatigue(int64_t n, const double* s1, const double* s2, double uts, const double* h_xs,
                   const double* h_ys, double endurance_mpa, double overload_mpa, double fos_ref_mpa,
                   double* mean, double* alt, double* eq, double* life, double* damage, double* fos,
                   hipStream_t st);

}  // namespace vox8
