// vox8 — MFMA trial of the element contraction (vox8_mfma_trial.hip).
#pragma once
#include "vox8_common.h"

namespace vox8 {
// y = K x (FP32, nrhs column-major vectors of n_dofs) with the element
// contraction on the matrix cores; Ke row-major 24 x 24 FP32, fK the element
// factors, fixed the per-node fixed-DOF bits.
int launch_apply_mfma_trial(const Grid& g, const float* Ke, const float* fK, const uint8_t* fixed,
                            const float* x, float* y, int nrhs, hipStream_t st);
}  // namespace vox8
