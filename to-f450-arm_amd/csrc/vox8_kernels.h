// vox8 — kernel launchers shared between translation units.
#pragma once
#include "vox8_common.h"

namespace vox8 {

enum Op : int { OP_K = 1, OP_M = 2, OP_KM = 3 };

// Optional fusion of a per-column dot product into a stencil launch (PCG p.Kp).
struct DotFuse {
  PcgScalars* scal = nullptr;     // per-column state; null = plain product
  double* partials = nullptr;     // [ncols * nblocks]
  unsigned int* counters = nullptr;  // [ncols]
};

// ---- stencil (vox8_stencil_dense.hip) ----
int launch_element_factors(const Grid& g, const double* rho, const uint8_t* active, double p,
                           double emin, double mthr, double mq, double* fK, double* fM,
                           float* fK32, float* fM32, hipStream_t st,
                           unsigned int* changed = nullptr);
template <typename T>
int launch_apply_dense(const Grid& g, int op, const T* fK, const T* fM, const T* Ke,
                       const T* Me, const uint8_t* fixed, const T* x, int64_t ldx, T* yK,
                       T* yM, int64_t ldy, int nrhs, hipStream_t st, DotFuse fuse = DotFuse());
template <typename T>
int launch_tip_mass(const int64_t* nodes, int n_tip, double m, const uint8_t* fixed, const T* x,
                    int64_t ldx, T* y, int64_t ldy, int nrhs, hipStream_t st);
int launch_jacobi(const Grid& g, const double* f, const double* Ae, const uint8_t* fixed,
                  int positive_guard, const int64_t* tip_nodes, int n_tip, double tip_m,
                  double* diag_out, double* dinv, hipStream_t st);

// ---- wave / block reductions (device helpers) ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum over a block of up to 1024 threads; result valid in every thread.
// `sh` must hold at least blockDim.x/64 doubles.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0.0;
  for (int i = 0; i < nw; ++i) t += sh[i];   // fixed order: deterministic
  return t;
}

// Block partial sums for a finishing kernel.  No grid-wide fence is needed:
// the kernel boundary orders the partials before the kernel that sums them.
// (A device-scope fence in every block is costly here: on this multi-XCD part
// it writes back and invalidates the L2 each time.)
template <int NV>
__device__ __forceinline__ void write_partials(const double* vals, double* partials, int bid) {
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) partials[(int64_t)bid * NV + v] = vals[v];
  }
}

// Fixed-order sum of nb block partials (stride NV) by one block: thread t sums
// blocks t, t + blockDim.x, ... in order, then block_sum.  Bitwise reproducible.
template <int NV>
__device__ __forceinline__ void sum_partials(const double* partials, int nb, double* out,
                                             double* sh) {
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) t += partials[(int64_t)b * NV + v];
    out[v] = block_sum(t, sh);
  }
}

// Finishing kernel of a fused K p . p product: scal[col].pq = sum of the nb
// partials of column col (partials + col * nb); grid (1, ncols), 256 threads.
int launch_pq_finish(const DotFuse& fuse, int nb, int ncols, hipStream_t st);
int launch_partials_finish(const double* partials, int nb, int ncols, double* out, hipStream_t st);

// "Last block finalises" grid reduction of NV values per block.  Each block
// passes its block sums in `vals` (thread 0 only reads them); partial sums go to
// partials[slot_base + blockIdx.x*NV + v].  Returns true in every thread of the
// last block to arrive, where `out[v]` then holds the total over blocks summed
// in a fixed order (bitwise reproducible).
template <int NV>
__device__ bool grid_reduce_last(const double* vals, double* partials, unsigned int* counter,
                                 int nblocks, int bid, double* out, double* sh) {
  __shared__ int s_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v)
      __hip_atomic_store(&partials[(int64_t)bid * NV + v], vals[v], __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    __threadfence();
    unsigned int prev = atomicAdd(counter, 1u);
    s_last = (prev == (unsigned)(nblocks - 1));
  }
  __syncthreads();
  if (!s_last) return false;
  __threadfence();
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x)
      t += __hip_atomic_load(&partials[(int64_t)b * NV + v], __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
    out[v] = block_sum(t, sh);
  }
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

}  // namespace vox8
