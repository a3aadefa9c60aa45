// vox8 — MFMA trial: K x in FP32 with the 24x24 element contraction on the
// matrix cores, for comparison with the structured stencil (k_apply_fast).
// Not used by the solvers; kept so the comparison can be re-measured.
//
// Tiles of 31 x 7 x 7 owned nodes.  A workgroup loads the 33 x 9 x 9 node box
// (owned nodes plus one halo layer) of one column into LDS, forms the
// contributions of the 32 x 8 x 8 elements touching its owned nodes (halo
// elements are recomputed by the neighbouring tiles, so no global atomics),
// accumulates them per node in LDS and writes the owned nodes.  Each wave takes
// x-rows of 32 elements: Y (24 x 32) = Ke (24 x 24) * (f_e U_e) (24 x 32) as a
// chain of 12 v_mfma_f32_32x32x2f32 (Ke rows padded to 32), then scatters its
// 24 x 32 outputs into the node accumulators with LDS atomic adds.  Lane layout
// (measured on gfx950): A[i][k] from lane i + 32 k, B[k][j] from lane j + 32 k,
// D[8 (r/4) + 4 (l/32) + r%4][l%32] in accumulator r of lane l.
// Output at fixed DOFs is zero (as the stencil).  The LDS additions are
// unordered, so the last bits of a node's sum can vary between runs.
#include "vox8_kernels.h"
#include "vox8_mfma_trial.h"

namespace vox8 {

typedef float v16f __attribute__((ext_vector_type(16)));

constexpr int MT_OX = 31, MT_OY = 7, MT_OZ = 7;                       // owned nodes
constexpr int MT_LX = MT_OX + 2, MT_LY = MT_OY + 2, MT_LZ = MT_OZ + 2;  // loaded box
constexpr int MT_EY = MT_OY + 1, MT_EZ = MT_OZ + 1;                     // element rows
constexpr int MT_NODES = MT_LX * MT_LY * MT_LZ;
constexpr int MT_THREADS = 256;
constexpr int MT_LOAD_IT = (MT_NODES * 3 + MT_THREADS - 1) / MT_THREADS;
constexpr int MT_GROUPS_PER_WAVE = MT_EY * MT_EZ / (MT_THREADS / 64);
static_assert(MT_EY * MT_EZ % (MT_THREADS / 64) == 0, "element rows split evenly over the waves");

__device__ __forceinline__ int mt_node_off(int b) {
  return local_off(b, 0) + MT_LX * (local_off(b, 1) + MT_LY * local_off(b, 2));
}

__global__ void __launch_bounds__(MT_THREADS)
k_apply_mfma_trial(Grid g, int tiles_x, int tiles_y, const float* __restrict__ Ke,
                   const float* __restrict__ fK, const uint8_t* __restrict__ fixed,
                   const float* __restrict__ x, float* __restrict__ y) {
  __shared__ float xs[MT_NODES * 3];
  __shared__ float acc[MT_NODES * 3];
  const int64_t col = blockIdx.y;
  x += col * g.n_dofs;
  y += col * g.n_dofs;
  int t = blockIdx.x;
  const int tx = t % tiles_x;
  t /= tiles_x;
  const int ty = t % tiles_y, tz = t / tiles_y;
  const int x0 = tx * MT_OX, y0 = ty * MT_OY, z0 = tz * MT_OZ;   // first owned node
  // node box [x0-1, x0+MT_OX+1) x [y0-1, ...) x [z0-1, ...): rows of MT_LX
  // nodes, 3 MT_LX contiguous values each.  All loads are issued before any
  // is consumed (out-of-box lanes read element 0 and discard it), so the
  // workgroup waits for HBM once, not once per value.
  float v[MT_LOAD_IT];
#pragma unroll
  for (int it = 0; it < MT_LOAD_IT; ++it) {
    const int i = threadIdx.x + it * MT_THREADS;
    const int row = i / (3 * MT_LX), w = i - row * (3 * MT_LX);
    const int gx = x0 - 1 + w / 3, gy = y0 - 1 + row % MT_LY, gz = z0 - 1 + row / MT_LY;
    const bool ok = i < MT_NODES * 3 && gx >= 0 && gx < g.NX && gy >= 0 && gy < g.NY && gz >= 0 &&
                    gz < g.NZ;
    const int64_t idx = ok ? 3 * (gx + (int64_t)g.NX * (gy + (int64_t)g.NY * gz)) + (w - 3 * (w / 3)) : 0;
    const float xv = x[idx];
    v[it] = ok ? xv : 0.0f;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ex = lane & 31, kk = lane >> 5;
  // element factors of this wave's element rows (grp = wave + 4 q), prefetched
  float fr[MT_GROUPS_PER_WAVE];
#pragma unroll
  for (int q = 0; q < MT_GROUPS_PER_WAVE; ++q) {
    const int grp = wave + q * (MT_THREADS / 64);
    const int gex = x0 - 1 + ex, gey = y0 - 1 + grp % MT_EY, gez = z0 - 1 + grp / MT_EY;
    const bool ok = gex >= 0 && gex < g.nx && gey >= 0 && gey < g.ny && gez >= 0 && gez < g.nz;
    const float fv = fK[ok ? gex + (int64_t)g.nx * (gey + (int64_t)g.ny * gez) : 0];
    fr[q] = ok ? fv : 0.0f;
  }
  float a[12];                       // A[i][k'] = Ke[i][2 s + k'], i = lane % 32 (< 24)
#pragma unroll
  for (int s = 0; s < 12; ++s) a[s] = ex < 24 ? Ke[ex * 24 + 2 * s + kk] : 0.0f;
#pragma unroll
  for (int it = 0; it < MT_LOAD_IT; ++it) {
    const int i = threadIdx.x + it * MT_THREADS;
    if (i < MT_NODES * 3) {
      xs[i] = v[it];
      acc[i] = 0.0f;
    }
  }
  __syncthreads();
  int boff[12], bcomp[12];           // B: local DOF 2 s + kk of the lane's element
#pragma unroll
  for (int s = 0; s < 12; ++s) {
    const int d = 2 * s + kk, b = d / 3;
    boff[s] = mt_node_off(b);
    bcomp[s] = d - 3 * b;
  }
#pragma unroll 2
  for (int q = 0; q < MT_GROUPS_PER_WAVE; ++q) {   // wave-uniform
    const int grp = wave + q * (MT_THREADS / 64);
    const int ey = grp % MT_EY, ez = grp / MT_EY;
    const float f = fr[q];
    const int base = ex + MT_LX * (ey + MT_LY * ez);
    v16f d = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f,
              0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int s = 0; s < 12; ++s) {
      const float bv = f * xs[3 * (base + boff[s]) + bcomp[s]];
      d = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], bv, d, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < 16; ++r) {
      const int i = 8 * (r >> 2) + 4 * kk + (r & 3);
      if (i < 24) {
        const int b = i / 3;
        atomicAdd(&acc[3 * (base + mt_node_off(b)) + (i - 3 * b)], d[r]);
      }
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 3 * MT_OX * MT_OY * MT_OZ; i += MT_THREADS) {
    const int row = i / (3 * MT_OX), w = i - row * (3 * MT_OX);
    const int oy = row % MT_OY, oz = row / MT_OY, ox = w / 3, c = w - 3 * ox;
    const int gx = x0 + ox, gy = y0 + oy, gz = z0 + oz;
    if (gx < g.NX && gy < g.NY && gz < g.NZ) {
      const int64_t gn = gx + (int64_t)g.NX * (gy + (int64_t)g.NY * gz);
      const float val = acc[3 * ((1 + ox) + MT_LX * ((1 + oy) + MT_LY * (1 + oz))) + c];
      y[3 * gn + c] = ((fixed[gn] >> c) & 1) ? 0.0f : val;
    }
  }
}

int launch_apply_mfma_trial(const Grid& g, const float* Ke, const float* fK, const uint8_t* fixed,
                            const float* x, float* y, int nrhs, hipStream_t st) {
  VOX8_CHECK(nrhs >= 1 && nrhs <= 65535, "nrhs out of range");
  const int tx = (g.NX + MT_OX - 1) / MT_OX, ty = (g.NY + MT_OY - 1) / MT_OY,
            tz = (g.NZ + MT_OZ - 1) / MT_OZ;
  hipLaunchKernelGGL(k_apply_mfma_trial, dim3((unsigned)(tx * ty * tz), (unsigned)nrhs),
                     dim3(MT_THREADS), 0, st, g, tx, ty, Ke, fK, fixed, x, y);
  VOX8_HIP(hipGetLastError());
  return OK;
}

}  // namespace vox8
