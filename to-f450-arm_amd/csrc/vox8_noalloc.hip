// vox8 — allocation-free entry point: block dot products with a fixed,
// library-owned reduction workspace.
//
// vox8_block_dots reduces through the context's workspace and reads the result
// back to the host.  Callers that cannot allocate (work recorded under HIP
// graph capture, serving loops) get vox8_block_dots_noalloc instead: no
// context, the per-block partial sums go to `scratch`, a fixed device array of
// this library (one copy per GPU, held until the process ends), and the result
// stays on the device.  Every call on a GPU shares that GPU's array, so calls
// that overlap on one GPU race; nothing in it needs clearing, as every call
// writes each partial it reads.
#include <mutex>

#include "../../include/vox8.h"
#include "vox8_kernels.h"

namespace vox8 {

constexpr int NOALLOC_MAX_K = 16;                       // vectors per call
constexpr int NOALLOC_BLOCKS = 512;                     // reduction blocks
constexpr int NOALLOC_VALS = 2 * NOALLOC_MAX_K + 1;

static __device__ __align__(256) double scratch[NOALLOC_BLOCKS * NOALLOC_VALS];

// Per-block partials of {X_j . v (k), X_j . X_j (k), v . v}, nv = 2k + 1 values
// per block at part[block * nv + j].
__global__ void __launch_bounds__(256)
k_bdots_part(int64_t n, int k, const double* __restrict__ X, int64_t ldx,
             const double* __restrict__ v, double* __restrict__ part) {
  __shared__ double sh[8];
  double acc[NOALLOC_VALS];
#pragma unroll
  for (int j = 0; j < NOALLOC_VALS; ++j) acc[j] = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double b = v[i];
#pragma unroll
    for (int j = 0; j < NOALLOC_MAX_K; ++j) {
      if (j < k) {
        const double a = X[j * ldx + i];
        acc[j] += a * b;
        acc[NOALLOC_MAX_K + j] += a * a;
      }
    }
    acc[2 * NOALLOC_MAX_K] += b * b;
  }
  const int nv = 2 * k + 1;
#pragma unroll
  for (int j = 0; j < NOALLOC_VALS; ++j) {
    const int slot = j < NOALLOC_MAX_K ? (j < k ? j : -1)
                   : (j < 2 * NOALLOC_MAX_K ? (j - NOALLOC_MAX_K < k ? k + j - NOALLOC_MAX_K : -1)
                                            : 2 * k);
    if (slot < 0) continue;
    const double t = block_sum(acc[j], sh);
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * nv + slot] = t;
  }
}

// out[j] = fixed-order sum of the nb block partials of value j (one block).
__global__ void __launch_bounds__(256)
k_bdots_finish(int nb, int nv, const double* __restrict__ part, double* __restrict__ out) {
  __shared__ double sh[8];
  for (int j = 0; j < nv; ++j) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) t += part[(int64_t)b * nv + j];
    t = block_sum(t, sh);
    if (threadIdx.x == 0) out[j] = t;
  }
}

// Device address of `scratch` on the current device, looked up once per
// device (hipGetSymbolAddress allocates nothing).
static int scratch_ptr(double** out) {
  static std::mutex mu;
  static double* ptrs[64] = {};
  int dev = 0;
  VOX8_HIP(hipGetDevice(&dev));
  VOX8_CHECK(dev >= 0 && dev < 64, "device index out of range");
  std::lock_guard<std::mutex> lock(mu);
  if (!ptrs[dev]) {
    void* p = nullptr;
    VOX8_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    VOX8_CHECK(p != nullptr, "hipGetSymbolAddress returned no address for the workspace");
    ptrs[dev] = static_cast<double*>(p);
  }
  *out = ptrs[dev];
  return OK;
}

}  // namespace vox8

using namespace vox8;

extern "C" int vox8_block_dots_noalloc(void* hip_stream, int64_t n, int k, const double* d_X,
                                       int64_t ld, const double* d_v, double* d_out) {
  VOX8_CHECK(d_X && d_v && d_out && n > 0 && ld >= n, "bad argument");
  if (k < 1 || k > NOALLOC_MAX_K) {
    set_error("vox8_block_dots_noalloc: k must be 1.." + std::to_string(NOALLOC_MAX_K) +
              " (the fixed workspace holds the partial sums of at most that many vectors)");
    return ERR_ARG;
  }
  double* part = nullptr;
  VOX8_TRY(scratch_ptr(&part));
  const hipStream_t st = (hipStream_t)hip_stream;
  int64_t nb = (n + 255) / 256;
  if (nb > NOALLOC_BLOCKS) nb = NOALLOC_BLOCKS;
  hipLaunchKernelGGL(k_bdots_part, dim3((unsigned)nb), dim3(256), 0, st, n, k, d_X, ld, d_v, part);
  hipLaunchKernelGGL(k_bdots_finish, dim3(1), dim3(256), 0, st, (int)nb, 2 * k + 1,
                     (const double*)part, d_out);
  VOX8_HIP(hipGetLastError());
  return OK;
}
