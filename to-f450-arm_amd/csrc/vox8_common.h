// vox8 — shared definitions for the gfx950 voxel hex8 kernels and the C-ABI.
//
// Layout in HBM (all device arrays):
//   nodes   n = ix + NX*(iy + NY*iz), NX = nx+1, NY = ny+1 (x fastest)
//   DOFs    3*n + c  (c = x,y,z)          -- identical to the reference layout, so
//                                            full-length vectors need no permutation
//   element e = ex + nx*(ey + ny*ez)     (x fastest; the reference's active order
//                                            for a full box)
//   blocks of k vectors are column-major: vector j starts at ptr + j*ld.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <string>
#include <vector>

namespace vox8 {

// Status codes returned by every C-ABI call.
enum Status : int {
  OK = 0,
  ERR_ARG = -1,
  ERR_HIP = -2,
  ERR_STATE = -3,
  ERR_NUMERIC = -4,
};

void set_error(const std::string& msg);

// Process-wide lock held around every HIP call that allocates, frees,
// creates or destroys a stream, event or graph, or waits on the host for the
// device (device, stream and event synchronisation), and around each stream
// capture.  Contexts may be driven from several host threads at once (one per
// design or per solve); the lock keeps those calls from overlapping another
// thread's capture, which the runtime does not allow.  Recursive: a locked
// section may call helpers that lock again.
std::recursive_mutex& hip_serial();
using HipSerial = std::lock_guard<std::recursive_mutex>;

#define VOX8_HIP(call)                                                        \
  do {                                                                        \
    hipError_t e_ = (call);                                                   \
    if (e_ != hipSuccess) {                                                   \
      ::vox8::set_error(std::string(#call " failed: ") + hipGetErrorString(e_) \
                        + " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
      return ::vox8::ERR_HIP;                                                 \
    }                                                                         \
  } while (0)

#define VOX8_CHECK(cond, msg)                                                 \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ::vox8::set_error(msg);                                                 \
      return ::vox8::ERR_ARG;                                                 \
    }                                                                         \
  } while (0)

#define VOX8_TRY(expr)            \
  do {                            \
    int s_ = (expr);              \
    if (s_ != 0) return s_;       \
  } while (0)

// Structured grid description passed by value to kernels.
struct Grid {
  int nx, ny, nz;      // elements per axis
  int NX, NY, NZ;      // nodes per axis
  int64_t n_nodes;
  int64_t n_dofs;
  int64_t n_elem;
};

// Node index -> (i, j, k) with 32-bit divisions (node counts are below 2^31,
// which vox8_create checks); 64-bit divisions cost several times as much.
#ifdef __HIPCC__
__device__ __forceinline__ void node_ijk(const Grid& g, int64_t n, int& i, int& j, int& k) {
  const uint32_t u = (uint32_t)n, t = u / (uint32_t)g.NX;
  i = (int)(u - t * (uint32_t)g.NX);
  k = (int)(t / (uint32_t)g.NY);
  j = (int)(t - (uint32_t)k * (uint32_t)g.NY);
}
#endif

inline Grid make_grid(int nx, int ny, int nz) {
  Grid g;
  g.nx = nx; g.ny = ny; g.nz = nz;
  g.NX = nx + 1; g.NY = ny + 1; g.NZ = nz + 1;
  g.n_nodes = (int64_t)g.NX * g.NY * g.NZ;
  g.n_dofs = 3 * g.n_nodes;
  g.n_elem = (int64_t)nx * ny * nz;
  return g;
}

// Reference local node order (bottom face CCW, then top face):
// local index -> (dx, dy, dz) offsets.
__host__ __device__ constexpr int local_off(int a, int axis) {
  // a: 0..7
  return axis == 2 ? (a >> 2)
       : axis == 0 ? ((a & 3) == 1 || (a & 3) == 2)
                   : ((a & 3) >= 2);
}
// (ox, oy, oz) -> local index
__host__ __device__ constexpr int local_idx(int ox, int oy, int oz) {
  return 4 * oz + (oy ? (ox ? 2 : 3) : (ox ? 1 : 0));
}

// Per-column scalar state of the batched Jacobi-PCG (one entry per RHS).
struct PcgScalars {
  double rr;        // ||r||^2
  double rz;        // r . D^-1 r   (current)
  double rz_prev;   // previous rz
  double pq;        // p . K p
  double atol2;     // (rtol * ||b||)^2
  double bnorm;     // ||b||
  int iters;        // iterations done
  int done;         // 1 converged, 0 running
  int pad0, pad1;
};

// Device-side reduction workspace for "last block finalises" reductions.
struct ReduceWs {
  double* partials;          // [max_blocks * max_vals]
  unsigned int* counter;     // one counter per concurrent reduction slot
  int max_blocks;
  int max_vals;
};

}  // namespace vox8
