// vox8 — geometric multigrid V-cycle preconditioner for K on the voxel grid.
//
// The reference preconditions its CG and LOBPCG with the Jacobi diagonal only
// (src/fem.py:441-443, 543-547; src/fem_gpu.py:61-62, 140-141).  Both solvers'
// stopping tests are unchanged when a stronger preconditioner is used; only
// the iteration count drops.  This V-cycle is a fixed symmetric operator:
//   * levels: the box grid coarsened 2:1 in every axis (ceil for odd counts),
//     coarse element = box of twice the size, so its element matrix is 2^l Ke
//     (exact for any constant material), factor = mean of the 8 children
//     (rediscretisation);
//   * transfer: trilinear prolongation P, restriction R = P^T, both masked by
//     the fixed DOFs of their level (a coarse DOF is fixed when a fine node of
//     its interpolation stencil has it fixed);
//   * smoother: Chebyshev polynomial of D^-1 K on [lam_max/alpha, lam_max]
//     (three-term recurrence on the iterates, Saad, Iterative Methods,
//     sec. 12.1), nu steps before and after; lam_max from Lanczos in the D
//     inner product, +10 % margin;
//   * coarsest level: dense inverse (in-place Gauss-Jordan of the SPD matrix).
// Smoothing and transfers run in FP32 on the structured stencil with fused
// Chebyshev epilogues; the outer Krylov method stays FP64.
#include <cmath>
#include <algorithm>
#include <cstdlib>


#include "vox8_mg.h"

namespace vox8 {

struct Diag24 { double v[24]; };

static unsigned nb(int64_t n, int cap = 2048) {
  int64_t b = (n + 255) / 256;
  return (unsigned)(b < cap ? (b < 1 ? 1 : b) : cap);
}

__global__ void k_mg_factors(Grid gf, Grid gc, const double* __restrict__ ff,
                             const uint8_t* __restrict__ af, double* __restrict__ fc64,
                             float* __restrict__ fc32, uint8_t* __restrict__ ac) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= gc.n_elem) return;
  const int X = (int)(e % gc.nx), Y = (int)((e / gc.nx) % gc.ny), Z = (int)(e / ((int64_t)gc.nx * gc.ny));
  double s = 0.0;
  uint8_t act = 0;
  for (int k = 0; k < 8; ++k) {
    const int x = 2 * X + (k & 1), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k >> 2);
    if (x < gf.nx && y < gf.ny && z < gf.nz) {
      const int64_t f = x + (int64_t)gf.nx * (y + (int64_t)gf.ny * z);
      s += ff[f];
      act |= af[f];
    }
  }
  fc64[e] = 0.125 * s;
  fc32[e] = (float)(0.125 * s);
  ac[e] = act;
}

__global__ void k_mg_jacobi(Grid g, const double* __restrict__ f, Diag24 dg,
                            const uint8_t* __restrict__ fixed, float* __restrict__ dinv) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes) return;
  const int ix = (int)(n % g.NX), iy = (int)((n / g.NX) % g.NY), iz = (int)(n / ((int64_t)g.NX * g.NY));
  double d[3] = {0, 0, 0};
  for (int s = 0; s < 8; ++s) {
    const int sx = s & 1, sy = (s >> 1) & 1, sz = s >> 2;
    const int ex = ix - 1 + sx, ey = iy - 1 + sy, ez = iz - 1 + sz;
    if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
    const double fe = f[ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez)];
    const int a = local_idx(1 - sx, 1 - sy, 1 - sz);
    for (int c = 0; c < 3; ++c) d[c] += fe * dg.v[3 * a + c];
  }
  const uint8_t fx = fixed[n];
  for (int c = 0; c < 3; ++c)
    dinv[3 * n + c] = ((fx >> c) & 1) ? 0.0f : (d[c] != 0.0 ? (float)(1.0 / d[c]) : 1.0f);
}


// A node's three components as one 12-byte access (global_load/store_dwordx3
// on 4-byte alignment): a wave then touches each cache line of the
// interleaved x, y, z layout once, where three strided scalar accesses each
// touched all of them.
struct F3 { float x, y, z; };
__device__ __forceinline__ F3 ld_f3(const float* p) { return *reinterpret_cast<const F3*>(p); }
__device__ __forceinline__ void st_f3(float* p, F3 v) { *reinterpret_cast<F3*>(p) = v; }

// coarse b = R r: weights 1 at the coincident fine node, 1/2 at the neighbours, per axis.
// Branch-free: a neighbour outside the fine grid is read at the clamped
// position and added with weight 0 (same terms, same order as skipping it).
__global__ void k_mg_restrict(Grid gf, Grid gc, const float* __restrict__ rf,
                              const uint8_t* __restrict__ fixc, float* __restrict__ bc,
                              const int* __restrict__ skip) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= gc.n_nodes) return;
  const int col = blockIdx.y;
  if (skip && skip[col]) return;
  rf += (int64_t)col * gf.n_dofs;
  bc += (int64_t)col * gc.n_dofs;
  int I, J, K;
  node_ijk(gc, n, I, J, K);
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
  for (int dz = -1; dz <= 1; ++dz) {
    const int z = 2 * K + dz;
    const bool okz = z >= 0 && z < gf.NZ;
    const int zc = min(max(z, 0), gf.NZ - 1);
    const float wz = okz ? (dz ? 0.5f : 1.0f) : 0.0f;
#pragma unroll
    for (int dy = -1; dy <= 1; ++dy) {
      const int y = 2 * J + dy;
      const bool oky = y >= 0 && y < gf.NY;
      const int yc = min(max(y, 0), gf.NY - 1);
      const float wy = oky ? wz * (dy ? 0.5f : 1.0f) : 0.0f;
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const int x = 2 * I + dx;
        const bool okx = x >= 0 && x < gf.NX;
        const int xc = min(max(x, 0), gf.NX - 1);
        const float w = okx ? wy * (dx ? 0.5f : 1.0f) : 0.0f;
        const F3 v = ld_f3(rf + 3 * (xc + (int64_t)gf.NX * (yc + (int64_t)gf.NY * zc)));
        s0 += w * v.x;
        s1 += w * v.y;
        s2 += w * v.z;
      }
    }
  }
  const uint8_t fx = fixc[n];
  st_f3(bc + 3 * n, F3{(fx & 1) ? 0.0f : s0, (fx & 2) ? 0.0f : s1, (fx & 4) ? 0.0f : s2});
}

// fine x += P coarse x (trilinear), fixed fine DOFs untouched.  Branch-free:
// on an axis where the fine node coincides with a coarse one the second coarse
// index equals the first and its term is not added.
__global__ void k_mg_prolong_add(Grid gc, Grid gf, const float* __restrict__ xc,
                                 const uint8_t* __restrict__ fixf, float* __restrict__ xf,
                                 const int* __restrict__ skip) {
  int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= gf.n_nodes) return;
  const int col = blockIdx.y;
  if (skip && skip[col]) return;
  xc += (int64_t)col * gc.n_dofs;
  xf += (int64_t)col * gf.n_dofs;
  int i, j, k;
  node_ijk(gf, n, i, j, k);
  const int x0 = i >> 1, x1 = (i + 1) >> 1, y0 = j >> 1, y1 = (j + 1) >> 1, z0 = k >> 1, z1 = (k + 1) >> 1;
  const float wx = (x0 == x1) ? 1.0f : 0.5f, wy = (y0 == y1) ? 1.0f : 0.5f, wz = (z0 == z1) ? 1.0f : 0.5f;
  float s0 = 0.f, s1 = 0.f, s2 = 0.f;
#pragma unroll
  for (int cz = 0; cz < 2; ++cz)
#pragma unroll
    for (int cy = 0; cy < 2; ++cy)
#pragma unroll
      for (int cx = 0; cx < 2; ++cx) {
        const bool use = (cz == 0 || z0 != z1) && (cy == 0 || y0 != y1) && (cx == 0 || x0 != x1);
        const int64_t c = (cx ? x1 : x0) + (int64_t)gc.NX * ((cy ? y1 : y0) + (int64_t)gc.NY * (cz ? z1 : z0));
        const F3 v = ld_f3(xc + 3 * c);
        s0 += use ? v.x : 0.0f;
        s1 += use ? v.y : 0.0f;
        s2 += use ? v.z : 0.0f;
      }
  const float w = wx * wy * wz;
  const uint8_t fx = fixf[n];
  float* p = xf + 3 * n;
  F3 o = ld_f3(p);
  o.x = (fx & 1) ? o.x : o.x + w * s0;
  o.y = (fx & 2) ? o.y : o.y + w * s1;
  o.z = (fx & 4) ? o.z : o.z + w * s2;
  st_f3(p, o);
}

// dense coarse matrix A[i][j] (identity rows/cols at fixed DOFs)
__global__ void k_mg_coarse_assemble(Grid g, const double* __restrict__ f, const uint8_t* __restrict__ fixed,
                                     const double* __restrict__ Ke, double scale, double* __restrict__ A) {
  const int64_t n = g.n_dofs;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * n;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = e / n, j = e % n;
    const int64_t ni = i / 3, nj = j / 3;
    const int ci = (int)(i % 3), cj = (int)(j % 3);
    if (((fixed[ni] >> ci) & 1) || ((fixed[nj] >> cj) & 1)) { A[e] = (i == j) ? 1.0 : 0.0; continue; }
    const int ix = (int)(ni % g.NX), iy = (int)((ni / g.NX) % g.NY), iz = (int)(ni / ((int64_t)g.NX * g.NY));
    const int jx = (int)(nj % g.NX), jy = (int)((nj / g.NX) % g.NY), jz = (int)(nj / ((int64_t)g.NX * g.NY));
    double s = 0.0;
    for (int q = 0; q < 8; ++q) {
      const int sx = q & 1, sy = (q >> 1) & 1, sz = q >> 2;
      const int ex = ix - 1 + sx, ey = iy - 1 + sy, ez = iz - 1 + sz;
      if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
      const int ox = jx - ex, oy = jy - ey, oz = jz - ez;
      if (ox < 0 || ox > 1 || oy < 0 || oy > 1 || oz < 0 || oz > 1) continue;
      const int a = local_idx(1 - sx, 1 - sy, 1 - sz), b = local_idx(ox, oy, oz);
      s += f[ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez)] * Ke[(3 * a + ci) * 24 + 3 * b + cj];
    }
    A[e] = scale * s;
  }
}

// Blocked Gauss-Jordan inverse of an SPD matrix (no pivoting).  For each
// block K of GJ_B consecutive pivots, with R the other indices and
// D = A_KK^-1:
//   A_KK <- D,  A_KR <- D A_KR,  A_RK <- -A_RK D,  A_RR <- A_RR - A_RK D A_KR,
// the scalar Gauss-Jordan step applied to a block; after the last block A
// holds its inverse.  Three launches per block: D by an in-LDS scalar
// Gauss-Jordan in one workgroup, W = A_:K D, then every entry from the
// previous iterate into the other buffer of a pair (no read/write hazard).
// The scalar pivots are the Schur-complement diagonals, positive for an SPD
// matrix; one below 1e-10 of the original diagonal marks the coarse matrix
// singular (flag = 1; the first block clears the flag).
constexpr int GJ_B = 32;

__global__ void k_gj_diag(int n, const double* __restrict__ A, double* __restrict__ diag0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) diag0[i] = fabs(A[(int64_t)i * n + i]);
}

__global__ void __launch_bounds__(256)
k_gjb_panel(int n, int k0, int b, const double* __restrict__ A, double* __restrict__ D,
            const double* __restrict__ diag0, double* __restrict__ flag) {
  __shared__ double m[2][GJ_B * GJ_B];
  __shared__ int bad;
  if (threadIdx.x == 0) bad = 0;
  for (int e = threadIdx.x; e < b * b; e += blockDim.x)
    m[0][e] = A[(int64_t)(k0 + e / b) * n + k0 + e % b];
  __syncthreads();
  int cur = 0;
  for (int p = 0; p < b; ++p) {
    const double* a = m[cur];
    double* o = m[cur ^ 1];
    const double app = a[p * b + p];
    const double q = 1.0 / app;
    for (int e = threadIdx.x; e < b * b; e += blockDim.x) {
      const int i = e / b, j = e % b;
      double v;
      if (i == p) v = (j == p) ? q : a[p * b + j] * q;
      else if (j == p) v = -a[i * b + p] * q;
      else v = a[e] - a[i * b + p] * a[p * b + j] * q;
      o[e] = v;
    }
    if (threadIdx.x == 0 && !(app > 1e-10 * diag0[k0 + p])) bad = 1;
    __syncthreads();
    cur ^= 1;
  }
  for (int e = threadIdx.x; e < b * b; e += blockDim.x) D[e] = m[cur][e];
  if (threadIdx.x == 0) {
    if (k0 == 0) *flag = bad ? 1.0 : 0.0;
    else if (bad) *flag = 1.0;
  }
}

// W[i][l] = sum_m A[i][k0 + m] D[m][l]
__global__ void k_gjb_w(int n, int k0, int b, const double* __restrict__ A, const double* __restrict__ D,
                        double* __restrict__ W) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * b) return;
  const int i = (int)(e / b), l = (int)(e % b);
  const double* ar = A + (int64_t)i * n + k0;
  double s = 0.0;
  for (int m = 0; m < b; ++m) s += ar[m] * D[m * b + l];
  W[e] = s;
}

// Every entry of the next iterate.  A thread owns column j of a group of
// GJ_ROWS rows (blockIdx.y): the panel rows A_K[:, j] are loaded once into
// registers and reused for every row of the group, while the row's W (or D)
// entries are the same for the whole wavefront (uniform loads); the sums run
// over m in the same order as a per-entry loop, so the values are unchanged.
constexpr int GJ_ROWS = 8;
__global__ void __launch_bounds__(256)
k_gjb_update(int n, int k0, int b, const double* __restrict__ A, double* __restrict__ B,
             const double* __restrict__ D, const double* __restrict__ W) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int i0 = blockIdx.y * GJ_ROWS;
  const bool jK = j >= k0 && j < k0 + b;
  double akj[GJ_B];
  if (!jK) {
#pragma unroll
    for (int m = 0; m < GJ_B; ++m) akj[m] = m < b ? A[(int64_t)(k0 + m) * n + j] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < GJ_ROWS; ++r) {
    const int i = i0 + r;
    if (i >= n) break;
    const bool iK = i >= k0 && i < k0 + b;
    double v;
    if (jK) {
      v = iK ? D[(i - k0) * b + (j - k0)] : -W[(int64_t)i * b + (j - k0)];
    } else if (iK) {
      v = 0.0;
#pragma unroll
      for (int m = 0; m < GJ_B; ++m)                    // b is uniform: a uniform guard
        if (m < b) v += D[(i - k0) * b + m] * akj[m];
    } else {
      v = A[(int64_t)i * n + j];
#pragma unroll
      for (int m = 0; m < GJ_B; ++m)
        if (m < b) v -= W[(int64_t)i * b + m] * akj[m];
    }
    B[(int64_t)i * n + j] = v;
  }
}

__global__ void k_to_f32_n(int64_t n, const double* __restrict__ a, float* __restrict__ b) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[i] = (float)a[i];
}

// x = Ainv b for up to CN columns at once (blockIdx.y selects the column
// group), so the inverse is read once per group (groups of 4 columns: more
// groups give more workgroups, which this latency-bound product needs more
// than fewer reads of the inverse).  A block computes 4 rows; its 4
// wavefronts split the n entries (wavefront w takes k = 64 w + lane + 256 m,
// coalesced reads of the rows and of the group's columns of b), reduce
// across their lanes, and wavefront 0 adds the 4 partial sums in a fixed
// order.
template <int CN>
__global__ void __launch_bounds__(256)
k_mg_coarse_apply(int n, int ncols, const float* __restrict__ Ainv, const float* __restrict__ b,
                  float* __restrict__ x, const int* __restrict__ skip) {
  __shared__ float part[4][4 * CN];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row0 = blockIdx.x * 4;
  const int c0 = blockIdx.y * CN;
  const int nc = ncols - c0 < CN ? ncols - c0 : CN;
  float s[4][CN];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < CN; ++c) s[i][c] = 0.0f;
#pragma unroll 4
  for (int k = 64 * w + lane; k < n; k += 256) {
    float bk[CN];
#pragma unroll
    for (int c = 0; c < CN; ++c) bk[c] = c < nc ? b[(int64_t)(c0 + c) * n + k] : 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const float a = row0 + i < n ? Ainv[(int64_t)(row0 + i) * n + k] : 0.0f;
#pragma unroll
      for (int c = 0; c < CN; ++c) s[i][c] += a * bk[c];
    }
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int c = 0; c < CN; ++c) {
      float v = s[i][c];
#pragma unroll
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
      if (lane == 0) part[w][i * CN + c] = v;
    }
  __syncthreads();
  if (threadIdx.x < 4 * CN) {
    const int i = threadIdx.x / CN, c = threadIdx.x % CN;
    const float v = ((part[0][threadIdx.x] + part[1][threadIdx.x]) + part[2][threadIdx.x]) +
                    part[3][threadIdx.x];
    const int col = c0 + c;
    if (row0 + i < n && c < nc && !(skip && skip[col])) x[(int64_t)col * n + row0 + i] = v;
  }
}

__global__ void k_f64_to_f32(int64_t n, const double* __restrict__ a, float* __restrict__ b,
                             const int* __restrict__ skip) {
  if (skip && skip[blockIdx.y]) return;
  const int64_t off = (int64_t)blockIdx.y * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[off + i] = (float)a[off + i];
}

__global__ void k_f32_to_f64(int64_t n, const float* __restrict__ a, double* __restrict__ b,
                             const int* __restrict__ skip) {
  if (skip && skip[blockIdx.y]) return;
  const int64_t off = (int64_t)blockIdx.y * n;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    b[off + i] = (double)a[off + i];
}

// lambda_max(D^-1 K) by Lanczos in the D inner product <a, b>_D = sum a_i b_i / dinv_i,
// in which D^-1 K is self-adjoint (fixed DOFs, dinv = 0, are outside the space).
// Lanczos vectors are FP32, every sum FP64; alpha_j, beta_j stay on the device
// (scal[slot], scal[slot + 1]) until the level's steps are done.

// start vector: a fixed hash of the DOF index (the estimate then depends only
// on the operator), zero on fixed DOFs; partial of ||q||_D^2
__global__ void __launch_bounds__(256)
k_lz_start(int64_t n, const float* __restrict__ dinv, float* __restrict__ q, double* __restrict__ part) {
  __shared__ double sh[8];
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    const float di = dinv[i];
    const float u = di > 0.0f ? (float)((z >> 40) * (1.0 / 16777216.0)) - 0.5f : 0.0f;
    q[i] = u;
    if (di > 0.0f) s += (double)u * u / di;
  }
  const double t = block_sum(s, sh);
  write_partials<1>(&t, part, blockIdx.x);
}

// scal[slot] = 1 / ||q||_D
__global__ void __launch_bounds__(256)
k_lz_start_finish(int nb, const double* __restrict__ part, double* __restrict__ scal, int slot) {
  __shared__ double sh[8];
  double tot;
  sum_partials<1>(part, nb, &tot, sh);
  if (threadIdx.x == 0) scal[slot] = tot > 0.0 ? 1.0 / sqrt(tot) : 0.0;
}

__global__ void k_lz_scale(int64_t n, const double* __restrict__ scal, int slot, float* __restrict__ q) {
  const float s = (float)scal[slot];
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    q[i] *= s;
}

// w = D^-1 (K q) - beta_prev q_prev, written over q_prev; partials of
// <w, q>_D, <w, w>_D, <q, q>_D.  beta_prev = null on the first step, where
// q_prev is not read (the buffer holds whatever its last user left).
__global__ void __launch_bounds__(256)
k_lz_update(int64_t n, const float* __restrict__ dinv, const float* __restrict__ kq,
            const float* __restrict__ q, float* __restrict__ qp, const double* __restrict__ beta_prev,
            double* __restrict__ part) {
  __shared__ double sh[8];
  const float bp = beta_prev ? (float)*beta_prev : 0.0f;
  double s1 = 0.0, s2 = 0.0, s3 = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const float di = dinv[i];
    if (di > 0.0f) {
      const float qi = q[i];
      const float w = beta_prev ? di * kq[i] - bp * qp[i] : di * kq[i];
      qp[i] = w;
      const double dd = 1.0 / (double)di;
      s1 += (double)w * qi * dd;
      s2 += (double)w * w * dd;
      s3 += (double)qi * qi * dd;
    } else {
      qp[i] = 0.0f;
    }
  }
  double t[3];
  t[0] = block_sum(s1, sh);
  t[1] = block_sum(s2, sh);
  t[2] = block_sum(s3, sh);
  write_partials<3>(t, part, blockIdx.x);
}

// alpha = <w, q>_D / <q, q>_D, beta = ||w - alpha q||_D -> scal[slot], scal[slot + 1]
__global__ void __launch_bounds__(256)
k_lz_finish(int nb, const double* __restrict__ part, double* __restrict__ scal, int slot) {
  __shared__ double sh[8];
  double t[3];
  sum_partials<3>(part, nb, t, sh);
  if (threadIdx.x == 0) {
    const double a = t[2] > 0.0 ? t[0] / t[2] : 0.0;
    const double b2 = t[2] > 0.0 ? t[1] - t[0] * t[0] / t[2] : 0.0;
    scal[slot] = a;
    scal[slot + 1] = b2 > 0.0 ? sqrt(b2) : 0.0;
  }
}

// next Lanczos vector (w - alpha q) / beta, in place over w
__global__ void k_lz_next(int64_t n, const double* __restrict__ scal, int slot,
                          const float* __restrict__ q, float* __restrict__ w) {
  const double a = scal[slot], b = scal[slot + 1];
  const float fa = (float)a, rb = b > 0.0 ? (float)(1.0 / b) : 0.0f;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    w[i] = (w[i] - fa * q[i]) * rb;
}

// Largest eigenvalue of the symmetric tridiagonal matrix diag(a[0..m)),
// off-diagonal b[0..m-1), by bisection on the Sturm count (Golub & Van Loan,
// Matrix Computations, sec. 8.4.1).
static double tridiag_max_eig(const double* a, const double* b, int m) {
  if (m <= 0) return 0.0;
  double lo = a[0], hi = a[0];
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  // count of eigenvalues < x
  auto below = [&](double x) {
    int c = 0;
    double d = 1.0;
    for (int i = 0; i < m; ++i) {
      const double bb = i > 0 ? b[i - 1] * b[i - 1] : 0.0;
      d = (a[i] - x) - (i > 0 ? bb / d : 0.0);
      if (d == 0.0) d = -1e-300;
      if (d < 0.0) ++c;
    }
    return c;
  };
  for (int it = 0; it < 200 && hi - lo > 1e-13 * std::max(std::fabs(lo), std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (below(mid) >= m) hi = mid; else lo = mid;
  }
  return hi;
}

// ---------------------------------------------------------------------------
static int level_apply(const MGLevel& L, const float* x, float* y, int ncols, hipStream_t st,
                       Epilogue<float> ep) {
  if (L.S32) return gal_apply(L.g, L.S32, L.s_by_entry, L.fixed, x, y, ncols, ep, st);
  return launch_apply_fast<float>(L.g, L.pattern, OP_K, L.c32, L.fK, nullptr, L.fixed, x,
                                  L.g.n_dofs, y, nullptr, L.g.n_dofs, ncols, st, DotFuse(), ep);
}

int mg_build(MultiGrid& mg, const Grid& g0, const uint8_t* h_active, const uint8_t* h_fixed,
             const double* Ke, int pattern0, const FastCoef<float>& c0, float* fK0_32,
             uint8_t* d_active0, uint8_t* d_fixed0, const MGAlloc& al, hipStream_t st) {
  VOX8_CHECK(pattern0 != PAT_DENSE, "multigrid needs the structured (box-element) stencil");
  mg.lv.clear();
  std::vector<uint8_t> act(h_active, h_active + g0.n_elem), fix(h_fixed, h_fixed + g0.n_nodes);
  Grid g = g0;
  double scale = 1.0;
  for (int l = 0; l < MG_MAX_LEVELS; ++l) {
    MGLevel L;
    L.g = g;
    L.scale = scale;
    L.pattern = pattern0;
    L.c32 = c0;
    for (int r = 0; r < 9; ++r)
      for (int q = 0; q < 9; ++q) { L.c32.lin[r][q] *= (float)scale; L.c32.bil[r][q] *= (float)scale; }
    for (int r = 0; r < 3; ++r)
      for (int q = 0; q < 3; ++q) L.c32.tri[r][q] *= (float)scale;
    for (int i = 0; i < 24; ++i) L.diag24[i] = scale * Ke[i * 25];
    void* p = nullptr;
    if (l == 0) {
      L.fK = fK0_32;
      L.active = d_active0;
      L.fixed = d_fixed0;
    } else {
      VOX8_TRY(al.dev(al.ctx, &p, sizeof(double) * g.n_elem)); L.fK64 = (double*)p;
      VOX8_TRY(al.dev(al.ctx, &p, sizeof(float) * g.n_elem)); L.fK = (float*)p;
      VOX8_TRY(al.dev(al.ctx, &p, g.n_elem)); L.active = (uint8_t*)p;
      VOX8_TRY(al.dev(al.ctx, &p, g.n_nodes)); L.fixed = (uint8_t*)p;
      VOX8_HIP(hipMemcpyAsync(L.active, act.data(), g.n_elem, hipMemcpyHostToDevice, st));
      VOX8_HIP(hipMemcpyAsync(L.fixed, fix.data(), g.n_nodes, hipMemcpyHostToDevice, st));
      VOX8_HIP(hipStreamSynchronize(st));   // act / fix change on the next level
      if (mg.galerkin) {
        if (!mg.operators_only) {
          VOX8_TRY(al.dev(al.ctx, &p, sizeof(double) * 576 * g.n_elem));
          L.E64 = (double*)p;
        }
        VOX8_TRY(al.dev(al.ctx, &p, sizeof(float) * 128 * g.n_nodes)); L.S32 = (float*)p;
        L.s_by_entry = g.n_nodes >= mg.lane_nodes;
      }
    }
    VOX8_TRY(al.dev(al.ctx, &p, sizeof(float) * g.n_dofs)); L.dinv = (float*)p;
    VOX8_TRY(al.dev(al.ctx, &p, sizeof(float) * g.n_dofs)); L.pv = (float*)p;
    mg.lv.push_back(L);
    // coarsen until the grid is tiny (dense inverse is then cheap) or a side
    // cannot be halved any more
    const bool small = g.n_dofs <= mg.coarse_dofs;
    if (small || g.nx < 2 || g.ny < 2 || g.nz < 2) break;
    // next level: masks on the host
    Grid gc = make_grid((g.nx + 1) / 2, (g.ny + 1) / 2, (g.nz + 1) / 2);
    std::vector<uint8_t> actc(gc.n_elem, 0), fixc(gc.n_nodes, 0), used(gc.n_nodes, 0);
    for (int Z = 0; Z < gc.nz; ++Z)
      for (int Y = 0; Y < gc.ny; ++Y)
        for (int X = 0; X < gc.nx; ++X) {
          uint8_t a = 0;
          for (int k = 0; k < 8; ++k) {
            const int x = 2 * X + (k & 1), y = 2 * Y + ((k >> 1) & 1), z = 2 * Z + (k >> 2);
            if (x < g.nx && y < g.ny && z < g.nz) a |= act[x + (int64_t)g.nx * (y + (int64_t)g.ny * z)];
          }
          actc[X + (int64_t)gc.nx * (Y + (int64_t)gc.ny * Z)] = a;
          if (a)
            for (int k = 0; k < 8; ++k)
              used[(X + (k & 1)) + (int64_t)gc.NX * ((Y + ((k >> 1) & 1)) + (int64_t)gc.NY * (Z + (k >> 2)))] = 1;
        }
    for (int K = 0; K < gc.NZ; ++K)
      for (int J = 0; J < gc.NY; ++J)
        for (int I = 0; I < gc.NX; ++I) {
          const int64_t c = I + (int64_t)gc.NX * (J + (int64_t)gc.NY * K);
          // a coarse DOF is fixed when any fine node it interpolates to (the
          // 3x3x3 fine neighbourhood of its position) has that DOF fixed, so a
          // clamp on odd-indexed nodes survives coarsening
          uint8_t f = 0;
          for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
              for (int dx = -1; dx <= 1; ++dx) {
                const int x = 2 * I + dx, y = 2 * J + dy, z = 2 * K + dz;
                if (x < 0 || y < 0 || z < 0 || x >= g.NX || y >= g.NY || z >= g.NZ) continue;
                f |= fix[x + (int64_t)g.NX * (y + (int64_t)g.NY * z)];
              }
          if (!used[c]) f = 7;
          fixc[c] = f;
        }
    act.swap(actc);
    fix.swap(fixc);
    g = gc;
    scale *= 2.0;
  }
  const MGLevel& C = mg.lv.back();
  mg.n_coarse = C.g.n_dofs <= mg.coarse_dofs ? (int)C.g.n_dofs : 0;
  void* p = nullptr;
  if (mg.n_coarse) {
    const size_t nn = (size_t)mg.n_coarse * mg.n_coarse;
    if (!mg.operators_only) {
      VOX8_TRY(al.dev(al.ctx, &p, sizeof(double) * nn)); mg.A = (double*)p;
      VOX8_TRY(al.dev(al.ctx, &p, sizeof(double) * nn)); mg.A2 = (double*)p;
    }
    VOX8_TRY(al.dev(al.ctx, &p, sizeof(float) * nn)); mg.Ainv = (float*)p;
  }
  VOX8_TRY(al.dev(al.ctx, &p, sizeof(double) * std::max<size_t>(6144, (size_t)mg.n_coarse * (GJ_B + 1)
                                                                       + GJ_B * GJ_B + 8)));
  mg.red = (double*)p;
  VOX8_TRY(al.dev(al.ctx, &p, sizeof(unsigned int) * 4)); mg.counter = (unsigned int*)p;
  VOX8_HIP(hipMemsetAsync(mg.counter, 0, sizeof(unsigned int) * 4, st));
  VOX8_TRY(al.dev(al.ctx, &p, sizeof(double) * MG_SCAL_SIZE)); mg.scal = (double*)p;
  VOX8_TRY(al.pinned(al.ctx, &p, sizeof(double) * MG_SCAL_SIZE)); mg.host = (double*)p;
  mg.ncols = 0;
  mg.built = true;
  mg.updated = false;
  ++mg.epoch;
  return OK;
}

int mg_reserve(MultiGrid& mg, int ncols, const MGAlloc& al) {
  if (mg.ncols >= ncols) return OK;
  for (auto& L : mg.lv) {
    void* p = nullptr;
    const size_t bytes = sizeof(float) * (size_t)ncols * L.g.n_dofs;
    VOX8_TRY(al.dev(al.ctx, &p, bytes)); L.x = (float*)p;
    VOX8_TRY(al.dev(al.ctx, &p, bytes)); L.r = (float*)p;
    VOX8_TRY(al.dev(al.ctx, &p, bytes)); L.b = (float*)p;
    VOX8_TRY(al.dev(al.ctx, &p, bytes)); L.d = (float*)p;
    VOX8_TRY(al.dev(al.ctx, &p, bytes)); L.d2 = (float*)p;
  }
  mg.ncols = ncols;
  ++mg.epoch;
  return OK;
}

int mg_copy_operators(MultiGrid& dst, const MultiGrid& src, hipStream_t st) {
  VOX8_CHECK(dst.built && src.built && src.updated, "shared multigrid: hierarchies not set up");
  bool same = dst.lv.size() == src.lv.size() && dst.n_coarse == src.n_coarse &&
              dst.galerkin == src.galerkin;
  for (size_t l = 0; same && l < dst.lv.size(); ++l) {
    const Grid &a = dst.lv[l].g, &b = src.lv[l].g;
    same = a.nx == b.nx && a.ny == b.ny && a.nz == b.nz && (dst.lv[l].S32 != nullptr) == (src.lv[l].S32 != nullptr);
  }
  VOX8_CHECK(same, "shared multigrid: the two contexts' hierarchies differ (grid, options or "
                   "coarse-level kind)");
  for (size_t l = 0; l < dst.lv.size(); ++l) {
    MGLevel& D = dst.lv[l];
    const MGLevel& S = src.lv[l];
    if (l > 0) {
      if (S.S32) {
        VOX8_HIP(hipMemcpyAsync(D.S32, S.S32, sizeof(float) * 128 * (size_t)D.g.n_nodes,
                                hipMemcpyDeviceToDevice, st));
      } else {
        VOX8_HIP(hipMemcpyAsync(D.fK64, S.fK64, sizeof(double) * (size_t)D.g.n_elem,
                                hipMemcpyDeviceToDevice, st));
        VOX8_HIP(hipMemcpyAsync(D.fK, S.fK, sizeof(float) * (size_t)D.g.n_elem, hipMemcpyDeviceToDevice,
                                st));
      }
    }
    VOX8_HIP(hipMemcpyAsync(D.dinv, S.dinv, sizeof(float) * (size_t)D.g.n_dofs, hipMemcpyDeviceToDevice,
                            st));
    D.lam_max = S.lam_max;
  }
  if (dst.n_coarse)
    VOX8_HIP(hipMemcpyAsync(dst.Ainv, src.Ainv, sizeof(float) * (size_t)dst.n_coarse * dst.n_coarse,
                            hipMemcpyDeviceToDevice, st));
  dst.updated = true;
  ++dst.updates;
  ++dst.epoch;
  return OK;
}

int mg_update(MultiGrid& mg, const double* fK0, const double* dKe, hipStream_t st) {
  // level factors and Jacobi diagonals
  for (size_t l = 0; l < mg.lv.size(); ++l) {
    MGLevel& L = mg.lv[l];
    const double* f64 = l == 0 ? fK0 : L.fK64;
    if (l > 0 && L.E64) {
      const MGLevel& F = mg.lv[l - 1];
      VOX8_TRY(gal_build_level(F.g, L.g, l == 1 ? fK0 : nullptr, dKe, l == 1 ? nullptr : F.E64,
                               F.fixed, L.fixed, L.E64, L.S32, L.s_by_entry, st));
      VOX8_TRY(gal_jacobi(L.g, L.E64, L.fixed, L.dinv, st));
      continue;
    }
    if (l > 0) {
      const MGLevel& F = mg.lv[l - 1];
      hipLaunchKernelGGL(k_mg_factors, dim3(nb(L.g.n_elem, 1 << 20)), dim3(256), 0, st, F.g, L.g,
                         (const double*)(l == 1 ? fK0 : F.fK64), (const uint8_t*)F.active, L.fK64, L.fK,
                         L.active);
      VOX8_HIP(hipGetLastError());
    }
    Diag24 dg;
    for (int i = 0; i < 24; ++i) dg.v[i] = L.diag24[i];
    hipLaunchKernelGGL(k_mg_jacobi, dim3(nb(L.g.n_nodes, 1 << 20)), dim3(256), 0, st, L.g, f64, dg,
                       (const uint8_t*)L.fixed, L.dinv);
    VOX8_HIP(hipGetLastError());
  }
  // lambda_max(D^-1 K) per smoothed level: MG_LANCZOS_STEPS of Lanczos from the
  // same start vector on every update, so the estimate is a function of the
  // current operator alone.  The largest Ritz value approaches lambda_max from
  // below (20 steps: within 1 % on SIMP density fields); +10 % margin.
  const size_t nsm = mg.n_coarse ? mg.lv.size() - 1 : mg.lv.size();
  for (size_t l = 0; l < nsm; ++l) {
    MGLevel& L = mg.lv[l];
    const int64_t n = L.g.n_dofs;
    const unsigned pb = nb(n, 1024);
    const int base = MG_SCAL_LZ + 2 * MG_LANCZOS_STEPS * (int)l;
    float* q = L.pv;
    float* qp = L.d;
    hipLaunchKernelGGL(k_lz_start, dim3(pb), dim3(256), 0, st, n, (const float*)L.dinv, q, mg.red);
    hipLaunchKernelGGL(k_lz_start_finish, dim3(1), dim3(256), 0, st, (int)pb, (const double*)mg.red,
                       mg.scal, MG_SCAL_NORM);
    hipLaunchKernelGGL(k_lz_scale, dim3(nb(n)), dim3(256), 0, st, n, (const double*)mg.scal,
                       MG_SCAL_NORM, q);
    for (int j = 0; j < MG_LANCZOS_STEPS; ++j) {
      VOX8_TRY(level_apply(L, q, L.r, 1, st, Epilogue<float>()));
      const int slot = base + 2 * j;
      hipLaunchKernelGGL(k_lz_update, dim3(pb), dim3(256), 0, st, n, (const float*)L.dinv,
                         (const float*)L.r, (const float*)q, qp,
                         (const double*)(j > 0 ? mg.scal + slot - 1 : nullptr), mg.red);
      hipLaunchKernelGGL(k_lz_finish, dim3(1), dim3(256), 0, st, (int)pb, (const double*)mg.red,
                         mg.scal, slot);
      hipLaunchKernelGGL(k_lz_next, dim3(nb(n)), dim3(256), 0, st, n, (const double*)mg.scal, slot,
                         (const float*)q, qp);
      std::swap(q, qp);
    }
    VOX8_HIP(hipGetLastError());
  }
  if (nsm) {
    const size_t cnt = 2 * MG_LANCZOS_STEPS * nsm;
    VOX8_HIP(hipMemcpyAsync(mg.host + MG_SCAL_LZ, mg.scal + MG_SCAL_LZ, sizeof(double) * cnt,
                            hipMemcpyDeviceToHost, st));
    VOX8_HIP(hipStreamSynchronize(st));
  }
  for (size_t l = 0; l < nsm; ++l) {
    MGLevel& L = mg.lv[l];
    const double* h = mg.host + MG_SCAL_LZ + 2 * MG_LANCZOS_STEPS * l;
    double a[MG_LANCZOS_STEPS], b[MG_LANCZOS_STEPS];
    int m = 0;
    double amax = 0.0;
    for (int j = 0; j < MG_LANCZOS_STEPS; ++j) {
      a[j] = h[2 * j];
      b[j] = h[2 * j + 1];
      amax = std::max(amax, std::fabs(a[j]));
      ++m;
      // the Krylov space is exhausted (levels with fewer free DOFs than steps)
      if (!(b[j] > 1e-7 * amax)) break;
    }
    L.lam_max = mg.lam_margin * tridiag_max_eig(a, b, m);
    if (!(L.lam_max > 0.0) || !std::isfinite(L.lam_max)) L.lam_max = 2.5;
  }
  // coarsest level: assemble and invert
  if (mg.n_coarse) {
    const MGLevel& C = mg.lv.back();
    const int n = mg.n_coarse;
    if (C.E64)
      VOX8_TRY(gal_coarse_assemble(C.g, C.E64, C.fixed, mg.A, st));
    else
      hipLaunchKernelGGL(k_mg_coarse_assemble, dim3(nb((int64_t)n * n, 4096)), dim3(256), 0, st, C.g,
                         (const double*)(mg.lv.size() == 1 ? fK0 : C.fK64), (const uint8_t*)C.fixed, dKe,
                         C.scale, mg.A);
    double* diag0 = mg.red;
    const unsigned gs = (unsigned)((n + 255) / 256);
    hipLaunchKernelGGL(k_gj_diag, dim3(gs), dim3(256), 0, st, n, (const double*)mg.A, diag0);
    double* a = mg.A;
    double* b = mg.A2;
    double* D = mg.red + n;                       // GJ_B^2 doubles
    double* W = mg.red + n + GJ_B * GJ_B;         // n * GJ_B doubles
    for (int k0 = 0; k0 < n; k0 += GJ_B) {
      const int bs = std::min(GJ_B, n - k0);
      hipLaunchKernelGGL(k_gjb_panel, dim3(1), dim3(256), 0, st, n, k0, bs, (const double*)a, D,
                         (const double*)diag0, mg.scal + 7);
      hipLaunchKernelGGL(k_gjb_w, dim3((unsigned)(((int64_t)n * bs + 255) / 256)), dim3(256), 0, st, n,
                         k0, bs, (const double*)a, (const double*)D, W);
      hipLaunchKernelGGL(k_gjb_update, dim3(gs, (unsigned)((n + GJ_ROWS - 1) / GJ_ROWS)), dim3(256), 0,
                         st, n, k0, bs, (const double*)a, b, (const double*)D, (const double*)W);
      std::swap(a, b);
    }
    hipLaunchKernelGGL(k_to_f32_n, dim3(nb((int64_t)n * n)), dim3(256), 0, st, (int64_t)n * n,
                       (const double*)a, mg.Ainv);
    VOX8_HIP(hipGetLastError());
    VOX8_HIP(hipMemcpyAsync(mg.host + 7, mg.scal + 7, sizeof(double), hipMemcpyDeviceToHost, st));
    VOX8_HIP(hipStreamSynchronize(st));
    if (mg.host[7] != 0.0) {
      set_error("multigrid: the coarsest-level stiffness matrix is singular (a part of the "
                "structure is not held by fixed DOFs on the coarse grid)");
      return ERR_NUMERIC;
    }
  }
  mg.updated = true;
  ++mg.updates;
  ++mg.epoch;
  return OK;
}

struct ChebCoef { double theta, delta, sigma; };

// nu Chebyshev steps (three-term recurrence).  from_zero: the iterate starts
// at x = 0, r = b and the first step (EPI_CHEB0) reads b and forms the first
// direction itself; otherwise the direction left in L.d by a residual pass is
// continued from the current x, r.
// pending (coarse levels): the post-smoothing sweep's last step xs += d is
// not launched; *pending receives d, and the prolongation that reads xs adds
// it on the fly.
// Chebyshev smoothing of the first kind on [lam_max / alpha, lam_max] in the
// three-term form on the iterates (EPI_X0 / EPI_X): with
// theta = (lmax + lmin) / 2, delta = (lmax - lmin) / 2, sigma = theta / delta,
// rho_0 = 1 / sigma and rho_k = 1 / (2 sigma - rho_{k-1}),
//   x_1     = x_0 + D^-1 (b - A x_0) / theta,
//   x_{k+1} = x_k + rho_k rho_{k-1} (x_k - x_{k-1}) + (2 rho_k / delta) D^-1 (b - A x_k).
// A step reads x_k (the product's input), x_{k-1} and b and writes x_{k+1}:
// no residual or direction vectors are kept.  The level's buffers x, d and
// d2 hold three consecutive iterates in turn.
struct ChebStep {
  double c1, c2;
};

static Epilogue<float> step_epilogue(const MGLevel& L, int mode, double theta, ChebStep cs,
                                     const int* skip) {
  Epilogue<float> ep;
  ep.mode = mode;
  ep.b = L.b;
  ep.dinv = L.dinv;
  ep.c0 = (float)(1.0 / theta);
  ep.c1 = (float)cs.c1;
  ep.c2 = (float)cs.c2;
  ep.ld = L.g.n_dofs;
  ep.skip = skip;
  return ep;
}

// The step coefficients (c1, c2) for k = 1 .. nu - 1 (index k - 1).
static void cheb_steps(double lmax, double alpha, int nu, double* theta, std::vector<ChebStep>& cs) {
  const double lmin = lmax / alpha;
  const double th = 0.5 * (lmax + lmin), delta = 0.5 * (lmax - lmin), sigma = th / delta;
  double rho = 1.0 / sigma;
  cs.clear();
  for (int k = 1; k < nu; ++k) {
    const double rho1 = 1.0 / (2.0 * sigma - rho);
    cs.push_back({rho1 * rho, 2.0 * rho1 / delta});
    rho = rho1;
  }
  *theta = th;
}

// nu steps from x_0 = 0 (x_1 = D^-1 b / theta is formed on load by the first
// product), leaving x_nu in L.x; with `resid`, then r = b - A x_nu in L.r.
static int presmooth(const MGLevel& L, int nu, double alpha, int ncols, hipStream_t st,
                     const int* skip, bool resid) {
  double theta;
  std::vector<ChebStep> cs;
  cheb_steps(L.lam_max, alpha, nu, &theta, cs);
  float* X[3] = {L.x, L.d, L.d2};
  auto slot = [&](int j) { return X[((j - nu) % 3 + 3) % 3]; };   // x_nu in L.x
  if (nu == 1) {
    VOX8_CHECK(resid, "smoothing from zero needs two steps or a residual");
    // x_1 = D^-1 b / theta and r = b - A x_1 (the first step of the residual form)
    Epilogue<float> ep = step_epilogue(L, EPI_CHEB0, theta, {0.0, 0.0}, skip);
    ep.r = L.r;
    ep.xs = L.x;
    return level_apply(L, L.b, nullptr, ncols, st, ep);
  }
  for (int k = 1; k < nu; ++k) {
    Epilogue<float> ep = step_epilogue(L, k == 1 ? EPI_X0 : EPI_X, theta, cs[k - 1], skip);
    ep.prev = k == 2 ? 2 : 1;                       // x_1 is never stored: c0 D^-1 b
    ep.xprev = k >= 3 ? slot(k - 1) : nullptr;
    ep.xs = slot(k + 1);
    VOX8_TRY(level_apply(L, k == 1 ? L.b : slot(k), nullptr, ncols, st, ep));
  }
  if (!resid) return OK;
  Epilogue<float> ep = step_epilogue(L, EPI_RESID, theta, {0.0, 0.0}, skip);
  ep.r = L.r;                                        // no direction: dnew stays null
  return level_apply(L, L.x, nullptr, ncols, st, ep);
}

// nu steps from x_0 = L.x.  The result goes to out64 (FP64, finest level) or
// stays in one of the level's buffers, returned in *xout.
static int postsmooth(const MGLevel& L, int nu, double alpha, int ncols, hipStream_t st,
                      const int* skip, double* out64, const float** xout) {
  VOX8_CHECK(!out64 || !L.S32, "FP64 smoothing output needs the structured stencil");
  double theta;
  std::vector<ChebStep> cs;
  cheb_steps(L.lam_max, alpha, nu, &theta, cs);
  float* X[3] = {L.x, L.d, L.d2};
  for (int k = 0; k < nu; ++k) {
    Epilogue<float> ep = step_epilogue(L, EPI_X, theta, k == 0 ? ChebStep{0.0, 1.0 / theta} : cs[k - 1],
                                       skip);
    ep.prev = k == 0 ? 0 : 1;
    ep.xprev = k == 0 ? nullptr : X[(k - 1) % 3];
    ep.xs = X[(k + 1) % 3];
    if (k == nu - 1 && out64) {
      ep.xs = nullptr;
      ep.out64 = out64;
    }
    VOX8_TRY(level_apply(L, X[k % 3], nullptr, ncols, st, ep));
  }
  if (xout) *xout = out64 ? nullptr : X[nu % 3];
  return OK;
}

// Cycle on level l: x = MG_l(L.b).  Pre-smoothing from zero, gamma
// coarse-grid corrections (each: restrict the residual, cycle on l+1, add the
// prolonged correction, and, before another correction, recompute the
// residual), post-smoothing.  Levels below w_min_dofs take one correction.
// The coarsest level is the dense inverse, or 8 nu smoothing steps from zero
// when it is too large for one.  With the same polynomial before and after
// and symmetric transfers the operator is symmetric, as PCG requires.
// out64 (finest level only): where the cycle's result goes as FP64.
// *xout: the level buffer holding the result otherwise.
static int mg_cycle(MultiGrid& mg, int l, int ncols, hipStream_t st, const int* skip,
                    double* out64, const float** xout) {
  const int nl = (int)mg.lv.size();
  MGLevel& L = mg.lv[l];
  if (l == nl - 1 && mg.n_coarse) {
    const unsigned rb = (unsigned)((mg.n_coarse + 3) / 4);
    hipLaunchKernelGGL(k_mg_coarse_apply<4>, dim3(rb, (ncols + 3) / 4), dim3(256), 0, st,
                       mg.n_coarse, ncols, (const float*)mg.Ainv, (const float*)L.b, L.x, skip);
    VOX8_HIP(hipGetLastError());
    *xout = L.x;
    return OK;
  }
  if (l == nl - 1) {                                 // no coarse grid below
    VOX8_TRY(presmooth(L, 8 * mg.nu, mg.alpha, ncols, st, skip, false));
    *xout = L.x;
    return OK;
  }
  const int nu = (l < 8 && mg.nu_level[l] > 0) ? mg.nu_level[l]
                 : (l >= mg.nu_coarse_level && mg.nu_coarse > 0) ? mg.nu_coarse : mg.nu;
  VOX8_TRY(presmooth(L, nu, mg.alpha, ncols, st, skip, true));
  MGLevel& Cn = mg.lv[l + 1];
  const bool exact_below = (l + 1 == nl - 1) && mg.n_coarse;
  int g = (exact_below || L.g.n_dofs < mg.w_min_dofs || l < mg.v_top) ? 1 : mg.gamma;
  if (l < 8 && mg.gamma_level[l] > 0) g = mg.gamma_level[l];
  for (int j = 0; j < g; ++j) {
    hipLaunchKernelGGL(k_mg_restrict, dim3((unsigned)((Cn.g.n_nodes + 255) / 256), ncols), dim3(256), 0, st,
                       L.g, Cn.g, (const float*)L.r, (const uint8_t*)Cn.fixed, Cn.b, skip);
    VOX8_HIP(hipGetLastError());
    const float* xc = nullptr;
    VOX8_TRY(mg_cycle(mg, l + 1, ncols, st, skip, nullptr, &xc));
    hipLaunchKernelGGL(k_mg_prolong_add, dim3((unsigned)((L.g.n_nodes + 255) / 256), ncols), dim3(256), 0,
                       st, Cn.g, L.g, xc, (const uint8_t*)L.fixed, L.x, skip);
    VOX8_HIP(hipGetLastError());
    if (j + 1 < g) {                                 // the next correction's residual
      double theta;
      std::vector<ChebStep> cs;
      cheb_steps(L.lam_max, mg.alpha, 1, &theta, cs);
      Epilogue<float> ep = step_epilogue(L, EPI_RESID, theta, {0.0, 0.0}, skip);
      ep.r = L.r;
      VOX8_TRY(level_apply(L, L.x, nullptr, ncols, st, ep));
    }
  }
  return postsmooth(L, nu, mg.alpha, ncols, st, skip, out64, xout);
}

static int mg_apply_launches(MultiGrid& mg, const double* r, double* z, int ncols, hipStream_t st,
                             const int* skip) {
  MGLevel& F = mg.lv[0];
  if (r) {
    hipLaunchKernelGGL(k_f64_to_f32, dim3(nb(F.g.n_dofs), ncols), dim3(256), 0, st, F.g.n_dofs, r, F.b,
                       skip);
    VOX8_HIP(hipGetLastError());
  }
  // With a coarser level below, the finest level's cycle ends with a
  // post-smoothing sweep, whose last step writes z directly.
  const bool fused_out = mg.lv.size() > 1;
  const float* xo = nullptr;
  VOX8_TRY(mg_cycle(mg, 0, ncols, st, skip, fused_out ? z : nullptr, &xo));
  if (!fused_out) {
    hipLaunchKernelGGL(k_f32_to_f64, dim3(nb(F.g.n_dofs), ncols), dim3(256), 0, st, F.g.n_dofs,
                       xo, z, skip);
    VOX8_HIP(hipGetLastError());
  }
  return OK;
}

void mg_drop_graphs(MultiGrid& mg) {
  HipSerial lock(hip_serial());
  for (auto& gr : mg.graphs)
    if (gr.exec) (void)hipGraphExecDestroy(gr.exec);
  mg.graphs.clear();
}

MultiGrid::~MultiGrid() {
  HipSerial lock(hip_serial());
  mg_drop_graphs(*this);
  if (cap_st) (void)hipStreamDestroy(cap_st);
}

// Capture the launches of one cycle and instantiate.  The capture runs on the
// context's stream, or, when that is the null stream (which cannot be
// captured), on a private stream created on first use: every stream alive in
// the process takes a share of the few hardware queues HIP maps streams onto
// (GPU_MAX_HW_QUEUES), and once they are shared, work meant to overlap
// serialises.  A capture holds the process-wide HIP lock, so no other
// thread allocates, frees or synchronises while it is open.  Returns false,
// with no error left pending, when the runtime rejects or invalidates the
// capture; the caller then launches the cycle directly.
static bool mg_capture(MultiGrid& mg, const double* r, double* z, int ncols, const int* skip,
                       hipStream_t st, hipGraphExec_t* exec) {
  HipSerial lock(hip_serial());
  *exec = nullptr;
  if (!st && !mg.cap_st && hipStreamCreateWithFlags(&mg.cap_st, hipStreamNonBlocking) != hipSuccess) {
    mg.cap_st = nullptr;
    (void)hipGetLastError();
    return false;
  }
  hipStream_t cs = st ? st : mg.cap_st;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  const int rc = mg_apply_launches(mg, r, z, ncols, cs, skip);
  hipGraph_t graph = nullptr;
  const hipError_t ec = hipStreamEndCapture(cs, &graph);
  bool ok = rc == OK && ec == hipSuccess && graph != nullptr;
  if (ok) ok = hipGraphInstantiateWithFlags(exec, graph, 0) == hipSuccess;
  if (graph) (void)hipGraphDestroy(graph);
  if (!ok) {
    *exec = nullptr;
    (void)hipGetLastError();
  }
  return ok;
}

int mg_apply(MultiGrid& mg, const double* r, double* z, int ncols, hipStream_t st, const int* skip) {
  VOX8_CHECK(mg.updated && ncols <= mg.ncols, "multigrid not set up for this density / width");
  if (!mg.use_graphs) return mg_apply_launches(mg, r, z, ncols, st, skip);
  MGGraph* hit = nullptr;
  for (auto& gr : mg.graphs)
    if (gr.r == r && gr.z == z && gr.ncols == ncols && gr.skip == skip && gr.epoch == mg.epoch)
      hit = &gr;
  if (!hit) {
    // drop stale entries; keep at most 8 live ones (least recently used out)
    for (size_t i = 0; i < mg.graphs.size();) {
      if (mg.graphs[i].epoch != mg.epoch) {
        HipSerial lock(hip_serial());
        (void)hipGraphExecDestroy(mg.graphs[i].exec);
        mg.graphs.erase(mg.graphs.begin() + i);
      } else {
        ++i;
      }
    }
    if (mg.graphs.size() >= 8) {
      size_t old = 0;
      for (size_t i = 1; i < mg.graphs.size(); ++i)
        if (mg.graphs[i].last_use < mg.graphs[old].last_use) old = i;
      {
        HipSerial lock(hip_serial());
        (void)hipGraphExecDestroy(mg.graphs[old].exec);
      }
      mg.graphs.erase(mg.graphs.begin() + old);
    }
    // FP32 stencil shapes of every smoothed level for this width, measured
    // before the cycle is captured on the level's work vectors (their
    // contents are dead between cycles; tuning synchronises; VOX8_TUNE=0
    // keeps the fixed rule)
    static const bool tune = [] {
      const char* e = getenv("VOX8_TUNE");
      return !(e && e[0] == '0');
    }();
    if (tune) {
      const size_t nsm = mg.n_coarse ? mg.lv.size() - 1 : mg.lv.size();
      for (size_t l = 0; l < nsm; ++l) {
        MGLevel& L = mg.lv[l];
        if (L.S32) continue;                  // node-block levels have no shape to tune
        VOX8_TRY(stencil_tune_fp32(L.g, L.pattern, L.c32, L.fK, L.fixed, L.dinv, L.d, L.r, L.x,
                                   L.d2, L.g.n_dofs, ncols, st));
      }
    }
    MGGraph gr;
    gr.r = r; gr.z = z; gr.ncols = ncols; gr.skip = skip; gr.epoch = mg.epoch;
    if (!mg_capture(mg, r, z, ncols, skip, st, &gr.exec)) {
      // after repeated rejections stop trying; either way run this cycle directly
      if (++mg.capture_failures >= 4) mg.use_graphs = false;
      return mg_apply_launches(mg, r, z, ncols, st, skip);
    }
    mg.capture_failures = 0;
    mg.graphs.push_back(gr);
    hit = &mg.graphs.back();
  }
  hit->last_use = ++mg.use_clock;
  VOX8_HIP(hipGraphLaunch(hit->exec, st));
  return OK;
}

}  // namespace vox8
