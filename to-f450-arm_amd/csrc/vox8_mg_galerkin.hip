// vox8 — Galerkin coarse operators for the multigrid preconditioner.
//
// A_{l+1} = P^T A_l P with the masked trilinear transfer of vox8_mg.hip.  On
// a 2:1-coarsened box grid every fine element lies inside one coarse element
// and its nodal values interpolate from that element's 8 corners only, so
// the triple product is element-local: the coarse element matrix is
//   E_c = sum over the (up to 8) children f of  P_f^T E_f P_f,
// with P_f the 24x24 interpolation of child f's DOFs from the parent's
// corners (weight per axis 1, 1/2 or 0), E_f = f_e Ke on the fine level
// (rows/columns of fixed DOFs zeroed) or the child's Galerkin matrix, and the
// rows/columns of the coarse fixed DOFs zeroed.  The element matrices (FP64)
// build the next level and the coarsest dense matrix; the smoother applies
// each coarse level in node-block form (the 27-point 3x3-block stencil,
// FP32, symmetric half stored) with the Chebyshev epilogues of the
// structured stencil.
#include "vox8_mg.h"

namespace vox8 {

struct F3 { float x, y, z; };      // one node's three components: one 12-byte access

// Buffer-resource loads: the per-lane byte offset is the node's, a uniform
// part (block entry, column) can go in the scalar offset, and an offset past
// the buffer (GAL_OOB) returns 0.  Byte offsets are 32-bit (the launcher
// checks the arrays are below 2^31 bytes).
using GalRsrc = __amdgpu_buffer_rsrc_t;
constexpr uint32_t GAL_OOB = 0x80000000u;
__device__ __forceinline__ GalRsrc gal_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ float gal_ld(GalRsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, 0));
}

// weight of parent corner A for the local node a of child c (per axis: the
// node sits at half-position p = child bit + node offset in {0, 1, 2}; the
// corner at 2 q, q in {0, 1})
__device__ __forceinline__ double gal_w(int c, int a, int A) {
  double w = 1.0;
#pragma unroll
  for (int ax = 0; ax < 3; ++ax) {
    const int cb = ax == 0 ? (c & 1) : (ax == 1 ? ((c >> 1) & 1) : (c >> 2));
    const int p = cb + local_off(a, ax);
    const int q = local_off(A, ax);
    w *= (p == 2 * q) ? 1.0 : (p == 1 ? 0.5 : 0.0);
  }
  return w;
}

__device__ __forceinline__ int64_t gal_node(const Grid& g, int ex, int ey, int ez, int a) {
  return (ex + local_off(a, 0)) + (int64_t)g.NX * ((ey + local_off(a, 1)) + (int64_t)g.NY * (ez + local_off(a, 2)));
}

// one workgroup per coarse element; FROM_FACTORS: children are fine-level
// elements f_e Ke (masked by the fine fixed DOFs), else stored child matrices
template <bool FROM_FACTORS>
__global__ void __launch_bounds__(256)
k_gal_build(Grid gf, Grid gc, const double* __restrict__ fK, const double* __restrict__ Ke,
            const double* __restrict__ Ef, const uint8_t* __restrict__ fixf,
            const uint8_t* __restrict__ fixc, double* __restrict__ Ec64) {
  __shared__ double acc[576], kc[576], tt[576];
  __shared__ uint8_t fxl[8];
  const int64_t E = blockIdx.x;
  const int X = (int)(E % gc.nx), Y = (int)((E / gc.nx) % gc.ny), Z = (int)(E / ((int64_t)gc.nx * gc.ny));
  for (int i = threadIdx.x; i < 576; i += blockDim.x) acc[i] = 0.0;
  for (int c = 0; c < 8; ++c) {
    const int fx = 2 * X + (c & 1), fy = 2 * Y + ((c >> 1) & 1), fz = 2 * Z + (c >> 2);
    if (fx >= gf.nx || fy >= gf.ny || fz >= gf.nz) continue;
    const int64_t f = fx + (int64_t)gf.nx * (fy + (int64_t)gf.ny * fz);
    double fe = 1.0;
    if (FROM_FACTORS) {
      fe = fK[f];
      if (fe == 0.0) continue;                 // inactive child: no contribution
      if (threadIdx.x < 8) fxl[threadIdx.x] = fixf[gal_node(gf, fx, fy, fz, threadIdx.x)];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 576; i += blockDim.x) {
      if (FROM_FACTORS) {
        const int r = i / 24, q = i % 24;
        const bool m = ((fxl[r / 3] >> (r % 3)) & 1) || ((fxl[q / 3] >> (q % 3)) & 1);
        kc[i] = m ? 0.0 : fe * Ke[i];
      } else {
        kc[i] = Ef[f * 576 + i];
      }
    }
    __syncthreads();
    // T = E_f P_c: T[i][3A+cA] = sum_b E_f[i][3b+cA] w(c, b, A)
    for (int i = threadIdx.x; i < 576; i += blockDim.x) {
      const int r = i / 24, J = i % 24, A = J / 3, cA = J % 3;
      double s = 0.0;
#pragma unroll
      for (int b = 0; b < 8; ++b) s += kc[r * 24 + 3 * b + cA] * gal_w(c, b, A);
      tt[i] = s;
    }
    __syncthreads();
    // acc += P_c^T T: acc[3A+cI][J] += sum_a w(c, a, A) T[3a+cI][J]
    for (int i = threadIdx.x; i < 576; i += blockDim.x) {
      const int I = i / 24, J = i % 24, A = I / 3, cI = I % 3;
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 8; ++a) s += gal_w(c, a, A) * tt[(3 * a + cI) * 24 + J];
      acc[i] += s;
    }
    __syncthreads();
  }
  if (threadIdx.x < 8) fxl[threadIdx.x] = fixc[gal_node(gc, X, Y, Z, threadIdx.x)];
  __syncthreads();
  for (int i = threadIdx.x; i < 576; i += blockDim.x) {
    const int r = i / 24, q = i % 24;
    const bool m = ((fxl[r / 3] >> (r % 3)) & 1) || ((fxl[q / 3] >> (q % 3)) & 1);
    Ec64[E * 576 + i] = m ? 0.0 : acc[i];
  }
}

// Jacobi inverse diagonal of an element-assembled level
__global__ void k_gal_jacobi(Grid g, const double* __restrict__ E64, const uint8_t* __restrict__ fixed,
                             float* __restrict__ dinv) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes) return;
  const int ix = (int)(n % g.NX), iy = (int)((n / g.NX) % g.NY), iz = (int)(n / ((int64_t)g.NX * g.NY));
  double d[3] = {0, 0, 0};
  for (int s = 0; s < 8; ++s) {
    const int sx = s & 1, sy = (s >> 1) & 1, sz = s >> 2;
    const int ex = ix - 1 + sx, ey = iy - 1 + sy, ez = iz - 1 + sz;
    if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
    const int64_t e = ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez);
    const int a = local_idx(1 - sx, 1 - sy, 1 - sz);
    for (int c = 0; c < 3; ++c) d[c] += E64[e * 576 + (3 * a + c) * 25];
  }
  const uint8_t fx = fixed[n];
  for (int c = 0; c < 3; ++c)
    dinv[3 * n + c] = ((fx >> c) & 1) ? 0.0f : (d[c] != 0.0 ? (float)(1.0 / d[c]) : 1.0f);
}

// dense matrix of an element-assembled level (identity rows/columns at fixed DOFs)
__global__ void k_gal_coarse_assemble(Grid g, const double* __restrict__ E64,
                                      const uint8_t* __restrict__ fixed, double* __restrict__ A) {
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
      s += E64[(ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez)) * 576 + (3 * a + ci) * 24 + 3 * b + cj];
    }
    A[e] = s;
  }
}

// Node-block form of an element-assembled level: per node i the 3x3 blocks
// A(i, i) and A(i, i + d) for the 13 "forward" offsets d (positive linear
// offset); A(i, i - d) = A(i - d, i)^T is read from node i - d.  Stored FP32,
// node-major S[i * 128 + blk * 9 + 3 r + c] (126 values padded to 128, half
// the 27-point stencil's 243): the lanes working on one node read one
// contiguous 512-byte record.
__constant__ int GAL_D[14][3] = {{0, 0, 0},
                                 {1, 0, 0},
                                 {-1, 1, 0}, {0, 1, 0}, {1, 1, 0},
                                 {-1, -1, 1}, {0, -1, 1}, {1, -1, 1},
                                 {-1, 0, 1}, {0, 0, 1}, {1, 0, 1},
                                 {-1, 1, 1}, {0, 1, 1}, {1, 1, 1}};

__global__ void k_gal_nodeblocks(Grid g, const double* __restrict__ E64, float* __restrict__ S,
                                 int64_t sn, int64_t se) {
  const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= g.n_nodes) return;
  const int ix = (int)(n % g.NX), iy = (int)((n / g.NX) % g.NY), iz = (int)(n / ((int64_t)g.NX * g.NY));
  for (int blk = 0; blk < 14; ++blk) {
    const int dx = GAL_D[blk][0], dy = GAL_D[blk][1], dz = GAL_D[blk][2];
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int s = 0; s < 8; ++s) {
      const int sx = s & 1, sy = (s >> 1) & 1, sz = s >> 2;
      const int ex = ix - 1 + sx, ey = iy - 1 + sy, ez = iz - 1 + sz;
      if (ex < 0 || ey < 0 || ez < 0 || ex >= g.nx || ey >= g.ny || ez >= g.nz) continue;
      const int ox = 1 - sx + dx, oy = 1 - sy + dy, oz = 1 - sz + dz;   // neighbour in the element
      if (ox < 0 || ox > 1 || oy < 0 || oy > 1 || oz < 0 || oz > 1) continue;
      const int a = local_idx(1 - sx, 1 - sy, 1 - sz), b = local_idx(ox, oy, oz);
      const double* Ee = E64 + (ex + (int64_t)g.nx * (ey + (int64_t)g.ny * ez)) * 576;
#pragma unroll
      for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[3 * r + c] += Ee[(3 * a + r) * 24 + 3 * b + c];
    }
#pragma unroll
    for (int q = 0; q < 9; ++q) S[n * sn + (blk * 9 + q) * se] = (float)acc[q];
  }
}

// y = A x for a node-block level with the Chebyshev/residual epilogues of
// the structured stencil (Epilogue in vox8_fast.h).  Eight lanes per node
// share its 27 block products (the diagonal block, 13 forward blocks stored
// at the node, 13 backward blocks A(n - d, n)^T stored at the neighbour):
// lane l takes products l, l + 8, ... for all NC columns of the group
// (blockIdx.y), so every block is loaded once, and the partial sums meet in
// a butterfly over the 8 lanes; lane l then runs the epilogue of column l.
// Workgroups are renumbered so that each XCD (they take workgroups
// round-robin) works on a contiguous range of nodes, keeping the y and z
// neighbours' blocks and vectors in its own L2.  The blocks are zero at
// fixed DOFs, so x needs no masking and the product is zero there.
template <int MODE, int NC>
__global__ void __launch_bounds__(256)
k_gal_apply(Grid g, int ncols, int wg_per_xcd, const float* __restrict__ S,
            const uint8_t* __restrict__ fixed, const float* __restrict__ x, float* __restrict__ y,
            Epilogue<float> ep) {
  constexpr int NPW = 256 / 8;                         // nodes per workgroup
  const int wg = (blockIdx.x % 8) * wg_per_xcd + blockIdx.x / 8;
  const int64_t n = (int64_t)wg * NPW + threadIdx.x / 8;
  const int l8 = threadIdx.x & 7;
  const int c0 = blockIdx.y * NC;
  const bool live = n < g.n_nodes;                    // whole 8-lane groups share this
  const int64_t ld = g.n_dofs, nn = g.n_nodes;
  const int64_t nc = live ? n : 0;
  int ix, iy, iz;
  node_ijk(g, nc, ix, iy, iz);
  const float c0f = ep.c0;
  float acc[NC][3];
#pragma unroll
  for (int k = 0; k < NC; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0f;
  // Items t = l8 + 8 i (i < 4, t < 27), unrolled so every load of the lane is
  // issued before the first is used.  Item t is the offset o of the 27 in the
  // order (dx + 1) + 3 (dy + 1) + 9 (dz + 1): forward block t at o = t + 13,
  // backward block t - 13 at o = 26 - t (GAL_D lists the forward offsets in
  // that order).  Loads go through buffer resources; an item past 26, a
  // neighbour outside the grid or a node past the end reads past the buffer,
  // which returns 0.
  const GalRsrc rs = gal_rsrc(S, 4u * 128u * (uint32_t)nn);
  const GalRsrc rx = gal_rsrc(x, 4u * (uint32_t)(ncols * ld));
  const GalRsrc rdi = gal_rsrc(ep.dinv, (MODE == EPI_CHEB0 || MODE == EPI_X0) ? 12u * (uint32_t)nn : 0u);
  uint32_t xoff[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) xoff[k] = 4u * (uint32_t)((c0 + k < ncols ? c0 + k : ncols - 1) * ld);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int t = l8 + 8 * i;
    const bool fwd = t < 14;
    const int o = fwd ? t + 13 : 26 - t;
    const int dx = o % 3 - 1, dy = (o / 3) % 3 - 1, dz = o / 9 - 1;
    const int blk = fwd ? t : t - 13;
    const int jx = ix + dx, jy = iy + dy, jz = iz + dz;
    const bool in = live && t < 27 && jx >= 0 && jy >= 0 && jz >= 0 && jx < g.NX && jy < g.NY &&
                    jz < g.NZ;
    const uint32_t j = (uint32_t)(jx + g.NX * (jy + g.NY * jz));
    const uint32_t vs = in ? 4u * ((fwd ? (uint32_t)n : j) * 128u + 9u * (uint32_t)blk) : GAL_OOB;
    const uint32_t vx = in ? 12u * j : GAL_OOB;
    float b[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) b[q] = gal_ld(rs, vs + 4u * q, 0u);
    if (!fwd) {                                        // backward: the transpose
      float u;
      u = b[1]; b[1] = b[3]; b[3] = u;
      u = b[2]; b[2] = b[6]; b[6] = u;
      u = b[5]; b[5] = b[7]; b[7] = u;
    }
    float s0 = 1.0f, s1 = 1.0f, s2 = 1.0f;
    if (MODE == EPI_CHEB0 || MODE == EPI_X0) {         // x holds b; the input is c0 D^-1 b
      s0 = c0f * gal_ld(rdi, vx, 0u);
      s1 = c0f * gal_ld(rdi, vx + 4u, 0u);
      s2 = c0f * gal_ld(rdi, vx + 8u, 0u);
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      const float x0 = gal_ld(rx, vx, xoff[k]) * s0;
      const float x1 = gal_ld(rx, vx + 4u, xoff[k]) * s1;
      const float x2 = gal_ld(rx, vx + 8u, xoff[k]) * s2;
      acc[k][0] += b[0] * x0 + b[1] * x1 + b[2] * x2;
      acc[k][1] += b[3] * x0 + b[4] * x1 + b[5] * x2;
      acc[k][2] += b[6] * x0 + b[7] * x1 + b[8] * x2;
    }
  }
  // sum over the node's 8 lanes (aligned groups of 8 within the wavefront)
#pragma unroll
  for (int k = 0; k < NC; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      float v = acc[k][c];
      v += __shfl_xor(v, 1, 64);
      v += __shfl_xor(v, 2, 64);
      v += __shfl_xor(v, 4, 64);
      acc[k][c] = v;
    }
  if (!live || l8 >= NC) return;
  const int col = c0 + l8;
  if (col >= ncols || (ep.skip && ep.skip[col])) return;
  float a0 = 0.0f, a1 = 0.0f, a2 = 0.0f;
#pragma unroll
  for (int k = 0; k < NC; ++k)
    if (k == l8) { a0 = acc[k][0]; a1 = acc[k][1]; a2 = acc[k][2]; }
  const uint8_t fxn = fixed[n];
  const int64_t o = (int64_t)col * ld + 3 * n;
  const float av[3] = {a0, a1, a2};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const bool fx = (fxn >> c) & 1;
    const float val = fx ? 0.0f : av[c];
    if (MODE == EPI_PLAIN) {
      y[o + c] = val;
    } else if (MODE == EPI_RESID) {
      const float r = fx ? 0.0f : ep.b[o + c] - val;
      ep.r[o + c] = r;
      if (ep.dnew) ep.dnew[o + c] = ep.c2 * ep.dinv[3 * n + c] * r;
    } else if (MODE == EPI_CHEB) {
      const float d = x[o + c];
      ep.xs[o + c] += d;
      const float r = ep.r[o + c] - val;
      ep.r[o + c] = r;
      ep.dnew[o + c] = ep.c1 * d + ep.c2 * ep.dinv[3 * n + c] * r;
    } else if (MODE == EPI_X0) {                           // x holds b; x_1 = c0 D^-1 b
      const float bv = x[o + c];
      const float x1 = c0f * ep.dinv[3 * n + c] * bv;
      const float r = fx ? 0.0f : bv - val;
      ep.xs[o + c] = fx ? 0.0f : x1 + ep.c1 * x1 + ep.c2 * ep.dinv[3 * n + c] * r;
    } else if (MODE == EPI_X) {                            // x holds x_k
      const float xk = x[o + c], bv = ep.b[o + c], di = ep.dinv[3 * n + c];
      const float xp = ep.prev == 1 ? ep.xprev[o + c] : (ep.prev == 2 ? c0f * di * bv : 0.0f);
      const float dx = ep.prev == 0 ? 0.0f : xk - xp;
      const float r = fx ? 0.0f : bv - val;
      ep.xs[o + c] = fx ? 0.0f : xk + ep.c1 * dx + ep.c2 * di * r;
    } else {                                               // EPI_CHEB0 (x holds b)
      const float bv = x[o + c];
      const float d = c0f * ep.dinv[3 * n + c] * bv;
      ep.xs[o + c] = d;
      const float r = fx ? 0.0f : bv - val;
      ep.r[o + c] = r;
      if (ep.dnew) ep.dnew[o + c] = ep.c1 * d + ep.c2 * ep.dinv[3 * n + c] * r;
    }
  }
}

// The same product with one node per lane, for levels large enough to fill
// the GPU that way: the blocks are stored entry-major (S[e * n_nodes + n],
// e = blk * 9 + 3 r + c), so each block entry and each neighbour's vector is
// one coalesced load of consecutive nodes, and the sum over the 27 products
// stays in the lane (no butterfly).  The block entry and the column are
// uniform, so they go in the loads' scalar offsets.

// Values of lane - 1 (wave_shr:1) and lane + 1 (wave_shl:1) of the
// wavefront; lane 0, respectively lane 63, keeps `edge`.
__device__ __forceinline__ float lane_prev_or(float edge, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                               __builtin_bit_cast(int, v), 0x138,
                                                               0xf, 0xf, false));
}
__device__ __forceinline__ float lane_next_or(float edge, float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                               __builtin_bit_cast(int, v), 0x130,
                                                               0xf, 0xf, false));
}

// The 27 offsets d = (dx, dy, dz) in the order o = (dx + 1) + 3 (dy + 1) +
// 9 (dz + 1) are GAL_D mirrored around the diagonal: block |o - 13|, read
// transposed at the neighbour when o < 13.  A loop trip takes one (dy, dz)
// row of three offsets: the lane loads the neighbour row's vector at its own
// x position, and its x - 1 / x + 1 values come from the adjacent lanes
// (consecutive lanes hold consecutive nodes); lanes 0 and 63 load the one
// value that lies outside the wavefront.
template <int MODE, int NC>
__global__ void __launch_bounds__(256)
k_gal_apply_lane(Grid g, int ncols, int wg_per_xcd, const float* __restrict__ S,
                 const uint8_t* __restrict__ fixed, const float* __restrict__ x,
                 float* __restrict__ y, Epilogue<float> ep) {
  const int wg = (blockIdx.x % 8) * wg_per_xcd + blockIdx.x / 8;
  const int64_t n_raw = (int64_t)wg * 256 + threadIdx.x;
  const bool live = n_raw < g.n_nodes;                  // every lane takes part in the shifts
  const int64_t n = live ? n_raw : g.n_nodes - 1;
  const int64_t ld = g.n_dofs;
  const uint32_t nn = (uint32_t)g.n_nodes, n32 = (uint32_t)n;
  int ix, iy, iz;
  node_ijk(g, n, ix, iy, iz);
  const int lane = threadIdx.x & 63;
  const int c0 = blockIdx.y * NC;
  const float c0f = ep.c0;
  const GalRsrc rs = gal_rsrc(S, 4u * 126u * nn);
  const GalRsrc rx = gal_rsrc(x, 4u * (uint32_t)(ncols * ld));
  const GalRsrc rdi = gal_rsrc(ep.dinv, (MODE == EPI_CHEB0 || MODE == EPI_X0) ? 12u * nn : 0u);
  uint32_t xoff[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) xoff[k] = 4u * (uint32_t)((c0 + k < ncols ? c0 + k : ncols - 1) * ld);
  float acc[NC][3];
#pragma unroll
  for (int k = 0; k < NC; ++k) acc[k][0] = acc[k][1] = acc[k][2] = 0.0f;
#pragma unroll 1
  for (int r = 0; r < 9; ++r) {
    const int dy = r % 3 - 1, dz = r / 3 - 1;
    const int jy = iy + dy, jz = iz + dz;
    const bool row_in = jy >= 0 && jz >= 0 && jy < g.NY && jz < g.NZ;
    const uint32_t jc = n32 + (uint32_t)(g.NX * (dy + g.NY * dz));    // (ix, jy, jz)
    bool in[3];
    in[0] = row_in && ix > 0;
    in[1] = row_in;
    in[2] = row_in && ix + 1 < g.NX;
    // the blocks of the three offsets
    float b[3][9];
#pragma unroll
    for (int t = 0; t < 3; ++t) {
      const int o = t + 3 * r;
      const int blk = o < 13 ? 13 - o : o - 13;
      const uint32_t j = jc + (uint32_t)(t - 1);
      const uint32_t vs = in[t] ? 4u * (o < 13 ? j : n32) : GAL_OOB;
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        // transposed read for o < 13 (blk and o are uniform: scalar offsets)
        const uint32_t e0 = 4u * (uint32_t)(blk * 9 + q) * nn;
        const uint32_t e1 = 4u * (uint32_t)(blk * 9 + 3 * (q % 3) + q / 3) * nn;
        b[t][q] = gal_ld(rs, vs, o < 13 ? e1 : e0);
      }
    }
    // the row's vector: own x position, then the lanes' neighbours; lane 0
    // fetches x - 1 and lane 63 x + 1 themselves
    const uint32_t vc = row_in ? 12u * jc : GAL_OOB;
    const uint32_t ve = (lane == 0 && in[0]) ? 12u * (jc - 1u)
                      : ((lane == 63 && in[2]) ? 12u * (jc + 1u) : GAL_OOB);
    float s[3][3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (MODE == EPI_CHEB0 || MODE == EPI_X0) {              // x holds b; the input is c0 D^-1 b
        const float dc = gal_ld(rdi, vc + 4u * c, 0u), de = gal_ld(rdi, ve + 4u * c, 0u);
        s[1][c] = c0f * dc;
        s[0][c] = c0f * lane_prev_or(de, dc);
        s[2][c] = c0f * lane_next_or(de, dc);
      } else {
        s[0][c] = s[1][c] = s[2][c] = 1.0f;
      }
    }
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      float xv[3][3];
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const float xc = gal_ld(rx, vc + 4u * c, xoff[k]), xe = gal_ld(rx, ve + 4u * c, xoff[k]);
        xv[1][c] = xc;
        xv[0][c] = lane_prev_or(xe, xc);
        xv[2][c] = lane_next_or(xe, xc);
      }
#pragma unroll
      for (int t = 0; t < 3; ++t) {
        const float x0 = in[t] ? xv[t][0] * s[t][0] : 0.0f;
        const float x1 = in[t] ? xv[t][1] * s[t][1] : 0.0f;
        const float x2 = in[t] ? xv[t][2] * s[t][2] : 0.0f;
        acc[k][0] += b[t][0] * x0 + b[t][1] * x1 + b[t][2] * x2;
        acc[k][1] += b[t][3] * x0 + b[t][4] * x1 + b[t][5] * x2;
        acc[k][2] += b[t][6] * x0 + b[t][7] * x1 + b[t][8] * x2;
      }
    }
  }
  if (!live) return;
  const uint8_t fxn = fixed[n];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int col = c0 + k;
    if (col >= ncols || (ep.skip && ep.skip[col])) continue;
    const int64_t o = (int64_t)col * ld + 3 * n;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const bool fx = (fxn >> c) & 1;
      const float val = fx ? 0.0f : acc[k][c];
      if (MODE == EPI_PLAIN) {
        y[o + c] = val;
      } else if (MODE == EPI_RESID) {
        const float r = fx ? 0.0f : ep.b[o + c] - val;
        ep.r[o + c] = r;
        if (ep.dnew) ep.dnew[o + c] = ep.c2 * ep.dinv[3 * n + c] * r;
      } else if (MODE == EPI_CHEB) {
        const float d = x[o + c];
        ep.xs[o + c] += d;
        const float r = ep.r[o + c] - val;
        ep.r[o + c] = r;
        ep.dnew[o + c] = ep.c1 * d + ep.c2 * ep.dinv[3 * n + c] * r;
      } else if (MODE == EPI_X0) {                           // x holds b; x_1 = c0 D^-1 b
        const float bv = x[o + c];
        const float x1 = c0f * ep.dinv[3 * n + c] * bv;
        const float r = fx ? 0.0f : bv - val;
        ep.xs[o + c] = fx ? 0.0f : x1 + ep.c1 * x1 + ep.c2 * ep.dinv[3 * n + c] * r;
      } else if (MODE == EPI_X) {                            // x holds x_k
        const float xk = x[o + c], bv = ep.b[o + c], di = ep.dinv[3 * n + c];
        const float xp = ep.prev == 1 ? ep.xprev[o + c] : (ep.prev == 2 ? c0f * di * bv : 0.0f);
        const float dx = ep.prev == 0 ? 0.0f : xk - xp;
        const float r = fx ? 0.0f : bv - val;
        ep.xs[o + c] = fx ? 0.0f : xk + ep.c1 * dx + ep.c2 * di * r;
      } else {                                               // EPI_CHEB0 (x holds b)
        const float bv = x[o + c];
        const float d = c0f * ep.dinv[3 * n + c] * bv;
        ep.xs[o + c] = d;
        const float r = fx ? 0.0f : bv - val;
        ep.r[o + c] = r;
        if (ep.dnew) ep.dnew[o + c] = ep.c1 * d + ep.c2 * ep.dinv[3 * n + c] * r;
      }
    }
  }
}


int gal_build_level(const Grid& gf, const Grid& gc, const double* fK0, const double* dKe,
                    const double* Ef64, const uint8_t* fixf, const uint8_t* fixc, double* Ec64,
                    float* S32, bool by_entry, hipStream_t st) {
  if (Ef64)
    hipLaunchKernelGGL(k_gal_build<false>, dim3((unsigned)gc.n_elem), dim3(256), 0, st, gf, gc,
                       (const double*)nullptr, (const double*)nullptr, Ef64, fixf, fixc, Ec64);
  else
    hipLaunchKernelGGL(k_gal_build<true>, dim3((unsigned)gc.n_elem), dim3(256), 0, st, gf, gc, fK0,
                       dKe, (const double*)nullptr, fixf, fixc, Ec64);
  // node-major records (8 lanes per node) or entry-major arrays (a node per lane)
  const int64_t sn = by_entry ? 1 : 128, se = by_entry ? gc.n_nodes : 1;
  hipLaunchKernelGGL(k_gal_nodeblocks, dim3((unsigned)((gc.n_nodes + 255) / 256)), dim3(256), 0, st, gc,
                     (const double*)Ec64, S32, sn, se);
  VOX8_HIP(hipGetLastError());
  return OK;
}

int gal_jacobi(const Grid& g, const double* E64, const uint8_t* fixed, float* dinv, hipStream_t st) {
  hipLaunchKernelGGL(k_gal_jacobi, dim3((unsigned)((g.n_nodes + 255) / 256)), dim3(256), 0, st, g, E64,
                     fixed, dinv);
  VOX8_HIP(hipGetLastError());
  return OK;
}

int gal_coarse_assemble(const Grid& g, const double* E64, const uint8_t* fixed, double* A,
                        hipStream_t st) {
  const int64_t nn = g.n_dofs * g.n_dofs;
  int64_t b = (nn + 255) / 256;
  if (b > 4096) b = 4096;
  hipLaunchKernelGGL(k_gal_coarse_assemble, dim3((unsigned)b), dim3(256), 0, st, g, E64, fixed, A);
  VOX8_HIP(hipGetLastError());
  return OK;
}

template <int NC>
static void gal_apply_nc(const Grid& g, const float* S32, bool by_entry, const uint8_t* fixed,
                         const float* x, float* y, int ncols, const Epilogue<float>& ep,
                         hipStream_t st) {
  const int npw = by_entry ? 256 : 32;                // nodes per workgroup
  const int nwg = (int)((g.n_nodes + npw - 1) / npw);
  const int per_xcd = (nwg + 7) / 8;
  const dim3 grid((unsigned)(8 * per_xcd), (unsigned)((ncols + NC - 1) / NC)), blk(256);
#define VOX8_GAL_LAUNCH(KERN)                                                                      \
  switch (ep.mode) {                                                                               \
    case EPI_RESID: hipLaunchKernelGGL((KERN<EPI_RESID, NC>), grid, blk, 0, st, g, ncols, per_xcd, S32, fixed, x, y, ep); break; \
    case EPI_CHEB: hipLaunchKernelGGL((KERN<EPI_CHEB, NC>), grid, blk, 0, st, g, ncols, per_xcd, S32, fixed, x, y, ep); break;   \
    case EPI_CHEB0: hipLaunchKernelGGL((KERN<EPI_CHEB0, NC>), grid, blk, 0, st, g, ncols, per_xcd, S32, fixed, x, y, ep); break; \
    case EPI_X0: hipLaunchKernelGGL((KERN<EPI_X0, NC>), grid, blk, 0, st, g, ncols, per_xcd, S32, fixed, x, y, ep); break;       \
    case EPI_X: hipLaunchKernelGGL((KERN<EPI_X, NC>), grid, blk, 0, st, g, ncols, per_xcd, S32, fixed, x, y, ep); break;         \
    default: hipLaunchKernelGGL((KERN<EPI_PLAIN, NC>), grid, blk, 0, st, g, ncols, per_xcd, S32, fixed, x, y, ep); break;        \
  }
  if (by_entry) {
    VOX8_GAL_LAUNCH(k_gal_apply_lane)
  } else {
    VOX8_GAL_LAUNCH(k_gal_apply)
  }
#undef VOX8_GAL_LAUNCH
}

int gal_apply(const Grid& g, const float* S32, bool by_entry, const uint8_t* fixed, const float* x,
              float* y, int ncols, const Epilogue<float>& ep, hipStream_t st) {
  // both kernels address the blocks and vectors with 32-bit byte offsets
  VOX8_CHECK((int64_t)128 * 4 * g.n_nodes < ((int64_t)1 << 31) &&
                 (int64_t)4 * ncols * g.n_dofs < ((int64_t)1 << 31),
             "Galerkin level too large for 32-bit offsets");
  // column groups of at most six: with an eight-column instance an 8-column
  // cycle on 440x64x48 took 3.2 ms, as two groups of four 2.7 ms, and twelve
  // columns 4.1 ms against 3.4 ms as two groups of six; five columns (the
  // five load cases) 1.96 ms in the six-column instance, 1.91 ms in their own
  if (ncols <= 1) gal_apply_nc<1>(g, S32, by_entry, fixed, x, y, ncols, ep, st);
  else if (ncols <= 2) gal_apply_nc<2>(g, S32, by_entry, fixed, x, y, ncols, ep, st);
  else if (ncols <= 4) gal_apply_nc<4>(g, S32, by_entry, fixed, x, y, ncols, ep, st);
  else if (ncols == 5) gal_apply_nc<5>(g, S32, by_entry, fixed, x, y, ncols, ep, st);   // the block PCG
  else if (ncols <= 6) gal_apply_nc<6>(g, S32, by_entry, fixed, x, y, ncols, ep, st);
  else if (ncols <= 8) gal_apply_nc<4>(g, S32, by_entry, fixed, x, y, ncols, ep, st);
  else gal_apply_nc<6>(g, S32, by_entry, fixed, x, y, ncols, ep, st);
  VOX8_HIP(hipGetLastError());
  return OK;
}

}  // namespace vox8
