// vox8 — structured matrix-free K.u / M.u for box elements (the hot kernel).
//
// Replaces CSR SpMV over assemble_K / assemble_M (src/fem.py:311-398).
//
// Element operator in the Hadamard (monomial) basis.  A trilinear field on an
// element is u = sum_k a_k m_k with m = [1, xi, eta, zeta, xi eta, eta zeta,
// xi zeta, xi eta zeta]; the coefficients are separable sums/differences of
// the 8 corner values (three butterfly stages, 24 adds per component).  For
// any constant material on a box element the stiffness in this basis couples
// only {xi, eta, zeta} (9x9 block), {xi eta, eta zeta, xi zeta} (9x9 block) and
// {xi eta zeta} (3x3 block), and an orthotropic/isotropic D leaves 45 non-zeros
// in those blocks; the consistent mass is diagonal.  K_e u_e therefore costs
// 72 + 45 + 21 + 72 operations instead of the 576 FMAs of a dense 24x24 product.
// The host (vox8_capi) derives the coefficients from the caller's Ke/Me and
// verifies the structure; otherwise the general dense kernel is used.
//
// Data movement: one workgroup owns a (TX-1) x (TY-1) node tile of an x-y
// plane and marches over a z range; thread (tx, ty) computes element
// (tx, ty) of the tile in every element layer.  A wavefront is one element row
// (TX = 64), so the x-neighbour exchange of nodal forces is a lane shuffle; the
// y-neighbour exchange goes through LDS (6 values per element and operator);
// the z-neighbour exchange stays in registers (the top face of layer L is
// carried to the bottom face of layer L+1).  Node planes (with a one-node halo)
// are staged in LDS, the next one prefetched into registers while the current
// layer computes.  Inputs are read once per tile plus the halo, outputs written
// once; the summation order is fixed, so results are bitwise reproducible.
// Tiles are numbered so that each XCD (workgroups are dealt round-robin to the
// 8 XCDs) receives a contiguous range, keeping halo re-reads in its L2.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "vox8_fast.h"

namespace vox8 {

// Element tile TX x TY = threads (TY = 8 or 12 element rows, one wavefront
// each); owned nodes (TX - 1) x (TY - 1); staged node plane (TX + 1) x (TY + 1).
constexpr int TX = 64, OX = TX - 1, PX = TX + 1;
constexpr int NXCD = 8;

template <typename T, int PAT>
__device__ __forceinline__ void block_products(const FastCoef<T>& C, const T al[9], const T ab[9],
                                               const T at[3], T gl[9], T gb[9], T gt[3]) {
  if (PAT == PAT_ORTHO) {
    // non-zeros of the orthotropic pattern (row: columns)
    gl[0] = C.lin[0][0] * al[0] + C.lin[0][4] * al[4] + C.lin[0][8] * al[8];
    gl[1] = C.lin[1][1] * al[1] + C.lin[1][3] * al[3];
    gl[2] = C.lin[2][2] * al[2] + C.lin[2][6] * al[6];
    gl[3] = C.lin[3][1] * al[1] + C.lin[3][3] * al[3];
    gl[4] = C.lin[4][0] * al[0] + C.lin[4][4] * al[4] + C.lin[4][8] * al[8];
    gl[5] = C.lin[5][5] * al[5] + C.lin[5][7] * al[7];
    gl[6] = C.lin[6][2] * al[2] + C.lin[6][6] * al[6];
    gl[7] = C.lin[7][5] * al[5] + C.lin[7][7] * al[7];
    gl[8] = C.lin[8][0] * al[0] + C.lin[8][4] * al[4] + C.lin[8][8] * al[8];
    gb[0] = C.bil[0][0] * ab[0] + C.bil[0][5] * ab[5];
    gb[1] = C.bil[1][1] * ab[1] + C.bil[1][8] * ab[8];
    gb[2] = C.bil[2][2] * ab[2] + C.bil[2][3] * ab[3] + C.bil[2][7] * ab[7];
    gb[3] = C.bil[3][2] * ab[2] + C.bil[3][3] * ab[3] + C.bil[3][7] * ab[7];
    gb[4] = C.bil[4][4] * ab[4] + C.bil[4][6] * ab[6];
    gb[5] = C.bil[5][0] * ab[0] + C.bil[5][5] * ab[5];
    gb[6] = C.bil[6][4] * ab[4] + C.bil[6][6] * ab[6];
    gb[7] = C.bil[7][2] * ab[2] + C.bil[7][3] * ab[3] + C.bil[7][7] * ab[7];
    gb[8] = C.bil[8][1] * ab[1] + C.bil[8][8] * ab[8];
    gt[0] = C.tri[0][0] * at[0];
    gt[1] = C.tri[1][1] * at[1];
    gt[2] = C.tri[2][2] * at[2];
  } else {
#pragma unroll
    for (int r = 0; r < 9; ++r) {
      T sl = 0, sb = 0;
#pragma unroll
      for (int c = 0; c < 9; ++c) { sl += C.lin[r][c] * al[c]; sb += C.bil[r][c] * ab[c]; }
      gl[r] = sl;
      gb[r] = sb;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) gt[r] = C.tri[r][0] * at[0] + C.tri[r][1] * at[1] + C.tri[r][2] * at[2];
  }
}

// corner values v[x + 2y + 4z] -> butterfly coefficients in monomial order
// [1, xi, eta, zeta, xi eta, eta zeta, xi zeta, xi eta zeta]
template <typename T>
__device__ __forceinline__ void fwd8(const T v[8], T m[8]) {
  const T s0 = v[1] + v[0], d0 = v[1] - v[0], s1 = v[3] + v[2], d1 = v[3] - v[2];
  const T s2 = v[5] + v[4], d2 = v[5] - v[4], s3 = v[7] + v[6], d3 = v[7] - v[6];
  const T ss0 = s1 + s0, sd0 = s1 - s0, ds0 = d1 + d0, dd0 = d1 - d0;
  const T ss1 = s3 + s2, sd1 = s3 - s2, ds1 = d3 + d2, dd1 = d3 - d2;
  m[0] = ss1 + ss0;  // 1
  m[1] = ds1 + ds0;  // xi
  m[2] = sd1 + sd0;  // eta
  m[3] = ss1 - ss0;  // zeta
  m[4] = dd1 + dd0;  // xi eta
  m[5] = sd1 - sd0;  // eta zeta
  m[6] = ds1 - ds0;  // xi zeta
  m[7] = dd1 - dd0;  // xi eta zeta
}

// transpose of fwd8: monomial values -> corner values
template <typename T>
__device__ __forceinline__ void inv8(const T m[8], T v[8]) {
  const T ss0 = m[0] - m[3], ss1 = m[0] + m[3];
  const T sd0 = m[2] - m[5], sd1 = m[2] + m[5];
  const T ds0 = m[1] - m[6], ds1 = m[1] + m[6];
  const T dd0 = m[4] - m[7], dd1 = m[4] + m[7];
  const T s0 = ss0 - sd0, s1 = ss0 + sd0, d0 = ds0 - dd0, d1 = ds0 + dd0;
  const T s2 = ss1 - sd1, s3 = ss1 + sd1, d2 = ds1 - dd1, d3 = ds1 + dd1;
  v[0] = s0 - d0; v[1] = s0 + d0; v[2] = s1 - d1; v[3] = s1 + d1;
  v[4] = s2 - d2; v[5] = s2 + d2; v[6] = s3 - d3; v[7] = s3 + d3;
}

// Buffer-resource access (raw buffer, byte offsets).  DWORD3 of the
// descriptor is the gfx9-family default for untyped 32-bit access.  Loads use
// only the 8- and 32-bit builtins (the compiler merges adjacent dwords into
// dwordx3 / dwordx4 + dwordx2 loads; element extraction from the results of
// the wider load builtins is miscompiled by this toolchain).
using Rsrc = __amdgpu_buffer_rsrc_t;
constexpr uint32_t OOB = 0x80000000u;   // past every array (all < 2^31 bytes)
__device__ __forceinline__ Rsrc make_rsrc(const void* p, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ uint32_t ld_u8(Rsrc r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b8(r, off, 0, 0);
}
__device__ __forceinline__ uint32_t ld_w(Rsrc r, uint32_t off) {
  return __builtin_amdgcn_raw_buffer_load_b32(r, off, 0, 0);
}
__device__ __forceinline__ void st_w(Rsrc r, uint32_t off, uint32_t v) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, off, 0, 0);
}
__device__ __forceinline__ float ld1(Rsrc r, uint32_t off, float) {
  return __builtin_bit_cast(float, ld_w(r, off));
}
__device__ __forceinline__ double ld1(Rsrc r, uint32_t off, double) {
  return __builtin_bit_cast(double, ((uint64_t)ld_w(r, off + 4u) << 32) | ld_w(r, off));
}
__device__ __forceinline__ void st1(Rsrc r, uint32_t off, float v) {
  st_w(r, off, __builtin_bit_cast(uint32_t, v));
}
__device__ __forceinline__ void st1(Rsrc r, uint32_t off, double v) {
  const uint64_t u = __builtin_bit_cast(uint64_t, v);
  st_w(r, off, (uint32_t)u);
  st_w(r, off + 4u, (uint32_t)(u >> 32));
}
// a node's three components
template <typename T>
__device__ __forceinline__ void ld3(Rsrc r, uint32_t off, T v[3]) {
#pragma unroll
  for (int c = 0; c < 3; ++c) v[c] = ld1(r, off + (uint32_t)(c * sizeof(T)), T(0));
}
// stores of a triple as one 96-bit (FP32) / 128 + 64-bit (FP64) access built
// from a vector value (the compiler does not always merge dword stores)
__device__ __forceinline__ void st3(Rsrc r, uint32_t off, const float v[3]) {
  typedef unsigned int u3 __attribute__((ext_vector_type(3)));
  const u3 w = {__builtin_bit_cast(uint32_t, v[0]), __builtin_bit_cast(uint32_t, v[1]),
                __builtin_bit_cast(uint32_t, v[2])};
  __builtin_amdgcn_raw_buffer_store_b96(w, r, off, 0, 0);
}
__device__ __forceinline__ void st3(Rsrc r, uint32_t off, const double v[3]) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  const uint64_t a = __builtin_bit_cast(uint64_t, v[0]), b = __builtin_bit_cast(uint64_t, v[1]),
                 c = __builtin_bit_cast(uint64_t, v[2]);
  const u4 w = {(uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32)};
  const u2 u = {(uint32_t)c, (uint32_t)(c >> 32)};
  __builtin_amdgcn_raw_buffer_store_b128(w, r, off, 0, 0);
  __builtin_amdgcn_raw_buffer_store_b64(u, r, off + 16u, 0, 0);
}

// Value of lane + 1 of the wavefront (DPP wave_shl:1, one VALU move per
// dword; lane 63 receives no value and its result is never used).
__device__ __forceinline__ float lane_next(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130,
                                                               0xf, 0xf, false));
}
__device__ __forceinline__ double lane_next(double v) {
  const int2 w = __builtin_bit_cast(int2, v);
  int2 r;
  r.x = __builtin_amdgcn_update_dpp(0, w.x, 0x130, 0xf, 0xf, false);
  r.y = __builtin_amdgcn_update_dpp(0, w.y, 0x130, 0xf, 0xf, false);
  return __builtin_bit_cast(double, r);
}

// corner k = ox + 2 oy + 4 oz of the element.  After the x exchange a thread
// holds, per face oz and row oy, the sum of its element's ox = 1 corner and the
// right neighbour's ox = 0 corner: s[oy][oz] belongs to node (ex + 1, ey + oy).
template <typename T>
__device__ __forceinline__ void x_exchange(const T f[3][8], T s[2][2][3]) {
#pragma unroll
  for (int oz = 0; oz < 2; ++oz)
#pragma unroll
    for (int oy = 0; oy < 2; ++oy)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const int k0 = 2 * oy + 4 * oz;
        s[oy][oz][c] = f[c][k0 + 1] + lane_next(f[c][k0]);
      }
}

// Element forces of one operator for the thread's element (monomials in, corner
// forces out).
template <typename T, int PAT>
__device__ __forceinline__ void element_K(const FastCoef<T>& C, T fk, const T mono[3][8], T f[3][8]) {
  if (PAT == PAT_CUBE) {
    // class values scaled by the element factor; al/ab index 3 * monomial + c
    const T A = fk * C.lin[0][0], B = fk * C.lin[0][4], Cs = fk * C.lin[1][1];
    const T D = fk * C.bil[0][0], E = fk * C.bil[0][5], F = fk * C.bil[2][2], G = fk * C.bil[2][3];
    const T Hs = fk * C.tri[0][0];
    // lin block: normal terms (xi_x, eta_y, zeta_z), shear pairs
    const T a0 = mono[0][1], a4 = mono[1][2], a8 = mono[2][3];
    const T S = a0 + a4 + a8, AB = A - B, BS = B * S;
    const T l0 = AB * a0 + BS, l4 = AB * a4 + BS, l8 = AB * a8 + BS;
    const T l13 = Cs * (mono[1][1] + mono[0][2]);   // (eta_x... ) pair: gl[1] = gl[3]
    const T l26 = Cs * (mono[2][1] + mono[0][3]);   // gl[2] = gl[6]
    const T l57 = Cs * (mono[2][2] + mono[1][3]);   // gl[5] = gl[7]
    // bil block (ab[3 m + c] = mono[c][4 + m])
    const T b0 = mono[0][4], b1 = mono[1][4], b2 = mono[2][4];
    const T b3 = mono[0][5], b4 = mono[1][5], b5 = mono[2][5];
    const T b6 = mono[0][6], b7 = mono[1][6], b8 = mono[2][6];
    const T g0 = D * b0 + E * b5, g5 = E * b0 + D * b5;
    const T g1 = D * b1 + E * b8, g8 = E * b1 + D * b8;
    const T g4 = D * b4 + E * b6, g6 = E * b4 + D * b6;
    const T FG = F - G, GS = G * (b2 + b3 + b7);
    const T g2 = FG * b2 + GS, g3 = FG * b3 + GS, g7 = FG * b7 + GS;
    T m[8];
    // component x: lin rows 0, 3, 6; bil rows 0, 3, 6; tri row 0
    m[0] = T(0); m[1] = l0; m[2] = l13; m[3] = l26; m[4] = g0; m[5] = g3; m[6] = g6;
    m[7] = Hs * mono[0][7];
    inv8(m, f[0]);
    // component y: lin rows 1, 4, 7; bil rows 1, 4, 7; tri row 1
    m[0] = T(0); m[1] = l13; m[2] = l4; m[3] = l57; m[4] = g1; m[5] = g4; m[6] = g7;
    m[7] = Hs * mono[1][7];
    inv8(m, f[1]);
    // component z: lin rows 2, 5, 8; bil rows 2, 5, 8; tri row 2
    m[0] = T(0); m[1] = l26; m[2] = l57; m[3] = l8; m[4] = g2; m[5] = g5; m[6] = g8;
    m[7] = Hs * mono[2][7];
    inv8(m, f[2]);
    return;
  }
  T al[9], ab[9], at[3], gl[9], gb[9], gt[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    al[c] = mono[c][1]; al[3 + c] = mono[c][2]; al[6 + c] = mono[c][3];
    ab[c] = mono[c][4]; ab[3 + c] = mono[c][5]; ab[6 + c] = mono[c][6];
    at[c] = mono[c][7];
  }
  block_products<T, PAT>(C, al, ab, at, gl, gb, gt);
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    T m[8];
    m[0] = T(0);
    m[1] = fk * gl[c]; m[2] = fk * gl[3 + c]; m[3] = fk * gl[6 + c];
    m[4] = fk * gb[c]; m[5] = fk * gb[3 + c]; m[6] = fk * gb[6 + c];
    m[7] = fk * gt[c];
    inv8(m, f[c]);
  }
}

template <typename T, int PAT>
__device__ __forceinline__ void element_M(const FastCoef<T>& C, T fm, const T mono[3][8], T f[3][8]) {
  if (PAT == PAT_CUBE) {
    const T w0 = fm * C.md[0], w1 = fm * C.md[1], w2 = fm * C.md[4], w3 = fm * C.md[7];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T m[8];
      m[0] = w0 * mono[c][0];
      m[1] = w1 * mono[c][1]; m[2] = w1 * mono[c][2]; m[3] = w1 * mono[c][3];
      m[4] = w2 * mono[c][4]; m[5] = w2 * mono[c][5]; m[6] = w2 * mono[c][6];
      m[7] = w3 * mono[c][7];
      inv8(m, f[c]);
    }
    return;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    T m[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) m[k] = fm * C.md[k] * mono[c][k];
    inv8(m, f[c]);
  }
}

// MODE (K only): EPI_PLAIN y = A x; EPI_RESID r = b - A x, d = c2 D^-1 r;
// EPI_CHEB one Chebyshev smoothing step with x holding the direction d:
//   xs += d, r -= A d, d' = c1 d + c2 D^-1 r   (all node-local once A d is known);
// EPI_CHEB0 the same step from xs = 0, r = b with x holding b: d = c0 D^-1 b
//   is formed while staging, then xs = d, r = b - A d, d' = c1 d + c2 D^-1 r;
// EPI_X0 / EPI_X (FP32) a step of the three-term iterate form (EpiMode): only
//   x_{k+1} is stored (into xs, or as FP64 into out64).
template <typename T, int PAT, bool DO_K, bool DO_M, bool FUSE, int MODE, int TY>
__global__ void __launch_bounds__(TX * TY) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 && !DO_M ? 6 : 1)))
k_apply_fast(Grid g, int zchunk, int tiles_x, int tiles_y, int ncols, FastCoef<T> C,
             const T* __restrict__ fK,
             const T* __restrict__ fM, const uint8_t* __restrict__ fixed,
             const T* __restrict__ x, int64_t ldx, T* __restrict__ yK, T* __restrict__ yM,
             int64_t ldy, DotFuse fuse, Epilogue<T> ep) {
  constexpr int NT = TX * TY, OY = TY - 1, PY = TY + 1, PN = PX * PY;
  constexpr int NLD = (PN + NT - 1) / NT;                  // plane nodes loaded per thread
  constexpr int NOP = (DO_K ? 1 : 0) + (DO_M ? 1 : 0);
  // FP32: three plane slots and two y-exchange buffers, one barrier per layer.
  // FP64 (twice the LDS per value): two slots, one buffer, two barriers.
  constexpr bool ONEBAR = sizeof(T) == 4;
  constexpr int NSLOT = ONEBAR ? 3 : 2, NXB = ONEBAR ? 2 : 1;
  __shared__ T pl[NSLOT][3][PN];       // node planes, component-major
  __shared__ T xch[NXB][NOP][2][3][NT];  // y exchange: s[0][oz][c] per element
  __shared__ double sh[NT / 64];
  // XCD-aware work order: workgroups are dealt round-robin to the 8 XCDs, so
  // XCD j = blockIdx.x % 8 takes the contiguous range [j q + min(j, r), ...)
  // of the nb = 8 q + r work items; an item is (tile, column) with the column
  // fastest, so the columns of a tile and neighbouring tiles share that XCD's
  // L2 (element factors, node masks, diagonal, halo nodes)
  const int nb = gridDim.x, bid = blockIdx.x;
  const int xq = nb / NXCD, xr = nb % NXCD, xj = bid % NXCD;
  const int item = xj * xq + min(xj, xr) + bid / NXCD;
  const int tile = item / ncols, col = item - tile * ncols;
  const int ntiles = nb / ncols;
  if (FUSE && fuse.scal[col].done) return;
  if (ep.skip && ep.skip[col]) return;
  constexpr bool XF = MODE == EPI_X0 || MODE == EPI_X;          // three-term iterate form
  constexpr bool LR = MODE == EPI_LOBRES;
  x += col * ldx;
  if (DO_K && (MODE == EPI_PLAIN || LR)) yK += col * ldy;
  if (DO_M) yM += col * ldy;
  const double lamc = LR ? ep.lam[col] : 0.0;
  const bool lockd = LR && ep.conv && ep.conv[col];
  if (LR) {
    if (ep.w32) ep.w32 += col * ep.ld;
    else ep.w += col * ep.ld;
  }
  if (MODE != EPI_PLAIN && !LR) {
    if (!XF) {
      ep.r += col * ep.ld;
      if (ep.dnew) ep.dnew += col * ep.ld;
    }
    if (MODE == EPI_RESID || MODE == EPI_X) ep.b += col * ep.ld;
    if (MODE != EPI_RESID && ep.xs) ep.xs += col * ep.ld;
    if (MODE == EPI_X && ep.prev == 1) ep.xprev += col * ep.ld;
    if (MODE == EPI_X && ep.out64) ep.out64 += col * ep.ld;
  }
  const int tx_t = tile % tiles_x, rest = tile / tiles_x;
  const int ty_t = rest % tiles_y, zc_t = rest / tiles_y;
  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
  const int nx0 = tx_t * OX, oy0 = ty_t * OY;                  // owned node origin
  const int zs = zc_t * zchunk;
  const int ze = min(zs + zchunk, g.NZ);                       // owned node planes [zs, ze)
  const int ex0 = nx0 - 1, ey0 = oy0 - 1;                      // element tile origin
  const bool owner = tx < OX && ty < OY;
  const int gx = nx0 + tx, gy = oy0 + ty;                      // owned node
  const bool own_in = owner && gx < g.NX && gy < g.NY;
  const int ex = ex0 + tx, ey = ey0 + ty;                      // element
  const bool exy_in = ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny;

  // Memory operations go through buffer resources (base and byte size in
  // scalar registers): the hardware range check makes a load at offset OOB
  // return 0 and drops a store there, so no global load or store sits in a
  // branch and lanes without an owned node need no separate path.  (The memory
  // counter covers loads and stores in issue order; with path-dependent counts
  // the compiler's waits for a load would also wait for just-issued stores.)
  // Byte offsets are 32-bit: the launcher checks every array < 2^31 bytes.
  // Every load of the loop is issued one layer before its use: at the top of
  // layer L the element/epilogue operands of layer L + 1 and the node plane
  // L + 3 (one barrier) or L + 2 (two barriers).
  const uint32_t vbytes = 3u * (uint32_t)g.n_nodes * (uint32_t)sizeof(T);
  const uint32_t ebytes = (uint32_t)g.n_elem * (uint32_t)sizeof(T);
  const Rsrc rx = make_rsrc(x, vbytes);
  const Rsrc rfix = make_rsrc(fixed, (uint32_t)g.n_nodes);
  const Rsrc rfk = make_rsrc(fK, DO_K ? ebytes : 0u);
  const Rsrc rfm = make_rsrc(fM, DO_M ? ebytes : 0u);
  // (a residual pass without a direction output does not read D^-1)
  const Rsrc rdi = make_rsrc(ep.dinv, (MODE != EPI_PLAIN && ep.dinv && (MODE != EPI_RESID || ep.dnew))
                                          ? vbytes : 0u);
  const Rsrc rb = make_rsrc((MODE == EPI_RESID || MODE == EPI_X) ? ep.b : (MODE == EPI_CHEB ? ep.r : x),
                            MODE == EPI_PLAIN ? 0u : vbytes);        // epilogue input a
  const Rsrc rxs = make_rsrc(ep.xs, (MODE == EPI_CHEB || MODE == EPI_CHEB0 || (XF && !ep.out64))
                                        ? vbytes : 0u);
  const Rsrc rxp = make_rsrc(ep.xprev, (MODE == EPI_X && ep.prev == 1) ? vbytes : 0u);
  const Rsrc r64 = make_rsrc(ep.out64, (MODE == EPI_X && ep.out64) ? 2u * vbytes : 0u);
  const Rsrc ry1 = make_rsrc((MODE == EPI_PLAIN || LR) ? yK : ep.r, (DO_K && !XF) ? vbytes : 0u);
  const Rsrc rw = LR && ep.w32 ? make_rsrc(ep.w32, vbytes / 2u) : make_rsrc(ep.w, LR ? vbytes : 0u);
  const Rsrc ry2 = make_rsrc(ep.dnew, (MODE != EPI_PLAIN && !XF && !LR && ep.dnew) ? vbytes : 0u);
  const Rsrc rym = make_rsrc(yM, DO_M ? vbytes : 0u);

  // plane staging: thread loads nodes q = t + i NT of the (PX x PY) plane;
  // the in-plane part of the node index is fixed per thread
  const uint32_t plane = (uint32_t)g.NX * (uint32_t)g.NY;
  uint32_t qoff[NLD];
  bool qin[NLD];
#pragma unroll
  for (int i = 0; i < NLD; ++i) {
    const int q = t + i * NT;
    const int py = q / PX, px = q - py * PX;
    const int nxg = ex0 + px, nyg = ey0 + py;
    qin[i] = q < PN && nxg >= 0 && nxg < g.NX && nyg >= 0 && nyg < g.NY;
    qoff[i] = qin[i] ? (uint32_t)nxg + (uint32_t)g.NX * (uint32_t)nyg : 0u;
  }
  T pre[NLD][3];
  T prd[NLD][3];                                               // CHEB0, X0: D^-1 of the plane
  uint32_t prf[NLD];
  auto fetch = [&](int iz) {
    const bool zin = iz >= 0 && iz < g.NZ;
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const bool in = zin && qin[i];
      const uint32_t n = qoff[i] + plane * (uint32_t)iz;
      const uint32_t o3 = in ? 3u * (uint32_t)sizeof(T) * n : OOB, o1 = in ? n : OOB;
      prf[i] = ld_u8(rfix, o1);
      ld3(rx, o3, pre[i]);
      if (MODE == EPI_CHEB0 || MODE == EPI_X0) ld3(rdi, o3, prd[i]);
    }
  };
  auto store = [&](int slot) {
#pragma unroll
    for (int i = 0; i < NLD; ++i) {
      const int q = t + i * NT;
      if (q < PN)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          T v = pre[i][c];
          if (MODE == EPI_CHEB0 || MODE == EPI_X0) v *= ep.c0 * prd[i][c];
          pl[slot][c][q] = ((prf[i] >> c) & 1u) ? T(0) : v;   // out of range: v = 0
        }
    }
  };
  const uint32_t exy = exy_in ? (uint32_t)ex + (uint32_t)g.nx * (uint32_t)ey : 0u;
  const uint32_t eplane = (uint32_t)g.nx * (uint32_t)g.ny;
  const uint32_t nxy = own_in ? (uint32_t)gx + (uint32_t)g.NX * (uint32_t)gy : 0u;

  // Element factors of element layer L + 1 are loaded during layer L.  The
  // operands of the owned node of plane L (mask, epilogue inputs, and the
  // byte offset the epilogue stores to: OOB when the layer owns no node) are
  // loaded once the element is computed (so they are not live during it) and
  // used after the barrier.
  struct NodeOps {
    uint32_t fnode, off;
    T a[3];    // RESID, X: b; CHEB: r; CHEB0, X0: b (= x)
    T di[3];   // D^-1
    T xs[3];   // CHEB: iterate; X: x_{k-1} (prev 1)
  };
  auto load_fac = [&](int L, T& fk, T& fm) {
    const bool ein = exy_in && L >= 0 && L < g.nz;
    const uint32_t oe = ein ? (exy + eplane * (uint32_t)L) * (uint32_t)sizeof(T) : OOB;
    if (DO_K) fk = ld1(rfk, oe, T(0));
    if (DO_M) fm = ld1(rfm, oe, T(0));
  };
  auto load_node = [&](int L, NodeOps& o) {
    const bool ok = own_in && L >= zs && L < ze;
    const uint32_t nn = nxy + plane * (uint32_t)L;
    o.off = ok ? 3u * (uint32_t)sizeof(T) * nn : OOB;
    o.fnode = ld_u8(rfix, ok ? nn : OOB);
    if (MODE != EPI_PLAIN && !LR) ld3(rb, o.off, o.a);
    if (MODE != EPI_PLAIN) ld3(rdi, o.off, o.di);
    if (MODE == EPI_CHEB) ld3(rxs, o.off, o.xs);
    if (MODE == EPI_X) ld3(rxp, o.off, o.xs);
  };
  // Hadamard coefficients of the x-y face of the element at (tx, ty) in a
  // node plane: per component [s+, s-, d+, d-] with s/d the x sum/difference
  // of rows ty and ty + 1 combined in y.  The top face of element layer L is
  // the bottom face of layer L + 1 and is carried, so each layer transforms
  // one face.
  auto face = [&](int slot, T fc[3][4]) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const T v0 = pl[slot][c][ty * PX + tx], v1 = pl[slot][c][ty * PX + tx + 1];
      const T v2 = pl[slot][c][(ty + 1) * PX + tx], v3 = pl[slot][c][(ty + 1) * PX + tx + 1];
      const T s0 = v1 + v0, d0 = v1 - v0, s1 = v3 + v2, d1 = v3 - v2;
      fc[c][0] = s1 + s0; fc[c][1] = s1 - s0; fc[c][2] = d1 + d0; fc[c][3] = d1 - d0;
    }
  };

  T carK[3] = {0, 0, 0}, carM[3] = {0, 0, 0};
  double dot = 0.0;
  T fk = T(0), fm = T(0), fk1 = T(0), fm1 = T(0);
  fetch(zs - 1);
  store(0);
  fetch(zs);
  store(1);
  if (ONEBAR) fetch(zs + 1);
  load_fac(zs - 1, fk, fm);
  // All loads of the prologue complete here, so none is pending at the loop
  // header.
  __builtin_amdgcn_s_waitcnt(0x0F70);                          // vmcnt(0)
  __syncthreads();
  const int tn = owner ? t + TX : t;                           // element (tx, ty + 1)
  // slots of planes L (sa), L + 1 (sb) and, one barrier: L + 2 (sc)
  int sa = 0, sb = 1, sc = 2, xb = 0;
  T fbot[3][4];                                                // bottom face of layer L
  face(sa, fbot);
  for (int L = zs - 1; L < ze; ++L) {
    // One barrier: plane L + 2 (loaded during layer L - 1) goes to the slot of
    // plane L - 1, which every thread finished reading before the previous
    // barrier; it is read after this layer's barrier.
    if (ONEBAR) store(sc);
    load_fac(L + 1, fk1, fm1);
    fetch(ONEBAR ? L + 3 : L + 2);
    T mono[3][8];
    {
      T ftop[3][4];
      face(sb, ftop);
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        // z stage: monomials [1, xi, eta, zeta, xi eta, eta zeta, xi zeta, xi eta zeta]
        mono[c][0] = ftop[c][0] + fbot[c][0];
        mono[c][3] = ftop[c][0] - fbot[c][0];
        mono[c][2] = ftop[c][1] + fbot[c][1];
        mono[c][5] = ftop[c][1] - fbot[c][1];
        mono[c][1] = ftop[c][2] + fbot[c][2];
        mono[c][6] = ftop[c][2] - fbot[c][2];
        mono[c][4] = ftop[c][3] + fbot[c][3];
        mono[c][7] = ftop[c][3] - fbot[c][3];
#pragma unroll
        for (int k = 0; k < 4; ++k) fbot[c][k] = ftop[c][k];
      }
    }
    T xin[3];                                                   // owned node's input value
#pragma unroll
    for (int c = 0; c < 3; ++c) xin[c] = pl[sa][c][(ty + 1) * PX + tx + 1];
    T sK[2][2][3], sM[2][2][3];
    if (DO_K) {
      T f[3][8];
      element_K<T, PAT>(C, fk, mono, f);                       // fk = 0 outside the grid
      x_exchange(f, sK);
#pragma unroll
      for (int oz = 0; oz < 2; ++oz)
#pragma unroll
        for (int c = 0; c < 3; ++c) xch[xb][0][oz][c][t] = sK[0][oz][c];
    }
    if (DO_M) {
      T f[3][8];
      element_M<T, PAT>(C, fm, mono, f);
      x_exchange(f, sM);
#pragma unroll
      for (int oz = 0; oz < 2; ++oz)
#pragma unroll
        for (int c = 0; c < 3; ++c) xch[xb][NOP - 1][oz][c][t] = sM[0][oz][c];
    }
    NodeOps cur;
    load_node(L, cur);
    __syncthreads();
    if (!ONEBAR) store(sa);                                     // plane L + 2 replaces plane L
    T bK[3], bM[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (DO_K) {
        bK[c] = carK[c] + sK[1][0][c] + xch[xb][0][0][c][tn];
        carK[c] = sK[1][1][c] + xch[xb][0][1][c][tn];
      }
      if (DO_M) {
        bM[c] = carM[c] + sM[1][0][c] + xch[xb][NOP - 1][0][c][tn];
        carM[c] = sM[1][1][c] + xch[xb][NOP - 1][1][c][tn];
      }
    }
    T vK[3], vM[3], o1[3], o2[3], o3[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const bool fx = (cur.fnode >> c) & 1;
      if (DO_K) {
        const T val = fx ? T(0) : bK[c];
        vK[c] = val;
        if (MODE == EPI_PLAIN) {
          if (FUSE) dot += cur.off != OOB ? (double)val * (double)xin[c] : 0.0;
        } else if (LR) {
          // formed below, once M x is known
        } else if (MODE == EPI_RESID) {
          const T r = fx ? T(0) : cur.a[c] - val;
          o1[c] = r;
          o2[c] = ep.c2 * cur.di[c] * r;
        } else if (MODE == EPI_CHEB) {
          o3[c] = cur.xs[c] + xin[c];
          const T r = cur.a[c] - val;
          o1[c] = r;
          o2[c] = ep.c1 * xin[c] + ep.c2 * cur.di[c] * r;
        } else if (MODE == EPI_X0) {                            // xin = x_1 = c0 D^-1 b
          const T r = fx ? T(0) : cur.a[c] - val;
          o3[c] = fx ? T(0) : xin[c] + ep.c1 * xin[c] + ep.c2 * cur.di[c] * r;
        } else if (MODE == EPI_X) {                             // xin = x_k
          const T xp = ep.prev == 2 ? ep.c0 * cur.di[c] * cur.a[c] : cur.xs[c];
          const T dx = ep.prev == 0 ? T(0) : xin[c] - xp;
          const T r = fx ? T(0) : cur.a[c] - val;
          o3[c] = fx ? T(0) : xin[c] + ep.c1 * dx + ep.c2 * cur.di[c] * r;
        } else {                                                // EPI_CHEB0
          o3[c] = xin[c];
          const T r = fx ? T(0) : cur.a[c] - val;
          o1[c] = r;
          o2[c] = ep.c1 * xin[c] + ep.c2 * cur.di[c] * r;
        }
      }
      if (DO_M) vM[c] = fx ? T(0) : bM[c];
      if constexpr (LR && DO_K && DO_M) {
        const double r = (double)vK[c] - lamc * (double)vM[c];
        dot += cur.off != OOB ? r * r : 0.0;
        o1[c] = lockd ? T(0) : (ep.dinv ? (T)(cur.di[c] * r) : (T)r);
      }
    }
    if (DO_K) {
      if (MODE == EPI_PLAIN || LR) {
        st3(ry1, cur.off, vK);
        if (LR && ep.w32) {
          const float w3[3] = {(float)o1[0], (float)o1[1], (float)o1[2]};
          st3(rw, cur.off == OOB ? OOB : cur.off / 2u, w3);
        } else if (LR) {
          st3(rw, cur.off, o1);
        }
      } else if (XF) {
        if (MODE == EPI_X && ep.out64) {
          const double w[3] = {(double)o3[0], (double)o3[1], (double)o3[2]};
          st3(r64, cur.off == OOB ? OOB : 2u * cur.off, w);
        } else {
          st3(rxs, cur.off, o3);
        }
      } else {
        st3(ry1, cur.off, o1);
        st3(ry2, cur.off, o2);
        if (MODE == EPI_CHEB || MODE == EPI_CHEB0) st3(rxs, cur.off, o3);
      }
    }
    if (DO_M) st3(rym, cur.off, vM);
    fk = fk1;
    fm = fm1;
    if (ONEBAR) {
      const int s_old = sa;
      sa = sb; sb = sc; sc = s_old;
      xb ^= 1;
    } else {
      const int s_old = sa;
      sa = sb; sb = s_old;
      __syncthreads();
    }
  }
  if (FUSE || LR) {
    const double v = block_sum(dot, sh);
    write_partials<1>(&v, (FUSE ? fuse.partials : ep.partials) + (int64_t)col * ntiles, tile);
  }
}

static int env_int(const char* name, int dflt) {
  const char* e = getenv(name);
  return e ? atoi(e) : dflt;
}

static int tiles_x(const Grid& g) { return (g.NX + OX - 1) / OX; }
static int tiles_y(const Grid& g, int ty) { return (g.NY + ty - 2) / (ty - 1); }

// Rows of the element tile: 12 where that needs fewer element rows over the
// grid's y extent (less halo work), else 8.  FP32 kernels only: the FP64 ones
// need more registers than two 12-wavefront workgroups per CU allow.
template <typename T>
static int tile_rows(const Grid& g) {
  if (sizeof(T) != 4) return 8;
  return tiles_y(g, 12) * 12 < tiles_y(g, 8) * 8 ? 12 : 8;
}

// Node planes per workgroup (the fixed rule; FP32 launches are autotuned
// instead, see below): halve the z range until the launch has 384 workgroups
// over all columns.  Grids with at least that many (tile, plane) pairs keep
// chunks of >= 4 planes (halo work stays modest; one column of the c5t grid
// gets 7-plane chunks, six columns the whole z range, both measured best);
// smaller grids, whose kernels are bound by the latency of the z march, go
// down to single planes.  VOX8_ZCHUNK fixes the chunk.
static int fast_zchunk(const Grid& g, int ty, int ncols) {
  static const int ov = env_int("VOX8_ZCHUNK", 0);
  if (ov > 0) return ov < g.NZ ? ov : g.NZ;
  const int64_t tiles = (int64_t)tiles_x(g) * tiles_y(g, ty);
  const int64_t wide = tiles * (ncols > 0 ? ncols : 1);
  const int zmin = tiles * g.NZ >= 384 ? 4 : 1;
  int z = g.NZ;
  while (z > zmin && wide * ((g.NZ + z - 1) / z) < 384) z = (z + 1) / 2;
  return z;
}

static int fast_blocks(const Grid& g, int ty, int zc) {
  return tiles_x(g) * tiles_y(g, ty) * ((g.NZ + zc - 1) / zc);
}

int64_t stencil_max_blocks(const Grid& g) {
  return (int64_t)std::max(fast_blocks(g, 8, 1), fast_blocks(g, 12, 1));
}

// ---- shape autotuner (FP32) ----------------------------------------------
// The (tile rows, z chunk) of an FP32 launch is measured rather than guessed:
// for a grid and column count, stencil_tune_fp32 times a Chebyshev smoothing
// step (the multigrid's kernel) of each candidate shape on the caller's work
// vectors and keeps the fastest.  An FP32 result does not depend on the shape (every node's sum
// is formed in the same order whatever the tiling), so tuning changes no
// value.  FP64 launches are measured the same way on first use, except the
// PCG's, whose fused dot products are summed per workgroup: they keep the
// fixed rule above, which keeps those results independent of timing.
struct ShapeKey {
  int nx, ny, nz, ncols;
  bool operator<(const ShapeKey& o) const {
    if (nx != o.nx) return nx < o.nx;
    if (ny != o.ny) return ny < o.ny;
    if (nz != o.nz) return nz < o.nz;
    return ncols < o.ncols;
  }
};
static std::mutex g_shape_mutex;
static std::map<ShapeKey, StencilShape> g_shapes;

template <typename T>
static StencilShape stencil_shape(const Grid& g, int ncols) {
  if (sizeof(T) == 4) {
    std::lock_guard<std::mutex> lock(g_shape_mutex);
    const auto it = g_shapes.find(ShapeKey{g.nx, g.ny, g.nz, ncols});
    if (it != g_shapes.end()) return it->second;
  }
  StencilShape sh;
  sh.ty = tile_rows<T>(g);
  sh.zc = fast_zchunk(g, sh.ty, ncols);
  return sh;
}

bool stencil_tuned_fp32(const Grid& g, int ncols) {
  std::lock_guard<std::mutex> lock(g_shape_mutex);
  return g_shapes.count(ShapeKey{g.nx, g.ny, g.nz, ncols}) != 0;
}

template <int TY>
static int time_shape(const Grid& g, int pattern, const FastCoef<float>& C, const float* fK,
                      const uint8_t* fixed, const float* d, const Epilogue<float>& ep, int ncols,
                      int zc, hipStream_t st, hipEvent_t e0, hipEvent_t e1, float* ms) {
  const int nb = fast_blocks(g, TY, zc);
  const int64_t ld = ep.ld;
#define VOX8_TUNE_LAUNCH(PAT)                                                                   \
  hipLaunchKernelGGL((k_apply_fast<float, PAT, true, false, false, EPI_X, TY>),                  \
                     dim3(nb * ncols), dim3(TX * TY), 0, st, g, zc, tiles_x(g), tiles_y(g, TY),  \
                     ncols, C, fK, (const float*)nullptr, fixed, d, ld, (float*)nullptr,         \
                     (float*)nullptr, ld, DotFuse(), ep)
  auto launch = [&]() {
    if (pattern == PAT_CUBE) VOX8_TUNE_LAUNCH(PAT_CUBE);
    else if (pattern == PAT_ORTHO) VOX8_TUNE_LAUNCH(PAT_ORTHO);
    else VOX8_TUNE_LAUNCH(PAT_BLOCK);
  };
#undef VOX8_TUNE_LAUNCH
  launch();                                               // warm-up
  VOX8_HIP(hipEventRecord(e0, st));
  constexpr int REPS = 3;
  for (int i = 0; i < REPS; ++i) launch();
  VOX8_HIP(hipEventRecord(e1, st));
  VOX8_HIP(hipGetLastError());
  VOX8_HIP(hipEventSynchronize(e1));
  VOX8_HIP(hipEventElapsedTime(ms, e0, e1));
  *ms /= REPS;
  return OK;
}

int stencil_tune_fp32(const Grid& g, int pattern, const FastCoef<float>& C, const float* fK,
                      const uint8_t* fixed, const float* dinv, float* d, float* r, float* xs,
                      float* dnew, int64_t ld, int ncols, hipStream_t st) {
  if (stencil_tuned_fp32(g, ncols)) return OK;
  static const char* pin = getenv("VOX8_FP32_SHAPE");
  if (pin) {
    // "rows/chunk", e.g. "12/999": every FP32 launch takes that tile-row
    // count and z chunk (clamped to the grid) without timing; for profiler
    // runs, whose counter collection distorts the timings the choice rests on
    int ty = 0, zc = 0;
    VOX8_CHECK(sscanf(pin, "%d/%d", &ty, &zc) == 2 && (ty == 8 || ty == 12) && zc >= 1,
               "VOX8_FP32_SHAPE must be rows/chunk with rows 8 or 12");
    std::lock_guard<std::mutex> lock(g_shape_mutex);
    g_shapes[ShapeKey{g.nx, g.ny, g.nz, ncols}] = StencilShape{ty, zc < g.NZ ? zc : g.NZ};
    return OK;
  }
  HipSerial hip_lock(hip_serial());     // event creation and host waits
  // the timed launch: the multigrid's step (three-term form): x_k in d,
  // x_{k-1} in r, b in xs, x_{k+1} to dnew
  Epilogue<float> ep;
  ep.mode = EPI_X;
  ep.b = xs;
  ep.xprev = r;
  ep.prev = 1;
  ep.xs = dnew;
  ep.dinv = dinv;
  ep.c1 = 0.5f;
  ep.c2 = 1e-3f;
  ep.ld = ld;
  VOX8_CHECK(pattern != PAT_DENSE, "shape tuning needs the structured stencil");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  VOX8_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error("hipEventCreate failed");
    return ERR_HIP;
  }
  struct Cand {
    int ty, zc;
    float ms;
  };
  std::vector<Cand> done;
  int rc = OK;
  for (int ty : {8, 12}) {
    // distinct z chunkings: the rule's choice and up to 6 others around it
    std::vector<int> zcs;
    for (int chunks = 1; chunks <= g.NZ && (int)zcs.size() < 12; ++chunks) {
      const int z = (g.NZ + chunks - 1) / chunks;
      if ((g.NZ + z - 1) / z != chunks) continue;
      const int64_t wgs = (int64_t)tiles_x(g) * tiles_y(g, ty) * chunks * ncols;
      if (wgs < 64 && z > 1) continue;            // far too few workgroups
      if (wgs > 16384) break;                     // far too many
      zcs.push_back(z);
    }
    for (int z : zcs) {
      float ms = 0.f;
      rc = ty == 12 ? time_shape<12>(g, pattern, C, fK, fixed, d, ep, ncols, z, st, e0, e1, &ms)
                    : time_shape<8>(g, pattern, C, fK, fixed, d, ep, ncols, z, st, e0, e1, &ms);
      if (rc != OK) break;
      done.push_back(Cand{ty, z, ms});
    }
    if (rc != OK) break;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != OK) return rc;
  // Candidates within 3 % of the fastest count as equal (run-to-run timing
  // noise); among those the one computing the fewest elements (least halo
  // work and traffic) wins, so the choice is stable across runs.
  StencilShape best{tile_rows<float>(g), fast_zchunk(g, tile_rows<float>(g), ncols)};
  float fastest = 1e30f;
  for (const Cand& c : done) fastest = std::min(fastest, c.ms);
  int64_t best_work = INT64_MAX;
  for (const Cand& c : done) {
    if (c.ms > 1.03f * fastest) continue;
    const int64_t work = (int64_t)tiles_x(g) * tiles_y(g, c.ty) * c.ty * ((g.NZ + c.zc - 1) / c.zc) *
                         (c.zc + 1);
    if (work < best_work || (work == best_work && c.zc > best.zc)) {
      best_work = work;
      best = StencilShape{c.ty, c.zc};
    }
  }
  static const bool log = env_int("VOX8_TUNE_LOG", 0) != 0;
  if (log) {
    fprintf(stderr, "[vox8 tune] %dx%dx%d x%d:", g.nx, g.ny, g.nz, ncols);
    for (const Cand& c : done) fprintf(stderr, " %d/%d:%.1fus", c.ty, c.zc, 1e3f * c.ms);
    fprintf(stderr, " -> %d/%d\n", best.ty, best.zc);
  }
  std::lock_guard<std::mutex> lock(g_shape_mutex);
  g_shapes[ShapeKey{g.nx, g.ny, g.nz, ncols}] = best;
  return OK;
}

// ---- FP64 launch shapes ------------------------------------------------------
// A launch without fused dot products gives bitwise the same values for every
// z chunk, so its chunk is measured on first use (per grid, columns and
// operator), like the FP32 shapes.  A PCG launch sums its dot products per
// workgroup, so its chunk must not depend on timing: it keeps the rule above.
// the distinct z chunkings of a grid (chunks of ceil(NZ / n) planes)
static std::vector<int> zchunk_candidates(const Grid& g, int ty, int ncols) {
  std::vector<int> zcs;
  for (int chunks = 1; chunks <= g.NZ && (int)zcs.size() < 12; ++chunks) {
    const int z = (g.NZ + chunks - 1) / chunks;
    if ((g.NZ + z - 1) / z != chunks) continue;
    const int64_t wgs = (int64_t)tiles_x(g) * tiles_y(g, ty) * chunks * ncols;
    if (wgs < 64 && z > 1) continue;
    if (wgs > 16384) break;
    zcs.push_back(z);
  }
  if (zcs.empty()) zcs.push_back(g.NZ);
  return zcs;
}

struct Shape64Key {
  int nx, ny, nz, ncols, op;
  bool operator<(const Shape64Key& o) const {
    if (nx != o.nx) return nx < o.nx;
    if (ny != o.ny) return ny < o.ny;
    if (nz != o.nz) return nz < o.nz;
    if (ncols != o.ncols) return ncols < o.ncols;
    return op < o.op;
  }
};
static std::map<Shape64Key, int> g_zc64;

// The measured chunk of an unfused FP64 launch: on the first launch of a key
// every candidate is timed (one warm-up, three timed launches) writing the
// caller's outputs, which the real launch then overwrites.  Under stream
// capture (no host waits allowed) the fixed rule stands in, unrecorded.
template <typename Launch>
static int tuned_zchunk64(const Grid& g, int op, int ncols, hipStream_t st, Launch& launch,
                          int* zc_out) {
  const Shape64Key key{g.nx, g.ny, g.nz, ncols, op};
  {
    std::lock_guard<std::mutex> lock(g_shape_mutex);
    const auto it = g_zc64.find(key);
    if (it != g_zc64.end()) { *zc_out = it->second; return OK; }
  }
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  VOX8_HIP(hipStreamIsCapturing(st, &cs));
  const int rule = fast_zchunk(g, 8, ncols);
  if (cs != hipStreamCaptureStatusNone) { *zc_out = rule; return OK; }
  HipSerial hip_lock(hip_serial());     // event creation and host waits
  hipEvent_t e0 = nullptr, e1 = nullptr;
  VOX8_HIP(hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    set_error("hipEventCreate failed");
    return ERR_HIP;
  }
  struct Cand { int zc; float ms; };
  std::vector<Cand> done;
  int rc = OK;
  for (int z : zchunk_candidates(g, 8, ncols)) {
    launch(8, z);                                        // warm-up
    if (hipEventRecord(e0, st) != hipSuccess) { rc = ERR_HIP; break; }
    for (int i = 0; i < 3; ++i) launch(8, z);
    float ms = 0.f;
    if (hipEventRecord(e1, st) != hipSuccess || hipGetLastError() != hipSuccess ||
        hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      rc = ERR_HIP;
      break;
    }
    done.push_back(Cand{z, ms / 3});
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != OK) {
    set_error("FP64 stencil shape timing failed");
    return rc;
  }
  // within 3 % of the fastest counts as equal; the least halo work then wins
  float fastest = 1e30f;
  for (const Cand& c : done) fastest = std::min(fastest, c.ms);
  int best = rule;
  int64_t best_work = INT64_MAX;
  for (const Cand& c : done) {
    if (c.ms > 1.03f * fastest) continue;
    const int64_t work = (int64_t)fast_blocks(g, 8, c.zc) * (c.zc + 1);
    if (work < best_work || (work == best_work && c.zc > best)) { best_work = work; best = c.zc; }
  }
  static const bool log = env_int("VOX8_TUNE_LOG", 0) != 0;
  if (log) {
    fprintf(stderr, "[vox8 tune64] %dx%dx%d x%d op %d:", g.nx, g.ny, g.nz, ncols, op);
    for (const Cand& c : done) fprintf(stderr, " %d:%.1fus", c.zc, 1e3f * c.ms);
    fprintf(stderr, " -> %d (rule %d)\n", best, rule);
  }
  std::lock_guard<std::mutex> lock(g_shape_mutex);
  g_zc64[key] = best;
  *zc_out = best;
  return OK;
}

template <typename T>
int launch_apply_fast(const Grid& g, int pattern, int op, const FastCoef<T>& C, const T* fK,
                      const T* fM, const uint8_t* fixed, const T* x, int64_t ldx, T* yK, T* yM,
                      int64_t ldy, int nrhs, hipStream_t st, DotFuse fuse, Epilogue<T> ep) {
  if (3 * g.n_nodes * (int64_t)sizeof(T) >= ((int64_t)1 << 31) ||
      g.n_elem * (int64_t)sizeof(T) >= ((int64_t)1 << 31)) {
    set_error("structured stencil: grid too large for 32-bit byte offsets (use the dense kernel)");
    return ERR_ARG;
  }
  const bool fz = fuse.scal != nullptr;
  if (ep.mode == EPI_LOBRES) {
    if (sizeof(T) != 8 || op != OP_KM || fz || !ep.lam || !(ep.w || ep.w32) || !ep.partials ||
        (int64_t)nrhs * stencil_max_blocks(g) > ep.partials_cap) {
      set_error("residual epilogue: FP64 K and M products with lam, w and partials");
      return ERR_ARG;
    }
  } else if (ep.mode != EPI_PLAIN && (op != OP_K || fz)) {
    set_error("stencil epilogue modes apply to K alone");
    return ERR_ARG;
  }
  if ((ep.mode == EPI_X0 || ep.mode == EPI_X) &&
      (sizeof(T) != 4 || (ep.out64 && 3 * g.n_nodes * (int64_t)sizeof(double) >= ((int64_t)1 << 31)))) {
    set_error("three-term Chebyshev epilogue: FP32 only, FP64 output below 2^31 bytes per column");
    return ERR_ARG;
  }
  // one launch at a given shape (the kernel instance follows the pattern,
  // operator, fusion, epilogue and tile rows)
  auto launch = [&](int ty, int zc) {
    const int nb = fast_blocks(g, ty, zc);   // workgroups per column
#define VOX8_LAUNCH_TY(PAT, K, M, F, MD, TYV)                                                     \
  hipLaunchKernelGGL((k_apply_fast<T, PAT, K, M, F, MD, TYV>), dim3(nb * nrhs), dim3(TX * TYV), 0, \
                     st, g, zc, tiles_x(g), tiles_y(g, TYV), nrhs, C, fK, fM, fixed, x, ldx, yK, yM,\
                     ldy, fuse, ep)
#define VOX8_LAUNCH(PAT, K, M, F, MD)                                                            \
  do {                                                                                           \
    if constexpr (sizeof(T) == 4) {                                                              \
      if (ty == 12) {                                                                            \
        VOX8_LAUNCH_TY(PAT, K, M, F, MD, 12);                                                    \
        break;                                                                                   \
      }                                                                                          \
    }                                                                                            \
    VOX8_LAUNCH_TY(PAT, K, M, F, MD, 8);                                                         \
  } while (0)
#define VOX8_PAT(PAT)                                                                   \
  if (op == OP_K) {                                                                     \
    if (ep.mode == EPI_RESID) VOX8_LAUNCH(PAT, true, false, false, EPI_RESID);          \
    else if (ep.mode == EPI_CHEB) VOX8_LAUNCH(PAT, true, false, false, EPI_CHEB);       \
    else if (ep.mode == EPI_CHEB0) VOX8_LAUNCH(PAT, true, false, false, EPI_CHEB0);     \
    else if (ep.mode == EPI_X0 || ep.mode == EPI_X) {                                   \
      if constexpr (sizeof(T) == 4) {                                                   \
        if (ep.mode == EPI_X0) VOX8_LAUNCH(PAT, true, false, false, EPI_X0);            \
        else VOX8_LAUNCH(PAT, true, false, false, EPI_X);                               \
      }                                                                                 \
    }                                                                                   \
    else if (fz) VOX8_LAUNCH(PAT, true, false, true, EPI_PLAIN);                         \
    else VOX8_LAUNCH(PAT, true, false, false, EPI_PLAIN);                               \
  } else if (op == OP_M) VOX8_LAUNCH(PAT, false, true, false, EPI_PLAIN);                \
  else if (ep.mode == EPI_LOBRES) {                                                     \
    if constexpr (sizeof(T) == 8) VOX8_LAUNCH(PAT, true, true, false, EPI_LOBRES);      \
  } else VOX8_LAUNCH(PAT, true, true, false, EPI_PLAIN);
    if (pattern == PAT_CUBE) { VOX8_PAT(PAT_CUBE) }
    else if (pattern == PAT_ORTHO) { VOX8_PAT(PAT_ORTHO) }
    else { VOX8_PAT(PAT_BLOCK) }
#undef VOX8_PAT
#undef VOX8_LAUNCH
#undef VOX8_LAUNCH_TY
    return nb;
  };
  StencilShape shp = stencil_shape<T>(g, nrhs);
  if constexpr (sizeof(T) == 8) {
    // the fused sums (PCG p.Kp, LOBPCG residual norms) are per workgroup:
    // their shape keeps the fixed rule, so the bits do not depend on timing
    if (!fz && ep.mode != EPI_LOBRES && env_int("VOX8_ZCHUNK", 0) <= 0)
      VOX8_TRY(tuned_zchunk64(g, op, nrhs, st, launch, &shp.zc));
  }
  const int nb = launch(shp.ty, shp.zc);
  VOX8_HIP(hipGetLastError());
  if (fz) VOX8_TRY(launch_pq_finish(fuse, nb, nrhs, st));
  if (ep.mode == EPI_LOBRES && ep.res) VOX8_TRY(launch_partials_finish(ep.partials, nb, nrhs, ep.res, st));
  return OK;
}

template int launch_apply_fast<double>(const Grid&, int, int, const FastCoef<double>&,
                                       const double*, const double*, const uint8_t*,
                                       const double*, int64_t, double*, double*, int64_t, int,
                                       hipStream_t, DotFuse, Epilogue<double>);
template int launch_apply_fast<float>(const Grid&, int, int, const FastCoef<float>&, const float*,
                                      const float*, const uint8_t*, const float*, int64_t, float*,
                                      float*, int64_t, int, hipStream_t, DotFuse, Epilogue<float>);

}  // namespace vox8

namespace vox8 {

// Monomial k evaluated at reference-local corner i.
static double mono_at(int i, int k) {
  static const int corner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                   {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  static const int expo[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                 {1, 1, 0}, {0, 1, 1}, {1, 0, 1}, {1, 1, 1}};
  double v = 1.0;
  for (int a = 0; a < 3; ++a)
    if (expo[k][a]) v *= corner[i][a];
  return v;
}

int fast_coefficients(const double* Ke, const double* Me, FastCoef<double>* c64,
                      FastCoef<float>* c32) {
  // Kt = (1/64) (H^T (x) I3) Ke (H (x) I3), indices 3k+c
  double Kt[24][24], Mt[24][24];
  double H[8][8];
  for (int i = 0; i < 8; ++i)
    for (int k = 0; k < 8; ++k) H[i][k] = mono_at(i, k);
  for (int p = 0; p < 24; ++p)
    for (int q = 0; q < 24; ++q) {
      const int k = p / 3, c = p % 3, l = q / 3, d = q % 3;
      double sk = 0.0, sm = 0.0;
      for (int i = 0; i < 8; ++i)
        for (int j = 0; j < 8; ++j) {
          const double w = H[i][k] * H[j][l];
          sk += w * Ke[(3 * i + c) * 24 + 3 * j + d];
          sm += w * Me[(3 * i + c) * 24 + 3 * j + d];
        }
      Kt[p][q] = sk / 64.0;
      Mt[p][q] = sm / 64.0;
    }
  double kmax = 0.0, mmax = 0.0;
  for (int p = 0; p < 24; ++p)
    for (int q = 0; q < 24; ++q) { kmax = fmax(kmax, fabs(Kt[p][q])); mmax = fmax(mmax, fabs(Mt[p][q])); }
  if (kmax == 0.0 || mmax == 0.0) return PAT_DENSE;
  const double tk = 1e-13 * kmax, tm = 1e-13 * mmax;
  auto grp = [](int p) { const int k = p / 3; return k == 0 ? 0 : (k <= 3 ? 1 : (k <= 6 ? 2 : 3)); };
  // block structure: no coupling across groups, nothing on the constant monomial
  for (int p = 0; p < 24; ++p)
    for (int q = 0; q < 24; ++q) {
      const bool inblk = grp(p) == grp(q) && grp(p) != 0;
      if (!inblk && fabs(Kt[p][q]) > tk) return PAT_DENSE;
      if (p != q && fabs(Mt[p][q]) > tm) return PAT_DENSE;
    }
  for (int k = 0; k < 8; ++k)
    for (int c = 1; c < 3; ++c)
      if (fabs(Mt[3 * k + c][3 * k + c] - Mt[3 * k][3 * k]) > tm) return PAT_DENSE;
  FastCoef<double> C{};
  for (int r = 0; r < 9; ++r)
    for (int s = 0; s < 9; ++s) {
      C.lin[r][s] = Kt[3 + r][3 + s];
      C.bil[r][s] = Kt[12 + r][12 + s];
    }
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s) C.tri[r][s] = Kt[21 + r][21 + s];
  for (int k = 0; k < 8; ++k) C.md[k] = Mt[3 * k][3 * k];
  // orthotropic pattern?
  bool in_lin[9][9] = {}, in_bil[9][9] = {};
  for (auto& e : ORTHO_LIN) in_lin[e[0]][e[1]] = true;
  for (auto& e : ORTHO_BIL) in_bil[e[0]][e[1]] = true;
  int pat = PAT_ORTHO;
  for (int r = 0; r < 9; ++r)
    for (int s = 0; s < 9; ++s) {
      if (!in_lin[r][s] && fabs(C.lin[r][s]) > tk) pat = PAT_BLOCK;
      if (!in_bil[r][s] && fabs(C.bil[r][s]) > tk) pat = PAT_BLOCK;
    }
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s)
      if (r != s && fabs(C.tri[r][s]) > tk) pat = PAT_BLOCK;
  if (pat == PAT_ORTHO) {
    // cubic isotropy: every entry equals its class representative and the
    // mass diagonal depends on the monomial degree only
    double vals[45];
    for (int i = 0; i < 21; ++i) vals[i] = C.lin[ORTHO_LIN[i][0]][ORTHO_LIN[i][1]];
    for (int i = 0; i < 21; ++i) vals[21 + i] = C.bil[ORTHO_BIL[i][0]][ORTHO_BIL[i][1]];
    for (int i = 0; i < 3; ++i) vals[42 + i] = C.tri[i][i];
    const double rep[8] = {C.lin[0][0], C.lin[0][4], C.lin[1][1], C.bil[0][0],
                           C.bil[0][5], C.bil[2][2], C.bil[2][3], C.tri[0][0]};
    bool cube = true;
    for (int i = 0; i < 45; ++i) cube = cube && fabs(vals[i] - rep[CUBE_CLASS[i]]) <= tk;
    for (int k = 1; k < 3; ++k) cube = cube && fabs(C.md[1 + k] - C.md[1]) <= tm &&
                                       fabs(C.md[4 + k] - C.md[4]) <= tm;
    if (cube) pat = PAT_CUBE;
  }
  *c64 = C;
  for (int r = 0; r < 9; ++r)
    for (int s = 0; s < 9; ++s) {
      c32->lin[r][s] = (float)C.lin[r][s];
      c32->bil[r][s] = (float)C.bil[r][s];
    }
  for (int r = 0; r < 3; ++r)
    for (int s = 0; s < 3; ++s) c32->tri[r][s] = (float)C.tri[r][s];
  for (int k = 0; k < 8; ++k) c32->md[k] = (float)C.md[k];
  return pat;
}

}  // namespace vox8
