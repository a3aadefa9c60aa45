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
// plane and marches over a z range.  Per element layer it stages one new node
// plane (with a one-node halo) in LDS, computes the TX x TY elements touching
// the tile (one per thread), writes their 24 nodal forces to LDS
// (component-major, conflict-free), and every owned node gathers its four
// bottom-face and four top-face contributions.  Each input value is read from
// HBM once per tile (plus the halo), each output written once, in x-contiguous
// rows.  Summation order is fixed, so results are bitwise reproducible.
#include "vox8_fast.h"

namespace vox8 {

constexpr int TX = 32, TY = 8, NT = TX * TY;            // element tile = threads
constexpr int PX = TX + 1, PY = TY + 1, PN = PX * PY;   // node plane incl. halo

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

// MODE (K only): EPI_PLAIN y = A x; EPI_RESID r = b - A x, d = c2 D^-1 r;
// EPI_CHEB one Chebyshev smoothing step with x holding the direction d:
//   xs += d, r -= A d, d' = c1 d + c2 D^-1 r   (all node-local once A d is known)
template <typename T, int PAT, bool DO_K, bool DO_M, bool FUSE, int MODE>
__global__ void __launch_bounds__(NT)
k_apply_fast(Grid g, int zchunk, FastCoef<T> C, const T* __restrict__ fK,
             const T* __restrict__ fM, const uint8_t* __restrict__ fixed,
             const T* __restrict__ x, int64_t ldx, T* __restrict__ yK, T* __restrict__ yM,
             int64_t ldy, DotFuse fuse, Epilogue<T> ep) {
  __shared__ T pl[2][3][PN];        // two node planes (component-major)
  __shared__ T buf[24][NT];          // nodal forces of the layer's elements
  __shared__ double sh[8];
  const int col = blockIdx.z;
  if (FUSE && fuse.scal[col].done) return;
  x += col * ldx;
  if (DO_K && MODE == EPI_PLAIN) yK += col * ldy;
  if (DO_M) yM += col * ldy;
  if (MODE != EPI_PLAIN) {
    ep.r += col * ep.ld;
    ep.dnew += col * ep.ld;
    if (MODE == EPI_RESID) ep.b += col * ep.ld;
    if (MODE == EPI_CHEB) ep.xs += col * ep.ld;
  }
  const int t = threadIdx.x;
  const int tx = t % TX, ty = t / TX;
  // blockIdx.x: tile column; blockIdx.y: (tile row, z chunk); blockIdx.z: vector
  const int ty_tiles = (g.NY + TY - 2) / (TY - 1);
  const int nx0 = blockIdx.x * (TX - 1);                      // owned node origin
  const int oy0 = (blockIdx.y % ty_tiles) * (TY - 1);
  const int zs = (blockIdx.y / ty_tiles) * zchunk;
  const int ze = min(zs + zchunk, g.NZ);                       // owned node planes [zs, ze)
  const int ex0 = nx0 - 1, ey0 = oy0 - 1;                      // element tile origin
  // owned node of this thread (gather role)
  const bool owner = tx < TX - 1 && ty < TY - 1;
  const int gx = nx0 + tx, gy = oy0 + ty;
  const bool own_in = owner && gx < g.NX && gy < g.NY;

  auto load_plane = [&](int slot, int iz) {
    for (int q = t; q < 3 * PN; q += NT) {
      const int row = q / (3 * PX), rem = q % (3 * PX);
      const int px = rem / 3, c = rem % 3;
      const int nxg = ex0 + px, nyg = ey0 + row;
      T v = T(0);
      if (iz >= 0 && iz < g.NZ && nxg >= 0 && nxg < g.NX && nyg >= 0 && nyg < g.NY) {
        const int64_t n = nxg + (int64_t)g.NX * (nyg + (int64_t)g.NY * iz);
        if (!((fixed[n] >> c) & 1)) v = x[3 * n + c];
      }
      pl[slot][c][row * PX + px] = v;
    }
  };

  T partK[3] = {0, 0, 0}, partM[3] = {0, 0, 0};
  double dot = 0.0;
  int cur = 0;
  load_plane(cur, zs - 1);
  for (int L = zs - 1; L < ze; ++L) {
    load_plane(cur ^ 1, L + 1);
    __syncthreads();
    // ---- element of this thread in layer L ----
    const int ex = ex0 + tx, ey = ey0 + ty;
    const bool elem_in = L >= 0 && L < g.nz && ex >= 0 && ex < g.nx && ey >= 0 && ey < g.ny;
    T fk = T(0), fm = T(0);
    if (elem_in) {
      const int64_t e = ex + (int64_t)g.nx * (ey + (int64_t)g.ny * L);
      if (DO_K) fk = fK[e];
      if (DO_M) fm = fM[e];
    }
    T mono[3][8];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      T v[8];
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        const int ox = k & 1, oy = (k >> 1) & 1, oz = k >> 2;
        v[k] = pl[oz ? (cur ^ 1) : cur][c][(ty + oy) * PX + tx + ox];
      }
      fwd8(v, mono[c]);
    }
    // ---- K part ----
    if (DO_K) {
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
        T m[8], v[8];
        m[0] = T(0);
        m[1] = fk * gl[c]; m[2] = fk * gl[3 + c]; m[3] = fk * gl[6 + c];
        m[4] = fk * gb[c]; m[5] = fk * gb[3 + c]; m[6] = fk * gb[6 + c];
        m[7] = fk * gt[c];
        inv8(m, v);
#pragma unroll
        for (int k = 0; k < 8; ++k) buf[3 * k + c][t] = v[k];
      }
      __syncthreads();
      if (owner) {
        T bot[3] = {partK[0], partK[1], partK[2]}, top[3] = {0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int sx = s & 1, sy = s >> 1;
          const int el = (ty + sy) * TX + tx + sx;
          const int cb = (1 - sx) + 2 * (1 - sy);      // corner of the node in that element
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            bot[c] += buf[3 * cb + c][el];
            top[c] += buf[3 * (cb + 4) + c][el];
          }
        }
        if (L >= zs && own_in) {
          const int64_t n = gx + (int64_t)g.NX * (gy + (int64_t)g.NY * L);
          const uint8_t f = fixed[n];
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            const bool fx = (f >> c) & 1;
            const T val = fx ? T(0) : bot[c];
            const T xin = pl[cur][c][(ty + 1) * PX + tx + 1];
            if (MODE == EPI_PLAIN) {
              yK[3 * n + c] = val;
              if (FUSE) dot += (double)val * (double)xin;
            } else if (MODE == EPI_RESID) {
              const T r = fx ? T(0) : ep.b[3 * n + c] - val;
              ep.r[3 * n + c] = r;
              ep.dnew[3 * n + c] = ep.c2 * ep.dinv[3 * n + c] * r;
            } else {
              ep.xs[3 * n + c] += xin;
              const T r = ep.r[3 * n + c] - val;
              ep.r[3 * n + c] = r;
              ep.dnew[3 * n + c] = ep.c1 * xin + ep.c2 * ep.dinv[3 * n + c] * r;
            }
          }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) partK[c] = top[c];
      }
      if (DO_M) __syncthreads();
    }
    // ---- M part (diagonal in the monomial basis) ----
    if (DO_M) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        T m[8], v[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) m[k] = fm * C.md[k] * mono[c][k];
        inv8(m, v);
#pragma unroll
        for (int k = 0; k < 8; ++k) buf[3 * k + c][t] = v[k];
      }
      __syncthreads();
      if (owner) {
        T bot[3] = {partM[0], partM[1], partM[2]}, top[3] = {0, 0, 0};
#pragma unroll
        for (int s = 0; s < 4; ++s) {
          const int sx = s & 1, sy = s >> 1;
          const int el = (ty + sy) * TX + tx + sx;
          const int cb = (1 - sx) + 2 * (1 - sy);
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            bot[c] += buf[3 * cb + c][el];
            top[c] += buf[3 * (cb + 4) + c][el];
          }
        }
        if (L >= zs && own_in) {
          const int64_t n = gx + (int64_t)g.NX * (gy + (int64_t)g.NY * L);
          const uint8_t f = fixed[n];
#pragma unroll
          for (int c = 0; c < 3; ++c) yM[3 * n + c] = ((f >> c) & 1) ? T(0) : bot[c];
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) partM[c] = top[c];
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  if (FUSE) {
    const double v = block_sum(dot, sh);
    double tot;
    const int nb = gridDim.x * gridDim.y;
    const int bid = blockIdx.y * gridDim.x + blockIdx.x;
    if (grid_reduce_last<1>(&v, fuse.partials + (int64_t)col * nb, fuse.counters + col, nb, bid,
                            &tot, sh)) {
      if (threadIdx.x == 0) fuse.scal[col].pq = tot;
    }
  }
}

int fast_zchunk(const Grid& g) {
  // enough blocks to fill 256 CUs several times over; at least 4 planes per chunk
  const int64_t tiles = (int64_t)((g.NX + TX - 2) / (TX - 1)) * ((g.NY + TY - 2) / (TY - 1));
  int z = g.NZ;
  while (z > 4 && tiles * ((g.NZ + z - 1) / z) < 2048) z = (z + 1) / 2;
  return z;
}

int fast_blocks(const Grid& g) {
  const int zc = fast_zchunk(g);
  return ((g.NX + TX - 2) / (TX - 1)) * ((g.NY + TY - 2) / (TY - 1)) * ((g.NZ + zc - 1) / zc);
}

template <typename T>
int launch_apply_fast(const Grid& g, int pattern, int op, const FastCoef<T>& C, const T* fK,
                      const T* fM, const uint8_t* fixed, const T* x, int64_t ldx, T* yK, T* yM,
                      int64_t ldy, int nrhs, hipStream_t st, DotFuse fuse, Epilogue<T> ep) {
  const int zc = fast_zchunk(g);
  const int gx = (g.NX + TX - 2) / (TX - 1);
  const int gy = ((g.NY + TY - 2) / (TY - 1)) * ((g.NZ + zc - 1) / zc);
  const dim3 grid(gx, gy, nrhs), blk(NT);
  const bool fz = fuse.scal != nullptr;
  if (ep.mode != EPI_PLAIN && (op != OP_K || fz)) {
    set_error("stencil epilogue modes apply to K alone");
    return ERR_ARG;
  }
#define VOX8_LAUNCH(PAT, K, M, F, MD)                                                            \
  hipLaunchKernelGGL((k_apply_fast<T, PAT, K, M, F, MD>), grid, blk, 0, st, g, zc, C, fK, fM, fixed, \
                     x, ldx, yK, yM, ldy, fuse, ep)
#define VOX8_PAT(PAT)                                                                   \
  if (op == OP_K) {                                                                     \
    if (ep.mode == EPI_RESID) VOX8_LAUNCH(PAT, true, false, false, EPI_RESID);          \
    else if (ep.mode == EPI_CHEB) VOX8_LAUNCH(PAT, true, false, false, EPI_CHEB);       \
    else if (fz) VOX8_LAUNCH(PAT, true, false, true, EPI_PLAIN);                         \
    else VOX8_LAUNCH(PAT, true, false, false, EPI_PLAIN);                               \
  } else if (op == OP_M) VOX8_LAUNCH(PAT, false, true, false, EPI_PLAIN);                \
  else VOX8_LAUNCH(PAT, true, true, false, EPI_PLAIN);
  if (pattern == PAT_ORTHO) { VOX8_PAT(PAT_ORTHO) } else { VOX8_PAT(PAT_BLOCK) }
#undef VOX8_PAT
#undef VOX8_LAUNCH
  VOX8_HIP(hipGetLastError());
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
