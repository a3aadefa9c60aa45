// vox8 — block LOBPCG for the k smallest modes of K x = lam M x on free DOFs.
//
// Replaces scipy.sparse.linalg.lobpcg as called by the reference
// (src/fem.py:436-454; GPU variant src/fem_gpu.py:126-152) with:
//   * absolute residual stop test ||K x_j - lam_j M x_j||_2 < tol per column,
//     converged columns stay converged (scipy lobpcg's sticky active mask);
//   * Jacobi preconditioner T = diag(K_ff)^-1 (guard diag > 0, src/fem.py:441-443);
//   * a final explicit Rayleigh-Ritz on the returned block (as scipy does).
// Design (written for this hardware, not a transcription of scipy):
//   * basis S = [X | W | P] and the products AS = K S, BS = M S are kept as
//     column-major blocks; every product is formed by the fused K&M stencil
//     (W, then X' and P' after the Ritz update), none is updated implicitly;
//   * the Rayleigh-Ritz problem is solved on the device in FP64 by one
//     workgroup: diagonal scaling, SVQB deflation of the Gram matrix of M
//     (Stathopoulos & Wu 2002) and cyclic Jacobi eigen-solves with a
//     round-robin parallel ordering (Golub & Van Loan, sec. 8.5);
//   * the Ritz update, residual, its norms and the preconditioned residual
//     are one pass over the blocks.
#include <cstdio>
#include <cstdlib>

#include "vox8_kernels.h"
#include "vox8_pcg.h"

namespace vox8 {

constexpr int MMAX = 36;        // largest Rayleigh-Ritz basis (k <= 12)

// ---- counter-based normal random numbers for the default start block --------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void k_random_block(int64_t n, int k, uint64_t seed, const uint8_t* __restrict__ fixed,
                               double* __restrict__ X) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const bool fx = (fixed[i / 3] >> (i % 3)) & 1;
    for (int j = 0; j < k; ++j) {
      const uint64_t h = mix64(seed ^ mix64((uint64_t)j * (uint64_t)n + (uint64_t)i));
      const double u1 = ((h >> 11) + 1.0) * (1.0 / 9007199254740994.0);   // (0,1]
      const double u2 = (double)(mix64(h) >> 11) * (1.0 / 9007199254740992.0);
      const double z = sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
      X[(int64_t)j * n + i] = fx ? 0.0 : z;
    }
  }
}

// ---- Gram matrices GA = S^T AS, GB = S^T BS over the first m columns ---------
// Each block stages GR-row chunks of all 3m columns in LDS (every HBM value is
// read once; the next chunk is prefetched into registers while the current
// one is reduced) and accumulates 3x3 tiles of the upper-triangular tile pairs
// of both matrices; a task (matrix, tile pair) is split over nsl row slices.
// Per-block partials [block][2][m][m]; off-diagonal tiles are mirrored.
constexpr int GR = 64;
constexpr int GPF = (MMAX + 3) / 4;        // columns per wave and matrix
size_t gram_lds_bytes(int m) {
  const size_t stage = (size_t)3 * m * (GR + 1), red = 256 * 9;
  return sizeof(double) * (stage > red ? stage : red);
}

__global__ void __launch_bounds__(256)
k_gram(int64_t n, int m, const double* __restrict__ S, const double* __restrict__ AS,
       const double* __restrict__ BS, double* __restrict__ partials) {
  extern __shared__ double gsh[];            // [3 m][GR + 1], then 256 x 9 for the reduction
  const int nt = (m + 2) / 3;
  const int ntp = nt * (nt + 1) / 2;
  const int ntask = 2 * ntp;
  const int nsl = ntask >= 256 ? 1 : 256 / ntask;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int task = t / nsl, sl = t - task * nsl;
  const bool active = task < ntask;
  int gsel = 0, ti = 0, tj = 0;
  if (active) {
    gsel = task / ntp;
    int p = task - gsel * ntp;
    while (p >= nt - ti) { p -= nt - ti; ++ti; }
    tj = ti + p;
  }
  double acc[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[i][j] = 0.0;
  // staging role: wave wv loads columns wv, wv + 4, ... of each matrix, row = lane
  double pf[3][GPF];
  auto fetch = [&](int64_t c0) {
    const int64_t r = c0 + lane;
#pragma unroll
    for (int mat = 0; mat < 3; ++mat) {
      const double* src = mat == 0 ? S : (mat == 1 ? AS : BS);
#pragma unroll
      for (int i = 0; i < GPF; ++i) {
        const int cc = wv + 4 * i;
        pf[mat][i] = (cc < m && r < n) ? src[(int64_t)cc * n + r] : 0.0;
      }
    }
  };
  const int64_t step = (int64_t)gridDim.x * GR;
  int64_t c0 = (int64_t)blockIdx.x * GR;
  if (c0 < n) fetch(c0);
  for (; c0 < n; c0 += step) {
#pragma unroll
    for (int mat = 0; mat < 3; ++mat)
#pragma unroll
      for (int i = 0; i < GPF; ++i) {
        const int cc = wv + 4 * i;
        if (cc < m) gsh[(mat * m + cc) * (GR + 1) + lane] = pf[mat][i];
      }
    __syncthreads();
    if (c0 + step < n) fetch(c0 + step);
    if (active) {
      const int rows = (int)(n - c0 < GR ? n - c0 : GR);
      const double* A0 = gsh + (3 * ti) * (GR + 1);
      const double* B0 = gsh + ((1 + gsel) * m + 3 * tj) * (GR + 1);
      for (int r = sl; r < rows; r += nsl) {
        double a[3], bv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          a[i] = 3 * ti + i < m ? A0[i * (GR + 1) + r] : 0.0;
          bv[i] = 3 * tj + i < m ? B0[i * (GR + 1) + r] : 0.0;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
          for (int j = 0; j < 3; ++j) acc[i][j] += a[i] * bv[j];
      }
    }
    __syncthreads();
  }
  double* red = gsh;                          // 256 x 9 doubles
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) red[t * 9 + 3 * i + j] = acc[i][j];
  __syncthreads();
  if (active && sl == 0) {
    double* out = partials + ((int64_t)blockIdx.x * 2 + gsel) * m * m;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double v = 0.0;
        for (int q = 0; q < nsl; ++q) v += red[(t + q) * 9 + 3 * i + j];
        const int ci = 3 * ti + i, cj = 3 * tj + j;
        if (ci < m && cj < m) {
          out[ci * m + cj] = v;
          if (ti != tj) out[cj * m + ci] = v;
        }
      }
  }
}

// Sum the block partials in a fixed order into G = [GA | GB] (m*m each): one
// wavefront per entry.
__global__ void __launch_bounds__(64)
k_gram_finish(int chunks, int m, const double* __restrict__ partials, double* __restrict__ G) {
  const int tot = 2 * m * m;
  const int e = blockIdx.x;
  double s = 0.0;
  for (int c = threadIdx.x; c < chunks; c += 64) s += partials[(int64_t)c * tot + e];
  s = wave_sum(s);
  if (threadIdx.x == 0) G[e] = s;
}

// ---- dense symmetric eigen-solver in one workgroup (parallel cyclic Jacobi) --
// A (m x m, row stride MMAX, in LDS) is diagonalised in place; V accumulates
// the rotations.  Parallel ordering: the circle method pairs all indices into
// disjoint rotations per round (Golub & Van Loan, sec. 8.5.5); an odd m gets a
// dummy index, whose partner is left unrotated that round.  One round is one
// pass: thread roles are the 2x2 blocks (pair a, pair b) of A, which become
// R_a^T A_ab R_b, and the (row, pair) segments of V, which become V_j,ab R_b;
// every entry belongs to exactly one role, so the pass has no hazards.
constexpr int RR_THREADS = 256;

// the threshold test of the Jacobi method below: rotate (p, q) when |a_pq|
// exceeds eps sqrt(|a_pp a_qq|), compared squared
__device__ __forceinline__ bool jacobi_rotates(double apq, double app, double aqq, double eps2) {
  return fabs(apq) > 1e-150 && apq * apq > eps2 * fabs(app * aqq);
}

__device__ void jacobi_eig(double* A, double* V, int m, double eps) {
  __shared__ double rc[MMAX / 2 + 1], rs[MMAX / 2 + 1];
  __shared__ int rp[MMAX / 2 + 1], rq[MMAX / 2 + 1];
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int e = tid; e < m * m; e += nth) V[(e / m) * MMAX + e % m] = (e / m == e % m) ? 1.0 : 0.0;
  const int mm = (m & 1) ? m + 1 : m;
  const int npair = mm / 2;
  const int nA = npair * npair, nroles = nA + m * npair;
  // threshold cyclic Jacobi (jacobi_rotates); the iteration ends before a
  // sweep that would apply no rotation (off-diagonal entries are then at the
  // rounding level of the diagonal).  Such a sweep is known before it starts:
  // nothing changes until some pair rotates, so it rotates none exactly when
  // no pair passes the test at its start, which all threads check at once.
  // The rotation's tangent t = sign(d h) |h| / (|d| + sqrt(d^2 +
  // h^2)), d = a_qq - a_pp, h = 2 a_pq, is computed in FP32 on d and h scaled
  // by a power of two, and c = 1/sqrt(1 + t^2), s = t c in FP64: the rotation
  // is orthogonal to FP64 rounding whatever t is, so only the angle is
  // approximate (relative 1e-7: the entry drops to 1e-7 of itself instead of
  // zero and the next sweep removes the rest; sweep counts are unchanged on
  // LOBPCG-like Grams of size 12-24).  That replaces three FP64 divides and two
  // square roots, the latency of every round.  Tangents the FP32 path would
  // flush (|h| < 1e-30 |d|) take the FP64 formula.
  const double eps2 = eps * eps;
  for (int sweep = 0; sweep < 40; ++sweep) {
    int need = 0;
    for (int e = tid; e < m * m; e += nth) {
      const int p = e / m, q = e - p * m;
      if (p < q) need |= jacobi_rotates(A[p * MMAX + q], A[p * MMAX + p], A[q * MMAX + q], eps2);
    }
    if (!__syncthreads_or(need)) break;
    for (int round = 0; round < mm - 1; ++round) {
      if (tid < npair) {
        const int i = tid;
        const int x = (i == 0) ? 0 : 1 + (i - 1 + round) % (mm - 1);
        const int y = 1 + (mm - 2 - i + round) % (mm - 1);
        int p = x < y ? x : y, q = x < y ? y : x;
        double c = 1.0, sn = 0.0;
        if (q < m) {
          const double apq = A[p * MMAX + q];
          const double app = A[p * MMAX + p], aqq = A[q * MMAX + q];
          if (jacobi_rotates(apq, app, aqq, eps2)) {
            const double d = aqq - app, h = 2.0 * apq;
            int ex;
            frexp(fmax(fabs(d), fabs(h)), &ex);
            const float ds = (float)ldexp(fabs(d), -ex), hs = (float)ldexp(fabs(h), -ex);
            const float tf = hs * __builtin_amdgcn_rcpf(ds + __builtin_amdgcn_sqrtf(ds * ds + hs * hs));
            double t;
            if (tf >= 1e-30f) {
              t = (d > 0.0 && h < 0.0) || (d < 0.0 && h > 0.0) ? -(double)tf : (double)tf;
            } else {
              const double tau = d / h;
              t = tau >= 0.0 ? 1.0 / (tau + sqrt(1.0 + tau * tau)) : -1.0 / (-tau + sqrt(1.0 + tau * tau));
            }
            c = rsqrt(1.0 + t * t);
            sn = t * c;
          }
        } else {
          q = -1;                           // p is paired with the dummy index
        }
        rc[i] = c; rs[i] = sn; rp[i] = p; rq[i] = q;
      }
      __syncthreads();
      for (int e = tid; e < nroles; e += nth) {
        if (e < nA) {                       // A block (a, b)
          const int ia = e / npair, ib = e - ia * npair;
          const int p1 = rp[ia], q1 = rq[ia], p2 = rp[ib], q2 = rq[ib];
          const double c1 = rc[ia], s1 = rs[ia], c2 = rc[ib], s2 = rs[ib];
          if (s1 == 0.0 && s2 == 0.0) continue;
          const double a_pp = A[p1 * MMAX + p2];
          const double a_pq = q2 >= 0 ? A[p1 * MMAX + q2] : 0.0;
          const double a_qp = q1 >= 0 ? A[q1 * MMAX + p2] : 0.0;
          const double a_qq = (q1 >= 0 && q2 >= 0) ? A[q1 * MMAX + q2] : 0.0;
          const double t_pp = c1 * a_pp - s1 * a_qp, t_qp = s1 * a_pp + c1 * a_qp;
          const double t_pq = c1 * a_pq - s1 * a_qq, t_qq = s1 * a_pq + c1 * a_qq;
          A[p1 * MMAX + p2] = c2 * t_pp - s2 * t_pq;
          if (q2 >= 0) A[p1 * MMAX + q2] = s2 * t_pp + c2 * t_pq;
          if (q1 >= 0) A[q1 * MMAX + p2] = c2 * t_qp - s2 * t_qq;
          if (q1 >= 0 && q2 >= 0) A[q1 * MMAX + q2] = s2 * t_qp + c2 * t_qq;
        } else {                            // V row j, pair b
          const int f = e - nA;
          const int j = f / npair, ib = f - j * npair;
          const int p = rp[ib], q = rq[ib];
          if (q < 0 || rs[ib] == 0.0) continue;
          const double c = rc[ib], sn = rs[ib];
          const double vp = V[j * MMAX + p], vq = V[j * MMAX + q];
          V[j * MMAX + p] = c * vp - sn * vq;
          V[j * MMAX + q] = sn * vp + c * vq;
        }
      }
      __syncthreads();
    }
  }
}

// Generalised Rayleigh-Ritz: GA c = lam GB c, k smallest lam.  The basis is
// orthonormalised in the K metric (GA), which separates directions the mass
// metric cannot (high-frequency content has large K-norm, small M-norm):
//   D = diag(GA)^-1/2, GA_s = D GA D = V diag(s) V^T, keep s > drop * max(s),
//   Q = V_keep diag(s^-1/2), H = Q^T (D GB D) Q = Y diag(mu) Y^T,
//   lam = 1/mu for the k largest mu, c = D Q y / sqrt(mu)  (M-normalised).
// Works on the leading m x m block of G (row stride ld).  Returns true on
// failure (fewer than k directions kept or a non-positive mu).
__device__ bool rr_solve(int m, int ld, int k, const double* __restrict__ G, double* __restrict__ C,
                         double* __restrict__ lam, double drop_tol, double jeps) {
  __shared__ double A[MMAX * MMAX], B[MMAX * MMAX], V[MMAX * MMAX], T1[MMAX * MMAX];
  __shared__ double dsc[MMAX], sig[MMAX], mu[MMAX];
  __shared__ int keep[MMAX], perm[MMAX];
  __shared__ int s_r, s_bad;
  const int tid = threadIdx.x, nth = blockDim.x;
  const double* GA = G;
  const double* GB = G + ld * ld;
  for (int i = tid; i < m; i += nth) {
    const double d = GA[i * ld + i];
    dsc[i] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
  }
  __syncthreads();
  for (int e = tid; e < m * m; e += nth) {
    const int i = e / m, j = e % m;
    const double sc = dsc[i] * dsc[j];
    A[i * MMAX + j] = 0.5 * (GA[i * ld + j] + GA[j * ld + i]) * sc;
    B[i * MMAX + j] = 0.5 * (GB[i * ld + j] + GB[j * ld + i]) * sc;
  }
  __syncthreads();
  jacobi_eig(A, V, m, jeps);
  if (tid == 0) {
    double smax = 0.0;
    for (int i = 0; i < m; ++i) { sig[i] = A[i * MMAX + i]; smax = fmax(smax, sig[i]); }
    int r = 0;
    for (int i = 0; i < m; ++i)
      if (sig[i] > drop_tol * smax) keep[r++] = i;
    s_r = r;
  }
  __syncthreads();
  const int r = s_r;
  for (int e = tid; e < m * r; e += nth) {                 // Q (m x r) in T1
    const int i = e / r, l = e % r;
    T1[i * MMAX + l] = V[i * MMAX + keep[l]] / sqrt(sig[keep[l]]);
  }
  __syncthreads();
  for (int e = tid; e < m * r; e += nth) {                 // V <- B_s Q
    const int i = e / r, l = e % r;
    double s = 0.0;
    for (int j = 0; j < m; ++j) s += B[i * MMAX + j] * T1[j * MMAX + l];
    V[i * MMAX + l] = s;
  }
  __syncthreads();
  for (int e = tid; e < r * r; e += nth) {                 // A <- Q^T B_s Q
    const int l = e / r, q = e % r;
    double s = 0.0;
    for (int i = 0; i < m; ++i) s += T1[i * MMAX + l] * V[i * MMAX + q];
    A[l * MMAX + q] = s;
  }
  __syncthreads();
  for (int e = tid; e < r * r; e += nth) {
    const int l = e / r, q = e % r;
    B[l * MMAX + q] = 0.5 * (A[l * MMAX + q] + A[q * MMAX + l]);
  }
  __syncthreads();
  jacobi_eig(B, V, r, jeps);                                // V: eigenvectors Y (r x r)
  for (int i = tid; i < r; i += nth) mu[i] = B[i * MMAX + i];
  __syncthreads();
  // descending mu by rank, ties by index; NaN ranks as -inf, so the ranks
  // always form a permutation (a NaN among the k largest fails the check below)
  for (int i = tid; i < r; i += nth) {
    const double v = mu[i] == mu[i] ? mu[i] : -INFINITY;
    int rank = 0;
    for (int j = 0; j < r; ++j) {
      const double u = mu[j] == mu[j] ? mu[j] : -INFINITY;
      rank += u > v || (u == v && j < i);
    }
    perm[rank] = i;
  }
  __syncthreads();
  if (tid == 0) {
    bool bad = r < k;
    for (int j = 0; j < k && j < r; ++j) bad = bad || !(mu[perm[j]] > 0.0);
    s_bad = bad;
  }
  __syncthreads();
  const bool bad = s_bad;
  if (!bad) {
    for (int e = tid; e < m * k; e += nth) {
      const int i = e / k, j = e % k;
      double s = 0.0;
      for (int l = 0; l < r; ++l) s += T1[i * MMAX + l] * V[l * MMAX + perm[j]];
      C[i * k + j] = s * dsc[i] / sqrt(mu[perm[j]]);
    }
    for (int j = tid; j < k; j += nth) lam[j] = 1.0 / mu[perm[j]];
  }
  __syncthreads();
  return bad;
}

// Rayleigh-Ritz on [X | W | P] (m = 3k) with the restart of LOBPCG done on
// the device: if the full basis fails, [X | W] is used and the P rows of C are
// zero (so the update is X' = X C_X + W C_W, P' = W C_W); if that fails too,
// C keeps X unchanged and status = 1.
// Out (global): C (m x k, row-major), lam (k), status (0 ok, 1 failure).
__global__ void __launch_bounds__(RR_THREADS)
k_rayleigh_ritz(int m, int k, const double* __restrict__ G, double* __restrict__ C,
                double* __restrict__ lam, double* __restrict__ status, double drop_tol,
                double jeps) {
  bool bad = rr_solve(m, m, k, G, C, lam, drop_tol, jeps);
  int used = m;
  if (bad && m > 2 * k) {
    bad = rr_solve(2 * k, m, k, G, C, lam, drop_tol, jeps);
    used = 2 * k;
  }
  if (bad) used = 0;
  for (int e = threadIdx.x + used * k; e < m * k; e += blockDim.x) C[e] = 0.0;
  __syncthreads();
  if (bad)
    for (int j = threadIdx.x; j < k; j += blockDim.x) C[j * k + j] = 1.0;
  if (threadIdx.x == 0) status[0] = bad ? 1.0 : 0.0;
}

// ---- Ritz update (templated on the block size) ------------------------------
// S: column-major blocks (ld n) holding [X | W | P]; the first NB*K columns
// form the current basis.  Writes X' = S C into the X slots and, for NB > 1,
// the new directions P' = [W P] C[K:] into the P slots.  No product is
// updated here: K X', M X', K P', M P' are recomputed exactly by the stencil.
// Implicit updates (K P' = [KW KP] C[K:]) carry the rounding of every earlier
// step, amplified by the cancellation in P' near convergence; the error grows
// geometrically and the Rayleigh-Ritz then picks diverging combinations.
// One row per thread (no grid-stride loop): with a loop the compiler keeps
// all of C in registers across iterations (M*K doubles: one wave per SIMD,
// and scratch spills from k = 8 on); here C is read from LDS as it is used.
// conv (may be null): pairs locked so far; their new direction P' is written
// as zero (their P, K P and M P stay zero: the K and M products skip them).
template <int K, int NB>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NB * K * K > 150 ? 2 : 4)))
k_ritz_combine(int64_t n, double* __restrict__ S, const double* __restrict__ Cg,
               const int* __restrict__ conv) {
  constexpr int M = NB * K;
  __shared__ double C[M * K];
  __shared__ int lk[K];
  for (int e = threadIdx.x; e < M * K; e += blockDim.x) C[e] = Cg[e];
  if (threadIdx.x < K) lk[threadIdx.x] = conv ? conv[threadIdx.x] : 0;
  __syncthreads();
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  double v[M];
#pragma unroll
  for (int i = 0; i < M; ++i) v[i] = S[(int64_t)i * n + r];
#pragma unroll
  for (int j = 0; j < K; ++j) {
    double x = 0.0, p = 0.0;
#pragma unroll
    for (int i = 0; i < K; ++i) x += v[i] * C[i * K + j];
#pragma unroll
    for (int i = K; i < M; ++i) p += v[i] * C[i * K + j];
    S[(int64_t)j * n + r] = x + p;
    if (NB > 1) S[(int64_t)(2 * K + j) * n + r] = lk[j] ? 0.0 : p;
  }
}

// R = AX - BX diag(lam) (exact products), ||R_j||^2 reduced, W = D^-1 R
// (or W = R when a preconditioner follows); W = 0 for the pairs in conv.
template <int K>
__global__ void __launch_bounds__(256)
k_residual(int64_t n, double* __restrict__ S, const double* __restrict__ AS,
           const double* __restrict__ BS, const double* __restrict__ lamg,
           const double* __restrict__ dinv, double* __restrict__ partials,
           unsigned int* __restrict__ counter, double* __restrict__ res_out,
           const int* __restrict__ conv) {
  __shared__ double lam[K], sh[8], bs[K];
  __shared__ int lk[K];
  const int tid = threadIdx.x;
  if (tid < K) {
    lam[tid] = lamg[tid];
    lk[tid] = conv ? conv[tid] : 0;
  }
  __syncthreads();
  double nrm[K];
#pragma unroll
  for (int j = 0; j < K; ++j) nrm[j] = 0.0;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + tid; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
    const double di = dinv ? dinv[r] : 1.0;
#pragma unroll
    for (int j = 0; j < K; ++j) {
      const double res = AS[(int64_t)j * n + r] - lam[j] * BS[(int64_t)j * n + r];
      nrm[j] += res * res;
      S[(int64_t)(K + j) * n + r] = lk[j] ? 0.0 : di * res;   // locked: W stays zero
    }
  }
#pragma unroll
  for (int j = 0; j < K; ++j) {
    const double t = block_sum(nrm[j], sh);
    if (tid == 0) bs[j] = t;
  }
  __syncthreads();
  if (tid == 0)
    for (int j = 0; j < K; ++j) partials[(int64_t)blockIdx.x * 12 + j] = bs[j];
}

// Residual norms from the k_residual partials (fixed order; one block).
__global__ void __launch_bounds__(256)
k_residual_finish(int k, int nb, const double* __restrict__ partials, double* __restrict__ res_out) {
  __shared__ double sh[8];
  for (int j = 0; j < k; ++j) {
    double t = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) t += partials[(int64_t)b * 12 + j];
    t = block_sum(t, sh);
    if (threadIdx.x == 0) res_out[j] = t;
  }
}

__global__ void k_copy_cols(int64_t n, int k, const double* __restrict__ a, double* __restrict__ b) {
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n * k;
       e += (int64_t)gridDim.x * blockDim.x)
    b[e] = a[e];
}

// ---------------------------------------------------------------------------
static unsigned grid_blocks(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (unsigned)(b < 1024 ? (b < 1 ? 1 : b) : 1024);
}

static int gram_and_rr(LobpcgWork& w, int m, int k, hipStream_t st, double drop_tol) {
  hipLaunchKernelGGL(k_gram, dim3(w.gram_chunks), dim3(256), gram_lds_bytes(m), st, w.n, m, w.S,
                     w.AS, w.BS, w.gram_partials);
  VOX8_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_gram_finish, dim3(2 * m * m), dim3(64), 0, st, w.gram_chunks, m,
                     (const double*)w.gram_partials, w.G);
  VOX8_HIP(hipGetLastError());
  double* C = w.G + 2 * MMAX * MMAX;
  double* lam = C + MMAX * 12;
  double* status = lam + 12;
  static const double jeps = [] {
    const char* e = getenv("VOX8_JACOBI_EPS");
    return e ? atof(e) : 1e-18;
  }();
  hipLaunchKernelGGL(k_rayleigh_ritz, dim3(1), dim3(RR_THREADS), 0, st, m, k, (const double*)w.G, C, lam,
                     status, drop_tol, jeps);
  VOX8_HIP(hipGetLastError());
  return OK;
}

template <int K>
static void launch_combine(const LobpcgWork& w, int nb, const double* C, hipStream_t st) {
  const dim3 g((unsigned)((w.n + 255) / 256)), b(256);
  const int* cv = w.conv;
  if (nb == 1) hipLaunchKernelGGL((k_ritz_combine<K, 1>), g, b, 0, st, w.n, w.S, C, cv);
  else if (nb == 2) hipLaunchKernelGGL((k_ritz_combine<K, 2>), g, b, 0, st, w.n, w.S, C, cv);
  else hipLaunchKernelGGL((k_ritz_combine<K, 3>), g, b, 0, st, w.n, w.S, C, cv);
}

// Soft locking: a pair whose residual norm fell below tol (sticky, as the
// host's convergence test) leaves the search space: its W and P columns and
// their K, M products are zeroed, so the Rayleigh-Ritz drops them, and the
// preconditioner and the K, M products of W skip the column.  Its
// X column keeps being refined by the other directions.  Without this the
// noise-level W of converged pairs, normalised by the K scaling, pollutes the
// basis and can stall the remaining pairs.
// Pairs locked in earlier iterations need nothing here: the residual and
// Ritz-update kernels write their W and P as zero and the products skip
// them.  Only a pair whose residual fell below tol in this iteration has its
// six columns zeroed; in an iteration with none the kernel reads the flags and
// returns.  The flags themselves are updated by k_lock_flags afterwards, so
// every block of this kernel sees the same old flags.
template <int K>
__global__ void __launch_bounds__(256)
k_lock(int64_t n, double* __restrict__ S, double* __restrict__ AS, double* __restrict__ BS,
       const double* __restrict__ res2, const int* __restrict__ conv, double tol2) {
  __shared__ int lk[K], any;
  if (threadIdx.x == 0) any = 0;
  __syncthreads();
  if (threadIdx.x < K) {
    const int j = threadIdx.x;
    lk[j] = !conv[j] && fabs(res2[j]) < tol2;             // newly converged
    if (lk[j]) any = 1;
  }
  __syncthreads();
  if (!any) return;
  for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
       r += (int64_t)gridDim.x * blockDim.x) {
#pragma unroll
    for (int j = 0; j < K; ++j)
      if (lk[j]) {
        S[(int64_t)(K + j) * n + r] = 0.0;
        S[(int64_t)(2 * K + j) * n + r] = 0.0;
        AS[(int64_t)(K + j) * n + r] = 0.0;
        BS[(int64_t)(K + j) * n + r] = 0.0;
        AS[(int64_t)(2 * K + j) * n + r] = 0.0;
        BS[(int64_t)(2 * K + j) * n + r] = 0.0;
      }
  }
}

// conv[j] |= (residual below tol): after k_lock, in stream order.
__global__ void k_lock_flags(int k, const double* __restrict__ res2, int* __restrict__ conv, double tol2) {
  const int j = threadIdx.x;
  if (j < k && fabs(res2[j]) < tol2) conv[j] = 1;
}

template <int K>
static void launch_lock(const LobpcgWork& w, double tol2, hipStream_t st) {
  hipLaunchKernelGGL((k_lock<K>), dim3(grid_blocks(w.n)), dim3(256), 0, st, w.n, w.S, w.AS, w.BS,
                     (const double*)w.res, (const int*)w.conv, tol2);
  hipLaunchKernelGGL(k_lock_flags, dim3(1), dim3(64), 0, st, K, (const double*)w.res, w.conv, tol2);
}

__global__ void k_zero_int(int n, int* __restrict__ a) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = 0;
}

template <int K>
static void launch_residual(const LobpcgWork& w, const double* lam, const double* dinv, hipStream_t st) {
  const unsigned nb = grid_blocks(w.n);
  hipLaunchKernelGGL((k_residual<K>), dim3(nb), dim3(256), 0, st, w.n, w.S,
                     (const double*)w.AS, (const double*)w.BS, lam, dinv, w.res_partials, w.counters,
                     w.res, (const int*)w.conv);
  hipLaunchKernelGGL(k_residual_finish, dim3(1), dim3(256), 0, st, K, (int)nb,
                     (const double*)w.res_partials, w.res);
}

#define VOX8_K_SWITCH(k, F, ...)                                                              \
  switch (k) {                                                                              \
    case 1: F<1>(__VA_ARGS__); break;  case 2: F<2>(__VA_ARGS__); break;                    \
    case 3: F<3>(__VA_ARGS__); break;  case 4: F<4>(__VA_ARGS__); break;                    \
    case 5: F<5>(__VA_ARGS__); break;  case 6: F<6>(__VA_ARGS__); break;                    \
    case 7: F<7>(__VA_ARGS__); break;  case 8: F<8>(__VA_ARGS__); break;                    \
    case 9: F<9>(__VA_ARGS__); break;  case 10: F<10>(__VA_ARGS__); break;                  \
    case 11: F<11>(__VA_ARGS__); break; default: F<12>(__VA_ARGS__); break;                 \
  }

// X' (and P') from the Ritz coefficients, their exact K and M products by
// the stencil, residual; tol2 > 0 also soft-locks converged pairs.  Pairs
// locked earlier skip the P products (their P, KP, MP are zeroed by the lock).
// w32 (may be null): the preconditioner's FP32 input, where the residual goes
// instead of W when the product forms it (*in32 is then true).
static int ritz_step(LobpcgWork& w, int nb, int k, const double* dinv,
                     const ApplyKM& applyKM, hipStream_t st, double tol2, float* w32 = nullptr,
                     bool* in32 = nullptr) {
  const double* C = w.G + 2 * MMAX * MMAX;
  const double* lam = C + MMAX * 12;
  VOX8_K_SWITCH(k, launch_combine, w, nb, C, st);
  VOX8_HIP(hipGetLastError());
  // the residual with the products of X' where the product kernel can form
  // it (W is free: P' = [W P] C was formed above)
  ResidualFuse rf{lam, w.conv, dinv, w.S + (int64_t)k * w.n, w32, w.res_partials,
                  w.res_partials_cap, w.res};
  bool fused = false;
  VOX8_TRY(applyKM(w.S, w.AS, w.BS, k, st, nullptr, &rf, &fused));
  if (in32) *in32 = fused && w32;
  if (nb > 1) {
    const int64_t off = (int64_t)2 * k * w.n;
    VOX8_TRY(applyKM(w.S + off, w.AS + off, w.BS + off, k, st, w.conv, nullptr, nullptr));
  }
  if (!fused) {
    VOX8_K_SWITCH(k, launch_residual, w, lam, dinv, st);
    VOX8_HIP(hipGetLastError());
  }
  if (tol2 > 0.0) {
    VOX8_K_SWITCH(k, launch_lock, w, tol2, st);
    VOX8_HIP(hipGetLastError());
  }
  return OK;
}

// host mirror layout: lam[12] | res[12] | status
static int read_state(LobpcgWork& w, int k, hipStream_t st) {
  double* C = w.G + 2 * MMAX * MMAX;
  double* lam = C + MMAX * 12;
  VOX8_HIP(hipMemcpyAsync(w.host, lam, sizeof(double) * 13, hipMemcpyDeviceToHost, st));
  VOX8_HIP(hipMemcpyAsync(w.host + 13, w.res, sizeof(double) * k, hipMemcpyDeviceToHost, st));
  VOX8_HIP(hipStreamSynchronize(st));
  return OK;
}

int lobpcg_fp64(const PcgOps<double>& A, const ApplyKM& applyKM,
                int kw, int guard, double tol, int maxiter, const double* X0, double* Xout,
                double* lam_out, LobpcgWork& w, hipStream_t st, LobpcgResult* res, uint64_t seed,
                const Precond<double>* M, float* M32) {
  const int k = kw + guard;                     // iterated block: wanted pairs + guard columns
  VOX8_CHECK(kw >= 1 && guard >= 0 && guard <= LOBPCG_GUARD_MAX && k <= 12 && k <= w.kmax,
             "lobpcg: block size must be 1..12");
  const int64_t n = w.n;
  double* X = w.S;
  double* W = w.S + (int64_t)k * n;
  double* AW = w.AS + (int64_t)k * n;
  double* BW = w.BS + (int64_t)k * n;
  // deflation: drop basis directions whose scaled K-Gram eigenvalue is below
  // drop_tol * max (numerically dependent in the energy norm)
  double drop_tol = 1e-13;
  if (const char* e = getenv("VOX8_LOBPCG_DROP")) drop_tol = atof(e);
  if (X0) {
    hipLaunchKernelGGL(k_copy_cols, dim3(grid_blocks(n * kw)), dim3(256), 0, st, n, kw, X0, X);
    VOX8_HIP(hipGetLastError());
    if (guard > 0) {
      // guard columns: the previous solve's, else random
      if (w.guard_cols >= guard)
        hipLaunchKernelGGL(k_copy_cols, dim3(grid_blocks(n * guard)), dim3(256), 0, st, n, guard,
                           (const double*)w.guard, X + (int64_t)kw * n);
      else
        hipLaunchKernelGGL(k_random_block, dim3(grid_blocks(n)), dim3(256), 0, st, n, guard,
                           seed ^ 0x5bd1e995ull, A.fixed, X + (int64_t)kw * n);
      VOX8_HIP(hipGetLastError());
    }
  } else {
    hipLaunchKernelGGL(k_random_block, dim3(grid_blocks(n)), dim3(256), 0, st, n, k, seed, A.fixed, X);
    VOX8_HIP(hipGetLastError());
  }
  // initial Rayleigh-Ritz on X alone
  VOX8_TRY(applyKM(X, w.AS, w.BS, k, st, nullptr, nullptr, nullptr));
  const double* Tdinv = M ? nullptr : A.dinv;
  hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, 12, w.conv);
  VOX8_HIP(hipGetLastError());
  const double tol2 = tol * tol;
  // with a preconditioner reading FP32 (M32), the residual of a fused
  // product goes straight there: W is then written only by the preconditioner
  // (locked pairs' W stays the zero the lock wrote)
  float* w32 = M ? M32 : nullptr;
  bool in32 = false;
  VOX8_TRY(gram_and_rr(w, k, k, st, drop_tol));
  VOX8_TRY(ritz_step(w, 1, k, Tdinv, applyKM, st, tol2, w32, &in32));
  if (M) VOX8_TRY((*M)(in32 ? nullptr : W, W, k, st, w.conv));
  VOX8_TRY(read_state(w, k, st));
  VOX8_CHECK(w.host[12] == 0.0, "lobpcg: linearly dependent initial block");
  std::vector<int> conv(kw, 0);
  auto update_conv = [&]() {
    bool all = true;
    for (int j = 0; j < kw; ++j) {
      if (sqrt(fabs(w.host[13 + j])) < tol) conv[j] = 1;
      all = all && conv[j];
    }
    return all;
  };
  bool all = update_conv();
  int it = 0;
  bool have_p = false;
  // VOX8_LOBPCG_LOG=1: one stderr line per iteration (status, lam, residual
  // norms, sticky convergence flags), from the state read back anyway
  static const bool log = [] {
    const char* e = getenv("VOX8_LOBPCG_LOG");
    return e && e[0] != '0';
  }();
  auto trace = [&]() {
    if (!log) return;
    fprintf(stderr, "[vox8 lobpcg] it %d status %g |", it, w.host[12]);
    for (int j = 0; j < k; ++j) fprintf(stderr, " %.10e", w.host[j]);
    fprintf(stderr, " |");
    for (int j = 0; j < k; ++j) fprintf(stderr, " %.2e%s", sqrt(fabs(w.host[13 + j])), j < kw && conv[j] ? "*" : "");
    fprintf(stderr, "\n");
  };
  trace();
  while (!all && it < maxiter) {
    VOX8_TRY(applyKM(W, AW, BW, k, st, w.conv, nullptr, nullptr));   // locked pairs: W, KW, MW stay zero
    const int m = have_p ? 3 * k : 2 * k;
    // the Rayleigh-Ritz kernel restarts without P on the device when needed
    VOX8_TRY(gram_and_rr(w, m, k, st, drop_tol));
    VOX8_TRY(ritz_step(w, m / k, k, Tdinv, applyKM, st, tol2, w32, &in32));
    if (M) VOX8_TRY((*M)(in32 ? nullptr : W, W, k, st, w.conv));
    have_p = true;
    ++it;
    VOX8_TRY(read_state(w, k, st));
    if (w.host[12] != 0.0) break;           // basis collapsed: X was kept, stop here
    all = update_conv();
    trace();
  }
  // final Rayleigh-Ritz on X alone (scipy lobpcg post-processing); AX, BX
  // are exact products of the current X
  VOX8_TRY(gram_and_rr(w, k, k, st, drop_tol));
  VOX8_TRY(ritz_step(w, 1, k, Tdinv, applyKM, st, 0.0));
  VOX8_TRY(read_state(w, k, st));
  hipLaunchKernelGGL(k_copy_cols, dim3(grid_blocks(n * kw)), dim3(256), 0, st, n, kw,
                     (const double*)X, Xout);
  VOX8_HIP(hipGetLastError());
  if (guard > 0) {
    hipLaunchKernelGGL(k_copy_cols, dim3(grid_blocks(n * guard)), dim3(256), 0, st, n, guard,
                       (const double*)(X + (int64_t)kw * n), w.guard);
    VOX8_HIP(hipGetLastError());
    w.guard_cols = guard;
  }
  for (int j = 0; j < k; ++j) w.ritz[j] = w.host[j];
  w.ritz_n = k;
  double mx = 0.0;
  for (int j = 0; j < kw; ++j) {
    lam_out[j] = w.host[j];
    mx = fmax(mx, sqrt(fabs(w.host[13 + j])));
  }
  VOX8_HIP(hipStreamSynchronize(st));
  res->iters = it;
  res->info = all ? 0 : (it > 0 ? it : 1);
  res->max_res = mx;
  return OK;
}

}  // namespace vox8
