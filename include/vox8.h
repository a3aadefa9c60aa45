/* vox8 — C ABI of the MI355X-native voxel hex8 FEM hot path.
 *
 * Plain C types only: device buffers are raw pointers (allocated by the caller,
 * e.g. torch tensors), host buffers are marked h_, device buffers d_.  Every
 * call returns 0 on success and a negative status on error; vox8_last_error()
 * then holds a message.  Solver calls report non-convergence through `info`
 * (scipy convention: 0 converged, >0 iterations done without converging), not
 * through the return status.  A context is bound to one HIP stream and is not
 * thread-safe; separate contexts are independent.
 *
 * Grid layout: nodes n = ix + (nx+1)(iy + (ny+1) iz), DOF 3n+c, elements
 * e = ex + nx(ey + ny ez); blocks of vectors are column-major with leading
 * dimension `ld` (>= 3 * n_nodes).
 *
 * The reference (NinhGhoster/TO-F450-arm) is NumPy/SciPy with a CuPy
 * monkey-patch; each entry point names the reference function it replaces.
 */
#ifndef VOX8_H
#define VOX8_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct vox8_ctx vox8_ctx;

enum { VOX8_OP_K = 1, VOX8_OP_M = 2, VOX8_OP_KM = 3 };
enum { VOX8_PREC_FP64 = 0, VOX8_PREC_MIXED = 1, VOX8_PREC_FP32 = 2 };

/* Last error message of the calling thread ("" if none). */
const char* vox8_last_error(void);
/* ABI version (major*100 + minor). */
int vox8_version(void);
/* Number of visible HIP devices (replaces cp.cuda.runtime.getDeviceCount,
 * src/fem_gpu.py:37, 42-43). */
int vox8_device_count(int* h_count);

/* Context for one nx x ny x nz voxel box on the current device, bound to
 * `hip_stream` (NULL = default stream).  Replaces FEM3D.__init__'s
 * per-problem state (src/fem.py:207-308); no COO/CSR pattern is built. */
int vox8_create(vox8_ctx** out, int nx, int ny, int nz, void* hip_stream);
int vox8_destroy(vox8_ctx* ctx);
int vox8_set_stream(vox8_ctx* ctx, void* hip_stream);

/* Unit-density element matrices, host 24x24 row-major, DOF order
 * [n0x,n0y,n0z,n1x,...] in the reference local node order.  Replaces
 * FEM3D.Ke_solid / Me_solid (src/fem.py:211-221); call again when a caller
 * mutates Ke_solid (anisotropy study).  Selects the structured stencil when Ke
 * has the box-element block structure, else the general 24x24 stencil. */
int vox8_set_element_matrices(vox8_ctx* ctx, const double* h_Ke, const double* h_Me);

/* Stencil selection: kind 0 = automatic (structured Hadamard-basis kernel
 * when Ke/Me have the box-element structure, else dense), 1 = always the
 * general dense 24x24 kernel.  *h_pattern (may be NULL) receives the kernel
 * in use: -1 dense, 0 orthotropic pattern (45 non-zeros), 1 general blocks. */
int vox8_set_stencil(vox8_ctx* ctx, int kind, int* h_pattern);

/* Active elements (h_elem_active[n_elem], 1 = in the model) and the fixed-DOF
 * bitmask per node (bit c = DOF 3n+c fixed; ghost nodes must be fixed).
 * Replaces the masks / fixed_dofs / ghost-DOF logic of src/fem.py:225-305. */
int vox8_set_topology(vox8_ctx* ctx, const uint8_t* h_elem_active, const uint8_t* h_node_fixed);

/* Lumped tip mass: mass_kg / n_nodes on the three translational DOFs of every
 * listed node (src/fem.py:383-398).  n_nodes = 0 clears it. */
int vox8_set_tip_mass(vox8_ctx* ctx, const int64_t* h_nodes, int n_nodes, double mass_kg);

/* Element densities d_rho[n_elem] (element grid order).  Computes the SIMP
 * factor e_min + (1-e_min) rho^p and the Pedersen mass factor on the device
 * (src/fem.py:316-320 assemble_K, 334-349 _mass_simp_factor).  Returns when
 * the factors are formed; factors equal to the current ones (the same
 * density set again) leave what is derived from them (multigrid hierarchy,
 * Jacobi diagonal) in place. */
int vox8_set_density(vox8_ctx* ctx, const double* d_rho, double p, double e_min,
                     double mass_threshold, double mass_q);

/* Y = K X and/or M X on the free DOFs for nrhs column-major vectors (X read as
 * zero at fixed DOFs, Y zero there).  prec FP64: double buffers; FP32: float
 * buffers.  Replaces assemble_K/assemble_M + CSR SpMV (src/fem.py:311-398,
 * K_ff @ x in scipy cg / lobpcg). */
int vox8_apply(vox8_ctx* ctx, int op, int prec, const void* d_X, void* d_YK, void* d_YM,
               int nrhs, int64_t ld);

/* Diagonal of K_ff or M_ff (zeros at fixed DOFs), K_ff.diagonal() of
 * src/fem.py:441,543. */
int vox8_diagonal(vox8_ctx* ctx, int op, double* d_diag);

/* Jacobi-PCG on K_ff U = F_f for nrhs load cases at once (independent CGs
 * sharing every launch).  d_U holds x0 when x0_given, else is overwritten.
 * Semantics of scipy.sparse.linalg.cg as called at src/fem.py:543-547 and
 * src/fem_gpu.py:61-81: stop when ||r|| < rtol ||F_f||, maxiter -> info.
 * prec MIXED: FP32 inner PCG inside FP64 iterative refinement on the FP64
 * residual (same stop test).  h_iters/h_info/h_relres: arrays of nrhs. */
int vox8_pcg(vox8_ctx* ctx, const double* d_F, double* d_U, int nrhs, int64_t ld, int x0_given,
             double rtol, int maxiter, int prec, int* h_iters, int* h_info, double* h_relres);

/* k smallest eigenpairs of K_ff x = lam M_ff x (M includes the tip lump) by
 * Jacobi-preconditioned LOBPCG; replaces scipy.sparse.linalg.lobpcg at
 * src/fem.py:450-454 and src/fem_gpu.py:149-152.  d_X0 (k column-major
 * vectors, may be NULL for a seeded random block), d_X receives M-orthonormal
 * modes; h_lam[k] ascending.  Stop: every column's absolute residual
 * ||K x - lam M x||_2 < tol (scipy semantics). */
int vox8_lobpcg(vox8_ctx* ctx, int k, double tol, int maxiter, int prec, const double* d_X0,
                double* d_X, double* h_lam, int* h_iters, int* h_info, double* h_max_res,
                uint64_t seed);
/* Guard columns of vox8_lobpcg (no reference counterpart; scipy's lobpcg has
 * none): `guard` (0..2) extra columns are iterated with the k wanted ones but
 * not returned or tested for convergence, so the last wanted pair converges
 * at the rate set by lambda_{k+guard+1} instead of lambda_{k+1}; a warm start
 * reuses the previous solve's guard columns.  -1 restores the default
 * (environment VOX8_LOBPCG_GUARD, else 0). */
int vox8_set_lobpcg_guard(vox8_ctx* ctx, int guard);
/* Ritz values of the last vox8_lobpcg solve's whole block, ascending: the k
 * wanted eigenvalues followed by those of the guard columns (k + guard
 * values; *n_out = 0 before the first solve).  Copies at most n_max into
 * h_out.  No reference counterpart: it lets a caller see the gap between the
 * last wanted eigenvalue and the next one, which decides whether guard
 * columns pay (vox8.topopt.SIMPOptimizer's "auto" policy). */
int vox8_lobpcg_ritz(vox8_ctx* ctx, double* h_out, int n_max, int* n_out);

/* The geometric-multigrid preconditioner used by the MIXED solvers:
 * z = cycle(r) (W-cycle by default) for nrhs FP64 column-major vectors (ld = n_dofs); r and z
 * may alias.  No reference counterpart: the reference preconditions with the
 * Jacobi diagonal (src/fem.py:441-443, 543-547). */
int vox8_precond_apply(vox8_ctx* ctx, const double* d_r, double* d_z, int nrhs);
/* One FP32 Chebyshev smoothing step on the fine grid, the building block of
 * the multigrid preconditioner: d_xs += d_d; d_r -= K d_d;
 * d_dnew = c1 d_d + c2 D^-1 d_r (D = diag K), nrhs column-major vectors
 * (ld = n_dofs).  No reference counterpart (the reference preconditions with
 * the Jacobi diagonal only, src/fem.py:543-547); exposed for custom smoothers
 * and for timing the dominant kernel.  Needs the structured stencil. */
int vox8_cheb_step(vox8_ctx* ctx, const float* d_d, float* d_r, float* d_xs, float* d_dnew,
                   double c1, double c2, int nrhs);
/* One FP32 Chebyshev smoothing step in the three-term form on the iterates,
 * the step the multigrid preconditioner runs on the fine grid:
 * d_xnew = d_xk + c1 (d_xk - d_xprev) + c2 D^-1 (d_b - K d_xk) (D = diag K;
 * d_xprev may be null: the c1 term is then dropped), zero at fixed DOFs;
 * nrhs column-major vectors (ld = n_dofs), d_xnew aliasing no input.  No
 * reference counterpart (the reference preconditions with the Jacobi
 * diagonal only, src/fem.py:543-547); exposed for custom smoothers and for
 * timing the dominant kernel.  Needs the structured stencil. */
int vox8_cheb_xstep(vox8_ctx* ctx, const float* d_xk, const float* d_xprev, const float* d_b,
                    float* d_xnew, double c1, double c2, int nrhs);

/* Level hierarchy: element counts per level (h_dims[3*l..3*l+2]) and the
 * smoother's lambda_max(D^-1 K) estimate per level. */
int vox8_mg_info(vox8_ctx* ctx, int* h_nlevels, int* h_dims, double* h_lam_max, int max_levels);
/* Multigrid options: Chebyshev steps before/after (nu), smoothing interval
 * [lam_max/alpha, lam_max], coarsest-level size (DOFs) for the dense inverse,
 * cycle index (1 V-cycle, 2 W-cycle) applied on levels with at least
 * w_min_dofs DOFs (smaller levels take one coarse correction).  Defaults:
 * nu 2, alpha 20, coarse_dofs 2000 (at most 4096), cycle 2, w_min_dofs 5000. */
int vox8_set_mg_options(vox8_ctx* ctx, int nu, double alpha, int coarse_dofs, int cycle,
                        int64_t w_min_dofs);
/* Chebyshev steps before/after on the levels from from_level on (level 0 is
 * the voxel grid), in place of nu there; nu_coarse 0 = nu on every level.
 * Default 4 from level 2: those levels' kernels are latency-bound, and the
 * better coarse corrections cut PCG and LOBPCG iterations by 10-15 % on the
 * 440x64x48 SIMP iterations.  No reference counterpart. */
int vox8_set_mg_coarse_smoothing(vox8_ctx* ctx, int nu_coarse, int from_level);
/* ctx's multigrid preconditioner takes the per-density operators of src's
 * hierarchy (Galerkin blocks or level factors, Jacobi diagonals, smoothing
 * bounds, coarsest inverse) by a device copy on ctx's stream instead of
 * computing its own: for a second context of the same model solving with the
 * same density at the same time (a concurrent modal solve).  The caller sets
 * the same densities on both and brings src up to date first
 * (vox8_prepare_multigrid); src must not be updated while ctx's solve runs,
 * and must outlive the sharing (src NULL ends it).  ctx keeps its own work
 * vectors and captured cycles.  No reference counterpart. */
int vox8_share_multigrid(vox8_ctx* ctx, vox8_ctx* src);
/* Set up (or update for the current density) the multigrid hierarchy with
 * work space for ncols columns; returns when the device work is done.  No
 * reference counterpart. */
int vox8_prepare_multigrid(vox8_ctx* ctx, int ncols);
/* Coarse levels of the multigrid preconditioner: 1 = Galerkin
 * A_{l+1} = P^T A_l P (per-element 24x24 matrices, applied in node-block
 * form; the default), 0 = rediscretised (the box element of each level with
 * the mean of its children's factors, applied by the matrix-free structured
 * stencil: cheaper passes, but at high density contrast the cycle can lose
 * its effect and PCG stalls).  No reference counterpart (the reference
 * preconditions with the Jacobi diagonal). */
int vox8_set_mg_galerkin(vox8_ctx* ctx, int enable);

/* Centre-point element stresses, Voigt order (xx, yy, zz, xy, yz, xz), linear
 * in the density: d_stress[6e + i] = rho[e] (D B u_e)_i for every element of
 * the grid (0 on inactive elements).  h_B is the 6x24 centre
 * strain-displacement matrix (row-major, columns in the element's local DOF
 * order), h_D the 6x6 constitutive matrix, as the reference builds them.
 * (FEM3D.compute_stress, src/fem.py:564-579) */
int vox8_stress(vox8_ctx* ctx, const double* d_u, const double* d_rho_grid, const double* h_B,
                const double* h_D, double* d_stress);

/* Von Mises stress and largest principal stress of n Voigt rows of any device
 * array (either output may be null), on the given HIP stream (null: the
 * default stream); no context needed.  (FEM3D.von_mises and
 * principal_stress_max, src/fem.py:582-602) */
int vox8_stress_invariants(void* hip_stream, int64_t n, const double* d_stress, double* d_vm,
                           double* d_pmax);

/* Goodman / S-N fatigue fields of two load-case stress fields (n elements,
 * Pa): mean, alternating, Goodman-equivalent alternating stress, life,
 * damage 1/max(life, 1) and factor of safety fos_ref_mpa*1e6/eq (outputs may be
 * null).  The S-N curve is a 7-point table of log10(stress, MPa), strictly
 * increasing, and log10(cycles); life = 1e9 below 0.99 endurance_mpa and 1
 * above 1.5 overload_mpa.  No context needed.  (fatigue_life_field and
 * sn_life, src/fatigue.py:24-121) */
int vox8_fatigue(void* hip_stream, int64_t n, const double* d_s1, const double* d_s2, double uts_pa,
                 const double* h_log_stress_mpa, const double* h_log_cycles, double endurance_mpa,
                 double overload_mpa, double fos_ref_mpa, double* d_mean, double* d_alt, double* d_eq,
                 double* d_life, double* d_damage, double* d_fos);

/* Replay each multigrid cycle as a captured HIP graph (default 1) or launch
 * its kernels one by one (0).  A cycle is captured once per (r, z, nrhs,
 * skip mask) and hierarchy state and replayed afterwards; results are the
 * same launches in the same order.  No counterpart in the reference (its
 * solvers are SciPy/CuPy calls). */
int vox8_set_mg_graphs(vox8_ctx* ctx, int enable);

/* Adjoint quantities for the current density:
 *   d_ce[e]  = sum_j w_j U_j,e^T Ke U_j,e      (elem_compliance, src/fem.py:605-615)
 *   d_dc[e]  = -p (1-e_min) rho^(p-1) d_ce[e]  (src/topopt.py:296-306)
 *   h_comp[j]= sum_e f_e U_j,e^T Ke U_j,e      (per-load-case compliance)
 * and, when d_phi != NULL, the eigenvalue sensitivity of (lam, phi)
 *   d_dlam[e] = ((1-e_min) p max(rho,1e-9)^(p-1) phi_e^T Ke phi_e
 *               - lam m'(rho) phi_e^T Me phi_e) / phi^T M phi   (src/fem.py:472-526)
 * with h_vmv = phi^T M phi (tip lump included).  Output pointers may be NULL. */
int vox8_sensitivities(vox8_ctx* ctx, const double* d_U, int nrhs, int64_t ld,
                       const double* h_w, const double* d_phi, double lam, const double* d_rho,
                       double p, double e_min, double mass_threshold, double mass_q,
                       double* d_ce, double* d_dc, double* d_dlam, double* h_comp,
                       double* h_vmv);

/* Cone density filter w(d) = max(0, r - d) on the element grid restricted to
 * d_mask (cone_kernel / filter_field, src/topopt.py:32-58).  setup builds the
 * taps and the masked denominator; filter applies it (d_in, d_out n_elem). */
int vox8_filter_setup(vox8_ctx* ctx, double r_min_vox, const uint8_t* d_mask);
int vox8_filter(vox8_ctx* ctx, const double* d_in, double* d_out);
/* Transpose of vox8_filter: d_out = H^T (d_in / den) on the mask, the exact
 * chain rule of the density filter (the reference filters sensitivities with
 * the forward filter instead, src/topopt.py:162; this entry backs the autograd
 * path and has no reference counterpart). */
int vox8_filter_adjoint(vox8_ctx* ctx, const double* d_in, double* d_out);

/* OC update with bisection on the multiplier (src/topopt.py:190-214) over
 * the elements with d_design[e] != 0; other entries of d_xnew copy d_x. */
int vox8_oc_update(vox8_ctx* ctx, const uint8_t* d_design, const double* d_x, const double* d_dc,
                   double V_e, double target_volume, double move, double* d_xnew,
                   double* h_lmid);

/* Loop bookkeeping reductions (src/topopt.py:351-354, 383-397, 178-188):
 * h_out4 = {sum a, sum |a|, max |a-b|, count} over d_mask (NULL = all). */
int vox8_stats(vox8_ctx* ctx, int64_t n, const uint8_t* d_mask, const double* d_a,
               const double* d_b, double* h_out4);
/* h_out[2k+1] = {X_j.v (k), X_j.X_j (k), v.v} for k column-major vectors. */
int vox8_block_dots(vox8_ctx* ctx, int64_t n, int k, const double* d_X, int64_t ld,
                    const double* d_v, double* h_out);
/* Allocation-free variant of vox8_block_dots.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (work
 * recorded under HIP graph capture, tight serving loops).  No context; the
 * same {X_j.v (k), X_j.X_j (k), v.v} go to DEVICE memory d_out (2k+1 doubles)
 * on `hip_stream` (NULL = default stream) with no host synchronisation.  The
 * reduction partials live in a fixed workspace of the library (one per GPU,
 * 512 x 33 doubles, kept until the process ends), so k is limited to 1..16
 * (larger k returns an error); calls that overlap on one GPU race. */
int vox8_block_dots_noalloc(void* hip_stream, int64_t n, int k, const double* d_X, int64_t ld,
                            const double* d_v, double* d_out);
/* y = a x + b y on d_mask (NULL = all). */
int vox8_axpby(vox8_ctx* ctx, int64_t n, const uint8_t* d_mask, double a, const double* d_x,
               double b, double* d_y);
/* out = design ? filt : (in_arm ? 1 : 0)  (src/topopt.py:273-281). */
int vox8_compose_density(vox8_ctx* ctx, const uint8_t* d_design, const uint8_t* d_in_arm,
                         const double* d_filt, double* d_out);

/* Device-resident tail of one SIMP iteration (src/topopt.py:327-399; replaces
 * the host-side mode tracking, penalty scaling and bookkeeping between
 * vox8_block_dots, vox8_sensitivities, vox8_filter, vox8_stats, vox8_axpby,
 * vox8_oc_update calls).  With k > 0 modes (columns of d_X, sorted by the
 * eigenvalues h_lam, clipped at 0): MAC tracking against d_vprev (have_prev;
 * else the lowest mode), the tracked mode copied into d_vprev, violation
 * max(0, 1 - lam/w_target); if positive, the eigenvalue sensitivity of the
 * tracked mode (tip mass as set) into d_dlam, filtered into d_dlam_f, and
 * d_sens_f += -mu viol (mean|sens_f| / max(mean|dlam_f|, 1e-30)) d_dlam_f over
 * the design.  Then the OC update of d_x into d_xnew, its filtered field into
 * d_check, and h_state[20] = {tracked mode, lam, violation, stats of sens_f
 * (4), of dlam_f (4), of d_check - d_phys_prev (4: [13] = change), of d_xnew
 * on the design (4: [15] = volume sum), penalty coefficient}.  One host
 * synchronisation; the launches are replayed from a captured HIP graph
 * (vox8_set_tail_graph). */
typedef struct vox8_tail_args {
  const double* d_X;         /* k column-major modes (ld ldx), sorted */
  const double* h_lam;       /* k eigenvalues, ascending, >= 0 */
  double* d_vprev;           /* n_dofs: previous tracked mode in, tracked mode out */
  const double* d_rho;       /* element-grid physical density */
  double* d_dlam;            /* n_elem scratch: eigenvalue sensitivity */
  double* d_dlam_f;          /* n_elem scratch: its filtered field */
  double* d_sens_f;          /* n_elem: filtered compliance sensitivity (updated) */
  const uint8_t* d_design;   /* n_elem design mask */
  const double* d_x;         /* n_elem design variables */
  double* d_xnew;            /* n_elem: OC result */
  double* d_check;           /* n_elem: filtered OC result */
  const double* d_phys_prev; /* n_elem: previous filtered design */
  int64_t ldx;
  double w_target, mu, p, e_min, mthr, mq, V_e, target_volume, move;
  int k, have_prev;
} vox8_tail_args;
int vox8_simp_tail(vox8_ctx* ctx, const vox8_tail_args* args, double* h_state);
/* Replay vox8_simp_tail's launches from captured HIP graphs (enable 1, the
 * default: direct launches if the runtime refuses the capture; 2: the same,
 * but a refused capture is an error) or launch them one by one (0).  A
 * graph is captured per set
 * of buffers and constants (at most four kept per context; the per-call
 * eigenvalues, penalty coefficient and mode-block address are uploaded before
 * each replay); the results are bitwise those of the direct launches.  No
 * reference counterpart: the reference's loop tail is host NumPy
 * (src/topopt.py:327-412). */
int vox8_set_tail_graph(vox8_ctx* ctx, int enable);

/* Modal effective mass of k modes (columns of d_X, leading dimension ld,
 * zero at fixed DOFs) in translation direction 0 (x), 1 (y) or 2 (z), with
 * the mass matrix of the current density and tip lump: r = unit translation
 * on the free DOFs,
 *   h_frac[j] = (X_j^T M r)^2 / (X_j^T M X_j) / (r^T M r),
 *   h_L[j] = X_j^T M r, h_gm[j] = X_j^T M X_j, *h_total = r^T M r
 * (any output may be NULL), FP64 reductions on the device.  Replaces
 * effective_mass_voxel, sims/modal_participation.py:37-57. */
int vox8_effective_mass(vox8_ctx* ctx, int k, const double* d_X, int64_t ld, int direction,
                        double* h_frac, double* h_L, double* h_gm, double* h_total);

/* MFMA trial (no reference counterpart; a measurement, not used by the
 * solvers): d_Y = K d_X in FP32 for nrhs column-major vectors, the 24x24
 * element contraction on the matrix cores (v_mfma_f32_32x32x2f32) with the
 * node assembly in LDS; zero at fixed DOFs.  Compare with vox8_apply(OP_K,
 * PREC_FP32), the structured stencil. */
int vox8_trial_mfma_apply(vox8_ctx* ctx, const float* d_X, float* d_Y, int nrhs);

#ifdef __cplusplus
}
#endif
#endif /* VOX8_H */
