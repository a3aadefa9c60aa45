// vox8 — launchers of the SIMP-loop kernels (vox8_topopt.hip).
// `gate` (optional, device): the launch does nothing unless *gate > 0, so a
// launch sequence can skip a branch decided on the device without a host read.
#pragma once
#include "vox8_fast.h"

namespace vox8 {
enum FilterMode { FILTER_FWD = 0, FILTER_ADJ = 1, FILTER_DEN = 2 };
int launch_filter(int nx, int ny, int nz, int ntap, const int* toff, const double* tw,
                  const uint8_t* mask, const double* den, const double* in, double* out,
                  int mode, hipStream_t st, const double* gate = nullptr);
int launch_quadforms(const Grid& g, const double* fK, const double* fM, const uint8_t* active,
                     const double* Ke, const double* Me, const double* U, int64_t ldu, int nrhs,
                     const double* wts, const double* phi, double* ce, double* kphi,
                     double* mphi, double* partials, unsigned int* counter, double* sums,
                     const int64_t* tip_nodes, int n_tip, double tip_m, hipStream_t st,
                     const double* gate = nullptr, const FastCoef<double>* mono = nullptr);
int launch_sens_finish(int64_t n, const uint8_t* active, const double* rho, double p, double emin,
                       double mthr, double mq, const double* ce, double* dc, const double* kphi,
                       const double* mphi, double lam, const double* vmv, double* dlam,
                       hipStream_t st, const double* lam_dev = nullptr, const double* gate = nullptr);
int launch_oc(int64_t n, const uint8_t* design, const double* x, const double* dc, double Ve,
              double target, double move, double* state, double* partials, unsigned int* counter,
              double* xnew, int max_steps, hipStream_t st);
int launch_stats(int64_t n, const uint8_t* mask, const double* a, const double* b, double* partials,
                 unsigned int* counter, double* out, hipStream_t st, const double* gate = nullptr);
// Xind (may be null): device address holding X, read by the kernel in place
// of X (a captured launch then follows the pointer stored there).
int launch_block_dots(int64_t n, int k, const double* X, int64_t ldx, const double* v,
                      double* partials, unsigned int* counter, double* out, hipStream_t st,
                      const double* const* Xind = nullptr);
int launch_axpby(int64_t n, const uint8_t* mask, double a, const double* x, double b, double* y,
                 hipStream_t st);
// Device-resident loop tail (state layout in vox8_topopt.hip).
constexpr int TAIL_STATE = 20;
// The values of one call that change from iteration to iteration, read by
// the tail's kernels from device memory, so the tail's launches can be
// captured once and replayed with new values.
struct TailPar {
  double lam[16];     // eigenvalues, ascending, clipped at 0
  double mu;          // frequency-penalty coefficient
  const double* X;    // the mode block (k column-major vectors)
};
struct TailSelect {
  double w_target;    // (2 pi f_target)^2
  int k;              // number of modes (0: no modal constraint)
  int have_prev;      // a previous tracked mode exists
};
int launch_tail_select(const TailSelect& s, const TailPar* par, const double* dots, double* state,
                       hipStream_t st);
int launch_gather_col(int64_t n, const TailPar* par, int64_t ldx, const double* state, double* v,
                      hipStream_t st);
int launch_tail_penalty(int64_t n, const TailPar* par, double* state, const double* x, double* y,
                        hipStream_t st);
int launch_compose(int64_t n, const uint8_t* design, const uint8_t* in_arm, const double* filt,
                   double* out, hipStream_t st);
// Modal effective mass: r = free-DOF unit translation; reductions of
// MX = M [X_0..X_{k-1}, r] (k+1 columns, ld 3 n_nodes) into out[3k+1]
// (fractions, L_j = X_j^T M r, g_j = X_j^T M X_j, r^T M r).  `partials`
// holds max_doubles doubles.
int launch_unit_translation(int64_t n_nodes, int dir, const uint8_t* fixed, double* r,
                            hipStream_t st);
int launch_effective_mass(int64_t n_nodes, int dir, int k, const double* X, int64_t ldx,
                          const double* MX, double* partials, int max_doubles, double* out,
                          hipStream_t st);
}  // namespace vox8
