// vox8 — geometric multigrid preconditioner on the structured voxel grid.
#pragma once
#include <vector>

#include "vox8_fast.h"

namespace vox8 {

// Layout of MultiGrid::scal: [MG_SCAL_NORM] start-vector scale, [7] coarse
// pivot flag, from MG_SCAL_LZ: per smoothed level l, MG_LANCZOS_STEPS pairs
// (alpha_j, beta_j) at MG_SCAL_LZ + 2 * MG_LANCZOS_STEPS * l + 2 j.
constexpr int MG_LANCZOS_STEPS = 20;
constexpr int MG_MAX_LEVELS = 16;
constexpr int MG_SCAL_NORM = 0;
constexpr int MG_SCAL_LZ = 8;
constexpr int MG_SCAL_SIZE = MG_SCAL_LZ + 2 * MG_LANCZOS_STEPS * MG_MAX_LEVELS;

struct MGLevel {
  Grid g{};
  double scale = 1.0;          // element matrix of this level = scale * Ke (box scaled by 2^l)
  int pattern = PAT_DENSE;
  FastCoef<float> c32{};
  double diag24[24]{};         // diagonal of the level's element matrix
  double* fK64 = nullptr;      // element factors (FP64 master copy)
  float* fK = nullptr;         // element factors used by the FP32 smoother
  uint8_t* active = nullptr;
  uint8_t* fixed = nullptr;    // per-node fixed-DOF bitmask
  float* dinv = nullptr;       // Jacobi inverse diagonal
  double lam_max = 2.0;        // estimate of lambda_max(D^-1 K)
  float* pv = nullptr;         // Lanczos vector (the other one is d)
  double* E64 = nullptr;       // Galerkin levels: per-element 24x24 matrices (FP64)
  float* S32 = nullptr;        //   and the FP32 node blocks the smoother applies,
  bool s_by_entry = false;     //   stored entry-major (a node per lane) or node-major
  // work vectors, ncols column-major blocks of n_dofs
  float *x = nullptr, *r = nullptr, *b = nullptr, *d = nullptr, *d2 = nullptr;
};

// One captured cycle: the launches of mg_apply for fixed (r, z, ncols, skip)
// replayed as a HIP graph (a W-cycle is ~150 launches, most of them tiny
// coarse-level kernels the host cannot issue as fast as the GPU runs them).
struct MGGraph {
  const double* r = nullptr;
  double* z = nullptr;
  int ncols = 0;
  const int* skip = nullptr;
  uint64_t epoch = 0;
  uint64_t last_use = 0;
  hipGraphExec_t exec = nullptr;
};

struct MultiGrid {
  std::vector<MGLevel> lv;
  int ncols = 0;               // capacity of the work vectors
  int nu = 2;                  // Chebyshev smoothing steps before and after
  int nu_coarse = 4;           // the same on levels >= nu_coarse_level (0: nu); the smaller
  int nu_coarse_level = 2;     // levels' kernels are latency-bound, so more steps there cost
                               // little against the coarse corrections they improve
  int nu_level[8] = {};        // per-level override of the two rules above (0: the rules)
  int gamma_level[8] = {};     // per-level override of the coarse-correction count (0: rules)
  int gamma = 2;               // coarse corrections per level (1 = V-cycle, 2 = W-cycle)
  int v_top = 0;               // the finest v_top levels take one correction (V) whatever gamma
  int64_t w_min_dofs = 5000;   // levels smaller than this take one correction (V) even when
                               // gamma = 2: their kernels are launch-latency bound
  int coarse_dofs = 2000;      // stop coarsening at this many DOFs (dense inverse)
  int64_t lane_nodes = 65536;  // Galerkin levels with at least this many nodes apply one node
                               // per lane; smaller ones spread a node over 8 lanes
  double alpha = 20.0;         // smoothing interval [lam_max/alpha, lam_max]
  bool galerkin = true;        // coarse levels P^T A P (element matrices); false: rediscretised
  bool operators_only = false; // the per-density operators come from another hierarchy
                               // (mg_copy_operators): no element matrices are allocated
  double lam_margin = 1.1;     // lambda_max = margin x the largest Lanczos Ritz value
  // coarsest level: dense inverse when small
  int n_coarse = 0;
  double* A = nullptr;         // n_coarse^2 FP64 (assembled, then inverted)
  double* A2 = nullptr;        // second n_coarse^2 buffer of the Gauss-Jordan ping-pong
  float* Ainv = nullptr;       // FP32 copy used in the V-cycle
  double* red = nullptr;       // reduction scratch
  unsigned int* counter = nullptr;
  double* scal = nullptr;      // MG_SCAL_SIZE device scalars (Lanczos, coarse pivot check)
  double* host = nullptr;      // pinned copy of scal
  float* tmp32 = nullptr;      // fine-level input/output staging (ncols * n0)
  bool built = false;
  bool updated = false;
  uint64_t updates = 0;        // completed mg_update / mg_copy_operators calls
  // captured cycles; epoch changes whenever a captured launch sequence would
  // (hierarchy, work vectors, coefficients, options)
  bool use_graphs = true;
  uint64_t epoch = 1;
  uint64_t use_clock = 0;
  int capture_failures = 0;
  std::vector<MGGraph> graphs;
  hipStream_t cap_st = nullptr;   // capture stream of a context on the null stream (which
                                  // cannot be captured); created on first capture
  MultiGrid() = default;
  MultiGrid(const MultiGrid&) = delete;
  MultiGrid& operator=(const MultiGrid&) = delete;
  ~MultiGrid();
};

// Galerkin coarse levels (vox8_mg_galerkin.hip).  gal_build_level: element
// matrices of coarse grid gc from fine grid gf (fK0/dKe for the finest
// level, else the fine level's element matrices Ef64).
int gal_build_level(const Grid& gf, const Grid& gc, const double* fK0, const double* dKe,
                    const double* Ef64, const uint8_t* fixf, const uint8_t* fixc, double* Ec64,
                    float* S32, bool by_entry, hipStream_t st);
int gal_jacobi(const Grid& g, const double* E64, const uint8_t* fixed, float* dinv, hipStream_t st);
int gal_coarse_assemble(const Grid& g, const double* E64, const uint8_t* fixed, double* A,
                        hipStream_t st);
int gal_apply(const Grid& g, const float* S32, bool by_entry, const uint8_t* fixed, const float* x,
              float* y, int ncols, const Epilogue<float>& ep, hipStream_t st);

// Destroy every captured cycle (device state is unchanged).
void mg_drop_graphs(MultiGrid& mg);

// Allocation helpers are supplied by the context (device/pinned, tracked).
struct MGAlloc {
  void* ctx;
  int (*dev)(void* ctx, void** p, size_t bytes);
  int (*pinned)(void* ctx, void** p, size_t bytes);
};

// Build the level hierarchy for the fine grid and its host masks.
int mg_build(MultiGrid& mg, const Grid& g0, const uint8_t* h_active, const uint8_t* h_fixed,
             const double* Ke, int pattern0, const FastCoef<float>& c0, float* fK0_32,
             uint8_t* d_active0, uint8_t* d_fixed0, const MGAlloc& al, hipStream_t st);
// Per-density setup: restrict factors, Jacobi diagonals, lambda_max, coarse inverse.
int mg_update(MultiGrid& mg, const double* fK0, const double* dKe, hipStream_t st);
// Ensure work vectors for ncols columns.
int mg_reserve(MultiGrid& mg, int ncols, const MGAlloc& al);
// Take the per-density operators of src (a hierarchy of the same grid, masks
// and options, updated for the density dst is to use): Galerkin node blocks
// or level factors, Jacobi diagonals, lambda_max and the coarsest inverse are
// copied on st, so dst needs no update of its own.
int mg_copy_operators(MultiGrid& dst, const MultiGrid& src, hipStream_t st);
// z = cycle(r) (V or W, see MultiGrid::gamma) for ncols FP64 column-major
// vectors (ld = fine n_dofs).  skip (device, per column, may be null):
// non-zero columns are left untouched (z keeps its contents).  r == nullptr:
// the input is already in the finest level's FP32 right-hand side lv[0].b
// (written there by the caller's last kernel), so no conversion runs.
int mg_apply(MultiGrid& mg, const double* r, double* z, int ncols, hipStream_t st,
             const int* skip = nullptr);

}  // namespace vox8
