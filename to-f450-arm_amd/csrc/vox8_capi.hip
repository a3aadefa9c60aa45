// vox8 — C-ABI entry points (declared in include/vox8.h) and the context object.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <new>

#include "../../include/vox8.h"
#include "vox8_fast.h"
#include "vox8_kernels.h"
#include "vox8_mfma_trial.h"
#include "vox8_mg.h"
#include "vox8_post.h"
#include "vox8_pcg.h"
#include "vox8_topopt.h"

namespace vox8 {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
std::recursive_mutex& hip_serial() {
  static std::recursive_mutex m;
  return m;
}
}  // namespace vox8

using namespace vox8;

struct vox8_ctx {
  Grid g{};
  hipStream_t st = nullptr;
  int device = 0;
  std::vector<void*> dev_allocs;
  std::vector<void*> host_allocs;
  double Ke[576]{}, Me[576]{};
  int pattern = PAT_DENSE;          // structure of Ke/Me found at set time
  int stencil_kind = 0;             // 0 auto, 1 force the dense 24x24 kernel
  FastCoef<double> c64{};
  FastCoef<float> c32{};
  bool have_mats = false, have_topo = false, have_density = false;
  double *dKe = nullptr, *dMe = nullptr;
  float *dKe32 = nullptr, *dMe32 = nullptr;
  uint8_t *d_active = nullptr, *d_fixed = nullptr;
  double *fK = nullptr, *fM = nullptr;
  float *fK32 = nullptr, *fM32 = nullptr;
  int64_t* d_tip = nullptr;
  int n_tip = 0, tip_cap = 0;
  double tip_m = 0.0;
  double* dinvK = nullptr;
  float* dinvK32 = nullptr;
  bool dinv_valid = false;
  int dinv_guard = -1;
  // reductions
  double* red_partials = nullptr;   // 4096 * 40 doubles
  unsigned int* red_counters = nullptr;  // 64
  double* red_out = nullptr;        // 64 doubles
  double* host = nullptr;           // pinned 64 doubles
  unsigned int* d_changed = nullptr;  // vox8_set_density: did any element factor change
  unsigned int* h_changed = nullptr;  // pinned mirror
  // PCG
  PcgWork<double> w64;
  PcgWork<float> w32;
  // LOBPCG
  LobpcgWork lw;
  // filter
  int* d_toff = nullptr;
  double* d_tw = nullptr;
  int ntap = 0;
  uint8_t* d_fmask = nullptr;
  double* d_den = nullptr;
  bool have_filter = false;
  // multigrid (built on first use; re-set up when the density changes)
  MultiGrid mg;
  std::vector<uint8_t> h_active, h_fixed;
  uint64_t density_version = 0, mg_version = ~0ull;
  vox8_ctx* mg_src = nullptr;       // vox8_share_multigrid: operators copied from this context
  // OC / sensitivities scratch
  double* d_oc = nullptr;
  double* d_tail = nullptr;         // TAIL_STATE doubles
  // vox8_simp_tail: the launches captured as HIP graphs (keyed by every value
  // the launches take by value; the per-call values go through d_tail_par)
  TailPar* d_tail_par = nullptr;
  TailPar* h_tail_par = nullptr;    // pinned staging of it
  int tail_graphs = 1;              // 0 direct launches, 1 graphs, 2 graphs or an error
  struct TailGraph {
    std::vector<unsigned char> key;
    hipGraphExec_t exec = nullptr;
    uint64_t last_use = 0;
  };
  std::vector<TailGraph> tail_cache;
  uint64_t tail_clock = 0;
  int lobpcg_guard = -1;            // guard columns of vox8_lobpcg (-1: environment or 0)
  double *d_kphi = nullptr, *d_mphi = nullptr, *d_ce = nullptr, *d_sums = nullptr,
         *d_wts = nullptr;
  // effective mass: M [X, r] and r (em_cols columns of n_dofs)
  double* d_em = nullptr;
  int em_cols = 0;
};

namespace {

template <typename T>
int dalloc(vox8_ctx* c, T** p, size_t count) {
  HipSerial lock(hip_serial());
  void* q = nullptr;
  VOX8_HIP(hipMalloc(&q, count * sizeof(T) + 256));
  c->dev_allocs.push_back(q);
  *p = static_cast<T*>(q);
  return OK;
}
template <typename T>
int halloc(vox8_ctx* c, T** p, size_t count) {
  HipSerial lock(hip_serial());
  void* q = nullptr;
  VOX8_HIP(hipHostMalloc(&q, count * sizeof(T) + 64, hipHostMallocDefault));
  c->host_allocs.push_back(q);
  *p = static_cast<T*>(q);
  return OK;
}
void dfree(vox8_ctx* c, void* p) {
  if (!p) return;
  HipSerial lock(hip_serial());
  for (auto it = c->dev_allocs.begin(); it != c->dev_allocs.end(); ++it)
    if (*it == p) { (void)hipFree(p); c->dev_allocs.erase(it); return; }
}
void hfree(vox8_ctx* c, void* p) {
  if (!p) return;
  HipSerial lock(hip_serial());
  for (auto it = c->host_allocs.begin(); it != c->host_allocs.end(); ++it)
    if (*it == p) { (void)hipHostFree(p); c->host_allocs.erase(it); return; }
}

int require(vox8_ctx* c, bool density) {
  VOX8_CHECK(c != nullptr, "null context");
  if (!c->have_mats) { set_error("element matrices not set (vox8_set_element_matrices)"); return ERR_STATE; }
  if (!c->have_topo) { set_error("topology not set (vox8_set_topology)"); return ERR_STATE; }
  if (density && !c->have_density) { set_error("density not set (vox8_set_density)"); return ERR_STATE; }
  VOX8_HIP(hipSetDevice(c->device));
  return OK;
}

int ensure_dinv(vox8_ctx* c, int guard) {
  if (c->dinv_valid && c->dinv_guard == guard) return OK;
  VOX8_TRY(launch_jacobi(c->g, c->fK, c->dKe, c->d_fixed, guard, nullptr, 0, 0.0, nullptr, c->dinvK,
                         c->st));
  VOX8_TRY(convert_to_float(c->dinvK, c->dinvK32, c->g.n_dofs, c->st));
  c->dinv_valid = true;
  c->dinv_guard = guard;
  return OK;
}

template <typename T>
int apply_op(vox8_ctx* c, int op, const T* x, int64_t ldx, T* yK, T* yM, int64_t ldy, int nrhs,
             DotFuse fuse, const int* skip = nullptr) {
  const T* fK = nullptr;
  const T* fM = nullptr;
  const T* Ke = nullptr;
  const T* Me = nullptr;
  if constexpr (sizeof(T) == 8) {
    fK = (const T*)c->fK; fM = (const T*)c->fM; Ke = (const T*)c->dKe; Me = (const T*)c->dMe;
  } else {
    fK = (const T*)c->fK32; fM = (const T*)c->fM32; Ke = (const T*)c->dKe32; Me = (const T*)c->dMe32;
  }
  if (c->pattern != PAT_DENSE && c->stencil_kind == 0) {
    Epilogue<T> ep;
    ep.skip = skip;
    if constexpr (sizeof(T) == 8)
      VOX8_TRY(launch_apply_fast<double>(c->g, c->pattern, op, c->c64, fK, fM, c->d_fixed, x, ldx,
                                         yK, yM, ldy, nrhs, c->st, fuse, ep));
    else
      VOX8_TRY(launch_apply_fast<float>(c->g, c->pattern, op, c->c32, fK, fM, c->d_fixed, x, ldx,
                                        yK, yM, ldy, nrhs, c->st, fuse, ep));
  } else {
    VOX8_TRY(launch_apply_dense<T>(c->g, op, fK, fM, Ke, Me, c->d_fixed, x, ldx, yK, yM, ldy, nrhs,
                                   c->st, fuse));
  }
  if ((op & OP_M) && c->n_tip > 0)
    VOX8_TRY(launch_tip_mass<T>(c->d_tip, c->n_tip, c->tip_m, c->d_fixed, x, ldx, yM, ldy, nrhs, c->st));
  return OK;
}

template <typename T>
int ensure_pcg_work(vox8_ctx* c, PcgWork<T>& w, int ncols) {
  if (w.ncols >= ncols) return OK;
  dfree(c, w.r); dfree(c, w.p); dfree(c, w.q); dfree(c, w.x); dfree(c, w.rhs64);
  dfree(c, w.scal); dfree(c, w.partials); dfree(c, w.counters); dfree(c, w.skip);
  hfree(c, w.host_scal); hfree(c, w.host_scale);
  const int64_t n = c->g.n_dofs;
  VOX8_TRY(dalloc(c, &w.r, (size_t)ncols * n));
  VOX8_TRY(dalloc(c, &w.p, (size_t)ncols * n));
  VOX8_TRY(dalloc(c, &w.q, (size_t)ncols * n));
  VOX8_TRY(dalloc(c, &w.x, (size_t)ncols * n));
  VOX8_TRY(dalloc(c, &w.rhs64, (size_t)ncols * n + ncols));
  VOX8_TRY(dalloc(c, &w.scal, ncols));
  VOX8_TRY(dalloc(c, &w.skip, ncols));
  const int64_t apply_blocks = (c->g.n_nodes + 255) / 256;
  const int64_t per_col = apply_blocks > 3 * 1024 ? apply_blocks : 3 * 1024;
  VOX8_TRY(dalloc(c, &w.partials, (size_t)(ncols * per_col)));
  VOX8_TRY(dalloc(c, &w.counters, ncols));
  VOX8_HIP(hipMemsetAsync(w.counters, 0, sizeof(unsigned int) * ncols, c->st));
  VOX8_TRY(halloc(c, &w.host_scal, ncols));
  VOX8_TRY(halloc(c, &w.host_scale, ncols));
  w.ncols = ncols;
  w.n = n;
  return OK;
}

int ensure_lobpcg_work(vox8_ctx* c, int k) {
  LobpcgWork& w = c->lw;
  if (w.kmax >= k) return OK;
  dfree(c, w.S); dfree(c, w.AS); dfree(c, w.BS); dfree(c, w.G); dfree(c, w.gram_partials);
  dfree(c, w.res); dfree(c, w.res_partials); dfree(c, w.counters); hfree(c, w.host);
  dfree(c, w.guard); dfree(c, w.conv);
  w.guard_cols = 0;
  const int64_t n = c->g.n_dofs;
  VOX8_TRY(dalloc(c, &w.S, (size_t)3 * k * n));
  VOX8_TRY(dalloc(c, &w.AS, (size_t)3 * k * n));
  VOX8_TRY(dalloc(c, &w.BS, (size_t)3 * k * n));
  VOX8_TRY(dalloc(c, &w.G, (size_t)(2 * 36 * 36 + 36 * 12 + 12 + 8)));
  int64_t chunks = (n + 511) / 512;          // Gram blocks: >= 8 row chunks of 64 each
  if (chunks > 512) chunks = 512;
  if (chunks < 1) chunks = 1;
  w.gram_chunks = (int)chunks;
  VOX8_TRY(dalloc(c, &w.gram_partials, (size_t)chunks * 2 * 36 * 36));
  VOX8_TRY(dalloc(c, &w.res, 16));
  // residual partials: the residual kernel's 1024 blocks x 12, or the product
  // kernel's workgroups x columns when the residual is formed there
  w.res_partials_cap = std::max<int64_t>((int64_t)1024 * 12, (int64_t)k * stencil_max_blocks(c->g));
  VOX8_TRY(dalloc(c, &w.res_partials, (size_t)w.res_partials_cap));
  VOX8_TRY(dalloc(c, &w.counters, 4));
  VOX8_HIP(hipMemsetAsync(w.counters, 0, sizeof(unsigned int) * 4, c->st));
  VOX8_TRY(halloc(c, &w.host, 32));
  VOX8_TRY(dalloc(c, &w.guard, (size_t)LOBPCG_GUARD_MAX * n));
  VOX8_TRY(dalloc(c, &w.conv, 16));
  w.kmax = k;
  w.n = n;
  return OK;
}

int mg_dev_alloc(void* ctx, void** p, size_t bytes) {
  uint8_t* q = nullptr;
  VOX8_TRY(dalloc((vox8_ctx*)ctx, &q, bytes));
  *p = q;
  return OK;
}
int mg_pinned_alloc(void* ctx, void** p, size_t bytes) {
  uint8_t* q = nullptr;
  VOX8_TRY(halloc((vox8_ctx*)ctx, &q, bytes));
  *p = q;
  return OK;
}

// Multigrid preconditioner for the current density, with work space for ncols.
int ensure_mg(vox8_ctx* c, int ncols) {
  if (c->pattern == PAT_DENSE || c->stencil_kind != 0) {
    set_error("the multigrid preconditioner needs the structured stencil (box-element Ke/Me)");
    return ERR_STATE;
  }
  MGAlloc al{c, mg_dev_alloc, mg_pinned_alloc};
  if (!c->mg.built) {
    VOX8_TRY(mg_build(c->mg, c->g, c->h_active.data(), c->h_fixed.data(), c->Ke, c->pattern, c->c32,
                      c->fK32, c->d_active, c->d_fixed, al, c->st));
    c->mg_version = ~0ull;
  }
  VOX8_TRY(mg_reserve(c->mg, ncols, al));
  if (c->mg_version != c->density_version) {
    if (c->mg_src) {
      const vox8_ctx* s = c->mg_src;
      VOX8_CHECK(s->mg.built && s->mg.updated && s->mg_version == s->density_version,
                 "shared multigrid: the source context's hierarchy is not set up for its current "
                 "density (vox8_prepare_multigrid on it first)");
      VOX8_TRY(mg_copy_operators(c->mg, s->mg, c->st));
    } else {
      VOX8_TRY(mg_update(c->mg, c->fK, c->dKe, c->st));
    }
    c->mg_version = c->density_version;
  }
  return OK;
}

// An environment switch set to "0" (tuning / A-B runs).
bool env_off(const char* name) {
  const char* e = getenv(name);
  return e && e[0] == '0';
}

int sync_read(vox8_ctx* c, const double* d, double* h, int count) {
  VOX8_HIP(hipMemcpyAsync(c->host, d, sizeof(double) * count, hipMemcpyDeviceToHost, c->st));
  VOX8_HIP(hipStreamSynchronize(c->st));
  for (int i = 0; i < count; ++i) h[i] = c->host[i];
  return OK;
}

}  // namespace

#define API_BEGIN try {
#define API_END                                                   \
  }                                                               \
  catch (const std::bad_alloc&) {                                 \
    set_error("host allocation failed");                          \
    return ERR_HIP;                                               \
  }                                                               \
  catch (const std::exception& e) {                               \
    set_error(std::string("exception: ") + e.what());             \
    return ERR_STATE;                                             \
  }

extern "C" {

const char* vox8_last_error(void) { return g_err.c_str(); }
int vox8_version(void) { return 100; }

int vox8_device_count(int* h_count) {
  VOX8_CHECK(h_count != nullptr, "null output");
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  *h_count = (e == hipSuccess) ? n : 0;
  return OK;
}

int vox8_create(vox8_ctx** out, int nx, int ny, int nz, void* hip_stream) {
  API_BEGIN
  HipSerial lock(hip_serial());
  VOX8_CHECK(out != nullptr, "null output");
  VOX8_CHECK(nx > 0 && ny > 0 && nz > 0, "grid dimensions must be positive");
  VOX8_CHECK((int64_t)(nx + 1) * (ny + 1) * (nz + 1) < (int64_t)1 << 31, "grid too large");
  vox8_ctx* c = new vox8_ctx();
  c->g = make_grid(nx, ny, nz);
  if (const char* e = getenv("VOX8_MG_WMIN")) c->mg.w_min_dofs = atoll(e);   // tuning overrides
  if (const char* e = getenv("VOX8_MG_COARSE")) c->mg.coarse_dofs = std::max(24, std::min(4096, atoi(e)));
  if (const char* e = getenv("VOX8_MG_GALERKIN")) c->mg.galerkin = e[0] != '0';
  if (const char* e = getenv("VOX8_MG_NU")) c->mg.nu = std::max(1, std::min(16, atoi(e)));
  if (const char* e = getenv("VOX8_MG_NUC")) c->mg.nu_coarse = std::max(0, std::min(16, atoi(e)));
  if (const char* e = getenv("VOX8_MG_NUC_L")) c->mg.nu_coarse_level = std::max(1, atoi(e));
  if (const char* e = getenv("VOX8_MG_ALPHA")) c->mg.alpha = std::max(1.5, atof(e));
  if (const char* e = getenv("VOX8_MG_MARGIN")) c->mg.lam_margin = std::max(1.0, atof(e));
  if (const char* e = getenv("VOX8_MG_VTOP")) c->mg.v_top = std::max(0, atoi(e));
  if (const char* e = getenv("VOX8_MG_GAMMA_LEVELS")) {       // "g0,g1,...": corrections per level
    int l = 0;
    for (const char* q = e; *q && l < 8; ++l) {
      c->mg.gamma_level[l] = std::max(0, std::min(2, atoi(q)));
      while (*q && *q != ',') ++q;
      if (*q == ',') ++q;
    }
  }
  if (const char* e = getenv("VOX8_MG_NU_LEVELS")) {          // "n0,n1,...": steps per level
    int l = 0;
    for (const char* q = e; *q && l < 8; ++l) {
      c->mg.nu_level[l] = std::max(0, std::min(16, atoi(q)));
      while (*q && *q != ',') ++q;
      if (*q == ',') ++q;
    }
  }
  if (const char* e = getenv("VOX8_MG_LANE_NODES")) c->mg.lane_nodes = std::max(1LL, atoll(e));
  c->st = (hipStream_t)hip_stream;
  int s = 0;
  auto fail = [&](int code) { vox8_destroy(c); *out = nullptr; return code; };
  if (hipGetDevice(&c->device) != hipSuccess) { set_error("no HIP device"); return fail(ERR_HIP); }
  const int64_t ne = c->g.n_elem, nn = c->g.n_nodes, nd = c->g.n_dofs;
  if ((s = dalloc(c, &c->dKe, 576)) || (s = dalloc(c, &c->dMe, 576)) ||
      (s = dalloc(c, &c->dKe32, 576)) || (s = dalloc(c, &c->dMe32, 576)) ||
      (s = dalloc(c, &c->d_active, ne)) || (s = dalloc(c, &c->d_fixed, nn)) ||
      (s = dalloc(c, &c->fK, ne)) || (s = dalloc(c, &c->fM, ne)) ||
      (s = dalloc(c, &c->fK32, ne)) || (s = dalloc(c, &c->fM32, ne)) ||
      (s = dalloc(c, &c->dinvK, nd)) || (s = dalloc(c, &c->dinvK32, nd)) ||
      (s = dalloc(c, &c->red_partials, 4096 * 40)) || (s = dalloc(c, &c->red_counters, 64)) ||
      (s = dalloc(c, &c->red_out, 64)) || (s = halloc(c, &c->host, 64)) ||
      (s = dalloc(c, &c->d_oc, 8)) || (s = dalloc(c, &c->d_tail, 32)) ||
      (s = dalloc(c, &c->d_kphi, ne)) ||
      (s = dalloc(c, &c->d_mphi, ne)) || (s = dalloc(c, &c->d_ce, ne)) ||
      (s = dalloc(c, &c->d_sums, 32)) || (s = dalloc(c, &c->d_wts, 32)) ||
      (s = dalloc(c, &c->d_den, ne)) || (s = dalloc(c, &c->d_fmask, ne)) ||
      (s = dalloc(c, &c->d_changed, 1)) || (s = halloc(c, &c->h_changed, 1)) ||
      (s = dalloc(c, &c->d_tail_par, 1)) || (s = halloc(c, &c->h_tail_par, 1)))
    return fail(s);
  // stream-ordered (a legacy-stream memset conflicts with graph captures in
  // progress on other threads)
  if (hipMemsetAsync(c->red_counters, 0, 64 * sizeof(unsigned int), c->st) != hipSuccess ||
      hipStreamSynchronize(c->st) != hipSuccess) {
    set_error("hipMemset failed");
    return fail(ERR_HIP);
  }
  *out = c;
  return OK;
  API_END
}

int vox8_destroy(vox8_ctx* c) {
  if (!c) return OK;
  HipSerial lock(hip_serial());
  (void)hipSetDevice(c->device);
  if (c->st) (void)hipStreamSynchronize(c->st); else (void)hipDeviceSynchronize();
  for (auto& tg : c->tail_cache) (void)hipGraphExecDestroy(tg.exec);
  c->tail_cache.clear();
  for (void* p : c->dev_allocs) (void)hipFree(p);
  for (void* p : c->host_allocs) (void)hipHostFree(p);
  delete c;
  return OK;
}

int vox8_set_stream(vox8_ctx* c, void* hip_stream) {
  VOX8_CHECK(c != nullptr, "null context");
  c->st = (hipStream_t)hip_stream;
  return OK;
}

int vox8_set_element_matrices(vox8_ctx* c, const double* h_Ke, const double* h_Me) {
  API_BEGIN
  VOX8_CHECK(c && h_Ke && h_Me, "null argument");
  VOX8_HIP(hipSetDevice(c->device));
  std::memcpy(c->Ke, h_Ke, sizeof(c->Ke));
  std::memcpy(c->Me, h_Me, sizeof(c->Me));
  float k32[576], m32[576];
  for (int i = 0; i < 576; ++i) { k32[i] = (float)h_Ke[i]; m32[i] = (float)h_Me[i]; }
  VOX8_HIP(hipMemcpyAsync(c->dKe, c->Ke, sizeof(c->Ke), hipMemcpyHostToDevice, c->st));
  VOX8_HIP(hipMemcpyAsync(c->dMe, c->Me, sizeof(c->Me), hipMemcpyHostToDevice, c->st));
  VOX8_HIP(hipMemcpyAsync(c->dKe32, k32, sizeof(k32), hipMemcpyHostToDevice, c->st));
  VOX8_HIP(hipMemcpyAsync(c->dMe32, m32, sizeof(m32), hipMemcpyHostToDevice, c->st));
  VOX8_HIP(hipStreamSynchronize(c->st));
  c->pattern = fast_coefficients(c->Ke, c->Me, &c->c64, &c->c32);
  c->mg.built = false;             // level operators follow Ke
  c->have_mats = true;
  c->dinv_valid = false;
  return OK;
  API_END
}

int vox8_set_stencil(vox8_ctx* c, int kind, int* h_pattern) {
  VOX8_CHECK(c != nullptr, "null context");
  VOX8_CHECK(kind == 0 || kind == 1, "kind must be 0 (auto) or 1 (dense)");
  c->stencil_kind = kind;
  c->mg.built = false;
  if (h_pattern) *h_pattern = (kind == 1) ? PAT_DENSE : c->pattern;
  return OK;
}

int vox8_precond_apply(vox8_ctx* c, const double* d_r, double* d_z, int nrhs) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(d_r && d_z && nrhs >= 1 && nrhs <= 64, "bad argument");
  VOX8_TRY(ensure_mg(c, nrhs));
  return mg_apply(c->mg, d_r, d_z, nrhs, c->st);
  API_END
}

int vox8_cheb_step(vox8_ctx* c, const float* d_d, float* d_r, float* d_xs, float* d_dnew, double c1,
                   double c2, int nrhs) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(d_d && d_r && d_xs && d_dnew && nrhs >= 1 && nrhs <= 64, "bad argument");
  VOX8_CHECK(d_dnew != d_d, "d_dnew must not alias d_d");
  if (c->pattern == PAT_DENSE || c->stencil_kind != 0) {
    set_error("vox8_cheb_step needs the structured stencil (box elements)");
    return ERR_STATE;
  }
  VOX8_TRY(ensure_dinv(c, 0));
  Epilogue<float> ep;
  ep.mode = EPI_CHEB;
  ep.r = d_r;
  ep.xs = d_xs;
  ep.dnew = d_dnew;
  ep.dinv = c->dinvK32;
  ep.c1 = (float)c1;
  ep.c2 = (float)c2;
  ep.ld = c->g.n_dofs;
  return launch_apply_fast<float>(c->g, c->pattern, OP_K, c->c32, c->fK32, nullptr, c->d_fixed, d_d,
                                  c->g.n_dofs, nullptr, nullptr, c->g.n_dofs, nrhs, c->st, DotFuse(),
                                  ep);
  API_END
}

int vox8_cheb_xstep(vox8_ctx* c, const float* d_xk, const float* d_xprev, const float* d_b,
                    float* d_xnew, double c1, double c2, int nrhs) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(d_xk && d_b && d_xnew && nrhs >= 1 && nrhs <= 64, "bad argument");
  VOX8_CHECK(d_xnew != d_xk && d_xnew != d_xprev && d_xnew != d_b, "d_xnew must not alias an input");
  if (c->pattern == PAT_DENSE || c->stencil_kind != 0) {
    set_error("vox8_cheb_xstep needs the structured stencil (box elements)");
    return ERR_STATE;
  }
  VOX8_TRY(ensure_dinv(c, 0));
  Epilogue<float> ep;
  ep.mode = EPI_X;
  ep.b = d_b;
  ep.xs = d_xnew;
  ep.xprev = d_xprev;
  ep.prev = d_xprev ? 1 : 0;
  ep.dinv = c->dinvK32;
  ep.c1 = (float)c1;
  ep.c2 = (float)c2;
  ep.ld = c->g.n_dofs;
  return launch_apply_fast<float>(c->g, c->pattern, OP_K, c->c32, c->fK32, nullptr, c->d_fixed, d_xk,
                                  c->g.n_dofs, nullptr, nullptr, c->g.n_dofs, nrhs, c->st, DotFuse(),
                                  ep);
  API_END
}

int vox8_mg_info(vox8_ctx* c, int* h_nlevels, int* h_dims, double* h_lam_max, int max_levels) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(h_nlevels != nullptr, "null output");
  VOX8_TRY(ensure_mg(c, 1));
  const int nl = (int)c->mg.lv.size();
  *h_nlevels = nl;
  for (int l = 0; l < nl && l < max_levels; ++l) {
    if (h_dims) { h_dims[3 * l] = c->mg.lv[l].g.nx; h_dims[3 * l + 1] = c->mg.lv[l].g.ny; h_dims[3 * l + 2] = c->mg.lv[l].g.nz; }
    if (h_lam_max) h_lam_max[l] = c->mg.lv[l].lam_max;
  }
  return OK;
  API_END
}

int vox8_set_mg_options(vox8_ctx* c, int nu, double alpha, int coarse_dofs, int cycle,
                        int64_t w_min_dofs) {
  VOX8_CHECK(c != nullptr, "null context");
  VOX8_CHECK(nu >= 1 && nu <= 16 && alpha > 1.0 && coarse_dofs >= 24 && coarse_dofs <= 4096 &&
                 cycle >= 1 && cycle <= 2 && w_min_dofs >= 0,
             "mg options out of range (nu 1..16, alpha > 1, coarse_dofs 24..4096, cycle 1 or 2, "
             "w_min_dofs >= 0)");
  c->mg.nu = nu;
  c->mg.gamma = cycle;
  c->mg.w_min_dofs = w_min_dofs;
  c->mg.alpha = alpha;
  if (c->mg.coarse_dofs != coarse_dofs) c->mg.built = false;
  c->mg.coarse_dofs = coarse_dofs;
  c->mg_version = ~0ull;
  ++c->mg.epoch;
  return OK;
}

int vox8_set_mg_coarse_smoothing(vox8_ctx* c, int nu_coarse, int from_level) {
  VOX8_CHECK(c != nullptr, "null context");
  VOX8_CHECK(nu_coarse >= 0 && nu_coarse <= 16 && from_level >= 1,
             "coarse smoothing out of range (nu_coarse 0..16, from_level >= 1)");
  c->mg.nu_coarse = nu_coarse;
  c->mg.nu_coarse_level = from_level;
  ++c->mg.epoch;                       // captured cycles hold the old step counts
  return OK;
}

int vox8_share_multigrid(vox8_ctx* c, vox8_ctx* src) {
  VOX8_CHECK(c != nullptr && c != src, "share_multigrid: null context or a context sharing with itself");
  VOX8_CHECK(src == nullptr || src->mg_src == nullptr, "share_multigrid: the source shares itself");
  if (src) {
    VOX8_CHECK(src->g.nx == c->g.nx && src->g.ny == c->g.ny && src->g.nz == c->g.nz &&
                   src->device == c->device,
               "share_multigrid: contexts of different grids or devices");
  }
  c->mg_src = src;
  if (c->mg.operators_only != (src != nullptr)) {
    c->mg.operators_only = src != nullptr;   // a sharing hierarchy holds no element matrices
    c->mg.built = false;
  }
  c->mg_version = ~0ull;               // the next solve takes the operators (or updates itself)
  return OK;
}

int vox8_prepare_multigrid(vox8_ctx* c, int ncols) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(ncols >= 1 && ncols <= 16, "prepare_multigrid: 1..16 columns");
  VOX8_TRY(ensure_mg(c, ncols));
  VOX8_HIP(hipStreamSynchronize(c->st));
  return OK;
  API_END
}

int vox8_set_mg_galerkin(vox8_ctx* c, int enable) {
  VOX8_CHECK(c != nullptr, "null context");
  if (c->mg.galerkin != (enable != 0)) {
    c->mg.galerkin = enable != 0;
    c->mg.built = false;           // the level operators change form
    c->mg_version = ~0ull;
  }
  return OK;
}

int vox8_set_mg_graphs(vox8_ctx* c, int enable) {
  VOX8_CHECK(c != nullptr, "null context");
  c->mg.use_graphs = enable != 0;
  if (!c->mg.use_graphs) mg_drop_graphs(c->mg);
  return OK;
}

int vox8_set_topology(vox8_ctx* c, const uint8_t* h_active, const uint8_t* h_fixed) {
  API_BEGIN
  VOX8_CHECK(c && h_active && h_fixed, "null argument");
  VOX8_HIP(hipSetDevice(c->device));
  VOX8_HIP(hipMemcpyAsync(c->d_active, h_active, c->g.n_elem, hipMemcpyHostToDevice, c->st));
  VOX8_HIP(hipMemcpyAsync(c->d_fixed, h_fixed, c->g.n_nodes, hipMemcpyHostToDevice, c->st));
  VOX8_HIP(hipStreamSynchronize(c->st));
  c->h_active.assign(h_active, h_active + c->g.n_elem);
  c->h_fixed.assign(h_fixed, h_fixed + c->g.n_nodes);
  c->mg.built = false;             // level hierarchy follows the topology
  c->lw.guard_cols = 0;            // guard columns of an earlier topology are not reused
  c->have_topo = true;
  c->have_density = false;
  c->dinv_valid = false;
  return OK;
  API_END
}

int vox8_set_tip_mass(vox8_ctx* c, const int64_t* h_nodes, int n_nodes, double mass_kg) {
  API_BEGIN
  VOX8_CHECK(c != nullptr, "null context");
  VOX8_CHECK(n_nodes >= 0, "negative node count");
  VOX8_HIP(hipSetDevice(c->device));
  if (n_nodes == 0 || mass_kg <= 0.0) { c->n_tip = 0; c->tip_m = 0.0; return OK; }
  VOX8_CHECK(h_nodes != nullptr, "null node list");
  for (int i = 0; i < n_nodes; ++i)
    VOX8_CHECK(h_nodes[i] >= 0 && h_nodes[i] < c->g.n_nodes, "tip node out of range");
  if (n_nodes > c->tip_cap) {
    dfree(c, c->d_tip);
    VOX8_TRY(dalloc(c, &c->d_tip, n_nodes));
    c->tip_cap = n_nodes;
  }
  VOX8_HIP(hipMemcpyAsync(c->d_tip, h_nodes, sizeof(int64_t) * n_nodes, hipMemcpyHostToDevice, c->st));
  VOX8_HIP(hipStreamSynchronize(c->st));
  c->n_tip = n_nodes;
  c->tip_m = mass_kg / n_nodes;
  return OK;
  API_END
}

int vox8_set_density(vox8_ctx* c, const double* d_rho, double p, double e_min, double mthr,
                     double mq) {
  API_BEGIN
  VOX8_TRY(require(c, false));
  VOX8_CHECK(d_rho != nullptr, "null density");
  // The factors are compared with the ones they replace: setting the same
  // density again (a caller that re-sends its density with every solve)
  // keeps the density version, so nothing derived from it (the multigrid
  // hierarchy, the Jacobi diagonal) is set up again.
  VOX8_HIP(hipMemsetAsync(c->d_changed, 0, sizeof(unsigned int), c->st));
  VOX8_TRY(launch_element_factors(c->g, d_rho, c->d_active, p, e_min, mthr, mq, c->fK, c->fM,
                                  c->fK32, c->fM32, c->st, c->d_changed));
  VOX8_HIP(hipMemcpyAsync(c->h_changed, c->d_changed, sizeof(unsigned int), hipMemcpyDeviceToHost,
                          c->st));
  VOX8_HIP(hipStreamSynchronize(c->st));
  if (!c->have_density || *c->h_changed) {
    c->dinv_valid = false;
    ++c->density_version;
  }
  c->have_density = true;
  return OK;
  API_END
}

int vox8_apply(vox8_ctx* c, int op, int prec, const void* d_X, void* d_YK, void* d_YM, int nrhs,
               int64_t ld) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(op >= 1 && op <= 3, "op must be K(1), M(2) or KM(3)");
  VOX8_CHECK(nrhs >= 1 && nrhs <= 65535, "nrhs out of range");
  VOX8_CHECK(ld >= c->g.n_dofs, "ld < n_dofs");
  VOX8_CHECK(d_X && (!(op & OP_K) || d_YK) && (!(op & OP_M) || d_YM), "null buffer");
  if (prec == VOX8_PREC_FP32)
    return apply_op<float>(c, op, (const float*)d_X, ld, (float*)d_YK, (float*)d_YM, ld, nrhs,
                           DotFuse());
  VOX8_CHECK(prec == VOX8_PREC_FP64, "prec must be FP64 (0) or FP32 (2) for vox8_apply");
  return apply_op<double>(c, op, (const double*)d_X, ld, (double*)d_YK, (double*)d_YM, ld, nrhs,
                          DotFuse());
  API_END
}

int vox8_diagonal(vox8_ctx* c, int op, double* d_diag) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(op == 1 || op == 2, "op must be K(1) or M(2)");
  VOX8_CHECK(d_diag != nullptr, "null output");
  if (op == 1)
    return launch_jacobi(c->g, c->fK, c->dKe, c->d_fixed, 0, nullptr, 0, 0.0, d_diag, nullptr, c->st);
  return launch_jacobi(c->g, c->fM, c->dMe, c->d_fixed, 0, c->d_tip, c->n_tip, c->tip_m, d_diag,
                       nullptr, c->st);
  API_END
}

int vox8_pcg(vox8_ctx* c, const double* d_F, double* d_U, int nrhs, int64_t ld, int x0_given,
             double rtol, int maxiter, int prec, int* h_iters, int* h_info, double* h_relres) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(d_F && d_U, "null buffer");
  VOX8_CHECK(nrhs >= 1 && nrhs <= 64, "nrhs must be 1..64");
  VOX8_CHECK(ld == c->g.n_dofs, "pcg: ld must equal n_dofs");
  VOX8_CHECK(rtol > 0.0 && maxiter >= 1, "rtol must be > 0 and maxiter >= 1");
  VOX8_TRY(ensure_dinv(c, 0));
  VOX8_TRY(ensure_pcg_work(c, c->w64, nrhs));
  std::vector<PcgResult> res(nrhs);
  PcgOps<double> A64;
  A64.g = &c->g;
  A64.fixed = c->d_fixed;
  A64.dinv = c->dinvK;
  A64.apply = [c](const double* x, int64_t ldx, double* y, int64_t ldy, int nr, hipStream_t,
                  DotFuse f) { return apply_op<double>(c, OP_K, x, ldx, y, nullptr, ldy, nr, f); };
  if (prec == VOX8_PREC_MIXED) {
    VOX8_TRY(ensure_mg(c, nrhs));
    Precond<double> M = [c](const double* r, double* z, int nr, hipStream_t st, const int* skip) {
      return mg_apply(c->mg, r, z, nr, st, skip);
    };
    // the residual update writes the cycle's FP32 right-hand side too
    float* m32 = env_off("VOX8_RESIDUAL_FP32") ? nullptr : c->mg.lv[0].b;
    VOX8_TRY(pcg_fp64(A64, d_F, ld, d_U, x0_given != 0, nrhs, rtol, maxiter, c->w64, c->st,
                      res.data(), &M, m32));
  } else {
    VOX8_CHECK(prec == VOX8_PREC_FP64, "prec must be FP64 (0) or MIXED (1)");
    VOX8_TRY(pcg_fp64(A64, d_F, ld, d_U, x0_given != 0, nrhs, rtol, maxiter, c->w64, c->st,
                      res.data()));
  }
  for (int j = 0; j < nrhs; ++j) {
    if (h_iters) h_iters[j] = res[j].iters;
    if (h_info) h_info[j] = res[j].info;
    if (h_relres) h_relres[j] = res[j].relres;
  }
  return OK;
  API_END
}

int vox8_lobpcg(vox8_ctx* c, int k, double tol, int maxiter, int prec, const double* d_X0,
                double* d_X, double* h_lam, int* h_iters, int* h_info, double* h_max_res,
                uint64_t seed) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(d_X && h_lam, "null buffer");
  VOX8_CHECK(k >= 1 && k <= 12, "k must be 1..12");
  VOX8_CHECK(tol > 0.0 && maxiter >= 0, "tol must be > 0");
  VOX8_CHECK(prec == VOX8_PREC_FP64 || prec == VOX8_PREC_MIXED, "prec must be FP64 or MIXED");
  VOX8_TRY(ensure_dinv(c, 1));
  // optional guard columns (iterated, not returned; tuning switch
  // VOX8_LOBPCG_GUARD): fewer iterations when lambda_k+1 is close to lambda_k,
  // at a higher cost per iteration; a warm start reuses the last solve's
  int guard = c->lobpcg_guard;
  if (guard < 0) {
    guard = 0;
    if (const char* e = getenv("VOX8_LOBPCG_GUARD")) guard = atoi(e);
  }
  guard = std::max(0, std::min(std::min(guard, LOBPCG_GUARD_MAX), 12 - k));
  VOX8_TRY(ensure_lobpcg_work(c, k + guard));
  if (!d_X0 || c->lw.guard_k != k) c->lw.guard_cols = 0;   // cold start: random guard columns
  c->lw.guard_k = k;
  PcgOps<double> A;
  A.g = &c->g;
  A.fixed = c->d_fixed;
  A.dinv = c->dinvK;
  auto km = [c](const double* x, double* yK, double* yM, int nr, hipStream_t, const int* skip,
                const ResidualFuse* rf, bool* fused) {
    // the residual goes into the structured stencil's epilogue unless a tip
    // mass is added to M x after that kernel, or the element kernel is dense
    if (rf && fused && c->pattern != PAT_DENSE && c->stencil_kind == 0 && c->n_tip == 0 &&
        !env_off("VOX8_FUSED_RESIDUAL")) {
      Epilogue<double> ep;
      ep.mode = EPI_LOBRES;
      ep.lam = rf->lam;
      ep.conv = rf->conv;
      ep.dinv = rf->dinv;
      ep.w = rf->w32 ? nullptr : rf->w;
      ep.w32 = rf->w32;
      ep.ld = c->g.n_dofs;
      ep.partials = rf->partials;
      ep.partials_cap = rf->partials_cap;
      ep.res = rf->res;
      *fused = true;
      return launch_apply_fast<double>(c->g, c->pattern, OP_KM, c->c64, c->fK, c->fM, c->d_fixed, x,
                                       c->g.n_dofs, yK, yM, c->g.n_dofs, nr, c->st, DotFuse(), ep);
    }
    if (fused) *fused = false;
    return apply_op<double>(c, OP_KM, x, c->g.n_dofs, yK, yM, c->g.n_dofs, nr, DotFuse(), skip);
  };
  LobpcgResult r{};
  if (prec == VOX8_PREC_MIXED) {
    VOX8_TRY(ensure_mg(c, k + guard));
    Precond<double> M = [c](const double* rr, double* z, int nr, hipStream_t st, const int* skip) {
      return mg_apply(c->mg, rr, z, nr, st, skip);
    };
    // the residual goes straight into the cycle's FP32 right-hand side
    float* m32 = env_off("VOX8_RESIDUAL_FP32") ? nullptr : c->mg.lv[0].b;
    VOX8_TRY(lobpcg_fp64(A, km, k, guard, tol, maxiter, d_X0, d_X, h_lam, c->lw, c->st, &r, seed, &M,
                         m32));
  } else {
    VOX8_TRY(lobpcg_fp64(A, km, k, guard, tol, maxiter, d_X0, d_X, h_lam, c->lw, c->st, &r, seed));
  }
  if (h_iters) *h_iters = r.iters;
  if (h_info) *h_info = r.info;
  if (h_max_res) *h_max_res = r.max_res;
  return OK;
  API_END
}

int vox8_set_lobpcg_guard(vox8_ctx* c, int guard) {
  VOX8_CHECK(c != nullptr, "null context");
  VOX8_CHECK(guard >= -1 && guard <= LOBPCG_GUARD_MAX, "guard must be -1 (environment) or 0..2");
  c->lobpcg_guard = guard;
  return OK;
}

int vox8_lobpcg_ritz(vox8_ctx* c, double* h_out, int n_max, int* n_out) {
  VOX8_CHECK(c != nullptr && n_out != nullptr && (h_out != nullptr || n_max == 0), "null argument");
  const int n = c->lw.ritz_n;
  for (int j = 0; j < n && j < n_max; ++j) h_out[j] = c->lw.ritz[j];
  *n_out = n;
  return OK;
}

int vox8_sensitivities(vox8_ctx* c, const double* d_U, int nrhs, int64_t ld, const double* h_w,
                       const double* d_phi, double lam, const double* d_rho, double p, double e_min,
                       double mthr, double mq, double* d_ce, double* d_dc, double* d_dlam,
                       double* h_comp, double* h_vmv) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(nrhs >= 0 && nrhs <= 16, "nrhs must be 0..16");
  VOX8_CHECK(nrhs == 0 || (d_U && h_w), "null U or weights");
  VOX8_CHECK(ld >= c->g.n_dofs || nrhs == 0, "ld < n_dofs");
  VOX8_CHECK((!d_dc && !d_dlam) || d_rho, "derivatives need d_rho");
  double w[16] = {0};
  for (int j = 0; j < nrhs; ++j) w[j] = h_w[j];
  VOX8_HIP(hipMemcpyAsync(c->d_wts, w, sizeof(w), hipMemcpyHostToDevice, c->st));
  double* ce = d_ce ? d_ce : c->d_ce;
  VOX8_TRY(launch_quadforms(c->g, c->fK, c->fM, c->d_active, c->dKe, c->dMe, d_U, ld, nrhs,
                            c->d_wts, d_phi, nrhs > 0 ? ce : nullptr, c->d_kphi, c->d_mphi,
                            c->red_partials, c->red_counters, c->d_sums, c->d_tip, c->n_tip,
                            c->tip_m, c->st, nullptr, c->pattern != PAT_DENSE ? &c->c64 : nullptr));
  if (d_dc || d_dlam)
    VOX8_TRY(launch_sens_finish(c->g.n_elem, c->d_active, d_rho, p, e_min, mthr, mq, ce,
                                nrhs > 0 ? d_dc : nullptr, c->d_kphi, c->d_mphi, lam,
                                c->d_sums + nrhs, d_phi ? d_dlam : nullptr, c->st));
  double h[17];
  VOX8_TRY(sync_read(c, c->d_sums, h, nrhs + 1));
  if (h_comp) for (int j = 0; j < nrhs; ++j) h_comp[j] = h[j];
  if (h_vmv) *h_vmv = d_phi ? h[nrhs] : 0.0;
  return OK;
  API_END
}

int vox8_filter_setup(vox8_ctx* c, double r_min_vox, const uint8_t* d_mask) {
  API_BEGIN
  VOX8_CHECK(c && d_mask, "null argument");
  VOX8_CHECK(r_min_vox > 0.0 && r_min_vox < 32.0, "filter radius must be in (0, 32) voxels");
  VOX8_HIP(hipSetDevice(c->device));
  const int r = (int)std::ceil(r_min_vox);
  std::vector<int> off;
  std::vector<double> w;
  for (int dz = -r; dz <= r; ++dz)
    for (int dy = -r; dy <= r; ++dy)
      for (int dx = -r; dx <= r; ++dx) {
        const double wt = r_min_vox - std::sqrt((double)(dx * dx + dy * dy + dz * dz));
        if (wt > 0.0) { off.push_back(dx); off.push_back(dy); off.push_back(dz); w.push_back(wt); }
      }
  dfree(c, c->d_toff);
  dfree(c, c->d_tw);
  c->d_toff = nullptr;
  c->d_tw = nullptr;
  c->ntap = (int)w.size();
  VOX8_TRY(dalloc(c, &c->d_toff, off.size()));
  VOX8_TRY(dalloc(c, &c->d_tw, w.size()));
  VOX8_HIP(hipMemcpyAsync(c->d_toff, off.data(), sizeof(int) * off.size(), hipMemcpyHostToDevice, c->st));
  VOX8_HIP(hipMemcpyAsync(c->d_tw, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice, c->st));
  VOX8_HIP(hipMemcpyAsync(c->d_fmask, d_mask, c->g.n_elem, hipMemcpyDeviceToDevice, c->st));
  VOX8_TRY(launch_filter(c->g.nx, c->g.ny, c->g.nz, c->ntap, c->d_toff, c->d_tw, c->d_fmask,
                         nullptr, nullptr, c->d_den, FILTER_DEN, c->st));
  VOX8_HIP(hipStreamSynchronize(c->st));
  c->have_filter = true;
  return OK;
  API_END
}

int vox8_stress(vox8_ctx* c, const double* d_u, const double* d_rho_grid, const double* h_B,
                const double* h_D, double* d_stress) {
  API_BEGIN
  VOX8_TRY(require(c, false));
  VOX8_CHECK(d_u && d_rho_grid && h_B && h_D && d_stress, "null argument");
  StressCoef sc;
  for (int i = 0; i < 6 * 24; ++i) sc.B[i] = h_B[i];
  for (int i = 0; i < 36; ++i) sc.D[i] = h_D[i];
  return launch_stress(c->g, sc, d_u, d_rho_grid, c->d_active, d_stress, c->st);
  API_END
}

int vox8_stress_invariants(void* hip_stream, int64_t n, const double* d_stress, double* d_vm,
                           double* d_pmax) {
  API_BEGIN
  VOX8_CHECK(n >= 0 && (n == 0 || d_stress), "bad argument");
  return launch_stress_invariants(n, d_stress, d_vm, d_pmax, (hipStream_t)hip_stream);
  API_END
}

int vox8_fatigue(void* hip_stream, int64_t n, const double* d_s1, const double* d_s2, double uts_pa,
                 const double* h_log_stress_mpa, const double* h_log_cycles, double endurance_mpa,
                 double overload_mpa, double fos_ref_mpa, double* d_mean, double* d_alt, double* d_eq,
                 double* d_life, double* d_damage, double* d_fos) {
  API_BEGIN
  VOX8_CHECK(n >= 0 && (n == 0 || (d_s1 && d_s2)) && h_log_stress_mpa && h_log_cycles && uts_pa > 0.0,
             "bad argument");
  for (int i = 1; i < 7; ++i)
    VOX8_CHECK(h_log_stress_mpa[i] > h_log_stress_mpa[i - 1], "S-N stress table must increase");
  return launch_fatigue(n, d_s1, d_s2, uts_pa, h_log_stress_mpa, h_log_cycles, endurance_mpa,
                        overload_mpa, fos_ref_mpa, d_mean, d_alt, d_eq, d_life, d_damage, d_fos,
                        (hipStream_t)hip_stream);
  API_END
}

int vox8_filter(vox8_ctx* c, const double* d_in, double* d_out) {
  API_BEGIN
  VOX8_CHECK(c && d_in && d_out, "null argument");
  VOX8_CHECK(d_in != d_out, "filter cannot run in place");
  if (!c->have_filter) { set_error("filter not set up (vox8_filter_setup)"); return ERR_STATE; }
  return launch_filter(c->g.nx, c->g.ny, c->g.nz, c->ntap, c->d_toff, c->d_tw, c->d_fmask, c->d_den,
                       d_in, d_out, FILTER_FWD, c->st);
  API_END
}

int vox8_filter_adjoint(vox8_ctx* c, const double* d_in, double* d_out) {
  API_BEGIN
  VOX8_CHECK(c && d_in && d_out, "null argument");
  VOX8_CHECK(d_in != d_out, "filter cannot run in place");
  if (!c->have_filter) { set_error("filter not set up (vox8_filter_setup)"); return ERR_STATE; }
  return launch_filter(c->g.nx, c->g.ny, c->g.nz, c->ntap, c->d_toff, c->d_tw, c->d_fmask, c->d_den,
                       d_in, d_out, FILTER_ADJ, c->st);
  API_END
}

int vox8_oc_update(vox8_ctx* c, const uint8_t* d_design, const double* d_x, const double* d_dc,
                   double V_e, double target, double move, double* d_xnew, double* h_lmid) {
  API_BEGIN
  VOX8_CHECK(c && d_design && d_x && d_dc && d_xnew, "null argument");
  VOX8_CHECK(V_e > 0.0, "V_e must be positive");
  VOX8_TRY(launch_oc(c->g.n_elem, d_design, d_x, d_dc, V_e, target, move, c->d_oc, c->red_partials,
                     c->red_counters, d_xnew, 100, c->st));
  if (h_lmid) {
    double h[5];
    VOX8_TRY(sync_read(c, c->d_oc, h, 5));
    *h_lmid = h[2];
  }
  return OK;
  API_END
}

int vox8_stats(vox8_ctx* c, int64_t n, const uint8_t* d_mask, const double* d_a, const double* d_b,
               double* h_out4) {
  API_BEGIN
  VOX8_CHECK(c && d_a && h_out4 && n > 0, "bad argument");
  VOX8_TRY(launch_stats(n, d_mask, d_a, d_b, c->red_partials, c->red_counters, c->red_out, c->st));
  return sync_read(c, c->red_out, h_out4, 4);
  API_END
}

int vox8_block_dots(vox8_ctx* c, int64_t n, int k, const double* d_X, int64_t ld, const double* d_v,
                    double* h_out) {
  API_BEGIN
  VOX8_CHECK(c && d_X && d_v && h_out && n > 0 && ld >= n, "bad argument");
  VOX8_TRY(launch_block_dots(n, k, d_X, ld, d_v, c->red_partials, c->red_counters, c->red_out, c->st));
  return sync_read(c, c->red_out, h_out, 2 * k + 1);
  API_END
}

int vox8_axpby(vox8_ctx* c, int64_t n, const uint8_t* d_mask, double a, const double* d_x, double b,
               double* d_y) {
  API_BEGIN
  VOX8_CHECK(c && d_x && d_y && n > 0, "bad argument");
  return launch_axpby(n, d_mask, a, d_x, b, d_y, c->st);
  API_END
}

namespace {

// The launches of the loop tail (src/topopt.py:327-412): mode tracking by
// MAC, the gated eigenvalue-sensitivity branch and frequency penalty, the OC
// bisection, the density filter and the change / volume statistics.  The
// per-call values (eigenvalues, mu, the mode block's address) are read from
// c->d_tail_par, so the sequence is the same for every call with the same
// buffers and constants.
int tail_launches(vox8_ctx* c, const vox8_tail_args* a, hipStream_t st) {
  const int64_t ne = c->g.n_elem, nd = c->g.n_dofs;
  double* state = c->d_tail;
  const TailPar* par = c->d_tail_par;
  TailSelect s{};
  s.k = a->k;
  s.have_prev = a->have_prev;
  s.w_target = a->w_target;
  if (a->k > 0) {
    if (a->have_prev)
      VOX8_TRY(launch_block_dots(nd, a->k, a->d_X, a->ldx, a->d_vprev, c->red_partials,
                                 c->red_counters, c->red_out, st, &par->X));
    VOX8_TRY(launch_tail_select(s, par, c->red_out, state, st));
    VOX8_TRY(launch_gather_col(nd, par, a->ldx, state, a->d_vprev, st));
    // eigenvalue sensitivity branch, gated on the violation (state[2])
    const double* gate = state + 2;
    VOX8_TRY(launch_quadforms(c->g, c->fK, c->fM, c->d_active, c->dKe, c->dMe, nullptr, nd, 0,
                              c->d_wts, a->d_vprev, nullptr, c->d_kphi, c->d_mphi, c->red_partials,
                              c->red_counters, c->d_sums, c->d_tip, c->n_tip, c->tip_m, st,
                              gate, c->pattern != PAT_DENSE ? &c->c64 : nullptr));
    VOX8_TRY(launch_sens_finish(ne, c->d_active, a->d_rho, a->p, a->e_min, a->mthr, a->mq, c->d_ce,
                                nullptr, c->d_kphi, c->d_mphi, 0.0, c->d_sums, a->d_dlam, st,
                                state + 1, gate));
    VOX8_TRY(launch_filter(c->g.nx, c->g.ny, c->g.nz, c->ntap, c->d_toff, c->d_tw, c->d_fmask,
                           c->d_den, a->d_dlam, a->d_dlam_f, FILTER_FWD, st, gate));
    VOX8_TRY(launch_stats(ne, a->d_design, a->d_sens_f, nullptr, c->red_partials, c->red_counters,
                          state + 3, st, gate));
    VOX8_TRY(launch_stats(ne, a->d_design, a->d_dlam_f, nullptr, c->red_partials, c->red_counters,
                          state + 7, st, gate));
    VOX8_TRY(launch_tail_penalty(ne, par, state, a->d_dlam_f, a->d_sens_f, st));
  } else {
    VOX8_TRY(launch_tail_select(s, par, nullptr, state, st));
  }
  VOX8_TRY(launch_oc(ne, a->d_design, a->d_x, a->d_sens_f, a->V_e, a->target_volume, a->move,
                     c->d_oc, c->red_partials, c->red_counters, a->d_xnew, 100, st));
  VOX8_TRY(launch_filter(c->g.nx, c->g.ny, c->g.nz, c->ntap, c->d_toff, c->d_tw, c->d_fmask, c->d_den,
                         a->d_xnew, a->d_check, FILTER_FWD, st));
  VOX8_TRY(launch_stats(ne, nullptr, a->d_check, a->d_phys_prev, c->red_partials, c->red_counters,
                        state + 11, st));
  VOX8_TRY(launch_stats(ne, a->d_design, a->d_xnew, nullptr, c->red_partials, c->red_counters,
                        state + 15, st));
  return OK;
}

// Every value tail_launches passes to a kernel by value, except those read
// from d_tail_par: two calls with equal keys record identical launches.
std::vector<unsigned char> tail_key(const vox8_ctx* c, const vox8_tail_args* a) {
  std::vector<unsigned char> k;
  auto put = [&](const auto& v) {
    const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
    k.insert(k.end(), p, p + sizeof(v));
  };
  put(a->k); put(a->have_prev); put(a->w_target); put(a->ldx);
  put(a->d_X != nullptr); put(a->d_vprev); put(a->d_rho); put(a->d_dlam); put(a->d_dlam_f);
  put(a->d_sens_f); put(a->d_design); put(a->d_x); put(a->d_xnew); put(a->d_check);
  put(a->d_phys_prev); put(a->p); put(a->e_min); put(a->mthr); put(a->mq); put(a->V_e);
  put(a->target_volume); put(a->move);
  put(c->g); put(c->fK); put(c->fM); put(c->d_active); put(c->dKe); put(c->dMe); put(c->d_wts);
  put(c->d_kphi); put(c->d_mphi); put(c->d_ce); put(c->d_sums); put(c->d_tip); put(c->n_tip);
  put(c->tip_m); put(c->pattern); put(c->c64); put(c->ntap); put(c->d_toff); put(c->d_tw);
  put(c->d_fmask); put(c->d_den); put(c->d_oc); put(c->d_tail); put(c->d_tail_par);
  put(c->red_partials); put(c->red_counters); put(c->red_out);
  return k;
}

// Capture tail_launches on the context's stream (or, when that is the null
// stream, which cannot be captured, on the multigrid's private capture
// stream) under the process-wide HIP lock, and instantiate.  Returns false,
// with no error pending, when the runtime rejects the capture.
bool tail_capture(vox8_ctx* c, const vox8_tail_args* a, hipGraphExec_t* exec) {
  HipSerial lock(hip_serial());
  *exec = nullptr;
  if (!c->st && !c->mg.cap_st &&
      hipStreamCreateWithFlags(&c->mg.cap_st, hipStreamNonBlocking) != hipSuccess) {
    c->mg.cap_st = nullptr;
    (void)hipGetLastError();
    return false;
  }
  hipStream_t cs = c->st ? c->st : c->mg.cap_st;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  const int rc = tail_launches(c, a, cs);
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

}  // namespace

int vox8_simp_tail(vox8_ctx* c, const vox8_tail_args* a, double* h_state) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(a && h_state, "null argument");
  VOX8_CHECK(a->k >= 0 && a->k <= 12, "tail: 0..12 modes");
  VOX8_CHECK(a->d_sens_f && a->d_design && a->d_x && a->d_xnew && a->d_check && a->d_phys_prev,
             "tail: null element field");
  VOX8_CHECK(a->V_e > 0.0, "V_e must be positive");
  if (!c->have_filter) { set_error("filter not set up (vox8_filter_setup)"); return ERR_STATE; }
  if (a->k > 0) {
    VOX8_CHECK(a->d_X && a->h_lam && a->d_vprev && a->d_rho && a->d_dlam && a->d_dlam_f,
               "tail: null modal argument");
    VOX8_CHECK(a->ldx >= c->g.n_dofs, "tail: ldx < n_dofs");
    VOX8_CHECK(a->w_target > 0.0, "tail: w_target must be positive");
  }
  // the per-call values, through pinned memory (the previous call's read
  // synchronised the stream, so the staging is free)
  TailPar* hp = c->h_tail_par;
  *hp = TailPar{};
  for (int j = 0; j < a->k; ++j) hp->lam[j] = a->h_lam[j];
  hp->mu = a->mu;
  hp->X = a->d_X;
  VOX8_HIP(hipMemcpyAsync(c->d_tail_par, hp, sizeof(TailPar), hipMemcpyHostToDevice, c->st));
  hipGraphExec_t exec = nullptr;
  if (c->tail_graphs) {
    std::vector<unsigned char> key = tail_key(c, a);
    for (auto& tg : c->tail_cache)
      if (tg.key == key) exec = tg.exec;
    if (!exec) {
      if (c->tail_cache.size() >= 4) {               // least recently used out
        size_t old = 0;
        for (size_t i = 1; i < c->tail_cache.size(); ++i)
          if (c->tail_cache[i].last_use < c->tail_cache[old].last_use) old = i;
        {
          HipSerial lock(hip_serial());
          (void)hipGraphExecDestroy(c->tail_cache[old].exec);
        }
        c->tail_cache.erase(c->tail_cache.begin() + old);
      }
      if (tail_capture(c, a, &exec)) {
        vox8_ctx::TailGraph tg;
        tg.key = std::move(key);
        tg.exec = exec;
        c->tail_cache.push_back(std::move(tg));
      } else if (c->tail_graphs == 2) {
        set_error("vox8_simp_tail: the runtime refused the graph capture");
        return ERR_HIP;
      } else {
        c->tail_graphs = 0;                           // the runtime refuses: launch directly
      }
    }
  }
  if (exec) {
    for (auto& tg : c->tail_cache)
      if (tg.exec == exec) tg.last_use = ++c->tail_clock;
    VOX8_HIP(hipGraphLaunch(exec, c->st));
  } else {
    VOX8_TRY(tail_launches(c, a, c->st));
  }
  return sync_read(c, c->d_tail, h_state, TAIL_STATE);
  API_END
}

int vox8_set_tail_graph(vox8_ctx* c, int enable) {
  VOX8_CHECK(c != nullptr, "null context");
  HipSerial lock(hip_serial());
  VOX8_CHECK(enable >= 0 && enable <= 2, "enable must be 0, 1 or 2");
  c->tail_graphs = enable;
  for (auto& tg : c->tail_cache) (void)hipGraphExecDestroy(tg.exec);
  c->tail_cache.clear();
  return OK;
}

int vox8_effective_mass(vox8_ctx* c, int k, const double* d_X, int64_t ld, int direction,
                        double* h_frac, double* h_L, double* h_gm, double* h_total) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(d_X && k >= 1 && k <= 16, "effective mass: 1..16 modes");
  VOX8_CHECK(ld >= c->g.n_dofs, "ld < n_dofs");
  VOX8_CHECK(direction >= 0 && direction <= 2, "direction must be 0 (x), 1 (y) or 2 (z)");
  const int64_t nd = c->g.n_dofs;
  if (c->em_cols < k + 2) {
    dfree(c, c->d_em);
    c->d_em = nullptr;
    c->em_cols = 0;
    VOX8_TRY(dalloc(c, &c->d_em, (size_t)(k + 2) * nd));
    c->em_cols = k + 2;
  }
  double* MX = c->d_em;
  double* r = c->d_em + (int64_t)(k + 1) * nd;
  VOX8_TRY(launch_unit_translation(c->g.n_nodes, direction, c->d_fixed, r, c->st));
  VOX8_TRY(apply_op<double>(c, OP_M, d_X, ld, nullptr, MX, nd, k, DotFuse()));
  VOX8_TRY(apply_op<double>(c, OP_M, r, nd, nullptr, MX + (int64_t)k * nd, nd, 1, DotFuse()));
  VOX8_TRY(launch_effective_mass(c->g.n_nodes, direction, k, d_X, ld, MX, c->red_partials,
                                 4096 * 40, c->red_out, c->st));
  double h[64];
  VOX8_TRY(sync_read(c, c->red_out, h, 3 * k + 1));
  for (int j = 0; j < k; ++j) {
    if (h_frac) h_frac[j] = h[j];
    if (h_L) h_L[j] = h[k + j];
    if (h_gm) h_gm[j] = h[2 * k + j];
  }
  if (h_total) *h_total = h[3 * k];
  return OK;
  API_END
}

int vox8_trial_mfma_apply(vox8_ctx* c, const float* d_X, float* d_Y, int nrhs) {
  API_BEGIN
  VOX8_TRY(require(c, true));
  VOX8_CHECK(d_X && d_Y, "null buffer");
  return launch_apply_mfma_trial(c->g, c->dKe32, c->fK32, c->d_fixed, d_X, d_Y, nrhs, c->st);
  API_END
}

int vox8_compose_density(vox8_ctx* c, const uint8_t* d_design, const uint8_t* d_in_arm,
                         const double* d_filt, double* d_out) {
  API_BEGIN
  VOX8_CHECK(c && d_design && d_in_arm && d_filt && d_out, "null argument");
  return launch_compose(c->g.n_elem, d_design, d_in_arm, d_filt, d_out, c->st);
  API_END
}

}  // extern "C"
