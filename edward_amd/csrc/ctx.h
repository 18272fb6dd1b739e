// Library context shared by the GLM engine, the surrogate and the
// hierarchical engine.  Host-side only (never passed to kernels).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/edhmc.h"

namespace edhmc {

struct Surrogate;   // mlp.hip
struct HierModel;   // hier.hip

// Kernel classes timed by the event profiler.
enum ProfKind {
  PK_GEMM_GRAD = 0,   // GEMM1 with residual epilogue (gradient pass)
  PK_GEMM_LL = 1,     // GEMM1 forward with log-likelihood epilogue only
  PK_GEMM_XTR = 2,    // GEMM2 (X^T residual, split-K)
  PK_CHAIN = 3,       // per-chain leapfrog / accept kernels
  PK_MLP = 4,         // surrogate MLP GEMMs (+ fused epilogues)
  PK_HIER = 5,        // hierarchical segmented reductions
  PK_COUNT = 6
};

// hipEvent pairs recorded around launches on the launch stream; summed at
// collection time.  Off unless enabled.
struct Profiler {
  bool on = false;
  uint32_t mask = ~0u;   // kernel classes (bit k = class k) that get events
  std::vector<hipEvent_t> pool;
  std::vector<std::pair<int, int>> marks;   // (kind, index of start event)
  double total_ms[PK_COUNT] = {};
  int64_t count[PK_COUNT] = {};
  size_t next = 0;
  hipEvent_t get() {
    if (next == pool.size()) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return nullptr;
      pool.push_back(e);
    }
    return pool[next++];
  }
};

// A device allocation owned by the context, or (view) a part of another one.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool view = false;
  void release() {
    if (p && !view) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    view = false;
  }
};

}  // namespace edhmc

struct edhmc_ctx {
  int device = 0;
  std::string err;

  // ---- generalised linear model ----
  bool bound = false;
  int cdt = EDHMC_F32;   // compute dtype of X / positions in the GEMMs
  int lik = EDHMC_LIK_BERNOULLI;
  float lik_scale = 1.0f;
  float lik_weight = 1.0f;   // likelihood multiplier (Edward's `scale` for the SG samplers)
  int64_t N = 0, N_pad = 0;
  int D = 0, D_pad = 0;
  edhmc::DevBuf Xr, Xc, y, ploc, pscale;
  edhmc::DevBuf Xf;   // bf16 X in MFMA fragment order (the W-stationary kernels' copy)
  edhmc::DevBuf pkind, plnorm;   // constrained prior families (has_kinds)
  bool has_kinds = false;

  // ---- chains ----
  int C = 0, C_pad = 0;
  bool state_valid = false;
  bool g_valid = false;   // g holds the exact gradient at z
  bool inner_exact = false;   // bf16: inner leapfrog gradients on both position planes
  edhmc::DevBuf z, r, g, z0, g0, z_lp, logp, K0, flag, ndiv;
  edhmc::DevBuf ratio;   // [C] FP64 Metropolis log ratio of the last HMC transition
  // ---- gradient scratch ----
  edhmc::DevBuf Rt, Gpart, llpart, ll;
  int ksplit = 1;
  int n_tiles1 = 0;      // GEMM1 column tiles (= llpart leading dim)
  // ---- moments ----
  bool mom_on = false;
  int64_t mom_count = 0;
  edhmc::DevBuf mom_mean, mom_m2;
  // ---- training-pair recording ----
  float* rec_W = nullptr;
  float* rec_G = nullptr;
  int64_t rec_cap = 0, rec_count = 0;
  int rec_stride = 1;
  // ---- misc scratch ----
  edhmc::DevBuf tmp_count, stats_part;
  edhmc::DevBuf pred_col;   // posterior-predictive column partials [m_tiles][N_pad]
  edhmc::DevBuf ll_split;   // log-likelihood range sums [16][C_pad] (fixed-order reduction)
  edhmc::DevBuf newton;          // scratch of edhmc_newton_terms

  edhmc::Profiler prof;
  edhmc::Surrogate* sur = nullptr;
  edhmc::HierModel* hier = nullptr;

  ~edhmc_ctx();
};

namespace edhmc {

inline int set_err(edhmc_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

inline int dev_alloc(edhmc_ctx* ctx, DevBuf& b, size_t bytes, bool zero = true) {
  if (b.p && b.bytes == bytes) {
    if (zero && hipMemset(b.p, 0, bytes) != hipSuccess)
      return set_err(ctx, EDHMC_EHIP, "hipMemset failed");
    return EDHMC_OK;
  }
  b.release();
  if (bytes == 0) return EDHMC_OK;
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    b.p = nullptr;
    return set_err(ctx, EDHMC_ENOMEM, "hipMalloc of %zu bytes failed", bytes);
  }
  b.bytes = bytes;
  if (zero && hipMemset(b.p, 0, bytes) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "hipMemset failed");
  return EDHMC_OK;
}

// Bracket a launch with profiler events when profiling is on.
struct ProfScope {
  edhmc_ctx* ctx;
  hipStream_t st;
  int kind;
  int idx = -1;
  ProfScope(edhmc_ctx* c, int k, hipStream_t s);
  ~ProfScope();
};

int prof_collect(edhmc_ctx* ctx);   // sync + fold recorded pairs into totals

inline int check_launch(edhmc_ctx* ctx, const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "%s: %s", what, hipGetErrorString(e));
  return EDHMC_OK;
}

inline int64_t round_up(int64_t x, int64_t m) { return (x + m - 1) / m * m; }

// RAII device guard: make ctx->device current for the call.
struct DeviceGuard {
  int prev = -1;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) prev = -1;
    if (prev != dev) (void)hipSetDevice(dev);
  }
  ~DeviceGuard() {
    int cur;
    if (prev >= 0 && hipGetDevice(&cur) == hipSuccess && cur != prev) (void)hipSetDevice(prev);
  }
};

#define EDHMC_TRY(expr)                 \
  do {                                  \
    int _rc = (expr);                   \
    if (_rc != EDHMC_OK) return _rc;    \
  } while (0)

// Implemented in the other translation units.
void surrogate_destroy(Surrogate* s);
void surrogate_invalidate(Surrogate* s);   // chain state changed under it
void hier_destroy(HierModel* h);

// GLM pieces reused by the surrogate transition (edhmc_glm.hip).
int glm_forward_ll(edhmc_ctx* ctx, hipStream_t st);   // llpart at z_lp (split)
int glm_write_lp(edhmc_ctx* ctx, hipStream_t st);     // z -> z_lp planes
int glm_refresh_logp(edhmc_ctx* ctx, hipStream_t st);

}  // namespace edhmc
