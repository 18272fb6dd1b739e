// Batched hierarchical logistic regression on MI355X (gfx950).
//
//   a_g ~ N(mu, exp(log_sigma)^2), mu ~ N(0,1), log_sigma ~ N(0,1),
//   w ~ N(0, I_D),  y_n ~ Bernoulli(a_{g(n)} + x_n . w),  rows sorted by group.
//
// In Edward this is Bernoulli(logits=tf.gather(a, group) + ed.dot(X, w)): the
// gather gives the IndexedSlices gradient of tests/inferences/hmc_test.py:105-123
// that hmc.py:203,208 densifies.  Latent vector per chain:
// [a (G), mu, log_sigma, w (D)].
//
// One gradient evaluation:
//   GEMM1  eta = X w + a_{g(n)} (fp32 MFMA; the intercept is gathered from the
//          chain state in the epilogue), residuals R^T, FP64 ll partials
//   GEMM2  grad_w = X^T R (split-K)
//   seg    grad_a[g] = sum of R over group g's rows: a ragged segmented
//          reduction over the sorted row ranges, thread per chain (coalesced
//          across chains in the K-blocked residual layout)
//   chain  hyperprior terms (reductions over the G intercepts), leapfrog /
//          Metropolis test as for the GLM.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "ctx.h"
#include "philox.h"
#include "transform.h"

namespace edhmc {

int glm_grad_eval_offset(edhmc_ctx* ctx, const float* goff, int64_t goff_ld, const int* gidx,
                         bool ll, hipStream_t st);
int glm_ensure_chains(edhmc_ctx* ctx, int C);

struct HierModel {
  int G = 0, D = 0, dim = 0, dim_pad = 0;
  // prior of the group-scale coordinate u (see HierArgs::skind); default:
  // u = log sigma ~ Normal(0, 1)
  int skind = kPriorNormal;
  float sp1 = 0.0f, sp2 = 1.0f;
  double slnorm = -0.5 * kLog2Pi;
  int64_t N = 0;
  DevBuf gidx, offsets, ones, zeros;
  int C = 0, C_pad = 0;
  bool state_valid = false;
  DevBuf th, r, g, th0, g0, ga, logp, K0;
};

void hier_destroy(HierModel* h) {
  if (!h) return;
  DevBuf* b[] = {&h->gidx, &h->offsets, &h->ones, &h->zeros, &h->th, &h->r, &h->g, &h->th0,
                 &h->g0, &h->ga, &h->logp, &h->K0};
  for (DevBuf* x : b) x->release();
  delete h;
}

constexpr int kHT = 256;   // threads per chain workgroup

__global__ void group_index_kernel(const int64_t* __restrict__ off, int G, int* __restrict__ gidx) {
  const int g = blockIdx.x;
  for (int64_t n = off[g] + threadIdx.x; n < off[g + 1]; n += blockDim.x) gidx[n] = g;
}

// grad_a[c, g] = sum_{n in group g} R[c, n]; R^T is fp32 K-blocked
// ([N_pad/16][C_pad][16]: 16 rows of one chain are 64 contiguous bytes).  One
// workgroup per group; four lanes per chain each load a float4 (rows
// 4q..4q+3 of a 16-row block), so a wave reads 16 chains x 64 B = 1 KiB
// contiguous per block.  The four lane sums are combined by lane shuffles in
// a fixed order (deterministic).
__global__ void __launch_bounds__(256)
seg_sum_kernel(const float* __restrict__ Rt, const int64_t* __restrict__ off, int C, int C_pad,
               int G, float* __restrict__ ga) {
  const int g = blockIdx.x;
  const int64_t b = off[g], e = off[g + 1];
  const int q = threadIdx.x & 3;
  const int64_t kb0 = b >> 4, kb1 = e > b ? ((e - 1) >> 4) : kb0 - 1;
  for (int c0 = 0; c0 < C; c0 += 64) {
    const int c = c0 + (threadIdx.x >> 2);
    float s = 0.0f;
    if (c < C) {
#pragma unroll 4
      for (int64_t kb = kb0; kb <= kb1; ++kb) {
        const f32x4_t v = *reinterpret_cast<const f32x4_t*>(Rt + (kb * C_pad + c) * 16 + 4 * q);
        const int64_t n0 = kb * 16 + 4 * q;
#pragma unroll
        for (int i = 0; i < 4; ++i) s += (n0 + i >= b && n0 + i < e) ? v[i] : 0.0f;
      }
    }
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (c < C && q == 0) ga[(int64_t)c * G + g] = s;
  }
}

struct HierArgs {
  int C, G, D, dim, dim_pad, D_pad, C_pad;
  float* th; float* r; float* g; float* th0; float* g0;
  const float* ga; const float* Gpart; int ksplit;
  const double* ll;
  double* logp; double* K0;
  float* zlp;              // w part for GEMM1 [C_pad x D_pad]
  float half, full;
  const float* r_fed; const double* u_fed; uint64_t seed, counter;
  float* sample_out; int32_t* accepted_out; int32_t* n_accept; int32_t* ndiv;
  double* ratio_out;
  const float* W_in; float* grad_out; double* logp_out;
  // group-scale coordinate u: sigma = exp(u) (skind >> 4 == 0) or
  // softplus(u) (== 1); its prior is the family skind & 15 of transform.h
  // in u-space with parameters sp1, sp2 and constant slnorm
  int skind; float sp1, sp2; double slnorm;
};

// sigma and d log sigma / du of the group-scale coordinate
__device__ inline void hier_sigma(const HierArgs& a, float u, float& sig, double& dlog) {
  if ((a.skind >> 4) == 0) {
    sig = expf(u);
    dlog = 1.0;
  } else {
    const double sp = softplus_f64((double)u);
    sig = (float)sp;
    dlog = sig_over_softplus_f64((double)u);
  }
}

__device__ inline double nlogpdf(double x, double m, double s) {
  const double t = (x - m) / s;
  return -0.5 * t * t - log(s) - 0.5 * kLog2Pi;
}

// Gradient of the log joint at th (into g) and, with want_lp, the log joint.
__device__ inline void hier_assemble(const HierArgs& a, int c, bool want_lp, double* red,
                                     double& logp) {
  const float* th = a.th + (int64_t)c * a.dim_pad;
  const float mu = th[a.G], ls = th[a.G + 1];
  float sig;
  double dlog;
  hier_sigma(a, ls, sig, dlog);
  const float inv2 = 1.0f / (sig * sig);
  const PriorTerm sprior = prior_term(a.skind & 15, (double)ls, a.sp1, a.sp2, a.slnorm);
  double s1 = 0.0, s2 = 0.0, plp = 0.0;
  for (int j = threadIdx.x; j < a.G; j += kHT) {
    const double z = ((double)th[j] - mu) / sig;
    s1 += z / sig;
    s2 += z * z;
  }
  s1 = block_sum_f64<kHT>(s1, red);
  s2 = block_sum_f64<kHT>(s2, red);
  float* g = a.g + (int64_t)c * a.dim_pad;
  const int64_t gstride = (int64_t)a.C_pad * a.D_pad;
  for (int j = threadIdx.x; j < a.dim; j += kHT) {
    float gv;
    if (j < a.G) {
      gv = a.ga[(int64_t)c * a.G + j] - (th[j] - mu) * inv2;
    } else if (j == a.G) {
      gv = (float)(s1 - (double)mu);
    } else if (j == a.G + 1) {
      // d/du sum_g log N(a_g; mu, sigma) = (s2 - G) d log sigma / du
      gv = (float)((s2 - (double)a.G) * dlog + sprior.dlp);
    } else {
      const int d = j - a.G - 2;
      const float* gp = a.Gpart + (int64_t)c * a.D_pad + d;
      float s = 0.0f;
      for (int q = 0; q < a.ksplit; ++q) s += gp[q * gstride];
      gv = s - th[j];
    }
    g[j] = gv;
    if (want_lp) {
      if (j < a.G) plp += nlogpdf(th[j], mu, sig);
      else if (j == a.G + 1) plp += sprior.lp;
      else plp += nlogpdf(th[j], 0.0, 1.0);
    }
  }
  if (want_lp) logp = a.ll[c] + block_sum_f64<kHT>(plp, red);
}

__device__ inline void hier_write_lp(const HierArgs& a, int c) {
  for (int d = threadIdx.x; d < a.D; d += kHT)
    a.zlp[(int64_t)c * a.D_pad + d] = a.th[(int64_t)c * a.dim_pad + a.G + 2 + d];
}

__global__ void __launch_bounds__(kHT) hier_load_kernel(HierArgs a, int* nonfinite) {
  const int c = blockIdx.x;
  int bad = 0;
  for (int j = threadIdx.x; j < a.dim; j += kHT) {
    const float v = a.W_in[(int64_t)c * a.dim + j];
    if (!isfinite(v)) bad = 1;
    a.th[(int64_t)c * a.dim_pad + j] = v;
  }
  if (bad) atomicAdd(nonfinite, 1);
  __syncthreads();
  hier_write_lp(a, c);
}

__global__ void __launch_bounds__(kHT) hier_eval_kernel(HierArgs a) {
  __shared__ double red[kHT / 64];
  const int c = blockIdx.x;
  double lp = 0.0;
  hier_assemble(a, c, true, red, lp);
  if (threadIdx.x == 0) {
    a.logp[c] = lp;
    if (a.logp_out) a.logp_out[c] = lp;
  }
  if (a.grad_out) {
    __syncthreads();
    for (int j = threadIdx.x; j < a.dim; j += kHT)
      a.grad_out[(int64_t)c * a.dim + j] = a.g[(int64_t)c * a.dim_pad + j];
  }
}

__global__ void __launch_bounds__(kHT) hier_begin_kernel(HierArgs a) {
  __shared__ double red[kHT / 64];
  const int c = blockIdx.x;
  double k = 0.0;
  const int nb = (a.dim + 3) / 4;
  for (int jb = threadIdx.x; jb < nb; jb += kHT) {
    f32x4_t rv;
    if (a.r_fed) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int j = 4 * jb + e;
        rv[e] = j < a.dim ? a.r_fed[(int64_t)c * a.dim + j] : 0.0f;
      }
    } else {
      rv = momentum4(a.seed, a.counter, (uint32_t)jb, (uint32_t)c);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int j = 4 * jb + e;
      if (j >= a.dim) continue;
      const int64_t idx = (int64_t)c * a.dim_pad + j;
      float r = rv[e];
      k += (double)r * (double)r;
      const float tv = a.th[idx], gv = a.g[idx];
      a.th0[idx] = tv;
      a.g0[idx] = gv;
      r = r + a.half * gv;
      a.r[idx] = r;
      a.th[idx] = tv + a.full * r;
    }
  }
  k = block_sum_f64<kHT>(k, red);
  if (threadIdx.x == 0) a.K0[c] = 0.5 * k;
  __syncthreads();
  hier_write_lp(a, c);
}

__global__ void __launch_bounds__(kHT) hier_mid_kernel(HierArgs a) {
  __shared__ double red[kHT / 64];
  const int c = blockIdx.x;
  double unused = 0.0;
  hier_assemble(a, c, false, red, unused);
  __syncthreads();
  for (int j = threadIdx.x; j < a.dim; j += kHT) {
    const int64_t idx = (int64_t)c * a.dim_pad + j;
    const float gv = a.g[idx];
    float r = a.r[idx];
    r = r + a.half * gv;
    r = r + a.half * gv;
    a.r[idx] = r;
    a.th[idx] = a.th[idx] + a.full * r;
  }
  __syncthreads();
  hier_write_lp(a, c);
}

__global__ void __launch_bounds__(kHT) hier_last_kernel(HierArgs a) {
  __shared__ double red[kHT / 64];
  __shared__ int acc_flag;
  const int c = blockIdx.x;
  double logpL = 0.0;
  hier_assemble(a, c, true, red, logpL);
  __syncthreads();
  double kl = 0.0;
  int bad = 0;
  for (int j = threadIdx.x; j < a.dim; j += kHT) {
    const int64_t idx = (int64_t)c * a.dim_pad + j;
    const float gv = a.g[idx];
    const float r = a.r[idx] + a.half * gv;
    a.r[idx] = r;
    kl += (double)r * (double)r;
    if (!isfinite(gv) || !isfinite(a.th[idx])) bad = 1;
  }
  kl = block_sum_f64<kHT>(kl, red);
  const double nbad = block_sum_f64<kHT>((double)bad, red);
  if (threadIdx.x == 0) {
    const double ratio = a.K0[c] - 0.5 * kl + logpL - a.logp[c];
    const double u = a.u_fed ? a.u_fed[c] : accept_uniform(a.seed, a.counter, (uint32_t)c);
    const int acc = (log(u) < ratio) && nbad == 0.0;
    acc_flag = acc;
    if (acc) a.logp[c] = logpL;
    if (a.accepted_out) a.accepted_out[c] = acc;
    if (a.n_accept) a.n_accept[c] += acc;
    if (a.ndiv && (nbad != 0.0 || !(ratio == ratio))) a.ndiv[c] += 1;
    if (a.ratio_out) a.ratio_out[c] = ratio;
  }
  __syncthreads();
  const int acc = acc_flag;
  for (int j = threadIdx.x; j < a.dim; j += kHT) {
    const int64_t idx = (int64_t)c * a.dim_pad + j;
    float tv;
    if (acc) {
      tv = a.th[idx];
    } else {
      tv = a.th0[idx];
      a.th[idx] = tv;
      a.g[idx] = a.g0[idx];
    }
    if (a.sample_out) a.sample_out[(int64_t)c * a.dim + j] = tv;
  }
}

// ---------------------------------------------------------------------------
static HierArgs hier_args(edhmc_ctx* ctx, HierModel* h) {
  HierArgs a = {};
  a.C = h->C; a.G = h->G; a.D = h->D; a.dim = h->dim; a.dim_pad = h->dim_pad;
  a.D_pad = ctx->D_pad; a.C_pad = ctx->C_pad;
  a.th = (float*)h->th.p; a.r = (float*)h->r.p; a.g = (float*)h->g.p;
  a.th0 = (float*)h->th0.p; a.g0 = (float*)h->g0.p;
  a.ga = (const float*)h->ga.p;
  a.Gpart = (const float*)ctx->Gpart.p; a.ksplit = ctx->ksplit;
  a.ll = (const double*)ctx->ll.p;
  a.logp = (double*)h->logp.p; a.K0 = (double*)h->K0.p;
  a.zlp = (float*)ctx->z_lp.p;
  a.ndiv = (int32_t*)ctx->ndiv.p;   // divergence counts share the GLM slot
  a.ratio_out = (double*)ctx->ratio.p;   // so does the last transition's log ratio
  a.skind = h->skind; a.sp1 = h->sp1; a.sp2 = h->sp2; a.slnorm = h->slnorm;
  return a;
}

static int hier_grad(edhmc_ctx* ctx, HierModel* h, bool ll, hipStream_t st) {
  EDHMC_TRY(glm_grad_eval_offset(ctx, (const float*)h->th.p, h->dim_pad, (const int*)h->gidx.p,
                                 ll, st));
  ProfScope ps(ctx, PK_HIER, st);
  hipLaunchKernelGGL(seg_sum_kernel, dim3(h->G), dim3(256), 0, st, (const float*)ctx->Rt.p,
                     (const int64_t*)h->offsets.p, h->C, ctx->C_pad, h->G, (float*)h->ga.p);
  return check_launch(ctx, "seg_sum");
}

static int hier_set_state(edhmc_ctx* ctx, const float* theta, int C, hipStream_t st) {
  HierModel* h = ctx->hier;
  if (!h) return set_err(ctx, EDHMC_ESTATE, "no hierarchical model bound");
  if (!theta) return set_err(ctx, EDHMC_EINVAL, "theta is NULL");
  EDHMC_TRY(glm_ensure_chains(ctx, C));
  if (h->C != C || h->C_pad != ctx->C_pad) {
    const size_t b = (size_t)ctx->C_pad * h->dim_pad * 4;
    EDHMC_TRY(dev_alloc(ctx, h->th, b));
    EDHMC_TRY(dev_alloc(ctx, h->r, b));
    EDHMC_TRY(dev_alloc(ctx, h->g, b));
    EDHMC_TRY(dev_alloc(ctx, h->th0, b));
    EDHMC_TRY(dev_alloc(ctx, h->g0, b));
    EDHMC_TRY(dev_alloc(ctx, h->ga, (size_t)C * h->G * 4));
    EDHMC_TRY(dev_alloc(ctx, h->logp, (size_t)C * 8));
    EDHMC_TRY(dev_alloc(ctx, h->K0, (size_t)C * 8));
    h->C = C;
    h->C_pad = ctx->C_pad;
  }
  EDHMC_TRY(dev_alloc(ctx, ctx->tmp_count, 4));
  HierArgs a = hier_args(ctx, h);
  a.W_in = theta;
  hipLaunchKernelGGL(hier_load_kernel, dim3(C), dim3(kHT), 0, st, a, (int*)ctx->tmp_count.p);
  EDHMC_TRY(check_launch(ctx, "hier_load"));
  int bad = 0;
  if (hipMemcpyAsync(&bad, ctx->tmp_count.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "hier set_state: copy failed");
  if (bad) {
    h->state_valid = false;
    return set_err(ctx, EDHMC_ENONFINITE, "theta contains %d non-finite rows", bad);
  }
  EDHMC_TRY(hier_grad(ctx, h, true, st));
  hipLaunchKernelGGL(hier_eval_kernel, dim3(C), dim3(kHT), 0, st, a);
  EDHMC_TRY(check_launch(ctx, "hier_eval"));
  h->state_valid = true;
  return EDHMC_OK;
}

static int hier_transition(edhmc_ctx* ctx, float eps, int L, const float* r0, const double* u,
                           uint64_t seed, uint64_t counter, int32_t* accepted_out,
                           float* sample_out, int32_t* n_accept, hipStream_t st) {
  HierModel* h = ctx->hier;
  if (!h || !h->state_valid) return set_err(ctx, EDHMC_ESTATE, "no hierarchical chain state");
  if (L < 1) return set_err(ctx, EDHMC_EINVAL, "n_steps must be >= 1, got %d", L);
  if (!std::isfinite(eps)) return set_err(ctx, EDHMC_EINVAL, "step_size must be finite");
  HierArgs a = hier_args(ctx, h);
  a.half = (float)(0.5 * (double)eps);
  a.full = eps;
  a.r_fed = r0;
  a.u_fed = u;
  a.seed = seed;
  a.counter = counter;
  {
    ProfScope ps(ctx, PK_CHAIN, st);
    hipLaunchKernelGGL(hier_begin_kernel, dim3(h->C), dim3(kHT), 0, st, a);
  }
  EDHMC_TRY(check_launch(ctx, "hier_begin"));
  for (int k = 1; k <= L; ++k) {
    const bool last = k == L;
    EDHMC_TRY(hier_grad(ctx, h, last, st));
    ProfScope ps(ctx, PK_CHAIN, st);
    if (!last) {
      hipLaunchKernelGGL(hier_mid_kernel, dim3(h->C), dim3(kHT), 0, st, a);
      EDHMC_TRY(check_launch(ctx, "hier_mid"));
    }
  }
  a.sample_out = sample_out;
  a.accepted_out = accepted_out;
  a.n_accept = n_accept;
  {
    ProfScope ps(ctx, PK_CHAIN, st);
    hipLaunchKernelGGL(hier_last_kernel, dim3(h->C), dim3(kHT), 0, st, a);
  }
  return check_launch(ctx, "hier_last");
}

}  // namespace edhmc

using namespace edhmc;

extern "C" {

int edhmc_bind_hier(edhmc_ctx* ctx, const float* X, const float* y, const int64_t* offsets,
                    int32_t G, int64_t N, int32_t D, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (!X || !y || !offsets) return set_err(ctx, EDHMC_EINVAL, "NULL argument");
  if (G <= 0 || N <= 0 || D <= 0) return set_err(ctx, EDHMC_EINVAL, "need G, N, D > 0");
  std::vector<int64_t> off(G + 1);
  if (hipMemcpyAsync(off.data(), offsets, (size_t)(G + 1) * 8, hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "bind_hier: copy of offsets failed");
  if (off[0] != 0 || off[G] != N)
    return set_err(ctx, EDHMC_EINVAL, "offsets must run from 0 to N=%lld", (long long)N);
  for (int g = 0; g < G; ++g)
    if (off[g + 1] < off[g]) return set_err(ctx, EDHMC_EINVAL, "offsets must be non-decreasing");
  if (ctx->hier) hier_destroy(ctx->hier);
  ctx->hier = nullptr;
  HierModel* h = new (std::nothrow) HierModel();
  if (!h) return set_err(ctx, EDHMC_ENOMEM, "host allocation");
  ctx->hier = h;
  h->G = G;
  h->D = D;
  h->N = N;
  h->dim = G + 2 + D;
  h->dim_pad = (int)round_up(h->dim, 4);
  // w ~ N(0, I): the GLM prior for the shared coefficients
  EDHMC_TRY(dev_alloc(ctx, h->ones, (size_t)D * 4, false));
  EDHMC_TRY(dev_alloc(ctx, h->zeros, (size_t)D * 4));
  std::vector<float> one(D, 1.0f);
  if (hipMemcpyAsync(h->ones.p, one.data(), (size_t)D * 4, hipMemcpyHostToDevice, st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "copy");
  EDHMC_TRY(edhmc_bind_glm(ctx, X, EDHMC_F32, y, N, D, EDHMC_LIK_BERNOULLI, 1.0f,
                           (const float*)h->zeros.p, (const float*)h->ones.p, EDHMC_F32, stream));
  EDHMC_TRY(dev_alloc(ctx, h->offsets, (size_t)(G + 1) * 8, false));
  if (hipMemcpyAsync(h->offsets.p, offsets, (size_t)(G + 1) * 8, hipMemcpyDeviceToDevice, st) !=
      hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "copy offsets");
  EDHMC_TRY(dev_alloc(ctx, h->gidx, (size_t)ctx->N_pad * 4));
  hipLaunchKernelGGL(group_index_kernel, dim3(G), dim3(256), 0, st,
                     (const int64_t*)h->offsets.p, G, (int*)h->gidx.p);
  EDHMC_TRY(check_launch(ctx, "group_index"));
  if (hipStreamSynchronize(st) != hipSuccess) return set_err(ctx, EDHMC_EHIP, "sync");
  return EDHMC_OK;
}

int edhmc_hier_set_scale_prior(edhmc_ctx* ctx, int32_t kind, double p1, double p2,
                               double lnorm) {
  if (!ctx) return EDHMC_EINVAL;
  HierModel* h = ctx->hier;
  if (!h) return set_err(ctx, EDHMC_ESTATE, "no hierarchical model bound");
  const int fam = kind & 15, map = kind >> 4;
  if (kind < 0 || map > 1 || fam > kPriorInvGamma || fam == kPriorBeta)
    return set_err(ctx, EDHMC_EINVAL, "scale prior kind %d: family Normal, Gamma or "
                   "InverseGamma, map exp (0) or softplus (1)", kind);
  if (!(p2 > 0.0) || !std::isfinite(p1) || !std::isfinite(p2) || !std::isfinite(lnorm) ||
      (fam != kPriorNormal && !(p1 > 0.0)))
    return set_err(ctx, EDHMC_EINVAL, "scale prior parameters must be finite and positive");
  h->skind = kind;
  h->sp1 = (float)p1;
  h->sp2 = (float)p2;
  h->slnorm = lnorm;
  h->state_valid = false;   // the carried log joint and gradient used the old prior
  return EDHMC_OK;
}

int edhmc_hier_set_state(edhmc_ctx* ctx, const float* theta, int32_t C, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  return hier_set_state(ctx, theta, C, (hipStream_t)stream);
}

int edhmc_hier_logp_grad(edhmc_ctx* ctx, const float* theta, int32_t C, float* grad_out,
                         double* logp_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  EDHMC_TRY(hier_set_state(ctx, theta, C, st));
  HierModel* h = ctx->hier;
  HierArgs a = hier_args(ctx, h);
  a.grad_out = grad_out;
  a.logp_out = logp_out;
  hipLaunchKernelGGL(hier_eval_kernel, dim3(C), dim3(kHT), 0, st, a);
  return check_launch(ctx, "hier_eval");
}

int edhmc_hier_transition(edhmc_ctx* ctx, float eps, int32_t n_steps, const float* r0,
                          const double* u, uint64_t seed, uint64_t counter, int32_t* accepted_out,
                          float* sample_out, int32_t* n_accept, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  return hier_transition(ctx, eps, n_steps, r0, u, seed, counter, accepted_out, sample_out,
                         n_accept, (hipStream_t)stream);
}

}  // extern "C"
