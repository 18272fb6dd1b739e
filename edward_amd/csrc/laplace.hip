// Newton terms of the GLM log joint at one position: log p, gradient and
// the negative Hessian P = X^T diag(s) X + diag(1 / scale^2), where s_n is
// the likelihood's curvature in eta_n (Bernoulli: sigma (1 - sigma);
// Normal: 1 / lik_scale^2).  These drive a Newton / Laplace warm start of
// the chains (the mode and the Gaussian approximation N(mode, P^-1) of the
// Laplace approximation, edward/inferences/laplace.py:17-137).
//
// Kernels (X as bound: row-major [N_pad, D_pad], fp32 or bf16):
//   newton_rows_kernel  one wave per row: eta = x.w (fp32 FMA, a lane holds
//                       8 consecutive columns), r = dll/deta, s, FP64 log-lik
//                       partial per workgroup.
//   newton_gram_kernel  64 x 64 blocks of sum_n s_n x_n x_n^T over a slice of
//                       the rows, staged through LDS in fp32 (a thread owns a
//                       4 x 4 sub-block); the blocks of the first column also
//                       sum r_n x_n (the gradient).  fp32 partials per slice.
//   newton_final_kernel FP64 sum of the slice partials, prior terms.
// Work: 2 N D_pad^2 flop for the Gram matrix (VALU, a few ms at N = 1e6,
// D = 512), one read of X for the rows.
#include <hip/hip_runtime.h>

#include <cmath>

#include "common.h"
#include "ctx.h"
#include "../../include/edhmc.h"

using namespace edhmc;

namespace {

constexpr int kGramBlock = 64;     // output block edge
constexpr int kGramRows = 32;      // rows staged per LDS batch
constexpr int kGramSlices = 64;    // row slices (split of the row sum)

template <typename T>
__device__ inline float load_x(const T* p);
template <>
__device__ inline float load_x<float>(const float* p) { return *p; }
template <>
__device__ inline float load_x<bf16_t>(const bf16_t* p) { return bf16_to_f32(*p); }

template <typename T>
__global__ void __launch_bounds__(256)
newton_rows_kernel(const T* __restrict__ X, const float* __restrict__ y, int64_t N, int D_pad,
                   const float* __restrict__ w, LikParams lp, float* __restrict__ r_out,
                   float* __restrict__ s_out, double* __restrict__ ll_part) {
  extern __shared__ float w_lds[];   // [D_pad]
  __shared__ double red[4];
  for (int d = threadIdx.x; d < D_pad; d += 256) w_lds[d] = w[d];
  __syncthreads();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double ll = 0.0;
  for (int64_t n = (int64_t)blockIdx.x * 4 + wave; n < N; n += (int64_t)gridDim.x * 4) {
    const T* row = X + n * D_pad;
    float acc = 0.0f;
    for (int d0 = lane * 8; d0 < D_pad; d0 += 512) {
#pragma unroll
      for (int e = 0; e < 8; ++e) acc = fmaf(load_x(row + d0 + e), w_lds[d0 + e], acc);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) {
      const float eta = acc, yv = y[n];
      float dll, s;
      double l;
      if (lp.lik == LIK_BERNOULLI) {
        const double e = exp(-fabs((double)eta));
        const double sig = eta >= 0.0f ? 1.0 / (1.0 + e) : e / (1.0 + e);
        dll = (float)((double)yv - sig);
        s = (float)(sig * (1.0 - sig));
        l = (double)yv * eta - (fmax((double)eta, 0.0) + log1p(e));
      } else {
        const double dd = (double)yv - eta;
        dll = (float)(dd * lp.inv_s2);
        s = lp.inv_s2;
        l = -0.5 * dd * dd * lp.inv_s2 + lp.ll_const;
      }
      r_out[n] = dll;
      s_out[n] = s;
      ll += l;
    }
  }
  ll = block_sum_f64<256>(ll, red);
  if (threadIdx.x == 0) ll_part[blockIdx.x] = ll;
}

// grid (nb, nb, kGramSlices), 256 threads: block (bi, bj) of the Gram matrix
// over rows [slice * rows_per, ...).  Thread (ty, tx) of a 16 x 16 layout
// owns rows bi*64 + 4 ty + i, columns bj*64 + 4 tx + j.
template <typename T>
__global__ void __launch_bounds__(256)
newton_gram_kernel(const T* __restrict__ X, int64_t N, int D_pad, const float* __restrict__ r,
                   const float* __restrict__ s, int64_t rows_per, float* __restrict__ H_part,
                   float* __restrict__ g_part) {
  __shared__ float A[kGramRows][kGramBlock + 4];
  __shared__ float B[kGramRows][kGramBlock + 4];
  __shared__ float S[kGramRows], Rv[kGramRows];
  const int bi = blockIdx.x, bj = blockIdx.y, slice = blockIdx.z;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int64_t n_lo = (int64_t)slice * rows_per;
  const int64_t n_hi = n_lo + rows_per < N ? n_lo + rows_per : N;
  float acc[4][4] = {};
  float gacc = 0.0f;   // gradient of column bi*64 + threadIdx.x (first 64 threads, bj == 0)
  for (int64_t n0 = n_lo; n0 < n_hi; n0 += kGramRows) {
    // stage 32 rows x 64 columns of both blocks: 2048 elements each, 8 per thread
    for (int q = threadIdx.x; q < kGramRows * kGramBlock; q += 256) {
      const int rr = q / kGramBlock, cc = q - rr * kGramBlock;
      const int64_t n = n0 + rr;
      float a = 0.0f, b = 0.0f;
      if (n < n_hi) {
        a = load_x(X + n * D_pad + bi * kGramBlock + cc);
        b = load_x(X + n * D_pad + bj * kGramBlock + cc);
      }
      A[rr][cc] = a;
      B[rr][cc] = b;
    }
    if (threadIdx.x < kGramRows) {
      const int64_t n = n0 + threadIdx.x;
      S[threadIdx.x] = n < n_hi ? s[n] : 0.0f;
      Rv[threadIdx.x] = n < n_hi ? r[n] : 0.0f;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < kGramRows; ++k) {
      const float sk = S[k];
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = A[k][4 * ty + i] * sk;
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = B[k][4 * tx + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    if (bj == 0 && threadIdx.x < kGramBlock) {
      for (int k = 0; k < kGramRows; ++k) gacc = fmaf(Rv[k], A[k][threadIdx.x], gacc);
    }
    __syncthreads();
  }
  float* out = H_part + (int64_t)slice * D_pad * D_pad;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
      out[(int64_t)(bi * kGramBlock + 4 * ty + i) * D_pad + bj * kGramBlock + 4 * tx + j] =
          acc[i][j];
  if (bj == 0 && threadIdx.x < kGramBlock)
    g_part[(int64_t)slice * D_pad + bi * kGramBlock + threadIdx.x] = gacc;
}

// P[i][j] = sum of slices + [i == j] / scale_i^2; grad_i = sum of slices -
// (w_i - loc_i) / scale_i^2; logp = sum of ll partials + prior (FP64).
__global__ void __launch_bounds__(256)
newton_final_kernel(const float* __restrict__ H_part, const float* __restrict__ g_part,
                    const double* __restrict__ ll_part, int n_ll, int D, int D_pad,
                    const float* __restrict__ w, const float* __restrict__ ploc,
                    const float* __restrict__ pscale, double* __restrict__ logp_out,
                    double* __restrict__ grad_out, double* __restrict__ P_out) {
  __shared__ double red[4];
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < (int64_t)D * D) {
    const int a = (int)(i / D), b = (int)(i - (int64_t)a * D);
    double v = 0.0;
    for (int q = 0; q < kGramSlices; ++q) v += H_part[(int64_t)q * D_pad * D_pad + (int64_t)a * D_pad + b];
    if (a == b) v += 1.0 / ((double)pscale[a] * (double)pscale[a]);
    P_out[i] = v;
  }
  if (blockIdx.x == 0) {
    double lp = 0.0;
    for (int d = threadIdx.x; d < D; d += 256) {
      double g = 0.0;
      for (int q = 0; q < kGramSlices; ++q) g += g_part[(int64_t)q * D_pad + d];
      const double sc = pscale[d], t = ((double)w[d] - (double)ploc[d]) / sc;
      grad_out[d] = g - t / sc;
      lp += -0.5 * t * t - log(sc) - 0.5 * kLog2Pi;
    }
    for (int q = threadIdx.x; q < n_ll; q += 256) lp += ll_part[q];
    lp = block_sum_f64<256>(lp, red);
    if (threadIdx.x == 0) *logp_out = lp;
  }
}

}  // namespace

extern "C" int edhmc_newton_terms(edhmc_ctx* ctx, const float* w, double* logp_out,
                                  double* grad_out, double* neg_hess_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!ctx->bound) return set_err(ctx, EDHMC_ESTATE, "no model bound");
  if (ctx->hier) return set_err(ctx, EDHMC_EUNSUPPORTED, "Newton terms are for the GLM");
  if (ctx->has_kinds)
    return set_err(ctx, EDHMC_EUNSUPPORTED, "Newton terms need Normal priors on every coordinate");
  if (!w || !logp_out || !grad_out || !neg_hess_out)
    return set_err(ctx, EDHMC_EINVAL, "NULL argument");
  // the likelihood weight multiplies the data terms (it is 1 for HMC)
  if (ctx->lik_weight != 1.0f)
    return set_err(ctx, EDHMC_EUNSUPPORTED, "Newton terms need likelihood weight 1");
  hipStream_t st = (hipStream_t)stream;
  const int D = ctx->D, D_pad = ctx->D_pad;
  const int64_t N = ctx->N;
  const int rows_grid = 2048;
  const int64_t rows_per = (N + kGramSlices - 1) / kGramSlices;
  // scratch: r and s [N_pad], slice partials of P and the gradient, log-lik
  // partials, the padded position
  const size_t rs_bytes = (size_t)ctx->N_pad * 8;
  const size_t h_bytes = (size_t)kGramSlices * D_pad * D_pad * 4;
  const size_t g_bytes = (size_t)kGramSlices * D_pad * 4;
  const size_t ll_bytes = (size_t)rows_grid * 8;
  const size_t w_bytes = (size_t)D_pad * 4;
  EDHMC_TRY(dev_alloc(ctx, ctx->newton, rs_bytes + h_bytes + g_bytes + ll_bytes + w_bytes, false));
  char* base = (char*)ctx->newton.p;
  float* r = (float*)base;
  float* s = r + ctx->N_pad;
  float* H = (float*)(base + rs_bytes);
  float* g = (float*)(base + rs_bytes + h_bytes);
  double* ll = (double*)(base + rs_bytes + h_bytes + g_bytes);
  float* wp = (float*)(base + rs_bytes + h_bytes + g_bytes + ll_bytes);
  if (hipMemsetAsync(wp, 0, w_bytes, st) != hipSuccess ||
      hipMemcpyAsync(wp, w, (size_t)D * 4, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "newton: copy of w failed");
  LikParams lp;
  lp.lik = ctx->lik;
  lp.inv_s2 = 1.0f / (ctx->lik_scale * ctx->lik_scale);
  lp.ll_const = -log((double)ctx->lik_scale) - 0.5 * kLog2Pi;
  const size_t shm = (size_t)D_pad * 4;
  const dim3 ggrid(D_pad / kGramBlock, D_pad / kGramBlock, kGramSlices);
  if (ctx->cdt == EDHMC_BF16) {
    hipLaunchKernelGGL(newton_rows_kernel<bf16_t>, dim3(rows_grid), dim3(256), shm, st,
                       (const bf16_t*)ctx->Xr.p, (const float*)ctx->y.p, N, D_pad, wp, lp, r, s,
                       ll);
    EDHMC_TRY(check_launch(ctx, "newton_rows"));
    hipLaunchKernelGGL(newton_gram_kernel<bf16_t>, ggrid, dim3(256), 0, st,
                       (const bf16_t*)ctx->Xr.p, N, D_pad, r, s, rows_per, H, g);
  } else {
    hipLaunchKernelGGL(newton_rows_kernel<float>, dim3(rows_grid), dim3(256), shm, st,
                       (const float*)ctx->Xr.p, (const float*)ctx->y.p, N, D_pad, wp, lp, r, s,
                       ll);
    EDHMC_TRY(check_launch(ctx, "newton_rows"));
    hipLaunchKernelGGL(newton_gram_kernel<float>, ggrid, dim3(256), 0, st,
                       (const float*)ctx->Xr.p, N, D_pad, r, s, rows_per, H, g);
  }
  EDHMC_TRY(check_launch(ctx, "newton_gram"));
  hipLaunchKernelGGL(newton_final_kernel, dim3((unsigned)(((int64_t)D * D + 255) / 256)),
                     dim3(256), 0, st, H, g, ll, rows_grid, D, D_pad, wp,
                     (const float*)ctx->ploc.p, (const float*)ctx->pscale.p, logp_out, grad_out,
                     neg_hess_out);
  return check_launch(ctx, "newton_final");
}
