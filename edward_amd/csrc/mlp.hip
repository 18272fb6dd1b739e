// Learned gradient surrogate: an MLP D -> H -> H -> D (tanh-GELU) whose output
// layer also sees the normalised input (a linear skip: out = W3 [h2 | x] + b3),
// trained on (w, grad log p(w | data)) pairs, then used as the leapfrog force.
// Near a log-concave posterior's mode the gradient is close to linear in w;
// the skip represents that part exactly and the hidden layers learn the rest.
//
// No reference counterpart (Edward's HMC always differentiates the log
// joint, hmc.py:199,206).  Exactness: leapfrog with any position-only force
// field is a composition of shears, so it is volume preserving and
// reversible; the Metropolis test still uses the exact log joint
// (hmc.py:100-111 with the exact _log_joint), so the chain targets the exact
// posterior (Neal 2011, "MCMC using Hamiltonian dynamics", sec. 5.5; Li et al.
// 2019, "Neural network gradient Hamiltonian Monte Carlo").
//
// Numerics: fp32 master weights, bf16 copies on the matrix cores (the same
// MFMA NT-GEMM as the log joint) with fp32 accumulation, fp32 Adam.
// Layouts are padded to 128 in every feature dimension and the batch; pads
// hold zeros, so they contribute nothing (GELU(0) = 0).
#include <algorithm>
#include <type_traits>
#include <cmath>

#include "common.h"
#include "ctx.h"
// the layer GEMMs' small tiles need few registers: allow four waves per SIMD
#define EDHMC_GEMM_WPE 4
#include "gemm_nt.h"
#include "philox.h"

namespace edhmc {

using CfgM = TileCfg<2, 2, 2, 2>;    // 128 x 128 tiles
// 128 x 128 tiles of eight waves (64 x 32 each) and 128-byte K stages, for
// 256 to 511 such tiles (the 4096-chain hidden layers: one workgroup per CU)
using CfgM8 = TileCfg<2, 1, 2, 4, 128>;
// 128 x 64 tiles of eight waves, for 128 to 255 128 x 128 tiles (the
// 4096-chain output layer, 4096 x 512: 256 workgroups)
using CfgM8N = TileCfg<1, 1, 4, 2, 128>;
// 64 x 64 tiles with 128-byte K stages for GEMMs with fewer output tiles
// (the training pass's weight gradients)
using CfgMS = TileCfg<1, 1, 2, 2, 128>;
#ifndef EDHMC_MLP_THRESH
#define EDHMC_MLP_THRESH 512
#endif

constexpr float kGeluK0 = 0.7978845608028654f;  // sqrt(2 / pi)
constexpr float kGeluK1 = 0.044715f;

// sigmoid(2 u) for u = k0 (x + k1 x^3): 1 / (1 + exp(-2 u)), from v_exp_f32
// and v_rcp_f32 (exp overflowing to +inf for very negative x gives 0).
__device__ inline float gelu_sig(float x) {
  const float p = fmaf(kGeluK1 * x * x, x, x);   // u / k0
  return fast_rcp(1.0f + __builtin_amdgcn_exp2f(p * (-2.0f * kGeluK0 * 1.4426950408889634f)));
}

// GELU(x) = x/2 (1 + tanh(k0 (x + k1 x^3)))  (Hendrycks & Gimpel 2016),
// evaluated as x sigmoid(2 k0 (x + k1 x^3))  (1 + tanh a = 2 sigmoid(2 a)).
__device__ inline float gelu(float x) { return x * gelu_sig(x); }

// d GELU / dx = s + 2 x s (1 - s) k0 (1 + 3 k1 x^2),  s = sigmoid(2 u).
__device__ inline float gelu_grad(float x) {
  const float sg = gelu_sig(x);
  return fmaf(2.0f * x * sg * (1.0f - sg), kGeluK0 * fmaf(3.0f * kGeluK1, x * x, 1.0f), sg);
}

struct Surrogate {
  int D = 0, H = 0, Bmax = 0;
  int Dp = 0, Hp = 0, Bp = 0;          // padded sizes
  int Kc = 0;                          // output-layer input width Hp + Dp ([h2 | x])
  int64_t n_params = 0;                // unpadded flat count
  // padded fp32 masters, grads, Adam moments: [W1 Hp x Dp | b1 Hp | W2 Hp x Hp | b2 Hp |
  //                                          W3 Dp x Kc | b3 Dp]
  int64_t oW1, ob1, oW2, ob2, oW3, ob3, n_pad;
  DevBuf P, Gr, Am, Av;
  DevBuf W1b, W2b, W3b, W2t, W3t;     // bf16 copies: forward [out x in], backward [in x out]
  DevBuf nrm;                          // in_shift, in_scale, out_shift, out_scale (each Dp)
  DevBuf x0, a1, h1, a2, h2, o;        // activations (batch Bp); h2 is [Bp x Kc] = [h2 | x]
  // second [h2 | x] buffer of the surrogate leapfrog: step k's output layer
  // reads buffer k & 1 while its epilogue writes the next input into the
  // other one, so no GEMM operand changes while the GEMM runs
  DevBuf h2x;
  DevBuf dO, da2, da1, h1t, h2t, dOt, da2t, da1t;
  DevBuf pool;   // one allocation: P .. lpart are views into it (edhmc_sur_create)
  DevBuf loss;                         // fp64 [1]
  DevBuf lpart;                        // fp64 [Bp / 32 x Dp / 32] loss partials (EpiLoss)
  // sampling state
  DevBuf ghat, ghat0;                  // [C x D_pad of the GLM]
  bool ghat_valid = false;
  bool norm_set = false;
};

void surrogate_invalidate(Surrogate* s) {
  if (s) s->ghat_valid = false;
}

void surrogate_destroy(Surrogate* s) {
  if (!s) return;
  DevBuf* b[] = {&s->P, &s->Gr, &s->Am, &s->Av, &s->W1b, &s->W2b, &s->W3b, &s->W2t, &s->W3t,
                 &s->nrm, &s->x0, &s->a1, &s->h1, &s->a2, &s->h2, &s->h2x, &s->o, &s->dO, &s->da2,
                 &s->da1, &s->h1t, &s->h2t, &s->dOt, &s->da2t, &s->da1t, &s->loss,
                 &s->lpart, &s->ghat, &s->ghat0, &s->pool};
  for (DevBuf* x : b) x->release();
  delete s;
}

// ---------------------------------------------------------------------------
// Epilogues
// ---------------------------------------------------------------------------
// a = acc + bias; store a (fp32, optional) and h = GELU(a) (bf16).
// Four bf16 values of rows m .. m + 3 (m a multiple of 4) in column n of a
// matrix, stored into its transpose out [N x ldt]: one 8-byte store.  In a
// 32 x 32 accumulator block, elements 4q .. 4q + 3 of a lane are such rows.
__device__ inline void store_t4(bf16_t* out, int64_t ldt, int n, int64_t m, const bf16_t (&v)[4]) {
  uint2 pk;
  pk.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
  pk.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
  *(uint2*)(out + (int64_t)n * ldt + m) = pk;
}

struct EpiBiasGelu {
  const float* bias;
  float* a_out;      // [M x ld] or null
  bf16_t* h_out;     // [M x ldh]
  int64_t ld, ldh;
  bf16_t* ht_out = nullptr;   // h transposed [N x ldt] (training), or null
  int64_t ldt = 0;
  template <typename Cfg>
  __device__ void apply(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p) const {
#pragma unroll
    for (int j = 0; j < Cfg::kWN; ++j) {
      const int n = p.n0 + 32 * j + (p.lane & 31);
      const float bj = bias[n];
#pragma unroll
      for (int i = 0; i < Cfg::kWM; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          bf16_t hb[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = 4 * q + v;
            const int64_t m = p.m0 + 32 * i + acc_row(r, p.lane);
            const float a = acc[i][j][r] + bj;
            if (a_out) a_out[m * ld + n] = a;
            hb[v] = f32_to_bf16(gelu(a));
            h_out[m * ldh + n] = hb[v];
          }
          if (ht_out) store_t4(ht_out, ldt, n, p.m0 + 32 * i + acc_row(4 * q, p.lane), hb);
        }
    }
  }
};

// da = acc * GELU'(a) (bf16); optionally also transposed.
struct EpiGeluBack {
  const float* a_in;
  bf16_t* da_out;
  int64_t ld;
  bf16_t* dat_out = nullptr;   // da transposed [N x ldt], or null
  int64_t ldt = 0;
  template <typename Cfg>
  __device__ void apply(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p) const {
#pragma unroll
    for (int j = 0; j < Cfg::kWN; ++j) {
      const int n = p.n0 + 32 * j + (p.lane & 31);
#pragma unroll
      for (int i = 0; i < Cfg::kWM; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          bf16_t db[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int r = 4 * q + v;
            const int64_t m = p.m0 + 32 * i + acc_row(r, p.lane);
            db[v] = f32_to_bf16(acc[i][j][r] * gelu_grad(a_in[m * ld + n]));
            da_out[m * ld + n] = db[v];
          }
          if (dat_out) store_t4(dat_out, ldt, n, p.m0 + 32 * i + acc_row(4 * q, p.lane), db);
        }
    }
  }
};

// Training output: o = acc + b3; loss += (o - t)^2; dO = 2 (o - t) / (B D),
// also transposed.  Target row m is G[idx[m]] when idx is given (a row
// index outside [0, P) reads NaN, so the batch's loss is NaN), else G[m].
struct EpiLoss {
  const float* bias;
  const float* G;            // targets [B x D] (raw gradients), or [P x D] with idx
  const int64_t* idx;        // batch rows of G, or null
  int64_t P;
  const float* out_shift; const float* out_scale;
  int B, D;
  int64_t ld;                // Dp
  float coef;                // 2 / (B D)
  bf16_t* dO;
  bf16_t* dOt;               // dO transposed [Dp x ldt]
  int64_t ldt;
  // per-wave FP64 loss partials, one slot per 32 x 32 block [Bp / 32 x
  // lcols] (a wave writes the slot of its first block; the rest stay 0),
  // summed in a fixed order by loss_sum_kernel
  double* lpart;
  int lcols;                 // Dp / 32
  template <typename Cfg>
  __device__ void apply(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p) const {
    double part = 0.0;
#pragma unroll
    for (int j = 0; j < Cfg::kWN; ++j) {
      const int n = p.n0 + 32 * j + (p.lane & 31);
      const float bj = bias[n];
      const bool nv = n < D;
      const float sh = nv ? out_shift[n] : 0.0f, isc = nv ? 1.0f / out_scale[n] : 0.0f;
#pragma unroll
      for (int i = 0; i < Cfg::kWM; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          bf16_t ob[4];
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int m = p.m0 + 32 * i + acc_row(4 * q + v, p.lane);
            float d = 0.0f;
            if (nv && m < B) {
              float g;
              if (idx) {
                const int64_t row = idx[m];
                g = (row >= 0 && row < P) ? G[row * D + n] : __builtin_nanf("");
              } else {
                g = G[(int64_t)m * D + n];
              }
              const float t = (g - sh) * isc;
              d = acc[i][j][4 * q + v] + bj - t;
              part += (double)d * (double)d;
            }
            ob[v] = f32_to_bf16(coef * d);
            dO[(int64_t)m * ld + n] = ob[v];
          }
          store_t4(dOt, ldt, n, p.m0 + 32 * i + acc_row(4 * q, p.lane), ob);
        }
    }
    part = wave_sum_f64(part) * (0.5 * (double)coef);   // mean over B x D
    if (p.lane == 0) lpart[(int64_t)(p.m0 >> 5) * lcols + (p.n0 >> 5)] = part;
  }
};

// Inference output: grad = (acc + b3) * out_scale + out_shift -> [B x D] fp32.
struct EpiGradOut {
  const float* bias; const float* out_shift; const float* out_scale;
  int B, D;
  float* out;                // [B x ldo]
  int64_t ldo;
  template <typename Cfg>
  __device__ void apply(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p) const {
#pragma unroll
    for (int j = 0; j < Cfg::kWN; ++j) {
      const int n = p.n0 + 32 * j + (p.lane & 31);
      if (n >= D) continue;
      const float bj = bias[n], sc = out_scale[n], sh = out_shift[n];
#pragma unroll
      for (int i = 0; i < Cfg::kWM; ++i)
#pragma unroll
        for (int r = 0; r < 16; ++r) {
          const int m = p.m0 + 32 * i + acc_row(r, p.lane);
          if (m < B) out[(int64_t)m * ldo + n] = (acc[i][j][r] + bj) * sc + sh;
        }
    }
  }
};

// Leapfrog step fused into the output layer: g = surrogate gradient at z;
// middle steps kick twice by eps/2 and drift, and write the next normalised
// input into the x columns of the *other* [h2 | x] buffer (the one this GEMM
// reads is an operand of sibling workgroups still running; the next step's
// first layer reads its input there); the last step kicks by eps/2, keeps g
// in ghat (the next trajectory's first half kick) and writes the hi/lo
// position planes for the exact log joint.  Middle steps do not store g:
// nothing reads it.
struct EpiLeapfrog {
  const float* bias; const float* out_shift; const float* out_scale;
  const float* in_shift; const float* in_scale;
  int C, D, Dz;              // chains, dims, row stride of z/r/ghat
  float half, full;
  int last;                  // 1: final half kick only
  int eval_only;             // 1: store ghat only (no kick / drift)
  float* z; float* r; float* ghat;
  bf16_t* xcat; int xcat_ld; // next MLP input: the x columns of the next step's [h2 | x]
  bf16_t* zlp; int64_t zlp_plane; int zlp_ld;   // hi/lo planes (last step)

  // z and r of the lane's accumulator elements: element (i, j, rr) is row
  // m0 + 32 i + acc_row(rr, lane), column n0 + 32 j + (lane & 31); one
  // outside the C x D state reads element 0 and is never stored.  Loaded
  // before the K loop (prefetch) where they cost at most 64 registers.
  template <typename Cfg>
  struct Pre {
    float z[Cfg::kWM][Cfg::kWN][16];
    float r[Cfg::kWM][Cfg::kWN][16];
  };
  template <typename Cfg, typename = std::enable_if_t<(Cfg::kWM * Cfg::kWN <= 2)>>
  __device__ Pre<Cfg> prefetch(const EpiPos& p) const {
    return load<Cfg>(p);
  }
  template <typename Cfg>
  __device__ void apply(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p) const {
    dispatch<Cfg, false>(acc, p, nullptr);
  }
  template <typename Cfg>
  __device__ void apply(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p,
                        const Pre<Cfg>& q) const {
    dispatch<Cfg, true>(acc, p, &q);
  }
  template <typename Cfg>
  __device__ Pre<Cfg> load(const EpiPos& p) const {
    Pre<Cfg> q;
#pragma unroll
    for (int j = 0; j < Cfg::kWN; ++j) {
      const int n = p.n0 + 32 * j + (p.lane & 31);
#pragma unroll
      for (int i = 0; i < Cfg::kWM; ++i)
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
          const int m = p.m0 + 32 * i + acc_row(rr, p.lane);
          const int64_t idx = (m < C && n < D) ? (int64_t)m * Dz + n : 0;
          q.z[i][j][rr] = z[idx];
          q.r[i][j][rr] = r[idx];
        }
    }
    return q;
  }

  // MODE 0: store g only; 1: last step; 2: middle step.  CHECK: the wave's
  // block reaches past C rows or D columns (elements outside are skipped).
  // PRE: z and r come from q (else they are loaded here).
  template <typename Cfg, int MODE, bool CHECK, bool PRE>
  __device__ void step(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p,
                       const Pre<Cfg>* q) const {
#pragma unroll
    for (int j = 0; j < Cfg::kWN; ++j) {
      const int n = p.n0 + 32 * j + (p.lane & 31);   // < Dp: the per-column vectors' length
      const float bj = bias[n], sc = out_scale[n], sh = out_shift[n];
      const float ish = in_shift[n], isc = in_scale[n];
#pragma unroll
      for (int i = 0; i < Cfg::kWM; ++i)
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) {
          const int m = p.m0 + 32 * i + acc_row(rr, p.lane);
          if (CHECK && !(m < C && n < D)) continue;
          const int64_t idx = (int64_t)m * Dz + n;
          const float g = (acc[i][j][rr] + bj) * sc + sh;
          const float rold = PRE ? q->r[i][j][rr] : (MODE ? r[idx] : 0.0f);
          const float zold = PRE ? q->z[i][j][rr] : (MODE ? z[idx] : 0.0f);
          if (MODE == 0) {
            ghat[idx] = g;
          } else if (MODE == 1) {
            ghat[idx] = g;
            r[idx] = rold + half * g;
            const float zv = zold;
            const bf16_t hi = f32_to_bf16(zv);
            zlp[(int64_t)m * zlp_ld + n] = hi;
            zlp[(int64_t)m * zlp_ld + n + zlp_plane] = f32_to_bf16(zv - bf16_to_f32(hi));
          } else {
            const float rv = (rold + half * g) + half * g;
            const float zn = zold + full * rv;
            r[idx] = rv;
            z[idx] = zn;
            xcat[(int64_t)m * xcat_ld + n] = f32_to_bf16((zn - ish) * isc);
          }
        }
    }
  }

  template <typename Cfg, bool PRE>
  __device__ void dispatch(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p,
                           const Pre<Cfg>* q) const {
    // without the prefetch the checked form keeps the loads next to their
    // use (batching them all would double the registers of large tiles)
    const bool check = !PRE || p.m0 + 32 * Cfg::kWM > C || p.n0 + 32 * Cfg::kWN > D;
    if (eval_only) {
      if (check) step<Cfg, 0, true, PRE>(acc, p, q);
      else step<Cfg, 0, false, PRE>(acc, p, q);
    } else if (last) {
      if (check) step<Cfg, 1, true, PRE>(acc, p, q);
      else step<Cfg, 1, false, PRE>(acc, p, q);
    } else {
      if (check) step<Cfg, 2, true, PRE>(acc, p, q);
      else step<Cfg, 2, false, PRE>(acc, p, q);
    }
  }
};

// ---------------------------------------------------------------------------
// Elementwise kernels
// ---------------------------------------------------------------------------
// x0 = bf16((w - in_shift) * in_scale) for rows < B, cols < D; zeros elsewhere.
__global__ void norm_in_kernel(const float* __restrict__ W, int B, int D, int64_t ldw,
                               const float* __restrict__ shift, const float* __restrict__ scale,
                               bf16_t* __restrict__ x0, int Bp, int Dp, bf16_t* __restrict__ xcat,
                               int xcat_ld, double* __restrict__ zero_out = nullptr,
                               int n_zero = 0, const int64_t* __restrict__ idx = nullptr,
                               int64_t P = 0) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n_zero) zero_out[i] = 0.0;   // the batch's loss partial slots
  if (i >= (int64_t)Bp * Dp) return;
  const int b = (int)(i / Dp), d = (int)(i - (int64_t)b * Dp);
  float v = 0.0f;
  if (b < B && d < D) {
    // row b of the batch: W[idx[b]] of the P rows (NaN outside [0, P)), or W[b]
    float w;
    if (idx) {
      const int64_t r = idx[b];
      w = (r >= 0 && r < P) ? W[r * ldw + d] : __builtin_nanf("");
    } else {
      w = W[(int64_t)b * ldw + d];
    }
    v = (w - shift[d]) * scale[d];
  }
  const bf16_t xb = f32_to_bf16(v);
  x0[i] = xb;
  xcat[(int64_t)b * xcat_ld + d] = xb;
}

// dst [B x D] = src rows idx[0..B) ([P x D]); a row whose index lies outside
// [0, P) is written as NaN, so that the loss of its batch is NaN.
__global__ void gather_rows_kernel(const float* __restrict__ src, int64_t P, int D,
                                   const int64_t* __restrict__ idx, int B, float* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)B * D) return;
  const int b = (int)(i / D), d = (int)(i - (int64_t)b * D);
  const int64_t r = idx[b];
  dst[i] = (r >= 0 && r < P) ? src[r * D + d] : __builtin_nanf("");
}

// *loss = the sum of the n per-wave partials in a fixed order (256 strided
// sums, then a fixed tree), so the loss has the same bits on every run.
__global__ void __launch_bounds__(256) loss_sum_kernel(const double* __restrict__ part, int n,
                                                       double* __restrict__ loss) {
  __shared__ double red[256];
  const int t = threadIdx.x;
  double x = 0.0;
  for (int k = t; k < n; k += 256) x += part[k];
  red[t] = x;
  __syncthreads();
#pragma unroll
  for (int w = 128; w > 0; w >>= 1) {
    if (t < w) red[t] += red[t + w];
    __syncthreads();
  }
  if (t == 0) *loss = red[0];
}

// out [C x R] = in [R x C]^T for bf16, 64 x 64 tiles (R, C multiples of 64);
// blockIdx.z picks one of two such transposes (tiles outside a matrix idle).
struct TransposeJob {
  const bf16_t* in;
  int R, Cc;
  bf16_t* out;
};

__global__ void __launch_bounds__(256)
transpose_bf16_kernel(TransposeJob j0, TransposeJob j1) {
  const TransposeJob& jb = blockIdx.z ? j1 : j0;
  __shared__ bf16_t t[64][66];
  const int r0 = blockIdx.y * 64, c0 = blockIdx.x * 64;
  if (r0 >= jb.R || c0 >= jb.Cc) return;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int k = ty; k < 64; k += 4) t[k][tx] = jb.in[(int64_t)(r0 + k) * jb.Cc + c0 + tx];
  __syncthreads();
  for (int k = ty; k < 64; k += 4) jb.out[(int64_t)(c0 + k) * jb.R + r0 + tx] = t[tx][k];
}

// Row sums of a bf16 matrix [R x K] -> fp32 (bias gradients from transposed
// activation gradients).  One wave per row, 16-byte loads (K is a multiple
// of 128: the padded batch).
__global__ void rowsum_bf16_kernel(const bf16_t* __restrict__ in, int R, int K, float* __restrict__ out) {
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= R) return;
  float s = 0.0f;
  for (int k = 8 * lane; k < K; k += 512) {
    const uint4 v = *(const uint4*)(in + (int64_t)row * K + k);
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int e = 0; e < 4; ++e)
      s += __builtin_bit_cast(float, w[e] << 16) + __builtin_bit_cast(float, w[e] & 0xffff0000u);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (lane == 0) out[row] = s;
}

// Least-squares Gram products of the output layer's linear skip over the
// rows idx[0..n) of (W, G): x = (w - in_shift) * in_scale and t = (g -
// out_shift) / out_scale in fp32, out = sum_rows x^T x (blockIdx.z = 0) or
// x^T t (z = 1), [D x D] FP64.  A 64 x 64 output block per workgroup, a
// thread owning 4 x 4 entries, rows staged 32 at a time through LDS.  An
// index outside [0, P) contributes NaN.
__global__ void __launch_bounds__(256)
skip_gram_kernel(const float* __restrict__ W, const float* __restrict__ G, int64_t P, int D,
                 const int64_t* __restrict__ idx, int n, const float* __restrict__ nrm, int Dp,
                 double* __restrict__ xtx, double* __restrict__ xtt) {
  __shared__ float la[32][64], lb[32][64];
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64, z = blockIdx.z;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const float* in_shift = nrm;
  const float* in_scale = nrm + Dp;
  const float* out_shift = nrm + 2 * Dp;
  const float* out_scale = nrm + 3 * Dp;
  double acc[4][4] = {};
  for (int r0 = 0; r0 < n; r0 += 32) {
    for (int q = threadIdx.x; q < 32 * 64; q += 256) {
      const int rr = q >> 6, c = q & 63, row = r0 + rr;
      float a = 0.0f, b = 0.0f;
      if (row < n && (idx[row] < 0 || idx[row] >= P)) {
        a = b = __builtin_nanf("");
      } else if (row < n) {
        const int64_t src = idx[row] * (int64_t)D;
        if (i0 + c < D) a = (W[src + i0 + c] - in_shift[i0 + c]) * in_scale[i0 + c];
        const int j = j0 + c;
        if (j < D) b = z == 0 ? (W[src + j] - in_shift[j]) * in_scale[j]
                              : (G[src + j] - out_shift[j]) / out_scale[j];
      }
      la[rr][c] = a;
      lb[rr][c] = b;
    }
    __syncthreads();
#pragma unroll 4
    for (int rr = 0; rr < 32; ++rr) {
      float av[4], bv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        av[u] = la[rr][ty * 4 + u];
        bv[u] = lb[rr][tx * 4 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] += (double)av[u] * (double)bv[v];
    }
    __syncthreads();
  }
  double* out = z == 0 ? xtx : xtt;
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int i = i0 + ty * 4 + u, j = j0 + tx * 4 + v;
      if (i < D && j < D) out[(int64_t)i * D + j] = acc[u][v];
    }
}

// out [n x D] = shift + E [n x D] R [D x D] (fp32; 64 x 64 output blocks, a
// thread owning 4 x 4 entries, K staged 32 at a time through LDS).
__global__ void __launch_bounds__(256)
affine_rows_kernel(const float* __restrict__ E, int64_t n, int D, const float* __restrict__ R,
                   const float* __restrict__ shift, float* __restrict__ out) {
  __shared__ float le[64][33], lr[32][64];
  const int64_t r0 = (int64_t)blockIdx.x * 64;
  const int c0 = blockIdx.y * 64;
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < D; k0 += 32) {
    for (int q = threadIdx.x; q < 64 * 32; q += 256) {
      const int rr = q >> 5, kk = q & 31;
      const int64_t row = r0 + rr;
      le[rr][kk] = (row < n && k0 + kk < D) ? E[row * D + k0 + kk] : 0.0f;
      const int k2 = q >> 6, cc = q & 63;
      lr[k2][cc] = (k0 + k2 < D && c0 + cc < D) ? R[(int64_t)(k0 + k2) * D + c0 + cc] : 0.0f;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < 32; ++kk) {
      float ev[4], rv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ev[u] = le[ty * 4 + u][kk];
        rv[u] = lr[kk][tx * 4 + u];
      }
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v) acc[u][v] = fmaf(ev[u], rv[v], acc[u][v]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 4; ++u)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int64_t row = r0 + ty * 4 + u;
      const int c = c0 + tx * 4 + v;
      if (row < n && c < D) out[row * D + c] = shift[c] + acc[u][v];
    }
}

// Adam (Kingma & Ba 2015) over the padded parameter vector.  With w1b..w3b
// given, the updated weights of the three layers are also written as the
// bf16 copies the GEMMs read (elements [oW_l, oW_l + n_l) of P).
struct AdamBf16Out {
  bf16_t* w[3];
  int64_t off[3], n[3];
};

__global__ void adam_kernel(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                            float* __restrict__ V, int64_t n, float lr, float b1, float b2,
                            float eps, float c1, float c2, AdamBf16Out wb) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = G[i];
  const float m = b1 * M[i] + (1.0f - b1) * g;
  const float v = b2 * V[i] + (1.0f - b2) * g * g;
  M[i] = m;
  V[i] = v;
  const float p = P[i] - lr * (m * c1) / (sqrtf(v * c2) + eps);
  P[i] = p;
#pragma unroll
  for (int l = 0; l < 3; ++l)
    if (wb.w[l] && i >= wb.off[l] && i < wb.off[l] + wb.n[l]) wb.w[l][i - wb.off[l]] = f32_to_bf16(p);
}

// fp32 [R x C] -> bf16 [R x C]
__global__ void to_bf16_kernel(const float* __restrict__ in, int64_t n, bf16_t* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f32_to_bf16(in[i]);
}

// Pack / unpack between the flat unpadded layout and the padded one.
struct PackLayout {
  int D, H, Dp, Hp;
  int64_t oW1, ob1, oW2, ob2, oW3, ob3;
};

__device__ inline int64_t padded_index(const PackLayout& L, int64_t f) {
  const int64_t sW1 = (int64_t)L.H * L.D, sW2 = (int64_t)L.H * L.H;
  const int64_t sW3 = (int64_t)L.D * (L.H + L.D);
  if (f < sW1) return L.oW1 + (f / L.D) * L.Dp + f % L.D;
  f -= sW1;
  if (f < L.H) return L.ob1 + f;
  f -= L.H;
  if (f < sW2) return L.oW2 + (f / L.H) * L.Hp + f % L.H;
  f -= sW2;
  if (f < L.H) return L.ob2 + f;
  f -= L.H;
  if (f < sW3) {   // W3 [D, H + D]: hidden columns then skip columns
    const int64_t row = f / (L.H + L.D), col = f % (L.H + L.D);
    return L.oW3 + row * (L.Hp + L.Dp) + (col < L.H ? col : L.Hp + (col - L.H));
  }
  f -= sW3;
  return L.ob3 + f;
}

__global__ void pack_kernel(PackLayout L, int64_t n, const float* __restrict__ flat,
                            float* __restrict__ padded, int to_padded) {
  const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= n) return;
  const int64_t pi = padded_index(L, f);
  if (to_padded) padded[pi] = flat[f];
  else ((float*)flat)[f] = padded[pi];
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static inline unsigned blocks_for(int64_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

static PackLayout layout_of(const Surrogate* s) {
  PackLayout L;
  L.D = s->D; L.H = s->H; L.Dp = s->Dp; L.Hp = s->Hp;
  L.oW1 = s->oW1; L.ob1 = s->ob1; L.oW2 = s->oW2; L.ob2 = s->ob2; L.oW3 = s->oW3; L.ob3 = s->ob3;
  return L;
}

template <typename Cfg, typename Epi>
static int mlp_gemm_cfg(edhmc_ctx* ctx, const void* A, int64_t lda_elems, const void* B,
                        int64_t ldb_elems, int M, int N, int K, Epi e, hipStream_t st) {
  GemmShape g = {};
  g.A = (const char*)A;
  g.lda_b = lda_elems * 2;
  g.a_stage_b = Cfg::KS;
  g.a_plane_b = 0;
  g.B = (const char*)B;
  g.ldb_b = ldb_elems * 2;
  g.b_stage_b = Cfg::KS;
  g.m_tiles = M / Cfg::BM;
  g.n_tiles = N / Cfg::BN;
  g.k_stages = K * 2 / Cfg::KS;
  g.ksplit = 1;
  g.stages_per_split = g.k_stages;
  g.swz = 0;
  if (M % Cfg::BM || N % Cfg::BN || (K * 2) % Cfg::KS)
    return set_err(ctx, EDHMC_EUNSUPPORTED, "MLP GEMM %d x %d x %d does not fit the tiles", M, N, K);
  if (!gemm_offsets_fit<Cfg, 1>(g)) return set_err(ctx, EDHMC_EUNSUPPORTED, "MLP tile too large");
  ProfScope ps(ctx, PK_MLP, st);
  hipLaunchKernelGGL((gemm_nt_kernel<bf16_t, Cfg, 1, Epi>), dim3(grid_size(g)),
                     dim3(Cfg::kThreads), 0, st, g, e);
  return check_launch(ctx, "mlp_gemm");
}

// The tile shape by the number of 128 x 128 output tiles (M and N are
// multiples of 128).  Measured on the 4096-chain layers (MI355X, us per
// launch, 64 x 64 / chosen): 4096 x 1024 x 512  13.8 / 9.8,
// 4096 x 1024 x 1024  19.2 / 14.4, 4096 x 512 x 1536  14.1 / 12.8.
template <typename Epi>
static int mlp_gemm(edhmc_ctx* ctx, const void* A, int64_t lda_elems, const void* B,
                    int64_t ldb_elems, int M, int N, int K, Epi e, hipStream_t st) {
  const int tiles = (M / 128) * (N / 128);
  if (tiles >= EDHMC_MLP_THRESH)
    return mlp_gemm_cfg<CfgM>(ctx, A, lda_elems, B, ldb_elems, M, N, K, e, st);
  if (tiles >= 256) return mlp_gemm_cfg<CfgM8>(ctx, A, lda_elems, B, ldb_elems, M, N, K, e, st);
  if (tiles >= 128) return mlp_gemm_cfg<CfgM8N>(ctx, A, lda_elems, B, ldb_elems, M, N, K, e, st);
  return mlp_gemm_cfg<CfgMS>(ctx, A, lda_elems, B, ldb_elems, M, N, K, e, st);
}

// One launch for one or two bf16 transposes (in [R x Cc] -> out [Cc x R]).
static int transpose(edhmc_ctx* ctx, TransposeJob a, const TransposeJob* b, hipStream_t st) {
  const TransposeJob bb = b ? *b : TransposeJob{a.in, 0, 0, a.out};
  const int gx = std::max(a.Cc, bb.Cc) / 64, gy = std::max(a.R, bb.R) / 64;
  hipLaunchKernelGGL(transpose_bf16_kernel, dim3(gx, gy, b ? 2 : 1), dim3(256), 0, st, a, bb);
  return check_launch(ctx, "transpose");
}

// The transposed bf16 weights the backward GEMMs read, from the forward copies.
static int transpose_weights(edhmc_ctx* ctx, Surrogate* s, hipStream_t st) {
  const TransposeJob w2{(const bf16_t*)s->W2b.p, s->Hp, s->Hp, (bf16_t*)s->W2t.p};
  const TransposeJob w3{(const bf16_t*)s->W3b.p, s->Dp, s->Kc, (bf16_t*)s->W3t.p};
  return transpose(ctx, w2, &w3, st);
}

// Refresh the bf16 weight copies (forward and transposed) from the masters.
static int refresh_weights(edhmc_ctx* ctx, Surrogate* s, hipStream_t st) {
  const float* P = (const float*)s->P.p;
  auto cvt = [&](int64_t off, int64_t n, DevBuf& out) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3(blocks_for(n)), dim3(256), 0, st, P + off, n,
                       (bf16_t*)out.p);
    return check_launch(ctx, "to_bf16");
  };
  EDHMC_TRY(cvt(s->oW1, (int64_t)s->Hp * s->Dp, s->W1b));
  EDHMC_TRY(cvt(s->oW2, (int64_t)s->Hp * s->Hp, s->W2b));
  EDHMC_TRY(cvt(s->oW3, (int64_t)s->Dp * s->Kc, s->W3b));
  EDHMC_TRY(transpose_weights(ctx, s, st));
  s->ghat_valid = false;
  return EDHMC_OK;
}

// Forward through the two hidden layers for the batch in s->x0 (Bp rows),
// or in xcat's x columns; keep_a: store pre-activations for the backward
// pass; h2 goes into the first Hp columns of xcat ([Bp x Kc]; s->h2 unless
// given).
static int forward_hidden(edhmc_ctx* ctx, Surrogate* s, int Bp, bool keep_a, hipStream_t st,
                          bf16_t* xcat = nullptr, bool x_from_xcat = false,
                          bf16_t* h1t = nullptr, bf16_t* h2t = nullptr) {
  const float* P = (const float*)s->P.p;
  EpiBiasGelu e1;
  e1.bias = P + s->ob1;
  e1.a_out = keep_a ? (float*)s->a1.p : nullptr;
  e1.h_out = (bf16_t*)s->h1.p;
  e1.ld = s->Hp;
  e1.ldh = s->Hp;
  e1.ht_out = h1t;
  e1.ldt = Bp;
  bf16_t* xc = xcat ? xcat : (bf16_t*)s->h2.p;
  if (x_from_xcat)
    EDHMC_TRY(mlp_gemm(ctx, xc + s->Hp, s->Kc, s->W1b.p, s->Dp, Bp, s->Hp, s->Dp, e1, st));
  else
    EDHMC_TRY(mlp_gemm(ctx, s->x0.p, s->Dp, s->W1b.p, s->Dp, Bp, s->Hp, s->Dp, e1, st));
  EpiBiasGelu e2;
  e2.bias = P + s->ob2;
  e2.a_out = keep_a ? (float*)s->a2.p : nullptr;
  e2.h_out = xc;
  e2.ld = s->Hp;
  e2.ldh = s->Kc;
  e2.ht_out = h2t;
  e2.ldt = Bp;
  return mlp_gemm(ctx, s->h1.p, s->Hp, s->W2b.p, s->Hp, Bp, s->Hp, s->Hp, e2, st);
}

static int check_sur(edhmc_ctx* ctx) {
  if (!ctx->sur) return set_err(ctx, EDHMC_ESTATE, "no surrogate: call edhmc_sur_create");
  return EDHMC_OK;
}

static int batch_ok(edhmc_ctx* ctx, Surrogate* s, int B) {
  if (B <= 0 || B > s->Bmax)
    return set_err(ctx, EDHMC_EINVAL, "batch %d outside 1..%d (max_batch)", B, s->Bmax);
  if (!s->norm_set) return set_err(ctx, EDHMC_ESTATE, "normalisation not set");
  return EDHMC_OK;
}

// Forward and backward over a batch: the parameter gradient into s->Gr, the
// mean loss into *loss (FP64, per-wave partials summed in a fixed order).  The
// batch is rows idx[0..B) of W and G ([P x D]) when idx is given, else their
// first B rows.  The forward and activation-gradient epilogues also write
// the transposed activations the weight-gradient GEMMs read (their K is the
// batch); h2t holds [h2 | x0]^T, its last Dp rows being x0^T.
static int loss_grad(edhmc_ctx* ctx, Surrogate* s, const float* W, const float* G, int B,
                     double* loss, hipStream_t st, const int64_t* idx = nullptr, int64_t P_rows = 0,
                     bool backward = true) {
  const int Bp = (int)round_up(B, 128);
  const float* P = (const float*)s->P.p;
  const float* nrm = (const float*)s->nrm.p;
  bf16_t* h2t = (bf16_t*)s->h2t.p;
  bf16_t* x0t = h2t + (int64_t)s->Hp * Bp;
  double* lpart = (double*)s->lpart.p;
  const int lcols = s->Dp / 32, n_slots = Bp / 32 * lcols;
  hipLaunchKernelGGL(norm_in_kernel, dim3(blocks_for((int64_t)Bp * s->Dp)), dim3(256), 0, st, W,
                     B, s->D, (int64_t)s->D, nrm, nrm + s->Dp, (bf16_t*)s->x0.p, Bp, s->Dp,
                     (bf16_t*)s->h2.p + s->Hp, s->Kc, lpart, n_slots, idx, P_rows);
  EDHMC_TRY(check_launch(ctx, "norm_in"));
  EDHMC_TRY(transpose(ctx, TransposeJob{(const bf16_t*)s->x0.p, Bp, s->Dp, x0t}, nullptr, st));
  EDHMC_TRY(forward_hidden(ctx, s, Bp, true, st, nullptr, false, (bf16_t*)s->h1t.p, h2t));
  EpiLoss el;
  el.bias = P + s->ob3;
  el.G = G;
  el.idx = idx;
  el.P = P_rows;
  el.out_shift = nrm + 2 * s->Dp;
  el.out_scale = nrm + 3 * s->Dp;
  el.B = B;
  el.D = s->D;
  el.ld = s->Dp;
  el.coef = 2.0f / ((float)B * (float)s->D);
  el.dO = (bf16_t*)s->dO.p;
  el.dOt = (bf16_t*)s->dOt.p;
  el.ldt = Bp;
  el.lpart = lpart;
  el.lcols = lcols;
  EDHMC_TRY(mlp_gemm(ctx, s->h2.p, s->Kc, s->W3b.p, s->Kc, Bp, s->Dp, s->Kc, el, st));
  hipLaunchKernelGGL(loss_sum_kernel, dim3(1), dim3(256), 0, st, (const double*)lpart, n_slots,
                     loss);
  EDHMC_TRY(check_launch(ctx, "loss_sum"));
  if (!backward) return EDHMC_OK;   // the loss alone
  float* Gr = (float*)s->Gr.p;
  // layer 3 weights: dW3 = dO^T [h2 | x0], db3 = column sums of dO
  EpiStoreF32 ew;
  ew.split_b = 0;
  ew.out = Gr + s->oW3;
  ew.ldc = s->Kc;
  EDHMC_TRY(mlp_gemm(ctx, s->dOt.p, Bp, h2t, Bp, s->Dp, s->Kc, Bp, ew, st));
  hipLaunchKernelGGL(rowsum_bf16_kernel, dim3(blocks_for(s->Dp, 4)), dim3(256), 0, st,
                     (const bf16_t*)s->dOt.p, s->Dp, Bp, Gr + s->ob3);
  EDHMC_TRY(check_launch(ctx, "rowsum"));
  // da2 = (dO W3) * GELU'(a2)
  EpiGeluBack eb2;
  eb2.a_in = (const float*)s->a2.p;
  eb2.da_out = (bf16_t*)s->da2.p;
  eb2.ld = s->Hp;
  eb2.dat_out = (bf16_t*)s->da2t.p;
  eb2.ldt = Bp;
  EDHMC_TRY(mlp_gemm(ctx, s->dO.p, s->Dp, s->W3t.p, s->Dp, Bp, s->Hp, s->Dp, eb2, st));
  ew.out = Gr + s->oW2;
  ew.ldc = s->Hp;
  EDHMC_TRY(mlp_gemm(ctx, s->da2t.p, Bp, s->h1t.p, Bp, s->Hp, s->Hp, Bp, ew, st));
  hipLaunchKernelGGL(rowsum_bf16_kernel, dim3(blocks_for(s->Hp, 4)), dim3(256), 0, st,
                     (const bf16_t*)s->da2t.p, s->Hp, Bp, Gr + s->ob2);
  EDHMC_TRY(check_launch(ctx, "rowsum"));
  // da1 = (da2 W2) * GELU'(a1)
  EpiGeluBack eb1;
  eb1.a_in = (const float*)s->a1.p;
  eb1.da_out = (bf16_t*)s->da1.p;
  eb1.ld = s->Hp;
  eb1.dat_out = (bf16_t*)s->da1t.p;
  eb1.ldt = Bp;
  EDHMC_TRY(mlp_gemm(ctx, s->da2.p, s->Hp, s->W2t.p, s->Hp, Bp, s->Hp, s->Hp, eb1, st));
  ew.out = Gr + s->oW1;
  ew.ldc = s->Dp;
  EDHMC_TRY(mlp_gemm(ctx, s->da1t.p, Bp, x0t, Bp, s->Hp, s->Dp, Bp, ew, st));
  hipLaunchKernelGGL(rowsum_bf16_kernel, dim3(blocks_for(s->Hp, 4)), dim3(256), 0, st,
                     (const bf16_t*)s->da1t.p, s->Hp, Bp, Gr + s->ob1);
  return check_launch(ctx, "rowsum");
}

// Surrogate gradient at the chains' current positions into s->ghat.
static int eval_ghat(edhmc_ctx* ctx, Surrogate* s, hipStream_t st);

// Per-chain begin of a surrogate trajectory.
struct SurChainArgs {
  int C, D, Dz;
  float* z; float* r; float* z0; float* ghat; float* ghat0;
  double* K0;
  const float* r_fed;
  uint64_t seed, counter;
  float half, full;
  const float* in_shift; const float* in_scale;
  bf16_t* xcat; int xcat_ld;   // first MLP input: the x columns of step 1's [h2 | x]
};

__global__ void __launch_bounds__(256) sur_begin_kernel(SurChainArgs a) {
  __shared__ double red[4];
  const int c = blockIdx.x;
  double k = 0.0;
  const int nb = (a.D + 3) / 4;
  for (int j = threadIdx.x; j < nb; j += 256) {
    f32x4_t rv;
    if (a.r_fed) {
#pragma unroll
      for (int e = 0; e < 4; ++e) {
        const int d = 4 * j + e;
        rv[e] = d < a.D ? a.r_fed[(int64_t)c * a.D + d] : 0.0f;
      }
    } else {
      rv = momentum4(a.seed, a.counter, (uint32_t)j, (uint32_t)c);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int d = 4 * j + e;
      if (d >= a.D) continue;
      const int64_t idx = (int64_t)c * a.Dz + d;
      float r = rv[e];
      k += (double)r * (double)r;
      const float zv = a.z[idx], gv = a.ghat[idx];
      a.z0[idx] = zv;
      a.ghat0[idx] = gv;
      r = r + a.half * gv;
      const float zn = zv + a.full * r;
      a.z[idx] = zn;
      a.r[idx] = r;
      a.xcat[(int64_t)c * a.xcat_ld + d] = f32_to_bf16((zn - a.in_shift[d]) * a.in_scale[d]);
    }
  }
  k = block_sum_f64<256>(k, red);
  if (threadIdx.x == 0) a.K0[c] = 0.5 * k;
}

struct SurAcceptArgs {
  int C, D, Dz;
  float* z; float* r; float* z0; float* ghat; float* ghat0;
  double* logp; const double* K0; const double* ll;
  float lik_weight;          // multiplies ll, as in the carried start-point log joint
  const float* ploc; const float* pscale;
  const double* u_fed; uint64_t seed, counter;
  float* sample_out; int32_t* accepted_out; int32_t* n_accept; int32_t* ndiv;
  double* ratio_out;
  double* mom_mean; double* mom_m2; int64_t mom_n;
};

__global__ void __launch_bounds__(256) sur_accept_kernel(SurAcceptArgs a) {
  __shared__ double red[4];
  __shared__ int acc_flag;
  const int c = blockIdx.x;
  double kl = 0.0, plp = 0.0;
  int bad = 0;
  for (int d = threadIdx.x; d < a.D; d += 256) {
    const int64_t idx = (int64_t)c * a.Dz + d;
    const float zv = a.z[idx], rv = a.r[idx];
    kl += (double)rv * (double)rv;
    const double s = a.pscale[d];
    const double t = ((double)zv - (double)a.ploc[d]) / s;
    plp += -0.5 * t * t - log(s) - 0.5 * kLog2Pi;
    if (!isfinite(zv) || !isfinite(rv)) bad = 1;
  }
  kl = block_sum_f64<256>(kl, red);
  plp = block_sum_f64<256>(plp, red);
  const double nbad = block_sum_f64<256>((double)bad, red);
  if (threadIdx.x == 0) {
    const double logpL = (double)a.lik_weight * a.ll[c] + plp;
    const double ratio = a.K0[c] - 0.5 * kl + logpL - a.logp[c];
    const double u = a.u_fed ? a.u_fed[c] : accept_uniform(a.seed, a.counter, (uint32_t)c);
    const int acc = (log(u) < ratio) && nbad == 0.0;
    acc_flag = acc;
    if (acc) a.logp[c] = logpL;
    if (a.accepted_out) a.accepted_out[c] = acc;
    if (a.n_accept) a.n_accept[c] += acc;
    if (a.ndiv && (nbad != 0.0 || !(ratio == ratio))) a.ndiv[c] += 1;
    a.ratio_out[c] = ratio;
  }
  __syncthreads();
  const int acc = acc_flag;
  for (int d = threadIdx.x; d < a.D; d += 256) {
    const int64_t idx = (int64_t)c * a.Dz + d;
    float zv;
    if (acc) {
      zv = a.z[idx];
    } else {
      zv = a.z0[idx];
      a.z[idx] = zv;
      a.ghat[idx] = a.ghat0[idx];
    }
    if (a.sample_out) a.sample_out[(int64_t)c * a.D + d] = zv;
    if (a.mom_mean) {
      const double x = zv;
      const double mo = a.mom_mean[idx];
      const double mn = mo + (x - mo) / (double)a.mom_n;
      a.mom_mean[idx] = mn;
      a.mom_m2[idx] += (x - mo) * (x - mn);
    }
  }
}

// The [h2 | x] buffer of leapfrog step k (k & 1 selects one of the two).
static bf16_t* xbuf(Surrogate* s, int k) {
  return (bf16_t*)((k & 1) ? s->h2x.p : s->h2.p);
}

// Epilogue of step k's output layer; a middle step writes the next input
// into step k + 1's buffer.
static EpiLeapfrog lf_epi(edhmc_ctx* ctx, Surrogate* s, float half, float full, int last,
                          int eval_only, int k) {
  const float* P = (const float*)s->P.p;
  const float* nrm = (const float*)s->nrm.p;
  EpiLeapfrog e;
  e.bias = P + s->ob3;
  e.in_shift = nrm;
  e.in_scale = nrm + s->Dp;
  e.out_shift = nrm + 2 * s->Dp;
  e.out_scale = nrm + 3 * s->Dp;
  e.C = ctx->C;
  e.D = ctx->D;
  e.Dz = ctx->D_pad;
  e.half = half;
  e.full = full;
  e.last = last;
  e.eval_only = eval_only;
  e.z = (float*)ctx->z.p;
  e.r = (float*)ctx->r.p;
  e.ghat = (float*)s->ghat.p;
  e.xcat = xbuf(s, k + 1) + s->Hp;
  e.xcat_ld = s->Kc;
  e.zlp = (bf16_t*)ctx->z_lp.p;
  e.zlp_plane = (int64_t)ctx->C_pad * ctx->D_pad;
  e.zlp_ld = ctx->D_pad;
  return e;
}

static int sur_state_buffers(edhmc_ctx* ctx, Surrogate* s) {
  if (!ctx->bound || !ctx->state_valid)
    return set_err(ctx, EDHMC_ESTATE, "no chain state: call edhmc_set_state first");
  if (ctx->cdt != EDHMC_BF16)
    return set_err(ctx, EDHMC_EUNSUPPORTED, "surrogate transitions need a bf16-bound model");
  if (ctx->D != s->D) return set_err(ctx, EDHMC_EINVAL, "surrogate D=%d, model D=%d", s->D, ctx->D);
  if (ctx->C > s->Bmax)
    return set_err(ctx, EDHMC_EINVAL, "%d chains exceed the surrogate batch %d", ctx->C, s->Bmax);
  const size_t bytes = (size_t)ctx->C * ctx->D_pad * 4;
  if (s->ghat.bytes != bytes) {
    EDHMC_TRY(dev_alloc(ctx, s->ghat, bytes));
    EDHMC_TRY(dev_alloc(ctx, s->ghat0, bytes));
    s->ghat_valid = false;
  }
  return EDHMC_OK;
}

static int eval_ghat(edhmc_ctx* ctx, Surrogate* s, hipStream_t st) {
  const int Bp = (int)round_up(ctx->C, 128);
  const float* nrm = (const float*)s->nrm.p;
  hipLaunchKernelGGL(norm_in_kernel, dim3(blocks_for((int64_t)Bp * s->Dp)), dim3(256), 0, st,
                     (const float*)ctx->z.p, ctx->C, s->D, (int64_t)ctx->D_pad, nrm, nrm + s->Dp,
                     (bf16_t*)s->x0.p, Bp, s->Dp,
                     (bf16_t*)s->h2.p + s->Hp, s->Kc);
  EDHMC_TRY(check_launch(ctx, "norm_in"));
  EDHMC_TRY(forward_hidden(ctx, s, Bp, false, st, nullptr, true));
  EpiLeapfrog e = lf_epi(ctx, s, 0.0f, 0.0f, 0, 1, 0);
  EDHMC_TRY(mlp_gemm(ctx, s->h2.p, s->Kc, s->W3b.p, s->Kc, Bp, s->Dp, s->Kc, e, st));
  s->ghat_valid = true;
  return EDHMC_OK;
}

static int sur_transition(edhmc_ctx* ctx, float eps, int L, const float* r0, const double* u,
                          uint64_t seed, uint64_t counter, int32_t* accepted_out,
                          float* sample_out, int32_t* n_accept, hipStream_t st) {
  Surrogate* s = ctx->sur;
  EDHMC_TRY(sur_state_buffers(ctx, s));
  if (L < 1) return set_err(ctx, EDHMC_EINVAL, "n_steps must be >= 1, got %d", L);
  if (!std::isfinite(eps)) return set_err(ctx, EDHMC_EINVAL, "step_size must be finite");
  if (!s->norm_set) return set_err(ctx, EDHMC_ESTATE, "normalisation not set");
  if (!s->ghat_valid) EDHMC_TRY(eval_ghat(ctx, s, st));
  // after plain transitions the carried log joint came from a gradient pass:
  // re-evaluate it with the forward pass the accept test below uses
  if (ctx->g_valid) EDHMC_TRY(glm_refresh_logp(ctx, st));
  const float half = (float)(0.5 * (double)eps), full = eps;
  const float* nrm = (const float*)s->nrm.p;
  SurChainArgs a;
  a.C = ctx->C; a.D = ctx->D; a.Dz = ctx->D_pad;
  a.z = (float*)ctx->z.p; a.r = (float*)ctx->r.p; a.z0 = (float*)ctx->z0.p;
  a.ghat = (float*)s->ghat.p; a.ghat0 = (float*)s->ghat0.p;
  a.K0 = (double*)ctx->K0.p;
  a.r_fed = r0; a.seed = seed; a.counter = counter;
  a.half = half; a.full = full;
  a.in_shift = nrm; a.in_scale = nrm + s->Dp;
  a.xcat = xbuf(s, 1) + s->Hp;
  a.xcat_ld = s->Kc;
  {
    ProfScope ps(ctx, PK_CHAIN, st);
    hipLaunchKernelGGL(sur_begin_kernel, dim3(ctx->C), dim3(256), 0, st, a);
  }
  EDHMC_TRY(check_launch(ctx, "sur_begin"));
  const int Bp = (int)round_up(ctx->C, 128);
  for (int k = 1; k <= L; ++k) {
    bf16_t* xk = xbuf(s, k);
    EDHMC_TRY(forward_hidden(ctx, s, Bp, false, st, xk, true));
    EpiLeapfrog e = lf_epi(ctx, s, half, full, k == L ? 1 : 0, 0, k);
    EDHMC_TRY(mlp_gemm(ctx, xk, s->Kc, s->W3b.p, s->Kc, Bp, s->Dp, s->Kc, e, st));
  }
  // exact log joint at the end point (z_lp planes written by the last step)
  EDHMC_TRY(glm_forward_ll(ctx, st));
  SurAcceptArgs b;
  b.C = ctx->C; b.D = ctx->D; b.Dz = ctx->D_pad;
  b.z = (float*)ctx->z.p; b.r = (float*)ctx->r.p; b.z0 = (float*)ctx->z0.p;
  b.ghat = (float*)s->ghat.p; b.ghat0 = (float*)s->ghat0.p;
  b.logp = (double*)ctx->logp.p; b.K0 = (const double*)ctx->K0.p;
  b.ll = (const double*)ctx->ll.p;
  b.lik_weight = ctx->lik_weight;
  b.ploc = (const float*)ctx->ploc.p; b.pscale = (const float*)ctx->pscale.p;
  b.u_fed = u; b.seed = seed; b.counter = counter;
  b.sample_out = sample_out; b.accepted_out = accepted_out; b.n_accept = n_accept;
  b.ndiv = (int32_t*)ctx->ndiv.p;
  b.ratio_out = (double*)ctx->ratio.p;
  b.mom_mean = nullptr; b.mom_m2 = nullptr; b.mom_n = 0;
  if (ctx->mom_on) {
    ctx->mom_count += 1;
    b.mom_mean = (double*)ctx->mom_mean.p;
    b.mom_m2 = (double*)ctx->mom_m2.p;
    b.mom_n = ctx->mom_count;
  }
  {
    ProfScope ps(ctx, PK_CHAIN, st);
    hipLaunchKernelGGL(sur_accept_kernel, dim3(ctx->C), dim3(256), 0, st, b);
  }
  EDHMC_TRY(check_launch(ctx, "sur_accept"));
  // the exact gradient at the new state is not known any more
  ctx->g_valid = false;
  return EDHMC_OK;
}

}  // namespace edhmc

using namespace edhmc;

extern "C" {

int edhmc_sur_create(edhmc_ctx* ctx, int32_t D, int32_t H, int32_t max_batch) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (D <= 0 || H <= 0 || max_batch <= 0 || D > 16384 || H > 16384)
    return set_err(ctx, EDHMC_EINVAL, "bad surrogate sizes D=%d H=%d batch=%d", D, H, max_batch);
  if (ctx->sur) surrogate_destroy(ctx->sur);
  Surrogate* s = new (std::nothrow) Surrogate();
  if (!s) return set_err(ctx, EDHMC_ENOMEM, "host allocation");
  ctx->sur = s;
  s->D = D; s->H = H; s->Bmax = max_batch;
  s->Dp = (int)round_up(D, 128);
  s->Hp = (int)round_up(H, 128);
  s->Bp = (int)round_up(max_batch, 128);
  s->Kc = s->Hp + s->Dp;
  s->n_params = (int64_t)H * D + H + (int64_t)H * H + H + (int64_t)D * (H + D) + D;
  s->oW1 = 0;
  s->ob1 = s->oW1 + (int64_t)s->Hp * s->Dp;
  s->oW2 = s->ob1 + s->Hp;
  s->ob2 = s->oW2 + (int64_t)s->Hp * s->Hp;
  s->oW3 = s->ob2 + s->Hp;
  s->ob3 = s->oW3 + (int64_t)s->Dp * s->Kc;
  s->n_pad = s->ob3 + s->Dp;
  const size_t pb = (size_t)s->n_pad * 4;
  const size_t hd = (size_t)s->Hp * s->Dp * 2, hh = (size_t)s->Hp * s->Hp * 2;
  const size_t dk = (size_t)s->Dp * s->Kc * 2, bk = (size_t)s->Bp * s->Kc * 2;
  const size_t bd = (size_t)s->Bp * s->Dp, bh = (size_t)s->Bp * s->Hp;
  // every buffer a view into one zeroed allocation (one hipMalloc and one
  // hipMemset instead of one of each per buffer)
  const std::pair<DevBuf*, size_t> parts[] = {
      {&s->P, pb}, {&s->Gr, pb}, {&s->Am, pb}, {&s->Av, pb},
      {&s->W1b, hd}, {&s->W2b, hh}, {&s->W3b, dk}, {&s->W2t, hh}, {&s->W3t, dk},
      {&s->nrm, (size_t)4 * s->Dp * 4},
      {&s->x0, bd * 2}, {&s->a1, bh * 4}, {&s->h1, bh * 2}, {&s->a2, bh * 4},
      {&s->h2, bk}, {&s->h2x, bk}, {&s->dO, bd * 2}, {&s->da2, bh * 2}, {&s->da1, bh * 2},
      {&s->h1t, bh * 2}, {&s->h2t, bk}, {&s->dOt, bd * 2}, {&s->da2t, bh * 2},
      {&s->da1t, bh * 2}, {&s->loss, 8}, {&s->lpart, (size_t)(s->Bp / 32) * (s->Dp / 32) * 8}};
  size_t total = 0;
  for (const auto& pt : parts) total += (size_t)round_up((int64_t)pt.second, 256);
  EDHMC_TRY(dev_alloc(ctx, s->pool, total));
  char* base = (char*)s->pool.p;
  for (const auto& pt : parts) {
    pt.first->p = base;
    pt.first->bytes = pt.second;
    pt.first->view = true;
    base += round_up((int64_t)pt.second, 256);
  }
  return EDHMC_OK;
}

int64_t edhmc_sur_num_params(edhmc_ctx* ctx) {
  if (!ctx || !ctx->sur) return -1;
  return ctx->sur->n_params;
}

int edhmc_sur_set_params(edhmc_ctx* ctx, const float* params, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(check_sur(ctx));
  Surrogate* s = ctx->sur;
  hipStream_t st = (hipStream_t)stream;
  if (!params) return set_err(ctx, EDHMC_EINVAL, "params is NULL");
  if (hipMemsetAsync(s->P.p, 0, s->P.bytes, st) != hipSuccess ||
      hipMemsetAsync(s->Am.p, 0, s->Am.bytes, st) != hipSuccess ||
      hipMemsetAsync(s->Av.p, 0, s->Av.bytes, st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "memset");
  hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(s->n_params)), dim3(256), 0, st, layout_of(s),
                     s->n_params, params, (float*)s->P.p, 1);
  EDHMC_TRY(check_launch(ctx, "pack"));
  return refresh_weights(ctx, s, st);
}

int edhmc_sur_get_params(edhmc_ctx* ctx, float* params_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(check_sur(ctx));
  Surrogate* s = ctx->sur;
  if (!params_out) return set_err(ctx, EDHMC_EINVAL, "params_out is NULL");
  hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(s->n_params)), dim3(256), 0,
                     (hipStream_t)stream, layout_of(s), s->n_params, params_out, (float*)s->P.p,
                     0);
  return check_launch(ctx, "unpack");
}

int edhmc_sur_set_norm(edhmc_ctx* ctx, const float* in_shift, const float* in_scale,
                       const float* out_shift, const float* out_scale, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(check_sur(ctx));
  Surrogate* s = ctx->sur;
  hipStream_t st = (hipStream_t)stream;
  if (!in_shift || !in_scale || !out_shift || !out_scale)
    return set_err(ctx, EDHMC_EINVAL, "NULL normalisation vector");
  float* n = (float*)s->nrm.p;
  if (hipMemsetAsync(n, 0, s->nrm.bytes, st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "memset");
  const float* src[4] = {in_shift, in_scale, out_shift, out_scale};
  for (int k = 0; k < 4; ++k)
    if (hipMemcpyAsync(n + k * s->Dp, src[k], (size_t)s->D * 4, hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      return set_err(ctx, EDHMC_EHIP, "copy normalisation");
  s->norm_set = true;
  s->ghat_valid = false;
  return EDHMC_OK;
}

int edhmc_sur_forward(edhmc_ctx* ctx, const float* W, int32_t B, float* grad_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(check_sur(ctx));
  Surrogate* s = ctx->sur;
  EDHMC_TRY(batch_ok(ctx, s, B));
  if (!W || !grad_out) return set_err(ctx, EDHMC_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  const int Bp = (int)round_up(B, 128);
  const float* nrm = (const float*)s->nrm.p;
  hipLaunchKernelGGL(norm_in_kernel, dim3(blocks_for((int64_t)Bp * s->Dp)), dim3(256), 0, st, W,
                     B, s->D, (int64_t)s->D, nrm, nrm + s->Dp, (bf16_t*)s->x0.p, Bp, s->Dp,
                     (bf16_t*)s->h2.p + s->Hp, s->Kc);
  EDHMC_TRY(check_launch(ctx, "norm_in"));
  EDHMC_TRY(forward_hidden(ctx, s, Bp, false, st));
  EpiGradOut e;
  e.bias = (const float*)s->P.p + s->ob3;
  e.out_shift = nrm + 2 * s->Dp;
  e.out_scale = nrm + 3 * s->Dp;
  e.B = B;
  e.D = s->D;
  e.out = grad_out;
  e.ldo = s->D;
  s->ghat_valid = false;   // x0/h buffers reused
  return mlp_gemm(ctx, s->h2.p, s->Kc, s->W3b.p, s->Kc, Bp, s->Dp, s->Kc, e, st);
}

int edhmc_sur_loss_grad(edhmc_ctx* ctx, const float* W, const float* G, int32_t B, double* loss_out,
                        float* param_grad_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(check_sur(ctx));
  Surrogate* s = ctx->sur;
  EDHMC_TRY(batch_ok(ctx, s, B));
  if (!W || !G) return set_err(ctx, EDHMC_EINVAL, "NULL argument");
  hipStream_t st = (hipStream_t)stream;
  s->ghat_valid = false;
  EDHMC_TRY(loss_grad(ctx, s, W, G, B, loss_out ? loss_out : (double*)s->loss.p, st, nullptr, 0,
                       param_grad_out != nullptr));
  if (param_grad_out) {
    hipLaunchKernelGGL(pack_kernel, dim3(blocks_for(s->n_params)), dim3(256), 0, st, layout_of(s),
                       s->n_params, param_grad_out, (float*)s->Gr.p, 0);
    EDHMC_TRY(check_launch(ctx, "unpack"));
  }
  return EDHMC_OK;
}

// One Adam step on a batch (rows idx[0..B) of the P rows of W and G, or the
// first B rows): loss and parameter gradient, the update, whose kernel also
// writes the layers' bf16 weight copies, and the transposed copies.
static int train_step(edhmc_ctx* ctx, Surrogate* s, const float* W, const float* G, int B,
                      float lr, float beta1, float beta2, float adam_eps, int step,
                      double* loss_out, hipStream_t st, const int64_t* idx, int64_t P_rows) {
  s->ghat_valid = false;
  EDHMC_TRY(loss_grad(ctx, s, W, G, B, loss_out ? loss_out : (double*)s->loss.p, st, idx, P_rows));
  const float c1 = (float)(1.0 / (1.0 - std::pow((double)beta1, step)));
  const float c2 = (float)(1.0 / (1.0 - std::pow((double)beta2, step)));
  AdamBf16Out wb;
  wb.w[0] = (bf16_t*)s->W1b.p; wb.off[0] = s->oW1; wb.n[0] = (int64_t)s->Hp * s->Dp;
  wb.w[1] = (bf16_t*)s->W2b.p; wb.off[1] = s->oW2; wb.n[1] = (int64_t)s->Hp * s->Hp;
  wb.w[2] = (bf16_t*)s->W3b.p; wb.off[2] = s->oW3; wb.n[2] = (int64_t)s->Dp * s->Kc;
  hipLaunchKernelGGL(adam_kernel, dim3(blocks_for(s->n_pad)), dim3(256), 0, st, (float*)s->P.p,
                     (const float*)s->Gr.p, (float*)s->Am.p, (float*)s->Av.p, s->n_pad, lr, beta1,
                     beta2, adam_eps, c1, c2, wb);
  EDHMC_TRY(check_launch(ctx, "adam"));
  return transpose_weights(ctx, s, st);
}

int edhmc_sur_train_step(edhmc_ctx* ctx, const float* W, const float* G, int32_t B, float lr,
                         float beta1, float beta2, float adam_eps, int32_t step, double* loss_out,
                         void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(check_sur(ctx));
  Surrogate* s = ctx->sur;
  if (step < 1) return set_err(ctx, EDHMC_EINVAL, "Adam step count starts at 1");
  EDHMC_TRY(batch_ok(ctx, s, B));
  if (!W || !G) return set_err(ctx, EDHMC_EINVAL, "NULL argument");
  return train_step(ctx, s, W, G, B, lr, beta1, beta2, adam_eps, step, loss_out,
                    (hipStream_t)stream, nullptr, 0);
}

int edhmc_sur_train_epoch(edhmc_ctx* ctx, const float* W, const float* G, int64_t P,
                          const int64_t* perm, int32_t batch, int32_t n_batches,
                          const float* lr_host, float beta1, float beta2, float adam_eps,
                          int32_t step0, double* losses_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(check_sur(ctx));
  Surrogate* s = ctx->sur;
  EDHMC_TRY(batch_ok(ctx, s, batch));
  if (!W || !G || !perm || !lr_host || !losses_out)
    return set_err(ctx, EDHMC_EINVAL, "NULL argument");
  if (n_batches < 0 || (int64_t)n_batches * batch > P)
    return set_err(ctx, EDHMC_EINVAL, "%d batches of %d exceed the %lld pairs", n_batches, batch,
                   (long long)P);
  if (step0 < 0) return set_err(ctx, EDHMC_EINVAL, "Adam step count must be >= 0");
  // each batch's rows are read through perm by the first kernels of the
  // step (no gathered copy)
  for (int b = 0; b < n_batches; ++b)
    EDHMC_TRY(train_step(ctx, s, W, G, batch, lr_host[b], beta1, beta2, adam_eps, step0 + b + 1,
                         losses_out + b, (hipStream_t)stream, perm + (int64_t)b * batch, P));
  return EDHMC_OK;
}

int edhmc_gather_rows(edhmc_ctx* ctx, const float* src, int64_t P, int32_t D,
                      const int64_t* idx, int32_t B, float* dst, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!src || !idx || !dst || P <= 0 || D <= 0 || B < 0)
    return set_err(ctx, EDHMC_EINVAL, "bad arguments");
  if (B == 0) return EDHMC_OK;
  hipLaunchKernelGGL(gather_rows_kernel, dim3(blocks_for((int64_t)B * D)), dim3(256), 0,
                     (hipStream_t)stream, src, P, D, idx, B, dst);
  return check_launch(ctx, "gather_rows");
}

int edhmc_sur_skip_gram(edhmc_ctx* ctx, const float* W, const float* G, int64_t P,
                        const int64_t* idx, int32_t n, double* xtx_out, double* xtt_out,
                        void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(check_sur(ctx));
  Surrogate* s = ctx->sur;
  if (!s->norm_set) return set_err(ctx, EDHMC_ESTATE, "normalisation not set");
  if (!W || !G || !idx || !xtx_out || !xtt_out || n <= 0 || P <= 0)
    return set_err(ctx, EDHMC_EINVAL, "bad arguments");
  const int nb = (s->D + 63) / 64;
  hipLaunchKernelGGL(skip_gram_kernel, dim3(nb, nb, 2), dim3(256), 0, (hipStream_t)stream, W, G,
                     P, s->D, idx, n, (const float*)s->nrm.p, s->Dp, xtx_out, xtt_out);
  return check_launch(ctx, "skip_gram");
}

int edhmc_affine_rows(edhmc_ctx* ctx, const float* E, int64_t n, int32_t D, const float* R,
                      const float* shift, float* out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!E || !R || !shift || !out || n < 0 || D <= 0)
    return set_err(ctx, EDHMC_EINVAL, "bad arguments");
  if (n == 0) return EDHMC_OK;
  hipLaunchKernelGGL(affine_rows_kernel, dim3((unsigned)((n + 63) / 64), (D + 63) / 64),
                     dim3(256), 0, (hipStream_t)stream, E, n, D, R, shift, out);
  return check_launch(ctx, "affine_rows");
}

int edhmc_transition_surrogate(edhmc_ctx* ctx, float eps, int32_t n_steps, const float* r0,
                               const double* u, uint64_t seed, uint64_t counter,
                               int32_t* accepted_out, float* sample_out, int32_t* n_accept,
                               void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(check_sur(ctx));
  if (ctx->has_kinds)
    return set_err(ctx, EDHMC_EUNSUPPORTED,
                   "the surrogate transition needs Normal priors on every coordinate");
  return sur_transition(ctx, eps, n_steps, r0, u, seed, counter, accepted_out, sample_out,
                        n_accept, (hipStream_t)stream);
}

}  // extern "C"
