// HMC engine for generalised linear models on MI355X (gfx950).
//
// One gradient evaluation of the log joint at the C chains' positions Z [C, D]
// (the work of hmc.py:199/206 tf.gradients over hmc.py:161-192) is
//   GEMM1  eta^T [C, N] = Z . X^T         (MFMA; epilogue: residual
//          R^T = dll/deta written chain-major, and the FP64 log-likelihood
//          partial sums per (chain, row tile) when the ratio needs them)
//   GEMM2  G [C, D] = R^T . X              (MFMA, split-K over rows)
//   chain  sum the split-K partials, add the prior gradient, then the
//          leapfrog kick/drift -- or, after the last step, the Metropolis
//          test of hmc.py:100-111 and the sample write of hmc.py:121-129.
// X is stored twice, row-major Xr [N, D] (GEMM1's B) and column-major
// Xc [D, N] (GEMM2's B), so every MFMA operand is K-contiguous.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <type_traits>

#include "common.h"
#include "ctx.h"
#include "gemm_nt.h"
#include "gemm2_dma.h"
#include "philox.h"
#include "transform.h"

namespace edhmc {

ProfScope::ProfScope(edhmc_ctx* c, int k, hipStream_t s) : ctx(c), st(s), kind(k) {
  if (!ctx->prof.on || !((ctx->prof.mask >> k) & 1u)) return;
  if (ctx->prof.marks.size() >= 4096) prof_collect(ctx);
  hipEvent_t e = ctx->prof.get();
  if (!e) return;
  idx = (int)ctx->prof.next - 1;
  (void)hipEventRecord(e, st);
}

ProfScope::~ProfScope() {
  if (idx < 0) return;
  hipEvent_t e = ctx->prof.get();
  if (!e) return;
  (void)hipEventRecord(e, st);
  ctx->prof.marks.push_back({kind, idx});
}

int prof_collect(edhmc_ctx* ctx) {
  Profiler& p = ctx->prof;
  for (auto& m : p.marks) {
    hipEvent_t a = p.pool[m.second], b = p.pool[m.second + 1];
    if (hipEventSynchronize(b) != hipSuccess) return set_err(ctx, EDHMC_EHIP, "event sync");
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, a, b) == hipSuccess) {
      p.total_ms[m.first] += ms;
      p.count[m.first] += 1;
    }
  }
  p.marks.clear();
  p.next = 0;
  return EDHMC_OK;
}

// Tile shapes (blocks per wave M, N; waves M, N).  The defaults can be
// overridden at compile time to sweep alternatives.
#ifndef EDHMC_CFG_G1
#define EDHMC_CFG_G1 2, 2, 2, 2
#endif
#ifndef EDHMC_CFG_G2
#define EDHMC_CFG_G2 2, 2, 4, 4
#endif
#ifndef EDHMC_CFG_P2
#define EDHMC_CFG_P2 1, 4, 2, 2, 64
#endif
#ifndef EDHMC_GEMM1_SWZ
#define EDHMC_GEMM1_SWZ 3
#endif
// GEMM1 tile order: 3 = each XCD keeps a slice of the chains (the W planes
// stay in its L2) and streams X tiles; 1 = each XCD keeps X row tiles and
// streams all chains.
constexpr int kGemm1Swz = EDHMC_GEMM1_SWZ;
using Cfg1 = TileCfg<EDHMC_CFG_G1>;    // GEMM1 (gradient pass, one plane)
// GEMM2 when D_pad % CfgG2::BN == 0: 256 x 256 tiles of 16 waves (64 x 64
// each), four waves per SIMD (4.33 -> 3.73 ms at N = 1e6, D = 512, 4096
// chains against 4 waves of 128 x 128, which hold one wave per SIMD)
using CfgG2 = TileCfg<EDHMC_CFG_G2>;
// bf16 GEMM2 with LDS-direct stage loads (gemm2_dma.h) where it applies
#ifndef EDHMC_G2_DMA
#define EDHMC_G2_DMA 1
#endif
constexpr bool kG2Dma = EDHMC_G2_DMA != 0;
using Cfg2S = TileCfg<1, 1, 4, 1>;     // 128 x 32, GEMM2 for narrow D
using CfgP2 = TileCfg<EDHMC_CFG_P2>;   // GEMM1 over hi/lo split positions
constexpr int cmax(int a, int b) { return a > b ? a : b; }
// chains are padded to a multiple of every M tile (and of the 128-row MLP
// batch tile); data rows to a multiple of every row tile
constexpr int kChainPad = cmax(cmax(Cfg1::BM, CfgG2::BM), cmax(CfgP2::BM, 128));
constexpr int kRowPad = cmax(cmax(Cfg1::BN, CfgP2::BN), 128);
static_assert(kChainPad % Cfg1::BM == 0 && kChainPad % CfgG2::BM == 0 &&
              kChainPad % CfgP2::BM == 0 && kChainPad % 128 == 0, "chain padding");
static_assert(kRowPad % Cfg1::BN == 0 && kRowPad % CfgP2::BN == 0, "row padding");
constexpr int kChainThreads = 256;

// ---------------------------------------------------------------------------
// GEMM1 epilogue: residuals and FP64 log-likelihood partials.
// Accumulator element (m, n) = eta of chain m at data row n.
// ---------------------------------------------------------------------------
template <typename T, bool WR, bool LL>
struct EpiGLM {
  const float* y;
  int64_t N;
  LikParams lp;
  T* Rt;              // K-blocked [N_pad / KE][C_pad][KE], KE = 64 bytes of T
  int64_t ldr;        // C_pad
  double* llpart;     // [n_tiles][C_pad]
  int64_t ll_ld;      // C_pad
  // Optional per-row offset added to eta: goff[m * goff_ld + gidx[n]] (the
  // group intercept a_{g(n)} of the hierarchical model); null for a GLM.
  const float* goff = nullptr;
  int64_t goff_ld = 0;
  const int* gidx = nullptr;

  // The residual tile goes through LDS: lanes write it row-major ([m][n],
  // padded rows), then each thread stores whole 16-byte pieces of the
  // K-blocked layout, in which the BM rows of one 64-byte K block are
  // contiguous -- instead of 2- or 4-byte scattered stores.  lds_bytes asks
  // the kernel for room for it beside the ll scratch and the staged group
  // intercepts.
  static constexpr int KG = 8;   // group intercepts staged per row tile
  template <typename Cfg>
  static constexpr int lds_bytes() {
    return WR ? ((Cfg::BM * (Cfg::BN * (int)sizeof(T) + 16) + 15) / 16 * 16 +
                 Cfg::kWavesN * Cfg::BM * 8 + Cfg::BM * KG * 4)
              : 0;
  }
  template <typename Cfg>
  __device__ void apply(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p) const {
    constexpr int WM = Cfg::kWM, WN = Cfg::kWN;
    constexpr int BM = Cfg::BM, BN = Cfg::BN, NT = Cfg::kThreads;
    constexpr int KE = kStageBytes / sizeof(T);
    constexpr int RROW = BN * (int)sizeof(T) + 16;          // LDS row stride (bytes)
    constexpr int RTILE = (BM * RROW + 15) / 16 * 16;
    constexpr int LLB = Cfg::kWavesN * BM * 8;
    const bool stage = WR && (RTILE + LLB <= p.smem_bytes);
    char* rt = p.smem;
    double* sm = reinterpret_cast<double*>(p.smem + (stage ? RTILE : 0));
    float yv[WN];
    bool valid[WN];
    int gj[WN];
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int64_t n = p.n0 + 32 * j + (p.lane & 31);
      valid[j] = n < N;
      yv[j] = valid[j] ? y[n] : 0.0f;
      gj[j] = (goff && valid[j]) ? gidx[n] : 0;
    }
    const int ml0 = p.m0 - p.tm * BM, nl0 = p.n0 - p.tn * BN;   // tile-local origin
    // Group intercepts: rows are sorted by group, so a row tile spans the
    // groups g_lo..g_hi; when that range is short, the tile's intercepts
    // a[m][g_lo..g_hi] are staged in LDS once (one coalesced pass) instead of
    // one scattered global load per accumulator element.
    float* gl = reinterpret_cast<float*>(p.smem + (stage ? RTILE : 0) + LLB);
    int g_lo = 0;
    bool g_lds = false;
    if (goff) {
      const int64_t nf = (int64_t)p.tn * BN;
      const int64_t nl = min((int64_t)N, nf + BN) - 1;
      if (nf < N) {
        g_lo = gidx[nf];
        const int ng = gidx[nl] - g_lo + 1;
        g_lds = ng <= KG && (stage ? RTILE : 0) + LLB + BM * KG * 4 <= p.smem_bytes;
        if (g_lds) {
          for (int t = threadIdx.x; t < BM * ng; t += NT) {
            const int m = t / ng, k = t - m * ng;
            gl[m * KG + k] = goff[(int64_t)(p.tm * BM + m) * goff_ld + g_lo + k];
          }
          __syncthreads();
        }
      }
    }
    // Forward-only Bernoulli pass: the epilogue sums -(softplus(eta) - y eta)
    // in fp32 over the lane's WN columns (non-negative terms, bern_nll_fast),
    // FP64 across them.
    const bool sp_only = LL && !WR && !goff && lp.lik == LIK_BERNOULLI;
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int mr = 32 * i + acc_row(r, p.lane);
        double llsum = 0.0;
        const float* grow = goff ? goff + (int64_t)(p.m0 + mr) * goff_ld : nullptr;
        if (sp_only) {
          float s32 = 0.0f;
#pragma unroll
          for (int j = 0; j < WN; ++j) {
            const float nl = bern_nll_fast(acc[i][j][r], yv[j]);   // unconditional: no branch
            s32 -= valid[j] ? nl : 0.0f;
          }
          llsum = (double)s32;
        } else {
#pragma unroll
          for (int j = 0; j < WN; ++j) {
            const int nc = 32 * j + (p.lane & 31);
            // padded rows (n >= N) carry gj = 0: index slot 0 for them, never
            // gj - g_lo, which is negative there and would address outside
            // the workgroup's LDS
            const float eta = acc[i][j][r] +
                              (grow ? (g_lds ? gl[(ml0 + mr) * KG + (valid[j] ? gj[j] - g_lo : 0)]
                                             : grow[gj[j]])
                                    : 0.0f);
            float dll;
            if (LL) {
              double ll;
              glm_both(lp, eta, yv[j], dll, ll);
              llsum += valid[j] ? ll : 0.0;
            } else {
              dll = glm_dll(lp, eta, yv[j]);
            }
            if (WR) {
              const T v = ElemOps<T>::from_f32(valid[j] ? dll : 0.0f);
              if (stage) {
                *(T*)(rt + (ml0 + mr) * RROW + (nl0 + nc) * (int)sizeof(T)) = v;
              } else {
                const int64_t m = p.m0 + mr, n = p.n0 + nc;
                Rt[((n / KE) * ldr + m) * KE + (n % KE)] = v;
              }
            }
          }
        }
        if (LL) {
#pragma unroll
          for (int off = 16; off > 0; off >>= 1) llsum += __shfl_xor(llsum, off, 64);
          if ((p.lane & 31) == 0) sm[p.wn * BM + ml0 + mr] = llsum;
        }
      }
    if (stage || LL) __syncthreads();
    if (stage) {
      // BN/KE K-blocks x BM rows x 4 pieces of 16 bytes
      constexpr int PIECES = (BN / KE) * BM * 4;
      char* base = (char*)Rt + ((int64_t)(p.tn * BN / KE) * ldr + (int64_t)p.tm * BM) * kStageBytes;
      for (int id = threadIdx.x; id < PIECES; id += NT) {
        const int part = id & 3, m = (id >> 2) % BM, nb = (id >> 2) / BM;
        const uint4 v = *(const uint4*)(rt + m * RROW + nb * kStageBytes + part * 16);
        *(uint4*)(base + ((int64_t)nb * ldr + m) * kStageBytes + part * 16) = v;
      }
    }
    if (LL) {
      for (int t = threadIdx.x; t < BM; t += NT) {
        double s = 0.0;
#pragma unroll
        for (int w = 0; w < Cfg::kWavesN; ++w) s += sm[w * BM + t];
        llpart[(int64_t)p.tn * ll_ld + p.tm * BM + t] = s;
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Posterior-predictive epilogue (criticism): accumulator element (m, n) =
// eta of posterior draw m at data row n.  Per element: the log-likelihood of
// y_n (summed per draw in FP64, as EpiGLM's partials), the mean response
// (sigmoid(eta) or eta) summed over the tile's valid draws into a per-tile
// column partial, and optionally the mean response itself into rep[m][n]
// (replicate_kernel then draws y_rep from it; keeping the RNG out of the
// epilogue keeps the accumulator loops fully unrolled, in registers).
// ---------------------------------------------------------------------------
template <typename T>
struct EpiPredict {
  const float* y;
  int64_t N;
  LikParams lp;
  double* llpart;       // [n_tiles][ll_ld]
  int64_t ll_ld;
  float* colsum;        // [m_tiles][N_pad] or null
  int64_t col_ld;       // N_pad
  float* rep;           // [m_valid][N] or null
  int m_valid;          // number of draws (rows >= m_valid are padding)

  template <typename Cfg>
  __device__ void apply(f32x16_t (&acc)[Cfg::kWM][Cfg::kWN], const EpiPos& p) const {
    constexpr int WM = Cfg::kWM, WN = Cfg::kWN, BM = Cfg::BM, NT = Cfg::kThreads;
    double* sm = reinterpret_cast<double*>(p.smem);                 // [kWavesN][BM]
    float* cs = reinterpret_cast<float*>(p.smem + Cfg::kWavesN * BM * 8);  // [kWavesM][BN]
    const int ml0 = p.m0 - p.tm * BM, nl0 = p.n0 - p.tn * Cfg::BN;
    float yv[WN], col[WN];
    bool valid[WN];
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      const int64_t n = p.n0 + 32 * j + (p.lane & 31);
      valid[j] = n < N;
      yv[j] = valid[j] ? y[n] : 0.0f;
      col[j] = 0.0f;
    }
#pragma unroll
    for (int i = 0; i < WM; ++i)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int mr = 32 * i + acc_row(r, p.lane);
        const int m = p.m0 + mr;
        const bool mv = m < m_valid;
        double llsum = 0.0;
#pragma unroll
        for (int j = 0; j < WN; ++j) {
          const float eta = acc[i][j][r];
          const int64_t n = p.n0 + 32 * j + (p.lane & 31);
          const double ll = glm_ll(lp, eta, yv[j]);
          llsum += valid[j] ? ll : 0.0;
          float mean;
          if (lp.lik == LIK_BERNOULLI) {
            const float e = __expf(-fabsf(eta));
            const float inv = fast_rcp(1.0f + e);
            mean = eta >= 0.0f ? inv : e * inv;
          } else {
            mean = eta;
          }
          col[j] += mv ? mean : 0.0f;
          // the mean response; replicate_kernel turns it into a draw
          if (rep && mv && valid[j]) rep[(int64_t)m * N + n] = mean;
        }
#pragma unroll
        for (int off = 16; off > 0; off >>= 1) llsum += __shfl_xor(llsum, off, 64);
        if ((p.lane & 31) == 0) sm[p.wn * BM + ml0 + mr] = llsum;
      }
    // column sums: the two lane halves hold different rows of the same columns
#pragma unroll
    for (int j = 0; j < WN; ++j) {
      col[j] += __shfl_xor(col[j], 32, 64);
      if (p.lane < 32) cs[p.wm * Cfg::BN + nl0 + 32 * j + p.lane] = col[j];
    }
    __syncthreads();
    for (int t = threadIdx.x; t < BM; t += NT) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < Cfg::kWavesN; ++w) s += sm[w * BM + t];
      llpart[(int64_t)p.tn * ll_ld + p.tm * BM + t] = s;
    }
    if (colsum) {
      for (int t = threadIdx.x; t < Cfg::BN; t += NT) {
        float s = 0.0f;
#pragma unroll
        for (int w = 0; w < Cfg::kWavesM; ++w) s += cs[w * Cfg::BN + t];
        colsum[(int64_t)p.tm * col_ld + (int64_t)p.tn * Cfg::BN + t] = s;
      }
    }
  }
};

// y_rep[s][n] from the mean response mu in place: Bernoulli u < mu with a
// 24-bit Philox uniform u; Normal mu + scale * z (Box-Muller).  Counter
// (s, n) on stream 1 of `seed`.
__global__ void __launch_bounds__(256)
replicate_kernel(float* __restrict__ rep, int S, int64_t N, int lik, float scale, uint64_t seed) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (int64_t)S * N) return;
  const int s = (int)(i / N);
  const int64_t n = i - (int64_t)s * N;
  const u32x4 x = philox_draw(seed, (uint64_t)s, (uint32_t)n, 0u, 1);
  const float mu = rep[i];
  if (lik == LIK_BERNOULLI) {
    const float u = (float)(x.x >> 8) * 5.9604644775390625e-8f;   // [0, 1)
    rep[i] = u < mu ? 1.0f : 0.0f;
  } else {
    float z0, z1;
    box_muller(x.x, x.y, z0, z1);
    rep[i] = mu + scale * z0;
  }
}

// mean_out[n] = (sum over M tiles of colsum[tm][n]) / S
__global__ void __launch_bounds__(256)
colsum_reduce_kernel(const float* __restrict__ part, int m_tiles, int64_t ld, int64_t N, float inv_s,
                     float* __restrict__ out) {
  const int64_t n = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (n >= N) return;
  float s = 0.0f;
  for (int t = 0; t < m_tiles; ++t) s += part[(int64_t)t * ld + n];
  out[n] = s * inv_s;
}

// ---------------------------------------------------------------------------
// Data layout: X [N, D] (fp32 or bf16 in) -> Xr [N_pad, D_pad], Xc [D_pad, N_pad]
// in the compute type; counts non-finite inputs.
// ---------------------------------------------------------------------------
template <typename Tin, typename T>
__global__ void __launch_bounds__(256)
layout_x_kernel(const Tin* __restrict__ X, int64_t N, int D, T* __restrict__ Xr,
                T* __restrict__ Xc, int64_t N_pad, int D_pad, int* nonfinite) {
  __shared__ float tile[32][33];
  const int64_t n0 = (int64_t)blockIdx.y * 32;
  const int d0 = blockIdx.x * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  int bad = 0;
  for (int k = ty; k < 32; k += 8) {
    const int64_t n = n0 + k;
    const int d = d0 + tx;
    float v = 0.0f;
    if (n < N && d < D) {
      if constexpr (sizeof(Tin) == 2) v = bf16_to_f32(((const bf16_t*)X)[n * D + d]);
      else v = ((const float*)X)[n * D + d];
      if (!isfinite(v)) { bad = 1; v = 0.0f; }
    }
    tile[k][tx] = v;
    Xr[n * D_pad + d] = ElemOps<T>::from_f32(v);
  }
  if (bad) atomicAdd(nonfinite, 1);
  __syncthreads();
  // Xc is K-blocked: [N_pad / KE][D_pad][KE] with KE = 64 bytes of T.
  constexpr int KE = kStageBytes / sizeof(T);
  for (int k = ty; k < 32; k += 8) {
    const int d = d0 + k;
    const int64_t n = n0 + tx;
    Xc[((n / KE) * D_pad + d) * KE + (n % KE)] = ElemOps<T>::from_f32(tile[tx][k]);
  }
}

// Xf: bf16 X in the operand order of v_mfma_f32_16x16x32_bf16, for the
// W-stationary kernels.  Block (b, s) -- rows 16 b .. 16 b + 15, columns
// 32 s .. 32 s + 31 -- is 1 KB at element offset (b * D_pad / 32 + s) * 512,
// and its 16-byte piece l (l < 64) holds row 16 b + (l & 15), columns
// 32 s + 8 (l >> 4) .. + 8: a wave's operand load for one 16-row block and
// one 32-column k step is then 1 KB contiguous.  One thread per piece.
__global__ void __launch_bounds__(256)
layout_frag_kernel(const bf16_t* __restrict__ Xr, int64_t pieces, int D_pad,
                   bf16_t* __restrict__ Xf) {
  const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (q >= pieces) return;
  const int KS = D_pad / 32;
  const int64_t blk = q >> 6;
  const int l = (int)(q & 63);
  const int64_t b = blk / KS;
  const int s = (int)(blk - b * KS);
  const int64_t n = 16 * b + (l & 15);
  const int d = 32 * s + 8 * (l >> 4);
  *(uint4*)(Xf + q * 8) = *(const uint4*)(Xr + n * D_pad + d);
}

// Deterministic reduction of the log-likelihood partials llpart[t][c] over
// the row tiles t, in two fixed-order passes: kLLSplit contiguous ranges of t
// per 64-chain block (4 interleaved strides inside a range, combined in a
// fixed order), then the kLLSplit range sums in order.  The split keeps
// thousands of threads in flight when C is small (a 512-chain reduction over
// 6400 tiles would otherwise run on 8 workgroups).
constexpr int kLLSplit = 16;

__global__ void __launch_bounds__(256)
ll_partial_kernel(const double* __restrict__ llpart, int n_tiles, int C_pad,
                  double* __restrict__ split) {
  __shared__ double part[4][64];
  const int tc = threadIdx.x & 63, tg = threadIdx.x >> 6;
  const int c = blockIdx.x * 64 + tc;
  const int per = (n_tiles + kLLSplit - 1) / kLLSplit;
  const int t0 = blockIdx.y * per, t1 = min(n_tiles, t0 + per);
  double s = 0.0;
  for (int t = t0 + tg; t < t1; t += 4) s += llpart[(int64_t)t * C_pad + c];
  part[tg][tc] = s;
  __syncthreads();
  if (tg == 0)
    split[(int64_t)blockIdx.y * C_pad + c] = ((part[0][tc] + part[1][tc]) + part[2][tc]) + part[3][tc];
}

__global__ void __launch_bounds__(256)
ll_final_kernel(const double* __restrict__ split, int C_pad, double* __restrict__ ll) {
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c >= C_pad) return;
  double s = 0.0;
#pragma unroll
  for (int k = 0; k < kLLSplit; ++k) s += split[(int64_t)k * C_pad + c];
  ll[c] = s;
}

__global__ void pad_copy_kernel(const float* __restrict__ src, int64_t n, float* __restrict__ dst,
                                int64_t n_pad, float fill, int* nonfinite) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_pad) return;
  float v = fill;
  if (i < n) {
    v = src[i];
    if (!isfinite(v)) atomicAdd(nonfinite, 1);
  }
  dst[i] = v;
}

// ---------------------------------------------------------------------------
// Chain kernels (one workgroup per chain).
// ---------------------------------------------------------------------------
struct ChainArgs {
  int C, D, D_pad, C_pad;
  int bf16;                 // z_lp holds hi/lo bf16 planes (else fp32)
  float* z; float* r; float* g; float* z0; float* g0;
  void* z_lp;
  double* logp; double* K0;
  const float* ploc; const float* pscale;
  const uint8_t* pkind;     // per-coordinate prior family | operand << 4 (null: Normal, identity)
  const double* plnorm;     // per-coordinate normalising constant of that family
  float half, full;         // fl32(eps/2), fl32(eps)
  const float* r_fed; const double* u_fed;
  uint64_t seed, counter;
  const float* Gpart; int ksplit;
  const double* ll;         // [C_pad] reduced log-likelihood
  float lik_weight;         // multiplies the likelihood (scale of the SG samplers; 1 for HMC)
  const float* beta;        // per-chain inverse temperature of the MH ratio (null = 1)
  float* sample_out; int32_t* accepted_out; int32_t* n_accept; int32_t* ndiv;
  double* ratio_out;        // per-chain log ratio of the HMC accept test (or null)
  double* mom_mean; double* mom_m2; int64_t mom_n;   // mom_n = count after update
  float* grad_out; double* logp_out;
  const float* W_in;        // set_state input [C, D]
  float* rec_W; float* rec_G;   // training-pair rows for this step (or null)
};

// The linear predictor's operand for position zv of coordinate d.
__device__ inline float operand(const ChainArgs& a, int d, float zv) {
  if (!a.pkind) return zv;
  const int op = a.pkind[d] >> 4;
  return op == kOpIdentity ? zv : (float)operand_f64(op, (double)zv);
}

__device__ inline void write_lp(const ChainArgs& a, int c, int d, float zv) {
  zv = operand(a, d, zv);
  const int64_t idx = (int64_t)c * a.D_pad + d;
  if (a.bf16) {
    bf16_t* p = (bf16_t*)a.z_lp;
    bf16_t hi = f32_to_bf16(zv);
    bf16_t lo = f32_to_bf16(zv - bf16_to_f32(hi));
    p[idx] = hi;
    p[idx + (int64_t)a.C_pad * a.D_pad] = lo;
  } else {
    ((float*)a.z_lp)[idx] = zv;
  }
}

__device__ inline float prior_grad(const ChainArgs& a, int d, float zv) {
  const float s = a.pscale[d];
  return -(zv - a.ploc[d]) / (s * s);
}

__device__ inline double prior_logp(const ChainArgs& a, int d, float zv) {
  if (a.pkind)
    return prior_term(a.pkind[d] & 15, (double)zv, a.ploc[d], a.pscale[d], a.plnorm[d]).lp;
  const double s = a.pscale[d];
  const double t = ((double)zv - (double)a.ploc[d]) / s;
  return -0.5 * t * t - log(s) - 0.5 * kLog2Pi;
}

// Gradient of the log joint in position space: gl = d loglik / d operand
// (already weighted), chained through the operand map, plus the prior term
// (with the log-Jacobian of a constrained prior).
__device__ inline float grad_u(const ChainArgs& a, int d, float zv, float gl) {
  if (!a.pkind) return gl + prior_grad(a, d, zv);
  const int k = a.pkind[d];
  const double u = zv;
  const PriorTerm t = prior_term(k & 15, u, a.ploc[d], a.pscale[d], a.plnorm[d]);
  return (float)((double)gl * d_operand_f64(k >> 4, u) + t.dlp);
}

// Momentum draw, K0 = |r0|^2 / 2, save (z0, g0), first half kick and drift.
__global__ void __launch_bounds__(kChainThreads) chain_begin_kernel(ChainArgs a) {
  __shared__ double red[kChainThreads / 64];
  const int c = blockIdx.x;
  double k = 0.0;
  const int nb = a.D_pad / 4;
  for (int j = threadIdx.x; j < nb; j += kChainThreads) {
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
      const int64_t idx = (int64_t)c * a.D_pad + d;
      float r = rv[e];
      k += (double)r * (double)r;
      const float zv = a.z[idx], gv = a.g[idx];
      a.z0[idx] = zv;
      a.g0[idx] = gv;
      r = r + a.half * gv;
      const float zn = zv + a.full * r;
      a.z[idx] = zn;
      a.r[idx] = r;
      write_lp(a, c, d, zn);
    }
  }
  k = block_sum_f64<kChainThreads>(k, red);
  if (threadIdx.x == 0) a.K0[c] = 0.5 * k;
}

__device__ inline float gsum(const ChainArgs& a, int c, int d) {
  const int64_t stride = (int64_t)a.C_pad * a.D_pad;
  const float* p = a.Gpart + (int64_t)c * a.D_pad + d;
  float s = 0.0f;
  for (int q = 0; q < a.ksplit; ++q) s += p[q * stride];
  return a.lik_weight * s;
}


// Middle of the trajectory: gradient at the new z, r += eps/2 g twice
// (end of this step, start of the next), drift.
__global__ void __launch_bounds__(kChainThreads) chain_mid_kernel(ChainArgs a) {
  const int c = blockIdx.x;
  for (int d = threadIdx.x; d < a.D; d += kChainThreads) {
    const int64_t idx = (int64_t)c * a.D_pad + d;
    const float zv = a.z[idx];
    const float gv = grad_u(a, d, zv, gsum(a, c, d));
    a.g[idx] = gv;
    if (a.rec_W) {
      a.rec_W[(int64_t)c * a.D + d] = zv;
      a.rec_G[(int64_t)c * a.D + d] = gv;
    }
    float r = a.r[idx];
    r = r + a.half * gv;
    r = r + a.half * gv;
    const float zn = zv + a.full * r;
    a.r[idx] = r;
    a.z[idx] = zn;
    write_lp(a, c, d, zn);
  }
}

// Shared end of trajectory: last half kick, ratio, accept, select, store.
// `g_at` supplies the gradient at the end point (exact or surrogate).
template <typename GradAt>
__device__ inline void chain_accept(const ChainArgs& a, GradAt g_at) {
  __shared__ double red[kChainThreads / 64];
  __shared__ int acc_flag;
  const int c = blockIdx.x;
  double kl = 0.0, plp = 0.0;
  int bad = 0;
  for (int d = threadIdx.x; d < a.D; d += kChainThreads) {
    const int64_t idx = (int64_t)c * a.D_pad + d;
    const float zv = a.z[idx];
    const float gv = g_at(c, d, idx, zv);
    a.g[idx] = gv;
    const float r = a.r[idx] + a.half * gv;
    a.r[idx] = r;
    kl += (double)r * (double)r;
    plp += prior_logp(a, d, zv);
    if (!isfinite(zv) || !isfinite(gv)) bad = 1;
  }
  kl = block_sum_f64<kChainThreads>(kl, red);
  plp = block_sum_f64<kChainThreads>(plp, red);
  const double ll = (double)a.lik_weight * a.ll[c];
  const double nbad = block_sum_f64<kChainThreads>((double)bad, red);
  if (threadIdx.x == 0) {
    const double logpL = ll + plp;
    const double ratio = a.K0[c] - 0.5 * kl + logpL - a.logp[c];
    const double u = a.u_fed ? a.u_fed[c] : accept_uniform(a.seed, a.counter, (uint32_t)c);
    // hmc.py:109: accept = log(u) < ratio (NaN never accepts).
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
  for (int d = threadIdx.x; d < a.D; d += kChainThreads) {
    const int64_t idx = (int64_t)c * a.D_pad + d;
    float zv;
    if (acc) {
      zv = a.z[idx];
    } else {
      zv = a.z0[idx];
      a.z[idx] = zv;
      a.g[idx] = a.g0[idx];
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

__global__ void __launch_bounds__(kChainThreads) chain_last_kernel(ChainArgs a) {
  chain_accept(a, [&](int c, int d, int64_t, float zv) {
    return grad_u(a, d, zv, gsum(a, c, d));
  });
}

// Evaluation at the current z: gradient + log joint into the state (and
// optional outputs).
__global__ void __launch_bounds__(kChainThreads) chain_eval_kernel(ChainArgs a) {
  __shared__ double red[kChainThreads / 64];
  const int c = blockIdx.x;
  double plp = 0.0;
  for (int d = threadIdx.x; d < a.D; d += kChainThreads) {
    const int64_t idx = (int64_t)c * a.D_pad + d;
    const float zv = a.z[idx];
    const float gv = grad_u(a, d, zv, gsum(a, c, d));
    a.g[idx] = gv;
    if (a.grad_out) a.grad_out[(int64_t)c * a.D + d] = gv;
    plp += prior_logp(a, d, zv);
  }
  plp = block_sum_f64<kChainThreads>(plp, red);
  const double ll = (double)a.lik_weight * a.ll[c];
  if (threadIdx.x == 0) {
    a.logp[c] = ll + plp;
    if (a.logp_out) a.logp_out[c] = ll + plp;
  }
}

// Load positions W [C, D] into z and z_lp; count non-finite entries.
__global__ void __launch_bounds__(kChainThreads) chain_load_kernel(ChainArgs a, int* nonfinite) {
  const int c = blockIdx.x;
  int bad = 0;
  for (int d = threadIdx.x; d < a.D; d += kChainThreads) {
    const int64_t idx = (int64_t)c * a.D_pad + d;
    float v = a.W_in[(int64_t)c * a.D + d];
    if (!isfinite(v)) bad = 1;
    a.z[idx] = v;
    write_lp(a, c, d, v);
  }
  if (bad) atomicAdd(nonfinite, 1);
}

// Positions W [C, D] into the z_lp planes only (forward-only evaluation).
__global__ void __launch_bounds__(kChainThreads) lp_load_kernel(ChainArgs a, int* nonfinite) {
  const int c = blockIdx.x;
  int bad = 0;
  for (int d = threadIdx.x; d < a.D; d += kChainThreads) {
    const float v = a.W_in[(int64_t)c * a.D + d];
    if (!isfinite(v)) bad = 1;
    write_lp(a, c, d, v);
  }
  if (bad) atomicAdd(nonfinite, 1);
}

// logp_out[c] = ll[c] + prior log density of W[c].
__global__ void __launch_bounds__(kChainThreads) logp_out_kernel(ChainArgs a) {
  __shared__ double red[kChainThreads / 64];
  const int c = blockIdx.x;
  double plp = 0.0;
  for (int d = threadIdx.x; d < a.D; d += kChainThreads)
    plp += prior_logp(a, d, a.W_in[(int64_t)c * a.D + d]);
  plp = block_sum_f64<kChainThreads>(plp, red);
  if (threadIdx.x == 0) a.logp_out[c] = (double)a.lik_weight * a.ll[c] + plp;
}

// Rewrite z_lp from z for every chain.
__global__ void __launch_bounds__(kChainThreads) chain_lp_kernel(ChainArgs a) {
  const int c = blockIdx.x;
  for (int d = threadIdx.x; d < a.D; d += kChainThreads)
    write_lp(a, c, d, a.z[(int64_t)c * a.D_pad + d]);
}

// Stochastic-gradient samplers: one workgroup per chain, the gradient at z in
// g (scaled log joint), N(0, 1) draws from the momentum stream of (seed,
// counter).  SGLD (edward/inferences/sgld.py:62-80):
//   z <- z + lr/2 g + sqrt(lr) xi.
// SGHMC (edward/inferences/sghmc.py:66-86), v a per-chain velocity:
//   sample = z + v;  v <- (1 - friction/2) v + lr g + sqrt(lr friction) xi.
// Both write the sample row and count it as accepted (n_accept += 1).
template <bool HAMILTONIAN>
__global__ void __launch_bounds__(kChainThreads)
sg_step_kernel(ChainArgs a, float lr, float friction, float noise_sd, float* v) {
  const int c = blockIdx.x;
  const int nb = (a.D + 3) / 4;
  for (int jb = threadIdx.x; jb < nb; jb += kChainThreads) {
    const f32x4_t xi = momentum4(a.seed, a.counter, (uint32_t)jb, (uint32_t)c);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int d = 4 * jb + e;
      if (d >= a.D) continue;
      const int64_t idx = (int64_t)c * a.D_pad + d;
      const float zv = a.z[idx], gv = a.g[idx];
      float zn;
      if (HAMILTONIAN) {
        const int64_t vi = (int64_t)c * a.D + d;
        const float vv = v[vi];
        zn = zv + vv;
        v[vi] = (1.0f - 0.5f * friction) * vv + lr * gv + noise_sd * xi[e];
      } else {
        zn = zv + 0.5f * lr * gv + noise_sd * xi[e];
      }
      a.z[idx] = zn;
      write_lp(a, c, d, zn);
      if (a.sample_out) a.sample_out[(int64_t)c * a.D + d] = zn;
    }
  }
  if (threadIdx.x == 0 && a.n_accept) a.n_accept[c] += 1;
}

// logp[c] = ll[c] + prior log density at z[c] (after a forward-only pass).
__global__ void __launch_bounds__(kChainThreads) chain_relogp_kernel(ChainArgs a) {
  __shared__ double red[kChainThreads / 64];
  const int c = blockIdx.x;
  double plp = 0.0;
  for (int d = threadIdx.x; d < a.D; d += kChainThreads)
    plp += prior_logp(a, d, a.z[(int64_t)c * a.D_pad + d]);
  plp = block_sum_f64<kChainThreads>(plp, red);
  if (threadIdx.x == 0) a.logp[c] = (double)a.lik_weight * a.ll[c] + plp;
}

// Random-walk Metropolis-Hastings (edward/inferences/metropolis_hastings.py:
// 64-156) with the proposal Normal(loc = z, scale = s): z0 <- z,
// z <- z + s xi (xi from the momentum stream), z_lp for the forward pass.
__global__ void __launch_bounds__(kChainThreads)
mh_propose_kernel(ChainArgs a, const float* __restrict__ scale) {
  const int c = blockIdx.x;
  const int nb = (a.D + 3) / 4;
  for (int jb = threadIdx.x; jb < nb; jb += kChainThreads) {
    const f32x4_t xi = momentum4(a.seed, a.counter, (uint32_t)jb, (uint32_t)c);
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      const int d = 4 * jb + e;
      if (d >= a.D) continue;
      const int64_t idx = (int64_t)c * a.D_pad + d;
      const float zv = a.z[idx];
      const float zn = zv + scale[d] * xi[e];
      a.z0[idx] = zv;
      a.z[idx] = zn;
      write_lp(a, c, d, zn);
    }
  }
}

// The symmetric proposal's log densities cancel (metropolis_hastings.py:
// 101-115), so ratio = log p(x, z') - log p(x, z); accept iff log u < ratio
// and the proposal is finite (a non-finite one is rejected and counted).
__global__ void __launch_bounds__(kChainThreads) mh_accept_kernel(ChainArgs a) {
  __shared__ double red[kChainThreads / 64];
  __shared__ int acc_flag;
  const int c = blockIdx.x;
  double plp = 0.0;
  int bad = 0;
  for (int d = threadIdx.x; d < a.D; d += kChainThreads) {
    const float zv = a.z[(int64_t)c * a.D_pad + d];
    plp += prior_logp(a, d, zv);
    if (!isfinite(zv)) bad = 1;
  }
  plp = block_sum_f64<kChainThreads>(plp, red);
  const double nbad = block_sum_f64<kChainThreads>((double)bad, red);
  if (threadIdx.x == 0) {
    const double lpn = (double)a.lik_weight * a.ll[c] + plp;
    // replica_exchange_mc.py:200: the whole MH ratio is scaled by the
    // replica's inverse temperature
    const double ratio = (lpn - a.logp[c]) * (a.beta ? (double)a.beta[c] : 1.0);
    const double u = a.u_fed ? a.u_fed[c] : accept_uniform(a.seed, a.counter, (uint32_t)c);
    const int acc = (log(u) < ratio) && nbad == 0.0;
    acc_flag = acc;
    if (acc) a.logp[c] = lpn;
    if (a.accepted_out) a.accepted_out[c] = acc;
    if (a.n_accept) a.n_accept[c] += acc;
    if (a.ndiv && (nbad != 0.0 || !(ratio == ratio))) a.ndiv[c] += 1;
  }
  __syncthreads();
  const int acc = acc_flag;
  for (int d = threadIdx.x; d < a.D; d += kChainThreads) {
    const int64_t idx = (int64_t)c * a.D_pad + d;
    float zv = a.z[idx];
    if (!acc) {
      zv = a.z0[idx];
      a.z[idx] = zv;
    }
    if (a.sample_out) a.sample_out[(int64_t)c * a.D + d] = zv;
  }
}

__global__ void copy_out_kernel(const float* __restrict__ src, int C, int D, int D_pad,
                                float* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)C * D) return;
  const int c = (int)(i / D), d = (int)(i - (int64_t)c * D);
  dst[i] = src[(int64_t)c * D_pad + d];
}

__global__ void copy_out_f64_kernel(const double* __restrict__ src, int C, int D, int D_pad,
                                    double* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)C * D) return;
  const int c = (int)(i / D), d = (int)(i - (int64_t)c * D);
  dst[i] = src[(int64_t)c * D_pad + d];
}

__global__ void rng_momenta_kernel(int C, int D, uint64_t seed, uint64_t counter, float* out) {
  const int c = blockIdx.x;
  const int nb = (D + 3) / 4;
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
    f32x4_t v = momentum4(seed, counter, (uint32_t)j, (uint32_t)c);
#pragma unroll
    for (int e = 0; e < 4; ++e)
      if (4 * j + e < D) out[(int64_t)c * D + 4 * j + e] = v[e];
  }
}

__global__ void rng_uniforms_kernel(int C, uint64_t seed, uint64_t counter, double* out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < C) out[c] = accept_uniform(seed, counter, (uint32_t)c);
}

// Autocovariance for ESS: thread per (chain, coordinate); lags in blocks of 16
// with a register window.
__global__ void __launch_bounds__(256)
autocov_kernel(const float* __restrict__ s, int64_t T, int C, int D, int max_lag,
               double* __restrict__ acov, double* __restrict__ cmean) {
  const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (id >= (int64_t)C * D) return;
  const int c = (int)(id / D), d = (int)(id - (int64_t)c * D);
  const int64_t stride = (int64_t)C * D;
  const float* x = s + (int64_t)c * D + d;
  double m = 0.0;
  for (int64_t i = 0; i < T; ++i) m += x[i * stride];
  m /= (double)T;
  if (cmean) cmean[id] = m;
  for (int l0 = 0; l0 <= max_lag; l0 += 16) {
    double acc[16];
    float win[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      acc[k] = 0.0;
      const int64_t j = l0 + k;
      win[k] = j < T ? x[j * stride] : 0.0f;
    }
    for (int64_t i = 0; i + l0 < T; ++i) {
      const double xi = (double)x[i * stride] - m;
#pragma unroll
      for (int k = 0; k < 16; ++k)
        if (i + l0 + k < T) acc[k] += xi * ((double)win[k] - m);
#pragma unroll
      for (int k = 0; k < 15; ++k) win[k] = win[k + 1];
      const int64_t j = i + l0 + 16;
      win[15] = j < T ? x[j * stride] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < 16; ++k)
      if (l0 + k <= max_lag) atomicAdd(&acov[(int64_t)(l0 + k) * D + d], acc[k] / (double)T);
  }
}

// Column statistics, pass 1: per (row chunk, column) FP64 sum and sum of
// squares; pass 2: fixed-order combination.
__global__ void __launch_bounds__(256)
col_stats_part_kernel(const float* __restrict__ X, int64_t P, int D, int chunks,
                      double* __restrict__ part) {
  const int d = blockIdx.x * 256 + threadIdx.x;
  const int ch = blockIdx.y;
  if (d >= D) return;
  const int64_t per = (P + chunks - 1) / chunks;
  const int64_t b = ch * per, e = min(P, b + per);
  double s = 0.0, q = 0.0;
  for (int64_t i = b; i < e; ++i) {
    const double v = X[i * D + d];
    s += v;
    q += v * v;
  }
  part[((int64_t)ch * D + d) * 2] = s;
  part[((int64_t)ch * D + d) * 2 + 1] = q;
}

__global__ void col_stats_final_kernel(const double* __restrict__ part, int64_t P, int D,
                                       int chunks, float* __restrict__ mean,
                                       float* __restrict__ sd) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= D) return;
  double s = 0.0, q = 0.0;
  for (int ch = 0; ch < chunks; ++ch) {
    s += part[((int64_t)ch * D + d) * 2];
    q += part[((int64_t)ch * D + d) * 2 + 1];
  }
  const double m = s / (double)P;
  const double v = fmax(q / (double)P - m * m, 0.0);
  mean[d] = (float)m;
  sd[d] = (float)sqrt(v);
}

// Empirical moments: thread per column, two passes over rows t0..T-1.
template <typename P>
__global__ void __launch_bounds__(256)
sample_moments_kernel(const P* __restrict__ x, int64_t T, int64_t K, int64_t t0,
                      double* __restrict__ mean, double* __restrict__ var) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= K) return;
  double s = 0.0;
  for (int64_t t = t0; t < T; ++t) s += (double)x[t * K + k];
  const double n = (double)(T - t0);
  const double m = s / n;
  double q = 0.0;
  for (int64_t t = t0; t < T; ++t) {
    const double d = (double)x[t * K + k] - m;
    q += d * d;
  }
  mean[k] = m;
  var[k] = q / n;
}

// ---------------------------------------------------------------------------
// Host-side orchestration
// ---------------------------------------------------------------------------
static int elem_bytes(const edhmc_ctx* ctx) { return ctx->cdt == EDHMC_BF16 ? 2 : 4; }

static LikParams lik_params(const edhmc_ctx* ctx) {
  LikParams p;
  p.lik = ctx->lik;
  p.inv_s2 = 1.0f / (ctx->lik_scale * ctx->lik_scale);
  p.ll_const = -log((double)ctx->lik_scale) - 0.5 * kLog2Pi;
  return p;
}

static ChainArgs chain_args(edhmc_ctx* ctx) {
  ChainArgs a = {};
  a.C = ctx->C; a.D = ctx->D; a.D_pad = ctx->D_pad; a.C_pad = ctx->C_pad;
  a.bf16 = ctx->cdt == EDHMC_BF16;
  a.z = (float*)ctx->z.p; a.r = (float*)ctx->r.p; a.g = (float*)ctx->g.p;
  a.z0 = (float*)ctx->z0.p; a.g0 = (float*)ctx->g0.p;
  a.z_lp = ctx->z_lp.p;
  a.logp = (double*)ctx->logp.p; a.K0 = (double*)ctx->K0.p;
  a.ploc = (const float*)ctx->ploc.p; a.pscale = (const float*)ctx->pscale.p;
  if (ctx->has_kinds) {
    a.pkind = (const uint8_t*)ctx->pkind.p;
    a.plnorm = (const double*)ctx->plnorm.p;
  }
  a.Gpart = (const float*)ctx->Gpart.p; a.ksplit = ctx->ksplit;
  a.ll = (const double*)ctx->ll.p;
  a.lik_weight = ctx->lik_weight;
  a.ndiv = (int32_t*)ctx->ndiv.p;
  a.ratio_out = (double*)ctx->ratio.p;
  return a;
}

// ll[c] = sum over the n_tiles row tiles of llpart[t][c] (fixed order).
static int ll_reduce(edhmc_ctx* ctx, int n_tiles, hipStream_t st) {
  hipLaunchKernelGGL(ll_partial_kernel, dim3(ctx->C_pad / 64, kLLSplit), dim3(256), 0, st,
                     (const double*)ctx->llpart.p, n_tiles, ctx->C_pad, (double*)ctx->ll_split.p);
  EDHMC_TRY(check_launch(ctx, "ll_partial"));
  hipLaunchKernelGGL(ll_final_kernel, dim3((ctx->C_pad + 255) / 256), dim3(256), 0, st,
                     (const double*)ctx->ll_split.p, ctx->C_pad, (double*)ctx->ll.p);
  return check_launch(ctx, "ll_final");
}

template <typename T, typename Cfg, int PLANES, bool WR, bool LL>
static int launch_gemm1(edhmc_ctx* ctx, hipStream_t st, const float* goff = nullptr,
                        int64_t goff_ld = 0, const int* gidx = nullptr) {
  const int eb = sizeof(T);
  GemmShape g = {};
  g.A = (const char*)ctx->z_lp.p;
  g.lda_b = (int64_t)ctx->D_pad * eb;
  g.a_stage_b = Cfg::KS;
  g.a_plane_b = (int64_t)ctx->C_pad * ctx->D_pad * eb;
  g.B = (const char*)ctx->Xr.p;
  g.ldb_b = (int64_t)ctx->D_pad * eb;
  g.b_stage_b = Cfg::KS;
  g.m_tiles = ctx->C_pad / Cfg::BM;
  g.n_tiles = (int)(ctx->N_pad / Cfg::BN);
  g.k_stages = ctx->D_pad * eb / Cfg::KS;
  g.ksplit = 1;
  g.stages_per_split = g.k_stages;
  g.swz = (kGemm1Swz == 3 && g.m_tiles % 8 == 0) ? 3 : 1;
  ProfScope ps(ctx, WR ? PK_GEMM_GRAD : PK_GEMM_LL, st);
  EpiGLM<T, WR, LL> e;
  e.y = (const float*)ctx->y.p;
  e.N = ctx->N;
  e.lp = lik_params(ctx);
  e.Rt = (T*)ctx->Rt.p;
  e.ldr = ctx->C_pad;
  e.llpart = (double*)ctx->llpart.p;
  e.ll_ld = ctx->C_pad;
  e.goff = goff;
  e.goff_ld = goff_ld;
  e.gidx = gidx;
  if (!gemm_offsets_fit<Cfg, PLANES>(g)) return set_err(ctx, EDHMC_EUNSUPPORTED, "GEMM1 tile too large");
  hipLaunchKernelGGL((gemm_nt_kernel<T, Cfg, PLANES, EpiGLM<T, WR, LL>>), dim3(grid_size(g)),
                     dim3(Cfg::kThreads), 0, st, g, e);
  EDHMC_TRY(check_launch(ctx, "gemm1"));
  if (LL) {
    EDHMC_TRY(ll_reduce(ctx, g.n_tiles, st));
  }
  return EDHMC_OK;
}

template <typename T, typename Cfg>
static int launch_gemm2_cfg(edhmc_ctx* ctx, hipStream_t st) {
  const int eb = sizeof(T);
  GemmShape g = {};
  g.A = (const char*)ctx->Rt.p;
  g.lda_b = kStageBytes;
  g.a_stage_b = (int64_t)ctx->C_pad * kStageBytes;
  g.a_plane_b = 0;
  g.B = (const char*)ctx->Xc.p;
  g.ldb_b = kStageBytes;
  g.b_stage_b = (int64_t)ctx->D_pad * kStageBytes;
  g.m_tiles = ctx->C_pad / Cfg::BM;
  g.n_tiles = ctx->D_pad / Cfg::BN;
  g.k_stages = (int)(ctx->N_pad * eb / kStageBytes);
  g.ksplit = ctx->ksplit;
  g.stages_per_split = (g.k_stages + g.ksplit - 1) / g.ksplit;
  g.swz = (g.ksplit % 8 == 0) ? 2 : 0;
  ProfScope ps(ctx, PK_GEMM_XTR, st);
  EpiStoreF32 e;
  e.out = (float*)ctx->Gpart.p;
  e.ldc = ctx->D_pad;
  e.split_b = (int64_t)ctx->C_pad * ctx->D_pad;
  if (!gemm_offsets_fit<Cfg, 1>(g)) return set_err(ctx, EDHMC_EUNSUPPORTED, "GEMM2 tile too large");
  hipLaunchKernelGGL((gemm_nt_kernel<T, Cfg, 1, EpiStoreF32>), dim3(grid_size(g)),
                     dim3(Cfg::kThreads), 0, st, g, e);
  return check_launch(ctx, "gemm2");
}

// bf16 GEMM2 on 256 x 256 tiles with stages loaded straight into LDS
// (gemm2_dma.h); same split-K and tile order as launch_gemm2_cfg, same bits.
static int launch_gemm2_dma(edhmc_ctx* ctx, hipStream_t st) {
  GemmShape g = {};
  g.A = (const char*)ctx->Rt.p;
  g.lda_b = kStageBytes;
  g.a_stage_b = (int64_t)ctx->C_pad * kStageBytes;
  g.B = (const char*)ctx->Xc.p;
  g.ldb_b = kStageBytes;
  g.b_stage_b = (int64_t)ctx->D_pad * kStageBytes;
  g.m_tiles = ctx->C_pad / kG2Tile;
  g.n_tiles = ctx->D_pad / kG2Tile;
  g.k_stages = (int)(ctx->N_pad * 2 / kStageBytes);
  g.ksplit = ctx->ksplit;
  g.stages_per_split = (g.k_stages + g.ksplit - 1) / g.ksplit;
  g.swz = (g.ksplit % 8 == 0) ? 2 : 0;
  ProfScope ps(ctx, PK_GEMM_XTR, st);
  hipLaunchKernelGGL(gemm2_dma_kernel, dim3(grid_size(g)), dim3(256), 0, st, g,
                     (float*)ctx->Gpart.p, (int64_t)ctx->D_pad,
                     (int64_t)ctx->C_pad * ctx->D_pad);
  return check_launch(ctx, "gemm2");
}

// Whether launch_gemm2_dma applies: bf16, whole 256 x 256 tiles, and every
// K split at least kG2Ring stages (the kernel's prologue fills its ring).
static bool gemm2_dma_ok(const edhmc_ctx* ctx) {
  if (!kG2Dma || ctx->cdt != EDHMC_BF16 || ctx->D_pad % kG2Tile || ctx->C_pad % kG2Tile)
    return false;
  const int64_t kst = ctx->N_pad * 2 / kStageBytes;
  const int64_t spl = (kst + ctx->ksplit - 1) / ctx->ksplit;
  return spl >= kG2Ring && kst - (int64_t)(ctx->ksplit - 1) * spl >= kG2Ring;
}

template <typename T>
static int launch_gemm2(edhmc_ctx* ctx, hipStream_t st) {
  if (std::is_same<T, bf16_t>::value && gemm2_dma_ok(ctx)) return launch_gemm2_dma(ctx, st);
  if (ctx->D_pad % CfgG2::BN == 0 && ctx->C_pad % CfgG2::BM == 0)
    return launch_gemm2_cfg<T, CfgG2>(ctx, st);
  return launch_gemm2_cfg<T, Cfg2S>(ctx, st);
}

// One gradient evaluation at z_lp.  split: use both bf16 planes; ll: write the
// log-likelihood partials.
static bool ws_forward_ok(const edhmc_ctx* ctx);
static int launch_ws_forward(edhmc_ctx* ctx, hipStream_t st, bool wr, int planes = 2,
                             bool ll = true);
static bool ws_grad_ok(const edhmc_ctx* ctx) {
#ifdef EDHMC_NO_WS_GRAD
  return false;
#endif
  return ws_forward_ok(ctx);
}

static int grad_eval(edhmc_ctx* ctx, bool split, bool ll, hipStream_t st) {
  if (ctx->cdt == EDHMC_BF16) {
    if (split) {
      if (ws_grad_ok(ctx)) {
        // the W-stationary kernel writes the residuals beside its
        // log-likelihood partials (the forward-only pass's formula; reduced
        // only when asked for)
        EDHMC_TRY(launch_ws_forward(ctx, st, true, 2, ll));
      } else if (ll) {
        EDHMC_TRY((launch_gemm1<bf16_t, CfgP2, 2, true, true>(ctx, st)));
      } else {
        EDHMC_TRY((launch_gemm1<bf16_t, CfgP2, 2, true, false>(ctx, st)));
      }
    } else {
      if (ll) EDHMC_TRY((launch_gemm1<bf16_t, Cfg1, 1, true, true>(ctx, st)));
      else if (ws_grad_ok(ctx)) EDHMC_TRY(launch_ws_forward(ctx, st, true, 1, false));
      else EDHMC_TRY((launch_gemm1<bf16_t, Cfg1, 1, true, false>(ctx, st)));
    }
    return launch_gemm2<bf16_t>(ctx, st);
  }
  if (ll) EDHMC_TRY((launch_gemm1<float, Cfg1, 1, true, true>(ctx, st)));
  else EDHMC_TRY((launch_gemm1<float, Cfg1, 1, true, false>(ctx, st)));
  return launch_gemm2<float>(ctx, st);
}

// fp32 gradient pass with a per-row gathered offset in eta (hierarchical
// model): residuals R^T, optional ll, then X^T R into the split-K partials.
#ifndef EDHMC_CFG_H
#define EDHMC_CFG_H 2, 2, 2, 2, 128
#endif
// GEMM1 of the hierarchical model (fp32, short K): 128-byte K stages halve
// the stage barriers of its 2-4 stages (1.11 vs 1.19 ms per pass at G = 2048,
// D = 64, 512 chains; 128 x 256 or 256 x 128 tiles leave one workgroup per
// CU for the residual tile's LDS: 3.4-3.9 ms)
using CfgH = TileCfg<EDHMC_CFG_H>;
int glm_grad_eval_offset(edhmc_ctx* ctx, const float* goff, int64_t goff_ld, const int* gidx,
                         bool ll, hipStream_t st) {
  if (ctx->cdt != EDHMC_F32) return set_err(ctx, EDHMC_EUNSUPPORTED, "offset path is fp32");
  if (ctx->C_pad % CfgH::BM || ctx->N_pad % CfgH::BN)
    return set_err(ctx, EDHMC_EUNSUPPORTED, "hierarchical GEMM1 tiles do not divide the padding");
  if (ll) EDHMC_TRY((launch_gemm1<float, CfgH, 1, true, true>(ctx, st, goff, goff_ld, gidx)));
  else EDHMC_TRY((launch_gemm1<float, CfgH, 1, true, false>(ctx, st, goff, goff_ld, gidx)));
  return launch_gemm2<float>(ctx, st);
}

// Buffers the chain-level engines above the GLM need to size their state.
int glm_ensure_chains(edhmc_ctx* ctx, int C);

// ---------------------------------------------------------------------------
// W-stationary forward log-likelihood (Bernoulli, bf16 X, hi/lo W planes):
// the sum over rows of y eta - softplus(eta) for every chain, eta = W X^T.
//
// A workgroup (8 wave64s, one per CU) owns 64 chains and a contiguous range
// of data rows.  The hi and lo planes of its 64 positions stay in LDS for
// the whole launch (rows padded by 16 bytes: a ds_read_b128 pass over eight
// or sixteen consecutive rows hits disjoint banks).  Each wave streams its
// own 64 rows of X per 512-row tile straight from memory into the operand
// registers of v_mfma_f32_16x16x32_bf16: for a 16-row block and a 64-byte
// k step, lane l holds row l & 15, bytes 16 (l >> 4) .. +16.  X is read from
// Xf, a copy in exactly that order (layout_frag_kernel), so each load
// instruction reads 1 KB contiguous -- eight whole cache lines, where the
// row-major copy gives sixteen half lines (12 % of the forward's time).
// kWsPre k steps of X are in flight, the prefetch running on into the next tile while the
// previous tile's epilogue runs; no LDS staging of X and no block barrier in
// the main loop.  Per k step a wave issues 32 MFMAs (2 planes x 4 chain
// blocks x 4 row blocks).  The epilogue of a tile adds y eta - softplus(eta) over the
// wave's valid rows into FP64 per-lane sums; at the end lanes, then waves,
// are summed in a fixed order into llpart[part][chain].
// ---------------------------------------------------------------------------
constexpr int kWsThreads = 512;
constexpr int kWsBM = 64;                  // chains per workgroup
constexpr int kWsWaveRows = 64;            // rows per wave per tile (4 blocks of 16)
constexpr int kWsTile = 8 * kWsWaveRows;   // rows per tile
#ifndef EDHMC_WS_PRE
#define EDHMC_WS_PRE 4
#endif
constexpr int kWsPre = EDHMC_WS_PRE;       // 64-byte k steps of X in flight per wave
#ifndef EDHMC_WS_PRE_GRAD
#define EDHMC_WS_PRE_GRAD 4
#endif
constexpr int kWsPreGrad = EDHMC_WS_PRE_GRAD;   // the same with the residual epilogue
#ifndef EDHMC_WS_PRE_GRAD1
#define EDHMC_WS_PRE_GRAD1 4
#endif
constexpr int kWsPreGrad1 = EDHMC_WS_PRE_GRAD1; // ... on one plane without the log-likelihood
#ifndef EDHMC_WS_WIDE1
#define EDHMC_WS_WIDE1 1
#endif
constexpr bool kWsWide1 = EDHMC_WS_WIDE1 != 0;   // one-plane pass on 128-chain tiles
constexpr int kWsMaxDPad = 512;
#ifndef EDHMC_WS_PAD
#define EDHMC_WS_PAD 16
#endif
constexpr int kWsPad = EDHMC_WS_PAD;       // bytes of padding per resident W row
constexpr int kWsRowBytes = 2 * kWsMaxDPad + kWsPad;

typedef __attribute__((ext_vector_type(4))) float f32x4v_t;

__device__ inline void mfma16(f32x4v_t& acc, const uint4& a, const uint4& b) {
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8_t, a),
                                                __builtin_bit_cast(bf16x8_t, b), acc, 0, 0, 0);
}

// WR: also the gradient pass's residuals r = y - sigmoid(eta) (bf16), written
// to Rt in the K-blocked layout [N_pad / 32][C_pad][32] that GEMM2 reads
// (0 for the padded rows N .. N_pad), beside the log-likelihood partials.
// PRE: 64-byte k steps of X in flight per wave (4 in every variant: 2 or 8
// measured slower, 8 spilling).
// PLANES: 2 (hi and lo planes of W, the exact log joint) or 1 (hi plane
// only: the one-plane inner leapfrog gradients); LL: sum y eta - softplus(eta)
// (bern_nll_fast's non-negative form).
template <bool WR, int PLANES, bool LL, int PRE, int NI = 4>
__global__ void __launch_bounds__(kWsThreads) __attribute__((amdgpu_waves_per_eu(1, 2)))
gemm_ll_ws_kernel(const bf16_t* __restrict__ Wp, int64_t plane_elems, int D_pad,
                  const bf16_t* __restrict__ X, int64_t N, int64_t N_pad, int64_t rows_per_part,
                  int parts, int mt_per_xcd, int part_per_xcd, double* __restrict__ llpart,
                  int64_t ll_ld, const float* __restrict__ y, bf16_t* __restrict__ Rt,
                  int64_t ldr) {
  // NI chain blocks of 16 per workgroup (BMC chains; 8 only for the one-plane
  // pass, whose single plane leaves room in LDS for 128 chains: each X
  // fragment then feeds twice the MFMAs)
  constexpr int BMC = 16 * NI;
  static_assert(PLANES * BMC <= 2 * kWsBM && (!LL || NI == 4), "chain tile");
  __shared__ __align__(16) char lds[2 * kWsBM * kWsRowBytes];
  // Consecutive workgroup ids go to the 8 XCDs in turn.  With part_per_xcd
  // (parts divides 8) every XCD streams one row range for 8 / parts slices
  // of its chain tiles, so all its workgroups read the same X rows through
  // its L2; otherwise an XCD owns mt_per_xcd chain tiles over all parts.
  const int id = blockIdx.x, xcd = id & 7, local = id >> 3;
  int tm, part;
  if (part_per_xcd) {
    part = xcd % parts;
    tm = (xcd / parts) * mt_per_xcd + local;
  } else {
    tm = xcd * mt_per_xcd + local % mt_per_xcd;
    part = local / mt_per_xcd;
  }
  if (part >= parts) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int rsb = 2 * D_pad + kWsPad;
  // resident W: LDS row p * BMC + c = plane p of chain BMC tm + c
  const int cpr = D_pad / 8;                 // 16-byte chunks per row
  for (int q = tid; q < PLANES * BMC * cpr; q += kWsThreads) {
    const int row = q / cpr, ch = q - row * cpr;
    const int pl = row / BMC, c = row - pl * BMC;
    const uint4 v = *(const uint4*)(Wp + pl * plane_elems + (int64_t)(tm * BMC + c) * D_pad +
                                    ch * 8);
    *(uint4*)(lds + row * rsb + ch * 16) = v;
  }
  __syncthreads();

  const int64_t n_begin = (int64_t)part * rows_per_part;
  const int64_t n_end = min(n_begin + rows_per_part, N);
  const int KS = D_pad / 32;                 // 64-byte k steps (a multiple of PRE)
  const int64_t n_tiles = n_end > n_begin ? (n_end - n_begin + kWsTile - 1) / kWsTile : 0;
  const int l16 = lane & 15, lq = (lane >> 4) * 16;
  // byte address of this lane's 16-byte piece of the Xf block of row block j
  // of the wave's rows in tile t at k step 0 (k step ks adds ks KB); blocks
  // past the padding are clamped (never used)
  const int64_t last_blk = N_pad / 16 - 1;
  auto row_ptr = [&](int64_t t, int j) -> const char* {
    int64_t b = (n_begin + t * kWsTile + wave * kWsWaveRows + j * 16) >> 4;
    b = b < last_blk ? b : last_blk;
    return (const char*)X + (b * KS) * 1024 + lane * 16;
  };
  // Forward (!WR): W is the MFMA's A operand and X its B operand, so
  // acc[i][j][v] is chain 16 i + 4 (lane >> 4) + v at data row 16 j + l16 of
  // the wave's rows; lls[i][v] sums that chain over this lane's rows.
  // Gradient pass (WR): the operands are swapped, acc[i][j][v] is chain
  // 16 i + l16 at row 16 j + 4 (lane >> 4) + v -- a lane's four values of
  // one chain are consecutive rows, stored as one 8-byte piece -- and
  // lls[i][0] sums chain 16 i + l16.  (The swapped layout costs the forward
  // 6 %; the residual stores gain more.)
  double lls[NI][4];
#pragma unroll
  for (int i = 0; i < NI; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) lls[i][v] = 0.0;
  if (n_tiles > 0) {
    uint4 xr[PRE][4];
    {
      const char* rp[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) rp[j] = row_ptr(0, j);
#pragma unroll
      for (int u = 0; u < PRE; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j) xr[u][j] = *(const uint4*)(rp[j] + u * 1024);
    }
    f32x4v_t acc[NI][4];
#pragma unroll
    for (int i = 0; i < NI; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = f32x4v_t{0.0f, 0.0f, 0.0f, 0.0f};
    // y of the wave's rows in the tile computed now (tile 0 first), lane l
    // holding row l
    float y_cur;
    {
      const int64_t n = n_begin + wave * kWsWaveRows + lane;
      y_cur = y[n < N_pad ? n : N_pad - 1];
    }
    // (t, ks0): the block of PRE k steps computed now; (tn, kn): the
    // block whose loads are issued meanwhile (the last block again past the
    // end: in bounds, never used)
    int64_t t = 0;
    int ks0 = 0;
    while (t < n_tiles) {
      int64_t tn = t;
      int kn = ks0 + PRE;
      if (kn == KS) { kn = 0; tn = t + 1; }
      int64_t tl = tn;
      int kl = kn;
      if (tl >= n_tiles) { tl = n_tiles - 1; kl = KS - PRE; }
      const char* rp[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) rp[j] = row_ptr(tl, j) + kl * 1024;
      // y of the wave's 64 rows of the next block's tile (lane l holding row
      // l, one coalesced load per k block; padded rows read y = 0).  It is
      // carried into the next iteration (y_cur below), whose epilogue uses
      // it: a load consumed in its own iteration's epilogue is placed inside
      // that branch and waited for together with every X prefetch in flight.
      float y_next;
      {
        const int64_t n = n_begin + tl * kWsTile + wave * kWsWaveRows + lane;
        y_next = y[n < N_pad ? n : N_pad - 1];
      }
#pragma unroll
      for (int u = 0; u < PRE; ++u) {
        const int ks = ks0 + u;
#pragma unroll
        for (int pl = 0; pl < PLANES; ++pl) {
          uint4 wf[NI];
#pragma unroll
          for (int i = 0; i < NI; ++i)
            wf[i] = *(const uint4*)(lds + (pl * BMC + i * 16 + l16) * rsb + ks * 64 + lq);
#pragma unroll
          for (int i = 0; i < NI; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
              if (WR) mfma16(acc[i][j], xr[u][j], wf[i]);
              else mfma16(acc[i][j], wf[i], xr[u][j]);
            }
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) xr[u][j] = *(const uint4*)(rp[j] + u * 1024);
      }
      if (ks0 + PRE == KS) {
        if (!WR) {
          // end of tile t: the log-likelihood terms of the wave's valid rows
          // (row j*16 + l16)
          const int64_t nb = n_begin + t * kWsTile + wave * kWsWaveRows + l16;
          bool valid[4];
          float yv[4];
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            valid[j] = nb + 16 * j < n_end;
            yv[j] = __shfl(y_cur, 16 * j + l16, 64);
          }
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              float s32 = 0.0f;
#pragma unroll
              for (int j = 0; j < 4; ++j) {
                const float nl = bern_nll_fast(acc[i][j][v], yv[j]);   // unconditional: no branch
                s32 -= valid[j] ? nl : 0.0f;
              }
              lls[i][v] += (double)s32;
            }
        } else {
          // end of tile t: this lane's rows are rb + 16 j + v
          // residuals: a lane's 4 rows of one chain are 8 contiguous bytes of
          // the chain's 64-byte K block; 32-bit offsets from the tile's first
          // K block (its 512 rows span 16 blocks: below 16 * C_pad * 32
          // elements).  Rt holds rows up to N_pad rounded up to a whole tile,
          // so the last tile's stores need no bounds test; rows from N on
          // get 0.  y comes from the lanes holding the rows (y_cur).
          const int64_t rb = n_begin + t * kWsTile + wave * kWsWaveRows + 4 * (lane >> 4);
          const int64_t row0 = n_begin + t * kWsTile;
          bf16_t* rt = Rt + (row0 >> 5) * ldr * 32;
          const uint32_t c0 = (uint32_t)(tm * BMC + l16) * 32u;
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const int64_t r0 = rb + 16 * j;
            const int64_t left = n_end - r0;
            float yv[4];
#pragma unroll
            for (int v = 0; v < 4; ++v) yv[v] = __shfl(y_cur, 16 * j + 4 * (lane >> 4) + v, 64);
            const int nl = (int)(r0 - row0);                       // row within the tile
            const uint32_t ro = (uint32_t)(nl >> 5) * (uint32_t)ldr * 32u + (uint32_t)(nl & 31) + c0;
#pragma unroll
            for (int i = 0; i < NI; ++i) {
              float s32 = 0.0f;
              float rv[4];
#pragma unroll
              for (int v = 0; v < 4; ++v) {
                const float eta = acc[i][j][v];
                // softplus and sigmoid from one t = exp(-|eta|)
                const float tt = __builtin_amdgcn_exp2f(-fabsf(eta) * 1.4426950408889634f);
                if (LL) {
                  // softplus(eta) - y eta as in bern_nll_fast
                  const float nl = fmaf(__builtin_amdgcn_logf(1.0f + tt), 0.6931471805599453f,
                                        fmaf(-yv[v], eta, fmaxf(eta, 0.0f)));
                  s32 -= v < left ? nl : 0.0f;
                }
                const float inv = fast_rcp(1.0f + tt);
                const float sig = eta >= 0.0f ? inv : tt * inv;
                // rows from n_end on: +0 by a bit mask (no branch around the
                // residual's arithmetic)
                const uint32_t keep = v < left ? 0xffffffffu : 0u;
                rv[v] = __builtin_bit_cast(float, __builtin_bit_cast(uint32_t, yv[v] - sig) & keep);
              }
              if (LL) lls[i][0] += (double)s32;
              uint2 pk;
              pk.x = f32x2_to_bf16x2(rv[0], rv[1]);
              pk.y = f32x2_to_bf16x2(rv[2], rv[3]);
              *(uint2*)(rt + ro + (uint32_t)(16 * i) * 32u) = pk;
            }
          }
        }
#pragma unroll
        for (int i = 0; i < NI; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = f32x4v_t{0.0f, 0.0f, 0.0f, 0.0f};
      }
      t = tn;
      ks0 = kn;
      y_cur = y_next;
    }
  }
  if (!LL) return;   // uniform across the workgroup: no barrier below is skipped by part of it
  if (!WR) {
    // lanes with equal lane >> 4 hold the same chains at different rows
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        double x = lls[i][v];
#pragma unroll
        for (int off = 1; off < 16; off <<= 1) x += __shfl_xor(x, off, 64);
        lls[i][v] = x;
      }
  } else {
    // lanes with equal l16 hold the same chain at different rows
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double x = lls[i][0];
      x += __shfl_xor(x, 16, 64);
      x += __shfl_xor(x, 32, 64);
      lls[i][0] = x;
    }
  }
  __syncthreads();   // the W tiles are no longer read: reuse LDS for the wave sums
  double* red = reinterpret_cast<double*>(lds);   // [8 waves][64 chains]
  if (!WR && l16 == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[wave * kWsBM + 16 * i + (lane >> 4) * 4 + v] = lls[i][v];
  }
  if (WR && lane < 16) {
#pragma unroll
    for (int i = 0; i < 4; ++i) red[wave * kWsBM + 16 * i + l16] = lls[i][0];
  }
  __syncthreads();
  if (tid < kWsBM) {
    double x = 0.0;
#pragma unroll
    for (int w = 0; w < kWsThreads / 64; ++w) x += red[w * kWsBM + tid];
    llpart[(int64_t)part * ll_ld + tm * kWsBM + tid] = x;
  }
}

// ---------------------------------------------------------------------------
// Pipelined W-stationary forward (the exact log joint of the hi/lo planes
// for D_pad = 32 KS): four waves per workgroup, one per SIMD, each with the
// whole tile body unrolled (KS k steps of 64 bytes).  A wave computes the
// MFMAs of tile t while it evaluates the log-likelihood terms of tile t - 1,
// 16 / KS (chain block, row block) pairs per k step, from a second
// accumulator set -- the epilogue's exp / log run between the matrix-core
// instructions of the same wave instead of after them.  W fragments are read
// from LDS one k step ahead; X (from Xf) PRE k steps ahead, running on into
// the next tile.  Per element the arithmetic and its order equal
// gemm_ll_ws_kernel<false, 2, true, .>'s (rows are split over 4 waves of 64
// rows per 256-row tile, so the FP64 partial sums group rows differently).
// ---------------------------------------------------------------------------
constexpr int kWpThreads = 256;
constexpr int kWpTile = 4 * kWsWaveRows;   // rows per tile

template <int KS, int PRE>
__global__ void __launch_bounds__(kWpThreads) __attribute__((amdgpu_waves_per_eu(1, 1)))
gemm_ll_wp_kernel(const bf16_t* __restrict__ Wp, int64_t plane_elems,
                  const bf16_t* __restrict__ X, int64_t N, int64_t N_pad, int64_t rows_per_part,
                  int parts, int mt_per_xcd, int part_per_xcd, double* __restrict__ llpart,
                  int64_t ll_ld, const float* __restrict__ y) {
  static_assert(KS % PRE == 0 && KS <= 16, "k steps per tile");
  constexpr int D_pad = 32 * KS;
  constexpr int RSB = 2 * D_pad + kWsPad;
  __shared__ __align__(16) char lds[2 * kWsBM * RSB];
  const int id = blockIdx.x, xcd = id & 7, local = id >> 3;
  int tm, part;
  if (part_per_xcd) {
    part = xcd % parts;
    tm = (xcd / parts) * mt_per_xcd + local;
  } else {
    tm = xcd * mt_per_xcd + local % mt_per_xcd;
    part = local / mt_per_xcd;
  }
  if (part >= parts) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  constexpr int CPR = D_pad / 8;
  for (int q = tid; q < 2 * kWsBM * CPR; q += kWpThreads) {
    const int row = q / CPR, ch = q - row * CPR;
    const int pl = row / kWsBM, c = row - pl * kWsBM;
    const uint4 v = *(const uint4*)(Wp + pl * plane_elems + (int64_t)(tm * kWsBM + c) * D_pad +
                                    ch * 8);
    *(uint4*)(lds + row * RSB + ch * 16) = v;
  }
  __syncthreads();

  const int64_t n_begin = (int64_t)part * rows_per_part;
  const int64_t n_end = min(n_begin + rows_per_part, N);
  const int64_t n_tiles = n_end > n_begin ? (n_end - n_begin + kWpTile - 1) / kWpTile : 0;
  const int l16 = lane & 15, lq = (lane >> 4) * 16;
  const int64_t last_blk = N_pad / 16 - 1;
  // this lane's 16-byte piece of the Xf block of row block j of the wave's
  // rows in tile t at k step 0 (blocks past the padding clamped, never used)
  auto blk_ptr = [&](int64_t t, int j) -> const char* {
    int64_t b = (n_begin + t * kWpTile + wave * kWsWaveRows + j * 16) >> 4;
    b = b < last_blk ? b : last_blk;
    return (const char*)X + (b * KS) * 1024 + lane * 16;
  };
  auto y_at = [&](int64_t t) -> float {
    const int64_t n = n_begin + t * kWpTile + wave * kWsWaveRows + lane;
    return y[n < N_pad ? n : N_pad - 1];
  };
  double lls[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) lls[i][v] = 0.0;
  if (n_tiles > 0) {
    uint4 xr[PRE][4];
#pragma unroll
    for (int u = 0; u < PRE; ++u)
#pragma unroll
      for (int j = 0; j < 4; ++j) xr[u][j] = *(const uint4*)(blk_ptr(0, j) + u * 1024);
    f32x4v_t acc[4][4], accE[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        acc[i][j] = f32x4v_t{0.0f, 0.0f, 0.0f, 0.0f};
        accE[i][j] = f32x4v_t{0.0f, 0.0f, 0.0f, 0.0f};
      }
    // W fragments of k step ks (both planes, 4 chain blocks), double-buffered
    uint4 wf[2][2][4];
    auto read_w = [&](int buf, int ks) {
#pragma unroll
      for (int pl = 0; pl < 2; ++pl)
#pragma unroll
        for (int i = 0; i < 4; ++i)
          wf[buf][pl][i] = *(const uint4*)(lds + (pl * kWsBM + i * 16 + l16) * RSB + ks * 64 + lq);
    };
    read_w(0, 0);
    float y_cur = y_at(0), yE = 0.0f;
    int64_t tE = -1;   // tile whose log-likelihood terms are pending (none yet)
    for (int64_t t = 0; t < n_tiles; ++t) {
      const int64_t tn = t + 1 < n_tiles ? t + 1 : n_tiles - 1;
      const float y_nx = y_at(tn);
      const char* pc[4];
      const char* pn[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        pc[j] = blk_ptr(t, j);
        pn[j] = blk_ptr(tn, j);
      }
      // tile tE's rows of this lane: 16 j + l16 of the wave's rows
      const int64_t nbE = n_begin + tE * kWpTile + wave * kWsWaveRows + l16;
      bool validE[4];
      float yEv[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        validE[j] = tE >= 0 && nbE + 16 * j < n_end;
        yEv[j] = __shfl(yE, 16 * j + l16, 64);
      }
      float s32[4];
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) {
        const int b = ks & 1;
        read_w(b ^ 1, (ks + 1) % KS);   // W is stationary: k step 0 again after the last
#pragma unroll
        for (int pl = 0; pl < 2; ++pl)
#pragma unroll
          for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) mfma16(acc[i][j], wf[b][pl][i], xr[ks % PRE][j]);
#pragma unroll
        for (int j = 0; j < 4; ++j)
          xr[ks % PRE][j] = ks + PRE < KS ? *(const uint4*)(pc[j] + (ks + PRE) * 1024)
                                          : *(const uint4*)(pn[j] + (ks + PRE - KS) * 1024);
        // log-likelihood terms of tile tE: (chain block, row block) pairs
        // [16 ks / KS, 16 (ks + 1) / KS), row blocks innermost
#pragma unroll
        for (int pq = 16 * ks / KS; pq < 16 * (ks + 1) / KS; ++pq) {
          const int i = pq >> 2, j = pq & 3;
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            if (j == 0) s32[v] = 0.0f;
            const float nl = bern_nll_fast(accE[i][j][v], yEv[j]);
            s32[v] -= validE[j] ? nl : 0.0f;
            if (j == 3) lls[i][v] += (double)s32[v];
          }
        }
        // keep each k step's loads in their step: hoisting them across the
        // unrolled body would outgrow the register file
        __builtin_amdgcn_sched_barrier(0);
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          accE[i][j] = acc[i][j];
          acc[i][j] = f32x4v_t{0.0f, 0.0f, 0.0f, 0.0f};
        }
      yE = y_cur;
      y_cur = y_nx;
      tE = t;
    }
    // the last tile's terms
    const int64_t nbE = n_begin + tE * kWpTile + wave * kWsWaveRows + l16;
    bool validE[4];
    float yEv[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      validE[j] = nbE + 16 * j < n_end;
      yEv[j] = __shfl(yE, 16 * j + l16, 64);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float s32 = 0.0f;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float nl = bern_nll_fast(accE[i][j][v], yEv[j]);
          s32 -= validE[j] ? nl : 0.0f;
        }
        lls[i][v] += (double)s32;
      }
  }
  // lanes with equal lane >> 4 hold the same chains at different rows
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      double x = lls[i][v];
#pragma unroll
      for (int off = 1; off < 16; off <<= 1) x += __shfl_xor(x, off, 64);
      lls[i][v] = x;
    }
  __syncthreads();   // the W tiles are no longer read: reuse LDS for the wave sums
  double* red = reinterpret_cast<double*>(lds);   // [4 waves][64 chains]
  if (l16 == 0) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) red[wave * kWsBM + 16 * i + (lane >> 4) * 4 + v] = lls[i][v];
  }
  __syncthreads();
  if (tid < kWsBM) {
    double x = 0.0;
#pragma unroll
    for (int w = 0; w < kWpThreads / 64; ++w) x += red[w * kWsBM + tid];
    llpart[(int64_t)part * ll_ld + tm * kWsBM + tid] = x;
  }
}

// Whether the bound data suit the W-stationary kernels (and bind makes Xf):
// Bernoulli, bf16, D_pad <= 512, enough rows to split.
static bool ws_data_ok(const edhmc_ctx* ctx) {
  return ctx->cdt == EDHMC_BF16 && ctx->lik == EDHMC_LIK_BERNOULLI && ctx->D_pad <= kWsMaxDPad &&
         (ctx->D_pad / 32) % kWsPre == 0 && (ctx->D_pad / 32) % kWsPreGrad == 0 &&
         (ctx->D_pad / 32) % kWsPreGrad1 == 0 && ctx->N >= 16 * kWsTile;
}

// Whether the W-stationary forward applies: the data suit it (Xf made) and
// the chain tiles are a multiple of 8 (one slice per XCD).
static bool ws_forward_ok(const edhmc_ctx* ctx) {
#ifdef EDHMC_NO_WS_FORWARD
  return false;
#endif
  return ws_data_ok(ctx) && ctx->C_pad % (8 * kWsBM) == 0;
}

static int launch_ws_forward(edhmc_ctx* ctx, hipStream_t st, bool wr, int planes, bool ll) {
  // the one-plane pass on 128-chain tiles where the chain tiles still split
  // evenly over the XCDs (each X fragment feeds 8 chain blocks, not 4)
  const bool wide = wr && planes == 1 && !ll && kWsWide1 && ctx->C_pad % (8 * 2 * kWsBM) == 0;
  const int m_tiles = ctx->C_pad / (wide ? 2 * kWsBM : kWsBM);
  // about one workgroup per CU: split the rows of each chain tile into parts
  int parts = std::max(1, 256 / m_tiles);
  int64_t rows_per_part = (ctx->N + parts - 1) / parts;
  rows_per_part = (rows_per_part + kWsTile - 1) / kWsTile * kWsTile;
  parts = (int)((ctx->N + rows_per_part - 1) / rows_per_part);
  if (parts > ctx->n_tiles1) return set_err(ctx, EDHMC_EUNSUPPORTED, "forward: too many parts");
#ifndef EDHMC_WS_PART_PER_XCD
#define EDHMC_WS_PART_PER_XCD 1
#endif
  const bool part_per_xcd = EDHMC_WS_PART_PER_XCD && 8 % parts == 0;
  // chain tiles per XCD: all of them split over the XCDs sharing a part, or
  // an eighth of them over all parts
  const int mt_per_xcd = part_per_xcd ? m_tiles / (8 / parts) : m_tiles / 8;
  const int grid = m_tiles * parts;
  {
    ProfScope ps(ctx, wr ? PK_GEMM_GRAD : PK_GEMM_LL, st);
    if (wide)
      hipLaunchKernelGGL((gemm_ll_ws_kernel<true, 1, false, kWsPreGrad1, 8>), dim3(grid),
                         dim3(kWsThreads), 0, st, (const bf16_t*)ctx->z_lp.p,
                         (int64_t)ctx->C_pad * ctx->D_pad, ctx->D_pad, (const bf16_t*)ctx->Xf.p,
                         ctx->N, ctx->N_pad, rows_per_part, parts, mt_per_xcd,
                         part_per_xcd ? 1 : 0, (double*)ctx->llpart.p, (int64_t)ctx->C_pad,
                         (const float*)ctx->y.p, (bf16_t*)ctx->Rt.p, (int64_t)ctx->C_pad);
    else if (wr && planes == 1 && !ll)
      hipLaunchKernelGGL((gemm_ll_ws_kernel<true, 1, false, kWsPreGrad1>), dim3(grid),
                         dim3(kWsThreads), 0, st, (const bf16_t*)ctx->z_lp.p,
                         (int64_t)ctx->C_pad * ctx->D_pad, ctx->D_pad, (const bf16_t*)ctx->Xf.p,
                         ctx->N, ctx->N_pad, rows_per_part, parts, mt_per_xcd,
                         part_per_xcd ? 1 : 0, (double*)ctx->llpart.p, (int64_t)ctx->C_pad,
                         (const float*)ctx->y.p, (bf16_t*)ctx->Rt.p, (int64_t)ctx->C_pad);
    else if (wr && planes == 2 && !ll)   // exact inner gradients: no log-likelihood terms
      hipLaunchKernelGGL((gemm_ll_ws_kernel<true, 2, false, kWsPreGrad>), dim3(grid),
                         dim3(kWsThreads), 0, st, (const bf16_t*)ctx->z_lp.p,
                         (int64_t)ctx->C_pad * ctx->D_pad, ctx->D_pad, (const bf16_t*)ctx->Xf.p,
                         ctx->N, ctx->N_pad, rows_per_part, parts, mt_per_xcd,
                         part_per_xcd ? 1 : 0, (double*)ctx->llpart.p, (int64_t)ctx->C_pad,
                         (const float*)ctx->y.p, (bf16_t*)ctx->Rt.p, (int64_t)ctx->C_pad);
    else if (wr && planes == 2)
      hipLaunchKernelGGL((gemm_ll_ws_kernel<true, 2, true, kWsPreGrad>), dim3(grid),
                         dim3(kWsThreads), 0, st, (const bf16_t*)ctx->z_lp.p,
                         (int64_t)ctx->C_pad * ctx->D_pad, ctx->D_pad, (const bf16_t*)ctx->Xf.p,
                         ctx->N, ctx->N_pad, rows_per_part, parts, mt_per_xcd,
                         part_per_xcd ? 1 : 0, (double*)ctx->llpart.p, (int64_t)ctx->C_pad,
                         (const float*)ctx->y.p, (bf16_t*)ctx->Rt.p, (int64_t)ctx->C_pad);
    else if (!wr && planes == 2 && ll && ctx->D_pad % 128 == 0 && ctx->D_pad >= 256) {
      // the pipelined forward, its tile body unrolled for D_pad = 32 KS (at
      // 4096 chains, N = 1e6: 3.64 / 4.93 / 6.1 ms against the eight-wave
      // kernel's 3.71 / 5.12 / 6.6 for D_pad = 256 / 384 / 512; at 128 the
      // eight-wave kernel is 2 % faster and keeps it)
      auto wp = [&](auto ks) {
        constexpr int KS = decltype(ks)::value;
        hipLaunchKernelGGL((gemm_ll_wp_kernel<KS, 4>), dim3(grid), dim3(kWpThreads), 0, st,
                           (const bf16_t*)ctx->z_lp.p, (int64_t)ctx->C_pad * ctx->D_pad,
                           (const bf16_t*)ctx->Xf.p, ctx->N, ctx->N_pad, rows_per_part, parts,
                           mt_per_xcd, part_per_xcd ? 1 : 0, (double*)ctx->llpart.p,
                           (int64_t)ctx->C_pad, (const float*)ctx->y.p);
      };
      switch (ctx->D_pad) {
        case 256: wp(std::integral_constant<int, 8>{}); break;
        case 384: wp(std::integral_constant<int, 12>{}); break;
        default: wp(std::integral_constant<int, 16>{}); break;   // 512 (ws_data_ok)
      }
    } else if (!wr && planes == 2 && ll)
      hipLaunchKernelGGL((gemm_ll_ws_kernel<false, 2, true, kWsPre>), dim3(grid), dim3(kWsThreads), 0, st,
                         (const bf16_t*)ctx->z_lp.p, (int64_t)ctx->C_pad * ctx->D_pad,
                         ctx->D_pad, (const bf16_t*)ctx->Xf.p, ctx->N, ctx->N_pad,
                         rows_per_part, parts, mt_per_xcd, part_per_xcd ? 1 : 0,
                         (double*)ctx->llpart.p, (int64_t)ctx->C_pad, (const float*)ctx->y.p,
                         nullptr, (int64_t)0);
    else
      return set_err(ctx, EDHMC_EUNSUPPORTED, "W-stationary pass: no variant for %d planes",
                     planes);
  }
  EDHMC_TRY(check_launch(ctx, "gemm_ll_ws"));
  return ll ? ll_reduce(ctx, parts, st) : EDHMC_OK;
}

// Forward only (log-likelihood partials, no residuals) at z_lp, exact form.
int glm_forward_ll(edhmc_ctx* ctx, hipStream_t st) {
  if (ws_forward_ok(ctx)) EDHMC_TRY(launch_ws_forward(ctx, st, false));
  else if (ctx->cdt == EDHMC_BF16) EDHMC_TRY((launch_gemm1<bf16_t, CfgP2, 2, false, true>(ctx, st)));
  else EDHMC_TRY((launch_gemm1<float, Cfg1, 1, false, true>(ctx, st)));
  return EDHMC_OK;
}

// Recompute the carried log joint at z with the forward-only pass (the
// formula the surrogate's Metropolis test evaluates proposals with), so both
// ends of its ratio use one formula after plain transitions.
int glm_refresh_logp(edhmc_ctx* ctx, hipStream_t st) {
  EDHMC_TRY(glm_write_lp(ctx, st));
  EDHMC_TRY(glm_forward_ll(ctx, st));
  ChainArgs a = chain_args(ctx);
  hipLaunchKernelGGL(chain_relogp_kernel, dim3(ctx->C), dim3(kChainThreads), 0, st, a);
  return check_launch(ctx, "chain_relogp");
}

int glm_write_lp(edhmc_ctx* ctx, hipStream_t st) {
  ChainArgs a = chain_args(ctx);
  hipLaunchKernelGGL(chain_lp_kernel, dim3(ctx->C), dim3(kChainThreads), 0, st, a);
  return check_launch(ctx, "chain_lp");
}

static int ensure_chains(edhmc_ctx* ctx, int C);

// Posterior-predictive pass over the bound rows for the S draws in z_lp.
template <typename T, typename Cfg, int PLANES>
static int launch_predict(edhmc_ctx* ctx, int S, bool want_mean, float* rep, uint64_t seed,
                          hipStream_t st) {
  const int eb = sizeof(T);
  GemmShape g = {};
  g.A = (const char*)ctx->z_lp.p;
  g.lda_b = (int64_t)ctx->D_pad * eb;
  g.a_stage_b = Cfg::KS;
  g.a_plane_b = (int64_t)ctx->C_pad * ctx->D_pad * eb;
  g.B = (const char*)ctx->Xr.p;
  g.ldb_b = (int64_t)ctx->D_pad * eb;
  g.b_stage_b = Cfg::KS;
  g.m_tiles = ctx->C_pad / Cfg::BM;
  g.n_tiles = (int)(ctx->N_pad / Cfg::BN);
  g.k_stages = ctx->D_pad * eb / Cfg::KS;
  g.ksplit = 1;
  g.stages_per_split = g.k_stages;
  g.swz = 1;
  if (!gemm_offsets_fit<Cfg, PLANES>(g)) return set_err(ctx, EDHMC_EUNSUPPORTED, "tile too large");
  if (want_mean) EDHMC_TRY(dev_alloc(ctx, ctx->pred_col, (size_t)g.m_tiles * ctx->N_pad * 4, false));
  EpiPredict<T> e;
  e.y = (const float*)ctx->y.p;
  e.N = ctx->N;
  e.lp = lik_params(ctx);
  e.llpart = (double*)ctx->llpart.p;
  e.ll_ld = ctx->C_pad;
  e.colsum = want_mean ? (float*)ctx->pred_col.p : nullptr;
  e.col_ld = ctx->N_pad;
  e.rep = rep;
  e.m_valid = S;
  {
    ProfScope ps(ctx, PK_GEMM_LL, st);
    hipLaunchKernelGGL((gemm_nt_kernel<T, Cfg, PLANES, EpiPredict<T>>), dim3(grid_size(g)),
                       dim3(Cfg::kThreads), 0, st, g, e);
  }
  EDHMC_TRY(check_launch(ctx, "predict"));
  EDHMC_TRY(ll_reduce(ctx, g.n_tiles, st));
  if (rep) {
    const int64_t total = (int64_t)S * ctx->N;
    hipLaunchKernelGGL(replicate_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st,
                       rep, S, ctx->N, ctx->lik, ctx->lik_scale, seed);
    EDHMC_TRY(check_launch(ctx, "replicate"));
  }
  return EDHMC_OK;
}

static int ensure_chains(edhmc_ctx* ctx, int C) {
  if (C <= 0) return set_err(ctx, EDHMC_EINVAL, "number of chains must be positive, got %d", C);
  if (ctx->C == C && ctx->z.p) return EDHMC_OK;
  const int C_pad = (int)round_up(C, kChainPad);
  const size_t cd = (size_t)C * ctx->D_pad;
  const int eb = elem_bytes(ctx);
  ctx->C = C;
  ctx->C_pad = C_pad;
  ctx->state_valid = false;
  ctx->mom_on = false;
  ctx->mom_mean.release();
  ctx->mom_m2.release();
  EDHMC_TRY(dev_alloc(ctx, ctx->z, cd * 4));
  EDHMC_TRY(dev_alloc(ctx, ctx->r, cd * 4));
  EDHMC_TRY(dev_alloc(ctx, ctx->g, cd * 4));
  EDHMC_TRY(dev_alloc(ctx, ctx->z0, cd * 4));
  EDHMC_TRY(dev_alloc(ctx, ctx->g0, cd * 4));
  const int planes = ctx->cdt == EDHMC_BF16 ? 2 : 1;
  EDHMC_TRY(dev_alloc(ctx, ctx->z_lp, (size_t)planes * C_pad * ctx->D_pad * eb));
  EDHMC_TRY(dev_alloc(ctx, ctx->logp, (size_t)C * 8));
  EDHMC_TRY(dev_alloc(ctx, ctx->K0, (size_t)C * 8));
  EDHMC_TRY(dev_alloc(ctx, ctx->ndiv, (size_t)C * 4));
  EDHMC_TRY(dev_alloc(ctx, ctx->ratio, (size_t)C * 8));
  // rows up to a whole W-stationary tile: its residual epilogue stores the
  // last tile's rows past N_pad (zeros) without a bounds test
  EDHMC_TRY(dev_alloc(ctx, ctx->Rt, (size_t)C_pad * round_up(ctx->N_pad, kWsTile) * eb, false));
  // split-K for GEMM2: aim for >= 2048 workgroups, >= 8 stages per split.
  const int bn2 = (ctx->D_pad % Cfg1::BN == 0) ? Cfg1::BN : Cfg2S::BN;
  const int tiles2 = (C_pad / 128) * (ctx->D_pad / bn2);
  const int kst = (int)(ctx->N_pad * eb / kStageBytes);
  int ks = (2048 + tiles2 - 1) / tiles2;
  ks = std::max(1, std::min(ks, std::min(64, kst / 8)));
  if (ks >= 8) ks = ks / 8 * 8;
  ctx->ksplit = ks;
  EDHMC_TRY(dev_alloc(ctx, ctx->Gpart, (size_t)ks * C_pad * ctx->D_pad * 4, false));
  ctx->n_tiles1 = (int)(ctx->N_pad / std::min(Cfg1::BN, CfgP2::BN));   // llpart rows
  EDHMC_TRY(dev_alloc(ctx, ctx->llpart, (size_t)C_pad * ctx->n_tiles1 * 8, false));
  EDHMC_TRY(dev_alloc(ctx, ctx->ll, (size_t)C_pad * 8));
  EDHMC_TRY(dev_alloc(ctx, ctx->ll_split, (size_t)kLLSplit * C_pad * 8, false));
  return EDHMC_OK;
}

int glm_ensure_chains(edhmc_ctx* ctx, int C) { return ensure_chains(ctx, C); }

static int do_set_state(edhmc_ctx* ctx, const float* W, int C, hipStream_t st) {
  if (!ctx->bound) return set_err(ctx, EDHMC_ESTATE, "no model bound");
  if (!W) return set_err(ctx, EDHMC_EINVAL, "W is NULL");
  EDHMC_TRY(ensure_chains(ctx, C));
  EDHMC_TRY(dev_alloc(ctx, ctx->tmp_count, 4));
  ChainArgs a = chain_args(ctx);
  a.W_in = W;
  hipLaunchKernelGGL(chain_load_kernel, dim3(C), dim3(kChainThreads), 0, st, a,
                     (int*)ctx->tmp_count.p);
  EDHMC_TRY(check_launch(ctx, "chain_load"));
  int bad = 0;
  if (hipMemcpyAsync(&bad, ctx->tmp_count.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "set_state: copy failed");
  if (bad) {
    ctx->state_valid = false;
    return set_err(ctx, EDHMC_ENONFINITE, "chain positions contain %d non-finite rows", bad);
  }
  EDHMC_TRY(grad_eval(ctx, true, true, st));
  a = chain_args(ctx);
  hipLaunchKernelGGL(chain_eval_kernel, dim3(C), dim3(kChainThreads), 0, st, a);
  EDHMC_TRY(check_launch(ctx, "chain_eval"));
  ctx->state_valid = true;
  ctx->g_valid = true;
  surrogate_invalidate(ctx->sur);
  return EDHMC_OK;
}

static int do_transition(edhmc_ctx* ctx, float eps, int L, const float* r0, const double* u,
                         uint64_t seed, uint64_t counter, int32_t* accepted_out,
                         float* sample_out, int32_t* n_accept, hipStream_t st) {
  if (!ctx->bound || !ctx->state_valid)
    return set_err(ctx, EDHMC_ESTATE, "no chain state: call edhmc_set_state first");
  if (L < 1) return set_err(ctx, EDHMC_EINVAL, "n_steps must be >= 1, got %d", L);
  if (!(std::isfinite(eps))) return set_err(ctx, EDHMC_EINVAL, "step_size must be finite");
  if (!ctx->g_valid) {
    // the previous transitions were surrogate-driven: exact gradient at z
    EDHMC_TRY(glm_write_lp(ctx, st));
    EDHMC_TRY(grad_eval(ctx, true, true, st));
    ChainArgs e = chain_args(ctx);
    hipLaunchKernelGGL(chain_eval_kernel, dim3(ctx->C), dim3(kChainThreads), 0, st, e);
    EDHMC_TRY(check_launch(ctx, "chain_eval"));
    ctx->g_valid = true;
  }
  surrogate_invalidate(ctx->sur);
  ChainArgs a = chain_args(ctx);
  a.half = (float)(0.5 * (double)eps);
  a.full = eps;
  a.r_fed = r0;
  a.u_fed = u;
  a.seed = seed;
  a.counter = counter;
  {
    ProfScope ps(ctx, PK_CHAIN, st);
    hipLaunchKernelGGL(chain_begin_kernel, dim3(ctx->C), dim3(kChainThreads), 0, st, a);
  }
  EDHMC_TRY(check_launch(ctx, "chain_begin"));
  for (int k = 1; k <= L; ++k) {
    const bool last = k == L;
    // bf16: the inner gradients see the hi plane of the positions unless
    // exact inner gradients were asked for (edhmc_set_inner_gradient)
    EDHMC_TRY(grad_eval(ctx, last || ctx->inner_exact, last, st));
    if (!last) {
      a.rec_W = a.rec_G = nullptr;
      if (ctx->rec_W && k % ctx->rec_stride == 0 && ctx->rec_count + ctx->C <= ctx->rec_cap) {
        a.rec_W = ctx->rec_W + ctx->rec_count * ctx->D;
        a.rec_G = ctx->rec_G + ctx->rec_count * ctx->D;
        ctx->rec_count += ctx->C;
      }
      {
        ProfScope ps(ctx, PK_CHAIN, st);
        hipLaunchKernelGGL(chain_mid_kernel, dim3(ctx->C), dim3(kChainThreads), 0, st, a);
      }
      EDHMC_TRY(check_launch(ctx, "chain_mid"));
    }
  }
  a.sample_out = sample_out;
  a.accepted_out = accepted_out;
  a.n_accept = n_accept;
  if (ctx->mom_on) {
    ctx->mom_count += 1;
    a.mom_mean = (double*)ctx->mom_mean.p;
    a.mom_m2 = (double*)ctx->mom_m2.p;
    a.mom_n = ctx->mom_count;
  }
  {
    ProfScope ps(ctx, PK_CHAIN, st);
    hipLaunchKernelGGL(chain_last_kernel, dim3(ctx->C), dim3(kChainThreads), 0, st, a);
  }
  return check_launch(ctx, "chain_last");
}

}  // namespace edhmc

using namespace edhmc;

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry points: one file-scope device array,
// one copy per GPU (held until the process ends).  Host code never uses the
// array's name as an address: its device address on each GPU comes from
// hipGetSymbolAddress, taken once per device at edhmc_create with that device
// current, and only that address is handed to kernels.  Calls that overlap
// on one GPU share the array and race.
// ---------------------------------------------------------------------------
constexpr int kStatsChunks = 256;      // row chunks of the column statistics
constexpr int kNoAllocMaxD = 4096;     // widest X of edhmc_col_stats_noalloc
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[(size_t)kStatsChunks * kNoAllocMaxD * 2];
static std::mutex g_scratch_mu;
static double* g_scratch[kMaxDevices] = {};

static int scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return EDHMC_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[device]) return EDHMC_OK;
  void* p = nullptr;
  if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
    (void)hipGetLastError();
    return EDHMC_EHIP;
  }
  g_scratch[device] = (double*)p;
  return EDHMC_OK;
}

// Column sums and sums of squares per row chunk into part [chunks][D][2]
// (FP64), then mean and standard deviation (divisor P) in a fixed order.
static int col_stats_launch(edhmc_ctx* ctx, const float* X, int64_t P, int D, int chunks,
                            double* part, float* mean_out, float* std_out, hipStream_t st) {
  hipLaunchKernelGGL(col_stats_part_kernel, dim3((D + 255) / 256, chunks), dim3(256), 0, st, X, P,
                     D, chunks, part);
  EDHMC_TRY(check_launch(ctx, "col_stats"));
  hipLaunchKernelGGL(col_stats_final_kernel, dim3((D + 255) / 256), dim3(256), 0, st,
                     (const double*)part, P, D, chunks, mean_out, std_out);
  return check_launch(ctx, "col_stats_final");
}

edhmc_ctx::~edhmc_ctx() {
  for (hipEvent_t e : prof.pool) (void)hipEventDestroy(e);
  if (sur) surrogate_destroy(sur);
  if (hier) hier_destroy(hier);
  DevBuf* bufs[] = {&Xr, &Xc, &Xf, &y, &ploc, &pscale, &z, &r, &g, &z0, &g0, &z_lp, &logp, &K0,
                    &flag, &ndiv, &ratio, &Rt, &Gpart, &llpart, &ll, &mom_mean, &mom_m2, &tmp_count,
                    &stats_part, &pred_col, &ll_split, &pkind, &plnorm,
                    &newton};
  for (DevBuf* b : bufs) b->release();
}

extern "C" {

const char* edhmc_version(void) { return "edhmc 0.1 (gfx950)"; }

int edhmc_create(int device, edhmc_ctx** out) {
  if (!out) return EDHMC_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return EDHMC_EHIP;
  if (device < 0 || device >= n) return EDHMC_EINVAL;
  DeviceGuard dg(device);
  if (scratch_init(device) != EDHMC_OK) return EDHMC_EHIP;
  edhmc_ctx* c = new (std::nothrow) edhmc_ctx();
  if (!c) return EDHMC_ENOMEM;
  c->device = device;
  *out = c;
  return EDHMC_OK;
}

int edhmc_destroy(edhmc_ctx* ctx) {
  if (!ctx) return EDHMC_OK;
  DeviceGuard dg(ctx->device);
  (void)hipDeviceSynchronize();
  delete ctx;
  return EDHMC_OK;
}

const char* edhmc_last_error(const edhmc_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int edhmc_bind_glm(edhmc_ctx* ctx, const void* X, int32_t x_in_dtype, const float* y, int64_t N,
                   int32_t D, int32_t likelihood, float lik_scale, const float* prior_loc,
                   const float* prior_scale, int32_t compute_dtype, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (!X || !y || !prior_loc || !prior_scale) return set_err(ctx, EDHMC_EINVAL, "NULL argument");
  if (N <= 0 || D <= 0) return set_err(ctx, EDHMC_EINVAL, "need N > 0 and D > 0 (N=%lld D=%d)",
                                       (long long)N, D);
  if (x_in_dtype != EDHMC_F32 && x_in_dtype != EDHMC_BF16)
    return set_err(ctx, EDHMC_EINVAL, "bad x_in_dtype %d", x_in_dtype);
  if (compute_dtype != EDHMC_F32 && compute_dtype != EDHMC_BF16)
    return set_err(ctx, EDHMC_EINVAL, "bad compute_dtype %d", compute_dtype);
  if (likelihood != EDHMC_LIK_BERNOULLI && likelihood != EDHMC_LIK_NORMAL)
    return set_err(ctx, EDHMC_EUNSUPPORTED, "unsupported likelihood %d", likelihood);
  if (likelihood == EDHMC_LIK_NORMAL && !(lik_scale > 0.0f && std::isfinite(lik_scale)))
    return set_err(ctx, EDHMC_EINVAL, "likelihood scale must be positive and finite");
  if (D > 65536) return set_err(ctx, EDHMC_EUNSUPPORTED, "D=%d too large", D);
  ctx->bound = false;
  ctx->state_valid = false;
  ctx->C = 0;
  ctx->z.release();
  ctx->cdt = compute_dtype;
  ctx->lik = likelihood;
  ctx->lik_scale = lik_scale;
  ctx->lik_weight = 1.0f;
  ctx->has_kinds = false;
  ctx->N = N;
  ctx->N_pad = round_up(N, kRowPad);
  ctx->D = D;
  ctx->D_pad = (int)round_up(D, 64);   // K of GEMM1 in whole 128-byte stages
  const int eb = elem_bytes(ctx);
  EDHMC_TRY(dev_alloc(ctx, ctx->Xr, (size_t)ctx->N_pad * ctx->D_pad * eb, false));
  EDHMC_TRY(dev_alloc(ctx, ctx->Xc, (size_t)ctx->N_pad * ctx->D_pad * eb, false));
  EDHMC_TRY(dev_alloc(ctx, ctx->y, (size_t)ctx->N_pad * 4, false));
  EDHMC_TRY(dev_alloc(ctx, ctx->ploc, (size_t)ctx->D_pad * 4, false));
  EDHMC_TRY(dev_alloc(ctx, ctx->pscale, (size_t)ctx->D_pad * 4, false));
  EDHMC_TRY(dev_alloc(ctx, ctx->tmp_count, 4));
  int* bad = (int*)ctx->tmp_count.p;
  dim3 grid(ctx->D_pad / 32, (unsigned)(ctx->N_pad / 32));
  if (ctx->N_pad / 32 > 2147483647LL) return set_err(ctx, EDHMC_EUNSUPPORTED, "N too large");
  if (x_in_dtype == EDHMC_F32) {
    if (compute_dtype == EDHMC_F32)
      hipLaunchKernelGGL((layout_x_kernel<float, float>), grid, dim3(256), 0, st, (const float*)X,
                         N, D, (float*)ctx->Xr.p, (float*)ctx->Xc.p, ctx->N_pad, ctx->D_pad, bad);
    else
      hipLaunchKernelGGL((layout_x_kernel<float, bf16_t>), grid, dim3(256), 0, st,
                         (const float*)X, N, D, (bf16_t*)ctx->Xr.p, (bf16_t*)ctx->Xc.p,
                         ctx->N_pad, ctx->D_pad, bad);
  } else {
    if (compute_dtype == EDHMC_F32)
      hipLaunchKernelGGL((layout_x_kernel<bf16_t, float>), grid, dim3(256), 0, st,
                         (const bf16_t*)X, N, D, (float*)ctx->Xr.p, (float*)ctx->Xc.p,
                         ctx->N_pad, ctx->D_pad, bad);
    else
      hipLaunchKernelGGL((layout_x_kernel<bf16_t, bf16_t>), grid, dim3(256), 0, st,
                         (const bf16_t*)X, N, D, (bf16_t*)ctx->Xr.p, (bf16_t*)ctx->Xc.p,
                         ctx->N_pad, ctx->D_pad, bad);
  }
  EDHMC_TRY(check_launch(ctx, "layout_x"));
  ctx->Xf.release();
  if (ws_data_ok(ctx)) {
    const int64_t pieces = ctx->N_pad * ctx->D_pad / 8;
    EDHMC_TRY(dev_alloc(ctx, ctx->Xf, (size_t)ctx->N_pad * ctx->D_pad * 2, false));
    hipLaunchKernelGGL(layout_frag_kernel, dim3((unsigned)((pieces + 255) / 256)), dim3(256), 0,
                       st, (const bf16_t*)ctx->Xr.p, pieces, ctx->D_pad, (bf16_t*)ctx->Xf.p);
    EDHMC_TRY(check_launch(ctx, "layout_frag"));
  }
  auto padc = [&](const float* src, int64_t n, DevBuf& dst, int64_t npad, float fill) {
    hipLaunchKernelGGL(pad_copy_kernel, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0, st,
                       src, n, (float*)dst.p, npad, fill, bad);
    return check_launch(ctx, "pad_copy");
  };
  EDHMC_TRY(padc(y, N, ctx->y, ctx->N_pad, 0.0f));
  EDHMC_TRY(padc(prior_loc, D, ctx->ploc, ctx->D_pad, 0.0f));
  EDHMC_TRY(padc(prior_scale, D, ctx->pscale, ctx->D_pad, 1.0f));
  int nbad = 0;
  if (hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "bind: copy failed");
  if (nbad) return set_err(ctx, EDHMC_ENONFINITE, "data contains non-finite values (%d blocks)", nbad);
  ctx->bound = true;
  return EDHMC_OK;
}

int edhmc_set_state(edhmc_ctx* ctx, const float* W, int32_t C, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  return do_set_state(ctx, W, C, (hipStream_t)stream);
}

int edhmc_get_state(edhmc_ctx* ctx, float* W_out, double* logp_out, float* grad_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (!ctx->state_valid) return set_err(ctx, EDHMC_ESTATE, "no chain state");
  const int64_t n = (int64_t)ctx->C * ctx->D;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (W_out) {
    hipLaunchKernelGGL(copy_out_kernel, dim3(blocks), dim3(256), 0, st, (const float*)ctx->z.p,
                       ctx->C, ctx->D, ctx->D_pad, W_out);
    EDHMC_TRY(check_launch(ctx, "copy_out"));
  }
  if (grad_out) {
    hipLaunchKernelGGL(copy_out_kernel, dim3(blocks), dim3(256), 0, st, (const float*)ctx->g.p,
                       ctx->C, ctx->D, ctx->D_pad, grad_out);
    EDHMC_TRY(check_launch(ctx, "copy_out"));
  }
  if (logp_out && hipMemcpyAsync(logp_out, ctx->logp.p, (size_t)ctx->C * 8,
                                 hipMemcpyDeviceToDevice, st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "get_state copy failed");
  return EDHMC_OK;
}

int edhmc_get_momentum(edhmc_ctx* ctx, float* r_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!ctx->state_valid) return set_err(ctx, EDHMC_ESTATE, "no chain state");
  if (!r_out) return set_err(ctx, EDHMC_EINVAL, "r_out is NULL");
  const int64_t n = (int64_t)ctx->C * ctx->D;
  hipLaunchKernelGGL(copy_out_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, (const float*)ctx->r.p, ctx->C, ctx->D, ctx->D_pad,
                     r_out);
  return check_launch(ctx, "copy_out");
}

int edhmc_logp_grad(edhmc_ctx* ctx, const float* W, int32_t C, float* grad_out, double* logp_out,
                    void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(do_set_state(ctx, W, C, (hipStream_t)stream));
  return edhmc_get_state(ctx, nullptr, logp_out, grad_out, stream);
}

int edhmc_logp(edhmc_ctx* ctx, const float* W, int32_t C, double* logp_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (!ctx->bound) return set_err(ctx, EDHMC_ESTATE, "no model bound");
  if (!W || !logp_out) return set_err(ctx, EDHMC_EINVAL, "NULL argument");
  EDHMC_TRY(ensure_chains(ctx, C));
  EDHMC_TRY(dev_alloc(ctx, ctx->tmp_count, 4));
  ChainArgs a = chain_args(ctx);
  a.W_in = W;
  a.logp_out = logp_out;
  hipLaunchKernelGGL(lp_load_kernel, dim3(C), dim3(kChainThreads), 0, st, a,
                     (int*)ctx->tmp_count.p);
  EDHMC_TRY(check_launch(ctx, "lp_load"));
  int bad = 0;
  if (hipMemcpyAsync(&bad, ctx->tmp_count.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "logp: copy failed");
  if (bad) return set_err(ctx, EDHMC_ENONFINITE, "positions contain %d non-finite rows", bad);
  EDHMC_TRY(glm_forward_ll(ctx, st));
  hipLaunchKernelGGL(logp_out_kernel, dim3(C), dim3(kChainThreads), 0, st, a);
  return check_launch(ctx, "logp_out");
}

int edhmc_predict(edhmc_ctx* ctx, const float* W, int32_t S, float* mean_out, double* loglik_out,
                  float* rep_out, uint64_t seed, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (!ctx->bound) return set_err(ctx, EDHMC_ESTATE, "no model bound");
  if (!W) return set_err(ctx, EDHMC_EINVAL, "W is NULL");
  EDHMC_TRY(ensure_chains(ctx, S));
  EDHMC_TRY(dev_alloc(ctx, ctx->tmp_count, 4));
  ChainArgs a = chain_args(ctx);
  a.W_in = W;
  hipLaunchKernelGGL(lp_load_kernel, dim3(S), dim3(kChainThreads), 0, st, a,
                     (int*)ctx->tmp_count.p);
  EDHMC_TRY(check_launch(ctx, "lp_load"));
  int bad = 0;
  if (hipMemcpyAsync(&bad, ctx->tmp_count.p, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "predict: copy failed");
  if (bad) return set_err(ctx, EDHMC_ENONFINITE, "draws contain %d non-finite rows", bad);
  if (ctx->cdt == EDHMC_BF16)
    EDHMC_TRY((launch_predict<bf16_t, CfgP2, 2>(ctx, S, mean_out != nullptr, rep_out, seed, st)));
  else
    EDHMC_TRY((launch_predict<float, Cfg1, 1>(ctx, S, mean_out != nullptr, rep_out, seed, st)));
  if (mean_out) {
    const int m_tiles = ctx->C_pad / (ctx->cdt == EDHMC_BF16 ? CfgP2::BM : Cfg1::BM);
    hipLaunchKernelGGL(colsum_reduce_kernel, dim3((unsigned)((ctx->N + 255) / 256)), dim3(256), 0,
                       st, (const float*)ctx->pred_col.p, m_tiles, ctx->N_pad, ctx->N,
                       1.0f / (float)S, mean_out);
    EDHMC_TRY(check_launch(ctx, "colsum_reduce"));
  }
  if (loglik_out && hipMemcpyAsync(loglik_out, ctx->ll.p, (size_t)S * 8, hipMemcpyDeviceToDevice,
                                   st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "predict: copy failed");
  return EDHMC_OK;
}

int edhmc_transition(edhmc_ctx* ctx, float eps, int32_t n_steps, const float* r0, const double* u,
                     uint64_t seed, uint64_t counter, int32_t* accepted_out, float* sample_out,
                     int32_t* n_accept, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  return do_transition(ctx, eps, n_steps, r0, u, seed, counter, accepted_out, sample_out,
                       n_accept, (hipStream_t)stream);
}

static int sg_step(edhmc_ctx* ctx, bool hamiltonian, float lr, float friction, float* v,
                   uint64_t seed, uint64_t counter, float* sample_out, int32_t* n_accept,
                   hipStream_t st) {
  if (!ctx->bound || !ctx->state_valid)
    return set_err(ctx, EDHMC_ESTATE, "no chain state: call edhmc_set_state first");
  if (!(lr >= 0.0f) || !std::isfinite(lr)) return set_err(ctx, EDHMC_EINVAL, "bad learning rate");
  if (hamiltonian && (!v || !(friction >= 0.0f) || !std::isfinite(friction)))
    return set_err(ctx, EDHMC_EINVAL, "SGHMC needs a velocity buffer and friction >= 0");
  if (!ctx->g_valid) {
    // gradient of the (scaled) log joint at the current positions
    EDHMC_TRY(glm_write_lp(ctx, st));
    EDHMC_TRY(grad_eval(ctx, true, true, st));
    ChainArgs e = chain_args(ctx);
    hipLaunchKernelGGL(chain_eval_kernel, dim3(ctx->C), dim3(kChainThreads), 0, st, e);
    EDHMC_TRY(check_launch(ctx, "chain_eval"));
  }
  surrogate_invalidate(ctx->sur);
  ChainArgs a = chain_args(ctx);
  a.seed = seed;
  a.counter = counter;
  a.sample_out = sample_out;
  a.n_accept = n_accept;
  const float sd = hamiltonian ? sqrtf(lr * friction) : sqrtf(lr);
  {
    ProfScope ps(ctx, PK_CHAIN, st);
    if (hamiltonian)
      hipLaunchKernelGGL(sg_step_kernel<true>, dim3(ctx->C), dim3(kChainThreads), 0, st, a, lr,
                         friction, sd, v);
    else
      hipLaunchKernelGGL(sg_step_kernel<false>, dim3(ctx->C), dim3(kChainThreads), 0, st, a, lr,
                         0.0f, sd, (float*)nullptr);
  }
  EDHMC_TRY(check_launch(ctx, "sg_step"));
  ctx->g_valid = false;   // positions moved; the next step re-evaluates the gradient
  return EDHMC_OK;
}

static int mh_step(edhmc_ctx* ctx, const float* proposal_scale, const float* beta,
                   const double* u, uint64_t seed, uint64_t counter, int32_t* accepted_out,
                   float* sample_out, int32_t* n_accept, hipStream_t st);

int edhmc_mh_step(edhmc_ctx* ctx, const float* proposal_scale, const double* u, uint64_t seed,
                  uint64_t counter, int32_t* accepted_out, float* sample_out, int32_t* n_accept,
                  void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  return mh_step(ctx, proposal_scale, nullptr, u, seed, counter, accepted_out, sample_out,
                 n_accept, (hipStream_t)stream);
}

// Adjacent-replica exchange (replica_exchange_mc.py:214-237), for each of the
// C / R independent replica ladders (chain = replica * n_ladders + ladder):
// with probability exchange_freq, from a random start i0 in {0, 1}, pairs
// (i, i+1), (i+2, i+3), ... swap their states iff
// log u < (logp_i - logp_{i+1}) (beta_{i+1} - beta_i), all decided on the
// log joints before any swap.  Draws: Philox on the uniform stream of
// (seed, counter) with a = 1 + ladder: chain slot 0 gives the coin
// ((x0 >> 8) 2^-24) and the start (x1 & 1), slot 1 + i the pair's uniform
// (53 bits, as the accept uniforms).
__global__ void __launch_bounds__(kChainThreads)
replica_exchange_kernel(ChainArgs a, int R, int ladders, const float* __restrict__ beta,
                        float exchange_freq, int32_t* __restrict__ swaps) {
  const int l = blockIdx.x;
  __shared__ int do_swap[64];
  if (threadIdx.x < R) do_swap[threadIdx.x] = 0;
  __syncthreads();
  if (threadIdx.x == 0) {
    const u32x4 x = philox_draw(a.seed, a.counter, 1u + (uint32_t)l, 0u, 1);
    const float coin = (float)(x.x >> 8) * 5.9604644775390625e-8f;
    if (coin < exchange_freq) {
      for (int i = (int)(x.y & 1u); i + 1 < R; i += 2) {
        const int ci = i * ladders + l, cj = (i + 1) * ladders + l;
        const double ratio = (a.logp[ci] - a.logp[cj]) * ((double)beta[cj] - (double)beta[ci]);
        const u32x4 y = philox_draw(a.seed, a.counter, 1u + (uint32_t)l, 1u + (uint32_t)i, 1);
        const uint64_t mbits = ((uint64_t)y.x << 21) | (uint64_t)(y.y >> 11);
        const double u = (double)mbits * 1.1102230246251565e-16;   // 2^-53
        if (log(u) < ratio) do_swap[i] = 1;
      }
    }
  }
  __syncthreads();
  for (int i = 0; i + 1 < R; ++i) {
    if (!do_swap[i]) continue;
    const int ci = i * ladders + l, cj = (i + 1) * ladders + l;
    for (int d = threadIdx.x; d < a.D; d += kChainThreads) {
      const int64_t pi = (int64_t)ci * a.D_pad + d, pj = (int64_t)cj * a.D_pad + d;
      const float t = a.z[pi];
      a.z[pi] = a.z[pj];
      a.z[pj] = t;
    }
    if (threadIdx.x == 0) {
      const double t = a.logp[ci];
      a.logp[ci] = a.logp[cj];
      a.logp[cj] = t;
      if (swaps) swaps[l] += 1;
    }
  }
  __syncthreads();
  if (a.sample_out)
    for (int i = 0; i < R; ++i) {
      const int c = i * ladders + l;
      for (int d = threadIdx.x; d < a.D; d += kChainThreads)
        a.sample_out[(int64_t)c * a.D + d] = a.z[(int64_t)c * a.D_pad + d];
    }
}

int edhmc_replica_exchange_step(edhmc_ctx* ctx, const float* proposal_scale, const float* beta,
                                int32_t n_replica, float exchange_freq, uint64_t seed,
                                uint64_t counter, int32_t* accepted_out, float* sample_out,
                                int32_t* n_accept, int32_t* swaps, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (!beta) return set_err(ctx, EDHMC_EINVAL, "beta is NULL");
  if (n_replica < 1 || n_replica > 64 || !ctx->C || ctx->C % n_replica)
    return set_err(ctx, EDHMC_EINVAL, "chains must be n_replica (1..64) ladders of equal size");
  EDHMC_TRY(mh_step(ctx, proposal_scale, beta, nullptr, seed, counter, accepted_out, nullptr,
                    n_accept, st));
  ChainArgs a = chain_args(ctx);
  a.seed = seed;
  a.counter = counter;
  a.sample_out = sample_out;
  const int ladders = ctx->C / n_replica;
  {
    ProfScope ps(ctx, PK_CHAIN, st);
    hipLaunchKernelGGL(replica_exchange_kernel, dim3(ladders), dim3(kChainThreads), 0, st, a,
                       n_replica, ladders, beta, exchange_freq, swaps);
  }
  EDHMC_TRY(check_launch(ctx, "replica_exchange"));
  return EDHMC_OK;
}

static int mh_step(edhmc_ctx* ctx, const float* proposal_scale, const float* beta,
                   const double* u, uint64_t seed, uint64_t counter, int32_t* accepted_out,
                   float* sample_out, int32_t* n_accept, hipStream_t st) {
  if (!ctx->bound || !ctx->state_valid)
    return set_err(ctx, EDHMC_ESTATE, "no chain state: call edhmc_set_state first");
  if (!proposal_scale) return set_err(ctx, EDHMC_EINVAL, "proposal_scale is NULL");
  // the carried log joint must come from the forward formula the ratio uses
  if (ctx->g_valid) EDHMC_TRY(glm_refresh_logp(ctx, st));
  surrogate_invalidate(ctx->sur);
  ChainArgs a = chain_args(ctx);
  a.seed = seed;
  a.counter = counter;
  a.u_fed = u;
  a.beta = beta;
  {
    ProfScope ps(ctx, PK_CHAIN, st);
    hipLaunchKernelGGL(mh_propose_kernel, dim3(ctx->C), dim3(kChainThreads), 0, st, a,
                       proposal_scale);
  }
  EDHMC_TRY(check_launch(ctx, "mh_propose"));
  ctx->g_valid = false;
  EDHMC_TRY(glm_forward_ll(ctx, st));
  a.sample_out = sample_out;
  a.accepted_out = accepted_out;
  a.n_accept = n_accept;
  {
    ProfScope ps(ctx, PK_CHAIN, st);
    hipLaunchKernelGGL(mh_accept_kernel, dim3(ctx->C), dim3(kChainThreads), 0, st, a);
  }
  return check_launch(ctx, "mh_accept");
}

int edhmc_set_lik_weight(edhmc_ctx* ctx, float w) {
  if (!ctx) return EDHMC_EINVAL;
  if (!(w >= 0.0f) || !std::isfinite(w)) return set_err(ctx, EDHMC_EINVAL, "weight must be >= 0");
  ctx->lik_weight = w;
  ctx->g_valid = false;
  return EDHMC_OK;
}

__global__ void prior_family_kernel(const uint8_t* __restrict__ kind,
                                    const double* __restrict__ lnorm, int D, int D_pad,
                                    uint8_t* __restrict__ kind_out, double* __restrict__ lnorm_out,
                                    int* bad) {
  const int d = blockIdx.x * blockDim.x + threadIdx.x;
  if (d >= D_pad) return;
  uint8_t k = 0;
  double l = 0.0;
  if (d < D) {
    k = kind[d];
    l = lnorm[d];
    if ((k & 15) > kPriorInvGamma || (k >> 4) > kOpSigmoid || !isfinite(l)) atomicAdd(bad, 1);
  }
  kind_out[d] = k;
  lnorm_out[d] = l;
}

int edhmc_set_prior_family(edhmc_ctx* ctx, const uint8_t* kind, const double* lnorm,
                           void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!ctx->bound) return set_err(ctx, EDHMC_ESTATE, "no model bound");
  hipStream_t st = (hipStream_t)stream;
  ctx->g_valid = false;
  surrogate_invalidate(ctx->sur);
  if (!kind) {
    ctx->has_kinds = false;
    return EDHMC_OK;
  }
  if (!lnorm) return set_err(ctx, EDHMC_EINVAL, "lnorm is NULL");
  EDHMC_TRY(dev_alloc(ctx, ctx->pkind, (size_t)ctx->D_pad, false));
  EDHMC_TRY(dev_alloc(ctx, ctx->plnorm, (size_t)ctx->D_pad * 8, false));
  EDHMC_TRY(dev_alloc(ctx, ctx->tmp_count, 4));
  int* bad = (int*)ctx->tmp_count.p;
  hipLaunchKernelGGL(prior_family_kernel, dim3((ctx->D_pad + 255) / 256), dim3(256), 0, st, kind,
                     lnorm, ctx->D, ctx->D_pad, (uint8_t*)ctx->pkind.p, (double*)ctx->plnorm.p,
                     bad);
  EDHMC_TRY(check_launch(ctx, "prior_family"));
  int nbad = 0;
  if (hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "set_prior_family: copy failed");
  if (nbad) {
    ctx->has_kinds = false;
    return set_err(ctx, EDHMC_EINVAL, "unknown prior family/operand code or non-finite constant");
  }
  ctx->has_kinds = true;
  return EDHMC_OK;
}

int edhmc_sgld_step(edhmc_ctx* ctx, float learning_rate, uint64_t seed, uint64_t counter,
                    float* sample_out, int32_t* n_accept, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  return sg_step(ctx, false, learning_rate, 0.0f, nullptr, seed, counter, sample_out, n_accept,
                 (hipStream_t)stream);
}

int edhmc_sghmc_step(edhmc_ctx* ctx, float learning_rate, float friction, float* v,
                     uint64_t seed, uint64_t counter, float* sample_out, int32_t* n_accept,
                     void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  return sg_step(ctx, true, learning_rate, friction, v, seed, counter, sample_out, n_accept,
                 (hipStream_t)stream);
}

int edhmc_run(edhmc_ctx* ctx, float eps, int32_t n_steps, float* params, int64_t T, int64_t t0,
              int64_t n_iter, uint64_t seed, uint64_t counter0, int32_t* n_accept, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!params) return set_err(ctx, EDHMC_EINVAL, "params is NULL");
  if (t0 < 0 || n_iter < 0 || t0 + n_iter > T)
    return set_err(ctx, EDHMC_EINVAL, "iterations [%lld, %lld) outside params of %lld rows",
                   (long long)t0, (long long)(t0 + n_iter), (long long)T);
  for (int64_t t = t0; t < t0 + n_iter; ++t) {
    EDHMC_TRY(do_transition(ctx, eps, n_steps, nullptr, nullptr, seed, counter0 + (uint64_t)t,
                            nullptr, params + t * (int64_t)ctx->C * ctx->D, n_accept,
                            (hipStream_t)stream));
  }
  return EDHMC_OK;
}

int edhmc_profile_enable(edhmc_ctx* ctx, int32_t enable) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(prof_collect(ctx));
  ctx->prof.on = enable != 0;
  for (int k = 0; k < PK_COUNT; ++k) {
    ctx->prof.total_ms[k] = 0.0;
    ctx->prof.count[k] = 0;
  }
  return EDHMC_OK;
}

int edhmc_profile_select(edhmc_ctx* ctx, uint32_t class_mask) {
  if (!ctx) return EDHMC_EINVAL;
  ctx->prof.mask = class_mask;
  return EDHMC_OK;
}

int edhmc_profile_read(edhmc_ctx* ctx, double* total_ms_out, int64_t* count_out) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  EDHMC_TRY(prof_collect(ctx));
  for (int k = 0; k < PK_COUNT; ++k) {
    if (total_ms_out) total_ms_out[k] = ctx->prof.total_ms[k];
    if (count_out) count_out[k] = ctx->prof.count[k];
  }
  return EDHMC_OK;
}

int edhmc_moments_enable(edhmc_ctx* ctx, int32_t enable) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!enable) {
    ctx->mom_on = false;
    ctx->mom_mean.release();
    ctx->mom_m2.release();
    return EDHMC_OK;
  }
  if (!ctx->z.p) return set_err(ctx, EDHMC_ESTATE, "no chains");
  const size_t bytes = (size_t)ctx->C * ctx->D_pad * 8;
  EDHMC_TRY(dev_alloc(ctx, ctx->mom_mean, bytes));
  EDHMC_TRY(dev_alloc(ctx, ctx->mom_m2, bytes));
  ctx->mom_count = 0;
  ctx->mom_on = true;
  return EDHMC_OK;
}

int edhmc_moments_get(edhmc_ctx* ctx, double* mean_out, double* m2_out, int64_t* count_out,
                      void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (!ctx->mom_on) return set_err(ctx, EDHMC_ESTATE, "moments not enabled");
  const int64_t n = (int64_t)ctx->C * ctx->D;
  const unsigned blocks = (unsigned)((n + 255) / 256);
  if (mean_out) {
    hipLaunchKernelGGL(copy_out_f64_kernel, dim3(blocks), dim3(256), 0, st,
                       (const double*)ctx->mom_mean.p, ctx->C, ctx->D, ctx->D_pad, mean_out);
    EDHMC_TRY(check_launch(ctx, "copy_out_f64"));
  }
  if (m2_out) {
    hipLaunchKernelGGL(copy_out_f64_kernel, dim3(blocks), dim3(256), 0, st,
                       (const double*)ctx->mom_m2.p, ctx->C, ctx->D, ctx->D_pad, m2_out);
    EDHMC_TRY(check_launch(ctx, "copy_out_f64"));
  }
  if (count_out) *count_out = ctx->mom_count;
  return EDHMC_OK;
}

int edhmc_divergences(edhmc_ctx* ctx, int32_t* out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!ctx->ndiv.p || !out) return set_err(ctx, EDHMC_ESTATE, "no chains");
  if (hipMemcpyAsync(out, ctx->ndiv.p, (size_t)ctx->C * 4, hipMemcpyDeviceToDevice,
                     (hipStream_t)stream) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "copy failed");
  return EDHMC_OK;
}

int edhmc_set_inner_gradient(edhmc_ctx* ctx, int32_t exact) {
  if (!ctx) return EDHMC_EINVAL;
  if (exact != 0 && exact != 1) return set_err(ctx, EDHMC_EINVAL, "exact must be 0 or 1");
  ctx->inner_exact = exact != 0;
  return EDHMC_OK;
}

int edhmc_last_ratio(edhmc_ctx* ctx, double* ratio_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!ctx->ratio.p || !ratio_out) return set_err(ctx, EDHMC_ESTATE, "no chains");
  if (hipMemcpyAsync(ratio_out, ctx->ratio.p, (size_t)ctx->C * 8, hipMemcpyDeviceToDevice,
                     (hipStream_t)stream) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "copy failed");
  return EDHMC_OK;
}

int edhmc_draw_momenta(edhmc_ctx* ctx, int32_t C, int32_t D, uint64_t seed, uint64_t counter,
                       float* r_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (C <= 0 || D <= 0 || !r_out) return set_err(ctx, EDHMC_EINVAL, "bad arguments");
  hipLaunchKernelGGL(rng_momenta_kernel, dim3(C), dim3(128), 0, (hipStream_t)stream, C, D, seed,
                     counter, r_out);
  return check_launch(ctx, "rng_momenta");
}

int edhmc_draw_uniforms(edhmc_ctx* ctx, int32_t C, uint64_t seed, uint64_t counter, double* u_out,
                        void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (C <= 0 || !u_out) return set_err(ctx, EDHMC_EINVAL, "bad arguments");
  hipLaunchKernelGGL(rng_uniforms_kernel, dim3((C + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     C, seed, counter, u_out);
  return check_launch(ctx, "rng_uniforms");
}

int edhmc_sample_moments(edhmc_ctx* ctx, const void* params, int32_t params_dtype, int64_t T,
                         int64_t K, int64_t t0, double* mean_out, double* var_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!params || !mean_out || !var_out || T <= 0 || K <= 0 || t0 < 0 || t0 >= T)
    return set_err(ctx, EDHMC_EINVAL, "bad arguments (T=%lld K=%lld t0=%lld)", (long long)T,
                   (long long)K, (long long)t0);
  if (params_dtype != EDHMC_F32 && params_dtype != 2)
    return set_err(ctx, EDHMC_EINVAL, "params dtype must be fp32 or fp64");
  hipStream_t st = (hipStream_t)stream;
  const unsigned blocks = (unsigned)((K + 255) / 256);
  if (params_dtype == EDHMC_F32)
    hipLaunchKernelGGL(sample_moments_kernel<float>, dim3(blocks), dim3(256), 0, st,
                       (const float*)params, T, K, t0, mean_out, var_out);
  else
    hipLaunchKernelGGL(sample_moments_kernel<double>, dim3(blocks), dim3(256), 0, st,
                       (const double*)params, T, K, t0, mean_out, var_out);
  return check_launch(ctx, "sample_moments");
}

int edhmc_record_pairs(edhmc_ctx* ctx, float* W_buf, float* G_buf, int64_t capacity,
                       int32_t stride) {
  if (!ctx) return EDHMC_EINVAL;
  if (stride < 1) return set_err(ctx, EDHMC_EINVAL, "stride must be >= 1");
  ctx->rec_stride = stride;
  if (W_buf && (!G_buf || capacity <= 0))
    return set_err(ctx, EDHMC_EINVAL, "recording needs both buffers and a positive capacity");
  ctx->rec_W = W_buf;
  ctx->rec_G = W_buf ? G_buf : nullptr;
  ctx->rec_cap = W_buf ? capacity : 0;
  if (W_buf) ctx->rec_count = 0;   // stopping keeps the count readable
  return EDHMC_OK;
}

int64_t edhmc_record_count(edhmc_ctx* ctx) { return ctx ? ctx->rec_count : -1; }

int edhmc_col_stats(edhmc_ctx* ctx, const float* X, int64_t P, int32_t D, float* mean_out,
                    float* std_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!X || !mean_out || !std_out || P <= 0 || D <= 0)
    return set_err(ctx, EDHMC_EINVAL, "bad arguments");
  const int chunks = (int)std::min<int64_t>(kStatsChunks, P);
  EDHMC_TRY(dev_alloc(ctx, ctx->stats_part, (size_t)chunks * D * 16, false));
  return col_stats_launch(ctx, X, P, D, chunks, (double*)ctx->stats_part.p, mean_out, std_out,
                          (hipStream_t)stream);
}

int edhmc_col_stats_noalloc(edhmc_ctx* ctx, const float* X, int64_t P, int32_t D,
                            float* mean_out, float* std_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  if (!X || !mean_out || !std_out || P <= 0 || D <= 0)
    return set_err(ctx, EDHMC_EINVAL, "bad arguments");
  if (D > kNoAllocMaxD)
    return set_err(ctx, EDHMC_EINVAL,
                   "D=%d exceeds the %d columns the allocation-free scratch holds", D,
                   kNoAllocMaxD);
  double* part = g_scratch[ctx->device];
  if (!part) return set_err(ctx, EDHMC_ESTATE, "no scratch address for device %d", ctx->device);
  // every (chunk, column) slot the final pass reads is written by the first
  // pass of this call: nothing needs clearing
  const int chunks = (int)std::min<int64_t>(kStatsChunks, P);
  return col_stats_launch(ctx, X, P, D, chunks, part, mean_out, std_out, (hipStream_t)stream);
}

int edhmc_autocov(edhmc_ctx* ctx, const float* s, int64_t T, int32_t C, int32_t D, int32_t max_lag,
                  double* acov_out, double* chain_mean_out, void* stream) {
  if (!ctx) return EDHMC_EINVAL;
  DeviceGuard dg(ctx->device);
  hipStream_t st = (hipStream_t)stream;
  if (!s || !acov_out || T < 2 || C <= 0 || D <= 0 || max_lag < 0)
    return set_err(ctx, EDHMC_EINVAL, "bad arguments");
  if (max_lag >= T) max_lag = (int)(T - 1);
  if (hipMemsetAsync(acov_out, 0, (size_t)(max_lag + 1) * D * 8, st) != hipSuccess)
    return set_err(ctx, EDHMC_EHIP, "memset failed");
  const int64_t n = (int64_t)C * D;
  hipLaunchKernelGGL(autocov_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, s, T, C,
                     D, max_lag, acov_out, chain_mean_out);
  return check_launch(ctx, "autocov");
}

}  // extern "C"
