// Shared device helpers for the MI355X (gfx950) HMC engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace edhmc {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8_t;
typedef __attribute__((ext_vector_type(16))) float f32x16_t;
typedef __attribute__((ext_vector_type(4))) float f32x4_t;

// bf16 is carried as its raw 16-bit pattern.
typedef uint16_t bf16_t;

constexpr double kLog2Pi = 1.8378770664093454835606594728112;

// Round-to-nearest-even float -> bf16 (NaN kept NaN).  On the device this is
// gfx950's v_cvt_pk_bf16_f32, which gives the host formula's result for
// every finite input (all 2^32 bit patterns compared on an MI355X) and a NaN
// for every NaN (whose payload may differ from the host's 0x7fc0).
__device__ __host__ inline bf16_t f32_to_bf16(float f) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_bit_cast(bf16_t, (__bf16)f);
#else
  uint32_t u = __builtin_bit_cast(uint32_t, f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)0x7fc0;
  u += 0x7fffu + ((u >> 16) & 1u);
  return (bf16_t)(u >> 16);
#endif
}

// Two floats to the bf16 pair (lo | hi << 16), as f32_to_bf16 of each; on
// the device one v_cvt_pk_bf16_f32.
__device__ inline uint32_t f32x2_to_bf16x2(float lo, float hi) {
  typedef __attribute__((ext_vector_type(2))) float f2_t;
  typedef __attribute__((ext_vector_type(2))) __bf16 b2_t;
  const f2_t v = {lo, hi};
  return __builtin_bit_cast(uint32_t, __builtin_convertvector(v, b2_t));
}

__device__ __host__ inline float bf16_to_f32(bf16_t h) {
  return __builtin_bit_cast(float, (uint32_t)h << 16);
}

template <typename T> struct ElemOps;
template <> struct ElemOps<float> {
  static constexpr int kBytes = 4;
  __device__ static inline float from_f32(float v) { return v; }
  __device__ static inline float to_f32(float v) { return v; }
};
template <> struct ElemOps<bf16_t> {
  static constexpr int kBytes = 2;
  __device__ static inline bf16_t from_f32(float v) { return f32_to_bf16(v); }
  __device__ static inline float to_f32(bf16_t v) { return bf16_to_f32(v); }
};

// ---------------------------------------------------------------------------
// Block reductions in FP64 with a fixed combination order (deterministic).
// ---------------------------------------------------------------------------
__device__ inline double wave_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// All threads of the block must call; returns the block total to every thread.
template <int NT>
__device__ inline double block_sum_f64(double v, double* sm /* >= NT/64 */) {
  v = wave_sum_f64(v);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sm[wave] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; ++w) t += sm[w];
  return t;
}

// ---------------------------------------------------------------------------
// Likelihood terms of the generalised-linear models.
//   Bernoulli(logits=eta):  ll = y eta - softplus(eta),  dll/deta = y - sigmoid(eta)
//   Normal(loc=eta, s):     ll = -(y-eta)^2/(2 s^2) - log s - log(2 pi)/2,
//                           dll/deta = (y - eta)/s^2
// ---------------------------------------------------------------------------
enum { LIK_BERNOULLI = 0, LIK_NORMAL = 1 };

struct LikParams {
  int lik;
  float inv_s2;    // 1/s^2 (normal)
  double ll_const; // -log s - log(2 pi)/2 (normal)
};

// The epilogues call only native transcendentals (v_exp_f32, v_log_f32,
// v_rcp_f32): an out-of-line math call inside an epilogue would force the
// live accumulators to scratch memory.
__device__ inline float fast_rcp(float x) { return __builtin_amdgcn_rcpf(x); }

// log(1 + e) for e in [0, 1]: u = 1 + e rounds, and log(u) * e / (u - 1)
// cancels that rounding (u - 1 is exact); e below half an ulp of 1 gives e.
__device__ inline float log1p_unit(float e) {
  const float u = 1.0f + e;
  const float l = __logf(u);
  return u == 1.0f ? e : l * (e * fast_rcp(u - 1.0f));
}

// Negative Bernoulli log-likelihood softplus(eta) - y eta for y in [0, 1]:
//   log1p(exp(-|eta|)) + (max(eta, 0) - y eta),
// log1p(t) for t = 2^(-|eta| log2 e) as ln2 log2(1 + t) (absolute error about
// 2^-24: 1 + t rounds; once it rounds to 1 the dropped t is below 2^-24 too),
// the second term one fma (exact for y in {0, 1}: 0 or |eta|).  Every term
// uses the same rounded eta, so summed over the rows a rounding error d of
// eta enters through (sigmoid(eta) - y) d, whose weights average to zero over
// the data at the posterior -- not through sigmoid(eta) d > 0, as it would if
// the linear part w . X^T y were added exactly beside a sum of softplus(eta).
__device__ inline float bern_nll_fast(float eta, float y) {
  const float t = __builtin_amdgcn_exp2f(-fabsf(eta) * 1.4426950408889634f);
  return fmaf(__builtin_amdgcn_logf(1.0f + t), 0.6931471805599453f,
              fmaf(-y, eta, fmaxf(eta, 0.0f)));
}

__device__ inline float glm_dll(const LikParams& p, float eta, float y) {
  if (p.lik == LIK_BERNOULLI) {
    float e = __expf(-fabsf(eta));
    float inv = fast_rcp(1.0f + e);
    float sig = eta >= 0.0f ? inv : e * inv;
    return y - sig;
  }
  return (y - eta) * p.inv_s2;
}

__device__ inline double glm_ll(const LikParams& p, float eta, float y) {
  if (p.lik == LIK_BERNOULLI) {
    float e = __expf(-fabsf(eta));
    float sp = fmaxf(eta, 0.0f) + log1p_unit(e);
    return (double)(y * eta) - (double)sp;
  }
  float d = y - eta;
  return -0.5 * (double)(d * d) * (double)p.inv_s2 + p.ll_const;
}

__device__ inline void glm_both(const LikParams& p, float eta, float y,
                                float& dll, double& ll) {
  if (p.lik == LIK_BERNOULLI) {
    float e = __expf(-fabsf(eta));
    float inv = fast_rcp(1.0f + e);
    float sig = eta >= 0.0f ? inv : e * inv;
    dll = y - sig;
    float sp = fmaxf(eta, 0.0f) + log1p_unit(e);
    ll = (double)(y * eta) - (double)sp;
  } else {
    float d = y - eta;
    dll = d * p.inv_s2;
    ll = -0.5 * (double)(d * d) * (double)p.inv_s2 + p.ll_const;
  }
}

}  // namespace edhmc
