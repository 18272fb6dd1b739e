// Constrained priors on unconstrained positions (auto-transform).
//
// Edward's HMC samples constrained latents in an unconstrained space
// (edward/inferences/inference.py:222-264, util/random_variables.py:856-917):
// a latent theta with support [0, inf) is theta = softplus(u), one with
// support (0, 1) is theta = sigmoid(u), and the log joint gains the log
// Jacobian of u -> theta (hmc.py:132-159):
//   log p_u(u) = log p_theta(T(u)) + log T'(u).
// Each coordinate d carries a kind byte: the low nibble names the prior
// family in the unconstrained space, the high nibble the operand that the
// linear predictor receives (theta itself, or u when the likelihood takes
// logit(theta) = u, as Bernoulli(probs=theta) with a sigmoid map does).
// p1, p2 are the family's two parameters; lnorm is its normalising constant,
// computed once on the host in FP64.
//
//   family                 log p_u(u) - lnorm                      d/du
//   0 Normal(p1, p2) on u  -(u-p1)^2 / 2 p2^2                      -(u-p1)/p2^2
//   1 Gamma(a=p1, b=p2)    (a-1) log th - b th + log s             ((a-1)/th - b) s + 1 - s
//     at th = softplus(u), s = sigmoid(u) = T'(u)
//   2 Beta(a=p1, b=p2)     a log s + b log(1-s)                    a (1-s) - b s
//     at th = s = sigmoid(u)  (T'(u) = s (1-s) folded in)
//   3 InvGamma(a=p1, b=p2) -(a+1) log th - b/th + log s            (-(a+1)/th + b/th^2) s + 1 - s
//     at th = softplus(u)
//
// Everything runs in FP64: these terms are O(C D) per gradient evaluation,
// next to the O(N D C) matrix products.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace edhmc {

enum : int { kPriorNormal = 0, kPriorGamma = 1, kPriorBeta = 2, kPriorInvGamma = 3 };
enum : int { kOpIdentity = 0, kOpSoftplus = 1, kOpSigmoid = 2 };

__device__ inline double softplus_f64(double x) {
  return x > 0.0 ? x + log1p(exp(-x)) : log1p(exp(x));
}

__device__ inline double sigmoid_f64(double x) {
  if (x >= 0.0) return 1.0 / (1.0 + exp(-x));
  const double e = exp(x);
  return e / (1.0 + e);
}

// log softplus(u); below u = -30, softplus(u) = e^u (1 - e^u/2 + ...) and
// log softplus(u) = u to 1e-13 (log of an underflowed softplus would be -inf).
__device__ inline double log_softplus_f64(double u) {
  return u < -30.0 ? u : log(softplus_f64(u));
}

// sigmoid(u) / softplus(u), which tends to 1 as u -> -inf.
__device__ inline double sig_over_softplus_f64(double u) {
  return u < -30.0 ? 1.0 : sigmoid_f64(u) / softplus_f64(u);
}

struct PriorTerm {
  double lp;    // log density of u (constants included)
  double dlp;   // its derivative
};

__device__ inline PriorTerm prior_term(int family, double u, double p1, double p2, double lnorm) {
  PriorTerm t;
  switch (family) {
    case kPriorGamma: {
      const double s = sigmoid_f64(u), th = softplus_f64(u);
      t.lp = (p1 - 1.0) * log_softplus_f64(u) - p2 * th - softplus_f64(-u) + lnorm;
      t.dlp = (p1 - 1.0) * sig_over_softplus_f64(u) - p2 * s + (1.0 - s);
      break;
    }
    case kPriorBeta: {
      const double s = sigmoid_f64(u);
      t.lp = -p1 * softplus_f64(-u) - p2 * softplus_f64(u) + lnorm;
      t.dlp = p1 * (1.0 - s) - p2 * s;
      break;
    }
    case kPriorInvGamma: {
      const double s = sigmoid_f64(u), th = softplus_f64(u), q = sig_over_softplus_f64(u);
      t.lp = -(p1 + 1.0) * log_softplus_f64(u) - p2 / th - softplus_f64(-u) + lnorm;
      t.dlp = -(p1 + 1.0) * q + p2 * q / th + (1.0 - s);
      break;
    }
    default: {
      const double z = (u - p1) / p2;
      t.lp = -0.5 * z * z + lnorm;
      t.dlp = -z / p2;
      break;
    }
  }
  return t;
}

// The operand the linear predictor receives, and its derivative.
__device__ inline double operand_f64(int op, double u) {
  switch (op) {
    case kOpSoftplus: return softplus_f64(u);
    case kOpSigmoid: return sigmoid_f64(u);
    default: return u;
  }
}

__device__ inline double d_operand_f64(int op, double u) {
  switch (op) {
    case kOpSoftplus: return sigmoid_f64(u);
    case kOpSigmoid: {
      const double s = sigmoid_f64(u);
      return s * (1.0 - s);
    }
    default: return 1.0;
  }
}

}  // namespace edhmc
