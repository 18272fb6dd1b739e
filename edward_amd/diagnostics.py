"""Effective sample size from device sample stores (no reference counterpart:
Edward reports only the acceptance rate, ``monte_carlo.py:152-158``).

The lag sums run on the GPU (``edhmc_autocov``); the truncation is done here.
``between=True`` (the default) is the multi-chain estimator of Gelman et al.
(BDA3, sec. 11.5): rho_t = 1 - (W - mean_m acov_m(t)) / var+, with var+ the
within-chain variance plus the spread of the chain means, summed with Geyer's
(1992) initial positive, monotone sequence, tau floored at 1/log10(M n),
ESS = M n / tau.  ``between=False`` leaves the spread of the chain means out,
so it credits only what the transitions themselves decorrelate: chains that
never move score 0 even when their starting points are spread over the
posterior."""

import math

import numpy as np
import torch

from edward_amd import _lib as L


def autocov(engine, samples, max_lag):
  """Sum over chains of each chain's biased autocovariance at lags
  0..max_lag, [max_lag+1, D] fp64, plus chain means [C, D]."""
  s = samples.contiguous()
  T, C, D = s.shape
  max_lag = int(min(max_lag, T - 1))
  acov = torch.empty(max_lag + 1, D, dtype=torch.float64, device=s.device)
  cmean = torch.empty(C, D, dtype=torch.float64, device=s.device)
  engine.check(engine.lib.edhmc_autocov(engine.h, L.ptr(s), T, C, D, max_lag, L.ptr(acov),
                                        L.ptr(cmean), engine.stream))
  return acov, cmean


def ess_from_acov(acov_sum, chain_means, n, between=False):
  """ESS per coordinate from summed autocovariances [lags, D] and chain means
  [M, D] of n draws per chain."""
  acov_sum = np.asarray(acov_sum, np.float64)
  chain_means = np.asarray(chain_means, np.float64)
  M, D = chain_means.shape
  lags = acov_sum.shape[0]
  acov_mean = acov_sum / M
  W = acov_mean[0] * n / (n - 1.0)
  var_plus = W * (n - 1.0) / n
  if between and M > 1:
    var_plus = var_plus + chain_means.var(axis=0, ddof=1)
  out = np.empty(D)
  for d in range(D):
    if not var_plus[d] > 0:
      out[d] = 0.0            # no movement: nothing was sampled
      continue
    rho = 1.0 - (W[d] - acov_mean[:, d]) / var_plus[d]
    rho[0] = 1.0
    total, prev, k = 0.0, float('inf'), 0
    truncated = False
    while 2 * k + 1 < lags:
      p = rho[2 * k] + rho[2 * k + 1]
      if p <= 0:
        truncated = True
        break
      p = min(p, prev)
      total += p
      prev = p
      k += 1
    if not truncated and 2 * k + 1 >= lags and lags < n:
      out[d] = float('nan')      # lag window too short: caller widens it
      continue
    tau = max(-1.0 + 2.0 * total, 1.0 / math.log10(M * n))
    out[d] = M * n / tau
  return out


def ess(engine, samples, max_lag=64, between=True):
  """Per-coordinate ESS of samples [T, C, D] (device, fp32); the lag window
  doubles until every coordinate's Geyer sequence is truncated."""
  T = samples.shape[0]
  if T < 4:
    return np.full(int(samples.shape[2]), np.nan)   # too few draws to estimate
  lag = min(max_lag, T - 1)
  while True:
    acov, cmean = autocov(engine, samples, lag)
    e = ess_from_acov(acov.cpu().numpy(), cmean.cpu().numpy(), T, between=between)
    if not np.any(np.isnan(e)) or lag >= T - 1:
      return np.where(np.isnan(e), 0.0, e)
    lag = min(2 * lag + 1, T - 1)


def min_ess(engine, samples, max_lag=64, between=True):
  return float(np.min(ess(engine, samples, max_lag, between)))
