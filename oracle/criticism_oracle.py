"""FP64 restatement of Edward's criticism functions for the GLM family.

Test infrastructure only (like the rest of ``oracle/``): the tests and the
benchmark's CPU baseline call it as the checker; the product path never does.

Follows ``edward/criticisms/evaluate.py`` and ``edward/criticisms/ppc.py``:
posterior draws w_s (one ``Empirical.sample()`` per draw, uniform over the
stored rows, ``models/empirical.py:98-110``) are pushed through the
likelihood; supervised metrics compare ``y_true`` with the prediction built
from the draws (``evaluate.py:128-150``) and ``log_lik`` averages the mean
row log density over the draws (``evaluate.py:199-204``).  The replicated
data sets of ``ppc.py:100-121`` use the engine's Philox stream so that the
device and this restatement draw the same y_rep.
"""

import math

import numpy as np

from oracle import hmc_oracle as O


def predictive(glm, W):
  """Mean response averaged over the draws W [S, D] (sigmoid(eta) for
  Bernoulli, eta for Normal), [N]; log-likelihood of glm.y per draw, [S]."""
  W = np.atleast_2d(np.asarray(W, np.float64))
  eta = glm.eta(W)
  ll = glm.loglik_terms(eta).sum(axis=1)
  mean = O.sigmoid(eta) if glm.likelihood == 'bernoulli' else eta
  return mean.mean(axis=0), ll, mean


def replicate(seed, mean, likelihood, scale=1.0):
  """y_rep [S, N] from mean responses [S, N]: Philox4x32-10 keyed by seed,
  counter (s, n) on the uniform stream (ctr = (n, 0, s_lo, s_hi | 2^31)).
  Bernoulli: y = 1 iff (x0 >> 8) 2^-24 < mu.  Normal: mu + scale z0 with z0
  the Box-Muller normal of (x0, x1) (as the momentum draw)."""
  mean = np.atleast_2d(np.asarray(mean, np.float64))
  S, N = mean.shape
  out = np.empty((S, N))
  nn = np.arange(N, dtype=np.uint64)
  for s in range(S):
    x = O.philox4x32_10(O._ctr(nn, np.zeros(N, np.uint64), s, O.STREAM_UNIFORM),
                        np.broadcast_to(O._key(seed), (N, 2)))
    a = np.floor(x[:, 0].astype(np.float64) / 256.0)
    if likelihood == 'bernoulli':
      mu32 = mean[s].astype(np.float32).astype(np.float64)
      out[s] = (a * 2.0 ** -24 < mu32).astype(np.float64)
    else:
      b = np.floor(x[:, 1].astype(np.float64) / 256.0)
      u1 = (a + 1.0) * 2.0 ** -24
      th = 2.0 * math.pi * b * 2.0 ** -24
      out[s] = mean[s] + scale * np.sqrt(-2.0 * np.log(u1)) * np.cos(th)
  return out


def binary_prediction(probs, rng):
  """``evaluate.py:136-141``: round the averaged probabilities, with a random
  prediction where they are exactly 0.5."""
  probs = np.asarray(probs, np.float64)
  tie = probs == 0.5
  return np.where(tie, np.round(rng.random(probs.shape)), np.round(probs))


def metric(name, y_true, y_pred):
  """The metric functions of ``evaluate.py:235-476`` on 1-D inputs (macro
  averaging of ``util/metrics.py:10-47`` reduces to a mean over elements)."""
  y_true = np.asarray(y_true, np.float64)
  y_pred = np.asarray(y_pred, np.float64)
  if name == 'binary_accuracy':
    return float(np.mean(y_true == y_pred))
  if name in ('log_loss', 'binary_crossentropy'):
    # sigmoid cross entropy with y_pred as logits (evaluate.py:290-303)
    return float(np.mean(np.maximum(y_pred, 0) - y_pred * y_true +
                         np.log1p(np.exp(-np.abs(y_pred)))))
  if name in ('mse', 'MSE', 'mean_squared_error'):
    return float(np.mean((y_pred - y_true) ** 2))
  if name in ('mae', 'MAE', 'mean_absolute_error'):
    return float(np.mean(np.abs(y_pred - y_true)))
  if name in ('mape', 'MAPE', 'mean_absolute_percentage_error'):
    return float(100.0 * np.mean(np.abs((y_true - y_pred) / np.clip(np.abs(y_true), 1e-8, None))))
  if name in ('msle', 'MSLE', 'mean_squared_logarithmic_error'):
    return float(np.mean((np.log(np.clip(y_pred, 1e-8, None) + 1.0) -
                          np.log(np.clip(y_true, 1e-8, None) + 1.0)) ** 2))
  if name == 'hinge':
    return float(np.mean(np.maximum(1.0 - y_true * y_pred, 0.0)))
  if name == 'squared_hinge':
    return float(np.mean(np.maximum(1.0 - y_true * y_pred, 0.0) ** 2))
  if name == 'poisson':
    return float(np.sum(y_pred - y_true * np.log(y_pred + 1e-8)))
  if name in ('cosine', 'cosine_proximity'):
    # tf.nn.l2_normalize: x / sqrt(max(sum x^2, 1e-12))
    a = y_true / math.sqrt(max(float(np.sum(y_true ** 2)), 1e-12))
    b = y_pred / math.sqrt(max(float(np.sum(y_pred ** 2)), 1e-12))
    return float(np.sum(a * b))
  raise NotImplementedError(name)
