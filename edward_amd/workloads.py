"""Synthetic workloads of the logistic-regression HMC benchmarks.

Host-generated with numpy's PCG64 (``default_rng(seed)``) so the CPU oracle
and the GPU see identical inputs.

  logreg_small  N=1000,  D=10,  1 chain,  eps=0.01,  L=10, T=5000, fp32
  logreg_mid    N=1e5,   D=256, 1024 chains, eps=0.002, L=20, T=2000, fp32
  logreg_large  N=1e6,   D=512, 4096 chains, eps=5e-4,  L=32, T=1000, X in bf16
  hier_groups   G=2048 groups of 1+Poisson(400) rows, D=64, 512 chains, L=16

The small workload follows Edward's logistic-regression example
(examples/bayesian_logistic_regression.py:41-48: Bernoulli(logits=ed.dot(X, w)),
Normal prior, Empirical storage); the paper's benchmark program
(docs/tex/iclr2017.tex:241-265) has the same model at N=581012, D=54.
"""

import numpy as np


def bf16_round(a):
  """Round float32 values to the nearest bf16 (ties to even), as float32."""
  a = np.ascontiguousarray(a, np.float32)
  u = a.view(np.uint32).astype(np.uint64)
  nan = (u & 0x7FFFFFFF) > 0x7F800000
  u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
  out = u.astype(np.uint32).view(np.float32).copy()
  out[nan] = np.nan
  return out


def _sigmoid(x):
  return 0.5 * (1.0 + np.tanh(0.5 * x))


class Workload(object):
  def __init__(self, **kw):
    self.__dict__.update(kw)

  def describe(self):
    return {k: v for k, v in self.__dict__.items()
            if isinstance(v, (int, float, str, bool))}


def logreg(N, D, C, eps, L, T, seed=0, scaled_columns=True, w_sd=0.1, bf16=False,
           init='normal', chunk=1 << 17):
  """Bayesian logistic regression with X ~ N(0,1) (columns scaled by U(0.5, 2)
  when ``scaled_columns``), w* ~ N(0, w_sd^2 I), y ~ Bernoulli(sigmoid(X w*)),
  prior w ~ N(0, I).  X is generated in row chunks as float32; with ``bf16``
  it is rounded to bf16 (the oracle then sees the rounded values)."""
  rng = np.random.default_rng(seed)
  scale = rng.uniform(0.5, 2.0, size=D).astype(np.float32) if scaled_columns \
      else np.ones(D, np.float32)
  w_true = (rng.normal(size=D) * w_sd).astype(np.float64)
  X = np.empty((N, D), np.float32)
  for s in range(0, N, chunk):
    e = min(N, s + chunk)
    X[s:e] = rng.standard_normal((e - s, D), dtype=np.float32) * scale
    if bf16:
      X[s:e] = bf16_round(X[s:e])
  eta = np.empty(N)
  for s in range(0, N, chunk):
    e = min(N, s + chunk)
    eta[s:e] = X[s:e].astype(np.float64) @ w_true
  y = (rng.random(N) < _sigmoid(eta)).astype(np.float32)
  if init == 'zeros':
    W0 = np.zeros((C, D), np.float32)
  else:
    W0 = rng.normal(size=(C, D)).astype(np.float32)
  return Workload(name='logreg', N=N, D=D, C=C, step_size=eps, n_steps=L, T=T, seed=seed,
                  X=X, y=y, w_true=w_true, W0=W0, prior_loc=np.zeros(D, np.float32),
                  prior_scale=np.ones(D, np.float32), bf16=bf16)


def logreg_small(seed=0):
  """N=1000, D=10, 1 chain started at 0 (Empirical(zeros) convention)."""
  rng = np.random.default_rng(seed)
  N, D = 1000, 10
  X = rng.standard_normal((N, D)).astype(np.float32)
  w_true = rng.normal(size=D)
  y = (rng.random(N) < _sigmoid(X.astype(np.float64) @ w_true)).astype(np.float32)
  return Workload(name='logreg_small', N=N, D=D, C=1, step_size=0.01, n_steps=10, T=5000,
                  seed=seed, X=X, y=y, w_true=w_true, W0=np.zeros((1, D), np.float32),
                  prior_loc=np.zeros(D, np.float32), prior_scale=np.ones(D, np.float32),
                  bf16=False)


def logreg_mid(seed=0, N=100000, D=256, C=1024):
  return logreg(N, D, C, 0.002, 20, 2000, seed=seed)


def logreg_large(seed=0, N=1000000, D=512, C=4096):
  return logreg(N, D, C, 5e-4, 32, 1000, seed=seed, bf16=True)


def hier_groups(seed=0, G=2048, D=64, C=512, mean_size=400):
  """Hierarchical logistic regression: group sizes 1 + Poisson(mean_size),
  rows sorted by group, X ~ N(0,1), mu ~ N(0,1), log sigma ~ N(0,1),
  a_g ~ N(mu, sigma^2), w ~ N(0, I), y ~ Bernoulli(a_g + x w)."""
  rng = np.random.default_rng(seed)
  sizes = 1 + rng.poisson(mean_size, size=G)
  offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
  N = int(offsets[-1])
  mu = rng.normal()
  log_sigma = rng.normal()
  a = rng.normal(mu, np.exp(log_sigma), size=G)
  w = rng.normal(size=D)
  X = rng.standard_normal((N, D)).astype(np.float32)
  group = np.repeat(np.arange(G), sizes)
  eta = a[group] + X.astype(np.float64) @ w
  y = (rng.random(N) < _sigmoid(eta)).astype(np.float32)
  dim = G + 2 + D
  # chains start near the generating values (the posterior of w has sd ~1e-3
  # at this N, so a cold start spends hundreds of transitions in transit)
  truth = np.concatenate([a, [mu, log_sigma], w])
  theta0 = (truth[None, :] + 0.002 * rng.normal(size=(C, dim))).astype(np.float32)
  return Workload(name='hier_groups', G=G, D=D, N=N, C=C, dim=dim, n_steps=16, T=1000,
                  step_size=0.001, seed=seed, X=X, y=y, offsets=offsets, theta0=theta0,
                  truth=dict(mu=mu, log_sigma=log_sigma, a=a, w=w))
