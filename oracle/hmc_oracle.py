"""FP64 NumPy restatement of Edward's HMC path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this module.  It is the checker the HIP engine is compared
against, never part of the product path.

Reference files restated here (paths relative to the Edward 1.3.5 tree):
  * ``edward/inferences/hmc.py:61-130``  transition: state gather, momentum,
    leapfrog, ratio, accept, scatter, ``n_accept``.
  * ``edward/inferences/hmc.py:161-192`` log joint = sum of prior and
    likelihood ``log_prob`` terms, normalising constants included,
    ``scale`` ignored.
  * ``edward/inferences/hmc.py:195-210`` kick-drift-kick leapfrog with L+1
    gradient evaluations.
  * ``edward/inferences/monte_carlo.py:95-150`` / ``inference.py:97-154``
    run loop: iteration t reads ``params[max(t-1, 0)]`` and writes
    ``params[t]``.
  * ``edward/models/empirical.py:87-96`` mean over axis 0, population (/T)
    standard deviation.
  * TensorFlow 1.5 ``Normal.log_prob`` and ``Bernoulli(logits).log_prob`` =
    ``-sigmoid_cross_entropy_with_logits`` (the third-party arithmetic the
    reference calls at ``hmc.py:185,190``), restated from their closed forms.

Build-only parts (no reference counterpart) restate published methods:
Philox4x32-10 (Salmon et al., SC'11), Box-Muller (1958), the multi-chain
effective sample size of Gelman et al. BDA3 sec. 11.5 with Geyer's (1992)
initial monotone sequence, and the hierarchical logistic log joint of the
batched-groups workload.

Pinning (what each part is checked against, see tests/test_oracle.py):
  leapfrog            closed-form Gaussian step matrix, reversibility,
                      volume preservation, O(eps^2) energy error.
  glm_logp_grad       central finite differences, brute-force loops on tiny
                      inputs, eta = 0 and large-|eta| special cases,
                      scipy.stats log-pdfs.
  hier_logp_grad      finite differences, brute-force loops.
  GLM.grad_bf16       the exact gradient where every residual is a bf16
                      number (eta = 0), the bf16 bound elsewhere.
  posterior_moments_is  quadrature of a skewed 2-D logistic posterior.
  hmc_transition      accept rule special cases (ratio >= 0, NaN, u = 0),
                      detailed balance on a 1-D target (exact posterior).
  run_hmc             index semantics (row 0 overwritten), n_accept.
  empirical moments   numpy reductions with /T.
  philox4x32_10       Random123 known-answer vectors.
  ess                 AR(1) closed form n (1 - rho) / (1 + rho); iid draws.
"""

import math
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

LOG_2PI = math.log(2.0 * math.pi)

# ----------------------------------------------------------------------------
# Distribution log-densities (TF 1.5 closed forms)
# ----------------------------------------------------------------------------


def normal_logp(z, loc, scale):
  """``Normal.log_prob``: -0.5 ((z - loc) / scale)^2 - log scale - 0.5 log 2pi.

  Used by ``hmc.py:185`` for the prior term of every latent variable.
  """
  z = np.asarray(z, np.float64)
  return -0.5 * ((z - loc) / scale) ** 2 - np.log(scale) - 0.5 * LOG_2PI


def normal_logp_grad(z, loc, scale):
  """d/dz of :func:`normal_logp`."""
  return -(np.asarray(z, np.float64) - loc) / np.asarray(scale, np.float64) ** 2


def softplus(x):
  """log(1 + e^x), overflow-free."""
  x = np.asarray(x, np.float64)
  return np.maximum(x, 0.0) + np.log1p(np.exp(-np.abs(x)))


def sigmoid(x):
  """1 / (1 + e^-x), overflow-free."""
  x = np.asarray(x, np.float64)
  e = np.exp(-np.abs(x))
  return np.where(x >= 0, 1.0 / (1.0 + e), e / (1.0 + e))


def bernoulli_logits_logp(y, eta):
  """``Bernoulli(logits=eta).log_prob(y)`` = -sigmoid_cross_entropy(y, eta).

  TF writes the cross entropy as max(eta, 0) - eta y + log1p(e^-|eta|); the
  negative of that equals y eta - softplus(eta).  Used by ``hmc.py:190``.
  """
  return np.asarray(y, np.float64) * eta - softplus(eta)


# ----------------------------------------------------------------------------
# Generalised-linear log joints
# ----------------------------------------------------------------------------


class GLM(object):
  """Log joint of ``y ~ Lik(dot(X, w))`` with ``w ~ Normal(loc, scale)``.

  ``likelihood`` is ``'bernoulli'`` (logits link; the logistic-regression
  workload, ``examples/bayesian_logistic_regression.py:41-44``) or
  ``'normal'`` with standard deviation ``lik_scale`` (the linear-regression
  model of ``tests/inferences/hmc_test.py:63-66``).  The log joint follows
  ``hmc.py:182-192``: sum over latent priors plus sum over the likelihood,
  constants included.  A bias term is expressed by a column of ones in X and
  one more latent coordinate with its own prior.
  """

  def __init__(self, X, y, prior_loc=0.0, prior_scale=1.0,
               likelihood='bernoulli', lik_scale=1.0):
    self.X = np.asarray(X, np.float64)
    self.y = np.asarray(y, np.float64)
    N, D = self.X.shape
    self.N, self.D = N, D
    self.prior_loc = np.broadcast_to(np.asarray(prior_loc, np.float64), (D,))
    self.prior_scale = np.broadcast_to(np.asarray(prior_scale, np.float64),
                                       (D,))
    if likelihood not in ('bernoulli', 'normal'):
      raise ValueError(likelihood)
    self.likelihood = likelihood
    self.lik_scale = float(lik_scale)

  def eta(self, W):
    return np.atleast_2d(np.asarray(W, np.float64)) @ self.X.T

  def loglik_terms(self, eta):
    if self.likelihood == 'bernoulli':
      return bernoulli_logits_logp(self.y, eta)
    return normal_logp(self.y, eta, self.lik_scale)

  def dloglik_deta(self, eta):
    if self.likelihood == 'bernoulli':
      return self.y - sigmoid(eta)
    return (self.y - eta) / self.lik_scale ** 2

  def logp_grad(self, W):
    """W [C, D] -> (logp [C], grad [C, D]) in FP64."""
    W = np.atleast_2d(np.asarray(W, np.float64))
    eta = self.eta(W)
    ll = self.loglik_terms(eta).sum(axis=1)
    lp = normal_logp(W, self.prior_loc, self.prior_scale).sum(axis=1)
    grad = self.dloglik_deta(eta) @ self.X + normal_logp_grad(
        W, self.prior_loc, self.prior_scale)
    return ll + lp, grad

  def logp(self, W):
    """Log joint only, [C]: the same sums as ``logp_grad`` with the
    elementwise likelihood terms evaluated over row blocks in threads (numpy
    releases the GIL in its ufuncs), blocks added in a fixed order."""
    W = np.atleast_2d(np.asarray(W, np.float64))
    eta = self.eta(W)
    n = self.N
    edges = np.linspace(0, n, min(n, 4 * (os.cpu_count() or 1)) + 1).astype(np.int64)

    def block(k):
      s, e = edges[k], edges[k + 1]
      if self.likelihood == 'bernoulli':
        t = bernoulli_logits_logp(self.y[s:e], eta[:, s:e])
      else:
        t = normal_logp(self.y[s:e], eta[:, s:e], self.lik_scale)
      return t.sum(axis=1)
    with ThreadPoolExecutor(os.cpu_count() or 1) as ex:
      parts = list(ex.map(block, range(len(edges) - 1)))
    ll = np.sum(parts, axis=0)
    return ll + normal_logp(W, self.prior_loc, self.prior_scale).sum(axis=1)

  def neg_hessian(self, w):
    """-d^2 logp / dw^2 at one position w [D]: X^T diag(s) X + diag(1 /
    scale^2), s the likelihood's curvature in eta (sigma (1 - sigma) for
    Bernoulli, 1 / lik_scale^2 for Normal)."""
    eta = self.X @ np.asarray(w, np.float64)
    if self.likelihood == 'bernoulli':
      p = sigmoid(eta)
      s = p * (1 - p)
    else:
      s = np.full(self.N, 1.0 / self.lik_scale ** 2)
    return (self.X * s[:, None]).T @ self.X + np.diag(1.0 / self.prior_scale ** 2)

  def loglik(self, W):
    """Likelihood part only (no prior), [C]."""
    return self.loglik_terms(self.eta(W)).sum(axis=1)

  def grad_bf16(self, W, planes=2):
    """The gradient with the roundings of the bf16 engine's gradient pass
    (no reference counterpart; it restates the precision choice of the
    device kernels so their trajectories can be followed closely): the
    position enters eta as its bf16 hi plane (planes=1, the inner leapfrog
    steps) or hi + lo planes (planes=2), the residuals y - sigmoid(eta) are
    rounded to bf16 before X^T r, the prior term uses the fp32 position.
    Products and sums in FP64."""
    W = np.atleast_2d(np.asarray(W, np.float64))
    w32 = W.astype(np.float32)
    hi = _bf16(w32)
    pos = hi if planes == 1 else hi + _bf16((w32 - hi.astype(np.float32)).astype(np.float32))
    eta = pos @ self.X.T
    res = _bf16(self.dloglik_deta(eta))
    return res @ self.X + normal_logp_grad(w32.astype(np.float64), self.prior_loc,
                                           self.prior_scale)


def _bf16(a):
  """Nearest bf16 number (ties to even) of the float32 value of a, as float64."""
  a = np.ascontiguousarray(np.asarray(a, np.float64).astype(np.float32))
  u = a.view(np.uint32).astype(np.uint64)
  u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
  return u.astype(np.uint32).view(np.float32).astype(np.float64)


class GaussianTarget(object):
  """Independent Gaussian N(mean, sd^2) per coordinate.

  Its leapfrog map is linear with a closed form (:func:`gaussian_leapfrog_matrix`),
  which pins the integrator.
  """

  def __init__(self, mean, sd):
    self.mean = np.asarray(mean, np.float64)
    self.sd = np.asarray(sd, np.float64)

  def logp_grad(self, W):
    W = np.atleast_2d(np.asarray(W, np.float64))
    return (normal_logp(W, self.mean, self.sd).sum(axis=1),
            normal_logp_grad(W, self.mean, self.sd))


class HierLogReg(object):
  """Batched hierarchical logistic regression (G groups, D shared features).

  Latent vector theta = [a (G), mu, log_sigma, w (D)]:
    a_g ~ Normal(mu, exp(log_sigma)), mu ~ Normal(0, 1),
    log_sigma ~ Normal(0, 1), w ~ Normal(0, 1),
    y_n ~ Bernoulli(logits = a_{g(n)} + x_n . w).
  In Edward this is ``Bernoulli(logits=tf.gather(a, group) + ed.dot(X, w))``;
  the gather form is the one ``tests/inferences/hmc_test.py:105-123`` builds
  (IndexedSlices gradient).  Rows are sorted by group, ``offsets[G+1]`` gives
  each group's row range.  The log joint is ``hmc.py:182-192``'s sum.
  """

  def __init__(self, X, y, offsets):
    self.X = np.asarray(X, np.float64)
    self.y = np.asarray(y, np.float64)
    self.offsets = np.asarray(offsets, np.int64)
    self.G = len(self.offsets) - 1
    self.N, self.D = self.X.shape
    assert self.offsets[0] == 0 and self.offsets[-1] == self.N
    self.group = np.repeat(np.arange(self.G), np.diff(self.offsets))
    self.dim = self.G + 2 + self.D

  def split(self, theta):
    theta = np.atleast_2d(np.asarray(theta, np.float64))
    G = self.G
    return theta[:, :G], theta[:, G], theta[:, G + 1], theta[:, G + 2:]

  def logp_grad(self, theta):
    a, mu, ls, w = self.split(theta)
    sig = np.exp(ls)
    eta = a[:, self.group] + w @ self.X.T
    ll = bernoulli_logits_logp(self.y, eta).sum(axis=1)
    res = self.y - sigmoid(eta)
    za = (a - mu[:, None]) / sig[:, None]
    lp = (normal_logp(a, mu[:, None], sig[:, None]).sum(axis=1)
          + normal_logp(mu, 0.0, 1.0) + normal_logp(ls, 0.0, 1.0)
          + normal_logp(w, 0.0, 1.0).sum(axis=1))
    C = a.shape[0]
    ga = np.zeros_like(a)
    for c in range(C):
      ga[c] = np.add.reduceat(res[c], self.offsets[:-1]) if self.N else 0.0
    # empty groups: reduceat returns the next element; zero them
    empty = np.diff(self.offsets) == 0
    ga[:, empty] = 0.0
    ga += -za / sig[:, None]
    gmu = (za / sig[:, None]).sum(axis=1) - mu
    gls = (-1.0 + za ** 2).sum(axis=1) - ls
    gw = res @ self.X - w
    grad = np.concatenate([ga, gmu[:, None], gls[:, None], gw], axis=1)
    return ll + lp, grad


# ----------------------------------------------------------------------------
# Leapfrog, transition, run loop
# ----------------------------------------------------------------------------


def leapfrog(z, r, eps, n_steps, logp_grad):
  """Kick-drift-kick leapfrog of ``hmc.py:195-210``.

  The gradient is evaluated at the start and after every drift (L+1
  evaluations, ``hmc.py:199,206``); no momentum flip.  Returns
  ``(zL, rL, logp(zL), grad(zL))``; the last two are the values the final
  gradient evaluation produced.
  """
  z = np.array(z, np.float64, copy=True)
  r = np.array(r, np.float64, copy=True)
  lp, g = logp_grad(z)
  for _ in range(n_steps):
    r = r + 0.5 * eps * g
    z = z + eps * r
    lp, g = logp_grad(z)
    r = r + 0.5 * eps * g
  return z, r, lp, g


def accept_rule(log_u, ratio):
  """``accept = log(u) < ratio`` (``hmc.py:108-109``).

  ratio >= 0 accepts whenever u < 1, u = 0 (log u = -inf) accepts any finite
  or +inf ratio, and a NaN ratio never accepts.
  """
  with np.errstate(invalid='ignore'):
    return np.asarray(log_u) < np.asarray(ratio)


def hmc_transition(z0, r0, u, eps, n_steps, logp_grad):
  """One HMC transition per chain with fed momenta and uniforms.

  Follows ``hmc.py:81-126``: leapfrog from (z0, r0), then
  ratio = 0.5|r0|^2 - 0.5|rL|^2 + logp(zL) - logp(z0), accept iff
  log(u) < ratio, keep zL or z0.  z0, r0: [C, D]; u: [C].
  """
  z0 = np.atleast_2d(np.asarray(z0, np.float64))
  r0 = np.atleast_2d(np.asarray(r0, np.float64))
  lp0, _ = logp_grad(z0)
  zL, rL, lpL, gL = leapfrog(z0, r0, eps, n_steps, logp_grad)
  ratio = (0.5 * (r0 ** 2).sum(axis=1) - 0.5 * (rL ** 2).sum(axis=1)
           + lpL - lp0)
  with np.errstate(divide='ignore'):
    log_u = np.log(np.asarray(u, np.float64))
  accept = accept_rule(log_u, ratio)
  z_new = np.where(accept[:, None], zL, z0)
  return dict(z=z_new, accept=accept, ratio=ratio, log_u=log_u, zL=zL, rL=rL,
              logpL=lpL, logp0=lp0)


def run_hmc(params, momenta, uniforms, eps, n_steps, logp_grad, t0=0,
            n_iter=None):
  """The ``Inference.run`` / ``MonteCarlo.update`` loop over a sample store.

  ``params`` [T, C, D] is updated in place.  Iteration t reads
  ``params[max(t-1, 0)]`` (``hmc.py:81-85``) and writes ``params[t]``
  (``hmc.py:121-126``), so row 0's initial value is overwritten by the first
  sample.  ``momenta[t]`` [C, D] and ``uniforms[t]`` [C] are the fed
  randomness.  Returns ``n_accept`` [C] (``hmc.py:129``).
  """
  T = params.shape[0]
  if n_iter is None:
    n_iter = T - t0
  C = params.shape[1]
  n_accept = np.zeros(C, np.int64)
  for t in range(t0, t0 + n_iter):
    z0 = params[max(t - 1, 0)]
    out = hmc_transition(z0, momenta[t], uniforms[t], eps, n_steps, logp_grad)
    params[t] = out['z']
    n_accept += out['accept']
  return n_accept


def gaussian_leapfrog_matrix(eps, sd, n_steps=1):
  """Exact linear map (z - mean, r) -> (z' - mean, r') of L leapfrog steps on
  a 1-D Gaussian N(mean, sd^2):

      one step = [[1 - e^2/(2 s^2),              e              ],
                  [-(e/s^2)(1 - e^2/(4 s^2)),    1 - e^2/(2 s^2)]]
  """
  a = 1.0 - eps ** 2 / (2.0 * sd ** 2)
  M = np.array([[a, eps], [-(eps / sd ** 2) * (1.0 - eps ** 2 / (4.0 * sd ** 2)),
                           a]])
  return np.linalg.matrix_power(M, n_steps)


# ----------------------------------------------------------------------------
# Empirical storage readout (edward/models/empirical.py:87-96)
# ----------------------------------------------------------------------------


def empirical_mean(params):
  """``Empirical.mean`` = reduce_mean over axis 0 (``empirical.py:87-88``)."""
  return np.asarray(params, np.float64).mean(axis=0)


def empirical_stddev(params):
  """``Empirical.stddev`` = sqrt(mean_t (p - mean)^2), divisor T
  (``empirical.py:90-93``)."""
  p = np.asarray(params, np.float64)
  return np.sqrt(((p - p.mean(axis=0)) ** 2).mean(axis=0))


def empirical_variance(params):
  """``Empirical.variance`` = stddev^2 (``empirical.py:95-96``)."""
  return empirical_stddev(params) ** 2


# ----------------------------------------------------------------------------
# Philox4x32-10 counter-based RNG (Salmon et al. 2011)
# ----------------------------------------------------------------------------

PHILOX_M0 = 0xD2511F53
PHILOX_M1 = 0xCD9E8D57
PHILOX_W0 = 0x9E3779B9
PHILOX_W1 = 0xBB67AE85
_MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(ctr, key):
  """Philox4x32 with 10 rounds.  ctr [..., 4], key [..., 2] (uint32) ->
  [..., 4] uint32.  Round: (hi0, lo0) = M0 * c0, (hi1, lo1) = M1 * c2,
  c' = (hi1 ^ c1 ^ k0, lo1, hi0 ^ c3 ^ k1, lo0); the key is bumped by the
  Weyl constants between rounds."""
  c = np.asarray(ctr, np.uint64) & _MASK32
  k = np.asarray(key, np.uint64) & _MASK32
  c0, c1, c2, c3 = c[..., 0], c[..., 1], c[..., 2], c[..., 3]
  k0, k1 = k[..., 0].copy(), k[..., 1].copy()
  for rnd in range(10):
    p0 = np.uint64(PHILOX_M0) * c0
    p1 = np.uint64(PHILOX_M1) * c2
    hi0, lo0 = p0 >> np.uint64(32), p0 & _MASK32
    hi1, lo1 = p1 >> np.uint64(32), p1 & _MASK32
    c0, c1, c2, c3 = (hi1 ^ c1 ^ k0), lo1, (hi0 ^ c3 ^ k1), lo0
    if rnd < 9:
      k0 = (k0 + np.uint64(PHILOX_W0)) & _MASK32
      k1 = (k1 + np.uint64(PHILOX_W1)) & _MASK32
  return np.stack([c0, c1, c2, c3], axis=-1).astype(np.uint32)


STREAM_MOMENTUM = 0
STREAM_UNIFORM = 1


def _key(seed):
  seed = int(seed) & 0xFFFFFFFFFFFFFFFF
  return np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint64)


def _ctr(a, b, counter, stream):
  counter = int(counter) & 0x7FFFFFFFFFFFFFFF
  lo = counter & 0xFFFFFFFF
  hi = (counter >> 32) | (0x80000000 if stream == STREAM_UNIFORM else 0)
  a, b = np.broadcast_arrays(np.asarray(a, np.uint64), np.asarray(b, np.uint64))
  return np.stack([a, b, np.full(a.shape, lo, np.uint64),
                   np.full(a.shape, hi, np.uint64)], axis=-1)


def momentum_normals(seed, counter, C, D):
  """The engine's momentum draw for transition ``counter``: r [C, D] ~ N(0, I).

  Chain c, coordinates 4j..4j+3 come from Philox(ctr=(j, c, counter_lo,
  counter_hi), key=seed); each 32-bit pair (x0, x1) gives, by Box-Muller,
  u1 = ((x0 >> 8) + 1) 2^-24, theta = 2 pi (x1 >> 8) 2^-24,
  (sqrt(-2 ln u1) cos theta, sqrt(-2 ln u1) sin theta).
  """
  nb = (D + 3) // 4
  jj, cc = np.meshgrid(np.arange(nb), np.arange(C))
  x = philox4x32_10(_ctr(jj, cc, counter, STREAM_MOMENTUM),
                    np.broadcast_to(_key(seed), jj.shape + (2,)))
  x = x.astype(np.float64)
  out = np.empty((C, nb, 4))
  for p in range(2):
    a = np.floor(x[..., 2 * p] / 256.0)
    b = np.floor(x[..., 2 * p + 1] / 256.0)
    u1 = (a + 1.0) * 2.0 ** -24
    th = 2.0 * math.pi * b * 2.0 ** -24
    rad = np.sqrt(-2.0 * np.log(u1))
    out[..., 2 * p] = rad * np.cos(th)
    out[..., 2 * p + 1] = rad * np.sin(th)
  return out.reshape(C, nb * 4)[:, :D]


def accept_uniforms(seed, counter, C):
  """The engine's accept uniform for transition ``counter``: u [C] in [0, 1)
  with 53 random bits, u = (x0 2^21 + (x1 >> 11)) 2^-53 from
  Philox(ctr=(0, c, counter_lo, counter_hi | 2^31), key=seed)."""
  cc = np.arange(C)
  x = philox4x32_10(_ctr(np.zeros(C, np.uint64), cc, counter, STREAM_UNIFORM),
                    np.broadcast_to(_key(seed), (C, 2)))
  x0 = x[:, 0].astype(np.uint64)
  x1 = x[:, 1].astype(np.uint64)
  m = (x0 << np.uint64(21)) | (x1 >> np.uint64(11))
  return m.astype(np.float64) * 2.0 ** -53


# ----------------------------------------------------------------------------
# Stochastic-gradient samplers
# ----------------------------------------------------------------------------


def sgld_learning_rate(step_size, t):
  """``sgld.py:66-67``: step_size / (t + 1)^0.55 (t = the update's index)."""
  return step_size / float(t + 1) ** 0.55


def sgld_step(z, grad, lr, xi):
  """``sgld.py:70-80``: z + lr/2 grad + sqrt(lr) xi."""
  return z + 0.5 * lr * grad + math.sqrt(lr) * xi


def sghmc_step(z, v, grad, lr, friction, xi):
  """``sghmc.py:73-86``: (z + v, (1 - friction/2) v + lr grad +
  sqrt(lr friction) xi) with lr = step_size * 0.01 given."""
  return z + v, (1.0 - 0.5 * friction) * v + lr * grad + math.sqrt(lr * friction) * xi


def mh_step(z, logp_z, scale, xi, u, logp_fn, beta=1.0):
  """Random-walk Metropolis-Hastings (``metropolis_hastings.py:97-156``) with
  proposal N(z, diag(scale^2)): z' = z + scale xi; the symmetric proposal
  terms cancel, ratio = beta (logp(z') - logp(z)) (beta = the replica's
  inverse temperature, ``replica_exchange_mc.py:200``; 1 for plain MH);
  accept iff log u < ratio.  Returns (z_new, logp_new, accept, ratio)."""
  zp = z + scale * xi
  lp = logp_fn(zp)
  ratio = (lp - logp_z) * beta
  with np.errstate(divide='ignore'):
    acc = accept_rule(np.log(u), ratio)
  return (np.where(acc[:, None], zp, z), np.where(acc, lp, logp_z), acc, ratio)


def replica_exchange(logp, betas, ladders, seed, counter, exchange_freq):
  """The exchange move of ``replica_exchange_mc.py:214-237`` for chains
  ordered replica-major (chain = replica * ladders + ladder): per ladder, a
  coin u0 = (x0 >> 8) 2^-24 < exchange_freq and a start i0 = x1 & 1 from
  Philox(ctr=(1 + ladder, 0, counter, stream 1)); pair (i, i+1) swaps iff
  log u < (logp_i - logp_{i+1}) (beta_{i+1} - beta_i) with u the 53-bit
  uniform of Philox(ctr=(1 + ladder, 1 + i, counter, stream 1)).  Returns the
  permutation new_index[chain] (the chain whose state each chain takes)."""
  R = len(betas)
  perm = np.arange(R * ladders)
  ll = np.arange(ladders, dtype=np.uint64)
  keys = np.broadcast_to(_key(seed), (ladders, 2))
  x = philox4x32_10(_ctr(1 + ll, np.zeros(ladders, np.uint64), counter, STREAM_UNIFORM), keys)
  active = np.floor(x[:, 0].astype(np.float64) / 256.0) * 2.0 ** -24 < exchange_freq
  start = (x[:, 1] & 1).astype(np.int64)
  for i in range(R - 1):
    y = philox4x32_10(_ctr(1 + ll, np.full(ladders, 1 + i, np.uint64), counter,
                           STREAM_UNIFORM), keys)
    u = ((y[:, 0].astype(np.uint64) << np.uint64(21)) |
         (y[:, 1].astype(np.uint64) >> np.uint64(11))).astype(np.float64) * 2.0 ** -53
    ci = i * ladders + np.arange(ladders)
    cj = ci + ladders
    ratio = (logp[ci] - logp[cj]) * (betas[i + 1] - betas[i])
    with np.errstate(divide='ignore'):
      swap = active & ((i - start) % 2 == 0) & (i >= start) & (np.log(u) < ratio)
    perm[ci[swap]], perm[cj[swap]] = cj[swap], ci[swap]
  return perm


def scaled_logp_grad(logp_grad_lik, prior_grad, lik_weight, prior_weight):
  """Gradient of ``inference.py:187-193``'s scaled log joint: lik_weight times
  the likelihood gradient plus prior_weight times the prior gradient."""
  return lik_weight * logp_grad_lik + prior_weight * prior_grad


# ----------------------------------------------------------------------------
# Effective sample size
# ----------------------------------------------------------------------------


def autocovariance(x):
  """Biased (/n) autocovariance of each row of x [M, n], by FFT."""
  x = np.asarray(x, np.float64)
  n = x.shape[-1]
  xc = x - x.mean(axis=-1, keepdims=True)
  nfft = 1 << int(math.ceil(math.log2(2 * n)))
  f = np.fft.rfft(xc, nfft, axis=-1)
  ac = np.fft.irfft(f * np.conj(f), nfft, axis=-1)[..., :n]
  return ac / n


def ess(x, between=True):
  """Multi-chain effective sample size of draws x [M chains, n draws].

  Gelman et al. (BDA3, sec. 11.5): rho_t = 1 - (W - mean_m acov_m(t)) / var+,
  with W the mean within-chain variance and var+ = (n-1)/n W + B/n.  The sum
  of autocorrelations is truncated by Geyer's (1992) initial positive
  sequence of pair sums rho_{2k} + rho_{2k+1}, made monotone, and
  tau = -1 + 2 sum_k P_k, floored at 1/log10(M n).  ESS = M n / tau.
  With between=False the between-chain term B/n is left out, so rho_t is the
  pooled within-chain autocorrelation and chains that never move give 0.
  """
  x = np.atleast_2d(np.asarray(x, np.float64))
  M, n = x.shape
  if n < 4:
    return float('nan')
  acov = autocovariance(x)
  chain_var = x.var(axis=1, ddof=1)     # exact (the FFT lag 0 carries rounding)
  chain_var[np.ptp(x, axis=1) == 0] = 0.0   # a constant chain has no variance
  acov[:, 0] = chain_var * (n - 1.0) / n
  W = chain_var.mean()
  var_plus = W * (n - 1.0) / n
  if between and M > 1:
    var_plus += x.mean(axis=1).var(ddof=1)
  if not var_plus > 0:
    return 0.0 if not between else float('nan')
  rho = 1.0 - (W - acov.mean(axis=0)) / var_plus
  rho[0] = 1.0
  # Geyer initial positive sequence over pairs P_k = rho_2k + rho_2k+1
  pairs = []
  k = 0
  while 2 * k + 1 < n:
    p = rho[2 * k] + rho[2 * k + 1]
    if p <= 0:
      break
    pairs.append(p)
    k += 1
  # initial monotone sequence
  for i in range(1, len(pairs)):
    if pairs[i] > pairs[i - 1]:
      pairs[i] = pairs[i - 1]
  tau = -1.0 + 2.0 * float(np.sum(pairs))
  tau = max(tau, 1.0 / math.log10(M * n))
  return M * n / tau


def min_ess(samples):
  """min over coordinates of :func:`ess`; samples [T, C, D]."""
  s = np.asarray(samples, np.float64)
  return min(ess(s[:, :, d].T) for d in range(s.shape[2]))


# ----------------------------------------------------------------------------
# Newton-Laplace posterior (FP64) for the large logistic workloads
# ----------------------------------------------------------------------------


def laplace_logreg(glm, w0=None, iters=50, tol=1e-10):
  """MAP and inverse Hessian of a Bernoulli GLM by Newton's method (FP64):
  the Gaussian N(MAP, H^-1) of ``edward/inferences/laplace.py:17-137``.

  Its moments differ from the posterior's by the posterior's skewness
  (relative to the posterior scale, O(D / sqrt(N)) in the worst case), so the
  moment oracle for the large workloads corrects them by importance sampling
  (``posterior_moments_is``).
  """
  assert glm.likelihood == 'bernoulli'
  D = glm.D
  w = np.zeros(D) if w0 is None else np.array(w0, np.float64)
  prec = 1.0 / glm.prior_scale ** 2
  for _ in range(iters):
    eta = glm.X @ w
    p = sigmoid(eta)
    g = glm.X.T @ (glm.y - p) - (w - glm.prior_loc) * prec
    H = (glm.X * (p * (1 - p))[:, None]).T @ glm.X + np.diag(prec)
    step = np.linalg.solve(H, g)
    w = w + step
    if np.max(np.abs(step)) < tol:
      break
  eta = glm.X @ w
  p = sigmoid(eta)
  H = (glm.X * (p * (1 - p))[:, None]).T @ glm.X + np.diag(prec)
  return w, np.linalg.inv(H)


def posterior_moments_is(glm, mode, cov, n_pairs, seed=0, chunk=256):
  """Posterior mean and variance of a GLM, FP64, with standard errors.

  Importance sampling from the Laplace approximation q = N(mode, cov) (its
  Gaussian is what ``edward/inferences/laplace.py:17-137`` fits) with
  antithetic pairs z = mode +- L xi, L L^T = cov, and self-normalised
  weights p(z) / q(z).  The proposal's own moments are exact, so only the
  weighted deviations from them are estimated (control variates): with
  per-pair weights b_i = w_i+ + w_i- the mean is mode + sum_i (w_i+ - w_i-)
  L xi_i / sum_i b_i, and E[(z - mode)^2] = diag(cov) + sum_i b_i ((L xi_i)^2
  - diag(cov)) / sum_i b_i.  Near a Gaussian posterior w_i+ ~ w_i-, so the
  first sum is small and its noise with it.  Standard errors are the ratio
  estimator's delta-method ones over the n_pairs independent pairs.

  No reference counterpart (Edward reports no posterior moments but the
  sample store's, ``edward/models/empirical.py:87-96``); pinned against
  quadrature of a skewed two-dimensional posterior (tests/test_oracle.py).
  Returns dict(mean, var, se_mean, se_var, weight_ess).
  """
  mode = np.asarray(mode, np.float64)
  D = mode.size
  Lc = np.linalg.cholesky(np.asarray(cov, np.float64))
  diag = np.einsum('ij,ij->i', Lc, Lc)
  rng = np.random.default_rng(seed)
  xi = rng.standard_normal((n_pairs, D))
  dz = xi @ Lc.T
  half_xi2 = 0.5 * (xi ** 2).sum(axis=1)
  lw = np.empty((2, n_pairs))
  for s in range(0, n_pairs, chunk):
    e = min(n_pairs, s + chunk)
    for k, sign in enumerate((1.0, -1.0)):
      # log p(z) - log q(z) up to a constant
      lw[k, s:e] = glm.logp(mode[None, :] + sign * dz[s:e]) + half_xi2[s:e]
  lw -= lw.max()
  w = np.exp(lw)
  b = w[0] + w[1]
  B = b.sum()
  a = (w[0] - w[1])[:, None] * dz
  delta = a.sum(axis=0) / B
  c = b[:, None] * (dz ** 2 - diag[None, :])
  e2 = c.sum(axis=0) / B
  se_mean = np.sqrt(((a - delta[None, :] * b[:, None]) ** 2).sum(axis=0)) / B
  se_var = np.sqrt(((c - e2[None, :] * b[:, None]) ** 2).sum(axis=0)) / B
  return dict(mean=mode + delta, var=diag + e2 - delta ** 2, se_mean=se_mean, se_var=se_var,
              weight_ess=float(B ** 2 / (w ** 2).sum()))


def logreg_grad_f32(X, y, W, threads=None):
  """One plain fp32 gradient evaluation of the logistic log joint for C
  chains, ``hmc.py:199``'s ``tf.gradients`` of ``_log_joint`` in closed
  form (prior N(0, I)): eta = W X^T, r = y - sigmoid(eta), grad = r X - W,
  every step in float32 -- the CPU restatement the benchmark times.  The
  GEMMs run in the host BLAS (all its threads); the elementwise sigmoid runs
  over column blocks in ``threads`` Python threads (numpy releases the GIL).
  X [N, D] float32, y [N] float32, W [C, D] float32."""
  X = np.asarray(X, np.float32)
  y = np.asarray(y, np.float32)
  W = np.asarray(W, np.float32)
  eta = W @ X.T
  n = eta.shape[1]
  nt = threads or (os.cpu_count() or 1)
  edges = np.linspace(0, n, 4 * nt + 1).astype(np.int64)

  def block(k):
    e = eta[:, edges[k]:edges[k + 1]]
    np.negative(e, out=e)
    with np.errstate(over='ignore'):     # exp(-eta) = inf gives sigmoid 0
      np.exp(e, out=e)
    e += np.float32(1.0)
    np.reciprocal(e, out=e)
    np.subtract(y[None, edges[k]:edges[k + 1]], e, out=e)
  with ThreadPoolExecutor(nt) as ex:
    list(ex.map(block, range(len(edges) - 1)))
  return eta @ X - W
