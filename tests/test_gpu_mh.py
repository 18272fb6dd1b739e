"""GPU parity of random-walk Metropolis-Hastings against the oracle
(oracle/hmc_oracle.py mh_step), replaying the device's Philox normals and
uniforms.  Bars: accept decisions identical where |ratio - log u| > 1e-3
(fp32 forward log joint vs FP64), positions within 1e-6 relative."""

import numpy as np
import pytest
import torch

import edward_amd as ed
from edward_amd.models import Bernoulli, Empirical, Normal
from oracle import hmc_oracle as O

pytestmark = pytest.mark.gpu


def _data(seed, N, D):
  rng = np.random.default_rng(seed)
  X = rng.normal(size=(N, D)).astype(np.float32)
  w = rng.normal(size=D) * 0.5
  y = (rng.random(N) < 1.0 / (1.0 + np.exp(-(X.astype(np.float64) @ w)))).astype(np.float32)
  return X, y, w


@pytest.mark.parametrize('precision', ['fp32', 'bf16'])
def test_mh_steps_replay_oracle(precision):
  from edward_amd.engine import GLMEngine
  from edward_amd.workloads import bf16_round
  X, y, w = _data(1, 800, 6)
  if precision == 'bf16':
    X = bf16_round(X)
  C, seed = 64, 21
  eng = GLMEngine(X, y, 0.0, 1.0, precision=precision)
  m = O.GLM(X.astype(np.float64), y, 0.0, 1.0)
  logp_fn = lambda Z: m.logp_grad(Z)[0]
  rng = np.random.default_rng(2)
  z = (w[None, :] + 0.05 * rng.normal(size=(C, 6))).astype(np.float32).astype(np.float64)
  eng.set_state(z.astype(np.float32))
  lz = logp_fn(z)
  scale = np.full(6, 0.04)
  out = torch.empty(C, 6, device='cuda')
  acc_d = torch.zeros(C, dtype=torch.int32, device='cuda')
  n_decided = 0
  for t in range(10):
    xi = O.momentum_normals(seed, t, C, 6)
    u = O.accept_uniforms(seed, t, C)
    eng.mh_step(scale, seed=seed, counter=t, sample_out=out, accepted_out=acc_d)
    zp = (z.astype(np.float32) + (scale * xi).astype(np.float32)).astype(np.float64)
    z_new, lz_new, acc, ratio = O.mh_step(z, lz, scale, xi, u, logp_fn)
    decided = np.abs(ratio - np.log(u)) > 1e-3
    n_decided += decided.sum()
    a = acc_d.cpu().numpy().astype(bool)
    np.testing.assert_array_equal(a[decided], acc[decided])
    zd = out.cpu().numpy().astype(np.float64)
    want = np.where(a[:, None], zp, z)      # follow the device's decisions
    np.testing.assert_allclose(zd, want, rtol=1e-6, atol=1e-7)
    z = zd
    lz = logp_fn(z)
  assert n_decided > 0.9 * 10 * C


def test_mh_surface_samples_the_posterior():
  """ed.MetropolisHastings with a Normal random-walk proposal on a small
  logistic regression: the pooled draws of 256 chains match the FP64
  Laplace posterior mean within 4 posterior sd."""
  N, D, T, C = 2000, 3, 400, 256
  X_np, y_np, _ = _data(5, N, D)
  X = ed.placeholder('float32', [N, D])
  w = Normal(loc=np.zeros(D, np.float32), scale=np.ones(D, np.float32))
  yv = Bernoulli(logits=ed.dot(X, w))
  qw = Empirical(np.zeros((T, C, D), np.float32))
  inf = ed.MetropolisHastings({w: qw}, {w: Normal(loc=w, scale=0.03)}, data={X: X_np, yv: y_np})
  inf.initialize(seed=9)
  for _ in range(T):
    info = inf.update()
  assert 0.1 < info['accept_rate'] < 0.95
  m = O.GLM(X_np.astype(np.float64), y_np, 0.0, 1.0)
  mode, cov = O.laplace_logreg(m)
  post = qw.params[T // 2:].double().mean((0, 1)).cpu().numpy()
  assert np.abs(post - mode).max() < 4 * np.sqrt(np.diag(cov)).max()
