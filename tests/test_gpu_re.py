"""GPU parity of replica-exchange Monte Carlo against the oracle
(oracle/hmc_oracle.py mh_step with per-chain inverse temperatures, then
replica_exchange), replaying the device's Philox draws.  Bars: tempered MH
decisions and exchange decisions identical where the ratio is more than 1e-3
from the threshold (fp32 forward log joint vs FP64); positions within 1e-6
relative."""

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


def test_replica_exchange_replays_oracle():
  from edward_amd.engine import GLMEngine
  X, y, w = _data(3, 600, 4)
  betas = np.array([1.0, 0.5, 0.2, 0.05])
  R, ladders, seed, freq = len(betas), 24, 31, 0.7
  C = R * ladders
  beta_c = np.repeat(betas, ladders)
  eng = GLMEngine(X, y, 0.0, 1.0)
  m = O.GLM(X.astype(np.float64), y, 0.0, 1.0)
  logp_fn = lambda Z: m.logp_grad(Z)[0]
  rng = np.random.default_rng(4)
  z = (w[None, :] + 0.1 * rng.normal(size=(C, 4))).astype(np.float32).astype(np.float64)
  eng.set_state(z.astype(np.float32))
  lz = logp_fn(z)
  scale = np.full(4, 0.05)
  out = torch.empty(C, 4, device='cuda')
  swaps = torch.zeros(ladders, dtype=torch.int32, device='cuda')
  checked, n_swaps = 0, 0
  for t in range(12):
    xi = O.momentum_normals(seed, t, C, 4)
    u = O.accept_uniforms(seed, t, C)
    eng.replica_exchange_step(scale, beta_c.astype(np.float32), R, freq, seed=seed, counter=t,
                              sample_out=out, swaps=swaps)
    z1, lz1, acc, ratio = O.mh_step(z, lz, scale, xi, u, logp_fn, beta_c)
    decided = np.all(np.abs(ratio - np.log(u)) > 1e-3)
    perm = O.replica_exchange(lz1, betas, ladders, seed, t, freq)
    # an exchange is decided when moving every log joint by up to 1e-3 (the
    # two alternating sign patterns bound the ratio's change) keeps the swaps
    alt = 1e-3 * (-1.0) ** np.repeat(np.arange(R), ladders)
    decided &= all(np.array_equal(perm, O.replica_exchange(lz1 + e, betas, ladders, seed, t, freq))
                   for e in (alt, -alt))
    zd = out.cpu().numpy().astype(np.float64)
    if decided:
      want = z1[perm]
      np.testing.assert_allclose(zd, want, rtol=1e-6, atol=1e-7)
      checked += 1
      n_swaps += int((perm != np.arange(C)).sum() // 2)
    z = zd
    lz = logp_fn(z)
  assert checked >= 8
  assert n_swaps > 0
  assert int(swaps.cpu().numpy().sum()) >= n_swaps


def test_replica_exchange_surface():
  N, D, T, C = 1500, 3, 300, 64
  X_np, y_np, _ = _data(6, N, D)
  X = ed.placeholder('float32', [N, D])
  w = Normal(loc=np.zeros(D, np.float32), scale=np.ones(D, np.float32))
  yv = Bernoulli(logits=ed.dot(X, w))
  qw = Empirical(np.zeros((T, C, D), np.float32))
  inf = ed.ReplicaExchangeMC({w: qw}, {w: Normal(loc=w, scale=0.04)}, data={X: X_np, yv: y_np},
                             inverse_temperatures=[1.0, 0.5, 0.25], exchange_freq=0.5)
  inf.initialize(seed=2)
  for _ in range(T):
    info = inf.update()
  assert 0.05 < info['accept_rate'] < 0.98
  assert inf.n_swaps.sum() > 0
  assert len(inf.replica_vars) == 3
  r2 = inf.replica_vars[2][w].params[T - 1].cpu().numpy()
  assert np.isfinite(r2).all() and np.abs(r2).max() > 0
  m = O.GLM(X_np.astype(np.float64), y_np, 0.0, 1.0)
  mode, cov = O.laplace_logreg(m)
  post = qw.params[T // 2:].double().mean((0, 1)).cpu().numpy()
  assert np.abs(post - mode).max() < 4 * np.sqrt(np.diag(cov)).max()
  # replica 0's store and the user's store hold the same draws
  np.testing.assert_array_equal(inf.replica_vars[0][w].params[T - 1].cpu().numpy(),
                                qw.params[T - 1].cpu().numpy())
