"""GPU parity of the posterior-predictive pass (edhmc_predict) and of
ed.evaluate / ed.ppc built on it, against oracle/criticism_oracle.py.

Bars: mean response over draws within 1e-5 absolute; per-draw log-likelihood
|d| <= 1e-6 |ll| + 1e-3 (fp32 GEMM, FP64 reduction); bf16 (X rounded on both
sides, hi/lo split draws) log-likelihood rtol 1e-5; replicated Bernoulli data
identical except where |u - mu| < 1e-6 (fp32 sigmoid against FP64; 2e-4 for
bf16 draws);
replicated Normal data within 1e-4; metrics equal to the oracle's on the same
draws.
"""

import numpy as np
import pytest
import torch

import edward_amd as ed
from edward_amd.models import Bernoulli, Empirical, Normal
from oracle import criticism_oracle as CO
from oracle import hmc_oracle as O

pytestmark = pytest.mark.gpu


def _data(seed, N, D, likelihood='bernoulli'):
  rng = np.random.default_rng(seed)
  X = rng.normal(size=(N, D)).astype(np.float32)
  w = rng.normal(size=D)
  if likelihood == 'bernoulli':
    y = (rng.random(N) < 1.0 / (1.0 + np.exp(-(X.astype(np.float64) @ w)))).astype(np.float32)
  else:
    y = (X.astype(np.float64) @ w + 0.5 * rng.normal(size=N)).astype(np.float32)
  return X, y, w


@pytest.mark.parametrize('likelihood,precision,N,D,S', [
    ('bernoulli', 'fp32', 1000, 10, 7), ('bernoulli', 'fp32', 3001, 45, 300),
    ('normal', 'fp32', 777, 12, 33), ('bernoulli', 'bf16', 5000, 64, 130)])
def test_predict_matches_oracle(likelihood, precision, N, D, S):
  from edward_amd.engine import GLMEngine
  from edward_amd.workloads import bf16_round
  X, y, w = _data(1, N, D, likelihood)
  if precision == 'bf16':
    X = bf16_round(X)
  scale = 0.5
  eng = GLMEngine(X, y, 0.0, 1.0, likelihood=likelihood, lik_scale=scale, precision=precision)
  rng = np.random.default_rng(2)
  W = (w[None, :] + 0.1 * rng.normal(size=(S, D))).astype(np.float32)
  mean, ll, rep = eng.predict(W, replicate=True, seed=99)
  glm = O.GLM(X.astype(np.float64), y, 0.0, 1.0, likelihood, scale)
  m_ref, ll_ref, mu = CO.predictive(glm, W.astype(np.float64))
  ll = ll.cpu().numpy()
  if precision == 'bf16':
    np.testing.assert_allclose(ll, ll_ref, rtol=1e-5)
  else:
    assert np.all(np.abs(ll - ll_ref) <= 1e-6 * np.abs(ll_ref) + 1e-3), (ll, ll_ref)
  np.testing.assert_allclose(mean.cpu().numpy(), m_ref, atol=1e-5 if likelihood == 'bernoulli'
                             else 1e-4 * np.abs(m_ref).max())
  r_ref = CO.replicate(99, mu, likelihood, scale)
  rep = rep.cpu().numpy()
  if likelihood == 'bernoulli':
    # the uniforms of the replicated draws (counter s, row n), for the tie mask
    u = np.stack([np.floor(O.philox4x32_10(
        O._ctr(np.arange(N, dtype=np.uint64), np.zeros(N, np.uint64), s, O.STREAM_UNIFORM),
        np.broadcast_to(O._key(99), (N, 2)))[:, 0] / 256.0) * 2.0 ** -24 for s in range(S)])
    # fp32: the device sigmoid against FP64; bf16: eta from the hi/lo split
    # (|w - hi - lo| <= 2^-18 |w|) moves the mean response by up to ~1e-4
    decided = np.abs(u - mu) > (2e-4 if precision == 'bf16' else 1e-6)
    assert decided.mean() > 0.99
    np.testing.assert_array_equal(rep[decided], r_ref[decided])
  else:
    np.testing.assert_allclose(rep, r_ref, atol=1e-4 * (1 + np.abs(r_ref).max()))


def _fitted(seed=5, N=400, D=3, T=50, const=False):
  X_np, y_np, w = _data(seed, N, D)
  X = ed.placeholder('float32', [N, D])
  wv = Normal(loc=np.zeros(D, np.float32), scale=np.ones(D, np.float32))
  yv = Bernoulli(logits=ed.dot(X, wv))
  rng = np.random.default_rng(seed + 1)
  rows = (w[None, :] + (0.0 if const else 0.2) * rng.normal(size=(T, D))).astype(np.float32)
  return X, X_np, y_np, wv, yv, Empirical(rows), rows


def test_evaluate_binary_accuracy_and_log_lik_constant_store():
  """All stored rows equal: every draw is w, so the metrics are those of w."""
  X, X_np, y_np, wv, yv, qw, rows = _fitted(const=True)
  y_post = ed.copy(yv, {wv: qw})
  acc = ed.evaluate('binary_accuracy', data={X: X_np, y_post: y_np}, n_samples=20, seed=3)
  glm = O.GLM(X_np.astype(np.float64), y_np, 0.0, 1.0)
  m_ref, ll_ref, _ = CO.predictive(glm, rows[:1].astype(np.float64))
  y_pred = CO.binary_prediction(m_ref, np.random.default_rng(0))
  assert acc == pytest.approx(CO.metric('binary_accuracy', y_np, y_pred), abs=1e-7)
  ll = ed.evaluate('log_lik', data={X: X_np, y_post: y_np}, n_samples=20, seed=3)
  assert ll == pytest.approx(ll_ref[0] / len(y_np), rel=1e-5)
  both = ed.evaluate(['binary_accuracy', 'log_loss'], data={X: X_np, y_post: y_np},
                     n_samples=5, seed=4)
  assert both[1] == pytest.approx(CO.metric('log_loss', y_np, y_pred), rel=1e-5)


def test_evaluate_matches_oracle_on_the_same_draws():
  X, X_np, y_np, wv, yv, qw, rows = _fitted(seed=8)
  y_post = ed.copy(yv, {wv: qw})
  S, seed = 64, 11
  acc, ll = ed.evaluate(['binary_accuracy', 'log_lik'], data={X: X_np, y_post: y_np},
                        n_samples=S, seed=seed)
  # the draws: one uniform row per sample (one chain in the store)
  rng = np.random.default_rng(seed)
  t = rng.integers(0, rows.shape[0], S)
  rng.integers(0, 1, S)
  rng.integers(0, 2 ** 63 - 1)
  W = rows[t].astype(np.float64)
  glm = O.GLM(X_np.astype(np.float64), y_np, 0.0, 1.0)
  m_ref, ll_ref, _ = CO.predictive(glm, W)
  y_pred = CO.binary_prediction(m_ref, rng)
  assert acc == pytest.approx(CO.metric('binary_accuracy', y_np, y_pred), abs=1.0 / len(y_np))
  assert ll == pytest.approx(np.mean(ll_ref / len(y_np)), rel=1e-5)
  assert acc > 0.5     # better than chance on this weakly separable data


def test_ppc_mean_statistic():
  X, X_np, y_np, wv, yv, qw, rows = _fitted(seed=12)
  y_post = ed.copy(yv, {wv: qw})
  S = 40
  T_rep, T_obs = ed.ppc(lambda xs, zs: torch.mean(xs[y_post]), data={X: X_np, y_post: y_np},
                        n_samples=S, seed=13)
  assert T_rep.shape == (S,) and T_obs.shape == (S,)
  np.testing.assert_allclose(T_obs, np.mean(y_np), rtol=1e-6)
  rng = np.random.default_rng(13)
  t = rng.integers(0, rows.shape[0], S)
  rng.integers(0, 1, S)
  pseed = int(rng.integers(0, 2 ** 63 - 1))
  glm = O.GLM(X_np.astype(np.float64), y_np, 0.0, 1.0)
  _, _, mu = CO.predictive(glm, rows[t].astype(np.float64))
  r_ref = CO.replicate(pseed, mu, 'bernoulli')
  np.testing.assert_allclose(T_rep, r_ref.mean(axis=1), atol=2.0 / len(y_np))
  # latent-variable discrepancy: zs carries the draw
  T_rep2, _ = ed.ppc(lambda xs, zs: torch.sum(zs[wv]), data={X: X_np, y_post: y_np},
                     latent_vars={wv: qw}, n_samples=5, seed=1)
  assert np.all(np.isfinite(T_rep2))
