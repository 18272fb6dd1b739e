"""Criticism (ed.evaluate / ed.ppc) on the CPU: the oracle restatement against
closed forms and brute force, the constant-parameter outputs of the
reference's own tests (tests/criticisms/evaluate_test.py:15-42), and the
argument checks of evaluate.py:101-118 and ppc.py:84-97."""

import math

import numpy as np
import pytest
from scipy import stats

import edward_amd as ed
from edward_amd.models import Bernoulli, Normal
from oracle import criticism_oracle as CO
from oracle import hmc_oracle as O


def test_oracle_metrics_closed_forms():
  yt = np.array([1.0, 0.0, 1.0, 1.0])
  yp = np.array([1.0, 1.0, 1.0, 0.0])
  assert CO.metric('binary_accuracy', yt, yp) == 0.5
  # sigmoid cross entropy with y_pred as logits: log(1 + e^-z) for y=1, log(1 + e^z) for y=0
  want = np.mean([math.log1p(math.exp(-1)), math.log1p(math.exp(1)), math.log1p(math.exp(-1)),
                  math.log(2.0)])
  assert CO.metric('log_loss', yt, yp) == pytest.approx(want, rel=1e-12)
  assert CO.metric('mse', yt, yp) == 0.5
  assert CO.metric('mae', [1.0, 2.0], [2.0, 4.0]) == 1.5
  assert CO.metric('mape', [1.0, 2.0], [2.0, 4.0]) == pytest.approx(100.0)
  assert CO.metric('cosine', [1.0, 0.0], [2.0, 0.0]) == pytest.approx(1.0)
  assert CO.metric('hinge', [1.0, -1.0], [0.5, 0.5]) == pytest.approx(1.0)
  with pytest.raises(NotImplementedError):
    CO.metric('hello world', yt, yp)


def test_oracle_predictive_matches_bruteforce():
  rng = np.random.default_rng(1)
  N, D, S = 30, 4, 5
  X = rng.normal(size=(N, D))
  y = (rng.random(N) < 0.5).astype(float)
  W = rng.normal(size=(S, D))
  glm = O.GLM(X, y, 0.0, 1.0)
  mean, ll, _ = CO.predictive(glm, W)
  for s in range(S):
    p = 1.0 / (1.0 + np.exp(-(X @ W[s])))
    assert ll[s] == pytest.approx(np.sum(stats.bernoulli.logpmf(y, p)), rel=1e-12)
  p_all = 1.0 / (1.0 + np.exp(-(W @ X.T)))
  np.testing.assert_allclose(mean, p_all.mean(axis=0), rtol=1e-12)
  g2 = O.GLM(X, X @ W[0] + 0.1, 0.0, 1.0, 'normal', 0.7)
  _, ll2, _ = CO.predictive(g2, W[:2])
  assert ll2[1] == pytest.approx(np.sum(stats.norm.logpdf(X @ W[0] + 0.1, X @ W[1], 0.7)),
                                 rel=1e-12)


def test_oracle_replicate_distribution():
  mu = np.full((4, 20000), 0.3)
  r = CO.replicate(7, mu, 'bernoulli')
  assert set(np.unique(r)) <= {0.0, 1.0}
  assert abs(r.mean() - 0.3) < 5 * math.sqrt(0.3 * 0.7 / r.size)
  # different (s, n) counters give different draws
  assert not np.array_equal(r[0], r[1])
  z = CO.replicate(8, np.zeros((2, 50000)), 'normal', 2.0)
  assert abs(z.mean()) < 5 * 2.0 / math.sqrt(z.size)
  assert abs(z.std() - 2.0) < 0.03


def test_evaluate_constant_bernoulli_like_the_reference():
  """tests/criticisms/evaluate_test.py:31-45 (Bernoulli cases)."""
  x = Bernoulli(probs=0.51)
  assert ed.evaluate('binary_accuracy', {x: np.array(1)}, n_samples=1) == pytest.approx(1.0)
  x = Bernoulli(probs=0.51, sample_shape=5)
  assert ed.evaluate('binary_accuracy', {x: np.array([1, 1, 1, 0, 0])},
                     n_samples=1) == pytest.approx(0.6)
  x = Bernoulli(probs=np.array([0.51, 0.49, 0.49]))
  assert ed.evaluate('binary_accuracy', {x: np.array([1, 0, 1])},
                     n_samples=1) == pytest.approx(2.0 / 3)


def test_evaluate_argument_checks_like_the_reference():
  """tests/criticisms/evaluate_test.py:15-26."""
  x = Normal(loc=0.0, scale=1.0)
  ed.evaluate('mean_squared_error', {x: np.array(0.0)}, n_samples=1)
  ed.evaluate(['mean_squared_error'], {x: np.array(0.0)}, n_samples=1)
  v = ed.evaluate(['mean_squared_error', 'mean_absolute_error'], {x: np.array(0.0)}, n_samples=1)
  assert len(v) == 2
  with pytest.raises(TypeError):
    ed.evaluate(x, {x: np.array(0.0)}, n_samples=1)
  with pytest.raises(NotImplementedError):
    ed.evaluate('hello world', {x: np.array(0.0)}, n_samples=1)
  with pytest.raises(TypeError):
    ed.evaluate('mse', {x: np.array(0.0)}, n_samples='1')


def test_evaluate_constant_normal_log_lik():
  x = Normal(loc=1.0, scale=2.0, sample_shape=3)
  d = np.array([0.0, 1.0, 3.0])
  v = ed.evaluate('log_lik', {x: d}, n_samples=4)
  assert v == pytest.approx(np.mean(stats.norm.logpdf(d, 1.0, 2.0)), rel=1e-6)


def test_ppc_argument_checks_like_the_reference():
  """tests/criticisms/ppc_test.py:21-23,52-53."""
  x = Normal(loc=0.0, scale=1.0)
  with pytest.raises(TypeError):
    ed.ppc(lambda xs, zs: 0.0, {'y': np.array(0.0)}, n_samples=1)
  with pytest.raises(TypeError):
    ed.ppc(lambda xs, zs: 0.0, {x: np.array(0.0)}, n_samples='1')
  with pytest.raises(TypeError):
    ed.ppc(1.0, {x: np.array(0.0)}, n_samples=1)
