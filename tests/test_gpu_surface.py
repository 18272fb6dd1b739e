"""The reference's HMC tests and example programs, run through
``import edward_amd as ed`` on the GPU.

* ``tests/inferences/hmc_test.py:14-46``: normal-normal (50 zeros, prior
  N(0,1), posterior N(0, 1/51)), dict and list forms, float32 and float64,
  ``t``, ``n_accept`` and ``reset`` -- same tolerances (rtol = atol = 0.1).
* ``hmc_test.py:48-91``: linear regression, step_size 0.01, atol 0.5.
* ``examples/bayesian_logistic_regression.py``: the 2-parameter model
  (w, b) with step_size 0.6 against posterior moments from exact 2-D
  quadrature on a grid (FP64).
Seeded with ``ed.set_seed`` (the reference seeds only under ``__main__``).
"""

import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def fresh_graph():
  from edward_amd import models
  models.reset_graph()
  yield
  models.reset_graph()


def _normal_normal(default, dtype, seed):
  import edward_amd as ed
  from edward_amd.models import Empirical, Normal
  ed.set_seed(seed)
  x_data = np.zeros(50, np.float32)
  mu = Normal(loc=np.array(0.0, dtype), scale=np.array(1.0, dtype))
  x = Normal(loc=mu, scale=np.array(1.0, dtype), sample_shape=50)
  n_samples = 2000
  if not default:
    qmu = Empirical(params=torch.ones(n_samples, dtype=getattr(torch, np.dtype(dtype).name)))
    inference = ed.HMC({mu: qmu}, data={x: x_data})
  else:
    inference = ed.HMC([mu], data={x: x_data})
    qmu = inference.latent_vars[mu]
  inference.run()
  return inference, qmu


@pytest.mark.parametrize('default', [True, False])
@pytest.mark.parametrize('dtype', [np.float32, np.float64])
def test_normal_normal(default, dtype):
  inference, qmu = _normal_normal(default, dtype, 42)
  np.testing.assert_allclose(qmu.mean().cpu().numpy(), 0, rtol=1e-1, atol=1e-1)
  np.testing.assert_allclose(qmu.stddev().cpu().numpy(), np.sqrt(1 / 51), rtol=1e-1, atol=1e-1)
  assert inference.t == (10000 if default else 2000)
  assert inference.n_accept > 0.1
  assert qmu.params.dtype == getattr(torch, np.dtype(dtype).name)
  inference.reset()
  assert inference.t == 0 and inference.n_accept == 0


@pytest.mark.parametrize('default', [True, False])
def test_linear_regression(default):
  import edward_amd as ed
  from edward_amd.models import Empirical, Normal
  ed.set_seed(1)
  rng = np.random.default_rng(1)

  def build_toy_dataset(N, w, noise_std=0.1):
    x = rng.normal(size=(N, len(w)))
    return x, x @ w + rng.normal(0, noise_std, size=N)
  N, D = 40, 10
  w_true = rng.normal(size=D)
  X_train, y_train = build_toy_dataset(N, w_true)
  X = ed.placeholder('float32', [N, D])
  w = Normal(loc=np.zeros(D, np.float32), scale=np.ones(D, np.float32))
  b = Normal(loc=np.zeros(1, np.float32), scale=np.ones(1, np.float32))
  y = Normal(loc=ed.dot(X, w) + b, scale=0.1 * np.ones(N, np.float32))
  n_samples = 2000
  if not default:
    qw = Empirical(torch.zeros(n_samples, D))
    qb = Empirical(torch.zeros(n_samples, 1))
    inference = ed.HMC({w: qw, b: qb}, data={X: X_train, y: y_train})
  else:
    inference = ed.HMC([w, b], data={X: X_train, y: y_train})
    qw, qb = inference.latent_vars[w], inference.latent_vars[b]
  inference.run(step_size=0.01)
  np.testing.assert_allclose(qw.mean().cpu().numpy(), w_true, rtol=5e-1, atol=5e-1)
  np.testing.assert_allclose(qb.mean().cpu().numpy(), [0.0], rtol=5e-1, atol=5e-1)
  assert inference.t == (10000 if default else n_samples)
  assert inference.n_accept > 0.1
  inference.reset()
  assert inference.t == 0 and inference.n_accept == 0


def _toy_logreg(N=40, noise_std=0.1, seed=42):
  rng = np.random.default_rng(seed)
  X = np.linspace(-6, 6, num=N)
  y = np.tanh(X) + rng.normal(0, noise_std, size=N)
  y = (y >= 0.5).astype(np.float32)
  X = ((X - 4.0) / 4.0).reshape(N, 1).astype(np.float32)
  return X, y


def _grid_moments(X, y, scale=3.0):
  """Exact posterior moments of (w, b) by quadrature on a 1201^2 grid."""
  w = np.linspace(-20, 20, 1201)
  b = np.linspace(-20, 20, 1201)
  Wg, Bg = np.meshgrid(w, b, indexing='ij')
  eta = Wg[..., None] * X[:, 0] + Bg[..., None]
  ll = (y * eta - np.logaddexp(0, eta)).sum(-1)
  lp = ll - 0.5 * (Wg / scale) ** 2 - 0.5 * (Bg / scale) ** 2
  p = np.exp(lp - lp.max())
  p /= p.sum()
  mw, mb = (p * Wg).sum(), (p * Bg).sum()
  vw, vb = (p * (Wg - mw) ** 2).sum(), (p * (Bg - mb) ** 2).sum()
  return np.array([mw, mb]), np.array([vw, vb])


def test_logistic_regression_example_matches_quadrature():
  """examples/bayesian_logistic_regression.py:34-81 with 64 chains (the
  Empirical stores hold [T, 64, ...]): HMC(step_size=0.6), the manual
  update/print_progress loop, moments against quadrature."""
  import edward_amd as ed
  from edward_amd.models import Bernoulli, Empirical, Normal
  ed.set_seed(42)
  N, D, T, C = 40, 1, 3000, 64
  X_train, y_train = _toy_logreg(N)
  X = ed.placeholder('float32', [N, D])
  w = Normal(loc=np.zeros(D, np.float32), scale=3.0 * np.ones(D, np.float32))
  b = Normal(loc=np.float32(0.0), scale=np.float32(3.0))
  y = Bernoulli(logits=ed.dot(X, w) + b)
  qw = Empirical(params=torch.zeros(T, C, D))
  qb = Empirical(params=torch.zeros(T, C))
  inference = ed.HMC({w: qw, b: qb}, data={X: X_train, y: y_train})
  inference.initialize(n_print=500, step_size=0.6)
  for _ in range(inference.n_iter):
    info = inference.update()
    inference.print_progress(info)
  assert inference.t == T and 0.3 < info['accept_rate'] <= 1.0
  mean_ref, var_ref = _grid_moments(X_train, y_train)
  sw = qw.params[500:].reshape(-1).double().cpu().numpy()
  sb = qb.params[500:].reshape(-1).double().cpu().numpy()
  n_eff = sw.size / 20.0
  for s, m, v in ((sw, mean_ref[0], var_ref[0]), (sb, mean_ref[1], var_ref[1])):
    assert abs(s.mean() - m) < 5 * math.sqrt(v / n_eff)
    assert s.var() == pytest.approx(v, rel=0.1)
  # Empirical readout through the device moments kernel, burn-in excluded
  mw = qw.mean(burn_in=500).double().cpu().numpy()
  np.testing.assert_allclose(mw.mean(), sw.mean(), rtol=1e-5)


def test_initialize_returns_self_and_run_keeps_hyperparameters():
  """Deliberate extension: ed.HMC(...).initialize(eps, L).run() keeps eps/L
  (the reference's run() re-initialises to the defaults 0.25/2)."""
  import edward_amd as ed
  from edward_amd.models import Empirical, Normal
  mu = Normal(loc=np.float32(0.0), scale=np.float32(1.0))
  x = Normal(loc=mu, scale=np.float32(1.0), sample_shape=50)
  qmu = Empirical(torch.zeros(200))
  inf = ed.HMC({mu: qmu}, data={x: np.zeros(50, np.float32)}).initialize(step_size=0.1,
                                                                       n_steps=3)
  assert isinstance(inf, ed.HMC)
  inf.run()
  assert inf.step_size == 0.1 and inf.n_steps == 3 and inf.t == 200
  inf.run(step_size=0.2)             # keyword arguments re-initialise (reference)
  assert inf.step_size == 0.2 and inf.n_steps == 2 and inf.t == 200


def test_debug_raises_on_non_finite_proposals():
  """A step size that overflows every proposal: rejected, counted, and with
  debug=True update() raises (the reference's check_numerics)."""
  import edward_amd as ed
  from edward_amd.models import Empirical, Normal
  mu = Normal(loc=np.float32(0.0), scale=np.float32(1.0))
  x = Normal(loc=mu, scale=np.float32(1.0), sample_shape=50)
  qmu = Empirical(torch.zeros(10))
  inf = ed.HMC({mu: qmu}, data={x: np.zeros(50, np.float32)})
  inf.initialize(step_size=1e30, n_steps=3)
  inf.update()
  assert inf.n_accept == 0 and inf.divergences.sum() == 1
  np.testing.assert_array_equal(qmu.params[0].cpu().numpy(), 0.0)   # kept z0
  inf.initialize(step_size=1e30, n_steps=3, debug=True)
  with pytest.raises(FloatingPointError):
    inf.update()
