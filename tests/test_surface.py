"""CPU tests of the Edward-compatible surface: argument checks, model
matching, seeding and progress output (the GPU runs of the same programs are
in test_gpu_surface.py).  Mirrors the reference's error behaviour:
``util/random_variables.py:21-83``, ``monte_carlo.py:78-91``,
``util/graphs.py:59-73``."""

import io

import numpy as np
import pytest
import torch

import edward_amd as ed
from edward_amd import compile as Cm
from edward_amd import models as M
from edward_amd.inferences import HMC, MonteCarlo, check_data, check_latent_vars
from edward_amd.models import Bernoulli, Empirical, Normal
from edward_amd.progbar import Progbar


def test_check_latent_vars_errors():
  with pytest.raises(TypeError):
    check_latent_vars([1, 2])
  w = Normal(loc=np.zeros(3, np.float32), scale=np.ones(3, np.float32))
  with pytest.raises(TypeError):
    check_latent_vars({w: 3})
  with pytest.raises(TypeError):
    check_latent_vars({w: Empirical(np.zeros((10, 4), np.float32))})
  with pytest.raises(TypeError):   # dtype mismatch
    check_latent_vars({w: Empirical(np.zeros((10, 3), np.float64))})
  check_latent_vars({w: Empirical(np.zeros((10, 3), np.float32))})
  check_latent_vars({w: Empirical(np.zeros((10, 7, 3), np.float32))})   # 7 chains


def test_check_data_errors():
  X = ed.placeholder('float32', [4, 2])
  w = Normal(loc=np.zeros(2), scale=np.ones(2))
  y = Bernoulli(logits=ed.dot(X, w))
  with pytest.raises(TypeError):
    check_data({X: w})
  with pytest.raises(TypeError):
    check_data({'X': np.zeros(3)})
  with pytest.raises(TypeError):
    check_data({y: object()})
  check_data({X: np.zeros((4, 2)), y: np.zeros(4)})


def test_montecarlo_rejects_non_empirical_and_builds_list_form():
  w = Normal(loc=np.zeros(3, np.float32), scale=np.ones(3, np.float32))
  qw = Normal(loc=np.zeros(3, np.float32), scale=np.ones(3, np.float32))
  with pytest.raises(TypeError):
    HMC({w: qw})
  inf = HMC([w])
  qz = inf.latent_vars[w]
  assert isinstance(qz, Empirical) and tuple(qz.params.shape) == (10000, 3)
  assert float(qz.params.abs().sum()) == 0.0     # zeros (monte_carlo.py:80-83)


def _logreg_model(N=50, D=3, T=20, C=None, bias=True):
  rng = np.random.default_rng(0)
  Xv = rng.normal(size=(N, D)).astype(np.float32)
  yv = (rng.random(N) < 0.5).astype(np.int32)
  X = ed.placeholder('float32', [N, D])
  w = Normal(loc=np.zeros(D, np.float32), scale=3.0 * np.ones(D, np.float32))
  lat = {w: Empirical(np.zeros((T, D) if C is None else (T, C, D), np.float32))}
  logits = ed.dot(X, w)
  if bias:
    b = Normal(loc=np.float32(0.0), scale=np.float32(3.0))
    logits = logits + b
    lat[b] = Empirical(np.zeros((T,) if C is None else (T, C), np.float32))
  y = Bernoulli(logits=logits)
  return lat, {X: Xv, y: yv}, Xv, yv


def test_glm_spec_of_the_logistic_regression_example():
  """examples/bayesian_logistic_regression.py:41-48: dot(X, w) + b."""
  lat, data, Xv, yv = _logreg_model()
  spec = Cm.glm_spec(lat, data)
  assert spec.likelihood == 'bernoulli'
  np.testing.assert_array_equal(spec.X[:, :3], Xv)
  np.testing.assert_array_equal(spec.X[:, 3], 1.0)
  np.testing.assert_array_equal(spec.y, yv)
  np.testing.assert_array_equal(spec.scale, 3.0)
  assert spec.dim == 4 and spec.n_chains == 1 and spec.n_rows == 20


def test_glm_spec_with_chains():
  lat, data, _, _ = _logreg_model(C=6)
  spec = Cm.glm_spec(lat, data)
  assert spec.n_chains == 6 and spec.dim == 4


def test_glm_spec_normal_normal_and_linear_regression():
  """hmc_test.py:18-21 (50 zeros, mu ~ N(0,1)) and :63-66 (y ~ N(Xw + b, 0.1))."""
  mu = Normal(loc=np.float32(0.0), scale=np.float32(1.0))
  x = Normal(loc=mu, scale=np.float32(1.0), sample_shape=50)
  spec = Cm.glm_spec({mu: Empirical(np.ones(30, np.float32))}, {x: np.zeros(50, np.float32)})
  assert spec.likelihood == 'normal' and spec.lik_scale == 1.0
  np.testing.assert_array_equal(spec.X, np.ones((50, 1)))
  N, D = 40, 10
  X = ed.placeholder('float32', [N, D])
  w = Normal(loc=np.zeros(D, np.float32), scale=np.ones(D, np.float32))
  b = Normal(loc=np.zeros(1, np.float32), scale=np.ones(1, np.float32))
  y = Normal(loc=ed.dot(X, w) + b, scale=0.1 * np.ones(N, np.float32))
  Xv = np.random.default_rng(1).normal(size=(N, D)).astype(np.float32)
  spec = Cm.glm_spec({w: Empirical(np.zeros((5, D), np.float32)),
                      b: Empirical(np.zeros((5, 1), np.float32))},
                     {X: Xv, y: np.zeros(N, np.float32)})
  assert spec.likelihood == 'normal' and spec.lik_scale == pytest.approx(0.1)
  assert spec.X.shape == (N, D + 1)


def test_unsupported_models_raise():
  w = Normal(loc=np.zeros(2, np.float32), scale=np.ones(2, np.float32))
  X = ed.placeholder('float32', [3, 2])
  y = Bernoulli(probs=ed.dot(X, w))
  with pytest.raises(NotImplementedError):
    Cm.glm_spec({w: Empirical(np.zeros((4, 2), np.float32))}, {X: np.ones((3, 2)), y: np.ones(3)})
  y2 = Bernoulli(logits=ed.dot(X, w) + 1.5)
  with pytest.raises(NotImplementedError):
    Cm.glm_spec({w: Empirical(np.zeros((4, 2), np.float32))}, {X: np.ones((3, 2)), y2: np.ones(3)})


def test_set_seed_refuses_after_model_nodes():
  M.reset_graph()
  ed.set_seed(42)
  assert ed.util.get_seed() == 42
  Normal(loc=0.0, scale=1.0)
  with pytest.raises(RuntimeError):
    ed.set_seed(1)
  M.reset_graph()


def test_progbar_format():
  out = io.StringIO()
  pb = Progbar(100, stream=out)
  pb.update(100, {'Acceptance Rate': 0.75})
  s = out.getvalue()
  assert '100/100 [100%]' in s and 'Acceptance Rate: 0.750' in s and 'Elapsed' in s


def test_empirical_moments_on_cpu_match_the_reference_definition():
  p = np.random.default_rng(3).normal(size=(7, 4)).astype(np.float32)
  q = Empirical(torch.tensor(p))
  np.testing.assert_allclose(q.mean().cpu().numpy(), p.mean(0), rtol=1e-6)
  np.testing.assert_allclose(q.stddev().cpu().numpy(), p.std(0), rtol=1e-5)
  np.testing.assert_allclose(q.mean(burn_in=3).cpu().numpy(), p[3:].mean(0), rtol=1e-6)


def test_replica_exchange_checks_the_temperatures():
  """``replica_exchange_mc.py:59-63``: the first inverse temperature must be 1."""
  import edward_amd as ed
  from edward_amd.models import Empirical, Normal
  z = Normal(0.0, 1.0)
  with pytest.raises(ValueError):
    ed.ReplicaExchangeMC({z: Empirical(np.zeros(10, np.float32))},
                         {z: Normal(loc=z, scale=0.5)}, inverse_temperatures=[0.5, 0.25])
  inf = ed.ReplicaExchangeMC({z: Empirical(np.zeros(10, np.float32))},
                             {z: Normal(loc=z, scale=0.5)}, inverse_temperatures=[1.0, 0.3])
  assert inf.n_replica == 2 and len(inf.replica_vars) == 2
