"""GPU parity of auto-transformed (constrained) priors: the device's
per-coordinate prior families and operand maps against
oracle/transform_oracle.py, and the reference's own auto-transform HMC tests
(``tests/inferences/inference_auto_transform_test.py:112-188``) through the
``ed.HMC`` surface.

Bars: log joint |dlogp| <= 1e-6 |logp| + 1e-3 and gradient within 1e-4 of
max|grad| (fp32 operands, FP64 prior terms); leapfrog end positions within
1e-4 relative with fed momenta; device-RNG replay positions within 1e-4
relative until the first near-threshold decision; posterior moments within
the reference tests' tolerances (1e-1 for the softplus-Normal cases, 5e-2 /
1e-2 for Beta-Bernoulli) and within 1e-2 over many chains."""

import numpy as np
import pytest
import scipy.stats as st
import torch

import edward_amd as ed
from edward_amd import bijectors as B
from edward_amd import compile as Cm
from edward_amd.models import Bernoulli, Beta, Empirical, Gamma, Normal, TransformedDistribution
from oracle import hmc_oracle as O
from oracle import transform_oracle as TO

pytestmark = pytest.mark.gpu

FAM = [Cm.FAMILY_NORMAL, Cm.FAMILY_GAMMA, Cm.FAMILY_BETA, Cm.FAMILY_INVGAMMA,
       Cm.FAMILY_NORMAL, Cm.FAMILY_GAMMA, Cm.FAMILY_BETA, Cm.FAMILY_NORMAL]
OP = [Cm.OP_IDENTITY, Cm.OP_SOFTPLUS, Cm.OP_SIGMOID, Cm.OP_SOFTPLUS,
      Cm.OP_SOFTPLUS, Cm.OP_SOFTPLUS, Cm.OP_IDENTITY, Cm.OP_SIGMOID]
P1 = [0.2, 2.0, 3.0, 2.5, 1.0, 0.7, 1.5, -0.3]
P2 = [1.3, 0.5, 2.0, 1.0, 0.8, 3.0, 2.5, 0.9]


def _model(seed, N, likelihood='bernoulli'):
  rng = np.random.default_rng(seed)
  D = len(FAM)
  X = rng.normal(size=(N, D)).astype(np.float32)
  if likelihood == 'bernoulli':
    y = (rng.random(N) < 0.4).astype(np.float32)
  else:
    y = (X.astype(np.float64) @ rng.normal(size=D) * 0.3 + rng.normal(size=N)).astype(np.float32)
  fams = [Cm.PriorFamily(f, a, b, o, 1) for f, a, b, o in zip(FAM, P1, P2, OP)]
  kind = np.concatenate([f.kind for f in fams])
  lnorm = np.concatenate([f.lnorm for f in fams])
  p1 = np.concatenate([f.p1 for f in fams])
  p2 = np.concatenate([f.p2 for f in fams])
  from edward_amd.engine import GLMEngine
  eng = GLMEngine(X, y, p1.astype(np.float32), p2.astype(np.float32), likelihood=likelihood,
                  lik_scale=0.9)
  eng.set_prior_family(kind, lnorm)
  m = TO.TransformedGLM(X.astype(np.float64), y, FAM, OP, p1, p2, likelihood, 0.9)
  return eng, m


def _check(lp, g, lp_ref, g_ref):
  lp = lp.cpu().numpy()
  g = g.cpu().numpy().astype(np.float64)
  assert np.all(np.abs(lp - lp_ref) <= 1e-6 * np.abs(lp_ref) + 1e-3), (lp, lp_ref)
  tol = 1e-4 * np.abs(g_ref).max()
  assert np.abs(g - g_ref).max() <= tol, (np.abs(g - g_ref).max(), tol)


def _rel_err(a, b):
  a, b = np.atleast_2d(a), np.atleast_2d(b)
  return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-6)


@pytest.mark.parametrize('likelihood', ['bernoulli', 'normal'])
def test_logp_grad_with_prior_families(likelihood):
  eng, m = _model(1, 900, likelihood)
  U = (np.random.default_rng(2).normal(size=(11, len(FAM))) * 1.5).astype(np.float32)
  U[0, :] = [-9.0, -9.0, 9.0, -9.0, 9.0, 12.0, -12.0, 0.0]     # far tails
  lp, g = eng.logp_grad(U)
  lp_ref, g_ref = m.logp_grad(U.astype(np.float64))
  _check(lp, g, lp_ref, g_ref)
  lpf = eng.logp(U).cpu().numpy()           # forward-only pass, same operands
  assert np.all(np.abs(lpf - lp_ref) <= 1e-6 * np.abs(lp_ref) + 1e-3)
  # clearing the families restores the plain Normal model
  eng.set_prior_family(None)
  lp2, _ = eng.logp_grad(U)
  plain = O.GLM(m.glm.X, m.glm.y, m.p1, m.p2, likelihood, 0.9).logp_grad(U.astype(np.float64))
  assert np.all(np.abs(lp2.cpu().numpy() - plain[0]) <= 1e-6 * np.abs(plain[0]) + 1e-3)


def test_transition_with_fed_randomness():
  eng, m = _model(3, 1200)
  rng = np.random.default_rng(4)
  C = 16
  U0 = (rng.normal(size=(C, len(FAM))) * 0.5).astype(np.float32)
  r0 = rng.normal(size=(C, len(FAM))).astype(np.float32)
  eng.set_state(U0)
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  eng.transition(0.02, 12, momenta=r0, uniforms=np.zeros(C), accepted_out=acc)
  z, lp, _ = eng.get_state()
  out = O.hmc_transition(U0.astype(np.float64), r0.astype(np.float64), np.zeros(C), 0.02, 12,
                         m.logp_grad)
  assert acc.cpu().numpy().all()
  assert _rel_err(z.cpu().numpy(), out['zL']).max() < 1e-4
  np.testing.assert_allclose(lp.cpu().numpy(), out['logpL'], rtol=1e-6, atol=1e-3)


def test_replay_with_device_rng():
  """60 transitions of 8 chains with the device's Philox draws, replayed by
  the oracle in the unconstrained space until the first near-threshold
  decision."""
  eng, m = _model(5, 400)
  C, T, seed, eps, L = 8, 60, 77, 0.05, 8
  params = torch.zeros(T, C, len(FAM), device='cuda')
  n_acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  eng.set_state(np.zeros((C, len(FAM)), np.float32))
  eng.run(params, eps, L, seed=seed, counter0=0, n_accept=n_acc)
  p = params.cpu().numpy().astype(np.float64)
  z = np.zeros((C, len(FAM)))
  n_ref = np.zeros(C, int)
  for t in range(T):
    out = O.hmc_transition(z, O.momentum_normals(seed, t, C, len(FAM)),
                           O.accept_uniforms(seed, t, C), eps, L, m.logp_grad)
    if np.any(np.abs(out['ratio'] - out['log_u']) < 1e-5):
      break
    n_ref += out['accept'].astype(int)
    z = out['z']
    assert _rel_err(p[t], z).max() < 1e-4, t
  else:
    np.testing.assert_array_equal(n_acc.cpu().numpy(), n_ref)
  assert t > 20


def _softplus_normal_moments(loc, scale):
  u = np.linspace(loc - 12 * scale, loc + 12 * scale, 200001)
  p = st.norm.pdf(u, loc, scale)
  th = O.softplus(u)
  m = np.trapezoid(p * th, u)
  return m, np.trapezoid(p * (th - m) ** 2, u)


def test_reference_hmc_custom():
  """``inference_auto_transform_test.py:112-136``."""
  x = TransformedDistribution(distribution=Normal(1.0, 1.0), bijector=B.Softplus())
  x.support = 'nonnegative'
  qx = Empirical(torch.as_tensor(np.random.default_rng(0).normal(size=1000).astype(np.float32)))
  inference = ed.HMC({x: qx})
  inference.initialize(auto_transform=True, step_size=0.8, seed=42)
  for _ in range(inference.n_iter):
    info_dict = inference.update()
  x_unconstrained = inference.transformations[x]
  qx_constrained_params = x_unconstrained.bijector.inverse(qx.params).cpu().numpy()
  x_mean, x_var = _softplus_normal_moments(1.0, 1.0)
  tail = qx_constrained_params[500:].astype(np.float64)
  np.testing.assert_allclose(tail.mean(), x_mean, rtol=1e-1, atol=1e-1)
  np.testing.assert_allclose(tail.var(), x_var, rtol=1e-1, atol=1e-1)
  assert 0.0 < info_dict['accept_rate'] <= 1.0


def test_reference_hmc_default():
  """``inference_auto_transform_test.py:138-160``: list form, zero start."""
  x = TransformedDistribution(distribution=Normal(1.0, 1.0), bijector=B.Softplus())
  x.support = 'nonnegative'
  inference = ed.HMC([x])
  inference.initialize(auto_transform=True, step_size=0.8, seed=3)
  for _ in range(inference.n_iter):
    info_dict = inference.update()
    inference.print_progress(info_dict)
  qx_constrained = inference.latent_vars[x]
  x_mean, x_var = _softplus_normal_moments(1.0, 1.0)
  tail = qx_constrained.params[500:].cpu().numpy().astype(np.float64)
  np.testing.assert_allclose(tail.mean(), x_mean, rtol=1e-1, atol=1e-1)
  np.testing.assert_allclose(tail.var(), x_var, rtol=1e-1, atol=1e-1)
  assert np.all(tail > 0)


def test_reference_hmc_betabernoulli():
  """``inference_auto_transform_test.py:162-188``: Beta(1, 1), ten
  Bernoulli draws; Beta(4, 8) posterior."""
  z = Beta(1., 1., name='z')
  xs = Bernoulli(probs=z, sample_shape=10)
  x_obs = np.asarray([0, 0, 1, 1, 0, 0, 0, 0, 0, 1], dtype=np.int32)
  qz_samples = torch.as_tensor(np.random.default_rng(1).random(1000).astype(np.float32))
  qz = Empirical(params=qz_samples, name='z_posterior')
  inference_hmc = ed.HMC({z: qz}, data={xs: x_obs})
  inference_hmc.run(step_size=1.0, n_steps=5, auto_transform=True, seed=11)
  z_unconstrained = inference_hmc.transformations[z]
  qz_constrained = z_unconstrained.bijector.inverse(qz.params).cpu().numpy().astype(np.float64)
  post = st.beta(4, 8)
  np.testing.assert_allclose(qz_constrained.mean(), post.mean(), rtol=5e-2, atol=5e-2)
  np.testing.assert_allclose(qz_constrained.var(), post.var(), rtol=1e-2, atol=1e-2)


def test_betabernoulli_many_chains():
  """The same model over 512 chains: posterior mean and variance of theta
  within 5e-3 / 1e-3 of Beta(4, 8), and the constrained view's moments."""
  z = Beta(1., 1.)
  xs = Bernoulli(probs=z, sample_shape=10)
  x_obs = np.asarray([0, 0, 1, 1, 0, 0, 0, 0, 0, 1], dtype=np.int32)
  qz = Empirical(np.zeros((400, 512), np.float32))
  inf = ed.HMC({z: qz}, data={xs: x_obs})
  inf.run(step_size=1.0, n_steps=5, auto_transform=True, seed=12)
  th = inf.latent_vars[z].params[100:].cpu().numpy().astype(np.float64)
  post = st.beta(4, 8)
  assert abs(th.mean() - post.mean()) < 5e-3
  assert abs(th.var() - post.var()) < 1e-3
  np.testing.assert_allclose(inf.latent_vars[z].mean(burn_in=100).cpu().numpy(),
                             th.mean(0), rtol=1e-5)


def test_gamma_weight_logistic_regression_matches_oracle_hmc():
  """A nonnegative weight (Gamma prior) beside Normal weights in a logistic
  regression: device HMC over 256 chains and the oracle's HMC give the same
  posterior means of the constrained weight within Monte Carlo error."""
  rng = np.random.default_rng(7)
  N = 500
  Xd = rng.normal(size=(N, 3)).astype(np.float32)
  eta = Xd.astype(np.float64) @ np.array([0.8, -0.5, 1.2])
  yd = (rng.random(N) < 1 / (1 + np.exp(-eta))).astype(np.int32)
  X = ed.placeholder('float32', [N, 2])
  w = Normal(loc=np.zeros(2, np.float32), scale=np.ones(2, np.float32))
  # the nonnegative coefficient multiplies the third column through a second design
  Xa = ed.placeholder('float32', [N, 1])
  A = Gamma(np.full(1, 2.0, np.float32), np.ones(1, np.float32))
  y = Bernoulli(logits=ed.dot(X, w) + ed.dot(Xa, A))
  C, T = 256, 300
  qw = Empirical(np.zeros((T, C, 2), np.float32))
  qa = Empirical(np.zeros((T, C, 1), np.float32))
  inf = ed.HMC({w: qw, A: qa}, data={X: Xd[:, :2], Xa: Xd[:, 2:], y: yd})
  inf.run(step_size=0.08, n_steps=10, auto_transform=True, seed=5)
  a_dev = inf.latent_vars[A].params[100:].cpu().numpy().astype(np.float64)[..., 0]
  w_dev = qw.params[100:].cpu().numpy().astype(np.float64)
  fam = [TO.FAMILY_NORMAL, TO.FAMILY_NORMAL, TO.FAMILY_GAMMA]
  op = [TO.OP_IDENTITY, TO.OP_IDENTITY, TO.OP_SOFTPLUS]
  m = TO.TransformedGLM(Xd.astype(np.float64), yd, fam, op, [0, 0, 2.0], [1, 1, 1.0])
  rng = np.random.default_rng(8)
  z = np.zeros((C, 3))
  draws = []
  for t in range(T):
    out = O.hmc_transition(z, rng.normal(size=(C, 3)), rng.random(C), 0.08, 10, m.logp_grad)
    z = out['z']
    if t >= 100:
      draws.append(z.copy())
  d = np.stack(draws)
  a_ref = O.softplus(d[..., 2])
  se = lambda x: x.std() * np.sqrt(20.0 / x.size) + 1e-3     # generous autocorrelation factor
  assert abs(a_dev.mean() - a_ref.mean()) < 4 * np.hypot(se(a_dev), se(a_ref))
  for j in range(2):
    assert abs(w_dev[..., j].mean() - d[..., j].mean()) < 4 * np.hypot(se(w_dev[..., j]),
                                                                    se(d[..., j]))
  assert np.all(a_dev > 0)


def test_fed_data_rebinds_with_the_prior_families():
  """A placeholder fed through update(feed_dict=...) rebinds the data; the
  constrained prior families are applied again (the log joint of the rebound
  engine equals the oracle's with the Gamma term)."""
  rng = np.random.default_rng(9)
  N = 300
  Xd = rng.normal(size=(N, 2)).astype(np.float32)
  yd = (rng.random(N) < 0.5).astype(np.int32)
  w = Normal(loc=np.zeros(1, np.float32), scale=np.ones(1, np.float32))
  A = Gamma(np.full(1, 2.0, np.float32), np.ones(1, np.float32))
  Xw = ed.placeholder('float32', [N, 1])
  Xa = ed.placeholder('float32', [N, 1])
  y = Bernoulli(logits=ed.dot(Xw, w) + ed.dot(Xa, A))
  qw = Empirical(np.zeros((5, 8, 1), np.float32))
  qa = Empirical(np.zeros((5, 8, 1), np.float32))
  inf = ed.HMC({w: qw, A: qa}, data={y: yd})
  inf.initialize(step_size=0.05, n_steps=3, auto_transform=True, seed=1)
  inf.update(feed_dict={Xw: Xd[:, :1], Xa: Xd[:, 1:]})
  inf.update(feed_dict={Xw: Xd[:, 1:], Xa: Xd[:, :1]})     # second feed: rebind
  U = np.array([[0.3, -0.5], [-1.0, 2.0]], np.float32)
  lp, _ = inf.engine.logp_grad(U)
  m = TO.TransformedGLM(np.concatenate([Xd[:, 1:], Xd[:, :1]], 1).astype(np.float64), yd,
                        [TO.FAMILY_NORMAL, TO.FAMILY_GAMMA], [TO.OP_IDENTITY, TO.OP_SOFTPLUS],
                        [0.0, 2.0], [1.0, 1.0])
  lp_ref, _ = m.logp_grad(U.astype(np.float64))
  np.testing.assert_allclose(lp.cpu().numpy(), lp_ref, rtol=1e-6, atol=1e-3)
