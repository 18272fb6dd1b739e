"""The auto-transform oracle (oracle/transform_oracle.py) and the host side
of the auto-transform (bijectors, ``transform``, the inference's
transformations, the compiled prior families), on the CPU.

What pins the oracle: scipy.stats log densities of the constrained priors;
quadrature of the unconstrained density (it integrates to one, which a
missing or wrong-signed log-Jacobian breaks, and reproduces the constrained
prior's closed-form moments); central finite differences of every gradient;
the Beta-Bernoulli conjugate posterior of the reference's
``tests/inferences/inference_auto_transform_test.py:167-188`` reached by an
oracle HMC run in the unconstrained space."""

import math

import numpy as np
import pytest
import scipy.stats as st
import torch

import edward_amd as ed
from edward_amd import bijectors as B
from edward_amd import compile as Cm
from edward_amd.models import (Bernoulli, Beta, Chi2, Empirical, Exponential, Gamma,
                               InverseGamma, Normal, TransformedDistribution)
from oracle import hmc_oracle as O
from oracle import transform_oracle as TO

FAMILIES = [
    (TO.FAMILY_GAMMA, 2.0, 0.5, 2.0 / 0.5, 2.0 / 0.25),
    (TO.FAMILY_GAMMA, 0.7, 3.0, 0.7 / 3.0, 0.7 / 9.0),
    (TO.FAMILY_INVGAMMA, 6.5, 2.0, 2.0 / 5.5, 4.0 / (5.5 ** 2 * 4.5)),
    (TO.FAMILY_BETA, 4.0, 8.0, 4.0 / 12.0, 32.0 / (144.0 * 13.0)),
    (TO.FAMILY_BETA, 0.8, 1.5, 0.8 / 2.3, 1.2 / (2.3 ** 2 * 3.3)),
]


def test_constrained_logpdf_matches_scipy():
  th = np.linspace(0.01, 0.99, 37)
  np.testing.assert_allclose(TO.constrained_logpdf(TO.FAMILY_BETA, th, 2.5, 0.7),
                             st.beta.logpdf(th, 2.5, 0.7), rtol=1e-12)
  x = np.linspace(0.05, 12.0, 41)
  np.testing.assert_allclose(TO.constrained_logpdf(TO.FAMILY_GAMMA, x, 2.0, 0.5),
                             st.gamma.logpdf(x, 2.0, scale=1 / 0.5), rtol=1e-12)
  np.testing.assert_allclose(TO.constrained_logpdf(TO.FAMILY_INVGAMMA, x, 3.5, 2.0),
                             st.invgamma.logpdf(x, 3.5, scale=2.0), rtol=1e-12)


@pytest.mark.parametrize('family,a,b,mean,var', FAMILIES)
def test_unconstrained_density_normalises_and_keeps_the_moments(family, a, b, mean, var):
  """exp(log p_u) over u must integrate to 1 and push forward to the
  constrained prior: E[theta] and Var[theta] equal the closed forms."""
  u = np.linspace(-60.0, 60.0, 400001)
  lp, _ = TO.prior_logp_grad(u, family, a, b)
  p = np.exp(lp)
  th = TO.to_constrained(family, u)
  Z = np.trapezoid(p, u)
  m = np.trapezoid(p * th, u)
  v = np.trapezoid(p * (th - m) ** 2, u)
  assert abs(Z - 1.0) < 1e-6
  np.testing.assert_allclose(m, mean, rtol=1e-5)
  np.testing.assert_allclose(v, var, rtol=1e-4)


def test_wrong_jacobian_breaks_normalisation():
  """The quadrature check has teeth: dropping the log-Jacobian term leaves a
  density that does not integrate to one."""
  u = np.linspace(-60.0, 60.0, 200001)
  th = TO.to_constrained(TO.FAMILY_GAMMA, u)
  p = np.exp(TO.constrained_logpdf(TO.FAMILY_GAMMA, th, 2.0, 0.5))
  assert abs(np.trapezoid(p, u) - 1.0) > 0.1


@pytest.mark.parametrize('family,a,b', [(TO.FAMILY_NORMAL, 0.3, 1.7), (TO.FAMILY_GAMMA, 2.0, 0.5),
                                        (TO.FAMILY_GAMMA, 0.6, 2.0),
                                        (TO.FAMILY_INVGAMMA, 3.5, 2.0),
                                        (TO.FAMILY_BETA, 4.0, 8.0), (TO.FAMILY_BETA, 0.5, 0.5)])
def test_prior_gradient_finite_differences(family, a, b):
  u = np.array([-12.0, -3.0, -0.7, 0.0, 0.4, 2.5, 9.0])
  h = 1e-5
  _, g = TO.prior_logp_grad(u, family, a, b)
  lp_p, _ = TO.prior_logp_grad(u + h, family, a, b)
  lp_m, _ = TO.prior_logp_grad(u - h, family, a, b)
  np.testing.assert_allclose(g, (lp_p - lp_m) / (2 * h), rtol=1e-6, atol=1e-6)


def test_transformed_glm_brute_force_and_gradient():
  rng = np.random.default_rng(0)
  N, D = 40, 5
  X = rng.normal(size=(N, D))
  y = (rng.random(N) < 0.5).astype(np.float64)
  fam = [TO.FAMILY_NORMAL, TO.FAMILY_GAMMA, TO.FAMILY_BETA, TO.FAMILY_INVGAMMA, TO.FAMILY_NORMAL]
  op = [TO.OP_IDENTITY, TO.OP_SOFTPLUS, TO.OP_SIGMOID, TO.OP_SOFTPLUS, TO.OP_SOFTPLUS]
  p1 = [0.1, 2.0, 3.0, 2.5, 1.0]
  p2 = [1.5, 0.5, 2.0, 1.0, 0.8]
  m = TO.TransformedGLM(X, y, fam, op, p1, p2)
  U = rng.normal(size=(3, D))
  lp, g = m.logp_grad(U)
  # brute force: loops over rows and coordinates with scipy densities
  for c in range(3):
    th = [U[c, 0], math.log1p(math.exp(U[c, 1])), 1 / (1 + math.exp(-U[c, 2])),
          math.log1p(math.exp(U[c, 3])), math.log1p(math.exp(U[c, 4]))]
    want = 0.0
    for n in range(N):
      eta = sum(X[n, d] * th[d] for d in range(D))
      want += st.bernoulli.logpmf(y[n], 1 / (1 + math.exp(-eta)))
    want += st.norm.logpdf(U[c, 0], 0.1, 1.5)
    want += st.gamma.logpdf(th[1], 2.0, scale=2.0) + math.log(1 / (1 + math.exp(-U[c, 1])))
    want += st.beta.logpdf(th[2], 3.0, 2.0) + math.log(th[2] * (1 - th[2]))
    want += st.invgamma.logpdf(th[3], 2.5, scale=1.0) + math.log(1 / (1 + math.exp(-U[c, 3])))
    want += st.norm.logpdf(U[c, 4], 1.0, 0.8)   # Normal pushed through softplus
    assert abs(lp[c] - want) < 1e-9 * max(1.0, abs(want))
  h = 1e-6
  for d in range(D):
    e = np.zeros(D)
    e[d] = h
    fd = (m.logp(U + e) - m.logp(U - e)) / (2 * h)
    np.testing.assert_allclose(g[:, d], fd, rtol=1e-6, atol=1e-6)


def test_beta_bernoulli_posterior_by_oracle_hmc():
  """``inference_auto_transform_test.py:167-188``: Beta(1, 1) prior, ten
  Bernoulli draws with three ones; HMC in the logit space (step 1.0, 5
  leapfrog steps) recovers Beta(4, 8): mean to 5e-2, variance to 1e-2."""
  x_obs = np.array([0, 0, 1, 1, 0, 0, 0, 0, 0, 1], np.float64)
  m = TO.TransformedGLM(np.ones((10, 1)), x_obs, TO.FAMILY_BETA, TO.OP_IDENTITY, 1.0, 1.0)
  rng = np.random.default_rng(3)
  C, T = 2000, 300
  z = rng.random((C, 1))
  draws = []
  for t in range(T):
    out = O.hmc_transition(z, rng.normal(size=(C, 1)), rng.random(C), 1.0, 5, m.logp_grad)
    z = out['z']
    if t >= 100:
      draws.append(O.sigmoid(z[:, 0]))
  d = np.concatenate(draws)
  post = st.beta(4, 8)
  assert abs(d.mean() - post.mean()) < 5e-3
  assert abs(d.var() - post.var()) < 1e-3


# ----------------------------------------------------------------------------
# bijectors and transform (host side of the product)
# ----------------------------------------------------------------------------


def _fd_log_jac(f, x, h=1e-6):
  x = np.asarray(x, np.float64)
  return np.log(np.abs((f(x + h) - f(x - h)) / (2 * h)))


def test_bijector_maps_and_log_jacobians():
  x = torch.linspace(-8, 8, 33, dtype=torch.float64)
  sp, sg = B.Softplus(), B.Sigmoid()
  np.testing.assert_allclose(sp.forward(x).numpy(), O.softplus(x.numpy()), rtol=1e-14)
  np.testing.assert_allclose(sp.inverse(sp.forward(x)).numpy(), x.numpy(), atol=1e-9)
  np.testing.assert_allclose(sg.inverse(sg.forward(x)).numpy(), x.numpy(), atol=1e-9)
  np.testing.assert_allclose(TO.softplus_inverse(O.softplus(x.numpy())), x.numpy(), atol=1e-9)
  np.testing.assert_allclose(TO.logit(O.sigmoid(x.numpy())), x.numpy(), atol=1e-9)
  np.testing.assert_allclose(sp.forward_log_det_jacobian(x).numpy(),
                             _fd_log_jac(O.softplus, x.numpy()), atol=1e-7)
  np.testing.assert_allclose(sg.forward_log_det_jacobian(x).numpy(),
                             _fd_log_jac(O.sigmoid, x.numpy()), atol=1e-6)
  y = sp.forward(x)
  np.testing.assert_allclose(sp.inverse_log_det_jacobian(y).numpy(),
                             -sp.forward_log_det_jacobian(x).numpy(), atol=1e-12)
  inv = B.Invert(sp)
  np.testing.assert_allclose(inv.forward(y).numpy(), x.numpy(), atol=1e-9)
  np.testing.assert_allclose(inv.forward_log_det_jacobian(y).numpy(),
                             sp.inverse_log_det_jacobian(y).numpy(), atol=1e-12)


def test_softmax_centered_log_jacobian_is_the_restricted_determinant():
  rng = np.random.default_rng(5)
  x = rng.normal(size=4)
  b = B.SoftmaxCentered()
  y = b.forward(torch.as_tensor(x)).numpy()
  np.testing.assert_allclose(y, TO.softmax_centered_forward(x), rtol=1e-13)
  assert abs(y.sum() - 1) < 1e-14 and y.shape == (5,)
  np.testing.assert_allclose(b.inverse(torch.as_tensor(y)).numpy(), x, atol=1e-12)
  h = 1e-6
  J = np.empty((4, 4))
  for j in range(4):
    e = np.zeros(4)
    e[j] = h
    J[:, j] = (TO.softmax_centered_forward(x + e)[:4] - TO.softmax_centered_forward(x - e)[:4]) / (2 * h)
  want = np.linalg.slogdet(J)[1]
  np.testing.assert_allclose(b.forward_log_det_jacobian(torch.as_tensor(x)).item(), want,
                             atol=1e-7)
  np.testing.assert_allclose(TO.softmax_centered_fldj(x), want, atol=1e-7)
  assert b.forward_event_shape((4,)) == (5,)


def test_transform_by_support():
  g = Gamma(2.0, 0.5)
  tg = ed.transform(g)
  assert isinstance(tg, TransformedDistribution) and tg.support == 'real'
  assert tg.bijector.name == 'invert_softplus'
  assert ed.transform(Beta(1.0, 1.0)).bijector.name == 'invert_sigmoid'
  n = Normal(0.0, 1.0)
  assert ed.transform(n) is n
  td = TransformedDistribution(Normal(0.0, 1.0), B.Softplus())
  with pytest.raises(AttributeError):
    ed.transform(td)           # no support assigned
  td.support = 'nonnegative'
  assert ed.transform(td).distribution is td
  assert ed.transform(n, bijector=B.Softplus()).bijector.name == 'softplus'
  with pytest.raises(ValueError):
    x = Bernoulli(probs=0.5)
    ed.transform(x)


def test_auto_transform_bookkeeping():
  """``inference.py:222-264``: transformations, latent_vars_unconstrained
  (the user's store) and latent_vars (the constrained view)."""
  x = TransformedDistribution(Normal(1.0, 1.0), B.Softplus())
  x.support = 'nonnegative'
  w = Normal(0.0, 1.0)
  qx = Empirical(np.array([0.0, -1.0, 2.0], np.float32))
  qw = Empirical(np.zeros(3, np.float32))
  inf = ed.HMC({x: qx, w: qw})
  inf._auto_transform(True)
  xu = inf.transformations[x]
  assert xu is not x and inf.transformations[w] is w
  assert inf.latent_vars_unconstrained[xu] is qx and inf.latent_vars_unconstrained[w] is qw
  assert inf.latent_vars[w] is qw
  view = inf.latent_vars[x]
  np.testing.assert_allclose(view.params.cpu().numpy(), O.softplus(np.array([0.0, -1.0, 2.0])),
                             rtol=1e-6)
  np.testing.assert_allclose(xu.bijector.inverse(qx.params).cpu().numpy(),
                             view.params.cpu().numpy(), rtol=1e-6)
  inf._auto_transform(False)
  assert not hasattr(inf, 'latent_vars_unconstrained') and inf.latent_vars[x] is qx


def test_compiled_prior_families():
  X = ed.placeholder('float32', [6, 3])
  a = Gamma(2.0, 0.5)
  b = Beta([3.0, 1.0], 2.0)
  e = Exponential(1.5)
  w = Normal(np.zeros(3, np.float32), np.ones(3, np.float32))
  yv = Bernoulli(logits=ed.dot(X, w) + a + e)
  lat = {w: Empirical(np.zeros((4, 3), np.float32)), a: Empirical(np.zeros(4, np.float32)),
         e: Empirical(np.zeros(4, np.float32))}
  spec = Cm.glm_spec(lat, {X: np.ones((6, 3)), yv: np.ones(6)},
                     transforms={a: 'softplus', e: 'softplus'})
  kind = spec.prior_kind
  assert list(kind) == [0, 0, 0, 1 | 16, 1 | 16]
  np.testing.assert_allclose(spec.loc[3:], [2.0, 1.0])
  np.testing.assert_allclose(spec.scale[3:], [0.5, 1.5])
  np.testing.assert_allclose(spec.prior_lnorm[3], 2 * math.log(0.5) - math.lgamma(2.0))
  np.testing.assert_allclose(spec.prior_lnorm[4], math.log(1.5))
  np.testing.assert_allclose(spec.prior_lnorm[:3], -0.5 * math.log(2 * math.pi))
  # Beta-Bernoulli: probs of a sigmoid-transformed latent -> logits = u
  z = Beta(1.0, 1.0)
  xs = Bernoulli(probs=z, sample_shape=10)
  spec = Cm.glm_spec({z: Empirical(np.zeros(5, np.float32))}, {xs: np.arange(10) % 2},
                     transforms={z: 'sigmoid'})
  assert spec.X.shape == (10, 1) and np.all(spec.X == 1)
  assert list(spec.prior_kind) == [2]                    # Beta family, operand u
  np.testing.assert_allclose(spec.prior_lnorm, [0.0], atol=1e-12)
  # vector probabilities tile the rows
  zb = Beta(np.ones(3, np.float32), 2.0)
  xb = Bernoulli(probs=zb, sample_shape=4)
  spec = Cm.glm_spec({zb: Empirical(np.zeros((5, 3), np.float32))}, {xb: np.zeros((4, 3))},
                     transforms={zb: 'sigmoid'})
  np.testing.assert_array_equal(spec.X, np.tile(np.eye(3), (4, 1)))
  # no observed variable: one zero row, likelihood weight 0
  t = TransformedDistribution(Normal(1.0, 1.0), B.Softplus())
  t.support = 'nonnegative'
  spec = Cm.glm_spec({t: Empirical(np.zeros(5, np.float32))}, {}, transforms={t: 'softplus'})
  assert spec.lik_weight == 0.0 and spec.X.shape == (1, 1) and not spec.X.any()
  assert list(spec.prior_kind) == [16]                   # Normal on u, softplus operand
  # Chi2 and InverseGamma
  c, ig = Chi2(3.0), InverseGamma(2.0, 1.0)
  yy = Normal(loc=ed.dot(X, w) + c + ig, scale=1.0)
  spec = Cm.glm_spec({w: Empirical(np.zeros((4, 3), np.float32)),
                      c: Empirical(np.zeros(4, np.float32)),
                      ig: Empirical(np.zeros(4, np.float32))},
                     {X: np.ones((6, 3)), yy: np.ones(6)},
                     transforms={c: 'softplus', ig: 'softplus'})
  assert list(spec.prior_kind[3:]) == [1 | 16, 3 | 16]
  np.testing.assert_allclose(spec.loc[3:], [1.5, 2.0])
  np.testing.assert_allclose(spec.scale[3:], [0.5, 1.0])


def test_constrained_latents_need_hmc_and_support():
  z = Beta(1.0, 1.0)
  xs = Bernoulli(probs=z, sample_shape=10)
  with pytest.raises(NotImplementedError):   # probs of an untransformed latent
    Cm.glm_spec({z: Empirical(np.zeros(5, np.float32))}, {xs: np.zeros(10)})
  d = ed.models.TransformedDistribution(Normal(0.0, 1.0), B.Sigmoid())
  d.support = 'nonnegative'
  yv = Normal(loc=d * 1.0, scale=1.0, sample_shape=3)
  with pytest.raises(NotImplementedError):   # bijector does not match the support
    Cm.glm_spec({d: Empirical(np.zeros(5, np.float32))}, {yv: np.zeros(3)},
                transforms={d: 'softplus'})
  inf = ed.SGLD({z: Empirical(np.zeros(5, np.float32))}, data={xs: np.zeros(10)})
  inf._auto_transform(True)
  with pytest.raises(NotImplementedError):
    inf._compile_transforms()


def _hier_data(seed=3, sizes=(5, 0, 9, 3), D=3):
  rng = np.random.default_rng(seed)
  off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
  N = int(off[-1])
  X = rng.standard_normal((N, D))
  y = (rng.random(N) < 0.5).astype(np.float64)
  return X, y, off


def test_scaled_hier_reduces_to_the_log_sigma_model():
  X, y, off = _hier_data()
  base = O.HierLogReg(X, y, off)
  m = TO.ScaledHierLogReg(X, y, off, 'exp', TO.FAMILY_NORMAL, 0.0, 1.0)
  th = np.random.default_rng(1).standard_normal((4, base.dim))
  lp0, g0 = base.logp_grad(th)
  lp1, g1 = m.logp_grad(th)
  np.testing.assert_allclose(lp1, lp0, rtol=1e-13)
  np.testing.assert_allclose(g1, g0, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('smap,family,a,b', [
    ('exp', TO.FAMILY_NORMAL, 0.3, 0.7),
    ('softplus', TO.FAMILY_GAMMA, 2.0, 3.0),
    ('softplus', TO.FAMILY_INVGAMMA, 3.0, 1.5),
    ('softplus', TO.FAMILY_NORMAL, -0.4, 0.6),
])
def test_scaled_hier_scipy_log_joint_and_gradient(smap, family, a, b):
  """log p against scipy.stats in the constrained space plus the log-Jacobian
  of sigma(u), and the gradient against central differences."""
  X, y, off = _hier_data()
  m = TO.ScaledHierLogReg(X, y, off, smap, family, a, b)
  G = m.G
  rng = np.random.default_rng(7)
  th = 0.6 * rng.standard_normal((3, m.dim))
  lp, g = m.logp_grad(th)
  for c in range(th.shape[0]):
    av, mu, u, w = th[c, :G], th[c, G], th[c, G + 1], th[c, G + 2:]
    sig = math.exp(u) if smap == 'exp' else math.log1p(math.exp(u))
    eta = av[m.group] + X @ w
    ref = st.bernoulli.logpmf(y, 1.0 / (1.0 + np.exp(-eta))).sum()
    ref += st.norm.logpdf(av, mu, sig).sum() + st.norm.logpdf(mu) + st.norm.logpdf(w).sum()
    if family == TO.FAMILY_NORMAL:
      ref += st.norm.logpdf(u, a, b)          # the density of u itself
    elif family == TO.FAMILY_GAMMA:
      ref += st.gamma.logpdf(sig, a, scale=1.0 / b) + math.log(1.0 / (1.0 + math.exp(-u)))
    else:
      ref += st.invgamma.logpdf(sig, a, scale=b) + math.log(1.0 / (1.0 + math.exp(-u)))
    np.testing.assert_allclose(lp[c], ref, rtol=1e-10)
  h = 1e-6
  for j in range(m.dim):
    e = np.zeros(m.dim)
    e[j] = h
    fd = (m.logp_grad(th + e)[0] - m.logp_grad(th - e)[0]) / (2 * h)
    np.testing.assert_allclose(g[:, j], fd, rtol=1e-6, atol=1e-6)


def test_hier_scale_prior_kinds():
  """The group-scale prior the hierarchical compiler hands the device:
  kind = family | map << 4 with map 0 for exp(log_sigma), 1 for a softplus-
  transformed positive latent, and the host-side normalising constant."""
  from edward_amd import hier as H
  ls = Normal(loc=0.5, scale=2.0)
  kind, p1, p2, lnorm = H._scale_prior(ed.exp(ls), {ls: None}, {})[1:]
  assert (kind, p1, p2) == (TO.FAMILY_NORMAL, 0.5, 2.0)
  np.testing.assert_allclose(lnorm, -math.log(2.0) - 0.5 * math.log(2 * math.pi), rtol=1e-6)
  for rv, fam, a, b in [(Gamma(concentration=2.0, rate=3.0), TO.FAMILY_GAMMA, 2.0, 3.0),
                        (Exponential(rate=1.5), TO.FAMILY_GAMMA, 1.0, 1.5),
                        (Chi2(df=3.0), TO.FAMILY_GAMMA, 1.5, 0.5),
                        (InverseGamma(concentration=3.0, rate=1.5), TO.FAMILY_INVGAMMA, 3.0,
                         1.5)]:
    sc, kind, p1, p2, lnorm = H._scale_prior(rv, {rv: None}, {rv: 'softplus'})
    assert sc is rv and kind == fam | (1 << 4) and (p1, p2) == (a, b)
    # lnorm + the u-density's other terms integrate to one over u
    # (the InverseGamma's sigma^-4 tail needs the long second range)
    mass = 0.0
    for u in (np.linspace(-40.0, 40.0, 400001), np.linspace(40.0, 4000.0, 400001)):
      mass += np.trapezoid(np.exp(TO.prior_logp_grad(u, fam, a, b)[0]), u)
    np.testing.assert_allclose(mass, 1.0, rtol=1e-6)
    np.testing.assert_allclose(lnorm, a * math.log(b) - math.lgamma(a), rtol=1e-6)
  sp = TransformedDistribution(distribution=Normal(loc=-0.4, scale=0.6), bijector=B.Softplus())
  assert H._scale_prior(sp, {sp: None}, {sp: 'softplus'})[1:3] == (TO.FAMILY_NORMAL | 16,
                                                                   np.float32(-0.4))
  g = Gamma(concentration=2.0, rate=3.0)
  with pytest.raises(NotImplementedError):
    H._scale_prior(g, {g: None}, {})                # not auto-transformed
  bt = Beta(concentration1=2.0, concentration0=2.0)
  with pytest.raises(NotImplementedError):
    H._scale_prior(bt, {bt: None}, {bt: 'sigmoid'})
  with pytest.raises(NotImplementedError):
    H._scale_prior(ed.exp(Gamma(concentration=2.0, rate=1.0)), {}, {})
