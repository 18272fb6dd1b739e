"""GPU parity of the hierarchical model's group-scale prior against the oracle.

The intercepts are a_g ~ N(mu, sigma^2) with sigma = exp(u) and
u = log_sigma ~ Normal(m, s), or sigma = softplus(u) for a positive latent
sigma (Gamma, InverseGamma, or a Normal pushed through Softplus) sampled in
its unconstrained coordinate u with the log-Jacobian
(``inference.py:222-264``, ``hmc.py:132-159``).  The oracle is
``oracle.transform_oracle.ScaledHierLogReg`` (pinned by scipy.stats and finite
differences in tests/test_transform_oracle.py).

Bars (those of tests/test_gpu_hier.py): log joint |dlogp| <= 1e-6 |logp| +
1e-3, gradient within 1e-4 of max|grad|, leapfrog end positions within 1e-4
relative per chain.
"""

import numpy as np
import pytest
import torch

from edward_amd import compile as Cm
from oracle import hmc_oracle as O
from oracle import transform_oracle as TO

pytestmark = pytest.mark.gpu

SCALE_PRIORS = [
    # (map, family, p1, p2)
    ('exp', TO.FAMILY_NORMAL, 0.3, 0.7),
    ('softplus', TO.FAMILY_GAMMA, 2.0, 3.0),
    ('softplus', TO.FAMILY_INVGAMMA, 3.0, 1.5),
    ('softplus', TO.FAMILY_NORMAL, -0.4, 0.6),
]


def _data(seed, sizes, D):
  rng = np.random.default_rng(seed)
  off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
  N = int(off[-1])
  X = rng.normal(size=(N, D)).astype(np.float32)
  y = (rng.random(N) < 0.5).astype(np.float32)
  return X, y, off


def _pair(X, y, off, smap, family, p1, p2):
  from edward_amd.hier import HierEngine
  eng = HierEngine(X, y, off)
  fam = Cm.PriorFamily(family, p1, p2, 0, 1)
  kind = family | ((1 if smap == 'softplus' else 0) << 4)
  eng.set_scale_prior(kind, float(fam.p1[0]), float(fam.p2[0]), float(fam.lnorm[0]))
  m = TO.ScaledHierLogReg(X.astype(np.float64), y, off, smap, family,
                          float(fam.p1[0]), float(fam.p2[0]))
  return eng, m


@pytest.mark.parametrize('smap,family,p1,p2', SCALE_PRIORS)
@pytest.mark.parametrize('sizes,D,C', [
    ([3, 0, 2, 4, 1], 3, 4),
    (list(np.random.default_rng(5).poisson(30, size=120)), 16, 37),
])
def test_hier_scale_prior_logp_grad(smap, family, p1, p2, sizes, D, C):
  X, y, off = _data(1, sizes, D)
  eng, m = _pair(X, y, off, smap, family, p1, p2)
  rng = np.random.default_rng(2)
  th = (rng.normal(size=(C, m.dim)) * 0.5).astype(np.float32)
  th[:, m.G + 1] = rng.uniform(-2.5, 2.5, size=C)   # small and large scales
  lp, g = eng.logp_grad(th)
  lp_ref, g_ref = m.logp_grad(th.astype(np.float64))
  lp = lp.cpu().numpy()
  g = g.cpu().numpy().astype(np.float64)
  assert np.all(np.abs(lp - lp_ref) <= 1e-6 * np.abs(lp_ref) + 1e-3), (lp, lp_ref)
  tol = 1e-4 * np.abs(g_ref).max()
  assert np.abs(g - g_ref).max() <= tol, (np.abs(g - g_ref).max(), tol)


@pytest.mark.parametrize('smap,family,p1,p2', SCALE_PRIORS[1:3])
def test_hier_scale_prior_transition_positions(smap, family, p1, p2):
  X, y, off = _data(7, list(np.random.default_rng(8).poisson(20, size=40)), 12)
  eng, m = _pair(X, y, off, smap, family, p1, p2)
  C, eps, L = 9, 0.01, 6
  rng = np.random.default_rng(9)
  th0 = (rng.normal(size=(C, m.dim)) * 0.3).astype(np.float32)
  r0 = rng.normal(size=(C, m.dim)).astype(np.float32)
  eng.set_state(th0)
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  out_t = torch.empty(C, m.dim, device='cuda')
  eng.transition(eps, L, momenta=r0, uniforms=np.zeros(C), accepted_out=acc, sample_out=out_t)
  ref = O.hmc_transition(th0.astype(np.float64), r0.astype(np.float64), np.zeros(C), eps, L,
                         m.logp_grad)
  assert acc.cpu().numpy().all()
  d = np.linalg.norm(out_t.cpu().numpy().astype(np.float64) - ref['zL'], axis=1)
  assert (d / np.linalg.norm(ref['zL'], axis=1)).max() < 1e-4


def test_hier_scale_prior_rejects_bad_kinds():
  from edward_amd._lib import EdhmcError
  X, y, off = _data(3, [4, 5], 2)
  eng, _ = _pair(X, y, off, 'exp', TO.FAMILY_NORMAL, 0.0, 1.0)
  for kind, p1, p2 in [(TO.FAMILY_BETA | 16, 2.0, 2.0),   # no Beta scale
                       (2 << 4, 0.0, 1.0),                # no sigmoid map
                       (TO.FAMILY_GAMMA | 16, -1.0, 1.0),  # concentration <= 0
                       (TO.FAMILY_NORMAL, 0.0, 0.0)]:     # scale <= 0
    with pytest.raises(EdhmcError):
      eng.set_scale_prior(kind, p1, p2, 0.0)


def test_hier_gamma_scale_edward_surface():
  """``sigma ~ Gamma`` as the intercepts' scale under ed.HMC: the positive
  latent is auto-transformed, the store holds u = softplus^-1(sigma), and the
  log joint of its last row on the engine equals the oracle's."""
  import edward_amd as ed
  from edward_amd.hier import HierEngine
  from edward_amd.models import Bernoulli, Empirical, Gamma, Normal
  rng = np.random.default_rng(31)
  G, D, N, C, T = 6, 3, 240, 4, 40
  group_np = rng.integers(0, G, size=N)
  X_np = rng.normal(size=(N, D)).astype(np.float32)
  y_np = (rng.random(N) < 0.5).astype(np.float32)
  X = ed.placeholder('float32', [N, D])
  group = ed.placeholder('int32', [N])
  mu = Normal(loc=0.0, scale=1.0)
  sigma = Gamma(concentration=2.0, rate=2.0)
  a = Normal(loc=mu, scale=sigma, sample_shape=G)
  w = Normal(loc=np.zeros(D, np.float32), scale=np.ones(D, np.float32))
  yv = Bernoulli(logits=ed.gather(a, group) + ed.dot(X, w))
  qa = Empirical(np.zeros((T, C, G), np.float32))
  qmu = Empirical(np.zeros((T, C), np.float32))
  qs = Empirical(np.zeros((T, C), np.float32))
  qw = Empirical(np.zeros((T, C, D), np.float32))
  inf = ed.HMC({a: qa, mu: qmu, sigma: qs, w: qw},
               data={X: X_np, group: group_np, yv: y_np})
  inf.initialize(step_size=0.05, n_steps=5, seed=3)
  for _ in range(T):
    inf.update()
  assert isinstance(inf.engine, HierEngine)
  u = qs.params[-1].cpu().numpy()
  assert np.isfinite(u).all() and np.abs(u).max() > 0
  # the pushed-forward view holds sigma = softplus(u) > 0
  sig_view = inf.latent_vars[sigma].params[-1].cpu().numpy()
  np.testing.assert_allclose(sig_view, np.log1p(np.exp(u.astype(np.float64))), rtol=1e-5)
  vals = torch.cat([qa.params[-1], qmu.params[-1][:, None], qs.params[-1][:, None],
                    qw.params[-1]], dim=1).cpu().numpy()
  order = np.argsort(group_np, kind='stable')
  off = np.concatenate([[0], np.cumsum(np.bincount(group_np, minlength=G))])
  m = TO.ScaledHierLogReg(X_np[order].astype(np.float64), y_np[order], off, 'softplus',
                          TO.FAMILY_GAMMA, 2.0, 2.0)
  lp, _ = inf.engine.logp_grad(vals)
  lp_ref, _ = m.logp_grad(vals.astype(np.float64))
  np.testing.assert_allclose(lp.cpu().numpy(), lp_ref, rtol=1e-6, atol=1e-3)
