"""Pins the FP64 oracle before it is trusted as the checker.

Each check is chosen so that a plausible slip in the oracle (a dropped term, a
wrong sign or index, a swapped kick/drift order) fails it.  The reference has
no golden vectors for this path (its HMC tests are statistical:
``tests/inferences/hmc_test.py:33-46,79-91``), so the pins are closed forms,
brute force on tiny inputs, finite differences, published known-answer
vectors, and the reference's own statistical tests run seeded.
"""

import math

import numpy as np
import pytest
from scipy import stats

from oracle import hmc_oracle as O


# --------------------------------------------------------------------------
# Philox4x32-10: Random123 known-answer vectors
# --------------------------------------------------------------------------

def test_philox_known_answers():
  kat = [
      ((0, 0, 0, 0), (0, 0),
       (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
      ((0xffffffff,) * 4, (0xffffffff,) * 2,
       (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
      ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
       (0xa4093822, 0x299f31d0),
       (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
  ]
  for ctr, key, want in kat:
    got = O.philox4x32_10(np.array(ctr, np.uint64), np.array(key, np.uint64))
    assert tuple(int(v) for v in got) == want


def test_momentum_and_uniform_streams_are_distributed_right():
  r = O.momentum_normals(seed=7, counter=3, C=64, D=250)
  assert r.shape == (64, 250)
  assert abs(r.mean()) < 0.02 and abs(r.var() - 1) < 0.03
  assert stats.kstest(r.ravel(), 'norm').pvalue > 1e-3
  u = O.accept_uniforms(seed=7, counter=3, C=20000)
  assert u.min() >= 0 and u.max() < 1
  assert stats.kstest(u, 'uniform').pvalue > 1e-3
  # different counters / streams give different draws
  assert not np.allclose(r, O.momentum_normals(7, 4, 64, 250))
  assert not np.allclose(u[:64], O.accept_uniforms(8, 3, 64))


# --------------------------------------------------------------------------
# Log joints
# --------------------------------------------------------------------------

def _tiny_logreg(seed=0, N=7, D=3, C=4):
  rng = np.random.default_rng(seed)
  X = rng.normal(size=(N, D))
  y = (rng.random(N) < 0.5).astype(np.float64)
  W = rng.normal(size=(C, D))
  loc = rng.normal(size=D) * 0.3
  scale = rng.uniform(0.5, 2.0, size=D)
  return X, y, W, loc, scale


def test_logreg_logp_matches_bruteforce_and_scipy():
  X, y, W, loc, scale = _tiny_logreg()
  m = O.GLM(X, y, loc, scale, 'bernoulli')
  lp, _ = m.logp_grad(W)
  for c in range(W.shape[0]):
    tot = 0.0
    for d in range(X.shape[1]):
      tot += stats.norm.logpdf(W[c, d], loc[d], scale[d])
    for n in range(X.shape[0]):
      eta = sum(X[n, d] * W[c, d] for d in range(X.shape[1]))
      p = 1.0 / (1.0 + math.exp(-eta))
      tot += stats.bernoulli.logpmf(int(y[n]), p)
    assert lp[c] == pytest.approx(tot, rel=1e-12, abs=1e-12)


def test_normal_likelihood_matches_scipy():
  X, _, W, loc, scale = _tiny_logreg(1)
  y = np.random.default_rng(2).normal(size=X.shape[0])
  m = O.GLM(X, y, loc, scale, 'normal', lik_scale=0.3)
  lp, _ = m.logp_grad(W)
  want = (stats.norm.logpdf(W, loc, scale).sum(1)
          + stats.norm.logpdf(y[None, :], W @ X.T, 0.3).sum(1))
  np.testing.assert_allclose(lp, want, rtol=1e-12)


@pytest.mark.parametrize('lik', ['bernoulli', 'normal'])
def test_glm_gradient_matches_finite_differences(lik):
  X, y, W, loc, scale = _tiny_logreg(3)
  if lik == 'normal':
    y = np.random.default_rng(4).normal(size=X.shape[0])
  m = O.GLM(X, y, loc, scale, lik, lik_scale=0.7)
  _, g = m.logp_grad(W)
  h = 1e-6
  for d in range(X.shape[1]):
    e = np.zeros_like(W)
    e[:, d] = h
    fd = (m.logp(W + e) - m.logp(W - e)) / (2 * h)
    np.testing.assert_allclose(g[:, d], fd, rtol=1e-6, atol=1e-6)


def test_logreg_special_cases():
  # eta = 0: each row contributes -log 2 whatever y; gradient rows y - 1/2
  X = np.zeros((5, 2))
  y = np.array([0, 1, 1, 0, 1.0])
  m = O.GLM(X, y, 0.0, 1.0)
  W = np.zeros((1, 2))
  lp, g = m.logp_grad(W)
  assert lp[0] == pytest.approx(-5 * math.log(2) - 2 * 0.5 * O.LOG_2PI)
  np.testing.assert_allclose(g, 0.0)
  # large |eta|: log-lik -> -|eta| for the wrong label, -> 0 for the right one
  Xb = np.array([[1.0]])
  for eta in (40.0, 700.0):
    m1 = O.GLM(Xb, [0.0], 0.0, 1e9)
    assert m1.loglik(np.array([[eta]]))[0] == pytest.approx(-eta, rel=1e-12)
    m2 = O.GLM(Xb, [1.0], 0.0, 1e9)
    assert abs(m2.loglik(np.array([[eta]]))[0]) < 1e-15
    assert np.isfinite(m2.logp_grad(np.array([[-eta]]))[1]).all()


def test_glm_grad_bf16_roundings():
  """At W = 0 every residual y - 1/2 is a bf16 number and the prior
  gradient vanishes: the rounded gradient is the exact one.  Elsewhere the
  rounded residuals move each gradient entry by at most 2^-8 sum |x r|, and
  the one-plane form differs from the two-plane form (its positions are
  rounded)."""
  rng = np.random.default_rng(9)
  X = rng.normal(size=(300, 7))
  y = (rng.random(300) < 0.4).astype(np.float64)
  m = O.GLM(X, y, 0.0, 1.0)
  W0 = np.zeros((1, 7))
  np.testing.assert_array_equal(m.grad_bf16(W0), m.logp_grad(W0)[1])
  W = rng.normal(size=(3, 7)) * 0.3
  _, g = m.logp_grad(W)
  res = y - O.sigmoid(W @ X.T)
  bound = 2.0 ** -8 * (np.abs(res) @ np.abs(X)) + 1e-3 * np.abs(W).max() * np.abs(X).sum(0)
  g2 = m.grad_bf16(W, planes=2)
  assert np.all(np.abs(g2 - g) <= bound)
  assert np.abs(g2 - g).max() > 0
  g1 = m.grad_bf16(W, planes=1)
  assert np.abs(g1 - g2).max() > 0


def test_hier_gradient_matches_finite_differences_and_loops():
  rng = np.random.default_rng(5)
  G, D = 4, 3
  sizes = np.array([3, 0, 2, 4])
  off = np.concatenate([[0], np.cumsum(sizes)])
  N = off[-1]
  X = rng.normal(size=(N, D))
  y = (rng.random(N) < 0.5).astype(float)
  m = O.HierLogReg(X, y, off)
  th = rng.normal(size=(2, m.dim)) * 0.5
  lp, g = m.logp_grad(th)
  # brute-force log joint
  for c in range(2):
    a, mu, ls, w = th[c, :G], th[c, G], th[c, G + 1], th[c, G + 2:]
    tot = stats.norm.logpdf(mu) + stats.norm.logpdf(ls)
    tot += stats.norm.logpdf(w).sum()
    tot += stats.norm.logpdf(a, mu, math.exp(ls)).sum()
    for gi in range(G):
      for n in range(off[gi], off[gi + 1]):
        eta = a[gi] + X[n] @ w
        tot += y[n] * eta - math.log1p(math.exp(eta))
    assert lp[c] == pytest.approx(tot, rel=1e-12)
  h = 1e-6
  for j in range(m.dim):
    e = np.zeros_like(th)
    e[:, j] = h
    fd = (m.logp_grad(th + e)[0] - m.logp_grad(th - e)[0]) / (2 * h)
    np.testing.assert_allclose(g[:, j], fd, rtol=1e-6, atol=1e-6)


# --------------------------------------------------------------------------
# Leapfrog
# --------------------------------------------------------------------------

@pytest.mark.parametrize('L', [1, 3, 10])
def test_leapfrog_matches_gaussian_closed_form(L):
  mean, sd, eps = 0.7, 1.3, 0.21
  tgt = O.GaussianTarget(mean, sd)
  z0, r0 = np.array([[2.0]]), np.array([[-0.4]])
  zL, rL, _, _ = O.leapfrog(z0, r0, eps, L, tgt.logp_grad)
  M = O.gaussian_leapfrog_matrix(eps, sd, L)
  want = M @ np.array([z0[0, 0] - mean, r0[0, 0]])
  assert zL[0, 0] - mean == pytest.approx(want[0], rel=1e-13)
  assert rL[0, 0] == pytest.approx(want[1], rel=1e-13)


def test_leapfrog_reversible_and_volume_preserving():
  X, y, W, loc, scale = _tiny_logreg(6, N=9, D=2, C=1)
  m = O.GLM(X, y, loc, scale)
  eps, L = 0.05, 7
  r0 = np.array([[0.3, -1.1]])
  zL, rL, _, _ = O.leapfrog(W, r0, eps, L, m.logp_grad)
  zB, rB, _, _ = O.leapfrog(zL, -rL, eps, L, m.logp_grad)
  np.testing.assert_allclose(zB, W, atol=1e-12)
  np.testing.assert_allclose(-rB, r0, atol=1e-12)
  # Jacobian determinant of (z, r) -> (zL, rL) by central differences
  x0 = np.concatenate([W[0], r0[0]])
  h = 1e-6

  def flow(x):
    z, r, _, _ = O.leapfrog(x[None, :2], x[None, 2:], eps, L, m.logp_grad)
    return np.concatenate([z[0], r[0]])
  J = np.stack([(flow(x0 + h * e) - flow(x0 - h * e)) / (2 * h)
                for e in np.eye(4)], axis=1)
  assert np.linalg.det(J) == pytest.approx(1.0, abs=1e-6)


def test_leapfrog_energy_error_is_second_order():
  X, y, W, loc, scale = _tiny_logreg(7, N=30, D=3, C=1)
  m = O.GLM(X, y, loc, scale)
  r0 = np.array([[0.5, -0.2, 0.9]])
  T = 0.5

  def herr(eps):
    L = int(round(T / eps))
    zL, rL, lpL, _ = O.leapfrog(W, r0, eps, L, m.logp_grad)
    H0 = -m.logp(W)[0] + 0.5 * (r0 ** 2).sum()
    H1 = -lpL[0] + 0.5 * (rL ** 2).sum()
    return abs(H1 - H0)
  e1, e2 = herr(0.02), herr(0.01)
  assert 2.5 < e1 / e2 < 6.0   # halving eps divides the error by ~4


def test_planted_kick_order_mistake_is_caught():
  """A drift-kick-drift integrator (a plausible slip) fails the closed form."""
  mean, sd, eps, L = 0.0, 1.0, 0.3, 4
  tgt = O.GaussianTarget(mean, sd)

  def wrong(z, r, eps, L, lg):
    for _ in range(L):
      z = z + 0.5 * eps * r
      r = r + eps * lg(z)[1]
      z = z + 0.5 * eps * r
    return z, r
  z0, r0 = np.array([[1.0]]), np.array([[0.5]])
  M = O.gaussian_leapfrog_matrix(eps, sd, L)
  want = M @ np.array([1.0, 0.5])
  zw, rw = wrong(z0, r0, eps, L, tgt.logp_grad)
  assert abs(zw[0, 0] - want[0]) > 1e-3 or abs(rw[0, 0] - want[1]) > 1e-3


# --------------------------------------------------------------------------
# Accept rule, transition and run-loop semantics
# --------------------------------------------------------------------------

def test_accept_rule_special_cases():
  assert O.accept_rule(math.log(0.999), 0.0)            # ratio >= 0 accepts
  assert not O.accept_rule(0.0, 0.0)                     # u = 1 boundary
  assert not O.accept_rule(-1.0, float('nan'))           # NaN rejects
  assert O.accept_rule(-math.inf, -1e30)                 # u = 0 accepts
  assert not O.accept_rule(-0.5, -0.5)                   # strict <


def test_transition_ratio_and_selection():
  X, y, W, loc, scale = _tiny_logreg(8, N=20, D=3, C=5)
  m = O.GLM(X, y, loc, scale)
  rng = np.random.default_rng(9)
  r0 = rng.normal(size=W.shape)
  u = np.array([0.0, 0.3, 0.999999, 0.5, 1e-300])
  out = O.hmc_transition(W, r0, u, 0.05, 5, m.logp_grad)
  zL, rL, lpL, _ = O.leapfrog(W, r0, 0.05, 5, m.logp_grad)
  ratio = 0.5 * (r0 ** 2).sum(1) - 0.5 * (rL ** 2).sum(1) + lpL - m.logp(W)
  np.testing.assert_allclose(out['ratio'], ratio, rtol=1e-13, atol=1e-12)
  for c in range(5):
    acc = math.log(u[c]) < ratio[c] if u[c] > 0 else True
    assert out['accept'][c] == acc
    np.testing.assert_array_equal(out['z'][c], zL[c] if acc else W[c])


def test_run_loop_overwrites_row0_and_counts_accepts():
  tgt = O.GaussianTarget(0.0, 1.0)
  T, C, D = 6, 3, 2
  params = np.full((T, C, D), 5.0)
  rng = np.random.default_rng(10)
  mom = rng.normal(size=(T, C, D))
  uni = np.ones((T, C)) * 0.5
  uni[2, 1] = 1.0          # log u = 0; rejects unless ratio > 0
  n_acc = O.run_hmc(params, mom, uni, 0.1, 3, tgt.logp_grad)
  # row 0 holds the first sample, not the initial 5.0 (hmc.py:81-85,121-126)
  first = O.hmc_transition(np.full((C, D), 5.0), mom[0], uni[0], 0.1, 3,
                           tgt.logp_grad)
  np.testing.assert_array_equal(params[0], first['z'])
  # row t was produced from row t-1
  for t in range(1, T):
    o = O.hmc_transition(params[t - 1], mom[t], uni[t], 0.1, 3, tgt.logp_grad)
    np.testing.assert_array_equal(params[t], o['z'])
  assert n_acc.shape == (C,) and (n_acc <= T).all() and n_acc.sum() > 0


def test_empirical_moments_divide_by_T():
  p = np.random.default_rng(11).normal(size=(9, 4, 3))
  np.testing.assert_allclose(O.empirical_mean(p), p.mean(0))
  np.testing.assert_allclose(O.empirical_stddev(p), p.std(0, ddof=0))
  np.testing.assert_allclose(O.empirical_variance(p), p.var(0, ddof=0))


def test_normal_normal_posterior_seeded():
  """Port of hmc_test.py:14-46, seeded: 50 zeros, prior N(0,1) -> N(0, 1/51).

  Run with many chains so the check is far tighter than the reference's
  rtol = atol = 0.1."""
  X = np.ones((50, 1))
  m = O.GLM(X, np.zeros(50), 0.0, 1.0, 'normal', lik_scale=1.0)
  T, C = 400, 64
  rng = np.random.default_rng(12)
  params = np.ones((T, C, 1))
  n_acc = O.run_hmc(params, rng.normal(size=(T, C, 1)), rng.random((T, C)),
                    0.25, 2, m.logp_grad)
  s = params[100:].reshape(-1)
  assert abs(s.mean()) < 0.01
  assert s.std() == pytest.approx(math.sqrt(1 / 51), rel=0.05)
  assert (n_acc > 0.1).all()


# --------------------------------------------------------------------------
# ESS
# --------------------------------------------------------------------------

@pytest.mark.parametrize('rho', [0.0, 0.5, 0.9, -0.3])
def test_ess_ar1_closed_form(rho):
  rng = np.random.default_rng(13)
  M, n = 8, 20000
  x = np.zeros((M, n))
  e = rng.normal(size=(M, n))
  x[:, 0] = e[:, 0] / math.sqrt(1 - rho ** 2)
  for t in range(1, n):
    x[:, t] = rho * x[:, t - 1] + e[:, t]
  want = M * n * (1 - rho) / (1 + rho)
  got = O.ess(x)
  if rho < 0:   # antithetic chains: ESS > M n, estimator is conservative
    assert got > M * n * 0.9
  else:
    assert got == pytest.approx(want, rel=0.1)


@pytest.mark.parametrize('rho', [0.0, 0.7])
def test_within_chain_ess_ar1_and_stuck_chains(rho):
  rng = np.random.default_rng(16)
  M, n = 16, 5000
  x = np.zeros((M, n))
  e = rng.normal(size=(M, n))
  x[:, 0] = e[:, 0] / math.sqrt(1 - rho ** 2)
  for t in range(1, n):
    x[:, t] = rho * x[:, t - 1] + e[:, t]
  assert O.ess(x, between=False) == pytest.approx(M * n * (1 - rho) / (1 + rho), rel=0.1)
  # chains frozen at spread-out starting points: no sampling happened
  frozen = np.repeat(rng.normal(size=(M, 1)), n, axis=1)
  assert O.ess(frozen, between=False) == 0.0
  assert O.ess(frozen, between=True) > 0     # the BDA3 form credits the spread


def test_min_ess_takes_worst_coordinate():
  rng = np.random.default_rng(14)
  T, C = 2000, 4
  s = rng.normal(size=(T, C, 3))
  s[:, :, 1] = np.cumsum(s[:, :, 1], axis=0) * 0.01 + s[:, :, 1]
  e = [O.ess(s[:, :, d].T) for d in range(3)]
  assert O.min_ess(s) == min(e) and e[1] < 0.5 * e[0]


def test_laplace_matches_long_hmc_on_small_logreg():
  rng = np.random.default_rng(15)
  N, D = 2000, 3
  X = rng.normal(size=(N, D))
  y = (rng.random(N) < O.sigmoid(X @ np.array([0.5, -1.0, 0.2]))).astype(float)
  m = O.GLM(X, y, 0.0, 1.0)
  w, cov = O.laplace_logreg(m)
  _, g = m.logp_grad(w[None])
  assert np.abs(g).max() < 1e-8
  T, C = 300, 32
  params = np.zeros((T, C, D)) + w
  O.run_hmc(params, rng.normal(size=(T, C, D)), rng.random((T, C)), 0.02, 10,
            m.logp_grad)
  s = params[50:].reshape(-1, D)
  se = np.sqrt(np.diag(cov) / (s.shape[0] / 10))
  assert np.all(np.abs(s.mean(0) - w) < 5 * se)
  np.testing.assert_allclose(s.var(0), np.diag(cov), rtol=0.15)


# --------------------------------------------------------------------------
# Stochastic-gradient samplers
# --------------------------------------------------------------------------

def test_sgld_gaussian_stationary_variance_closed_form():
  """Constant-rate SGLD on N(0, s^2) is the linear recursion
  z' = (1 - lr/(2 s^2)) z + sqrt(lr) xi with stationary variance
  s^2 / (1 - lr/(4 s^2)); the oracle step must reproduce it."""
  rng = np.random.default_rng(3)
  s2, lr, C = 2.0, 0.4, 200000
  z = np.zeros(C)
  for _ in range(200):
    z = O.sgld_step(z, -z / s2, lr, rng.normal(size=C))
  want = s2 / (1.0 - lr / (4.0 * s2))
  assert abs(z.var() / want - 1.0) < 0.02
  assert O.sgld_learning_rate(0.25, 0) == 0.25
  assert O.sgld_learning_rate(1.0, 3) == pytest.approx(4.0 ** -0.55)


def test_sghmc_gaussian_stationary_moments():
  """SGHMC with friction f on N(0, s^2): (z, v) is a linear AR(1) system;
  its stationary covariance solves the discrete Lyapunov equation."""
  from scipy import linalg
  rng = np.random.default_rng(4)
  s2, lr, f, C = 1.5, 0.05, 0.3, 100000
  z, v = np.zeros(C), np.zeros(C)
  for _ in range(2000):
    z, v = O.sghmc_step(z, v, -z / s2, lr, f, rng.normal(size=C))
  # state (z, v): z' = z + v ; v' = -lr/s2 z + (1 - f/2) v + sqrt(lr f) xi
  A = np.array([[1.0, 1.0], [-lr / s2, 1.0 - 0.5 * f]])
  Q = np.array([[0.0, 0.0], [0.0, lr * f]])
  P = linalg.solve_discrete_lyapunov(A, Q)
  assert abs(z.var() / P[0, 0] - 1.0) < 0.03
  assert abs(v.var() / P[1, 1] - 1.0) < 0.03


def test_mh_random_walk_targets_the_gaussian():
  """Random-walk MH on N(0, s^2 I) keeps the target: the stationary variance
  of many chains after burn-in is s^2; the symmetric proposal's density terms
  cancel, so the oracle's ratio is the log-density difference alone."""
  rng = np.random.default_rng(7)
  s2, C, D = 0.5, 20000, 2
  logp = lambda z: -0.5 * (z ** 2).sum(axis=1) / s2
  z = np.zeros((C, D))
  lz = logp(z)
  acc_tot = 0
  for _ in range(300):
    z, lz, acc, _ = O.mh_step(z, lz, 0.6, rng.normal(size=(C, D)), rng.random(C), logp)
    acc_tot += acc.mean()
  assert abs(z.var() / s2 - 1.0) < 0.03
  assert 0.2 < acc_tot / 300 < 0.95


def test_replica_exchange_keeps_the_tempered_targets():
  """MH at inverse temperatures (1, 0.25) on N(0, 1) plus the exchange move:
  replica i must stay distributed as N(0, 1/beta_i) (the exchange ratio's
  sign or temperature factor wrong would distort both), and exchanges must
  happen."""
  rng = np.random.default_rng(11)
  betas = np.array([1.0, 0.25])
  L = 4000
  C = 2 * L
  beta_c = np.repeat(betas, L)
  logp = lambda z: -0.5 * (z ** 2).sum(axis=1)
  z = np.zeros((C, 1))
  lz = logp(z)
  swaps = 0
  for t in range(400):
    z, lz, _, _ = O.mh_step(z, lz, 1.2 / np.sqrt(beta_c)[:, None], rng.normal(size=(C, 1)),
                            rng.random(C), logp, beta_c)
    perm = O.replica_exchange(lz, betas, L, 5, t, 0.5)
    swaps += int((perm != np.arange(C)).sum())
    z, lz = z[perm], lz[perm]
  assert abs(z[:L].var() - 1.0) < 0.08
  assert abs(z[L:].var() - 4.0) < 0.35
  assert swaps > 0


@pytest.mark.parametrize('likelihood', ['bernoulli', 'normal'])
def test_glm_neg_hessian_is_minus_the_gradient_jacobian(likelihood):
  """The Newton terms of the warm start: -Hessian by central differences of
  the oracle's gradient; symmetric positive definite."""
  rng = np.random.default_rng(21)
  N, D = 60, 5
  X = rng.normal(size=(N, D))
  y = (rng.random(N) < 0.5).astype(np.float64) if likelihood == 'bernoulli' else rng.normal(size=N)
  m = O.GLM(X, y, rng.normal(size=D) * 0.1, rng.uniform(0.5, 2.0, size=D), likelihood, 0.7)
  w = rng.normal(size=D) * 0.3
  P = m.neg_hessian(w)
  h = 1e-6
  J = np.empty((D, D))
  for j in range(D):
    e = np.zeros(D)
    e[j] = h
    J[:, j] = (m.logp_grad(w + e)[1][0] - m.logp_grad(w - e)[1][0]) / (2 * h)
  np.testing.assert_allclose(P, -J, rtol=1e-6, atol=1e-6)
  np.testing.assert_allclose(P, P.T, atol=1e-12)
  assert np.linalg.eigvalsh(P).min() > 0


def test_posterior_moments_is_matches_quadrature():
  """A skewed two-dimensional logistic posterior (30 rows, large weights):
  the importance-sampled moments agree with grid quadrature within 4
  standard errors, while the Laplace approximation's own moments are off
  by many (a dropped weight would not pass)."""
  rng = np.random.default_rng(21)
  N = 30
  X = rng.normal(size=(N, 2)) * np.array([1.0, 2.0])
  w_true = np.array([1.5, -1.0])
  y = (rng.random(N) < O.sigmoid(X @ w_true)).astype(np.float64)
  m = O.GLM(X, y, 0.0, 1.0)
  mode, cov = O.laplace_logreg(m)
  sd = np.sqrt(np.diag(cov))
  # quadrature on a grid of +-9 Laplace sd
  n = 601
  g0 = np.linspace(mode[0] - 9 * sd[0], mode[0] + 9 * sd[0], n)
  g1 = np.linspace(mode[1] - 9 * sd[1], mode[1] + 9 * sd[1], n)
  G = np.stack(np.meshgrid(g0, g1, indexing='ij'), axis=-1).reshape(-1, 2)
  lp = m.logp(G)
  p = np.exp(lp - lp.max())
  p /= p.sum()
  q_mean = p @ G
  q_var = p @ (G - q_mean) ** 2
  est = O.posterior_moments_is(m, mode, cov, 20000, seed=3)
  assert est['weight_ess'] > 0.5 * 40000
  assert np.all(np.abs(est['mean'] - q_mean) < 4 * est['se_mean']), (est['mean'], q_mean)
  assert np.all(np.abs(est['var'] - q_var) < 4 * est['se_var']), (est['var'], q_var)
  assert np.all(est['se_mean'] < 0.02 * np.sqrt(q_var))
  # the Laplace moments themselves are far outside those errors
  assert np.any(np.abs(mode - q_mean) > 10 * est['se_mean'])
  assert np.any(np.abs(np.diag(cov) - q_var) > 4 * est["se_var"])


def test_bf16_inner_positions_amplify_rounding_differences():
  """Why a few chains of the bf16 engine leave the rounded oracle's leapfrog
  by 1e-5..1e-4 while the rest agree to ~1e-7: the inner gradients are
  evaluated at the bf16 hi plane of the position, so where a coordinate lies
  within rounding distance of a bf16 tie, positions that differ by one fp32
  ulp (the GPU's fp32 and the oracle's FP64 arithmetic) round to
  neighbouring bf16 numbers and the evaluated position moves by 2^-8 of that
  coordinate.  Shown with the oracle alone (N=20000, D=128, 128 chains, eps
  = 0.002, L = 8): starts differing by one fp32 ulp in half the coordinates
  stay 4e-8 apart under the exact-gradient leapfrog, while under the rounded
  leapfrog (GLM.grad_bf16) one chain ends ~1e-4 apart and the median stays
  below 1e-7."""
  from edward_amd import workloads as W
  N, D, C, eps, L = 20000, 128, 128, 0.002, 8
  wl = W.logreg(N, D, C, eps, L, 10, seed=41, bf16=True)
  m = O.GLM(wl.X.astype(np.float64), wl.y, wl.prior_loc, wl.prior_scale)
  rng = np.random.default_rng(3)
  z0 = (wl.w_true[None, :] + 0.01 * rng.normal(size=(C, D))).astype(np.float32)
  r0 = rng.normal(size=(C, D))
  z0b = np.where(rng.random((C, D)) < 0.5, np.nextafter(z0, np.float32(np.inf)), z0)
  e32 = float(np.float32(eps))

  def leapfrog(z, rounded):
    z = z.astype(np.float64)
    r = r0.copy()
    g = m.grad_bf16(z, 2) if rounded else m.logp_grad(z)[1]
    for k in range(L):
      r = r + 0.5 * e32 * g
      z = z + e32 * r
      g = m.grad_bf16(z, 2 if k == L - 1 else 1) if rounded else m.logp_grad(z)[1]
      r = r + 0.5 * e32 * g
    return z

  def rel(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
  d_exact = rel(leapfrog(z0b, False), leapfrog(z0, False))
  d_round = rel(leapfrog(z0b, True), leapfrog(z0, True))
  assert d_exact.max() < 1e-7
  assert np.median(d_round) < 1e-7
  assert d_round.max() > 1e-5 and d_round.max() > 100 * np.median(d_round)
  assert (d_round > 1e-5).sum() <= 4          # isolated chains


def test_logreg_grad_f32_matches_the_fp64_gradient():
  """The fp32 CPU restatement the benchmark times equals the FP64 gradient to
  fp32 accuracy (threads only split the elementwise part)."""
  rng = np.random.default_rng(0)
  X = rng.normal(size=(3000, 16)).astype(np.float32)
  y = (rng.random(3000) < 0.5).astype(np.float32)
  W = (rng.normal(size=(5, 16)) * 0.3).astype(np.float32)
  g = O.logreg_grad_f32(X, y, W, threads=3)
  _, ref = O.GLM(X.astype(np.float64), y, 0.0, 1.0).logp_grad(W.astype(np.float64))
  assert g.dtype == np.float32
  assert np.abs(g - ref).max() < 1e-5 * np.abs(ref).max()
