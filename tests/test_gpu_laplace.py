"""GPU parity of the Newton terms (edhmc_newton_terms) and the Laplace warm
start against the oracle's FP64 log joint, gradient and negative Hessian
(oracle/hmc_oracle.py GLM.logp_grad, GLM.neg_hessian, laplace_logreg).

Bars: log p |d| <= 1e-6 |logp| + 1e-3; gradient within 1e-4 of max|grad|;
negative Hessian within 1e-5 of max|P| (fp32 products, fp32 sums over row
slices, FP64 across slices); bf16 X rounded on both sides.  Mode within
1e-5 of the oracle's mode (max norm, relative to its largest entry)."""

import numpy as np
import pytest
import torch

from oracle import hmc_oracle as O

pytestmark = pytest.mark.gpu


def _case(seed, N, D, likelihood='bernoulli', bf16=False):
  rng = np.random.default_rng(seed)
  X = rng.normal(size=(N, D)).astype(np.float32)
  if bf16:
    X = torch.as_tensor(X).to(torch.bfloat16).float().numpy()
  w_true = rng.normal(size=D) * 0.3
  eta = X.astype(np.float64) @ w_true
  if likelihood == 'bernoulli':
    y = (rng.random(N) < 1 / (1 + np.exp(-eta))).astype(np.float32)
  else:
    y = (eta + 0.5 * rng.normal(size=N)).astype(np.float32)
  loc = (rng.normal(size=D) * 0.1).astype(np.float32)
  scale = rng.uniform(0.5, 2.0, size=D).astype(np.float32)
  from edward_amd.engine import GLMEngine
  Xin = torch.as_tensor(X).to(torch.bfloat16) if bf16 else X
  eng = GLMEngine(Xin, y, loc, scale, likelihood=likelihood, lik_scale=0.5,
                  precision='bf16' if bf16 else 'fp32')
  m = O.GLM(X.astype(np.float64), y, loc, scale, likelihood, 0.5)
  return eng, m, w_true


@pytest.mark.parametrize('N,D,likelihood,bf16', [(1000, 10, 'bernoulli', False),
                                                 (5003, 77, 'bernoulli', False),
                                                 (4000, 33, 'normal', False),
                                                 (20000, 130, 'bernoulli', True)])
def test_newton_terms_match_oracle(N, D, likelihood, bf16):
  eng, m, w_true = _case(1, N, D, likelihood, bf16)
  w = (w_true + 0.05 * np.random.default_rng(2).normal(size=D)).astype(np.float32)
  lp, g, P = eng.newton_terms(w)
  lp_ref, g_ref = m.logp_grad(w.astype(np.float64))
  P_ref = m.neg_hessian(w.astype(np.float64))
  assert abs(lp - lp_ref[0]) <= 1e-6 * abs(lp_ref[0]) + 1e-3
  assert np.abs(g - g_ref[0]).max() <= 1e-4 * np.abs(g_ref).max()
  assert np.abs(P - P_ref).max() <= 1e-5 * np.abs(P_ref).max()


def test_laplace_mode_matches_oracle():
  eng, m, _ = _case(3, 6000, 40)
  mode, P, lp, it = eng.laplace()
  ref_mode, ref_cov = O.laplace_logreg(m)
  assert it < 15
  assert np.abs(mode - ref_mode).max() <= 1e-5 * np.abs(ref_mode).max()
  np.testing.assert_allclose(np.linalg.inv(P), ref_cov, rtol=1e-4, atol=1e-4 * np.abs(ref_cov).max())
  # the gradient vanishes at the mode
  _, g, _ = eng.newton_terms(mode)
  assert np.abs(g).max() < 1e-3 * np.sqrt(np.diag(P)).max()


def test_laplace_draws_and_gradient_pairs():
  """Draws of N(mode, P^-1) have its mean and covariance (within Monte Carlo
  error); the pairs hold the oracle's gradient at each draw."""
  from edward_amd.warmstart import LaplaceApproximation
  eng, m, _ = _case(4, 3000, 12)
  lap = LaplaceApproximation(eng)
  n = 40000
  W = lap.sample(n, seed=1).cpu().numpy().astype(np.float64)
  cov = np.linalg.inv(lap.precision)
  sd = np.sqrt(np.diag(cov))
  assert np.all(np.abs(W.mean(0) - lap.mode) < 5 * sd / np.sqrt(n))
  corr_err = (np.cov(W.T) - cov) / np.outer(sd, sd)
  assert np.abs(corr_err).max() < 5 * np.sqrt(2.0 / n)
  Wp, Gp = lap.gradient_pairs(1000, 256, seed=2)
  _, g_ref = m.logp_grad(Wp.cpu().numpy().astype(np.float64))
  assert np.abs(Gp.cpu().numpy() - g_ref).max() <= 1e-4 * np.abs(g_ref).max()


def test_newton_terms_reject_what_they_do_not_model():
  """Constrained prior families and a likelihood weight other than 1 are not
  modelled by the Newton terms: the call raises instead of returning a wrong
  Hessian."""
  from edward_amd import _lib as L
  eng, m, w_true = _case(5, 2000, 8)
  eng.set_lik_weight(0.5)
  with pytest.raises(L.EdhmcError) as err:
    eng.newton_terms(w_true.astype(np.float32))
  assert err.value.code == L.EUNSUPPORTED
  eng.set_lik_weight(1.0)
  eng.set_prior_family(np.full(8, 16, np.uint8), np.full(8, -0.9189385332046727))
  with pytest.raises(L.EdhmcError) as err:
    eng.newton_terms(w_true.astype(np.float32))
  assert err.value.code == L.EUNSUPPORTED
  eng.set_prior_family(None)
  lp, g, P = eng.newton_terms(w_true.astype(np.float32))
  assert np.isfinite(lp) and np.all(np.isfinite(P))
