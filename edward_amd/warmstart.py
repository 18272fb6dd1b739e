"""Warm start of many chains from the Laplace approximation of a GLM.

The mode of the log joint and the precision matrix there come from damped
Newton steps whose terms (log joint, gradient, X^T diag(s) X) are computed
by ``edhmc_newton_terms``; N(mode, P^-1) is the Gaussian that Edward's
``Laplace`` inference fits (``edward/inferences/laplace.py:17-137``).  Chains
started from its draws begin near the typical set, so the HMC burn-in that
would carry them there from the prior is not needed, and exact gradients at
its draws are training pairs for the gradient surrogate (positions along
exact Hamiltonian trajectories have the posterior as their marginal, which
the Gaussian approximates).
"""

import numpy as np
import torch


class LaplaceApproximation(object):
  """N(mode, P^-1) of a bound ``GLMEngine``'s log joint."""

  def __init__(self, engine, w0=None, max_iter=30, tol=1e-6):
    self.engine = engine
    self.mode, self.precision, self.logp_mode, self.iterations = engine.laplace(
        w0, max_iter=max_iter, tol=tol)
    chol = np.linalg.cholesky(self.precision)
    # a draw is mode + L^-T e (row form: e^T L^-1), Cov = (L L^T)^-1 = P^-1
    root = np.linalg.solve(chol, np.eye(chol.shape[0]))
    dev = engine.device
    self._mode = torch.as_tensor(self.mode, dtype=torch.float32, device=dev)
    self._root = torch.as_tensor(root, dtype=torch.float32, device=dev)

  def sample(self, n, seed=0):
    """n draws [n, D] (float32, on the engine's device): standard normal rows
    from the library's Philox stream keyed by ``seed``, mapped by
    mode + e L^-1 on the device (``edhmc_affine_rows``)."""
    e = self.engine.draw_momenta(int(n), int(seed), 0)
    return self.engine.affine_rows(e, self._root, self._mode)

  def gradient_pairs(self, n, batch, seed=0):
    """(W, G): n draws and the exact gradient of the log joint at each,
    evaluated ``batch`` positions at a time.  Uses the engine's chain state
    as scratch (set the chains' start afterwards)."""
    D = self._mode.numel()
    W = torch.empty(n, D, device=self._mode.device)
    G = torch.empty(n, D, device=self._mode.device)
    for i, b0 in enumerate(range(0, n, batch)):
      b1 = min(n, b0 + batch)
      w = self.sample(b1 - b0, seed=int(seed) * 1000003 + i)
      _, g = self.engine.logp_grad(w)
      W[b0:b1] = w
      G[b0:b1] = g
    return W, G
