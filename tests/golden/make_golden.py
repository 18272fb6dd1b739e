"""Writes the golden vectors under tests/golden/ from the FP64 oracle.

The reference ships no golden vectors for the HMC path (its tests are
statistical: tests/inferences/hmc_test.py:33-35,79-80), so these fixtures hold
oracle outputs on fixed inputs, pinned in turn by the oracle's closed-form,
finite-difference and brute-force tests (tests/test_oracle.py).

  logreg_c1.npz  c1 shape: X[1000x10] ~ N(0,1), w* ~ N(0,I), y ~ Bernoulli,
                 prior N(0,I); 4 chains, fed momenta/uniforms, eps=0.01, L=10.
  hier_small.npz hierarchical logistic regression, 9 ragged groups (empty and
                 single-row ones included), D=5, 3 chains, eps=0.02, L=6.

Run:  python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import hmc_oracle as O  # noqa: E402


def logreg_c1():
  rng = np.random.default_rng(0)
  N, D, C = 1000, 10, 4
  X = rng.standard_normal((N, D)).astype(np.float32)
  w_true = rng.normal(size=D)
  y = (rng.random(N) < O.sigmoid(X.astype(np.float64) @ w_true)).astype(np.float32)
  W0 = (w_true[None, :] + 0.05 * rng.normal(size=(C, D))).astype(np.float32)
  r0 = rng.normal(size=(C, D)).astype(np.float32)
  u = rng.random(C)
  eps, L = 0.01, 10
  m = O.GLM(X.astype(np.float64), y, 0.0, 1.0)
  lp0, g0 = m.logp_grad(W0.astype(np.float64))
  out = O.hmc_transition(W0.astype(np.float64), r0.astype(np.float64), u, eps, L, m.logp_grad)
  np.savez_compressed(os.path.join(HERE, 'logreg_c1.npz'), X=X, y=y, W0=W0, r0=r0, u=u,
                      eps=eps, L=L, logp0=lp0, grad0=g0, zL=out['zL'], rL=out['rL'],
                      logpL=out['logpL'], ratio=out['ratio'], accept=out['accept'],
                      z=out['z'])


def hier_small():
  rng = np.random.default_rng(1)
  sizes = np.array([12, 0, 1, 30, 7, 0, 22, 3, 15])
  off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
  N, D, C, G = int(off[-1]), 5, 3, len(sizes)
  X = rng.standard_normal((N, D)).astype(np.float32)
  a = rng.normal(0.2, 0.7, size=G)
  w = rng.normal(size=D) * 0.5
  eta = a[np.repeat(np.arange(G), sizes)] + X.astype(np.float64) @ w
  y = (rng.random(N) < O.sigmoid(eta)).astype(np.float32)
  m = O.HierLogReg(X.astype(np.float64), y, off)
  th0 = (rng.normal(size=(C, m.dim)) * 0.3).astype(np.float32)
  r0 = rng.normal(size=(C, m.dim)).astype(np.float32)
  u = rng.random(C)
  eps, L = 0.02, 6
  lp0, g0 = m.logp_grad(th0.astype(np.float64))
  out = O.hmc_transition(th0.astype(np.float64), r0.astype(np.float64), u, eps, L,
                         m.logp_grad)
  np.savez_compressed(os.path.join(HERE, 'hier_small.npz'), X=X, y=y, offsets=off, th0=th0,
                      r0=r0, u=u, eps=eps, L=L, logp0=lp0, grad0=g0, zL=out['zL'],
                      rL=out['rL'], logpL=out['logpL'], ratio=out['ratio'],
                      accept=out['accept'], z=out['z'])


if __name__ == '__main__':
  logreg_c1()
  hier_small()
  print('wrote', sorted(f for f in os.listdir(HERE) if f.endswith('.npz')))
