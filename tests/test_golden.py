"""Golden vectors (tests/golden/*.npz, written by tests/golden/make_golden.py).

CPU: the oracle reproduces the committed vectors (rtol 1e-12), so an edit
that changes the oracle's arithmetic shows up here.
GPU: libedhmc.so on the same inputs against the committed vectors, with the
bars of the device parity tests: log joint |d| <= 1e-6 |logp| + 1e-3,
gradient within 1e-4 of max|grad|, leapfrog end positions within 1e-4
relative per chain, identical accept decisions.
"""

import os

import numpy as np
import pytest

from oracle import hmc_oracle as O

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _load(name):
  return dict(np.load(os.path.join(HERE, name)))


def _rel(a, b):
  return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-6)


def test_oracle_reproduces_logreg_golden():
  d = _load('logreg_c1.npz')
  m = O.GLM(d['X'].astype(np.float64), d['y'], 0.0, 1.0)
  lp, g = m.logp_grad(d['W0'].astype(np.float64))
  np.testing.assert_allclose(lp, d['logp0'], rtol=1e-12)
  np.testing.assert_allclose(g, d['grad0'], rtol=1e-12, atol=1e-12)
  out = O.hmc_transition(d['W0'].astype(np.float64), d['r0'].astype(np.float64), d['u'],
                         float(d['eps']), int(d['L']), m.logp_grad)
  np.testing.assert_allclose(out['zL'], d['zL'], rtol=1e-12, atol=1e-14)
  np.testing.assert_array_equal(out['accept'], d['accept'])


def test_oracle_reproduces_hier_golden():
  d = _load('hier_small.npz')
  m = O.HierLogReg(d['X'].astype(np.float64), d['y'], d['offsets'])
  lp, g = m.logp_grad(d['th0'].astype(np.float64))
  np.testing.assert_allclose(lp, d['logp0'], rtol=1e-12)
  np.testing.assert_allclose(g, d['grad0'], rtol=1e-12, atol=1e-12)
  out = O.hmc_transition(d['th0'].astype(np.float64), d['r0'].astype(np.float64), d['u'],
                         float(d['eps']), int(d['L']), m.logp_grad)
  np.testing.assert_allclose(out['zL'], d['zL'], rtol=1e-12, atol=1e-14)
  np.testing.assert_array_equal(out['accept'], d['accept'])


def _check_lp_g(lp, g, d):
  lp = lp.cpu().numpy()
  g = g.cpu().numpy().astype(np.float64)
  assert np.all(np.abs(lp - d['logp0']) <= 1e-6 * np.abs(d['logp0']) + 1e-3)
  assert np.abs(g - d['grad0']).max() <= 1e-4 * np.abs(d['grad0']).max()


@pytest.mark.gpu
def test_device_logreg_golden():
  import torch
  from edward_amd.engine import GLMEngine
  d = _load('logreg_c1.npz')
  C, D = d['W0'].shape
  eng = GLMEngine(d['X'], d['y'], 0.0, 1.0)
  _check_lp_g(*eng.logp_grad(d['W0']), d)
  eng.set_state(d['W0'])
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  out = torch.empty(C, D, device='cuda')
  eng.transition(float(d['eps']), int(d['L']), momenta=d['r0'], uniforms=d['u'],
                 accepted_out=acc, sample_out=out)
  np.testing.assert_array_equal(acc.cpu().numpy().astype(bool), d['accept'])
  assert _rel(out.cpu().numpy().astype(np.float64), d['z']).max() < 1e-4


@pytest.mark.gpu
def test_device_hier_golden():
  import torch
  from edward_amd.hier import HierEngine
  d = _load('hier_small.npz')
  C, dim = d['th0'].shape
  eng = HierEngine(d['X'], d['y'], d['offsets'])
  _check_lp_g(*eng.logp_grad(d['th0']), d)
  eng.set_state(d['th0'])
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  out = torch.empty(C, dim, device='cuda')
  eng.transition(float(d['eps']), int(d['L']), momenta=d['r0'], uniforms=d['u'],
                 accepted_out=acc, sample_out=out)
  np.testing.assert_array_equal(acc.cpu().numpy().astype(bool), d['accept'])
  assert _rel(out.cpu().numpy().astype(np.float64), d['z']).max() < 1e-4
