"""Pins the FP64 MLP oracle: backward vs finite differences, GELU' vs a finite
difference of GELU, Adam's first steps written out."""

import numpy as np
import pytest

from oracle import mlp_oracle as MO


def _setup(seed=0, D=5, H=7, B=6):
  rng = np.random.default_rng(seed)
  p = {'W1': rng.normal(size=(H, D)) * 0.5, 'b1': rng.normal(size=H) * 0.1,
       'W2': rng.normal(size=(H, H)) * 0.4, 'b2': rng.normal(size=H) * 0.1,
       'W3': rng.normal(size=(D, H + D)) * 0.4, 'b3': rng.normal(size=D) * 0.1}
  norm = {'in_shift': rng.normal(size=D), 'in_scale': rng.uniform(0.5, 2, size=D),
          'out_shift': rng.normal(size=D), 'out_scale': rng.uniform(0.5, 2, size=D)}
  W = rng.normal(size=(B, D))
  G = rng.normal(size=(B, D))
  return p, norm, W, G


def test_gelu_grad_matches_finite_difference():
  x = np.linspace(-6, 6, 101)
  h = 1e-6
  fd = (MO.gelu(x + h) - MO.gelu(x - h)) / (2 * h)
  np.testing.assert_allclose(MO.gelu_grad(x), fd, atol=1e-8)
  # GELU(0) = 0 (padding contributes nothing), GELU ~ x for large x
  assert MO.gelu(0.0) == 0.0
  assert abs(MO.gelu(8.0) - 8.0) < 1e-9


def test_backward_matches_finite_differences():
  p, norm, W, G = _setup()
  loss, g = MO.loss_and_grad(p, W, G, norm)
  flat = MO.flatten(p)
  gflat = MO.flatten(g)
  h = 1e-6
  rng = np.random.default_rng(1)
  for i in rng.choice(flat.size, 40, replace=False):
    e = np.zeros_like(flat)
    e[i] = h
    lp = MO.loss_and_grad(MO.unflatten(flat + e, 5, 7), W, G, norm)[0]
    lm = MO.loss_and_grad(MO.unflatten(flat - e, 5, 7), W, G, norm)[0]
    assert gflat[i] == pytest.approx((lp - lm) / (2 * h), rel=1e-6, abs=1e-9)


def test_planted_missing_gelu_derivative_is_caught():
  """Dropping GELU' in the backward (a plausible slip) fails the FD check."""
  p, norm, W, G = _setup(2)
  f = MO.forward(p, W, norm)
  t = (G - norm['out_shift']) / norm['out_scale']
  dO = 2 * (f['o'] - t) / t.size
  wrong_W2 = (dO @ p['W3'][:, :7]).T @ f['h1']   # missing * gelu_grad(a2)
  _, g = MO.loss_and_grad(p, W, G, norm)
  assert np.abs(wrong_W2 - g['W2']).max() > 1e-3


def test_predict_inverts_the_output_normalisation():
  p, norm, W, G = _setup(3)
  o = MO.forward(p, W, norm)['o']
  np.testing.assert_allclose(MO.predict(p, W, norm), o * norm['out_scale'] + norm['out_shift'])


def test_adam_first_steps():
  p = np.array([1.0, -2.0])
  g = np.array([0.5, -0.1])
  m = np.zeros(2)
  v = np.zeros(2)
  p1, m, v = MO.adam(p, g, m, v, 1, lr=0.1)
  # step 1: m_hat = g, v_hat = g^2 -> p -= lr * sign(g) (up to eps)
  np.testing.assert_allclose(p1, p - 0.1 * np.sign(g), atol=1e-6)
  p2, m2, v2 = MO.adam(p1, g, m, v, 2, lr=0.1)
  np.testing.assert_allclose(p2, p1 - 0.1 * np.sign(g), atol=1e-6)


def test_bf16_round_ties_to_even_and_bound():
  one = 1.0
  # 1 + 2^-8 lies halfway between 1 and 1 + 2^-7: ties to the even mantissa (1)
  assert MO.bf16_round(one + 2.0 ** -8) == one
  # 1 + 3 * 2^-8 lies halfway between 1 + 2^-7 (odd) and 1 + 2^-6 (even)
  assert MO.bf16_round(one + 3 * 2.0 ** -8) == one + 2.0 ** -6
  # just above a tie rounds up, just below rounds down
  assert MO.bf16_round(one + 2.0 ** -8 + 2.0 ** -20) == one + 2.0 ** -7
  assert MO.bf16_round(one + 2.0 ** -8 - 2.0 ** -20) == one
  assert MO.bf16_round(-(one + 3 * 2.0 ** -8)) == -(one + 2.0 ** -6)
  x = np.random.default_rng(3).normal(size=10000) * 10.0 ** np.random.default_rng(4).uniform(
      -5, 5, size=10000)
  r = MO.bf16_round(x)
  bits = r.astype(np.float32).view(np.uint32)
  assert np.all(bits & 0xFFFF == 0)                       # a bf16 number
  # unit roundoff of an 8-bit significand: 2^-8
  assert np.all(np.abs(r - x.astype(np.float32)) <= 2.0 ** -8 * np.abs(x) * (1 + 1e-6))
  np.testing.assert_array_equal(MO.bf16_round(r), r)      # idempotent


def test_predict_bf16_is_predict_up_to_the_roundings():
  p, norm, W, _ = _setup(5, D=6, H=9, B=40)
  # general case: within the bf16 bound of the FP64 predictor
  a, b = MO.predict_bf16(p, W, norm), MO.predict(p, W, norm)
  assert np.linalg.norm(a - b) / np.linalg.norm(b) < 2e-2
  assert np.linalg.norm(a - b) / np.linalg.norm(b) > 1e-6   # the roundings do act
  # every rounded value already a bf16 number: zero hidden layers (GELU(0) =
  # 0), bf16 weights, inputs and normalisation giving bf16 inputs; only the
  # fp32 output arithmetic remains
  q = dict(p)
  q['W1'] = np.zeros_like(p['W1'])
  q['b1'] = np.zeros_like(p['b1'])
  q['W2'] = np.zeros_like(p['W2'])
  q['b2'] = np.zeros_like(p['b2'])
  q['W3'] = MO.bf16_round(p['W3'])
  nz = dict(norm)
  nz['in_shift'] = np.zeros(6)
  nz['in_scale'] = np.ones(6)
  Wb = MO.bf16_round(W)
  np.testing.assert_allclose(MO.predict_bf16(q, Wb, nz), MO.predict(q, Wb, nz), rtol=1e-6,
                             atol=1e-6)
