"""GPU checks of the gradient surrogate (libedhmc.so MLP kernels) against the
FP64 MLP oracle, and of surrogate-driven HMC against the exact posterior.

Bars: the GPU computes with bf16 weights/activations on the matrix cores and
fp32 accumulation, so loss and parameter gradients are compared to the FP64
oracle at 3e-2 relative (norm-wise), the forward pass at 2e-2; Adam on the
same gradient to 1e-5.  Surrogate HMC must reproduce the exact posterior
moments within Monte Carlo error even with an untrained network (the exact
log joint decides acceptance)."""

import math

import numpy as np
import pytest
import torch

from oracle import hmc_oracle as O
from oracle import mlp_oracle as MO

pytestmark = pytest.mark.gpu


def _engine(N=2000, D=24, seed=0, C=64):
  from edward_amd import workloads as W
  from edward_amd.engine import GLMEngine
  wl = W.logreg(N, D, C, 0.01, 10, 10, seed=seed, bf16=True, w_sd=0.5)
  eng = GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale, precision='bf16')
  return wl, eng


def _surrogate(eng, H=64, B=512, seed=0):
  from edward_amd.surrogate import Surrogate
  return Surrogate(eng, hidden=H, max_batch=B, seed=seed)


def _norm_np(rng, D):
  return {'in_shift': rng.normal(size=D) * 0.1, 'in_scale': rng.uniform(0.5, 2, size=D),
          'out_shift': rng.normal(size=D), 'out_scale': rng.uniform(0.5, 3, size=D)}


def _set_norm(sur, norm):
  sur.set_normalisation(*[norm[k].astype(np.float32) for k in
                          ('in_shift', 'in_scale', 'out_shift', 'out_scale')])


def _rel(a, b):
  return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.mark.parametrize('D,H,B', [(24, 64, 300), (40, 200, 129), (24, 1000, 300),
                                   (512, 1024, 512)])
def test_loss_grad_and_forward_match_oracle(D, H, B):
  """The last case is the headline network (512 -> 1024 -> 1024 -> 512) on a
  512-row batch."""
  print('case D=%d H=%d B=%d' % (D, H, B), flush=True)
  _, eng = _engine(D=D)
  sur = _surrogate(eng, H=H, B=512, seed=1)
  rng = np.random.default_rng(2)
  norm = _norm_np(rng, D)
  _set_norm(sur, norm)
  W = rng.normal(size=(B, D)).astype(np.float32)
  G = rng.normal(size=(B, D)).astype(np.float32) * 3
  loss, grad = sur.loss_grad(torch.tensor(W), torch.tensor(G))
  # a second evaluation gives the same bits (per-wave FP64 partials summed in
  # a fixed order), the loss alone (forward only, as for held-out batches)
  # the same value up to the rounding of its size-weighted mean
  loss2, grad2 = sur.loss_grad(torch.tensor(W), torch.tensor(G))
  assert loss2.item() == loss.item()
  np.testing.assert_array_equal(grad2.cpu().numpy(), grad.cpu().numpy())
  assert sur.validation_loss(torch.tensor(W), torch.tensor(G)) == pytest.approx(loss.item(),
                                                                                rel=1e-15)
  p = MO.unflatten(sur.get_params().cpu().numpy(), D, H)
  loss_ref, g_ref = MO.loss_and_grad(p, W.astype(np.float64), G.astype(np.float64), norm)
  assert loss.item() == pytest.approx(loss_ref, rel=3e-2)
  gg = MO.unflatten(grad.cpu().numpy(), D, H)
  for k in ('W1', 'b1', 'W2', 'b2', 'W3', 'b3'):
    assert _rel(gg[k], g_ref[k]) < 3e-2, k
  pred = sur.forward(torch.tensor(W)).cpu().numpy()
  assert _rel(pred, MO.predict(p, W.astype(np.float64), norm)) < 2e-2


@pytest.mark.parametrize('D,H,B', [(24, 64, 256), (512, 1024, 512)])
def test_adam_step_matches_oracle_on_the_same_gradient(D, H, B):
  _, eng = _engine(D=D)
  sur = _surrogate(eng, H=H, B=B, seed=3)
  rng = np.random.default_rng(4)
  _set_norm(sur, _norm_np(rng, D))
  W = torch.tensor(rng.normal(size=(B, D)).astype(np.float32), device='cuda')
  G = torch.tensor(rng.normal(size=(B, D)).astype(np.float32), device='cuda')
  p0 = sur.get_params().cpu().numpy().astype(np.float64)
  _, grad = sur.loss_grad(W, G)
  sur.train_step(W, G, lr=1e-2)
  p1 = sur.get_params().cpu().numpy()
  want, _, _ = MO.adam(p0, grad.cpu().numpy().astype(np.float64), 0.0, 0.0, 1, lr=1e-2)
  np.testing.assert_allclose(p1, want, atol=1e-5, rtol=1e-5)


def _field_problem(D=32, seed=6, amp=1.0, freq=2.0):
  rng = np.random.default_rng(seed)
  A = rng.normal(size=(D, D)) / math.sqrt(D)
  P = A @ A.T + np.eye(D)
  mu = rng.normal(size=D)
  cov = np.linalg.inv(P)

  def field(w):
    return -(w - mu) @ P + amp * np.tanh(freq * (w - mu))
  return rng, mu, cov, field


def test_adam_training_matches_oracle_step_by_step():
  """30 Adam steps on the same batches from the same start: the GPU losses
  follow the FP64 oracle's to bf16 accuracy (5%), and the weights end within
  2% (norm-wise) of the oracle's."""
  D, H, B = 32, 128, 512
  _, eng = _engine(D=D)
  sur = _surrogate(eng, H=H, B=B, seed=5)
  rng, mu, cov, field = _field_problem(D)
  Wtr = rng.multivariate_normal(mu, cov, size=B * 6)
  Gtr = field(Wtr)
  norm = {'in_shift': Wtr.mean(0), 'in_scale': 1 / Wtr.std(0), 'out_shift': Gtr.mean(0),
          'out_scale': Gtr.std(0)}
  _set_norm(sur, norm)
  flat = sur.get_params().cpu().numpy().astype(np.float64)
  m = np.zeros_like(flat)
  v = np.zeros_like(flat)
  Wt = torch.tensor(Wtr, dtype=torch.float32, device='cuda')
  Gt = torch.tensor(Gtr, dtype=torch.float32, device='cuda')
  gl, ol = [], []
  for step in range(1, 31):
    sl = slice(((step - 1) % 6) * B, ((step - 1) % 6 + 1) * B)
    gl.append(sur.train_step(Wt[sl].contiguous(), Gt[sl].contiguous(), lr=3e-3).item())
    loss, g = MO.loss_and_grad(MO.unflatten(flat, D, H), Wtr[sl], Gtr[sl], norm)
    ol.append(loss)
    flat, m, v = MO.adam(flat, MO.flatten(g), m, v, step, 3e-3)
  np.testing.assert_allclose(gl, ol, rtol=5e-2)
  assert ol[-1] < 0.5 * ol[0]
  assert _rel(sur.get_params().cpu().numpy(), flat) < 2e-2


def test_train_epoch_equals_single_steps():
  """One edhmc_sur_train_epoch call (batch rows read through a permutation on
  the device, per-step learning rates) leaves the same weights, bit for bit, as
  the same steps taken one by one on index_select batches, and the same
  losses, bit for bit (per-wave FP64 partials summed in a fixed order); the
  oracle pins the single steps (the test above)."""
  D, H, B, nb = 32, 128, 256, 5
  _, eng = _engine(D=D)
  rng, mu, cov, field = _field_problem(D, seed=9)
  Wtr = rng.multivariate_normal(mu, cov, size=B * nb + 37)
  Gtr = field(Wtr)
  norm = {'in_shift': Wtr.mean(0), 'in_scale': 1 / Wtr.std(0), 'out_shift': Gtr.mean(0),
          'out_scale': Gtr.std(0)}
  Wt = torch.tensor(Wtr, dtype=torch.float32, device='cuda')
  Gt = torch.tensor(Gtr, dtype=torch.float32, device='cuda')
  perm = torch.tensor(rng.permutation(Wtr.shape[0]), dtype=torch.int64, device='cuda')
  lrs = [3e-3 * 0.9 ** b for b in range(nb)]
  a = _surrogate(eng, H=H, B=B, seed=11)
  _set_norm(a, norm)
  la = a.train_epoch(Wt, Gt, perm, B, lrs).cpu().numpy()
  pa = a.get_params().cpu().numpy()
  b = _surrogate(eng, H=H, B=B, seed=11)
  _set_norm(b, norm)
  lb = []
  for k in range(nb):
    idx = perm[k * B:(k + 1) * B]
    lb.append(b.train_step(Wt.index_select(0, idx).contiguous(),
                           Gt.index_select(0, idx).contiguous(), lr=lrs[k]).item())
  np.testing.assert_array_equal(la, np.array(lb))
  np.testing.assert_array_equal(pa, b.get_params().cpu().numpy())
  assert a.step_count == b.step_count == nb
  # an index outside the pairs reads a NaN row: that step's loss is NaN
  bad = perm.clone()
  bad[3] = Wtr.shape[0]
  c = _surrogate(eng, H=H, B=B, seed=11)
  _set_norm(c, norm)
  lc = c.train_epoch(Wt, Gt, bad, B, lrs[:1]).cpu().numpy()
  assert np.isnan(lc[0])


def test_hidden_layers_learn_beyond_the_linear_skip():
  """Field -P (w - mu) + tanh(2 (w - mu)): the least-squares skip captures the
  linear part; training must cut the remaining error by 40% (the FP64
  oracle trained the same way reaches 0.47 of the linear-only error)."""
  D = 32
  _, eng = _engine(D=D)
  sur = _surrogate(eng, H=256, B=1024, seed=5)
  rng, mu, cov, field = _field_problem(D)
  Wtr = torch.tensor(rng.multivariate_normal(mu, cov, size=20000), dtype=torch.float32)
  Gtr = torch.tensor(field(Wtr.double().numpy()), dtype=torch.float32)
  Wte = rng.multivariate_normal(mu, cov, size=2000)
  sur.normalise_from(Wtr.cuda(), Gtr.cuda())
  sur.init_skip_least_squares(Wtr.cuda(), Gtr.cuda())
  err_linear = _rel(sur.forward(torch.tensor(Wte, dtype=torch.float32)).cpu().numpy(),
                    field(Wte))
  sur.fit(Wtr, Gtr, epochs=80, batch=1024, lr=5e-3, lr_final=1e-4)
  err = _rel(sur.forward(torch.tensor(Wte, dtype=torch.float32)).cpu().numpy(), field(Wte))
  assert err < 0.6 * err_linear, (err_linear, err)


def test_surrogate_trajectory_matches_oracle_leapfrog():
  """Fed momenta, u = 0 (accept): the end point of the surrogate-driven
  leapfrog equals the oracle's leapfrog under the oracle MLP with the same
  weights, to the bf16 precision of the network (5e-3 relative)."""
  D, C = 24, 64
  wl, eng = _engine(D=D, C=C)
  sur = _surrogate(eng, H=64, B=C, seed=7)
  rng = np.random.default_rng(8)
  norm = {'in_shift': np.zeros(D), 'in_scale': np.ones(D) * 2.0,
          'out_shift': np.zeros(D), 'out_scale': np.ones(D) * 5.0}
  _set_norm(sur, norm)
  W0 = (wl.w_true + 0.05 * rng.normal(size=(C, D))).astype(np.float32)
  r0 = rng.normal(size=(C, D)).astype(np.float32)
  eng.set_state(W0)
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  sur.transition(0.01, 6, momenta=r0, uniforms=np.zeros(C), accepted_out=acc)
  z, _, _ = eng.get_state()
  p = MO.unflatten(sur.get_params().cpu().numpy(), D, 64)

  def field(w):
    return None, MO.predict(p, w, norm)
  zL, rL, _, _ = O.leapfrog(W0.astype(np.float64), r0.astype(np.float64), 0.01, 6, field)
  err = np.linalg.norm(z.cpu().numpy() - zL, axis=1) / np.linalg.norm(zL, axis=1)
  assert acc.cpu().numpy().all()
  assert err.max() < 5e-3


def _posterior_moments(params):
  s = params.reshape(-1, params.shape[-1])
  return s.mean(0), s.var(0)


def test_untrained_surrogate_still_samples_the_exact_posterior():
  """Exactness: with a random (untrained) network as the force field, the
  Metropolis test on the exact log joint keeps the FP64 posterior (Laplace on
  N=4000 rows; its O(1/N) bias is far below the tolerance)."""
  D, C, T = 6, 512, 300
  from edward_amd import workloads as Wk
  from edward_amd.engine import GLMEngine
  wl = Wk.logreg(4000, D, C, 0.01, 10, 10, seed=9, bf16=True, w_sd=0.5)
  eng = GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale, precision='bf16')
  sur = _surrogate(eng, H=64, B=C, seed=10)
  m = O.GLM(wl.X.astype(np.float64), wl.y, wl.prior_loc, wl.prior_scale)
  w_map, cov = O.laplace_logreg(m)
  sd = np.sqrt(np.diag(cov))
  # a crude, untrained force: identity network scaled by the posterior size
  _set_norm(sur, {'in_shift': w_map, 'in_scale': 1.0 / sd, 'out_shift': np.zeros(D),
                  'out_scale': 1.0 / sd})
  rng = np.random.default_rng(11)
  eng.set_state((w_map + sd * rng.normal(size=(C, D))).astype(np.float32))
  params = torch.zeros(T, C, D, device='cuda')
  n_acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  for t in range(T):
    sur.transition(0.3 * sd.min(), 4, seed=12, counter=t, sample_out=params[t], n_accept=n_acc)
  acc_rate = n_acc.sum().item() / (C * T)
  assert 0.01 < acc_rate < 0.99
  s = params[T // 3:].cpu().numpy().astype(np.float64)
  mean, var = _posterior_moments(s)
  n_eff = s.shape[0] * C * max(acc_rate, 0.05) / 10   # conservative ESS
  assert np.all(np.abs(mean - w_map) < 5 * sd / math.sqrt(n_eff))
  np.testing.assert_allclose(var, sd ** 2, rtol=0.2)


def test_trained_surrogate_gains_over_untrained():
  """Acceptance of surrogate HMC with a network trained on recorded warm-up
  pairs versus the same network untrained, at the same eps and L."""
  D, C = 24, 256
  wl, eng = _engine(N=4000, D=D, C=C, seed=13)
  rng = np.random.default_rng(14)
  m = O.GLM(wl.X.astype(np.float64), wl.y, wl.prior_loc, wl.prior_scale)
  w_map, cov = O.laplace_logreg(m)
  sd = np.sqrt(np.diag(cov))
  eps, L = 0.5 * sd.min(), 10
  eng.set_state((w_map + sd * rng.normal(size=(C, D))).astype(np.float32))
  cap = C * (L - 1) * 8
  Wb = torch.empty(cap, D, device='cuda')
  Gb = torch.empty(cap, D, device='cuda')
  for t in range(10):
    if t == 2:
      eng.record_pairs(Wb, Gb)
    eng.transition(eps, L, seed=15, counter=t)
  eng.record_pairs(None)
  torch.cuda.synchronize()
  sur = _surrogate(eng, H=128, B=C, seed=16)
  untrained = sur.get_params().clone()

  def acceptance():
    n = torch.zeros(C, dtype=torch.int32, device='cuda')
    for t in range(20):
      sur.transition(eps, L, seed=17, counter=1000 + t, n_accept=n)
    return n.sum().item() / (20.0 * C)
  sur.normalise_from(Wb, Gb)
  a_untrained = acceptance()
  sur.set_params(untrained.cpu().numpy())
  sur.fit(Wb, Gb, epochs=40, batch=min(1024, C * 4), lr=2e-3, lr_final=1e-4)
  a_trained = acceptance()
  assert a_trained > 0.6 and a_trained > a_untrained + 0.1, (a_untrained, a_trained)


def test_surrogate_ratio_applies_the_likelihood_weight():
  """With a likelihood weight w != 1 (edhmc_set_lik_weight) both ends of the
  surrogate transition's Metropolis ratio carry w * loglik + log prior: the
  device ratio (edhmc_last_ratio) equals K(r0) - K(rL) + lw(zL) - lw(z0) of
  the FP64 oracle at the device's own end point and momenta (edhmc_get_momentum),
  within 1e-6 relative of the log likelihood."""
  D, C, w = 16, 64, 0.37
  wl, eng = _engine(N=3000, D=D, C=C, seed=21)
  sur = _surrogate(eng, H=64, B=C, seed=22)
  rng = np.random.default_rng(23)
  m = O.GLM(wl.X.astype(np.float64), wl.y, wl.prior_loc, wl.prior_scale)
  w_map, cov = O.laplace_logreg(m)
  sd = np.sqrt(np.diag(cov))
  _set_norm(sur, {'in_shift': w_map, 'in_scale': 1.0 / sd, 'out_shift': np.zeros(D),
                  'out_scale': 1.0 / sd})
  W0 = (w_map + sd * rng.normal(size=(C, D))).astype(np.float32)
  r0 = rng.normal(size=(C, D)).astype(np.float32)
  eng.set_lik_weight(w)
  eng.set_state(W0)
  sur.transition(0.3 * sd.min(), 5, momenta=r0, uniforms=np.zeros(C))
  z, _, _ = eng.get_state()
  zL = z.cpu().numpy().astype(np.float64)
  rL = eng.last_momentum().cpu().numpy().astype(np.float64)
  ratio = eng.last_ratio().cpu().numpy()

  def lw(W):
    return w * m.loglik(W) + O.normal_logp(W, m.prior_loc, m.prior_scale).sum(1)
  want = (0.5 * (r0.astype(np.float64) ** 2).sum(1) - 0.5 * (rL ** 2).sum(1)
          + lw(zL) - lw(W0.astype(np.float64)))
  np.testing.assert_allclose(ratio, want, rtol=0, atol=1e-6 * np.abs(m.loglik(zL)).max() + 1e-6)
  eng.set_lik_weight(1.0)
