"""GPU parity of the hierarchical logistic regression engine against the oracle.

Model: a_g ~ N(mu, exp(log_sigma)^2), mu, log_sigma ~ N(0,1), w ~ N(0, I),
y_n ~ Bernoulli(a_{g(n)} + x_n . w), latent [a (G), mu, log_sigma, w (D)].

Bars: log joint |dlogp| <= 1e-6 |logp| + 1e-3 and gradient within 1e-4 of
max|grad| (fp32 GEMMs, FP64 reductions); leapfrog end positions within 1e-4
relative per chain; accept decisions equal except where
|ratio - log u| < 1e-6 in the oracle.  Device momenta are checked through the
oracle's Philox restatement (a whole run replayed on the CPU).
"""

import numpy as np
import pytest
import torch

from oracle import hmc_oracle as O

pytestmark = pytest.mark.gpu


def _data(seed, sizes, D):
  rng = np.random.default_rng(seed)
  sizes = np.asarray(sizes)
  off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
  N = int(off[-1])
  X = rng.normal(size=(N, D)).astype(np.float32)
  G = len(sizes)
  a = rng.normal(0.3, 0.8, size=G)
  w = rng.normal(size=D) * 0.5
  group = np.repeat(np.arange(G), sizes)
  eta = a[group] + X.astype(np.float64) @ w
  y = (rng.random(N) < 1.0 / (1.0 + np.exp(-eta))).astype(np.float32)
  return X, y, off


def _engine(X, y, off):
  from edward_amd.hier import HierEngine
  return HierEngine(X, y, off)


def _rel_err(a, b):
  d = np.linalg.norm(a - b, axis=1)
  return d / np.maximum(np.linalg.norm(b, axis=1), 1e-6)


CASES = [
    # group sizes (ragged, with empty and single-row groups), D, C
    ([3, 0, 2, 4, 1], 3, 2),
    ([1] * 17 + [40, 0, 0, 65], 7, 5),
    (list(np.random.default_rng(3).poisson(30, size=97) + 1), 16, 33),
    (list(np.random.default_rng(4).poisson(50, size=300)), 64, 130),
]


@pytest.mark.parametrize('sizes,D,C', CASES)
def test_hier_logp_grad(sizes, D, C):
  X, y, off = _data(1, sizes, D)
  eng = _engine(X, y, off)
  m = O.HierLogReg(X.astype(np.float64), y, off)
  rng = np.random.default_rng(2)
  th = (rng.normal(size=(C, m.dim)) * 0.5).astype(np.float32)
  lp, g = eng.logp_grad(th)
  lp_ref, g_ref = m.logp_grad(th.astype(np.float64))
  lp = lp.cpu().numpy()
  g = g.cpu().numpy().astype(np.float64)
  assert np.all(np.abs(lp - lp_ref) <= 1e-6 * np.abs(lp_ref) + 1e-3), (lp, lp_ref)
  tol = 1e-4 * np.abs(g_ref).max()
  assert np.abs(g - g_ref).max() <= tol, (np.abs(g - g_ref).max(), tol)


def test_hier_gradient_directional_derivative():
  """The device gradient against central differences of the device log joint
  (FP64 reductions make a step of 1e-2 along a unit direction resolvable)."""
  X, y, off = _data(5, [20, 1, 0, 35, 12, 7], 9)
  eng = _engine(X, y, off)
  dim = len(off) - 1 + 2 + 9
  rng = np.random.default_rng(6)
  th = (rng.normal(size=(4, dim)) * 0.4).astype(np.float64)
  v = rng.normal(size=(4, dim))
  v /= np.linalg.norm(v, axis=1, keepdims=True)
  h = 1e-2
  _, g = eng.logp_grad(th.astype(np.float32))
  lp_p, _ = eng.logp_grad((th + h * v).astype(np.float32))
  lp_m, _ = eng.logp_grad((th - h * v).astype(np.float32))
  fd = (lp_p.cpu().numpy() - lp_m.cpu().numpy()) / (2 * h)
  dd = (g.cpu().numpy().astype(np.float64) * v).sum(axis=1)
  np.testing.assert_allclose(dd, fd, rtol=2e-3, atol=2e-2)


@pytest.mark.parametrize('eps,L', [(0.01, 1), (0.02, 5), (0.005, 16)])
def test_hier_transition_positions(eps, L):
  X, y, off = _data(7, list(np.random.default_rng(8).poisson(20, size=40)), 12)
  eng = _engine(X, y, off)
  m = O.HierLogReg(X.astype(np.float64), y, off)
  C = 9
  rng = np.random.default_rng(9)
  th0 = (rng.normal(size=(C, m.dim)) * 0.3).astype(np.float32)
  r0 = rng.normal(size=(C, m.dim)).astype(np.float32)
  eng.set_state(th0)
  u = np.zeros(C)
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  out_t = torch.empty(C, m.dim, device='cuda')
  eng.transition(eps, L, momenta=r0, uniforms=u, accepted_out=acc, sample_out=out_t)
  ref = O.hmc_transition(th0.astype(np.float64), r0.astype(np.float64), u, eps, L, m.logp_grad)
  assert acc.cpu().numpy().all()
  assert _rel_err(out_t.cpu().numpy().astype(np.float64), ref['zL']).max() < 1e-4


def test_hier_accept_decisions():
  X, y, off = _data(11, list(np.random.default_rng(12).poisson(25, size=60) + 1), 8)
  eng = _engine(X, y, off)
  m = O.HierLogReg(X.astype(np.float64), y, off)
  C = 200
  rng = np.random.default_rng(13)
  # equilibrate on the CPU first, so that a moderate step gives mixed ratios
  th = rng.normal(size=(C, m.dim)) * 0.2
  for _ in range(40):
    th = O.hmc_transition(th, rng.normal(size=th.shape), rng.random(C), 0.03, 8,
                          m.logp_grad)['z']
  th0 = th.astype(np.float32)
  r0 = rng.normal(size=(C, m.dim)).astype(np.float32)
  eps, L = 0.08, 6
  ref = O.hmc_transition(th0.astype(np.float64), r0.astype(np.float64), np.ones(C), eps, L,
                         m.logp_grad)
  u = np.clip(np.exp(np.minimum(ref['ratio'], 0.0)) * rng.uniform(0.5, 1.5, size=C),
              1e-12, 1 - 1e-12)
  ref = O.hmc_transition(th0.astype(np.float64), r0.astype(np.float64), u, eps, L, m.logp_grad)
  eng.set_state(th0)
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  eng.transition(eps, L, momenta=r0, uniforms=u, accepted_out=acc)
  acc = acc.cpu().numpy().astype(bool)
  decided = np.abs(ref['ratio'] - np.log(u)) > 1e-6
  assert 0.2 < ref['accept'].mean() < 0.9
  np.testing.assert_array_equal(acc[decided], ref['accept'][decided])


def test_hier_run_replays_device_rng():
  """Several transitions with device momenta and uniforms, replayed on the CPU
  with the oracle's Philox draws: identical accept sequence, positions within
  1e-4 relative."""
  X, y, off = _data(21, [15, 3, 0, 30, 9, 1, 22, 40], 6)
  eng = _engine(X, y, off)
  m = O.HierLogReg(X.astype(np.float64), y, off)
  C, eps, L, seed = 6, 0.04, 4, 1234
  rng = np.random.default_rng(22)
  th = (rng.normal(size=(C, m.dim)) * 0.3).astype(np.float32)
  eng.set_state(th)
  z = th.astype(np.float64)
  out_t = torch.empty(C, m.dim, device='cuda')
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  for t in range(6):
    eng.transition(eps, L, seed=seed, counter=t, sample_out=out_t, accepted_out=acc)
    r0 = O.momentum_normals(seed, t, C, m.dim)
    u = O.accept_uniforms(seed, t, C)
    ref = O.hmc_transition(z, r0, u, eps, L, m.logp_grad)
    decided = np.abs(ref['ratio'] - ref['log_u']) > 1e-6
    assert decided.all(), 'undecided accept in the replay; pick another seed'
    np.testing.assert_array_equal(acc.cpu().numpy().astype(bool), ref['accept'])
    zd = out_t.cpu().numpy().astype(np.float64)
    assert _rel_err(zd, ref['z']).max() < 1e-4
    z = zd   # continue from the device state (no drift of rounding across steps)
  assert eng.divergences().cpu().numpy().sum() == 0


def test_hier_rejects_bad_offsets():
  from edward_amd._lib import EdhmcError
  from edward_amd.hier import HierEngine
  X = np.zeros((5, 2), np.float32)
  y = np.zeros(5, np.float32)
  with pytest.raises(EdhmcError):
    HierEngine(X, y, np.array([0, 3, 2, 5]))
  with pytest.raises(EdhmcError):
    HierEngine(X, y, np.array([0, 3, 4]))


def test_hier_edward_surface():
  """ed.HMC over the gather model compiles to the hierarchical engine; the
  rows need not be sorted by group (the engine reorders them)."""
  import edward_amd as ed
  from edward_amd.models import Bernoulli, Empirical, Normal
  rng = np.random.default_rng(31)
  G, D, N, C, T = 5, 3, 200, 4, 30
  group_np = rng.integers(0, G, size=N)
  X_np = rng.normal(size=(N, D)).astype(np.float32)
  y_np = (rng.random(N) < 0.5).astype(np.float32)
  X = ed.placeholder('float32', [N, D])
  group = ed.placeholder('int32', [N])
  mu = Normal(loc=0.0, scale=1.0)
  log_sigma = Normal(loc=0.0, scale=1.0)
  a = Normal(loc=mu, scale=ed.exp(log_sigma), sample_shape=G)
  w = Normal(loc=np.zeros(D, np.float32), scale=np.ones(D, np.float32))
  yv = Bernoulli(logits=ed.gather(a, group) + ed.dot(X, w))
  qa = Empirical(np.zeros((T, C, G), np.float32))
  qmu = Empirical(np.zeros((T, C), np.float32))
  qls = Empirical(np.zeros((T, C), np.float32))
  qw = Empirical(np.zeros((T, C, D), np.float32))
  inf = ed.HMC({a: qa, mu: qmu, log_sigma: qls, w: qw}, data={X: X_np, group: group_np, yv: y_np})
  inf.initialize(step_size=0.05, n_steps=5, seed=3)
  for _ in range(T):
    inf.update()
  from edward_amd.hier import HierEngine
  assert isinstance(inf.engine, HierEngine)
  assert qa.params.shape == (T, C, G)
  vals = torch.cat([qa.params[-1], qmu.params[-1][:, None], qls.params[-1][:, None],
                    qw.params[-1]], dim=1).cpu().numpy()
  assert np.isfinite(vals).all() and np.abs(vals).max() > 0
  # the log joint of the store's last row on the engine equals the oracle's on
  # the unsorted data (the likelihood is a sum over rows)
  order = np.argsort(group_np, kind='stable')
  off = np.concatenate([[0], np.cumsum(np.bincount(group_np, minlength=G))])
  m = O.HierLogReg(X_np[order].astype(np.float64), y_np[order], off)
  lp, _ = inf.engine.logp_grad(vals)
  lp_ref, _ = m.logp_grad(vals.astype(np.float64))
  np.testing.assert_allclose(lp.cpu().numpy(), lp_ref, rtol=1e-6, atol=1e-3)


def test_hier_many_groups_partial_last_tile():
  """1600 groups of 16-24 rows: every row tile spans few groups (the
  epilogue stages their intercepts in LDS) and the partial last tile starts
  at a group index near 1600, where the padded rows must not index the
  staged tile with gidx - g_lo < 0."""
  sizes = list(np.random.default_rng(41).integers(16, 25, size=1600))
  X, y, off = _data(42, sizes, 8)
  eng = _engine(X, y, off)
  m = O.HierLogReg(X.astype(np.float64), y, off)
  th = (np.random.default_rng(43).normal(size=(5, m.dim)) * 0.3).astype(np.float32)
  lp, g = eng.logp_grad(th)
  lp_ref, g_ref = m.logp_grad(th.astype(np.float64))
  assert np.all(np.abs(lp.cpu().numpy() - lp_ref) <= 1e-6 * np.abs(lp_ref) + 1e-3)
  assert np.abs(g.cpu().numpy() - g_ref).max() <= 1e-4 * np.abs(g_ref).max()


def test_hier_full_size_logp_grad_and_transition():
  """The hierarchical workload at its full size (workloads.hier_groups: G = 2048
  groups of 1 + Poisson(400) rows, N ~ 821k, D = 64, 512 chains of 2114
  coordinates, eps = 1e-3, L = 16): log joint and gradient of all chains on
  the GPU, 4 of them against the oracle (bars of this file); one L = 16
  transition with fed momenta: end positions of those chains within 1e-4
  relative, the Metropolis log ratio (edhmc_last_ratio) against the
  oracle's (measured error printed, bar 1e-2 nats: fp32 dynamics over 2114
  coordinates move the kinetic energy by ~|r|^2 x 1e-6), and decisions with
  uniforms placed at ratio +- m (m = 4 x the measured error, at least 1e-5
  nats) equal to the oracle's."""
  from edward_amd import workloads as W
  wl = W.hier_groups()
  print('\ncase: hier_groups G=%d N=%d D=%d C=%d' % (wl.G, wl.N, wl.D, wl.C), flush=True)
  eng = _engine(wl.X, wl.y, wl.offsets)
  m = O.HierLogReg(wl.X.astype(np.float64), wl.y, wl.offsets)
  C, dim, eps, L = wl.C, wl.dim, wl.step_size, wl.n_steps
  eps = float(np.float32(eps))
  sel = np.array([0, 170, 341, 511])
  th0 = wl.theta0
  lp, g = eng.logp_grad(th0)
  lp_ref, g_ref = m.logp_grad(th0[sel].astype(np.float64))
  lps = lp.cpu().numpy()[sel]
  gs = g.cpu().numpy()[sel].astype(np.float64)
  print('measured: log joint err max %.3e nats (|logp| %.3e), gradient err max %.3e of '
        'max|grad| %.3e' % (np.abs(lps - lp_ref).max(), np.abs(lp_ref).max(),
                            np.abs(gs - g_ref).max(), np.abs(g_ref).max()), flush=True)
  assert np.all(np.abs(lps - lp_ref) <= 1e-6 * np.abs(lp_ref) + 1e-3), (lps, lp_ref)
  assert np.abs(gs - g_ref).max() <= 1e-4 * np.abs(g_ref).max()

  rng = np.random.default_rng(51)
  r0 = rng.normal(size=(C, dim)).astype(np.float32)
  eng.set_state(th0)
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  out_t = torch.empty(C, dim, device='cuda')
  eng.transition(eps, L, momenta=r0, uniforms=np.zeros(C), accepted_out=acc, sample_out=out_t)
  ratio = eng.last_ratio().cpu().numpy()
  assert acc.cpu().numpy().all()
  ref = O.hmc_transition(th0[sel].astype(np.float64), r0[sel].astype(np.float64),
                         np.zeros(len(sel)), eps, L, m.logp_grad)
  err_z = _rel_err(out_t.cpu().numpy()[sel].astype(np.float64), ref['zL'])
  e_rat = ratio[sel] - ref['ratio']
  print('measured: position rel err max %.2e; ratio err max %.3e nats (ratios %s)'
        % (err_z.max(), np.abs(e_rat).max(), np.round(ref['ratio'], 4)), flush=True)
  assert err_z.max() < 1e-4
  assert np.abs(e_rat).max() < 1e-2

  # decisions: even chains of sel rejected (where a uniform below 1 can do
  # it), odd ones accepted
  mgn = max(4.0 * np.abs(e_rat).max(), 1e-5)
  rat = ref['ratio']
  want = np.arange(len(sel)) % 2 == 1
  logu = np.where(want, np.minimum(rat, 0.0) - mgn, rat + mgn)
  want = want | (logu >= 0.0)
  logu = np.where(logu >= 0.0, np.minimum(rat, 0.0) - mgn, logu)
  u = rng.uniform(0.05, 0.95, size=C)
  u[sel] = np.exp(logu)
  eng.set_state(th0)
  acc.zero_()
  eng.transition(eps, L, momenta=r0, uniforms=u, accepted_out=acc)
  np.testing.assert_array_equal(eng.last_ratio().cpu().numpy(), ratio)
  np.testing.assert_array_equal(acc.cpu().numpy()[sel].astype(bool), want)
