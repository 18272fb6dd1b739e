"""GPU parity of the GLM engine (libedhmc.so via its C ABI) against the oracle.

Bars (stated per test):
  * fp32 log joint: |dlogp| <= 1e-6 |logp| + 1e-3; gradient within 1e-4 of
    max|grad|.  bf16 (X rounded to bf16 for both sides, hi/lo split
    positions): log joint rtol 1e-5, gradient within 1e-2 of max|grad| (bf16
    residuals).
  * Transition with fed momenta/uniforms: leapfrog end positions within 1e-4
    relative (||dz|| / max(||z||, 1e-6) per chain); accept decisions equal
    except where |ratio - log u| < 1e-6 in the oracle.
  * Device RNG: Philox-derived uniforms bit-exact, normals within 2e-6.
"""

import math

import numpy as np
import pytest
import torch

from oracle import hmc_oracle as O

pytestmark = pytest.mark.gpu


def _engine(wl, precision='fp32', likelihood='bernoulli', lik_scale=1.0):
  from edward_amd.engine import GLMEngine
  return GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale, likelihood=likelihood,
                   lik_scale=lik_scale, precision=precision)


def _oracle(wl, likelihood='bernoulli', lik_scale=1.0):
  return O.GLM(wl.X.astype(np.float64), wl.y, wl.prior_loc, wl.prior_scale, likelihood,
               lik_scale)


def _check_logp_grad(lp, g, lp_ref, g_ref, bf16=False):
  lp = lp.cpu().numpy()
  g = g.cpu().numpy().astype(np.float64)
  if bf16:
    np.testing.assert_allclose(lp, lp_ref, rtol=1e-5)
    tol = 1e-2 * np.abs(g_ref).max()
  else:
    assert np.all(np.abs(lp - lp_ref) <= 1e-6 * np.abs(lp_ref) + 1e-3), (lp, lp_ref)
    tol = 1e-4 * np.abs(g_ref).max()
  assert np.abs(g - g_ref).max() <= tol, (np.abs(g - g_ref).max(), tol)


def _small(seed=0, N=1000, D=10, C=3, likelihood='bernoulli'):
  from edward_amd import workloads as W
  wl = W.logreg(N, D, C, 0.01, 10, 10, seed=seed, scaled_columns=False, w_sd=1.0)
  if likelihood == 'normal':
    rng = np.random.default_rng(seed + 1)
    wl.y = (wl.X.astype(np.float64) @ rng.normal(size=D) + 0.3 * rng.normal(size=N)
            ).astype(np.float32)
  return wl


@pytest.mark.parametrize('N,D,C', [(1000, 10, 1), (1000, 10, 3), (3001, 45, 37),
                                   (257, 130, 129), (5000, 256, 200)])
def test_logp_grad_fp32(N, D, C):
  wl = _small(1, N, D, C)
  eng = _engine(wl)
  W0 = np.random.default_rng(2).normal(size=(C, D)).astype(np.float32) * 0.3
  lp, g = eng.logp_grad(W0)
  lp_ref, g_ref = _oracle(wl).logp_grad(W0.astype(np.float64))
  _check_logp_grad(lp, g, lp_ref, g_ref)


def test_logp_grad_normal_likelihood():
  wl = _small(3, 500, 12, 5, likelihood='normal')
  eng = _engine(wl, likelihood='normal', lik_scale=0.7)
  W0 = np.random.default_rng(4).normal(size=(5, 12)).astype(np.float32)
  lp, g = eng.logp_grad(W0)
  lp_ref, g_ref = _oracle(wl, 'normal', 0.7).logp_grad(W0.astype(np.float64))
  _check_logp_grad(lp, g, lp_ref, g_ref)


@pytest.mark.parametrize('N,D,C', [(1000, 10, 2), (4000, 96, 130)])
def test_logp_grad_bf16(N, D, C):
  from edward_amd import workloads as W
  wl = W.logreg(N, D, C, 0.01, 10, 10, seed=5, bf16=True)
  eng = _engine(wl, precision='bf16')
  W0 = np.random.default_rng(6).normal(size=(C, D)).astype(np.float32) * 0.2
  lp, g = eng.logp_grad(W0)
  lp_ref, g_ref = _oracle(wl).logp_grad(W0.astype(np.float64))
  _check_logp_grad(lp, g, lp_ref, g_ref, bf16=True)


@pytest.mark.parametrize('precision', ['fp32', 'bf16'])
def test_forward_only_logp(precision):
  """edhmc_logp (no gradient, the Metropolis test's forward) against the
  oracle (fp32 rtol 1e-6, bf16 rtol 1e-5 with X rounded to bf16 on both
  sides) and against the log joint of edhmc_logp_grad."""
  from edward_amd import workloads as W
  wl = W.logreg(3000, 70, 150, 0.01, 10, 10, seed=15, bf16=(precision == 'bf16'))
  eng = _engine(wl, precision=precision)
  W0 = (np.random.default_rng(16).normal(size=(150, 70)) * 0.2).astype(np.float32)
  lp = eng.logp(W0).cpu().numpy()
  lp_ref, _ = _oracle(wl).logp_grad(W0.astype(np.float64))
  np.testing.assert_allclose(lp, lp_ref, rtol=1e-5 if precision == 'bf16' else 1e-6)
  # the gradient pass sums y eta - softplus(eta) per element; the forward pass
  # sums -softplus(eta) and adds w . X^T y in FP64: both within the bar above
  lp2, _ = eng.logp_grad(W0)
  np.testing.assert_allclose(lp, lp2.cpu().numpy(), rtol=2e-6)


def test_nonfinite_data_and_positions_raise():
  from edward_amd._lib import NonFiniteError
  wl = _small(7, 300, 5, 1)
  X = wl.X.copy()
  X[17, 3] = np.inf
  from edward_amd.engine import GLMEngine
  with pytest.raises(NonFiniteError):
    GLMEngine(X, wl.y)
  eng = _engine(wl)
  W0 = np.zeros((2, 5), np.float32)
  W0[1, 2] = np.nan
  with pytest.raises(NonFiniteError):
    eng.set_state(W0)


def test_device_rng_matches_oracle_philox():
  wl = _small(8, 100, 250, 1)
  eng = _engine(wl)
  r = eng.draw_momenta(64, seed=7, counter=3).cpu().numpy()
  np.testing.assert_allclose(r, O.momentum_normals(7, 3, 64, 250), atol=2e-6, rtol=2e-6)
  u = eng.draw_uniforms(1000, seed=(1 << 40) + 9, counter=(1 << 33) + 5).cpu().numpy()
  np.testing.assert_array_equal(u, O.accept_uniforms((1 << 40) + 9, (1 << 33) + 5, 1000))


def _rel_err(a, b):
  a = np.asarray(a, np.float64)
  b = np.asarray(b, np.float64)
  return np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-6)


@pytest.mark.parametrize('N,D,C,eps,L', [(1000, 10, 1, 0.01, 10), (2000, 40, 64, 0.005, 7),
                                         (3000, 130, 33, 0.002, 3)])
def test_transition_positions_with_fed_randomness(N, D, C, eps, L):
  wl = _small(9, N, D, C)
  eng = _engine(wl)
  m = _oracle(wl)
  rng = np.random.default_rng(10)
  W0 = (rng.normal(size=(C, D)) * 0.3).astype(np.float32)
  r0 = rng.normal(size=(C, D)).astype(np.float32)
  eng.set_state(W0)
  u = np.zeros(C)   # log 0 = -inf: every finite proposal is accepted
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  eng.transition(eps, L, momenta=r0, uniforms=u, accepted_out=acc)
  z, lp, g = eng.get_state()
  out = O.hmc_transition(W0.astype(np.float64), r0.astype(np.float64), u, eps, L, m.logp_grad)
  assert acc.cpu().numpy().all()
  assert _rel_err(z.cpu().numpy(), out['zL']).max() < 1e-4
  np.testing.assert_allclose(lp.cpu().numpy(), out['logpL'], rtol=1e-6, atol=1e-3)


@pytest.mark.parametrize('precision', ['fp32', 'bf16'])
def test_accept_decisions_with_fed_randomness(precision):
  from edward_amd import workloads as W
  C, D = 256, 20
  wl = W.logreg(2000, D, C, 0.01, 10, 10, seed=11, bf16=(precision == 'bf16'))
  eng = _engine(wl, precision=precision)
  m = _oracle(wl)
  rng = np.random.default_rng(12)
  W0 = (rng.normal(size=(C, D)) * 0.05 + wl.w_true).astype(np.float32)
  r0 = rng.normal(size=(C, D)).astype(np.float32)
  out = O.hmc_transition(W0.astype(np.float64), r0.astype(np.float64), np.ones(C), 0.02, 8,
                         m.logp_grad)
  # uniforms placed so that accept probabilities are mixed
  u = np.exp(np.minimum(out['ratio'], 0.0)) * rng.uniform(0.5, 1.5, size=C)
  u = np.clip(u, 1e-12, 1 - 1e-12)
  out = O.hmc_transition(W0.astype(np.float64), r0.astype(np.float64), u, 0.02, 8, m.logp_grad)
  eng.set_state(W0)
  acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  eng.transition(0.02, 8, momenta=r0, uniforms=u, accepted_out=acc)
  acc = acc.cpu().numpy().astype(bool)
  margin = np.abs(out['ratio'] - np.log(u))
  # fp32 follows the oracle's dynamics (bar 1e-6); bf16 integrates with bf16
  # residuals and hi-only inner gradients, so its ratio differs from the FP64
  # oracle's by up to ~1e-2 nats: its decisions are compared outside 0.1.
  decided = margin > (1e-6 if precision == 'fp32' else 1e-1)
  assert decided.sum() > C // 2
  assert 0 < out['accept'].sum() < C
  np.testing.assert_array_equal(acc[decided], out['accept'][decided])
  z, _, _ = eng.get_state()
  z = z.cpu().numpy()
  for c in np.where(decided)[0]:
    want = out['zL'][c] if out['accept'][c] else W0[c]
    tol = 1e-4 if precision == 'fp32' else 5e-3
    assert _rel_err(z[c], want) < tol


def test_replay_small_workload_with_device_rng():
  """N=1000, D=10, 1 chain from 0, eps=0.01, L=10: 300 transitions with the
  device's Philox draws, replayed by the oracle from the same draws, until
  the first near-threshold accept (none expected)."""
  from edward_amd import workloads as W
  wl = W.logreg_small()
  eng = _engine(wl)
  m = _oracle(wl)
  T, seed = 300, 1234
  params = torch.zeros(T, 1, wl.D, device='cuda')
  n_acc = torch.zeros(1, dtype=torch.int32, device='cuda')
  eng.set_state(wl.W0)
  eng.run(params, wl.step_size, wl.n_steps, seed=seed, counter0=0, n_accept=n_acc)
  p = params.cpu().numpy().astype(np.float64)
  z = wl.W0.astype(np.float64)
  n_ref = 0
  for t in range(T):
    r0 = O.momentum_normals(seed, t, 1, wl.D)
    u = O.accept_uniforms(seed, t, 1)
    out = O.hmc_transition(z, r0, u, wl.step_size, wl.n_steps, m.logp_grad)
    if abs(out['ratio'][0] - out['log_u'][0]) < 1e-6:
      break
    n_ref += int(out['accept'][0])
    z = out['z']
    assert _rel_err(p[t], z)[0] < 1e-4, t
  else:
    assert int(n_acc.item()) == n_ref


def test_run_semantics_and_moments():
  wl = _small(13, 500, 6, 4)
  eng = _engine(wl)
  T = 25
  params = torch.full((T, 4, 6), 7.0, device='cuda')
  eng.set_state(np.zeros((4, 6), np.float32))
  eng.moments_enable(True)
  n_acc = torch.zeros(4, dtype=torch.int32, device='cuda')
  eng.run(params, 0.05, 5, seed=3, counter0=100, n_accept=n_acc)
  p = params.cpu().numpy().astype(np.float64)
  assert not np.any(p == 7.0)        # every row written, row 0 included
  z, _, _ = eng.get_state()
  np.testing.assert_array_equal(z.cpu().numpy(), p[-1])
  mean, m2, n = eng.moments()
  assert n == T
  np.testing.assert_allclose(mean.cpu().numpy(), p.mean(0), rtol=1e-12, atol=1e-12)
  np.testing.assert_allclose(m2.cpu().numpy() / T, p.var(0), rtol=1e-9, atol=1e-14)
  # rows are distinct transitions: a continuation from params[T-1] with the
  # next counters reproduces a single longer run
  params2 = torch.zeros(2 * T, 4, 6, device='cuda')
  eng.set_state(np.zeros((4, 6), np.float32))
  eng.run(params2, 0.05, 5, seed=3, counter0=100)
  np.testing.assert_array_equal(params2[:T].cpu().numpy(), p)
  assert n_acc.sum().item() > 0


@pytest.mark.parametrize('N,D,C', [(65536, 512, 512), (100000, 256, 256), (65536, 768, 256)])
def test_bf16_gradient_on_the_direct_load_gemm2(N, D, C):
  """bf16 gradients at shapes whose X^T R product runs on the 256 x 256
  direct-to-LDS kernel (whole tiles, every K split >= 5 stages): 64 splits
  of 32 stages (a whole number of five-stage groups plus a partial one),
  64 splits of 49 stages with a last split of 41 (stages past a split's end
  read as zeros), and three column tiles at D = 768 (GEMM1 there on the
  LDS-staged kernel), against the oracle's gradient with the pass's own
  roundings (GLM.grad_bf16, two position planes, bf16 residuals, FP64
  sums): max error within 1e-4 of max|grad| (measured 2.7e-5, 1.4e-5, 3.4e-5:
  fp32 sums over up to 100,000 rows, and residuals whose fp32 eta rounds to
  the other side of a bf16 tie), two evaluations bitwise equal."""
  from edward_amd import workloads as W
  print('case N=%d D=%d C=%d' % (N, D, C), flush=True)
  wl = W.logreg(N, D, C, 0.002, 8, 10, seed=41, bf16=True)
  eng = _engine(wl, precision='bf16')
  rng = np.random.default_rng(9)
  W0 = (wl.w_true[None, :] + 0.01 * rng.normal(size=(C, D))).astype(np.float32)
  _, g1 = eng.logp_grad(W0)
  _, g2 = eng.logp_grad(W0)
  g1, g2 = g1.cpu().numpy(), g2.cpu().numpy()
  np.testing.assert_array_equal(g1, g2)
  g_ref = _oracle(wl).grad_bf16(W0.astype(np.float64), planes=2)
  err = np.abs(g1 - g_ref).max() / np.abs(g_ref).max()
  print('measured: max gradient error %.2e of max|grad|' % err, flush=True)
  assert err < 1e-4, err


def test_mid_workload_full_size_logp_grad():
  """N=1e5, D=256 (column-scaled X), all 1024 chains on the GPU; 12 of them
  checked against the FP64 oracle."""
  from edward_amd import workloads as W
  wl = W.logreg_mid()
  eng = _engine(wl)
  lp, g = eng.logp_grad(wl.W0)
  idx = np.array([0, 1, 127, 128, 500, 511, 512, 700, 1000, 1021, 1022, 1023])
  lp_ref, g_ref = _oracle(wl).logp_grad(wl.W0[idx].astype(np.float64))
  _check_logp_grad(lp[idx], g[idx], lp_ref, g_ref)


def test_large_workload_bf16_logp_grad_and_directional_derivative():
  """N=1e6, D=512, 4096 chains, X in bf16.  3 chains against the FP64 oracle
  (on the same bf16-rounded X); the gradient of all chains checked against a
  central difference of the engine's own FP64 log joint along a random
  direction (size-independent property)."""
  from edward_amd import workloads as W
  wl = W.logreg_large()
  eng = _engine(wl, precision='bf16')
  rng = np.random.default_rng(14)
  Wp = (wl.w_true[None, :] + 0.003 * rng.normal(size=(wl.C, wl.D))).astype(np.float32)
  lp, g = eng.logp_grad(Wp)
  idx = np.array([0, 2048, 4095])
  lp_ref, g_ref = _oracle(wl).logp_grad(Wp[idx].astype(np.float64))
  lpi = lp[idx].cpu().numpy()
  assert np.all(np.abs(lpi - lp_ref) < 1e-6 * np.abs(lp_ref) + 0.05), (lpi - lp_ref)
  gi = g[idx].cpu().numpy()
  assert np.abs(gi - g_ref).max() <= 2e-2 * np.abs(g_ref).max()
  v = rng.normal(size=(wl.C, wl.D)).astype(np.float32)
  v /= np.linalg.norm(v, axis=1, keepdims=True)
  # FP32 accumulation of eta over 512 products leaves ~2e-3 nats of rounding
  # noise in the log joint at N=1e6; h is chosen so 2 h |g.v| >> that.
  h = 2e-3
  lpp, _ = eng.logp_grad(Wp + h * v)
  lpm, _ = eng.logp_grad(Wp - h * v)
  fd = (lpp - lpm).cpu().numpy() / (2 * h)
  gv = (g.cpu().numpy().astype(np.float64) * v).sum(1)
  assert np.median(np.abs(fd - gv) / np.abs(gv)) < 1e-2


@pytest.mark.parametrize('between', [True, False])
def test_device_ess_matches_oracle(between):
  """ESS from the device autocovariances against the oracle's FFT-based
  multi-chain estimator on AR(1) chains of several correlations."""
  from edward_amd import diagnostics
  wl = _small(3, 200, 4, 2)
  eng = _engine(wl)
  rng = np.random.default_rng(40)
  T, C, D = 400, 8, 6
  rho = np.array([0.0, 0.3, 0.6, 0.9, 0.95, 0.5])
  x = np.zeros((T, C, D))
  x[0] = rng.normal(size=(C, D)) / np.sqrt(1 - rho ** 2)
  for t in range(1, T):
    x[t] = rho * x[t - 1] + rng.normal(size=(C, D))
  x = x + rng.normal(size=(1, C, D)) * 0.2    # chain offsets (between-chain spread)
  xf = x.astype(np.float32)
  dev = diagnostics.ess(eng, torch.as_tensor(xf, device='cuda'), between=between)
  ref = np.array([O.ess(xf[:, :, d].astype(np.float64).T, between=between) for d in range(D)])
  np.testing.assert_allclose(dev, ref, rtol=1e-6)


@pytest.mark.parametrize('N,D,C', [(40000, 200, 512), (20000, 512, 1024), (20011, 500, 512),
                                   (20000, 384, 512), (30000, 128, 2048), (30000, 128, 4096)])
def test_forward_only_logp_chain_stationary_path(N, D, C):
  """Sizes that take the W-stationary kernel (bf16, Bernoulli, chain tiles a
  multiple of 8, D_pad <= 512; the forward takes its pipelined four-wave
  form for D_pad = 256, 384 and 512, the eight-wave one for 128; ragged last tiles,
  padded columns at D = 200 and 500; the last two split the rows into 8 and 4
  parts, one per XCD or pair of XCDs), forward-only and with the residuals
  of the gradient pass: log joint against the oracle (rtol 1e-5, X rounded
  to bf16 on both sides) and against the log joint of edhmc_logp_grad;
  gradient against the oracle's (within 1e-2 of max|grad|, the bf16 bar)."""
  from edward_amd import workloads as W
  print('case N=%d D=%d C=%d' % (N, D, C), flush=True)
  wl = W.logreg(N, D, C, 0.01, 10, 10, seed=31, bf16=True)
  eng = _engine(wl, precision='bf16')
  W0 = (np.random.default_rng(32).normal(size=(C, D)) * 0.05).astype(np.float32)
  lp = eng.logp(W0).cpu().numpy()
  sel = np.arange(0, C, max(1, C // 24))
  lp_ref, g_ref = _oracle(wl).logp_grad(W0[sel].astype(np.float64))
  np.testing.assert_allclose(lp[sel], lp_ref, rtol=1e-5)
  lpg, g = eng.logp_grad(W0)
  np.testing.assert_allclose(lp, lpg.cpu().numpy(), rtol=2e-6)
  _check_logp_grad(lpg.cpu()[sel], g.cpu()[sel], lp_ref, g_ref, bf16=True)


def _bf16_round(x):
  u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
  u = (u + 0x7fff + ((u >> 16) & 1)) >> 16 << 16
  return u.astype(np.uint32).view(np.float32).astype(np.float64)


def test_plain_bf16_transition_on_the_chain_stationary_path():
  """A plain bf16 transition at a size the W-stationary kernel takes (its
  one-plane residual pass for the inner leapfrog steps, its two-plane pass
  with the log-likelihood for the first and last gradients): end positions of
  chains spread over the chain tiles against the oracle's leapfrog with the
  inner gradients at bf16-rounded positions (the hi plane), within 1e-3
  relative per chain (bf16 residuals; measured ~1e-4)."""
  from edward_amd import workloads as W
  N, D, C, eps, L = 20000, 128, 512, 0.002, 8
  wl = W.logreg(N, D, C, eps, L, 10, seed=41, bf16=True)
  eng = _engine(wl, precision='bf16')
  m = _oracle(wl)
  rng = np.random.default_rng(3)
  W0 = (wl.w_true[None, :] + 0.01 * rng.normal(size=(C, D))).astype(np.float32)
  r0 = rng.normal(size=(C, D)).astype(np.float32)
  eng.set_state(W0)
  out = torch.empty(C, D, device='cuda')
  eng.transition(eps, L, momenta=r0, uniforms=np.zeros(C), sample_out=out)
  sel = np.array([0, 77, 130, 257, 390, 511])
  z, r = W0[sel].astype(np.float64), r0[sel].astype(np.float64)
  _, g = m.logp_grad(z)
  for k in range(L):
    r = r + 0.5 * eps * g
    z = z + eps * r
    _, g = m.logp_grad(z if k == L - 1 else _bf16_round(z))
    r = r + 0.5 * eps * g
  zd = out.cpu().numpy()[sel].astype(np.float64)
  err = np.linalg.norm(zd - z, axis=1) / np.linalg.norm(z, axis=1)
  assert err.max() < 1e-3, err


@pytest.mark.parametrize('N,D,C,path', [(20000, 128, 512, 'w-stationary'),
                                        (20000, 512, 512, 'w-stationary'),
                                        (20000, 256, 1024, 'w-stationary, 128-chain tiles'),
                                        (20000, 640, 256, 'lds-staged'),
                                        (6000, 96, 130, 'lds-staged')])
def test_plain_bf16_transition_follows_the_rounded_oracle(N, D, C, path):
  """A plain bf16 transition against the oracle's leapfrog with the gradient
  pass's own roundings (GLM.grad_bf16: hi + lo position planes at the ends,
  the hi plane for the inner steps, bf16 residuals), on the W-stationary
  kernels (bf16, D_pad <= 512, enough rows; D = 128 and 512, the latter in
  20 row parts with a ragged last tile; at 1024 chains the one-plane inner
  gradients on 128-chain tiles) and on the LDS-staged GEMMs
  (D_pad = 640 > 512; N below 16 row tiles): two runs bitwise equal; end
  positions of every chain: median within 1e-5 relative (measured 1e-7 to
  7e-7), every chain within 1e-3 (measured up to 1.6e-4 on isolated chains
  of either kernel path: where a coordinate lies within fp32 rounding of a
  bf16 tie the two sides' hi planes differ by one bf16 ulp, which
  test_oracle.py::test_bf16_inner_positions_amplify_rounding_differences
  reproduces with the oracle alone); the log ratio at the GPU's own end
  point (edhmc_last_ratio against the FP64 log joint at the end point and
  the end momenta of edhmc_get_momentum) within 5e-3 nats (measured up to
  1.2e-3 at D = 640: the hi + lo split's 2^-17 remainder of w times the
  gradient)."""
  from edward_amd import workloads as W
  print('case N=%d D=%d C=%d (%s)' % (N, D, C, path), flush=True)
  eps, L = 0.002, 8
  wl = W.logreg(N, D, C, eps, L, 10, seed=41, bf16=True)
  eng = _engine(wl, precision='bf16')
  m = _oracle(wl)
  rng = np.random.default_rng(3)
  W0 = (wl.w_true[None, :] + 0.01 * rng.normal(size=(C, D))).astype(np.float32)
  r0 = rng.normal(size=(C, D)).astype(np.float32)
  outs, ratios = [], []
  for _ in range(2):
    eng.set_state(W0)
    out = torch.empty(C, D, device='cuda')
    eng.transition(eps, L, momenta=r0, uniforms=np.zeros(C), sample_out=out)
    outs.append(out.cpu().numpy())
    ratios.append(eng.last_ratio().cpu().numpy())
  np.testing.assert_array_equal(outs[0], outs[1])
  np.testing.assert_array_equal(ratios[0], ratios[1])
  rL = eng.last_momentum().cpu().numpy().astype(np.float64)
  sel = np.arange(C)
  e32 = float(np.float32(eps))
  z, r = W0[sel].astype(np.float64), r0[sel].astype(np.float64)
  g = m.grad_bf16(z, planes=2)
  for k in range(L):
    r = r + 0.5 * e32 * g
    z = z + e32 * r
    g = m.grad_bf16(z, planes=2 if k == L - 1 else 1)
    r = r + 0.5 * e32 * g
  zd = outs[0][sel].astype(np.float64)
  err = np.linalg.norm(zd - z, axis=1) / np.linalg.norm(z, axis=1)
  rat_same = (0.5 * (r0[sel].astype(np.float64) ** 2).sum(1) - 0.5 * (rL[sel] ** 2).sum(1)
              + m.logp(zd) - m.logp(W0[sel].astype(np.float64)))
  e_rat = ratios[0][sel] - rat_same
  print('measured: position rel err max %.2e median %.2e (%d chains above 1e-5); ratio err '
        'at the end point max %.2e nats' % (err.max(), np.median(err), (err > 1e-5).sum(),
                                            np.abs(e_rat).max()), flush=True)
  assert np.median(err) < 1e-5, err
  assert err.max() < 1e-3, err
  assert np.abs(e_rat).max() < 5e-3, e_rat


@pytest.mark.parametrize('N,D,C', [(20000, 128, 512), (20000, 640, 256)])
def test_plain_bf16_transition_with_exact_inner_gradients(N, D, C):
  """edhmc_set_inner_gradient(1): every leapfrog gradient on both position
  planes (W-stationary and LDS-staged kernels): end positions of every chain
  against the oracle's leapfrog with GLM.grad_bf16(planes=2) at every step,
  median within 1e-5 and every chain within 1e-4 relative (measured median
  1.0e-6 / 1.5e-6, max 2.8e-6 / 3.1e-6: no hi-plane rounding flips remain,
  bf16 residual flips do)."""
  from edward_amd import workloads as W
  print('case N=%d D=%d C=%d' % (N, D, C), flush=True)
  eps, L = 0.002, 8
  wl = W.logreg(N, D, C, eps, L, 10, seed=43, bf16=True)
  eng = _engine(wl, precision='bf16')
  eng.set_inner_gradient(True)
  m = _oracle(wl)
  rng = np.random.default_rng(4)
  W0 = (wl.w_true[None, :] + 0.01 * rng.normal(size=(C, D))).astype(np.float32)
  r0 = rng.normal(size=(C, D)).astype(np.float32)
  eng.set_state(W0)
  out = torch.empty(C, D, device='cuda')
  eng.transition(eps, L, momenta=r0, uniforms=np.zeros(C), sample_out=out)
  e32 = float(np.float32(eps))
  z, r = W0.astype(np.float64), r0.astype(np.float64)
  g = m.grad_bf16(z, planes=2)
  for _ in range(L):
    r = r + 0.5 * e32 * g
    z = z + e32 * r
    g = m.grad_bf16(z, planes=2)
    r = r + 0.5 * e32 * g
  zd = out.cpu().numpy().astype(np.float64)
  err = np.linalg.norm(zd - z, axis=1) / np.linalg.norm(z, axis=1)
  print('measured: position rel err max %.2e median %.2e' % (err.max(), np.median(err)),
        flush=True)
  assert np.median(err) < 1e-5, err
  assert err.max() < 1e-4, err
  eng.set_inner_gradient(False)
