"""Benchmark: min-ESS/sec of Bayesian logistic-regression HMC on one MI355X,
surrogate-driven bf16 HMC versus plain fp32 HMC.

Workload (default): N = 1e6 rows, D = 512 features (X ~ N(0,1) with column
scales U(0.5, 2), stored bf16), 4096 chains from N(0, I), eps = 5e-4,
L = 32 leapfrog steps; the surrogate MLP 512 -> 1024 -> 1024 -> 512 (+ linear
skip) trained on ~200k (w, grad log p) pairs (196,608) from the plain
sampler's own warm-up trajectories -- the large logistic-regression configuration with the
learned-surrogate recipe.  A step is one HMC transition of all chains.

Flow (one JSON line on stdout at the end; progress on stderr):
  0. untimed: --warmup plain bf16 transitions (kernel load, first touch) and
     one epoch of a throwaway surrogate on random pairs (the one-time loading
     of the training path), then the chains are reset to their N(0, I) start;
  timed region (everything the surrogate method costs):
  1. warm-up: plain bf16 HMC from N(0, I) until the mean log joint settles,
     (w, exact gradient) pairs recorded along its last trajectories
     (--init laplace --pairs laplace instead: device Newton steps for the
     Laplace approximation, chains and pairs drawn from it);
  2. surrogate training on the GPU (own fwd/bwd/Adam kernels, 5% held out);
  3. K = --steps surrogate transitions (exact bf16/FP64 log joint in every
     Metropolis test), samples kept on the device;
  `value` = mixing min-ESS of the K samples / seconds of 1-3, where the
  mixing ESS of a coordinate is its multi-chain ESS minus the floor
  M n / (2n - 1) that chains which never move score (the spread of their
  start points), min over the D coordinates.
  Then, untimed or timed separately and reported beside `value`:
  4. plain fp32 HMC for --plain-steps transitions from the same states
     (sampling time only; its own warm-up is reported as a derived figure);
  5. posterior means and variances of the surrogate samples against the
     FP64 oracle's (tests/golden/posterior_large.npz), in units of the
     combined Monte Carlo error;
  6. the Laplace-start variant, the small / mid / hierarchical
     configurations, and the CPU restatement timed on the host cores.
`roofline` is the dominant kernel of the surrogate transitions; its launch
time is measured inside the timed region with hipEvents the library records
on its launch stream around that kernel only (per-class times come from 5
untimed transitions afterwards).
"""

import argparse
import math
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
  sys.path.insert(0, ROOT)

PEAK = {'bf16': 2.5e15, 'fp32': 157.3e12, 'hbm': 8.0e12}


def log(msg):
  sys.stderr.write('[bench %7.1fs] %s\n' % (time.time() - T0, msg))
  sys.stderr.flush()


T0 = time.time()


def parse():
  ap = argparse.ArgumentParser()
  ap.add_argument('--gpus', type=int, default=1)
  ap.add_argument('--steps', type=int, default=1000,
                  help='surrogate transitions kept as samples (timed)')
  ap.add_argument('--warmup', type=int, default=2,
                  help='untimed plain bf16 transitions (kernel load, first touch)')
  ap.add_argument('--burn-in', type=int, default=80,
                  help='max plain bf16 descent transitions before pair recording starts')
  ap.add_argument('--burn-in-steps', type=int, default=4,
                  help='leapfrog steps of the descent transitions')
  ap.add_argument('--rec-start-tol', type=float, default=100.0,
                  help='pair recording starts when the mean log joint rises less than this '
                       'over 5 descent transitions')
  ap.add_argument('--init', default='prior', choices=['laplace', 'prior'],
                  help="chain start: the workload's N(0, I) draws (the recipe), or draws from "
                       'the Laplace approximation N(mode, P^-1) found by device Newton steps')
  ap.add_argument('--laplace-burn-in', type=int, default=2,
                  help='transitions before pair recording starts after a Laplace start '
                       '(with --pairs trajectory)')
  ap.add_argument('--pairs', default='trajectory', choices=['laplace', 'trajectory'],
                  help='surrogate training pairs: recorded along warm-up trajectories (the '
                       'recipe), or exact gradients at Laplace draws (with --init laplace)')
  ap.add_argument('--n-pairs', type=int, default=114688,
                  help='training pairs drawn from the Laplace approximation')
  ap.add_argument('--burn-in-adapt', type=int, default=1,
                  help='1: the burn-in step-size centre follows the acceptance rate')
  ap.add_argument('--plain-steps', type=int, default=16, help='timed plain fp32 transitions')
  ap.add_argument('--workload', default='large',
                  choices=['large', 'hier', 'copy', 'forward', 'gradient', 'surrogate'],
                  help="'hier': the hierarchical logistic regression (G=2048 groups of "
                       "1+Poisson(400) rows, D=64, 512 chains, L=16), its own JSON line; "
                       "'copy': 1 GiB device-to-device copies, the byte-counter calibration "
                       "of HBM traffic profiles; 'forward': --steps exact forward passes "
                       "(the gemm_ll kernel alone) on the large workload, for profiling; "
                       "'gradient': --steps plain bf16 transitions of 4 leapfrog steps (the "
                       "gradient-pass kernels); "
                       "'surrogate': --steps surrogate transitions after a short warm start, "
                       "for profiling")
  ap.add_argument('--hidden', type=int, default=1024)
  ap.add_argument('--epochs', type=int, default=16)
  ap.add_argument('--lr', type=float, default=3e-3)
  ap.add_argument('--rec-stride', type=int, default=1)
  ap.add_argument('--rec-transitions', type=int, default=2)
  ap.add_argument('--rec-steps', type=int, default=25,
                  help='leapfrog steps of the recording transitions (0: the sampler\'s L); '
                       '2 x 24 inner points x 4096 chains = 196,608 pairs')
  ap.add_argument('--surrogate-only', action='store_true',
                  help='stop after the surrogate phase (tuning runs)')
  ap.add_argument('--seed', type=int, default=0)
  ap.add_argument('--jitter', type=float, default=0.2,
                  help='step size drawn from eps * U(1-j, 1+j) per transition (both samplers)')
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--no-extra', action='store_true',
                  help='skip the Laplace-start variant and the small / mid / hierarchical '
                       'configurations')
  ap.add_argument('--cpu-seconds', type=float, default=15.0)
  ap.add_argument('--cpu-chains', type=int, default=512)
  return ap.parse_args()


def cpu_model():
  try:
    for line in open('/proc/cpuinfo'):
      if line.startswith('model name'):
        return line.split(':', 1)[1].strip()
  except OSError:
    pass
  return 'unknown'


def cpu_baseline(wl, budget_s, ess_per_transition_plain, chains=512):
  """The oracle's fp32 restatement of one plain transition's work (L
  gradient evaluations: eta = W X^T, r = y - sigmoid(eta), g = r X - W, all
  float32) timed on a block of `chains` chains at the full N on the host
  cores, scaled to the workload's chains; CPU min-ESS/s = the GPU fp32
  run's ESS per transition / CPU seconds per transition (derived)."""
  from oracle import hmc_oracle as O  # the CPU restatement (checker) only
  cores = len(os.sched_getaffinity(0))
  Cs = min(int(chains), wl.C)
  W = wl.W0[:Cs].astype(np.float32)
  X = np.ascontiguousarray(wl.X, np.float32)
  y = np.asarray(wl.y, np.float32)
  reps, t = 0, 0.0
  while (t < budget_s or reps < 1) and reps < 50:
    t0 = time.perf_counter()
    O.logreg_grad_f32(X, y, W, threads=cores)
    t += time.perf_counter() - t0
    reps += 1
  sec_per_eval = t / reps * (wl.C / Cs)
  sec_per_transition = sec_per_eval * wl.n_steps
  return {
      'value': ess_per_transition_plain / sec_per_transition,
      'unit': 'min-ESS/s (derived: GPU fp32 mixing ESS per transition / CPU sec per '
              'transition)',
      'cores': cores,
      'kind': 'port',
      'cpu_model': cpu_model(),
      'sample': ('%d fp32 gradient evaluations of %d of %d chains at N=%d, D=%d (numpy '
                 'sgemm on %d threads, elementwise in %d threads), %.2f s each; scaled x%g '
                 'chains x L=%d -> %.1f s per transition'
                 % (reps, Cs, wl.C, wl.N, wl.D, cores, cores, t / reps, wl.C / Cs,
                    wl.n_steps, sec_per_transition)),
      'sec_per_transition': sec_per_transition,
  }


def ess_summary(engine, samples):
  """Per-coordinate multi-chain ESS of samples [n, M, D] and what the
  headline uses from it: the floor M n / (2n - 1) that never-moving chains
  score, the mixing ESS (multi-chain minus the floor, min over coordinates)
  and the within-chain ESS (the chain-mean spread left out)."""
  from edward_amd import diagnostics
  n, M = int(samples.shape[0]), int(samples.shape[1])
  e = diagnostics.ess(engine, samples)
  floor = M * n / (2.0 * n - 1.0)
  within = diagnostics.ess(engine, samples, between=False)
  return {'min_ess': float(e.min()), 'ess_floor': floor,
          'min_ess_mixing': float(np.maximum(e - floor, 0.0).min()),
          'min_ess_within_chain': float(within.min()), 'per_coord': e}


def data_matches_fixture(f, X, y):
  """Whether the workload (X, y) is the one a posterior fixture was computed
  on: the fingerprints tests/golden/make_posterior_large.py stores."""
  n = np.arange(X.shape[0], dtype=np.float64)
  return bool(np.isclose(float(X.sum(dtype=np.float64)), float(f['x_sum']), rtol=1e-9,
                         atol=1e-6) and
              np.isclose(float(n @ X[:, 0].astype(np.float64)), float(f['x_col0_moment']),
                         rtol=1e-9) and
              float(y.sum(dtype=np.float64)) == float(f['y_sum']) and
              np.array_equal(np.asarray(X[:4, :8], np.float32), f['x_head']))


def moment_check(samples, ess_d, path, wl):
  """Posterior means and variances of samples [T, C, D] against the FP64
  oracle's (importance-sampled Laplace, tests/golden/posterior_large.npz):
  z = (device - oracle) / sqrt(se_device^2 + se_oracle^2), se_device from
  the per-coordinate ESS (mean: sd / sqrt(ESS); variance: var sqrt(2 / ESS))."""
  if not os.path.exists(path):
    return None
  f = np.load(path)
  if not data_matches_fixture(f, wl.X, wl.y):
    return {'oracle': path, 'data_check': False}
  T, C, D = samples.shape
  mean = samples.double().mean((0, 1)).cpu().numpy()
  var = samples.double().var((0, 1), unbiased=False).cpu().numpy()
  ess_d = np.maximum(np.asarray(ess_d, np.float64), 1.0)
  se_m = np.sqrt(var / ess_d)
  se_v = var * np.sqrt(2.0 / ess_d)
  zm = (mean - f['mean']) / np.sqrt(se_m ** 2 + f['se_mean'] ** 2)
  zv = (var - f['var']) / np.sqrt(se_v ** 2 + f['se_var'] ** 2)
  sd = np.sqrt(f['var'])
  return {'oracle': 'FP64 Laplace + importance sampling, %d antithetic pairs '
                    '(tests/golden/posterior_large.npz)' % int(f['n_pairs']),
          'mean_abs_err_max_in_sd': float(np.abs(mean - f['mean']).max() / sd.min()),
          'mean_z_max': float(np.abs(zm).max()), 'mean_z_rms': float(np.sqrt((zm ** 2).mean())),
          'var_rel_err_max': float(np.abs(var / f['var'] - 1.0).max()),
          'var_z_max': float(np.abs(zv).max()), 'var_z_rms': float(np.sqrt((zv ** 2).mean())),
          'frac_abs_z_above_3': float(((np.abs(zm) > 3) | (np.abs(zv) > 3)).mean()),
          'laplace_mean_z_max': float(np.abs((f['mode'] - f['mean']) /
                                             np.sqrt(se_m ** 2 + f['se_mean'] ** 2)).max()),
          'data_check': True}


def laplace_variant(bf, wl, args, K):
  """The Laplace-start recipe: device Newton steps, chains and n_pairs
  training pairs drawn from N(mode, P^-1), training, K surrogate
  transitions; inclusive mixing min-ESS/s."""
  import torch
  from edward_amd.engine import jittered_step
  from edward_amd.surrogate import Surrogate
  from edward_amd.warmstart import LaplaceApproximation
  C, D, eps, L = wl.C, wl.D, wl.step_size, wl.n_steps
  dev = torch.device('cuda', 0)
  torch.cuda.synchronize()
  t0 = time.time()
  lap = LaplaceApproximation(bf)
  Wp, Gp = lap.gradient_pairs(args.n_pairs, C, seed=args.seed + 5)
  bf.set_state(lap.sample(C, seed=args.seed + 3))
  torch.cuda.synchronize()
  warm_s = time.time() - t0
  sur = Surrogate(bf, hidden=args.hidden, max_batch=C, seed=args.seed)
  # 32 epochs over its 114,688 pairs: about as many Adam steps as the
  # recipe's 16 epochs over ~200k
  sur.fit(Wp, Gp, epochs=32, batch=C, lr=args.lr, lr_final=5e-5, seed=args.seed,
          val_frac=0.05)
  torch.cuda.synchronize()
  train_s = time.time() - t0 - warm_s
  del Wp, Gp
  samples = torch.empty(K, C, D, device=dev)
  n_acc = torch.zeros(C, dtype=torch.int32, device=dev)
  ts = time.time()
  for k in range(K):
    sur.transition(jittered_step(eps, args.jitter, args.seed + 2, k), L, seed=args.seed + 2,
                   counter=k, sample_out=samples[k], n_accept=n_acc)
  torch.cuda.synchronize()
  samp_s = time.time() - ts
  es = ess_summary(bf, samples)
  incl = time.time() - t0
  del samples
  torch.cuda.empty_cache()
  return {'min_ess_mixing_per_s': es['min_ess_mixing'] / incl, 'seconds': incl,
          'warm_start_s': warm_s, 'train_s': train_s, 'sampling_s': samp_s,
          'newton_iterations': lap.iterations, 'train_pairs': args.n_pairs,
          'min_ess': es['min_ess'], 'min_ess_mixing': es['min_ess_mixing'],
          'ess_floor': es['ess_floor'], 'transitions': K,
          'accept': float(n_acc.cpu().numpy().mean()) / K}


def small_config(args):
  """The small configuration in full: N=1000, D=10, 1 chain from 0, eps=0.01,
  L=10, 5000 transitions (fp32), on the GPU (one edhmc_run call) and as the
  oracle's CPU restatement (numpy FP64, the same number of transitions)."""
  import torch
  from edward_amd import workloads as Wk
  from edward_amd.engine import GLMEngine
  from oracle import hmc_oracle as O  # the CPU restatement (checker) only
  wl = Wk.logreg_small()
  eng = GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale)
  eng.set_state(wl.W0)
  params = torch.zeros(wl.T, 1, wl.D, device='cuda')
  n_acc = torch.zeros(1, dtype=torch.int32, device='cuda')
  eng.run(params[:50], wl.step_size, wl.n_steps, seed=args.seed + 11)   # load the kernels
  eng.set_state(wl.W0)
  torch.cuda.synchronize()
  t0 = time.time()
  eng.run(params, wl.step_size, wl.n_steps, seed=args.seed + 12, n_accept=n_acc)
  torch.cuda.synchronize()
  gpu_s = time.time() - t0
  burn = wl.T // 10
  e_gpu = ess_summary(eng, params[burn:])
  acc = float(n_acc.item()) / wl.T
  del eng
  m = O.GLM(wl.X.astype(np.float64), wl.y, wl.prior_loc, wl.prior_scale)
  rng = np.random.default_rng(args.seed + 13)
  mom = rng.standard_normal((wl.T, 1, wl.D))
  uni = rng.random((wl.T, 1))
  cparams = np.zeros((wl.T, 1, wl.D))
  t0 = time.time()
  O.run_hmc(cparams, mom, uni, wl.step_size, wl.n_steps, m.logp_grad)
  cpu_s = time.time() - t0
  cpu_ess = min(O.ess(cparams[burn:, 0, d][None, :], between=False) for d in range(wl.D))
  return {'workload': 'logreg N=1000 D=10, 1 chain from 0, eps=0.01, L=10, T=5000, fp32',
          'gpu_wall_s': gpu_s, 'gpu_ms_per_transition': gpu_s / wl.T * 1e3,
          'gpu_min_ess_within_chain': e_gpu['min_ess_within_chain'],
          'gpu_min_ess_per_s': e_gpu['min_ess_within_chain'] / gpu_s, 'gpu_accept': acc,
          'cpu_wall_s': cpu_s, 'cpu_min_ess_within_chain': float(cpu_ess),
          'cpu_min_ess_per_s': float(cpu_ess) / cpu_s,
          'cpu_kind': 'port (oracle FP64 numpy, 1 thread), all 5000 transitions',
          'ess_window': 'transitions %d..%d' % (burn, wl.T)}


def mid_config(args):
  """The mid configuration, fp32: N=1e5, D=256 (column-scaled X), 1024
  chains from N(0, I), eps=0.002, L=20: warm-up until the mean log joint
  settles, then timed transitions; ms per transition, mixing min-ESS/s and
  the fp32 GEMM roofline of one gradient pass."""
  import torch
  from edward_amd import workloads as Wk
  from edward_amd.engine import GLMEngine
  wl = Wk.logreg_mid()
  C, D, eps, L = wl.C, wl.D, wl.step_size, wl.n_steps
  eng = GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale, precision='fp32')
  eng.set_state(wl.W0)
  means = []
  t = 0
  t0 = time.time()
  while t < 80:
    eng.transition(eps, L, seed=args.seed + 21, counter=t)
    _, lp, _ = eng.get_state()
    means.append(float(lp.cpu().numpy().mean()))
    t += 1
    if t >= 15 and abs(means[-1] - means[-6]) < 5.0:
      break
  torch.cuda.synchronize()
  warm_s = time.time() - t0
  K = 40
  samples = torch.empty(K, C, D, device='cuda')
  n_acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  torch.cuda.synchronize()
  t0 = time.time()
  for k in range(K):
    eng.transition(eps, L, seed=args.seed + 21, counter=t + k, sample_out=samples[k],
                   n_accept=n_acc)
  torch.cuda.synchronize()
  secs = time.time() - t0
  es = ess_summary(eng, samples)
  eng.profile(True)
  for k in range(3):
    eng.transition(eps, L, seed=args.seed + 21, counter=t + K + k)
  prof = eng.profile_read()
  eng.profile(False)
  g_ms = sum(prof[k][0] for k in ('gemm_grad', 'gemm_ll', 'gemm_xtr'))
  n_pass = prof['gemm_grad'][1] + prof['gemm_ll'][1]
  achieved = 4.0 * wl.N * D * C / (g_ms / max(n_pass, 1) * 1e-3)
  del samples, eng
  torch.cuda.empty_cache()
  return {'workload': 'logreg N=1e5 D=256 1024 chains eps=0.002 L=20 fp32 (from N(0, I))',
          'warmup_transitions': t, 'warmup_s': warm_s, 'transitions': K,
          'ms_per_transition': secs / K * 1e3, 'accept': float(n_acc.cpu().numpy().mean()) / K,
          'min_ess_mixing_per_s': es['min_ess_mixing'] / secs, 'min_ess': es['min_ess'],
          'ess_floor': es['ess_floor'],
          'roofline': {'bound': 'mfma', 'kernel': 'one fp32 gradient pass (eta = W X^T, X^T R)',
                       'achieved': achieved / 1e12, 'peak': PEAK['fp32'] / 1e12,
                       'unit': 'TFLOP/s', 'frac': achieved / PEAK['fp32']}}


def hier_config(args):
  """The hierarchical configuration (G=2048 groups of 1+Poisson(400) rows,
  D=64, 512 chains, eps=1e-3, L=16, fp32) from near the generating values:
  ms per transition and acceptance over a few timed transitions."""
  import torch
  from edward_amd import workloads as Wk
  from edward_amd.hier import HierEngine
  wl = Wk.hier_groups(seed=args.seed)
  eng = HierEngine(wl.X, wl.y, wl.offsets)
  eng.set_state(wl.theta0)
  for t in range(2):
    eng.transition(wl.step_size, wl.n_steps, seed=args.seed + 31, counter=t)
  n_acc = torch.zeros(wl.C, dtype=torch.int32, device='cuda')
  K = 10
  torch.cuda.synchronize()
  t0 = time.time()
  for t in range(K):
    eng.transition(wl.step_size, wl.n_steps, seed=args.seed + 31, counter=2 + t, n_accept=n_acc)
  torch.cuda.synchronize()
  secs = time.time() - t0
  del eng
  torch.cuda.empty_cache()
  return {'workload': 'hier_logreg G=%d N=%d D=%d chains=%d eps=%g L=%d fp32'
                      % (wl.G, wl.N, wl.D, wl.C, wl.step_size, wl.n_steps),
          'transitions': K, 'ms_per_transition': secs / K * 1e3,
          'accept': float(n_acc.cpu().numpy().mean()) / K}


def bench_hier(args):
  """Plain fp32 HMC on the hierarchical model: transitions/s over an
  unprofiled timed region, then the kernel classes' times from a few more
  transitions with every class bracketed by events."""
  import torch
  from edward_amd import workloads as W
  from edward_amd.hier import HierEngine
  wl = W.hier_groups(seed=args.seed)
  log('hier workload G=%d N=%d D=%d C=%d L=%d' % (wl.G, wl.N, wl.D, wl.C, wl.n_steps))
  eng = HierEngine(wl.X, wl.y, wl.offsets)
  eng.set_state(wl.theta0)
  C, L, eps = wl.C, wl.n_steps, wl.step_size
  dev = torch.device('cuda', 0)
  n_acc = torch.zeros(C, dtype=torch.int32, device=dev)
  for t in range(args.warmup):
    eng.transition(eps, L, seed=args.seed, counter=t, n_accept=n_acc)
  torch.cuda.synchronize(dev)
  n_acc.zero_()
  K = args.steps
  t0 = time.time()
  for t in range(K):
    eng.transition(eps, L, seed=args.seed, counter=args.warmup + t, n_accept=n_acc)
  torch.cuda.synchronize(dev)
  secs = time.time() - t0
  acc = float(n_acc.double().mean().item()) / K
  n_break = 3
  eng.profile(True)
  for t in range(n_break):
    eng.transition(eps, L, seed=args.seed, counter=args.warmup + K + t)
  prof = eng.profile_read()
  eng.profile(False)
  # one gradient pass: eta = X w (2NDC) and X^T R (2NDC) flops
  flops = 4.0 * wl.N * wl.D * C
  g_ms, g_n = prof['gemm_grad']
  x_ms, x_n = prof['gemm_xtr']
  ll_ms, ll_n = prof['gemm_ll']
  n_grad = g_n + ll_n
  gemm_ms = (g_ms + ll_ms + x_ms) / max(n_grad, 1)
  achieved = flops / (gemm_ms * 1e-3)
  log('hier: %.2f ms/transition, accept %.3f, profile %s' % (secs / K * 1e3, acc, prof))
  out = {
      'metric': 'hierarchical logreg HMC transitions/sec (G=2048, D=64, 512 chains, L=16)',
      'value': K / secs, 'unit': 'transitions/s', 'n_gpus': 1, 'steps': K,
      'warmup': args.warmup, 'ms_per_step': secs / K * 1e3, 'higher_is_better': True,
      'scaling': 'weak', 'vs_baseline': None, 'dtype': 'fp32', 'data': 'synthetic',
      'config': {'workload': 'hier_logreg G=%d N=%d D=%d chains=%d eps=%g L=%d'
                             % (wl.G, wl.N, wl.D, C, eps, L)},
      'accept': acc,
      'profile_ms_per_transition': {k: v[0] / n_break for k, v in prof.items()},
      'profile_from': '%d transitions after the timed region, every kernel class bracketed '
                      'by events' % n_break,
      'roofline': {'bound': 'mfma', 'kernel': 'gemm_grad+gemm_xtr (one gradient pass)',
                   'achieved': achieved / 1e12, 'peak': PEAK['fp32'] / 1e12,
                   'unit': 'TFLOP/s', 'frac': achieved / PEAK['fp32'], 'traffic': None},
  }
  print(json.dumps(out))


def measured_traffic(kernel, workload):
  """HBM bytes per launch of `kernel` from the committed rocprofv3 byte
  counters (profiles/traffic.json: FETCH_SIZE and WRITE_SIZE per dispatch,
  each scaled by its reading on a copy of known size), or None."""
  path = os.path.join(ROOT, 'profiles', 'traffic.json')
  if workload != 'large' or not os.path.exists(path):
    return None
  rec = json.load(open(path)).get(kernel)
  if not rec:
    return None
  return 1024.0 * (rec['fetch_kib_per_dispatch'] * rec['fetch_scale'] +
                   rec['write_kib_per_dispatch'] * rec['write_scale'])


def bench_copy(args):
  """Copies of a 1 GiB buffer: a kernel of known traffic (1 GiB read and
  1 GiB written per launch) against which the profiler's byte counters are
  calibrated."""
  import torch
  n = 1 << 28
  a = torch.ones(n, device='cuda')
  b = torch.empty_like(a)
  torch.cuda.synchronize()
  t0 = time.time()
  for _ in range(args.steps):
    b.copy_(a)
  torch.cuda.synchronize()
  secs = time.time() - t0
  print(json.dumps({'metric': 'device copy bandwidth', 'value': 2.0 * 4 * n * args.steps / secs
                    / 1e9, 'unit': 'GB/s', 'bytes_read_per_launch': 4 * n,
                    'bytes_written_per_launch': 4 * n, 'launches': args.steps}))


def bench_forward(args):
  """--steps exact forward log-joint evaluations (gemm_ll) of 4096 chains on
  the large logistic-regression workload: the dominant kernel of the
  surrogate transition in isolation, for counter profiles."""
  import torch
  from edward_amd import workloads as W
  from edward_amd.engine import GLMEngine
  wl = W.logreg_large()
  rng = np.random.default_rng(args.seed)
  Wp = torch.tensor((wl.w_true[None, :] + 0.003 * rng.normal(size=(wl.C, wl.D)))
                    .astype(np.float32), device='cuda')
  eng = GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale, precision='bf16')
  eng.logp(Wp)
  torch.cuda.synchronize()
  eng.profile(True)
  t0 = time.time()
  for _ in range(args.steps):
    eng.logp(Wp)
  torch.cuda.synchronize()
  secs = time.time() - t0
  ms, n = eng.profile_read()['gemm_ll']
  flop = 2.0 * wl.N * wl.D * wl.C
  print(json.dumps({'metric': 'exact forward log joint, N=%d D=%d chains=%d' % (wl.N, wl.D, wl.C),
                    'value': args.steps / secs, 'unit': 'evaluations/s', 'steps': args.steps,
                    'gemm_ll_ms': ms / max(n, 1),
                    'achieved_tflops': flop / (ms / max(n, 1) * 1e-3) / 1e12}))


def bench_gradient(args):
  """--steps plain bf16 HMC transitions of 4 leapfrog steps (the warm-up's
  descent transitions: three one-plane inner gradient passes and one exact
  two-plane pass with the log-likelihood, each GEMM1 with the residual
  epilogue plus X^T R) of 4096 chains near w* on the large workload, for
  counter profiles of the gradient-pass kernels."""
  import torch
  from edward_amd import workloads as W
  from edward_amd.engine import GLMEngine
  wl = W.logreg_large()
  rng = np.random.default_rng(args.seed)
  Wp = torch.tensor((wl.w_true[None, :] + 0.003 * rng.normal(size=(wl.C, wl.D)))
                    .astype(np.float32), device='cuda')
  eng = GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale, precision='bf16')
  eng.set_state(Wp)
  L = 4
  eng.transition(wl.step_size, L, seed=args.seed)
  torch.cuda.synchronize()
  eng.profile(True)
  t0 = time.time()
  for k in range(args.steps):
    eng.transition(wl.step_size, L, seed=args.seed, counter=k + 1)
  torch.cuda.synchronize()
  secs = time.time() - t0
  pr = eng.profile_read()
  flop = 2.0 * wl.N * wl.D * wl.C
  g1_ms, g1_n = pr['gemm_grad']
  g2_ms, g2_n = pr['gemm_xtr']
  print(json.dumps({'metric': 'plain bf16 HMC transitions of L=%d, N=%d D=%d chains=%d'
                              % (L, wl.N, wl.D, wl.C),
                    'value': args.steps / secs, 'unit': 'transitions/s', 'steps': args.steps,
                    'gemm_grad_ms_per_transition': g1_ms / args.steps,
                    'gemm_xtr_ms': g2_ms / max(g2_n, 1),
                    'gemm_xtr_tflops': flop / (g2_ms / max(g2_n, 1) * 1e-3) / 1e12}))


def bench_surrogate(args):
  """--steps surrogate transitions on the large workload after a short
  Laplace warm start and two training epochs: the sampling window's kernels
  (the exact forward and the MLP layers with their leapfrog epilogues) with
  little else around them, for counter profiles."""
  import torch
  from edward_amd import workloads as Wk
  from edward_amd.engine import GLMEngine
  from edward_amd.surrogate import Surrogate
  from edward_amd.warmstart import LaplaceApproximation
  wl = Wk.logreg_large()
  C = wl.C
  eng = GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale, precision='bf16')
  lap = LaplaceApproximation(eng)
  Wp, Gp = lap.gradient_pairs(4 * C, C, seed=args.seed + 1)
  sur = Surrogate(eng, hidden=args.hidden, max_batch=C, seed=args.seed)
  sur.fit(Wp, Gp, epochs=2, batch=C, lr=args.lr, lr_final=1e-3, seed=args.seed)
  eng.set_state(lap.sample(C, seed=args.seed + 2))
  n_acc = torch.zeros(C, dtype=torch.int32, device='cuda')
  torch.cuda.synchronize()
  t0 = time.time()
  for k in range(args.steps):
    sur.transition(wl.step_size, wl.n_steps, seed=args.seed, counter=k, n_accept=n_acc)
  torch.cuda.synchronize()
  secs = time.time() - t0
  # per-class kernel times over 5 more transitions, every class bracketed
  eng.profile(True)
  for k in range(5):
    sur.transition(wl.step_size, wl.n_steps, seed=args.seed, counter=args.steps + k)
  torch.cuda.synchronize()
  per = {kind: ms / 5 for kind, (ms, n) in eng.profile_read().items() if n}
  eng.profile(False)
  print(json.dumps({'metric': 'surrogate transitions, N=%d D=%d chains=%d L=%d H=%d'
                              % (wl.N, wl.D, C, wl.n_steps, args.hidden),
                    'value': args.steps / secs, 'unit': 'transitions/s', 'steps': args.steps,
                    'ms_per_transition': secs / args.steps * 1e3,
                    'accept': float(n_acc.cpu().numpy().mean()) / args.steps,
                    'per_transition_ms': per}))


def main():
  args = parse()
  if args.gpus != 1:
    print('bench.py measures one GPU: this build runs one process on one MI355X; '
          'multi-GPU scaling is not part of it (requested --gpus %d)' % args.gpus,
          file=sys.stderr)
    sys.exit(2)
  if args.workload == 'hier':
    return bench_hier(args)
  if args.workload == 'copy':
    return bench_copy(args)
  if args.workload == 'forward':
    return bench_forward(args)
  if args.workload == 'surrogate':
    return bench_surrogate(args)
  if args.workload == 'gradient':
    return bench_gradient(args)
  import torch
  from edward_amd import workloads as Wk
  from edward_amd.engine import GLMEngine, jittered_step
  from edward_amd.surrogate import Surrogate

  wl = Wk.logreg_large()
  N, D, C, eps, L = wl.N, wl.D, wl.C, wl.step_size, wl.n_steps
  log('workload N=%d D=%d C=%d eps=%g L=%d' % (N, D, C, eps, L))
  dev = torch.device('cuda', 0)
  bf = GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale, precision='bf16')
  bf.set_state(wl.W0)
  n_acc = torch.zeros(C, dtype=torch.int32, device=dev)
  for t in range(args.warmup):     # untimed
    bf.transition(eps, L, seed=args.seed + 99, counter=t, n_accept=n_acc)
  bf.set_state(wl.W0)
  # one epoch of a throwaway surrogate on random pairs (enough that the
  # skip's least squares takes its full row sample): the first use of the
  # training path's kernels and library calls costs ~0.3 s per process
  gen = torch.Generator(device=dev)
  gen.manual_seed(args.seed + 98)
  nw = 9 * C
  wsur = Surrogate(bf, hidden=args.hidden, max_batch=C, seed=args.seed + 97)
  wsur.fit(torch.randn(nw, D, generator=gen, device=dev),
           torch.randn(nw, D, generator=gen, device=dev), epochs=1, batch=C, lr=args.lr,
           lr_final=5e-5, seed=args.seed + 97, val_frac=0.05)
  del wsur
  # the same for the Laplace warm start's kernels (Newton terms, draws)
  from edward_amd.warmstart import LaplaceApproximation
  LaplaceApproximation(bf, max_iter=2).sample(C, seed=args.seed + 96)
  bf.set_state(wl.W0)
  torch.cuda.synchronize()

  # timed region: warm-up (+ pair recording) + training + K sampling transitions
  t_warm0 = time.time()
  burn_in = args.burn_in
  laplace_info = None
  pairs_from_laplace = args.init == 'laplace' and args.pairs == 'laplace'
  if args.init == 'laplace':
    from edward_amd.warmstart import LaplaceApproximation
    lap = LaplaceApproximation(bf)
    laplace_info = {'newton_iterations': lap.iterations, 'seconds': time.time() - t_warm0,
                    'logp_mode': lap.logp_mode}
    log('laplace start: %d Newton steps, %.2f s, logp(mode) %.2f'
        % (lap.iterations, laplace_info['seconds'], lap.logp_mode))
    if pairs_from_laplace:
      Wrec, Grec = lap.gradient_pairs(args.n_pairs, C, seed=args.seed + 5)
      torch.cuda.synchronize()
      laplace_info['pairs_seconds'] = time.time() - t_warm0 - laplace_info['seconds']
      log('training pairs: %d exact gradients at Laplace draws, %.2f s'
          % (args.n_pairs, laplace_info['pairs_seconds']))
    bf.set_state(lap.sample(C, seed=args.seed + 3))
    burn_in = args.laplace_burn_in
  # warm-up with plain bf16 HMC in two phases.  Descent: short trajectories
  # (--burn-in-steps leapfrog steps, hi-plane inner gradients): every
  # momentum refresh discards the kinetic energy the fall toward the mode
  # built up, so short trajectories shed the start's excess energy in fewer
  # gradient evaluations.  It ends once the mean log joint rises by less
  # than --rec-start-tol over 5 transitions (or after --burn-in transitions).
  # Recording: --rec-transitions transitions of the sampler's own L and step
  # size with exact (two-plane) gradients throughout, (w, grad) pairs taken
  # at every --rec-stride-th inner leapfrog point.
  means = []
  rec_stride = args.rec_stride
  n_rec = args.rec_transitions
  Lr = args.rec_steps if args.rec_steps > 0 else L
  if not pairs_from_laplace:
    cap = ((Lr - 1) // rec_stride) * C * n_rec
    Wrec = torch.empty(cap, D, device=dev)
    Grec = torch.empty(cap, D, device=dev)
  t = 0
  n_descent = 0
  jit = np.random.default_rng(args.seed + 7)
  log_c = math.log(eps)
  Lb = args.burn_in_steps
  while not pairs_from_laplace and t < burn_in:
    # the step size is jittered log-uniformly over [c/2, 2c]: with a fixed
    # eps L, directions whose frequency makes eps L a multiple of half a
    # period keep their energy through every momentum refresh.  With
    # --burn-in-adapt the centre c follows the acceptance rate within
    # [eps/4, 4 eps].
    eps_t = math.exp(log_c) * 2.0 ** jit.uniform(-1.0, 1.0)
    n_acc.zero_()
    bf.transition(eps_t, Lb, seed=args.seed, counter=t, n_accept=n_acc)
    _, lp, _ = bf.get_state()
    means.append(float(lp.cpu().numpy().mean()))
    acc_t = float(n_acc.cpu().numpy().mean())
    log('descent %d eps %.2e L %d mean logp %.2f accept %.3f' % (t, eps_t, Lb, means[-1], acc_t))
    if args.burn_in_adapt:
      log_c = min(max(log_c + 0.6 * (acc_t - 0.7), math.log(eps / 4)), math.log(4 * eps))
    t += 1
    if t >= 6 and means[-1] - means[-6] < args.rec_start_tol:
      break
  n_descent = t
  if not pairs_from_laplace:
    bf.set_inner_gradient(True)
    bf.record_pairs(Wrec, Grec, stride=rec_stride)
    for k in range(n_rec):
      eps_t = jittered_step(eps, args.jitter, args.seed, t)
      n_acc.zero_()
      bf.transition(eps_t, Lr, seed=args.seed, counter=t, n_accept=n_acc)
      _, lp, _ = bf.get_state()
      means.append(float(lp.cpu().numpy().mean()))
      log('recording %d eps %.2e L %d mean logp %.2f accept %.3f'
          % (t, eps_t, Lr, means[-1], float(n_acc.cpu().numpy().mean())))
      t += 1
    bf.set_inner_gradient(False)
  torch.cuda.synchronize()
  if pairs_from_laplace:
    n_pairs, n_rec = args.n_pairs, 0
  else:
    bf.record_pairs(None)
    n_pairs = bf.record_count()
  n_warm = t
  warm_s = time.time() - t_warm0
  log('warm-up done: %d transitions, %d pairs, %.1f s' % (t, n_pairs, warm_s))

  # train the surrogate
  t_tr0 = time.time()
  sur = Surrogate(bf, hidden=args.hidden, max_batch=C, seed=args.seed)
  losses = sur.fit(Wrec[:n_pairs], Grec[:n_pairs], epochs=args.epochs, batch=C, lr=args.lr,
                   lr_final=5e-5, seed=args.seed, log=log, val_frac=0.05)
  torch.cuda.synchronize()
  train_s = time.time() - t_tr0
  del Wrec, Grec
  log('training done: %.1f s, loss %.4f -> %.4f' % (train_s, losses[0], losses[-1]))

  # timed surrogate sampling
  K = args.steps
  samples = torch.empty(K, C, D, device=dev)
  n_acc.zero_()
  # inside the timed region only the exact forward (the dominant kernel) is
  # bracketed by events: every timed launch adds two event records, ~10 us
  # between dependent launches
  bf.profile(True, kinds=('gemm_ll',))
  torch.cuda.synchronize()
  ts0 = time.time()
  for k in range(K):
    sur.transition(jittered_step(eps, args.jitter, args.seed, t), L, seed=args.seed,
                   counter=t, sample_out=samples[k], n_accept=n_acc)
    t += 1
  torch.cuda.synchronize()
  t_end = time.time()
  sur_s = t_end - ts0
  incl_s = t_end - t_warm0
  prof = bf.profile_read()
  bf.profile(False)
  acc_sur = float(n_acc.cpu().numpy().mean()) / K
  es_sur = ess_summary(bf, samples)
  Kp = args.plain_steps
  Kw = min(K, Kp)
  es_sur_kw = ess_summary(bf, samples[:Kw])
  log('surrogate: %d transitions %.2f s, accept %.3f, min-ESS %.1f (floor %.1f, mixing %.1f)'
      % (K, sur_s, acc_sur, es_sur['min_ess'], es_sur['ess_floor'], es_sur['min_ess_mixing']))
  moments = moment_check(samples, es_sur['per_coord'],
                         os.path.join(ROOT, 'tests', 'golden', 'posterior_large.npz'), wl)
  log('moments vs the FP64 oracle: %s' % moments)
  state, _, _ = bf.get_state()
  # per-class kernel times from a short untimed run with every class timed
  n_break = 5
  bf.profile(True)
  for k in range(n_break):
    sur.transition(jittered_step(eps, args.jitter, args.seed, t + k), L, seed=args.seed,
                   counter=t + k)
  breakdown = bf.profile_read()
  bf.profile(False)
  del samples
  torch.cuda.empty_cache()

  if args.surrogate_only:
    print(json.dumps({'surrogate_min_ess_mixing_per_s': es_sur['min_ess_mixing'] / incl_s,
                      'accept': acc_sur, 'final_loss': losses[-1],
                      'ms_per_step': incl_s / K * 1e3, 'moments': moments}))
    return

  # plain fp32 from the same states
  f32 = GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale, precision='fp32')
  f32.set_state(state)
  psamples = torch.empty(Kp, C, D, device=dev)
  n_acc.zero_()
  torch.cuda.synchronize()
  tp0 = time.time()
  for k in range(Kp):
    f32.transition(jittered_step(eps, args.jitter, args.seed + 1, t + k), L,
                   seed=args.seed + 1, counter=t + k, sample_out=psamples[k],
                   n_accept=n_acc)
    torch.cuda.synchronize()
    log('plain fp32 transition %d' % k)
  plain_s = time.time() - tp0
  acc_plain = float(n_acc.cpu().numpy().mean()) / Kp
  es_plain = ess_summary(f32, psamples)
  log('plain fp32: %d transitions %.2f s, accept %.3f, min-ESS %.1f (floor %.1f, mixing %.1f)'
      % (Kp, plain_s, acc_plain, es_plain['min_ess'], es_plain['ess_floor'],
         es_plain['min_ess_mixing']))
  del psamples, f32
  torch.cuda.empty_cache()

  value = es_sur['min_ess_mixing'] / incl_s
  sur_rate = es_sur['min_ess_mixing'] / sur_s
  plain_rate = es_plain['min_ess_mixing'] / plain_s
  sur_rate_kw = es_sur_kw['min_ess_mixing'] / (sur_s * Kw / K)
  plain_s_per_t = plain_s / Kp
  # the plain method's own warm-up from N(0, I), priced at its measured
  # seconds per L-step transition over the gradient evaluations the bf16
  # warm-up took (descent transitions of --burn-in-steps, then n_rec of Lr)
  warm_equiv = (n_descent * args.burn_in_steps + n_rec * Lr) / float(L)
  plain_incl_derived = es_plain['min_ess_mixing'] / (plain_s + warm_equiv * plain_s_per_t)

  # roofline of the dominant kernel of the surrogate transitions (dominance
  # from the all-class breakdown; its launch time from the timed region)
  kinds = {k: v for k, v in breakdown.items() if v[1] > 0}
  dom = max(kinds, key=lambda k: kinds[k][0])
  ms, cnt = prof[dom] if prof[dom][1] > 0 else kinds[dom]
  issued = None
  if dom == 'gemm_ll':
    # algorithmic work of the exact forward: eta = W X^T, 2 N D C flop; the
    # kernel issues twice that on the matrix cores (hi and lo bf16 planes of W)
    flop = 2.0 * N * D * C
    issued = 2.0 * flop
  elif dom == 'mlp':
    H = args.hidden
    flop = 2.0 * C * (D * H + H * H + H * D) / 3.0   # mean over the 3 layer GEMMs
  else:
    flop = 2.0 * N * D * C
  achieved = flop / (ms / cnt * 1e-3)
  roof = {'bound': 'mfma', 'kernel': dom, 'achieved': achieved / 1e12,
          'peak': PEAK['bf16'] / 1e12, 'unit': 'TFLOP/s', 'frac': achieved / PEAK['bf16'],
          'traffic': measured_traffic(dom, 'large'), 'avg_launch_ms': ms / cnt,
          'flop_per_launch': flop,
          'mfma_issued_frac': None if issued is None else issued / (ms / cnt * 1e-3) / PEAK['bf16'],
          'per_transition_ms': {k: v[0] / n_break for k, v in kinds.items()},
          'per_transition_ms_from': '%d untimed transitions after the timed region, every '
                                    'kernel class bracketed by events' % n_break}

  extra = {}
  if not args.no_extra:
    if args.init != 'laplace':
      log('laplace-start variant')
      extra['laplace_start'] = laplace_variant(bf, wl, args, K)
      log('laplace-start variant: %s' % extra['laplace_start'])
    del bf
    torch.cuda.empty_cache()
    log('small configuration')
    extra['small_c1'] = small_config(args)
    log('mid configuration')
    extra['mid_c2'] = mid_config(args)
    log('hierarchical configuration')
    extra['hier_c4'] = hier_config(args)

  cpu = None
  if not args.no_cpu:
    cpu = cpu_baseline(wl, args.cpu_seconds, es_plain['min_ess_mixing'] / Kp,
                       chains=args.cpu_chains)
    log('cpu baseline: %s' % cpu['sample'])

  strip = lambda e: {k: v for k, v in e.items() if k != 'per_coord'}
  out = {
      'metric': 'min-ESS/sec, logreg HMC N=1M D=512 4096 chains: surrogate+BF16 vs plain FP32',
      'value': value,
      'unit': 'min-ESS/s',
      'n_gpus': 1,
      'steps': K,
      'warmup': args.warmup,
      'ms_per_step': incl_s / K * 1e3,
      'timed_region': ('Laplace start (device Newton steps, chains drawn from N(mode, P^-1)) '
                       '+ %d exact gradients at Laplace draws as training pairs' % n_pairs
                       if pairs_from_laplace else
                       '%swarm-up from %s (%d plain bf16 transitions of L=%d, then %d of '
                       'L=%d with exact gradients recording %d pairs)'
                       % ('Laplace start (device Newton steps) + ' if laplace_info else '',
                          'N(mode, P^-1) draws' if laplace_info else 'N(0, I)', n_descent,
                          args.burn_in_steps, n_rec, Lr, n_pairs)) +
                      ' + surrogate training + %d surrogate transitions: value = mixing '
                      'min-ESS of the %d samples / all of it' % (K, K),
      'warm_start': laplace_info if laplace_info else 'N(0, I) draws, plain bf16 warm-up',
      'higher_is_better': True,
      'scaling': 'weak',
      'vs_baseline': None,
      'dtype': 'bf16',
      'data': 'synthetic',
      'config': {'workload': 'logreg_hmc N=%d D=%d chains=%d eps=%g (x U(%.2f,%.2f) per '
                             'transition) L=%d X=bf16 surrogate MLP %d-%d-%d-%d+skip'
                             % (N, D, C, eps, 1 - args.jitter, 1 + args.jitter, L, D,
                                args.hidden, args.hidden, D),
                 'global_batch': C, 'seq_len': None, 'parallelism': 'none (1 GPU)'},
      'ess_estimator': 'multi-chain (Gelman et al. BDA3 11.5, Geyer initial monotone '
                       'sequence) per coordinate, minus the floor M n / (2n - 1) that '
                       'chains which never move score; min over the D coordinates',
      'surrogate_bf16': dict(strip(es_sur), min_ess_mixing_per_s_inclusive=value,
                             min_ess_mixing_per_s_sampling_only=sur_rate, transitions=K,
                             seconds_inclusive=incl_s, sampling_s=sur_s, accept=acc_sur,
                             ms_per_transition=sur_s / K * 1e3, warmup_s=warm_s,
                             warmup_transitions=n_warm, train_s=train_s, train_pairs=n_pairs,
                             final_loss=losses[-1],
                             **{'min_ess_mixing_first_%d' % Kw: es_sur_kw['min_ess_mixing']}),
      'plain_fp32': dict(strip(es_plain), min_ess_mixing_per_s=plain_rate, transitions=Kp,
                         seconds=plain_s, accept=acc_plain,
                         min_ess_mixing_per_s_with_warmup_derived=plain_incl_derived,
                         warmup_derived=('its own warm-up priced as the bf16 warm-up\'s '
                                         'gradient evaluations, %.1f transitions of L=%d x '
                                         '%.2f s' % (warm_equiv, L, plain_s_per_t))),
      'speedup_vs_plain_fp32': value / plain_rate if plain_rate > 0 else None,
      'speedup_vs_plain_fp32_with_its_warmup': (value / plain_incl_derived
                                                if plain_incl_derived > 0 else None),
      'speedup_sampling_equal_window': sur_rate_kw / plain_rate if plain_rate > 0 else None,
      'speedup_note': ('value (warm-up + training + sampling) against the plain fp32 '
                       'sampling rate from equilibrium states (its warm-up not charged); '
                       'with its warm-up priced in; and sampling only over the first %d '
                       'transitions of each' % Kw),
      'posterior_moments_vs_fp64_oracle': moments,
      'roofline': roof,
      'cpu_baseline': None if cpu is None else {k: cpu[k] for k in
                                                ('value', 'unit', 'cores', 'kind', 'sample',
                                                 'cpu_model')},
  }
  out.update(extra)
  print(json.dumps(out))


if __name__ == '__main__':
  main()
