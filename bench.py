"""Benchmark: min-ESS/sec of Bayesian logistic-regression HMC on one MI355X,
surrogate-driven bf16 HMC versus plain fp32 HMC.

Workload (default): N = 1e6 rows, D = 512 features (X ~ N(0,1) with column
scales U(0.5, 2), stored bf16), 4096 chains from N(0, I), eps = 5e-4,
L = 32 leapfrog steps -- the large logistic-regression configuration.

Flow (one JSON line on stdout at the end; progress on stderr):
  0. untimed: W plain bf16 transitions (kernel load, first touch), one draw
     through torch's Generator and matmul and one epoch of a throwaway
     surrogate on random pairs (the one-time loading of the library code the
     timed steps call), then the chains are reset to their N(0, I) start;
  timed region (everything the surrogate method costs):
  1. warm start (default --init laplace --pairs laplace): damped Newton
     steps on the device find the mode and precision P of the log joint; the
     chains start from draws of N(mode, P^-1), and --n-pairs exact gradients
     at further draws are the surrogate's training pairs.  (--init prior
     --pairs trajectory: plain bf16 HMC burn-in from N(0, I) until the mean
     log joint settles, pairs recorded along its last trajectories);
  2. surrogate training on the GPU (own fwd/bwd/Adam kernels, 5% held-out
     pairs);
  3. K surrogate transitions (exact bf16/FP64 log joint in every Metropolis
     test), samples kept on the device;
  `value` = min-ESS of the K samples (multi-chain, min over the D
  coordinates) / seconds of 1-3.
  4. comparator: Kp plain fp32 transitions from the same equilibrium states
     (sampling time only, which favours the comparator);
  5. the CPU restatement timed on a bounded sample on the host cores.
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
  ap.add_argument('--burn-in', type=int, default=40,
                  help='max plain bf16 burn-in transitions (timed, part of the method)')
  ap.add_argument('--rec-start-tol', type=float, default=300.0,
                  help='pair recording starts when the mean log joint moves less than this '
                       'over 5 burn-in transitions')
  ap.add_argument('--init', default='laplace', choices=['laplace', 'prior'],
                  help='chain start: draws from the Laplace approximation N(mode, P^-1) found '
                       'by device Newton steps (timed as warm-up), or the workload\'s N(0, I) '
                       'draws')
  ap.add_argument('--laplace-burn-in', type=int, default=2,
                  help='transitions before pair recording starts after a Laplace start '
                       '(with --pairs trajectory)')
  ap.add_argument('--pairs', default='laplace', choices=['laplace', 'trajectory'],
                  help='surrogate training pairs: exact gradients at Laplace draws (with '
                       '--init laplace; no burn-in transitions), or recorded along burn-in '
                       'trajectories')
  ap.add_argument('--n-pairs', type=int, default=114688,
                  help='training pairs drawn from the Laplace approximation')
  ap.add_argument('--burn-in-adapt', type=int, default=1,
                  help='1: the burn-in step-size centre follows the acceptance rate')
  ap.add_argument('--plain-steps', type=int, default=8, help='timed plain fp32 transitions')
  ap.add_argument('--workload', default='large',
                  choices=['large', 'mid', 'small', 'hier', 'copy', 'forward'],
                  help="'hier': the hierarchical logistic regression (G=2048 groups of "
                       "1+Poisson(400) rows, D=64, 512 chains, L=16), its own JSON line; "
                       "'copy': 1 GiB device-to-device copies, the byte-counter calibration "
                       "of HBM traffic profiles; 'forward': --steps exact forward passes "
                       "(the gemm_ll kernel alone) on the large workload, for profiling")
  ap.add_argument('--hidden', type=int, default=1024)
  ap.add_argument('--epochs', type=int, default=32)
  ap.add_argument('--lr', type=float, default=3e-3)
  ap.add_argument('--rec-stride', type=int, default=4)
  ap.add_argument('--rec-transitions', type=int, default=8)
  ap.add_argument('--surrogate-only', action='store_true',
                  help='stop after the surrogate phase (tuning runs)')
  ap.add_argument('--seed', type=int, default=0)
  ap.add_argument('--jitter', type=float, default=0.2,
                  help='step size drawn from eps * U(1-j, 1+j) per transition (both samplers)')
  ap.add_argument('--no-cpu', action='store_true')
  ap.add_argument('--cpu-seconds', type=float, default=15.0)
  return ap.parse_args()


def make_workload(name):
  from edward_amd import workloads as W
  if name == 'large':
    return W.logreg_large()
  if name == 'mid':
    wl = W.logreg(100000, 256, 1024, 0.002, 20, 2000, bf16=True)
    return wl
  wl = W.logreg(20000, 64, 512, 0.005, 16, 1000, bf16=True)
  return wl


def cpu_baseline(wl, budget_s, ess_per_transition_plain):
  """The oracle's fp32 restatement of one plain transition's work: L
  gradient evaluations (eta = W X^T, R = y - sigmoid(eta), G = R X), timed
  on a sample of chains at the full N on the host cores (OpenBLAS)."""
  from oracle import hmc_oracle as O  # the CPU restatement (checker) only
  cores = len(os.sched_getaffinity(0))
  X = wl.X
  y = wl.y
  Cs = 16
  W = wl.W0[:Cs].astype(np.float32)
  reps, t = 0, 0.0
  while t < budget_s and reps < 50:
    t0 = time.perf_counter()
    eta = W @ X.T
    R = (y[None, :] - O.sigmoid(eta)).astype(np.float32)
    G = R @ X
    t += time.perf_counter() - t0
    reps += 1
  sec_per_eval = t / reps * (wl.C / Cs)
  sec_per_transition = sec_per_eval * wl.n_steps
  return {
      'value': ess_per_transition_plain / sec_per_transition,
      'unit': 'min-ESS/s (derived: GPU fp32 ESS per transition / CPU sec per transition)',
      'cores': cores,
      'kind': 'port',
      'sample': ('%d gradient evaluations of %d of %d chains at N=%d, D=%d in fp32 numpy '
                 '(OpenBLAS, %d threads); scaled x%d chains x L=%d -> %.1f s per transition'
                 % (reps, Cs, wl.C, wl.N, wl.D, cores, wl.C // Cs, wl.n_steps,
                    sec_per_transition)),
      'sec_per_transition': sec_per_transition,
  }


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
  import torch
  from edward_amd import diagnostics
  from edward_amd.engine import GLMEngine
  from edward_amd.surrogate import Surrogate
  from edward_amd.engine import jittered_step

  wl = make_workload(args.workload)
  N, D, C, eps, L = wl.N, wl.D, wl.C, wl.step_size, wl.n_steps
  log('workload N=%d D=%d C=%d eps=%g L=%d' % (N, D, C, eps, L))
  dev = torch.device('cuda', 0)
  bf = GLMEngine(wl.X, wl.y, wl.prior_loc, wl.prior_scale, precision='bf16')
  bf.set_state(wl.W0)
  n_acc = torch.zeros(C, dtype=torch.int32, device=dev)
  for t in range(args.warmup):     # untimed
    bf.transition(eps, L, seed=args.seed + 99, counter=t, n_accept=n_acc)
  bf.set_state(wl.W0)
  # the Laplace draws go through a torch Generator and an fp32 matmul: one
  # unrelated draw here keeps the libraries' one-time initialisation (~0.28 s
  # on first use in a process) out of the timed region, as the transitions
  # above do for the engine's kernels
  gen = torch.Generator(device=dev)
  gen.manual_seed(args.seed + 98)
  (torch.randn(C, D, generator=gen, device=dev) @ torch.eye(D, device=dev)).sum().item()
  # likewise one epoch of a throwaway surrogate on random pairs (enough that
  # the skip's least squares takes its full row sample): the first use of the
  # training path's kernels and library calls costs ~0.3 s per process
  nw = 9 * C
  wsur = Surrogate(bf, hidden=args.hidden, max_batch=C, seed=args.seed + 97)
  wsur.fit(torch.randn(nw, D, generator=gen, device=dev),
           torch.randn(nw, D, generator=gen, device=dev), epochs=1, batch=C, lr=args.lr,
           lr_final=5e-5, seed=args.seed + 97, val_frac=0.05)
  del wsur
  torch.cuda.synchronize()

  # timed region: warm start + burn-in + pair recording + training + K sampling
  # transitions
  # 0. warm start: the Laplace approximation by device Newton steps, chains
  #    drawn from N(mode, P^-1)
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
  # 1. burn-in with plain bf16 HMC, then record (w, grad) pairs
  means = []
  # training pairs: every rec_stride-th inner leapfrog point of n_rec
  # transitions (~250k pairs from more independent trajectories than a
  # dense record of fewer transitions)
  rec_stride = args.rec_stride
  n_rec = args.rec_transitions
  if not pairs_from_laplace:
    cap = ((L - 1) // rec_stride) * C * n_rec
    Wrec = torch.empty(cap, D, device=dev)
    Grec = torch.empty(cap, D, device=dev)
  t = 0
  converged_at = None
  jit = np.random.default_rng(args.seed + 7)
  log_c = math.log(eps)
  recording, n_recorded = False, 0
  while not pairs_from_laplace and t < burn_in + n_rec:
    # Pair recording overlaps the end of the burn-in: it starts once the mean
    # log joint moves by less than --rec-start-tol over 5 transitions (near
    # equilibrium), or after --burn-in transitions, and records n_rec
    # transitions; the burn-in ends when the mean moves by less than 20 over 5
    # transitions (or --burn-in + n_rec transitions ran) and the recording is
    # full.
    if not recording and (t >= burn_in or
                          (t >= 10 and abs(means[-1] - means[-6]) < args.rec_start_tol)):
      bf.record_pairs(Wrec, Grec, stride=rec_stride)
      recording = True
    # the step size is jittered log-uniformly over [c/2, 2c] during burn-in:
    # with a fixed eps L, posterior directions whose frequency makes eps L a
    # multiple of half a period keep their energy through every momentum
    # refresh and approach the typical set very slowly.  With --burn-in-adapt
    # the centre c follows the acceptance rate (up while the chains still
    # fall toward the mode and accept everything, down near equilibrium),
    # within [eps/4, 4 eps]; the kept samples use eps itself.  Recording
    # transitions use the sampling step size.
    if recording:
      eps_t = jittered_step(eps, args.jitter, args.seed, t)
    else:
      eps_t = math.exp(log_c) * 2.0 ** jit.uniform(-1.0, 1.0)
    n_acc.zero_()
    bf.transition(eps_t, L, seed=args.seed, counter=t, n_accept=n_acc)
    _, lp, _ = bf.get_state()
    means.append(lp.mean().item())
    acc_t = n_acc.float().mean().item()
    log('warm-up %d eps %.2e mean logp %.2f accept %.3f%s'
        % (t, eps_t, means[-1], acc_t, ' (recording)' if recording else ''))
    if args.burn_in_adapt and not recording:
      log_c = min(max(log_c + 0.6 * (acc_t - 0.7), math.log(eps / 4)), math.log(4 * eps))
    t += 1
    if recording:
      n_recorded += 1
    if converged_at is None and t >= 10 and abs(means[-1] - means[-6]) < 20.0:
      converged_at = t
    if n_recorded >= n_rec and (converged_at is not None or t >= burn_in + n_rec):
      break
  torch.cuda.synchronize()
  if pairs_from_laplace:
    n_pairs, n_rec = args.n_pairs, 0
  else:
    bf.record_pairs(None)
    n_pairs = bf.record_count()
  n_warm = t
  warm_s = time.time() - t_warm0
  log('warm-up done: %d transitions, %d pairs, %.1f s' % (t, n_pairs, warm_s))

  # 2. train the surrogate
  t_tr0 = time.time()
  sur = Surrogate(bf, hidden=args.hidden, max_batch=C, seed=args.seed)
  losses = sur.fit(Wrec[:n_pairs], Grec[:n_pairs], epochs=args.epochs, batch=C, lr=args.lr,
                   lr_final=5e-5, seed=args.seed, log=log, val_frac=0.05)
  torch.cuda.synchronize()
  train_s = time.time() - t_tr0
  del Wrec, Grec
  log('training done: %.1f s, loss %.4f -> %.4f' % (train_s, losses[0], losses[-1]))

  # 3. timed surrogate sampling
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
  sur_s = time.time() - ts0
  prof = bf.profile_read()
  bf.profile(False)
  acc_sur = n_acc.float().mean().item() / K
  ess_sur = diagnostics.ess(bf, samples)
  mess_sur = float(ess_sur.min())
  mess_sur_within = float(diagnostics.ess(bf, samples, between=False).min())
  Kp = args.plain_steps
  mess_sur_kp = float(diagnostics.ess(bf, samples[:Kp]).min())
  log('surrogate: %d transitions %.2f s, accept %.3f, min-ESS %.1f (median %.1f)'
      % (K, sur_s, acc_sur, mess_sur, float(np.median(ess_sur))))
  state, _, _ = bf.get_state()
  # per-class kernel times from a short untimed run with every class timed
  n_break = 5
  bf.profile(True)
  for k in range(n_break):
    sur.transition(jittered_step(eps, args.jitter, args.seed, t + k), L, seed=args.seed,
                   counter=t + k)
  breakdown = bf.profile_read()
  bf.profile(False)
  sur_mean = samples.double().mean((0, 1)).cpu().numpy()
  sur_var = samples.double().var((0, 1)).cpu().numpy()
  del samples
  torch.cuda.empty_cache()

  if args.surrogate_only:
    print(json.dumps({'surrogate_min_ess_per_s': mess_sur / sur_s, 'accept': acc_sur,
                      'final_loss': losses[-1], 'ms_per_step': sur_s / K * 1e3}))
    return

  # 4. timed plain fp32 from the same equilibrium states
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
  acc_plain = n_acc.float().mean().item() / Kp
  ess_plain = diagnostics.ess(f32, psamples)
  mess_plain = float(ess_plain.min())
  mess_plain_within = float(diagnostics.ess(f32, psamples, between=False).min())
  log('plain fp32: %d transitions %.2f s, accept %.3f, min-ESS %.1f'
      % (Kp, plain_s, acc_plain, mess_plain))
  plain_mean = psamples.double().mean((0, 1)).cpu().numpy()
  del psamples, f32
  torch.cuda.empty_cache()

  sur_rate = mess_sur / sur_s
  plain_rate = mess_plain / plain_s
  # equal window: the surrogate's first Kp transitions against the plain Kp
  sur_rate_kp = mess_sur_kp / (sur_s * Kp / K)
  incl_s = sur_s + warm_s + train_s
  incl_rate = mess_sur / incl_s

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
          'traffic': measured_traffic(dom, args.workload), 'avg_launch_ms': ms / cnt,
          'flop_per_launch': flop,
          'mfma_issued_frac': None if issued is None else issued / (ms / cnt * 1e-3) / PEAK['bf16'],
          'per_transition_ms': {k: v[0] / n_break for k, v in kinds.items()},
          'per_transition_ms_from': '%d untimed transitions after the timed region, every '
                                    'kernel class bracketed by events' % n_break}

  cpu = None
  if not args.no_cpu:
    cpu = cpu_baseline(wl, args.cpu_seconds, mess_plain / Kp)
    log('cpu baseline: %s' % cpu['sample'])

  out = {
      'metric': 'min-ESS/sec, logreg HMC N=1M D=512 4096 chains: surrogate+BF16 vs plain FP32',
      'value': incl_rate,
      'unit': 'min-ESS/s',
      'n_gpus': 1,
      'steps': K,
      'warmup': args.warmup,
      'ms_per_step': incl_s / K * 1e3,
      'timed_region': ('Laplace start (device Newton steps, chains drawn from N(mode, P^-1)) '
                       '+ %d exact gradients at Laplace draws as training pairs' % n_pairs
                       if pairs_from_laplace else
                       '%sburn-in (%d plain bf16 transitions, pairs recorded over the last %d)'
                       % ('Laplace start (device Newton steps) + ' if laplace_info else '',
                          n_warm, n_rec)) +
                      ' + surrogate training + %d surrogate transitions: value = min-ESS of the '
                      '%d samples / all of it' % (K, K),
      'warm_start': laplace_info if laplace_info else 'N(0, I) draws',
      'higher_is_better': True,
      'scaling': 'weak',
      'vs_baseline': None,
      'dtype': 'bf16',
      'data': 'synthetic',
      'config': {'workload': 'logreg_hmc N=%d D=%d chains=%d eps=%g (x U(%.2f,%.2f) per '
                             'transition) L=%d X=bf16 surrogate MLP %d-%d-%d-%d+skip'
                             % (N, D, C, eps, 1 - args.jitter, 1 + args.jitter, L, D, args.hidden,
                                                            args.hidden, D),
                 'global_batch': C, 'seq_len': None, 'parallelism': 'none (1 GPU)'},
      'ess_estimator': 'multi-chain (Gelman et al. BDA3 11.5, Geyer initial monotone '
                       'sequence) over the timed transitions, min over the D coordinates',
      'plain_fp32': {'min_ess_per_s': plain_rate, 'min_ess': mess_plain, 'transitions': Kp,
                     'seconds': plain_s, 'accept': acc_plain,
                     'min_ess_within_chain': mess_plain_within},
      'surrogate_bf16': {'min_ess_per_s': sur_rate, 'min_ess': mess_sur, 'transitions': K,
                         'seconds': sur_s, 'accept': acc_sur,
                         'min_ess_per_s_sampling_only': sur_rate,
                         'ms_per_transition': sur_s / K * 1e3,
                         'warmup_s': warm_s, 'train_s': train_s, 'train_pairs': n_pairs,
                         'final_loss': losses[-1], 'min_ess_within_chain': mess_sur_within,
                         'min_ess_first_%d' % Kp: mess_sur_kp},
      'speedup_vs_plain_fp32': incl_rate / plain_rate if plain_rate > 0 else None,
      'speedup_sampling_only': sur_rate / plain_rate if plain_rate > 0 else None,
      'speedup_equal_window': sur_rate_kp / plain_rate if plain_rate > 0 else None,
      'posterior_mean_max_abs_diff_sur_vs_plain': float(np.abs(sur_mean - plain_mean).max()),
      'posterior_sd_median': float(np.median(np.sqrt(sur_var))),
      'roofline': roof,
      'cpu_baseline': None if cpu is None else {k: cpu[k] for k in
                                                ('value', 'unit', 'cores', 'kind', 'sample')},
  }
  print(json.dumps(out))


if __name__ == '__main__':
  main()
