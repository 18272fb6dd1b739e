"""Device engines: thin owners of an edhmc context.

torch holds device memory, names the stream and copies buffers; the
arithmetic of the sampler, the warm start and the surrogate's training runs
in libedhmc.so's kernels (small D-sized steps -- a least-squares solve, a
Cholesky factor -- on the host in FP64).
"""

import ctypes

import numpy as np
import torch

from edward_amd import _lib as L


def _dev_tensor(x, dtype, device):
  if isinstance(x, torch.Tensor):
    return x.to(device=device, dtype=dtype).contiguous()
  a = np.asarray(x)
  if not a.flags.writeable:
    a = a.copy()      # broadcast views are read-only; torch wants its own buffer
  return torch.as_tensor(a, dtype=dtype, device=device).contiguous()


def jittered_step(step_size, jitter, seed, counter):
  """Step size of transition `counter`: step_size * U(1 - jitter, 1 + jitter),
  a deterministic function of (seed, counter).  A fixed eps L makes posterior
  directions whose leapfrog rotation over L steps is a multiple of 2 pi
  return to their start every transition; drawing eps afresh each transition
  (Neal 2011, "MCMC using Hamiltonian dynamics", sec. 5.4.2) removes that,
  and keeps the chain exact (the draw is independent of the state)."""
  if not jitter:
    return float(step_size)
  u = np.random.default_rng([int(seed) & 0xFFFFFFFF, int(counter)]).random()
  return float(step_size) * (1.0 + float(jitter) * (2.0 * u - 1.0))


def require_gpu():
  if not torch.cuda.is_available():
    raise RuntimeError('edward_amd needs a ROCm GPU (MI355X/gfx950); none is visible')


class Context(object):
  """An edhmc_ctx on one device."""

  def __init__(self, device=0):
    require_gpu()
    self.lib = L.load()
    self.device = torch.device('cuda', device)
    h = ctypes.c_void_p()
    L.check(self.lib.edhmc_create(device, ctypes.byref(h)))
    self.h = h

  def close(self):
    if getattr(self, 'h', None):
      torch.cuda.synchronize(self.device)
      self.lib.edhmc_destroy(self.h)
      self.h = None

  def __del__(self):
    try:
      self.close()
    except Exception:
      pass

  def check(self, rc):
    L.check(rc, self.h)

  PROFILE_KINDS = ('gemm_grad', 'gemm_ll', 'gemm_xtr', 'chain', 'mlp', 'hier')

  def profile(self, on=True, kinds=None):
    """Turn the library's per-launch hipEvent timing on (clears totals) or off.
    ``kinds``: the classes of ``PROFILE_KINDS`` to time (default all); each
    timed launch adds two event records to its stream."""
    if kinds is None:
      mask = 0xffffffff
    else:
      mask = 0
      for k in kinds:
        mask |= 1 << self.PROFILE_KINDS.index(k)
    self.check(self.lib.edhmc_profile_select(self.h, mask))
    self.check(self.lib.edhmc_profile_enable(self.h, 1 if on else 0))

  def profile_read(self):
    """{kind: (total_ms, launches)} summed over the launches since enabling."""
    ms = (ctypes.c_double * 6)()
    n = (ctypes.c_int64 * 6)()
    self.check(self.lib.edhmc_profile_read(self.h, ms, n))
    return {k: (ms[i], int(n[i])) for i, k in enumerate(self.PROFILE_KINDS)}

  @property
  def stream(self):
    return L.stream_ptr(torch.cuda.current_stream(self.device))


class GLMEngine(Context):
  """HMC over a GLM  y ~ Lik(X w),  w ~ Normal(loc, scale)  for C chains.

  precision: 'fp32' (fp32 X and MFMA) or 'bf16' (bf16 X, bf16 MFMA with fp32
  accumulation, hi/lo split positions for the log joint).
  """

  def __init__(self, X, y, prior_loc=0.0, prior_scale=1.0, likelihood='bernoulli',
               lik_scale=1.0, precision='fp32', device=0):
    super(GLMEngine, self).__init__(device)
    self.precision = precision
    self.C = None
    self.bind(X, y, prior_loc, prior_scale, likelihood, lik_scale)

  def bind(self, X, y, prior_loc=0.0, prior_scale=1.0, likelihood='bernoulli', lik_scale=1.0):
    """(Re)bind the data and prior; discards the chain state (a new
    minibatch of the stochastic-gradient samplers rebinds every update)."""
    dev = self.device
    if isinstance(X, torch.Tensor) and X.dtype == torch.bfloat16:
      Xd, xin = X.to(dev).contiguous(), L.BF16
    else:
      Xd, xin = _dev_tensor(X, torch.float32, dev), L.F32
    if Xd.dim() != 2:
      raise ValueError('X must be 2-D, got shape %s' % (tuple(Xd.shape),))
    N, D = Xd.shape
    yd = _dev_tensor(y, torch.float32, dev).reshape(-1)
    if yd.numel() != N:
      raise ValueError('y has %d entries for %d rows of X' % (yd.numel(), N))
    loc = _dev_tensor(np.broadcast_to(np.asarray(
        prior_loc.cpu() if isinstance(prior_loc, torch.Tensor) else prior_loc,
        np.float32), (D,)), torch.float32, dev)
    scale = _dev_tensor(np.broadcast_to(np.asarray(
        prior_scale.cpu() if isinstance(prior_scale, torch.Tensor) else prior_scale,
        np.float32), (D,)), torch.float32, dev)
    lik = {'bernoulli': L.LIK_BERNOULLI, 'normal': L.LIK_NORMAL}[likelihood]
    cdt = {'fp32': L.F32, 'bf16': L.BF16}[self.precision]
    self.N, self.D = int(N), int(D)
    self.likelihood = likelihood
    self.C = None
    self.check(self.lib.edhmc_bind_glm(self.h, L.ptr(Xd), xin, L.ptr(yd), N, D, lik,
                                       float(lik_scale), L.ptr(loc), L.ptr(scale), cdt,
                                       self.stream))
    torch.cuda.synchronize(dev)

  def set_prior_family(self, kind, lnorm=None):
    """Per-coordinate constrained prior families and operand maps (kind
    byte = family | operand << 4, lnorm = normalising constants, see
    ``csrc/transform.h``); ``kind=None`` restores Normal priors."""
    if kind is None:
      self._prior_family = None
      self.check(self.lib.edhmc_set_prior_family(self.h, None, None, self.stream))
      return
    k = _dev_tensor(np.asarray(kind, np.uint8).reshape(-1), torch.uint8, self.device)
    ln = _dev_tensor(np.asarray(lnorm, np.float64).reshape(-1), torch.float64, self.device)
    if k.numel() != self.D or ln.numel() != self.D:
      raise ValueError('prior family needs %d coordinates' % self.D)
    self._prior_family = (k, ln)
    self.check(self.lib.edhmc_set_prior_family(self.h, L.ptr(k), L.ptr(ln), self.stream))

  def newton_terms(self, w):
    """(log p, gradient [D], negative Hessian [D, D]) of the log joint at one
    position w [D], in float64 (``edhmc_newton_terms``)."""
    w = _dev_tensor(np.asarray(w.cpu() if isinstance(w, torch.Tensor) else w,
                               np.float32).reshape(-1), torch.float32, self.device)
    if w.numel() != self.D:
      raise ValueError('position has %d coordinates, model has %d' % (w.numel(), self.D))
    lp = torch.empty(1, dtype=torch.float64, device=self.device)
    g = torch.empty(self.D, dtype=torch.float64, device=self.device)
    P = torch.empty(self.D, self.D, dtype=torch.float64, device=self.device)
    self.check(self.lib.edhmc_newton_terms(self.h, L.ptr(w), L.ptr(lp), L.ptr(g), L.ptr(P),
                                           self.stream))
    return float(lp.item()), g.cpu().numpy(), P.cpu().numpy()

  def laplace(self, w0=None, max_iter=30, tol=1e-6):
    """Mode of the log joint by damped Newton steps from w0 (default 0) and
    the precision matrix there: the Laplace approximation N(mode, P^-1)
    (``edward/inferences/laplace.py:17-137`` fits the same Gaussian).  A
    step is halved until the log joint does not decrease; the loop stops
    when the Newton decrement g^T P^-1 g / 2 falls below ``tol``.  Returns
    (mode, P, logp at the mode, iterations)."""
    w = np.zeros(self.D) if w0 is None else np.asarray(w0, np.float64).reshape(-1).copy()
    lp, g, P = self.newton_terms(w)
    it = 0
    for it in range(1, max_iter + 1):
      step = np.linalg.solve(P, g)
      dec = 0.5 * float(g @ step)
      if dec < tol:
        break
      t = 1.0
      while True:
        wn = w + t * step
        lpn, gn, Pn = self.newton_terms(wn)
        if lpn >= lp - 1e-9 * abs(lp) or t < 1e-6:
          break
        t *= 0.5
      w, lp, g, P = wn, lpn, gn, Pn
    return w, P, lp, it

  def set_inner_gradient(self, exact):
    """bf16: inner leapfrog gradients on both position planes (exact=True)
    or on the bf16-rounded positions (False, the default)."""
    self.check(self.lib.edhmc_set_inner_gradient(self.h, 1 if exact else 0))

  def set_lik_weight(self, w):
    """Multiplier of the likelihood in the log joint (Edward's ``scale`` of
    the observed variable, used by the stochastic-gradient samplers)."""
    self.check(self.lib.edhmc_set_lik_weight(self.h, float(w)))

  def sgld_step(self, learning_rate, seed=0, counter=0, sample_out=None, n_accept=None):
    self.check(self.lib.edhmc_sgld_step(self.h, float(learning_rate), int(seed), int(counter),
                                        L.ptr(sample_out), L.ptr(n_accept), self.stream))

  def mh_step(self, proposal_scale, uniforms=None, seed=0, counter=0, sample_out=None,
              n_accept=None, accepted_out=None):
    """Random-walk Metropolis-Hastings step with proposal N(z, diag(s^2))."""
    s = _dev_tensor(np.broadcast_to(np.asarray(
        proposal_scale.cpu() if isinstance(proposal_scale, torch.Tensor) else proposal_scale,
        np.float32), (self.D,)), torch.float32, self.device)
    self._mh_scale = s
    u = None if uniforms is None else _dev_tensor(uniforms, torch.float64, self.device)
    self.check(self.lib.edhmc_mh_step(self.h, L.ptr(s), L.ptr(u), int(seed), int(counter),
                                      L.ptr(accepted_out), L.ptr(sample_out), L.ptr(n_accept),
                                      self.stream))

  def replica_exchange_step(self, proposal_scale, beta, n_replica, exchange_freq, seed=0,
                            counter=0, sample_out=None, n_accept=None, accepted_out=None,
                            swaps=None):
    """Tempered random-walk MH on every chain, then adjacent-replica exchanges
    (chains ordered replica-major: chain = replica * ladders + ladder)."""
    s = _dev_tensor(np.broadcast_to(np.asarray(
        proposal_scale.cpu() if isinstance(proposal_scale, torch.Tensor) else proposal_scale,
        np.float32), (self.D,)), torch.float32, self.device)
    b = _dev_tensor(beta, torch.float32, self.device)
    self._re_keep = (s, b)
    self.check(self.lib.edhmc_replica_exchange_step(
        self.h, L.ptr(s), L.ptr(b), int(n_replica), float(exchange_freq), int(seed),
        int(counter), L.ptr(accepted_out), L.ptr(sample_out), L.ptr(n_accept), L.ptr(swaps),
        self.stream))

  def sghmc_step(self, learning_rate, friction, v, seed=0, counter=0, sample_out=None,
                 n_accept=None):
    self.check(self.lib.edhmc_sghmc_step(self.h, float(learning_rate), float(friction), L.ptr(v),
                                         int(seed), int(counter), L.ptr(sample_out),
                                         L.ptr(n_accept), self.stream))

  # -- state ---------------------------------------------------------------
  def set_state(self, W):
    W = _dev_tensor(W, torch.float32, self.device)
    if W.dim() == 1:
      W = W[None]
    if W.shape[1] != self.D:
      raise ValueError('positions have %d coordinates, model has %d' % (W.shape[1], self.D))
    self.C = int(W.shape[0])
    self.check(self.lib.edhmc_set_state(self.h, L.ptr(W), self.C, self.stream))

  def get_state(self):
    W = torch.empty(self.C, self.D, device=self.device)
    lp = torch.empty(self.C, dtype=torch.float64, device=self.device)
    g = torch.empty(self.C, self.D, device=self.device)
    self.check(self.lib.edhmc_get_state(self.h, L.ptr(W), L.ptr(lp), L.ptr(g), self.stream))
    return W, lp, g

  def logp_grad(self, W):
    W = _dev_tensor(W, torch.float32, self.device)
    if W.dim() == 1:
      W = W[None]
    C = int(W.shape[0])
    g = torch.empty(C, self.D, device=self.device)
    lp = torch.empty(C, dtype=torch.float64, device=self.device)
    self.C = C
    self.check(self.lib.edhmc_logp_grad(self.h, L.ptr(W), C, L.ptr(g), L.ptr(lp), self.stream))
    return lp, g

  def logp(self, W):
    """Log joint only (the forward pass of the Metropolis test) at W [C, D]."""
    W = _dev_tensor(W, torch.float32, self.device)
    if W.dim() == 1:
      W = W[None]
    C = int(W.shape[0])
    lp = torch.empty(C, dtype=torch.float64, device=self.device)
    if self.C != C:
      self.C = C
    self.check(self.lib.edhmc_logp(self.h, L.ptr(W), C, L.ptr(lp), self.stream))
    return lp

  def predict(self, W, mean=True, loglik=True, replicate=False, seed=0):
    """Posterior-predictive pass over the bound rows for draws W [S, D]:
    (mean [N] of E[y | w_s] over the draws, loglik [S] fp64, y_rep [S, N]),
    entries None where not asked for."""
    W = _dev_tensor(W, torch.float32, self.device)
    if W.dim() == 1:
      W = W[None]
    S = int(W.shape[0])
    m = torch.empty(self.N, device=self.device) if mean else None
    ll = torch.empty(S, dtype=torch.float64, device=self.device) if loglik else None
    rep = torch.empty(S, self.N, device=self.device) if replicate else None
    self.C = S
    self.check(self.lib.edhmc_predict(self.h, L.ptr(W), S, L.ptr(m), L.ptr(ll), L.ptr(rep),
                                      int(seed), self.stream))
    return m, ll, rep

  # -- sampling ------------------------------------------------------------
  def transition(self, step_size, n_steps, momenta=None, uniforms=None, seed=0, counter=0,
                 sample_out=None, n_accept=None, accepted_out=None):
    r0 = None if momenta is None else _dev_tensor(momenta, torch.float32, self.device)
    u = None if uniforms is None else _dev_tensor(uniforms, torch.float64, self.device)
    self.check(self.lib.edhmc_transition(
        self.h, float(step_size), int(n_steps), L.ptr(r0), L.ptr(u), int(seed), int(counter),
        L.ptr(accepted_out), L.ptr(sample_out), L.ptr(n_accept), self.stream))

  def run(self, params, step_size, n_steps, t0=0, n_iter=None, seed=0, counter0=0,
          n_accept=None):
    T = int(params.shape[0])
    if n_iter is None:
      n_iter = T - t0
    self.check(self.lib.edhmc_run(self.h, float(step_size), int(n_steps), L.ptr(params), T,
                                  int(t0), int(n_iter), int(seed), int(counter0),
                                  L.ptr(n_accept), self.stream))

  def moments_enable(self, on=True):
    self.check(self.lib.edhmc_moments_enable(self.h, 1 if on else 0))

  def moments(self):
    mean = torch.empty(self.C, self.D, dtype=torch.float64, device=self.device)
    m2 = torch.empty(self.C, self.D, dtype=torch.float64, device=self.device)
    n = ctypes.c_int64()
    self.check(self.lib.edhmc_moments_get(self.h, L.ptr(mean), L.ptr(m2), ctypes.byref(n),
                                          self.stream))
    return mean, m2, int(n.value)

  def divergences(self):
    out = torch.empty(self.C, dtype=torch.int32, device=self.device)
    self.check(self.lib.edhmc_divergences(self.h, L.ptr(out), self.stream))
    return out

  def record_pairs(self, W_buf=None, G_buf=None, stride=1):
    """Record (position, exact gradient) at every `stride`-th inner leapfrog
    point of the next plain transitions into W_buf/G_buf [capacity, D];
    None stops."""
    cap = 0 if W_buf is None else int(W_buf.shape[0])
    self._rec = (W_buf, G_buf)   # keep the buffers alive while recording
    self.check(self.lib.edhmc_record_pairs(self.h, L.ptr(W_buf), L.ptr(G_buf), cap,
                                           int(stride)))

  def record_count(self):
    return int(self.lib.edhmc_record_count(self.h))

  def col_stats(self, X):
    """Column mean and standard deviation (divisor P) of X [P, D] on device."""
    X = X.contiguous()
    P, D = X.shape
    mean = torch.empty(D, device=self.device)
    sd = torch.empty(D, device=self.device)
    self.check(self.lib.edhmc_col_stats(self.h, L.ptr(X), int(P), int(D), L.ptr(mean), L.ptr(sd),
                                        self.stream))
    return mean, sd

  def last_ratio(self):
    """FP64 log ratio K(r0) - K(rL) + log p(zL) - log p(z0) of every
    chain's last transition (plain or surrogate), [C] device tensor."""
    out = torch.empty(self.C, dtype=torch.float64, device=self.device)
    self.check(self.lib.edhmc_last_ratio(self.h, L.ptr(out), self.stream))
    return out

  def last_momentum(self):
    """Momenta at the end of every chain's last trajectory, [C, D]."""
    out = torch.empty(self.C, self.D, device=self.device)
    self.check(self.lib.edhmc_get_momentum(self.h, L.ptr(out), self.stream))
    return out

  def gather_rows(self, src, idx):
    """src [P, D] rows idx (device int64) -> a new [len(idx), D] tensor."""
    src = src.contiguous()
    P, D = src.shape
    out = torch.empty(int(idx.numel()), D, device=self.device)
    self.check(self.lib.edhmc_gather_rows(self.h, L.ptr(src), int(P), int(D), L.ptr(idx),
                                          int(idx.numel()), L.ptr(out), self.stream))
    return out

  def affine_rows(self, E, R, shift):
    """shift + E R for E [n, D], R [D, D], shift [D] (device fp32)."""
    n, D = E.shape
    out = torch.empty(n, D, device=self.device)
    self.check(self.lib.edhmc_affine_rows(self.h, L.ptr(E), int(n), int(D), L.ptr(R),
                                          L.ptr(shift), L.ptr(out), self.stream))
    return out

  def draw_momenta(self, C, seed, counter):
    out = torch.empty(C, self.D, device=self.device)
    self.check(self.lib.edhmc_draw_momenta(self.h, C, self.D, int(seed), int(counter),
                                           L.ptr(out), self.stream))
    return out

  def draw_uniforms(self, C, seed, counter):
    out = torch.empty(C, dtype=torch.float64, device=self.device)
    self.check(self.lib.edhmc_draw_uniforms(self.h, C, int(seed), int(counter), L.ptr(out),
                                            self.stream))
    return out
