"""Inference -> MonteCarlo -> HMC with Edward 1.3.5's control flow.

The class hierarchy, constructor checks, ``initialize``/``update``/``run``/
``print_progress``/``finalize`` and the ``t``/``n_accept``/``reset`` state
follow ``edward/inferences/inference.py:54-338``,
``monte_carlo.py:61-170`` and ``hmc.py:42-130``.  ``HMC.build_update`` does
not build a graph: it compiles the model to a device engine, and ``update``
runs one fused transition of every chain through the C ABI.

Deliberate differences (each has a test in tests/test_surface.py):
  * ``initialize`` returns ``self``, so ``HMC(...).initialize(eps, L).run()``
    chains; ``run()`` without keyword arguments after an explicit
    ``initialize`` keeps those hyperparameters (the reference's ``run`` always
    re-initialises, ``inference.py:125``, silently resetting eps/L to
    0.25/2).
  * ``accept_rate`` is ``n_accept / t`` read after the increment (the
    reference fetches it in the same session call as the increment,
    ``monte_carlo.py:139-140``, which races and divides by 0 at t = 0).
  * ``Empirical(params[T, C, D])`` for a ``[D]`` latent runs C independent
    chains (a shape error in the reference, ``util/random_variables.py:78-83``).
  * A non-finite proposal is rejected and counted (``divergences``); with
    ``debug=True`` ``update`` raises ``FloatingPointError`` (the reference's
    ``tf.add_check_numerics_ops``, ``inference.py:279-281``).
"""

import json
import os
import time

import numpy as np
import torch

from edward_amd import models as M
from edward_amd import compile as _compile
from edward_amd.progbar import Progbar


def check_latent_vars(latent_vars):
  """``util/random_variables.py:64-83``."""
  if not isinstance(latent_vars, dict):
    raise TypeError('latent_vars must have type dict.')
  for key, value in latent_vars.items():
    if not isinstance(key, (M.RandomVariable, M.Node)):
      raise TypeError('Latent variable key has an invalid type: {}'.format(type(key)))
    if not isinstance(value, (M.RandomVariable, M.Node)):
      raise TypeError('Latent variable value has an invalid type: {}'.format(type(value)))
    if isinstance(value, M.Empirical) and isinstance(key, M.RandomVariable):
      ev, ks = tuple(value.event_shape), tuple(key.shape)
      if ev != ks and not (len(ev) == len(ks) + 1 and ev[1:] == ks):
        raise TypeError('Key-value pair in latent_vars does not have same shape: '
                        '{}, {}'.format(ks, ev))
      if value.dtype != key.dtype:
        raise TypeError('Key-value pair in latent_vars does not have same dtype: '
                        '{}, {}'.format(key.dtype, value.dtype))


def check_data(data):
  """``util/random_variables.py:21-61``."""
  if not isinstance(data, dict):
    raise TypeError('data must have type dict.')
  for key, value in data.items():
    if isinstance(key, M.Placeholder):
      if isinstance(value, (M.RandomVariable, M.Node)):
        raise TypeError('The value of a feed cannot be a ed.RandomVariable object. '
                        'Acceptable feed values include Python scalars, strings, lists, '
                        'numpy ndarrays, or TensorHandles.')
    elif isinstance(key, M.RandomVariable):
      if isinstance(value, (M.RandomVariable, M.Node)):
        continue
      if isinstance(value, (float, int, list, np.ndarray, np.number, torch.Tensor)):
        shp = tuple(value.shape) if hasattr(value, 'shape') else np.shape(value)
        if key.shape and tuple(shp) != tuple(key.shape):
          raise TypeError('Key-value pair in data does not have same shape: '
                          '{}, {}'.format(key.shape, shp))
      else:
        raise TypeError('Data value has an invalid type: {}'.format(type(value)))
    else:
      raise TypeError('Data key has an invalid type: {}'.format(type(key)))


class Inference(object):
  """``edward/inferences/inference.py:19-338``."""

  def __init__(self, latent_vars=None, data=None):
    if latent_vars is None:
      latent_vars = {}
    if data is None:
      data = {}
    check_latent_vars(latent_vars)
    self.latent_vars = latent_vars
    check_data(data)
    self.data = dict(data)
    self._initialized = False

  def run(self, variables=None, use_coordinator=True, *args, **kwargs):
    """``inference.py:97-154``: initialize, (re)initialise variables, loop
    ``update``/``print_progress`` n_iter times, ``finalize``."""
    if args or kwargs or not self._initialized:
      self.initialize(*args, **kwargs)
    self._initialize_variables(variables)
    for _ in range(self.n_iter):
      info_dict = self.update()
      self.print_progress(info_dict)
    self.finalize()

  def initialize(self, n_iter=1000, n_print=None, scale=None, auto_transform=True,
                 logdir=None, log_timestamp=True, log_vars=None, debug=False):
    """``inference.py:157-285``."""
    self.n_iter = int(n_iter)
    self.n_print = int(n_iter / 100) if n_print is None else int(n_print)
    self.progbar = Progbar(self.n_iter)
    self.t = 0
    if scale is None:
      scale = {}
    elif not isinstance(scale, dict):
      raise TypeError('scale must be a dict object.')
    self.scale = scale
    self._auto_transform(auto_transform)
    self.logging = logdir is not None
    if self.logging:
      if log_timestamp:
        logdir = os.path.join(os.path.expanduser(logdir),
                              time.strftime('%Y%m%d_%H%M%S', time.gmtime()))
      os.makedirs(logdir, exist_ok=True)
      self._log_path = os.path.join(logdir, 'metrics.jsonl')
    self.debug = debug
    self._initialized = True
    return self

  def _initialize_variables(self, variables):
    pass

  def update(self, feed_dict=None):
    raise NotImplementedError()

  def print_progress(self, info_dict):
    raise NotImplementedError()

  def finalize(self):
    """``inference.py:334-338``."""
    if getattr(self, 'logging', False):
      pass

  def _auto_transform(self, auto_transform):
    """``inference.py:222-264``: with ``auto_transform``, a latent whose
    support differs from its approximation's is paired with its unconstrained
    version ``ed.transform(z)`` (``transformations``); an ``Empirical``
    approximation (support ``'points'``) then holds the unconstrained values,
    and ``latent_vars`` maps the latent to the pushed-forward view whose
    ``params`` are the constrained samples."""
    from edward_amd import bijectors
    if not hasattr(self, '_user_latent_vars'):
      self._user_latent_vars = dict(self.latent_vars)
    latent_vars = self._user_latent_vars
    self.transformations = {}
    self._transforms = {}
    if not auto_transform:
      self.latent_vars = dict(latent_vars)
      if hasattr(self, 'latent_vars_unconstrained'):
        del self.latent_vars_unconstrained
      return
    self.latent_vars = {}
    self.latent_vars_unconstrained = {}
    for z, qz in latent_vars.items():
      zs, qs = getattr(z, 'support', None), getattr(qz, 'support', None)
      if zs is not None and qs is not None and zs != qs and qs != 'point':
        z_u = bijectors.transform(z)
        self.transformations[z] = z_u
        qz_u = qz if qs == 'points' else bijectors.transform(qz)
        self.latent_vars_unconstrained[z_u] = qz_u
        if z_u is not z:
          self._transforms[z] = z_u.bijector.bijector.name   # the map u -> theta
          qz_c = M.ConstrainedEmpirical(qz_u, z_u.bijector)
        else:
          qz_c = qz_u
        self.latent_vars[z] = qz_c
      else:
        self.latent_vars[z] = qz
        self.latent_vars_unconstrained[z] = qz

  def _log(self, rec):
    if getattr(self, 'logging', False):
      with open(self._log_path, 'a') as f:
        f.write(json.dumps(rec) + '\n')


class MonteCarlo(Inference):
  """``edward/inferences/monte_carlo.py:15-170``."""

  def __init__(self, latent_vars=None, data=None):
    if isinstance(latent_vars, list):
      latent_vars = {z: M.Empirical(params=torch.zeros(
          (10000,) + tuple(z.shape), dtype=getattr(torch, z.dtype)))
          for z in latent_vars}
    elif isinstance(latent_vars, dict):
      for qz in latent_vars.values():
        if not isinstance(qz, M.Empirical):
          raise TypeError('Posterior approximation must consist of only Empirical '
                          'random variables.')
        elif len(qz.sample_shape) != 0:
          raise ValueError('Empirical posterior approximations must have a scalar '
                           'sample shape.')
    super(MonteCarlo, self).__init__(latent_vars, data)

  def initialize(self, *args, **kwargs):
    kwargs['n_iter'] = int(np.amin([qz.params.shape[0] for qz in self._stores().values()]))
    super(MonteCarlo, self).initialize(*args, **kwargs)
    self.train = self.build_update()
    self._reset_counts()
    return self

  def _reset_counts(self):
    self.t = 0
    self._n_accept = torch.zeros(self.n_chains, dtype=torch.int32, device=self._device)

  @property
  def n_accept(self):
    """Accepted transitions: an int for one chain, an int array over chains."""
    a = self._n_accept.cpu().numpy()
    return int(a[0]) if self.n_chains == 1 else a

  def reset(self):
    """The reference's ``reset`` ops (``inference.py:285``,
    ``monte_carlo.py:104``): t = 0 and n_accept = 0."""
    self._reset_counts()
    self._state_t = None

  # -- model binding (shared by the samplers) ------------------------------
  _scale_aware = False     # SGLD / SGHMC honour `scale`; HMC ignores it (hmc.py:161-192)

  _constrained_ok = False  # only HMC samples auto-transformed latents (hmc.py:132-159)

  def _stores(self):
    """The user's ``Empirical`` stores (for an auto-transformed latent they
    hold the unconstrained positions, as in the reference)."""
    return getattr(self, '_user_latent_vars', self.latent_vars)

  def _compile_transforms(self):
    tr = getattr(self, '_transforms', {})
    if tr and not self._constrained_ok:
      raise NotImplementedError(
          '%s cannot sample constrained latents: the reference writes its samples into '
          'the derived constrained tensor, which is not a store; use HMC, or give the '
          'latent unconstrained support' % type(self).__name__)
    return tr

  def _prepare(self):
    """Chain count and device from the stores; compile the model now if
    every placeholder it needs is bound, else at the first fed update."""
    from edward_amd.engine import require_gpu
    require_gpu()
    stores = self._stores()
    first = next(iter(stores))
    self.n_chains = _compile._chains_of(first, stores[first])
    self._device = torch.device('cuda', 0)
    self.program = None
    self.engine = None
    self._state_t = None
    try:
      self._compile(self.data)
    except _compile.UnboundPlaceholder:
      pass

  def _compile(self, data):
    prog = _compile.compile_model(self._stores(), data, precision=self.precision,
                                  scale=self.scale if self._scale_aware else None,
                                  transforms=self._compile_transforms())
    self.program = prog
    self.engine = prog.engine
    self.n_chains = prog.n_chains
    self._device = self.engine.device
    self._row = torch.empty(prog.n_chains, prog.dim, device=self._device)
    self._on_compiled()

  def _on_compiled(self):
    pass

  def _feed(self, feed_dict):
    """Bind fed placeholder values (``monte_carlo.py:134-136``): the first
    feed compiles the model, later feeds rebind the data on the device."""
    data = dict(self.data)
    data.update(feed_dict)
    if self.program is None:
      self._compile(data)
    else:
      _compile.rebind_model(self.program, self._stores(), data,
                            self.scale if self._scale_aware else None,
                            self._compile_transforms())
    self._state_t = None

  def _check_bound(self):
    if self.program is None:
      raise ValueError('the model has unbound placeholders: pass them in '
                       'update(feed_dict=...)')
    T = self.program.n_rows
    if self.t >= T:
      raise IndexError('t=%d is past the %d rows of the Empirical store' % (self.t, T))

  def _load_state(self):
    need = max(self.t - 1, 0)
    if self._state_t != need:
      self.engine.set_state(self.program.gather(need))

  def update(self, feed_dict=None):
    """``monte_carlo.py:111-150``: one transition, then t += 1."""
    if feed_dict:
      self._feed(feed_dict)
    self._check_bound()
    self._step()
    self.t += 1
    acc = self._n_accept.to(torch.float64).mean().item() / self.t
    info = {'t': self.t, 'accept_rate': acc}
    if self.debug:
      nd = int(self.engine.divergences().sum().item())
      if nd:
        raise FloatingPointError('%d chain(s) produced non-finite proposals' % nd)
    if self.logging and self.n_print != 0 and (self.t == 1 or self.t % self.n_print == 0):
      self._log(info)
    return info

  def print_progress(self, info_dict):
    """``monte_carlo.py:152-158``."""
    if self.n_print != 0:
      t = info_dict['t']
      if t == 1 or t % self.n_print == 0:
        self.progbar.update(t, {'Acceptance Rate': info_dict['accept_rate']})

  def build_update(self):
    raise NotImplementedError()


class HMC(MonteCarlo):
  """Hamiltonian Monte Carlo (``edward/inferences/hmc.py:14-130``).

  ``HMC(latent_vars={w: Empirical(params)}, data={X: X_data, y: y_data})``;
  ``initialize(step_size=0.25, n_steps=2, ...)``; ``run()`` or ``update()``.
  Build-only keyword arguments of ``initialize``:
    precision  'fp32' (default) or 'bf16' (X in bf16 on the matrix cores)
    seed       Philox key of the device momenta/uniforms (default: the
               ``ed.set_seed`` value, else 0)
    surrogate  a trained ``edward_amd.surrogate.Surrogate`` to drive the
               leapfrog (the exact log joint still decides acceptance)
    step_size_jitter  j in [0, 1): transition t uses step_size * U(1-j, 1+j)
               (default 0: the reference's fixed step size)
  """

  _constrained_ok = True

  def __init__(self, *args, **kwargs):
    super(HMC, self).__init__(*args, **kwargs)

  def initialize(self, step_size=0.25, n_steps=2, *args, **kwargs):
    """``hmc.py:45-59``; returns ``self``."""
    self.step_size = step_size
    self.n_steps = int(n_steps)
    self.precision = kwargs.pop('precision', 'fp32')
    from edward_amd import util
    seed = kwargs.pop('seed', None)
    self.seed = util.get_seed() if seed is None else int(seed)
    self.surrogate = kwargs.pop('surrogate', None)
    self.step_size_jitter = float(kwargs.pop('step_size_jitter', 0.0))
    if not 0.0 <= self.step_size_jitter < 1.0:
      raise ValueError('step_size_jitter must be in [0, 1)')
    return super(HMC, self).initialize(*args, **kwargs)

  def build_update(self):
    """Compile the model (``hmc.py:61-130`` builds the transition graph)."""
    if not hasattr(self, 'latent_vars_unconstrained'):
      raise ValueError('This implementation of HMC requires that all variables have '
                       'unconstrained support. Please initialize with auto_transform=True '
                       'to ensure this. (if your variables already have unconstrained '
                       'support then doing this is a no-op).')
    self._prepare()
    return self

  def _on_compiled(self):
    if self.surrogate is not None:
      self.surrogate.attach(self.engine)

  def _step(self):
    t = self.t
    self._load_state()
    from edward_amd.engine import jittered_step
    eps = jittered_step(self.step_size, self.step_size_jitter, self.seed, t)
    if self.surrogate is None:
      self.engine.transition(eps, self.n_steps, seed=self.seed, counter=t,
                             sample_out=self._row, n_accept=self._n_accept)
    else:
      self.surrogate.transition(eps, self.n_steps, seed=self.seed, counter=t,
                                sample_out=self._row, n_accept=self._n_accept)
    self.program.scatter(t, self._row)
    self._state_t = t

  @property
  def divergences(self):
    return self.engine.divergences().cpu().numpy()


class _SGMonteCarlo(MonteCarlo):
  """Shared set-up of the stochastic-gradient samplers: build-only keyword
  arguments ``precision`` and ``seed`` as for HMC; ``scale`` multiplies the
  observed variable's log likelihood and each latent's log prior (Edward's
  minibatch scaling, ``inference.py:187-193``)."""

  _scale_aware = True

  def _common_initialize(self, kwargs):
    self.precision = kwargs.pop('precision', 'fp32')
    from edward_amd import util
    seed = kwargs.pop('seed', None)
    self.seed = util.get_seed() if seed is None else int(seed)

  def build_update(self):
    self._prepare()
    return self


class SGLD(_SGMonteCarlo):
  """Stochastic gradient Langevin dynamics (``edward/inferences/sgld.py:10-121``,
  Welling and Teh 2011): z_t = z_{t-1} + lr/2 grad log p(x, z_{t-1}) +
  N(0, lr) with lr = step_size / (t + 1)^0.55; every draw is kept
  (n_accept counts each update)."""

  def initialize(self, step_size=0.25, *args, **kwargs):
    """``sgld.py:45-54``; returns ``self``."""
    self.step_size = step_size
    self._common_initialize(kwargs)
    return super(SGLD, self).initialize(*args, **kwargs)

  def _step(self):
    t = self.t
    self._load_state()
    lr = float(self.step_size) / float(t + 1) ** 0.55
    self.engine.sgld_step(lr, seed=self.seed, counter=t, sample_out=self._row,
                          n_accept=self._n_accept)
    self.program.scatter(t, self._row)
    self._state_t = t


class SGHMC(_SGMonteCarlo):
  """Stochastic gradient Hamiltonian Monte Carlo
  (``edward/inferences/sghmc.py:11-135``, Chen et al. 2014 eq. 15) with a
  velocity per chain that starts at zero (``sghmc.py:52-56``):
  z_t = z_{t-1} + v, v <- (1 - friction/2) v + lr grad + N(0, lr friction),
  lr = step_size * 0.01."""

  def initialize(self, step_size=0.25, friction=0.1, *args, **kwargs):
    """``sghmc.py:42-56``; returns ``self``."""
    self.step_size = step_size
    self.friction = friction
    self._common_initialize(kwargs)
    return super(SGHMC, self).initialize(*args, **kwargs)

  def _on_compiled(self):
    self.v = torch.zeros(self.program.n_chains, self.program.dim, device=self._device)

  def _step(self):
    t = self.t
    self._load_state()
    lr = float(self.step_size) * 0.01
    self.engine.sghmc_step(lr, float(self.friction), self.v, seed=self.seed, counter=t,
                           sample_out=self._row, n_accept=self._n_accept)
    self.program.scatter(t, self._row)
    self._state_t = t


class MetropolisHastings(MonteCarlo):
  """Metropolis-Hastings (``edward/inferences/metropolis_hastings.py:15-156``)
  with random-walk proposals: ``proposal_vars[z]`` must be
  ``Normal(loc=z, scale=s)`` with a constant s (scalar or per coordinate),
  a symmetric proposal whose log densities cancel in the ratio.  Latents
  without a proposal stay at their value (as in the reference, where only
  ``proposal_vars`` move).  Build-only ``initialize`` keywords: ``precision``,
  ``seed``."""

  def __init__(self, latent_vars, proposal_vars, data=None):
    check_latent_vars(proposal_vars)
    self.proposal_vars = proposal_vars
    super(MetropolisHastings, self).__init__(latent_vars, data)

  def initialize(self, *args, **kwargs):
    self.precision = kwargs.pop('precision', 'fp32')
    from edward_amd import util
    seed = kwargs.pop('seed', None)
    self.seed = util.get_seed() if seed is None else int(seed)
    return super(MetropolisHastings, self).initialize(*args, **kwargs)

  def build_update(self):
    for z, g in self.proposal_vars.items():
      if z not in self.latent_vars:
        raise NotImplementedError('proposal for %s, which is not a latent variable' % z.name)
      if not isinstance(g, M.Normal) or g.loc is not z:
        raise NotImplementedError('proposals must be Normal(loc=z, scale=s) random walks')
      if _compile._resolve(g.scale, {}) is None:
        raise NotImplementedError('the proposal scale must be a constant')
    self._prepare()
    return self

  def _on_compiled(self):
    scale = np.zeros(self.program.dim, np.float32)
    for lat in self.program.latents:
      g = self.proposal_vars.get(lat.rv)
      if g is None:
        continue
      s = np.asarray(_compile._resolve(g.scale, {}), np.float64)
      scale[lat.offset:lat.offset + lat.k] = np.broadcast_to(s.reshape(-1) if s.ndim else s,
                                                             (lat.k,))
    if np.any(scale < 0) or not np.all(np.isfinite(scale)):
      raise ValueError('proposal scales must be finite and non-negative')
    self._proposal_scale = torch.as_tensor(scale, device=self._device)

  def _step(self):
    t = self.t
    self._load_state()
    self.engine.mh_step(self._proposal_scale, seed=self.seed, counter=t, sample_out=self._row,
                        n_accept=self._n_accept)
    self.program.scatter(t, self._row)
    self._state_t = t


class ReplicaExchangeMC(MetropolisHastings):
  """Replica-exchange Monte Carlo (``edward/inferences/replica_exchange_mc.py:
  14-276``): replica i runs random-walk Metropolis-Hastings on the log joint
  tempered by ``inverse_temperatures[i]`` (the first must be 1); with
  probability ``exchange_freq`` per update adjacent replicas exchange states
  (Metropolis criterion on the untempered log joints).  Every replica keeps
  its own ``Empirical`` store (``replica_vars``, zeros at the start, where
  each replica's chain starts); the user's store receives replica 0.  A
  ``[T, C, D]`` store runs C independent replica ladders."""

  def __init__(self, latent_vars, proposal_vars, data=None,
               inverse_temperatures=np.logspace(0, -2, 5), exchange_freq=0.1):
    betas = np.asarray(inverse_temperatures, np.float64).reshape(-1)
    if betas.size < 1 or betas[0] != 1:
      raise ValueError('inverse_temperatures[0] must be 1.')
    super(ReplicaExchangeMC, self).__init__(latent_vars, proposal_vars, data)
    self.n_replica = int(betas.size)
    self.inverse_temperatures = betas
    self.exchange_freq = float(exchange_freq)
    self.replica_vars = [{z: M.Empirical(torch.zeros_like(qz.params)) for z, qz in
                          self.latent_vars.items()} for _ in range(self.n_replica)]

  def initialize(self, *args, **kwargs):
    """``replica_exchange_mc.py:74-76`` (auto_transform off)."""
    kwargs['auto_transform'] = False
    return super(ReplicaExchangeMC, self).initialize(*args, **kwargs)

  def _on_compiled(self):
    super(ReplicaExchangeMC, self)._on_compiled()
    R, C = self.n_replica, self.program.n_chains
    self._beta = torch.as_tensor(np.repeat(self.inverse_temperatures, C).astype(np.float32),
                                 device=self._device)
    self._row_all = torch.empty(R * C, self.program.dim, device=self._device)
    self._n_accept_all = torch.zeros(R * C, dtype=torch.int32, device=self._device)
    self._swaps = torch.zeros(C, dtype=torch.int32, device=self._device)

  def _reset_counts(self):
    super(ReplicaExchangeMC, self)._reset_counts()
    if getattr(self, '_n_accept_all', None) is not None:
      self._n_accept_all.zero_()

  def _gather_replicas(self, t):
    C = self.program.n_chains
    rows = []
    for rep in self.replica_vars:
      rows.append(torch.cat([rep[l.rv].params[t].reshape(C, l.k).to(torch.float32)
                             for l in self.program.latents], dim=1))
    return torch.cat(rows, dim=0).contiguous()

  def _scatter_replicas(self, t, rows):
    C = self.program.n_chains
    for i, rep in enumerate(self.replica_vars):
      block = rows[i * C:(i + 1) * C]
      for l in self.program.latents:
        dst = rep[l.rv].params[t]
        dst.copy_(block[:, l.offset:l.offset + l.k].reshape(dst.shape))
    self.program.scatter(t, rows[:C])

  def _step(self):
    t = self.t
    need = max(t - 1, 0)
    if self._state_t != need:
      self.engine.set_state(self._gather_replicas(need))
    self.engine.replica_exchange_step(self._proposal_scale, self._beta, self.n_replica,
                                      self.exchange_freq, seed=self.seed, counter=t,
                                      sample_out=self._row_all, n_accept=self._n_accept_all,
                                      swaps=self._swaps)
    self._n_accept.copy_(self._n_accept_all[:self.program.n_chains])
    self._scatter_replicas(t, self._row_all)
    self._state_t = t

  @property
  def n_swaps(self):
    """Exchanges per replica ladder since the start (no reference counterpart)."""
    return self._swaps.cpu().numpy()
