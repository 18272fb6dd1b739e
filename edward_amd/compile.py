"""Match an Edward model to a device engine.

The reference evaluates the log joint by copying the model graph at the
proposed sample (``hmc.py:161-192`` via ``util/random_variables.py:139-445``).
Here the model DAG is matched once to a kernel family:

  * GLM: one observed ``Bernoulli(logits=E)`` or ``Normal(loc=E, scale=s)``
    with E a sum of ``ed.dot(X, w)`` terms, bare latents broadcast over rows
    (a bias, or the mean of ``Normal(loc=mu, sample_shape=n)``) and constant
    multiples of these; every latent ``Normal(loc, scale)`` with constant
    parameters.  The latents are concatenated into one vector and the
    design matrix gets one column per coordinate (ones for broadcast terms).
  * Hierarchical: ``Bernoulli(logits=ed.gather(a, group) + ed.dot(X, w))``
    with ``a ~ Normal(mu, exp(log_sigma))``, ``mu, log_sigma ~ N(0, 1)``,
    ``w ~ N(0, I)``.

Anything else raises ``NotImplementedError``.
"""

import numpy as np
import torch

from edward_amd import models as M


def _array(v):
  if isinstance(v, torch.Tensor):
    return v.detach().cpu().numpy()
  return np.asarray(v)


class UnboundPlaceholder(ValueError):
  """A placeholder the model needs has no value yet (it is fed later through
  ``update(feed_dict=...)``)."""


def _resolve(node, data):
  """Constant value of a data node (placeholder bound in data, or array)."""
  if isinstance(node, M.Placeholder):
    if node not in data:
      raise UnboundPlaceholder('placeholder %s is not bound in data' % node.name)
    return _array(data[node]).astype(np.float64)
  if isinstance(node, (M.RandomVariable, M.Node)):
    return None
  return _array(node).astype(np.float64)


class _Term(object):
  def __init__(self, rv, cols):
    self.rv = rv        # latent RandomVariable
    self.cols = cols    # [N, k] coefficient block (or callable(N) -> block)


def _linear_terms(E, data, latents):
  """Flatten E into (latent, coefficient block) terms."""
  if isinstance(E, M.Add):
    return _linear_terms(E.a, data, latents) + _linear_terms(E.b, data, latents)
  if isinstance(E, M.Mul):
    ca, cb = _resolve(E.a, data), _resolve(E.b, data)
    if ca is not None and cb is None:
      const, other = ca, E.b
    elif cb is not None and ca is None:
      const, other = cb, E.a
    else:
      raise NotImplementedError('only constant * latent-expression products are supported')
    if np.ndim(const) != 0:
      raise NotImplementedError('only scalar multipliers are supported')
    terms = _linear_terms(other, data, latents)
    for t in terms:
      t.cols = _scaled(t.cols, float(const))
    return terms
  if isinstance(E, M.Dot):
    x, y = E.x, E.y
    if isinstance(y, M.RandomVariable) and y in latents:
      mat = _resolve(x, data)
      if mat is None or mat.ndim != 2:
        raise NotImplementedError('ed.dot(X, w) needs a data matrix X')
      return [_Term(y, mat)]
    if isinstance(x, M.RandomVariable) and x in latents:
      mat = _resolve(y, data)
      if mat is None or mat.ndim != 2:
        raise NotImplementedError('ed.dot(w, X) needs a data matrix X')
      return [_Term(x, mat.T)]
    raise NotImplementedError('ed.dot must pair a data matrix with a latent vector')
  if isinstance(E, M.RandomVariable):
    if E not in latents:
      raise NotImplementedError('%s is neither a latent nor data' % E.name)
    k = int(np.prod(E.shape)) if E.shape else 1
    if k != 1:
      raise NotImplementedError('only scalar latents can be broadcast over rows')
    return [_Term(E, 'ones')]
  c = _resolve(E, data)
  if c is not None:
    if np.any(c != 0):
      raise NotImplementedError('constant offsets in the linear predictor are not supported')
    return []
  raise NotImplementedError('unsupported node %r in the linear predictor' % (E,))


def _scaled(cols, s):
  if isinstance(cols, str):
    return ('ones', s)
  if isinstance(cols, tuple):
    return (cols[0], cols[1] * s)
  return cols * s


def _block(cols, N):
  if isinstance(cols, str):
    return np.ones((N, 1))
  if isinstance(cols, tuple):
    return np.full((N, 1), cols[1])
  if cols.shape[0] != N:
    raise ValueError('design block has %d rows, observations have %d' % (cols.shape[0], N))
  return cols


class Latent(object):
  def __init__(self, rv, qz, k, offset, n_chains):
    self.rv, self.qz, self.k, self.offset, self.n_chains = rv, qz, k, offset, n_chains


class Program(object):
  """A compiled model: the engine plus the map between the latent vector and
  the Empirical stores."""

  def __init__(self, engine, latents, n_rows, n_chains, dim):
    self.engine = engine
    self.latents = latents
    self.n_rows = n_rows
    self.n_chains = n_chains
    self.dim = dim

  def gather(self, t):
    """``params[t]`` of every latent as one [C, dim] float32 block."""
    parts = [l.qz.params[t].reshape(self.n_chains, l.k).to(torch.float32)
             for l in self.latents]
    return torch.cat(parts, dim=1).contiguous()

  def scatter(self, t, row):
    for l in self.latents:
      dst = l.qz.params[t]
      dst.copy_(row[:, l.offset:l.offset + l.k].reshape(dst.shape))


def _chains_of(rv, qz):
  ev = tuple(qz.event_shape)
  ks = tuple(rv.shape)
  if ev == ks:
    return 1
  if len(ev) == len(ks) + 1 and ev[1:] == ks:
    return ev[0]
  raise TypeError('Empirical of shape %s does not fit latent of shape %s' % (ev, ks))


def _normal_prior(rv, k):
  if not isinstance(rv, M.Normal):
    raise NotImplementedError('latent %s must be a Normal' % rv.name)
  loc, scale = _resolve(rv.loc, {}), _resolve(rv.scale, {})
  if loc is None or scale is None:
    return None
  shape = rv.shape if rv.shape else ()
  loc = np.broadcast_to(loc, shape).reshape(-1) if shape else np.full(1, float(loc))
  scale = np.broadcast_to(scale, shape).reshape(-1) if shape else np.full(1, float(scale))
  if loc.size != k:
    loc = np.resize(loc, k)
    scale = np.resize(scale, k)
  if np.any(scale <= 0):
    raise ValueError('prior scale of %s must be positive' % rv.name)
  return loc, scale


class GLMSpec(object):
  """What the GLM engine is bound with: design matrix, observations, prior,
  likelihood, and the latent layout."""

  def __init__(self, **kw):
    self.__dict__.update(kw)


def compile_model(latent_vars, data, precision='fp32', scale=None):
  """Bind the model to a device engine.  ``scale`` (Edward's per-variable
  multipliers, inference.py:187-193) is honoured by the stochastic-gradient
  samplers only; HMC passes None (hmc.py:161-192 ignores it)."""
  spec = glm_spec(latent_vars, data, scale)
  if not isinstance(spec, GLMSpec):
    return spec          # another engine family built its own program
  from edward_amd.engine import GLMEngine
  eng = GLMEngine(spec.X, spec.y, spec.loc, spec.scale, likelihood=spec.likelihood,
                  lik_scale=spec.lik_scale, precision=precision)
  if spec.lik_weight != 1.0:
    eng.set_lik_weight(spec.lik_weight)
  return Program(eng, spec.latents, spec.n_rows, spec.n_chains, spec.dim)


def rebind_model(program, latent_vars, data, scale=None):
  """Bind new data values (a fed minibatch) to a compiled program's engine;
  the latent layout must not change.  Discards the engine's chain state."""
  spec = glm_spec(latent_vars, data, scale)
  if not isinstance(spec, GLMSpec):
    raise NotImplementedError('feeding data is supported for the GLM family only')
  if spec.dim != program.dim or spec.n_chains != program.n_chains:
    raise ValueError('fed data changes the latent layout')
  eng = program.engine
  eng.bind(spec.X, spec.y, spec.loc, spec.scale, spec.likelihood, spec.lik_scale)
  if spec.lik_weight != 1.0:
    eng.set_lik_weight(spec.lik_weight)


def _scalar_scale(scale, rv):
  w = (scale or {}).get(rv, 1.0)
  w = np.asarray(w.detach().cpu().numpy() if isinstance(w, torch.Tensor) else w, np.float64)
  if w.size != 1:
    raise NotImplementedError('only scalar scale factors are supported')
  w = float(w.reshape(-1)[0])
  if not w > 0:
    raise ValueError('scale factors must be positive')
  return w


def _observed_values(obs, data):
  v = data[obs]
  if isinstance(v, M.Placeholder):
    if v not in data:
      raise UnboundPlaceholder('placeholder %s is not bound: feed it in '
                               'update(feed_dict=...)' % v.name)
    v = data[v]
  return _array(v).astype(np.float64).reshape(-1)


def glm_spec(latent_vars, data, scale=None):
  """Match the model to the GLM family; returns a GLMSpec (no device work)."""
  observed = [k for k in data if isinstance(k, M.RandomVariable)]
  if len(observed) != 1:
    raise NotImplementedError('exactly one observed random variable is supported, got %d'
                              % len(observed))
  obs = observed[0]
  yv = _observed_values(obs, data)
  latents = list(latent_vars.keys())
  if isinstance(obs, M.Bernoulli) and isinstance(obs.logits, M.Add) and \
          _has_gather(obs.logits):
    if scale:
      raise NotImplementedError('scale is not supported for the hierarchical model')
    from edward_amd import hier
    return hier.compile_hier(latent_vars, data, obs, yv)
  if isinstance(obs, M.Bernoulli):
    if obs.logits is None:
      raise NotImplementedError('Bernoulli(probs=...) likelihoods are not supported; '
                                'use logits')
    E, likelihood, lik_scale = obs.logits, 'bernoulli', 1.0
  elif isinstance(obs, M.Normal):
    s = _resolve(obs.scale, data)
    if s is None or np.ndim(s) > 1 or np.any(s != np.ravel(s)[0]):
      raise NotImplementedError('Normal likelihood needs one constant scale')
    E, likelihood, lik_scale = obs.loc, 'normal', float(np.ravel(s)[0])
  else:
    raise NotImplementedError('unsupported likelihood %s' % type(obs).__name__)
  terms = _linear_terms(E, data, set(latents))
  N = yv.size
  # Latent vector layout follows latent_vars' order.
  blocks, locs, scales, lat = [], [], [], []
  off = 0
  n_chains = None
  for rv in latents:
    qz = latent_vars[rv]
    k = int(np.prod(rv.shape)) if rv.shape else 1
    C = _chains_of(rv, qz)
    if n_chains is None:
      n_chains = C
    elif C != n_chains:
      raise ValueError('all Empirical stores must hold the same number of chains')
    pr = _normal_prior(rv, k)
    if pr is None:
      raise NotImplementedError('prior parameters of %s must be constants' % rv.name)
    locs.append(pr[0])
    # a scale factor w on log N(z; m, s) has the gradient of N(z; m, s / sqrt(w))
    # (the samplers that honour scale use gradients only)
    scales.append(pr[1] / np.sqrt(_scalar_scale(scale, rv)))
    blk = np.zeros((N, k))
    for t in terms:
      if t.rv is rv:
        b = _block(t.cols, N)
        if b.shape[1] != k:
          raise ValueError('ed.dot operand of %d columns for latent of size %d'
                           % (b.shape[1], k))
        blk = blk + b
    blocks.append(blk)
    lat.append(Latent(rv, qz, k, off, C))
    off += k
  X = np.concatenate(blocks, axis=1).astype(np.float32)
  T = int(min(l.qz.params.shape[0] for l in lat))
  return GLMSpec(X=X, y=yv.astype(np.float32), loc=np.concatenate(locs),
                 scale=np.concatenate(scales), likelihood=likelihood, lik_scale=lik_scale,
                 lik_weight=_scalar_scale(scale, obs), latents=lat, n_rows=T,
                 n_chains=n_chains, dim=off)


def _has_gather(E):
  if isinstance(E, M.Gather):
    return True
  if isinstance(E, M.Add):
    return _has_gather(E.a) or _has_gather(E.b)
  return False
