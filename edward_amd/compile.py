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

Latents with a constrained prior (Gamma, Exponential, Chi2, InverseGamma,
Beta, or a ``TransformedDistribution`` of a Normal by Softplus/Sigmoid with
that support) are sampled in the unconstrained space when the inference
auto-transforms them: the device evaluates their prior with the
log-Jacobian per coordinate (``csrc/transform.h``) and feeds the linear
predictor the constrained value.  ``Bernoulli(probs=theta)`` with a
sigmoid-transformed latent theta is the GLM with logits = u.  A model with
no observed variable binds one zero row with likelihood weight 0.

Anything else raises ``NotImplementedError``.
"""

import math


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


# prior families and operand maps of csrc/transform.h
FAMILY_NORMAL, FAMILY_GAMMA, FAMILY_BETA, FAMILY_INVGAMMA = 0, 1, 2, 3
OP_IDENTITY, OP_SOFTPLUS, OP_SIGMOID = 0, 1, 2
_HALF_LOG_2PI = 0.5 * math.log(2.0 * math.pi)


def _f32(a):
  """Parameters as the device holds them (float32), in float64."""
  return np.asarray(a, np.float64).astype(np.float32).astype(np.float64)


def _lgamma(a):
  return np.vectorize(math.lgamma, otypes=[np.float64])(a)


class PriorFamily(object):
  """Per-coordinate prior of the unconstrained position: family, its two
  parameters, normalising constant and the operand map."""

  def __init__(self, family, p1, p2, op, k):
    self.family = np.full(k, family, np.uint8)
    self.p1 = _f32(np.resize(np.asarray(p1, np.float64).reshape(-1), k))
    self.p2 = _f32(np.resize(np.asarray(p2, np.float64).reshape(-1), k))
    self.op = np.full(k, op, np.uint8)
    if family == FAMILY_NORMAL:
      if np.any(self.p2 <= 0):
        raise ValueError('prior scale must be positive')
      self.lnorm = -np.log(self.p2) - _HALF_LOG_2PI
    elif family in (FAMILY_GAMMA, FAMILY_INVGAMMA):
      if np.any(self.p1 <= 0) or np.any(self.p2 <= 0):
        raise ValueError('concentration and rate must be positive')
      self.lnorm = self.p1 * np.log(self.p2) - _lgamma(self.p1)
    else:
      if np.any(self.p1 <= 0) or np.any(self.p2 <= 0):
        raise ValueError('Beta concentrations must be positive')
      self.lnorm = _lgamma(self.p1 + self.p2) - _lgamma(self.p1) - _lgamma(self.p2)

  @property
  def kind(self):
    return (self.family | (self.op << 4)).astype(np.uint8)


def _constrained_family(rv, k, transform):
  """Prior family of an auto-transformed latent (``transform`` names the map
  u -> theta: 'softplus' or 'sigmoid')."""
  if isinstance(rv, M.TransformedDistribution):
    base, bij = rv.distribution, rv.bijector
    want = {'softplus': 'softplus', 'sigmoid': 'sigmoid'}.get(transform)
    if not isinstance(base, M.Normal) or getattr(bij, 'name', None) != want:
      raise NotImplementedError('TransformedDistribution latents must be a Normal pushed '
                                'through the bijector of their support (softplus for '
                                "'nonnegative', sigmoid for '01')")
    loc, scale = _resolve(base.loc, {}), _resolve(base.scale, {})
    if loc is None or scale is None:
      raise NotImplementedError('prior parameters of %s must be constants' % rv.name)
    # log p_theta(T(u)) + log T'(u) is the base Normal's density at u
    return PriorFamily(FAMILY_NORMAL, np.broadcast_to(loc, rv.shape or ()),
                       np.broadcast_to(scale, rv.shape or ()),
                       OP_SOFTPLUS if want == 'softplus' else OP_SIGMOID, k)
  if transform == 'softplus':
    if isinstance(rv, M.Gamma):
      a, b = rv.param_values()
      return PriorFamily(FAMILY_GAMMA, a, b, OP_SOFTPLUS, k)
    if isinstance(rv, M.Exponential):
      (b,) = rv.param_values()
      return PriorFamily(FAMILY_GAMMA, np.ones_like(b), b, OP_SOFTPLUS, k)
    if isinstance(rv, M.Chi2):
      (df,) = rv.param_values()
      return PriorFamily(FAMILY_GAMMA, 0.5 * df, np.full_like(df, 0.5), OP_SOFTPLUS, k)
    if isinstance(rv, M.InverseGamma):
      a, b = rv.param_values()
      return PriorFamily(FAMILY_INVGAMMA, a, b, OP_SOFTPLUS, k)
  if transform == 'sigmoid' and isinstance(rv, M.Beta):
    a, b = rv.param_values()
    return PriorFamily(FAMILY_BETA, a, b, OP_SIGMOID, k)
  raise NotImplementedError('no device prior for %s under the %s transform'
                            % (type(rv).__name__, transform))


def _prior_family(rv, k, transform):
  if transform is None:
    pr = _normal_prior(rv, k)
    if pr is None:
      raise NotImplementedError('prior parameters of %s must be constants' % rv.name)
    return PriorFamily(FAMILY_NORMAL, pr[0], pr[1], OP_IDENTITY, k)
  if transform == 'softmax_centered':
    raise NotImplementedError('simplex-valued latents are not supported by the device '
                              'samplers (the unconstrained position has one coordinate '
                              'fewer than the Empirical store)')
  return _constrained_family(rv, k, transform)


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


def _apply_spec(eng, spec):
  if spec.lik_weight != 1.0:
    eng.set_lik_weight(spec.lik_weight)
  if spec.prior_kind is not None:
    eng.set_prior_family(spec.prior_kind, spec.prior_lnorm)


def compile_model(latent_vars, data, precision='fp32', scale=None, transforms=None):
  """Bind the model to a device engine.  ``scale`` (Edward's per-variable
  multipliers, inference.py:187-193) is honoured by the stochastic-gradient
  samplers only; HMC passes None (hmc.py:161-192 ignores it).  ``transforms``
  maps auto-transformed latents to their map u -> theta ('softplus',
  'sigmoid')."""
  spec = glm_spec(latent_vars, data, scale, transforms)
  if not isinstance(spec, GLMSpec):
    return spec          # another engine family built its own program
  from edward_amd.engine import GLMEngine
  eng = GLMEngine(spec.X, spec.y, spec.loc, spec.scale, likelihood=spec.likelihood,
                  lik_scale=spec.lik_scale, precision=precision)
  _apply_spec(eng, spec)
  return Program(eng, spec.latents, spec.n_rows, spec.n_chains, spec.dim)


def rebind_model(program, latent_vars, data, scale=None, transforms=None):
  """Bind new data values (a fed minibatch) to a compiled program's engine;
  the latent layout must not change.  Discards the engine's chain state."""
  spec = glm_spec(latent_vars, data, scale, transforms)
  if not isinstance(spec, GLMSpec):
    raise NotImplementedError('feeding data is supported for the GLM family only')
  if spec.dim != program.dim or spec.n_chains != program.n_chains:
    raise ValueError('fed data changes the latent layout')
  eng = program.engine
  eng.bind(spec.X, spec.y, spec.loc, spec.scale, spec.likelihood, spec.lik_scale)
  _apply_spec(eng, spec)


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


def _probs_block(obs, rv, k, N):
  """Design block of ``Bernoulli(probs=theta)`` for a sigmoid-transformed
  latent theta of k coordinates: the logits are u, coordinate (row mod k)."""
  if N % k:
    raise ValueError('%d observations do not tile the %d probabilities' % (N, k))
  return np.tile(np.eye(k), (N // k, 1))


def glm_spec(latent_vars, data, scale=None, transforms=None):
  """Match the model to the GLM family; returns a GLMSpec (no device work)."""
  transforms = transforms or {}
  observed = [k for k in data if isinstance(k, M.RandomVariable)]
  if len(observed) > 1:
    raise NotImplementedError('exactly one observed random variable is supported, got %d'
                              % len(observed))
  latents = list(latent_vars.keys())
  probs_rv = None
  if not observed:
    # prior only: one zero row whose likelihood has weight 0
    obs, yv, E, likelihood, lik_scale, lik_weight = None, np.zeros(1), None, 'bernoulli', 1.0, 0.0
  else:
    obs = observed[0]
    yv = _observed_values(obs, data)
    lik_weight = _scalar_scale(scale, obs)
    if isinstance(obs, M.Bernoulli) and isinstance(obs.logits, M.Add) and \
            _has_gather(obs.logits):
      if scale:
        raise NotImplementedError('scale is not supported for the hierarchical model')
      from edward_amd import hier
      return hier.compile_hier(latent_vars, data, obs, yv, transforms)
    if isinstance(obs, M.Bernoulli):
      if obs.logits is None:
        if not (isinstance(obs.probs, M.RandomVariable) and obs.probs in latent_vars and
                transforms.get(obs.probs) == 'sigmoid'):
          raise NotImplementedError('Bernoulli(probs=...) is supported for a latent with '
                                    "'01' support under auto_transform; otherwise use logits")
        probs_rv, E = obs.probs, None
      else:
        E = obs.logits
      likelihood, lik_scale = 'bernoulli', 1.0
    elif isinstance(obs, M.Normal):
      s = _resolve(obs.scale, data)
      if s is None or np.ndim(s) > 1 or np.any(s != np.ravel(s)[0]):
        raise NotImplementedError('Normal likelihood needs one constant scale')
      E, likelihood, lik_scale = obs.loc, 'normal', float(np.ravel(s)[0])
    else:
      raise NotImplementedError('unsupported likelihood %s' % type(obs).__name__)
  terms = _linear_terms(E, data, set(latents)) if E is not None else []
  N = yv.size
  # Latent vector layout follows latent_vars' order.
  blocks, fams, lat = [], [], []
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
    fam = _prior_family(rv, k, transforms.get(rv))
    if fam.family[0] == FAMILY_NORMAL and fam.op[0] == OP_IDENTITY:
      # a scale factor w on log N(z; m, s) has the gradient of N(z; m, s / sqrt(w))
      # (the samplers that honour scale use gradients only)
      w = _scalar_scale(scale, rv)
      if w != 1.0:
        fam = PriorFamily(FAMILY_NORMAL, fam.p1, fam.p2 / np.sqrt(w), OP_IDENTITY, k)
    elif scale and rv in scale:
      raise NotImplementedError('scale factors on constrained latents are not supported')
    blk = np.zeros((N, k))
    if rv is probs_rv:
      blk = _probs_block(obs, rv, k, N)
      fam.op[:] = OP_IDENTITY       # logit(sigmoid(u)) = u
    for t in terms:
      if t.rv is rv:
        b = _block(t.cols, N)
        if b.shape[1] != k:
          raise ValueError('ed.dot operand of %d columns for latent of size %d'
                           % (b.shape[1], k))
        blk = blk + b
    blocks.append(blk)
    fams.append(fam)
    lat.append(Latent(rv, qz, k, off, C))
    off += k
  X = np.concatenate(blocks, axis=1).astype(np.float32)
  T = int(min(l.qz.params.shape[0] for l in lat))
  kind = np.concatenate([f.kind for f in fams])
  plain = not np.any(kind)
  return GLMSpec(X=X, y=yv.astype(np.float32), loc=np.concatenate([f.p1 for f in fams]),
                 scale=np.concatenate([f.p2 for f in fams]), likelihood=likelihood,
                 lik_scale=lik_scale, lik_weight=lik_weight, latents=lat, n_rows=T,
                 n_chains=n_chains, dim=off, prior_kind=None if plain else kind,
                 prior_lnorm=None if plain else np.concatenate([f.lnorm for f in fams]))


def _has_gather(E):
  if isinstance(E, M.Gather):
    return True
  if isinstance(E, M.Add):
    return _has_gather(E.a) or _has_gather(E.b)
  return False
