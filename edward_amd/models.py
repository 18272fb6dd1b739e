"""Random variables for the HMC surface, without TensorFlow.

Edward 1.3.5 wraps every ``tf.contrib.distributions`` class as a
``RandomVariable`` that is also a graph node
(``edward/models/random_variable.py:18-138``, ``random_variables.py:12-43``),
and HMC re-evaluates the model by copying that graph at each proposal
(``edward/util/random_variables.py:139-445``).  Here a model is a small
expression DAG over the same objects (``Normal``, ``Bernoulli``, ``Empirical``,
``placeholder``, ``dot``, ``+``, ``*``); ``edward_amd.compile`` matches it to a
kernel family instead of copying it.  ``support`` values follow
``random_variables.py:30-43`` and ``empirical.py:124``.
"""

import itertools

import numpy as np
import torch

_ids = itertools.count()
_created = [0]


def _as_array(x):
  if isinstance(x, torch.Tensor):
    return x.detach().cpu().numpy()
  return np.asarray(x)


def _dtype_name(dtype):
  if dtype is None:
    return 'float32'
  if isinstance(dtype, torch.dtype):
    return {torch.float32: 'float32', torch.float64: 'float64'}.get(dtype, str(dtype))
  return np.dtype(dtype).name


class Node(object):
  """Base of the expression DAG."""

  def __add__(self, other):
    return Add(self, other)

  def __radd__(self, other):
    return Add(other, self)

  def __mul__(self, other):
    return Mul(self, other)

  def __rmul__(self, other):
    return Mul(other, self)


class Placeholder(Node):
  """Stand-in for ``tf.placeholder(dtype, shape)``: a data slot bound through
  the inference's ``data`` dict."""

  def __init__(self, dtype='float32', shape=None, name=None):
    self.dtype = _dtype_name(dtype)
    self.shape = None if shape is None else tuple(int(s) for s in shape)
    self.name = name or 'Placeholder_%d' % next(_ids)

  def __repr__(self):
    return '<placeholder %s %s>' % (self.name, self.shape)


def placeholder(dtype='float32', shape=None, name=None):
  return Placeholder(dtype, shape, name)


class Dot(Node):
  """``ed.dot(x, y)`` (``edward/util/tensorflow.py:10-45``): matrix-vector
  product; which operand is the matrix is decided by rank."""

  def __init__(self, x, y):
    self.x, self.y = x, y


class Add(Node):
  def __init__(self, a, b):
    self.a, self.b = a, b


class Mul(Node):
  def __init__(self, a, b):
    self.a, self.b = a, b


class Gather(Node):
  """``tf.gather(params, indices)`` of a latent vector by a data index vector
  (the group-intercept lookup of hierarchical models)."""

  def __init__(self, params, indices):
    self.params, self.indices = params, indices


def gather(params, indices):
  return Gather(params, indices)


class Exp(Node):
  """``tf.exp`` of a latent (a scale parameterised on the log scale)."""

  def __init__(self, x):
    self.x = x


def exp(x):
  return Exp(x)


def dot(x, y):
  """``ed.dot``: the finite check of ``util/tensorflow.py:33-36`` happens when
  data are bound (non-finite data raise ``NonFiniteError``), and on constant
  operands immediately."""
  for v in (x, y):
    if not isinstance(v, Node):
      a = _as_array(v)
      if a.dtype.kind == 'f' and not np.all(np.isfinite(a)):
        raise ValueError('ed.dot: operand has non-finite values')
  return Dot(x, y)


class RandomVariable(Node):
  """Base random variable: keeps its constructor arguments (as Edward does for
  ``copy``, ``random_variable.py:101-110``), its shape and dtype."""

  support = None

  def __init__(self, *args, **kwargs):
    self._args = args
    self._kwargs = dict(kwargs)
    ss = kwargs.pop('sample_shape', ())
    self.sample_shape = tuple(int(s) for s in np.atleast_1d(ss)) if np.size(ss) else ()
    self.name = kwargs.pop('name', None) or '%s_%d' % (type(self).__name__, next(_ids))
    _created[0] += 1

  @property
  def shape(self):
    return tuple(self.sample_shape) + tuple(self.batch_shape) + tuple(self.event_shape)

  @property
  def event_shape(self):
    return ()

  def get_shape(self):
    return self.shape

  def __repr__(self):
    return '<ed.RandomVariable %r shape=%s dtype=%s>' % (self.name, self.shape, self.dtype)

  def __hash__(self):
    return id(self)

  def __eq__(self, other):
    return self is other


def _param_shape(p):
  if isinstance(p, RandomVariable):
    return p.shape
  if isinstance(p, Node):
    return None
  return np.shape(_as_array(p))


def _param_dtype(p):
  if isinstance(p, RandomVariable) or isinstance(p, Placeholder):
    return p.dtype
  if isinstance(p, Node) or isinstance(p, (bool, int, float)):
    return None   # Python scalars take the default float32, as TF converts them
  a = _as_array(p)
  return a.dtype.name if a.dtype.kind == 'f' else None


class Normal(RandomVariable):
  """Normal(loc, scale); ``support = 'real'`` (``random_variables.py:42``)."""

  support = 'real'

  def __init__(self, loc=0.0, scale=1.0, sample_shape=(), name=None, **kw):
    super(Normal, self).__init__(loc=loc, scale=scale, sample_shape=sample_shape, name=name)
    self.loc, self.scale = loc, scale
    dts = [d for d in (_param_dtype(loc), _param_dtype(scale)) if d]
    self.dtype = kw.get('dtype') or (dts[0] if dts else 'float32')
    shapes = [s for s in (_param_shape(loc), _param_shape(scale)) if s is not None]
    self.batch_shape = tuple(np.broadcast_shapes(*shapes)) if shapes else ()

  def log_prob_np(self, x, loc=None, scale=None):
    loc = _as_array(self.loc) if loc is None else loc
    scale = _as_array(self.scale) if scale is None else scale
    x = np.asarray(x, np.float64)
    return -0.5 * ((x - loc) / scale) ** 2 - np.log(scale) - 0.5 * np.log(2 * np.pi)


class Bernoulli(RandomVariable):
  """Bernoulli(logits=...) or Bernoulli(probs=...); ``support = 'binary'``
  (``random_variables.py:30``)."""

  support = 'binary'

  def __init__(self, logits=None, probs=None, sample_shape=(), name=None, dtype='int32'):
    super(Bernoulli, self).__init__(logits=logits, probs=probs, sample_shape=sample_shape,
                                    name=name)
    if (logits is None) == (probs is None):
      raise ValueError('Bernoulli needs exactly one of logits and probs')
    self.logits, self.probs = logits, probs
    self.dtype = _dtype_name(dtype) if not isinstance(dtype, str) else dtype
    p = logits if logits is not None else probs
    s = _param_shape(p)
    self.batch_shape = s if s is not None else ()


class _Parametric(RandomVariable):
  """A distribution given by named parameters (constants or nodes); batch
  shape is their broadcast shape, dtype the first floating parameter's."""

  _params = ()

  def __init__(self, *args, **kwargs):
    names = self._params
    if len(args) > len(names):
      raise TypeError('%s takes %d parameters' % (type(self).__name__, len(names)))
    vals = dict(zip(names, args))
    for n in names:
      if n in kwargs:
        vals[n] = kwargs.pop(n)
      if n not in vals:
        raise TypeError('%s needs the parameter %r' % (type(self).__name__, n))
    dtype = kwargs.pop('dtype', None)
    ss = kwargs.pop('sample_shape', ())
    name = kwargs.pop('name', None)
    for k in ('validate_args', 'allow_nan_stats'):
      kwargs.pop(k, None)
    if kwargs:
      raise TypeError('unexpected arguments %s' % sorted(kwargs))
    super(_Parametric, self).__init__(sample_shape=ss, name=name, **vals)
    for n in names:
      setattr(self, n, vals[n])
    dts = [d for d in (_param_dtype(vals[n]) for n in names) if d]
    self.dtype = dtype or (dts[0] if dts else 'float32')
    shapes = [sh for sh in (_param_shape(vals[n]) for n in names) if sh is not None]
    self.batch_shape = tuple(np.broadcast_shapes(*shapes)) if shapes else ()

  def param_values(self):
    """The parameters as float64 arrays of the batch shape (constants only)."""
    out = []
    for n in self._params:
      v = getattr(self, n)
      if isinstance(v, Node):
        raise NotImplementedError('parameter %s of %s must be a constant' % (n, self.name))
      out.append(np.broadcast_to(_as_array(v).astype(np.float64), self.batch_shape))
    return out


class Gamma(_Parametric):
  """Gamma(concentration, rate); ``support = 'nonnegative'``
  (``random_variables.py:37``)."""

  support = 'nonnegative'
  _params = ('concentration', 'rate')


class Exponential(_Parametric):
  """Exponential(rate) = Gamma(1, rate); ``support = 'nonnegative'``
  (``random_variables.py:36``)."""

  support = 'nonnegative'
  _params = ('rate',)


class Chi2(_Parametric):
  """Chi2(df) = Gamma(df / 2, 1 / 2); ``support = 'nonnegative'``
  (``random_variables.py:34``)."""

  support = 'nonnegative'
  _params = ('df',)


class InverseGamma(_Parametric):
  """InverseGamma(concentration, rate); ``support = 'nonnegative'``
  (``random_variables.py:38``)."""

  support = 'nonnegative'
  _params = ('concentration', 'rate')


class Beta(_Parametric):
  """Beta(concentration1, concentration0); ``support = '01'``
  (``random_variables.py:31``)."""

  support = '01'
  _params = ('concentration1', 'concentration0')


class TransformedDistribution(RandomVariable):
  """``TransformedDistribution(distribution, bijector)``: the law of
  ``bijector.forward(x)`` for ``x ~ distribution``.  It has no ``support``
  unless one is assigned (as Edward's generated class, which is not in the
  support table of ``random_variables.py:29-43``; ``util/random_variables.py:
  856-917`` sets it on the results of ``transform``)."""

  def __init__(self, distribution, bijector, name=None, sample_shape=(), **kw):
    super(TransformedDistribution, self).__init__(distribution=distribution, bijector=bijector,
                                                  name=name, sample_shape=sample_shape)
    self.distribution, self.bijector = distribution, bijector
    self.dtype = getattr(distribution, 'dtype', 'float32')
    base = tuple(getattr(distribution, 'batch_shape', ())) + \
        tuple(getattr(distribution, 'event_shape', ()))
    self.batch_shape = tuple(bijector.forward_event_shape(base))


class Empirical(RandomVariable):
  """Sample store ``Empirical(params[T, ...])`` (``edward/models/empirical.py``).

  ``params`` lives on the device as a torch tensor (a numpy array or tensor is
  copied there).  Iteration t of HMC writes ``params[t]``.  ``mean`` and
  ``stddev`` reduce over axis 0 with divisor T (``empirical.py:87-96``); a
  ``[T, C, D]`` store for a ``[D]`` latent holds C independent chains and its
  moments are per chain ``[C, D]``.  ``support = 'points'``.
  """

  support = 'points'

  def __init__(self, params, name=None, device=None):
    super(Empirical, self).__init__(params=params, name=name)
    if isinstance(params, torch.Tensor):
      t = params
    else:
      t = torch.as_tensor(np.asarray(params))
    if not t.is_floating_point():
      t = t.float()
    if torch.cuda.is_available():
      t = t.to(device or 'cuda')
    self._params = t.contiguous()
    self.dtype = _dtype_name(self._params.dtype)
    self.batch_shape = ()

  @property
  def params(self):
    return self._params

  @params.setter
  def params(self, value):
    self._params = value

  @property
  def n(self):
    return int(self._params.shape[0])

  @property
  def event_shape(self):
    return tuple(self._params.shape[1:])

  @property
  def shape(self):
    return self.event_shape

  def _moments(self, burn_in=0):
    from edward_amd import readout
    return readout.empirical_moments(self._params, burn_in)

  def mean(self, burn_in=0):
    """Mean over samples t >= burn_in (``empirical.py:87-88``)."""
    return self._moments(burn_in)[0].to(self._params.dtype)

  def variance(self, burn_in=0):
    return self._moments(burn_in)[1].to(self._params.dtype)

  def stddev(self, burn_in=0):
    """Population standard deviation, divisor T (``empirical.py:90-93``)."""
    return torch.sqrt(self._moments(burn_in)[1]).to(self._params.dtype)

  def sample(self, n=1, seed=None):
    """Uniform draws of stored rows (``empirical.py:98-110``)."""
    g = torch.Generator(device='cpu')
    if seed is not None:
      g.manual_seed(int(seed))
    idx = torch.randint(0, self.n, (int(n),), generator=g).to(self._params.device)
    return self._params.index_select(0, idx)


class ConstrainedEmpirical(TransformedDistribution):
  """An ``Empirical`` of unconstrained samples seen through the inverse of an
  auto-transform (``inference.py:240-255``): ``params`` are the constrained
  samples ``bijector.forward(qz.params)``, recomputed from the store on every
  access, with the store's ``mean``/``variance``/``stddev``/``sample`` over
  them."""

  def __init__(self, qz_unconstrained, z_bijector, name=None):
    from edward_amd import bijectors
    super(ConstrainedEmpirical, self).__init__(qz_unconstrained,
                                               bijectors.Invert(z_bijector), name=name)
    self.batch_shape = ()

  @property
  def params(self):
    p = self.distribution.params
    return self.bijector.forward(p).to(p.dtype)

  @property
  def n(self):
    return self.distribution.n

  @property
  def event_shape(self):
    return self.distribution.event_shape

  def mean(self, burn_in=0):
    return self.params[burn_in:].double().mean(0).to(self.distribution.params.dtype)

  def variance(self, burn_in=0):
    p = self.params[burn_in:].double()
    return ((p - p.mean(0)) ** 2).mean(0).to(self.distribution.params.dtype)

  def stddev(self, burn_in=0):
    return torch.sqrt(self.variance(burn_in))

  def sample(self, n=1, seed=None):
    return self.bijector.forward(self.distribution.sample(n, seed))


def reset_graph():
  """Forget that random variables were created (lets ``set_seed`` run again)."""
  _created[0] = 0
