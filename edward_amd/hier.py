"""Hierarchical logistic regression engine (batched over chains).

Model (latent vector per chain ``[a (G), mu, u, w (D)]``)::

  mu ~ Normal(0, 1);  log_sigma ~ Normal(m, s)   (u = log_sigma)
  a  ~ Normal(mu, exp(log_sigma))          # one intercept per group, shape [G]
  w  ~ Normal(0, 1)                        # shape [D]
  y  ~ Bernoulli(logits=ed.gather(a, group) + ed.dot(X, w))

or, with a positive group scale auto-transformed (u = softplus^-1(sigma),
``inference.py:222-264``), ``a ~ Normal(mu, sigma)`` with ``sigma`` ~ Gamma,
Exponential, Chi2, InverseGamma, or a Normal pushed through Softplus.

This is the gather model of ``tests/inferences/hmc_test.py:105-123`` grown to
a multilevel regression; its log joint is ``hmc.py:182-192``'s sum and the
gather's IndexedSlices gradient is densified as ``hmc.py:203,208`` do.  Rows
are reordered by group on the host once, so each group's rows form one
contiguous range ``offsets[g]:offsets[g+1]`` of the device design matrix.
"""

import numpy as np
import torch

from edward_amd import _lib as L
from edward_amd import models as M
from edward_amd.engine import Context, _dev_tensor


class HierEngine(Context):
  """HMC over the hierarchical logistic regression for C chains.

  Args:
    X: [N, D] design matrix (rows sorted by group).
    y: [N] 0/1 observations.
    offsets: [G + 1] int64 row offsets of the groups (0 ... N).
  """

  def __init__(self, X, y, offsets, device=0):
    super(HierEngine, self).__init__(device)
    dev = self.device
    Xd = _dev_tensor(X, torch.float32, dev)
    if Xd.dim() != 2:
      raise ValueError('X must be 2-D, got shape %s' % (tuple(Xd.shape),))
    N, D = Xd.shape
    yd = _dev_tensor(y, torch.float32, dev).reshape(-1)
    off = _dev_tensor(offsets, torch.int64, dev).reshape(-1)
    if yd.numel() != N:
      raise ValueError('y has %d entries for %d rows of X' % (yd.numel(), N))
    self.N, self.D, self.G = int(N), int(D), int(off.numel()) - 1
    self.dim = self.G + 2 + self.D
    self.precision = 'fp32'
    self.C = None
    self.check(self.lib.edhmc_bind_hier(self.h, L.ptr(Xd), L.ptr(yd), L.ptr(off), self.G, N, D,
                                        self.stream))
    torch.cuda.synchronize(dev)

  def set_scale_prior(self, kind, p1, p2, lnorm):
    """Prior of the group-scale coordinate (``edhmc_hier_set_scale_prior``):
    kind = family | map << 4 (map 0: sigma = exp(u), 1: sigma = softplus(u))."""
    self.check(self.lib.edhmc_hier_set_scale_prior(self.h, int(kind), float(p1), float(p2),
                                                   float(lnorm)))

  def _theta(self, theta):
    t = _dev_tensor(theta, torch.float32, self.device)
    if t.dim() == 1:
      t = t[None]
    if t.shape[1] != self.dim:
      raise ValueError('theta has %d coordinates, model has %d' % (t.shape[1], self.dim))
    return t

  def set_state(self, theta):
    t = self._theta(theta)
    self.C = int(t.shape[0])
    self.check(self.lib.edhmc_hier_set_state(self.h, L.ptr(t), self.C, self.stream))

  def logp_grad(self, theta):
    t = self._theta(theta)
    C = int(t.shape[0])
    self.C = C
    g = torch.empty(C, self.dim, device=self.device)
    lp = torch.empty(C, dtype=torch.float64, device=self.device)
    self.check(self.lib.edhmc_hier_logp_grad(self.h, L.ptr(t), C, L.ptr(g), L.ptr(lp),
                                             self.stream))
    return lp, g

  def transition(self, step_size, n_steps, momenta=None, uniforms=None, seed=0, counter=0,
                 sample_out=None, n_accept=None, accepted_out=None):
    r0 = None if momenta is None else _dev_tensor(momenta, torch.float32, self.device)
    u = None if uniforms is None else _dev_tensor(uniforms, torch.float64, self.device)
    self.check(self.lib.edhmc_hier_transition(
        self.h, float(step_size), int(n_steps), L.ptr(r0), L.ptr(u), int(seed), int(counter),
        L.ptr(accepted_out), L.ptr(sample_out), L.ptr(n_accept), self.stream))

  def divergences(self):
    out = torch.empty(self.C, dtype=torch.int32, device=self.device)
    self.check(self.lib.edhmc_divergences(self.h, L.ptr(out), self.stream))
    return out

  def last_ratio(self):
    """FP64 Metropolis log ratio of every chain's last transition, [C]."""
    out = torch.empty(self.C, dtype=torch.float64, device=self.device)
    self.check(self.lib.edhmc_last_ratio(self.h, L.ptr(out), self.stream))
    return out


def sort_by_group(X, y, group, G=None):
  """Stable reorder of the rows by group; returns (X, y, offsets[G+1])."""
  group = np.asarray(group).reshape(-1).astype(np.int64)
  if G is None:
    G = int(group.max()) + 1 if group.size else 0
  if group.size and (group.min() < 0 or group.max() >= G):
    raise ValueError('group indices must lie in [0, %d)' % G)
  order = np.argsort(group, kind='stable')
  counts = np.bincount(group, minlength=G)
  offsets = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
  return np.asarray(X)[order], np.asarray(y).reshape(-1)[order], offsets


def _is_std_normal(rv):
  def const(v, c):
    if isinstance(v, M.Node):
      return False
    a = np.asarray(v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else v, np.float64)
    return bool(np.all(a == c))
  return isinstance(rv, M.Normal) and const(rv.loc, 0.0) and const(rv.scale, 1.0)


def _scale_prior(sc, latent_vars, transforms):
  """(latent of the group-scale coordinate, kind, p1, p2, lnorm) of the
  scale argument of the intercepts' Normal."""
  from edward_amd import compile as Cm
  if isinstance(sc, M.Exp) and isinstance(sc.x, M.Normal):
    ls = sc.x
    if transforms.get(ls) is not None:
      raise NotImplementedError('log_sigma must have real support')
    loc, scale = Cm._resolve(ls.loc, {}), Cm._resolve(ls.scale, {})
    if loc is None or scale is None or np.size(loc) != 1 or np.size(scale) != 1:
      raise NotImplementedError('log_sigma needs a scalar Normal prior with constant '
                                'parameters')
    fam = Cm.PriorFamily(Cm.FAMILY_NORMAL, np.ravel(loc), np.ravel(scale), 0, 1)
    return ls, int(fam.family[0]), float(fam.p1[0]), float(fam.p2[0]), float(fam.lnorm[0])
  if isinstance(sc, M.RandomVariable) and sc in latent_vars:
    if transforms.get(sc) != 'softplus':
      raise NotImplementedError('a positive group scale must be auto-transformed '
                                '(initialize with auto_transform=True)')
    if sc.shape:
      raise NotImplementedError('the group scale must be a scalar')
    fam = Cm._constrained_family(sc, 1, 'softplus')
    return (sc, int(fam.family[0]) | (1 << 4), float(fam.p1[0]), float(fam.p2[0]),
            float(fam.lnorm[0]))
  raise NotImplementedError('a must be Normal(mu, exp(log_sigma)) with log_sigma ~ Normal, '
                            'or Normal(mu, sigma) with a positive latent sigma')


def compile_hier(latent_vars, data, obs, yv, transforms=None):
  """Match ``Bernoulli(logits=gather(a, group) + dot(X, w))`` and build the
  engine; raises NotImplementedError for any other gather model."""
  from edward_amd.compile import Latent, Program, _array, _chains_of, _resolve
  transforms = transforms or {}
  E = obs.logits
  parts = [E.a, E.b]
  ga = [p for p in parts if isinstance(p, M.Gather)]
  dt = [p for p in parts if isinstance(p, M.Dot)]
  if len(ga) != 1 or len(dt) != 1:
    raise NotImplementedError('the gather model must be gather(a, group) + dot(X, w)')
  ga, dt = ga[0], dt[0]
  a = ga.params
  group = _resolve(ga.indices, data)
  if group is None or not isinstance(a, M.Normal):
    raise NotImplementedError('gather(a, group) needs a Normal latent a and data indices')
  if isinstance(dt.y, M.RandomVariable):
    X, w = _resolve(dt.x, data), dt.y
  else:
    X, w = _resolve(dt.y, data), dt.x
    X = None if X is None else X.T
  if X is None or X.ndim != 2 or not _is_std_normal(w):
    raise NotImplementedError('dot(X, w) needs a data matrix and w ~ Normal(0, 1)')
  mu = a.loc
  if not (isinstance(mu, M.Normal) and _is_std_normal(mu)):
    raise NotImplementedError('the intercepts\' mean must be mu ~ Normal(0, 1)')
  for rv in (a, mu, w):
    if transforms.get(rv) is not None:
      raise NotImplementedError('%s must have real support' % rv.name)
  ls, skind, sp1, sp2, slnorm = _scale_prior(a.scale, latent_vars, transforms)
  for rv in (a, mu, ls, w):
    if rv not in latent_vars:
      raise NotImplementedError('%s must be a latent variable of the inference' % rv.name)
  if len(latent_vars) != 4:
    raise NotImplementedError('the hierarchical model has exactly the latents a, mu, '
                              'the group scale, w')
  G = int(np.prod(a.shape))
  D = int(np.prod(w.shape))
  if X.shape[1] != D:
    raise ValueError('X has %d columns, w has %d coordinates' % (X.shape[1], D))
  Xs, ys, offsets = sort_by_group(X, yv, group.astype(np.int64), G)
  lat, off, n_chains = [], 0, None
  for rv in (a, mu, ls, w):
    qz = latent_vars[rv]
    k = int(np.prod(rv.shape)) if rv.shape else 1
    C = _chains_of(rv, qz)
    if n_chains is None:
      n_chains = C
    elif C != n_chains:
      raise ValueError('all Empirical stores must hold the same number of chains')
    lat.append(Latent(rv, qz, k, off, C))
    off += k
  eng = HierEngine(Xs.astype(np.float32), ys.astype(np.float32), offsets)
  eng.set_scale_prior(skind, sp1, sp2, slnorm)
  T = int(min(l.qz.params.shape[0] for l in lat))
  return Program(eng, lat, T, n_chains, off)
