"""Criticism of fitted GLMs on the device: ``ed.copy``, ``ed.evaluate``, ``ed.ppc``.

Edward builds the posterior predictive with ``ed.copy(y, {w: qw})``
(``util/random_variables.py:139-445``) and evaluates it by running the copied
graph once per posterior draw (``criticisms/evaluate.py:128-150,199-204``,
``criticisms/ppc.py:100-121``).  Here the copy records the substitution, and
all draws go through one forward pass of the GLM kernel family
(``edhmc_predict``: eta = W X^T on the matrix cores, FP64 log-likelihood per
draw, mean response per row, replicated data from the engine's Philox stream).

Supported: outputs ``Bernoulli(logits=E)`` and ``Normal(loc=E, scale=s)``
with E linear in the latents (the models ``ed.HMC`` runs on the GLM engine).
Each draw takes one ``Empirical.sample()`` per latent, independently, as the
copied graph does.  ``ppc``'s discrepancy ``T(xs, zs)`` receives torch
tensors (device) where Edward's receives TF tensors.
"""

import numpy as np
import torch

from edward_amd import compile as _compile
from edward_amd import inferences as _inf
from edward_amd import models as M


class PosteriorPredictive(M.RandomVariable):
  """The value of ``ed.copy(y, dict_swap)`` for an observed GLM output y:
  y's distribution with each swapped latent replaced by its posterior."""

  def __init__(self, org, swap, name=None):
    super(PosteriorPredictive, self).__init__(name=name or (org.name + '_copied'))
    self.org = org
    self.swap = dict(swap)
    self.dtype = org.dtype
    self.batch_shape = org.batch_shape
    self.sample_shape = org.sample_shape

  @property
  def likelihood(self):
    return 'bernoulli' if isinstance(self.org, M.Bernoulli) else 'normal'


def copy(org_instance, dict_swap=None, scope='copied', replace_itself=False, copy_q=False,
         copy_parent_rvs=True):
  """``ed.copy`` (``util/random_variables.py:139-175``) for the posterior
  predictive of a GLM output: ``dict_swap`` maps latent variables to their
  posterior approximations (``Empirical``) or to constants."""
  if not isinstance(org_instance, (M.Bernoulli, M.Normal)):
    raise NotImplementedError('ed.copy supports the Bernoulli / Normal output of a GLM')
  dict_swap = dict(dict_swap or {})
  for k, v in dict_swap.items():
    if not isinstance(k, (M.RandomVariable, M.Node)):
      raise TypeError('dict_swap keys must be random variables or nodes')
  return PosteriorPredictive(org_instance, dict_swap)


# -- shared machinery --------------------------------------------------------

def _linear_predictor(out):
  org = out.org if isinstance(out, PosteriorPredictive) else out
  if isinstance(org, M.Bernoulli):
    if org.logits is None:
      raise NotImplementedError('Bernoulli(probs=...) outputs are not supported; use logits')
    return org, org.logits, 'bernoulli', 1.0
  if isinstance(org, M.Normal):
    s = _compile._resolve(org.scale, {})
    if s is None or np.ndim(s) > 1 or np.any(np.ravel(s) != np.ravel(s)[0]):
      raise NotImplementedError('Normal outputs need one constant scale')
    return org, org.loc, 'normal', float(np.ravel(s)[0])
  raise NotImplementedError('unsupported output %s' % type(org).__name__)


def _latents_of(E, data):
  seen = []

  def walk(e):
    if isinstance(e, M.RandomVariable):
      if e not in seen:
        seen.append(e)
      return
    for attr in ('a', 'b', 'x', 'y'):
      v = getattr(e, attr, None)
      if isinstance(v, (M.RandomVariable, M.Node)) and not isinstance(v, M.Placeholder):
        walk(v)
  walk(E)
  return seen


class _Design(object):
  """Design matrix of the output over the bound data and the latent layout."""

  def __init__(self, out, data):
    org, E, self.likelihood, self.scale = _linear_predictor(out)
    self.latents = _latents_of(E, data)
    terms = _compile._linear_terms(E, data, set(self.latents))
    n = None
    for t in terms:
      if not isinstance(t.cols, (str, tuple)):
        n = t.cols.shape[0]
    if n is None:
      shape = org.shape
      n = int(np.prod(shape)) if shape else 1
    blocks, self.sizes = [], []
    for rv in self.latents:
      k = int(np.prod(rv.shape)) if rv.shape else 1
      blk = np.zeros((n, k))
      for t in terms:
        if t.rv is rv:
          blk = blk + _compile._block(t.cols, n)
      blocks.append(blk)
      self.sizes.append(k)
    self.X = np.concatenate(blocks, axis=1).astype(np.float32)
    self.N = n


def _draws(design, swap, n_samples, rng, device):
  """[S, D] draws: one independent Empirical row per latent per draw (a
  constant swap value is used as is; a latent left unswapped is drawn from
  its Normal prior)."""
  parts = []
  for rv, k in zip(design.latents, design.sizes):
    q = swap.get(rv)
    if isinstance(q, M.Empirical):
      P = q.params.reshape(q.n, -1, k)           # [T, chains, k]
      t = torch.as_tensor(rng.integers(0, P.shape[0], n_samples), device=P.device)
      c = torch.as_tensor(rng.integers(0, P.shape[1], n_samples), device=P.device)
      parts.append(P[t, c].to(device, torch.float32))
    elif q is not None and not isinstance(q, M.RandomVariable):
      v = torch.as_tensor(np.broadcast_to(np.asarray(
          q.cpu() if isinstance(q, torch.Tensor) else q, np.float32).reshape(-1), (k,)))
      parts.append(v.to(device)[None].expand(n_samples, k))
    else:
      pr = _compile._normal_prior(rv, k)
      if pr is None:
        raise NotImplementedError('%s is neither swapped nor a constant-parameter Normal'
                                  % rv.name)
      z = rng.normal(size=(n_samples, k)) * pr[1] + pr[0]
      parts.append(torch.as_tensor(z, dtype=torch.float32, device=device))
  return torch.cat(parts, dim=1).contiguous()


def _engine(design, y_true):
  from edward_amd.engine import GLMEngine
  D = design.X.shape[1]
  return GLMEngine(design.X, np.asarray(y_true, np.float32).reshape(-1), np.zeros(D, np.float32),
                   np.ones(D, np.float32), likelihood=design.likelihood,
                   lik_scale=design.scale, precision='fp32')


def _placeholder_feed(data):
  return {k: v for k, v in data.items() if isinstance(k, M.Placeholder)}


# -- evaluate ------------------------------------------------------------------



def _metric(name, y_true, y_pred):
  """``evaluate.py:235-476`` on 1-D device tensors (macro averaging of
  ``util/metrics.py:10-47`` is a mean over the elements here)."""
  yt, yp = y_true.to(torch.float32), y_pred.to(torch.float32)
  if name == 'binary_accuracy':
    return (yt == yp).to(torch.float32).mean()
  if name in ('log_loss', 'binary_crossentropy'):
    return torch.nn.functional.binary_cross_entropy_with_logits(yp, yt)
  if name in ('mse', 'MSE', 'mean_squared_error'):
    return ((yp - yt) ** 2).mean()
  if name in ('mae', 'MAE', 'mean_absolute_error'):
    return (yp - yt).abs().mean()
  if name in ('mape', 'MAPE', 'mean_absolute_percentage_error'):
    return 100.0 * ((yt - yp) / yt.abs().clamp(min=1e-8)).abs().mean()
  if name in ('msle', 'MSLE', 'mean_squared_logarithmic_error'):
    return ((torch.log(yp.clamp(min=1e-8) + 1.0) - torch.log(yt.clamp(min=1e-8) + 1.0)) ** 2).mean()
  if name == 'hinge':
    return torch.clamp(1.0 - yt * yp, min=0.0).mean()
  if name == 'squared_hinge':
    return (torch.clamp(1.0 - yt * yp, min=0.0) ** 2).mean()
  if name == 'poisson':
    return (yp - yt * torch.log(yp + 1e-8)).sum()
  if name in ('cosine', 'cosine_proximity'):
    a = yt / torch.sqrt(torch.clamp((yt ** 2).sum(), min=1e-12))
    b = yp / torch.sqrt(torch.clamp((yp ** 2).sum(), min=1e-12))
    return (a * b).sum()
  raise NotImplementedError('Metric is not implemented: {}'.format(name))


def evaluate(metrics, data, n_samples=500, output_key=None, seed=None):
  """``ed.evaluate`` (``criticisms/evaluate.py:20-232``) for GLM outputs."""
  if isinstance(metrics, str) or callable(metrics):
    metrics = [metrics]
  elif not isinstance(metrics, list):
    raise TypeError('metrics must have type str or list, or be callable.')
  _inf.check_data(data)
  if not isinstance(n_samples, int):
    raise TypeError('n_samples must have type int.')
  if output_key is None:
    keys = [k for k in data if not isinstance(k, M.Placeholder)]
    if len(keys) == 1:
      output_key = keys[0]
    else:
      raise KeyError('User must specify output_key.')
  elif not isinstance(output_key, M.RandomVariable):
    raise TypeError('output_key must have type RandomVariable.')
  feed = _placeholder_feed(data)
  y_true_np = np.asarray(data[output_key].cpu() if isinstance(data[output_key], torch.Tensor)
                         else data[output_key], np.float32).reshape(-1)
  rng = np.random.default_rng(seed)
  const = _constant_output(output_key, y_true_np.size)
  if const is not None:
    return _evaluate_constant(metrics, const, y_true_np, n_samples, rng)
  design = _Design(output_key, feed)
  eng = _engine(design, y_true_np)
  swap = output_key.swap if isinstance(output_key, PosteriorPredictive) else {}
  W = _draws(design, swap, n_samples, rng, eng.device)
  supervised = [m for m in metrics if (m[0] if isinstance(m, tuple) else m) not in
                ('log_lik', 'log_likelihood')]
  y_true = torch.as_tensor(y_true_np, device=eng.device)
  need_rep = bool(supervised) and design.likelihood == 'normal'
  mean, ll, rep = eng.predict(W, mean=True, loglik=True, replicate=need_rep,
                              seed=int(rng.integers(0, 2 ** 63 - 1)))
  y_pred = None
  if supervised:
    if design.likelihood == 'bernoulli':
      # average the probabilities over the draws, round, random at exactly 0.5
      u = torch.as_tensor(rng.random(design.N), dtype=torch.float32, device=eng.device)
      y_pred = torch.round(torch.where(mean == 0.5, u, mean))
    else:
      y_pred = rep.mean(dim=0)     # Monte Carlo mean of the posterior predictive
  out = []
  for m in metrics:
    name, params = (m if isinstance(m, tuple) else (m, {}))
    if name in ('accuracy', 'crossentropy'):
      if float(y_true.max()) <= 1:
        name = 'binary_' + name
      else:
        raise NotImplementedError('sparse categorical metrics need a categorical output')
    if name in ('log_lik', 'log_likelihood'):
      # mean over draws of the mean row log density (evaluate.py:199-204)
      out.append(float((ll / design.N).mean()))
    elif callable(name):
      out.append(name(y_true, y_pred, **params))
    else:
      if params.get('average', 'macro') not in (None, 'macro', 'micro'):
        raise ValueError('average has to be one of (None, micro, macro)')
      out.append(float(_metric(name, y_true, y_pred)))
  return out[0] if len(out) == 1 else out


def _constant_output(out, n):
  """(likelihood, parameter [n], scale) of an output whose parameters are
  constants (no latent to draw), else None."""
  org = out.org if isinstance(out, PosteriorPredictive) else out
  def const(v):
    if v is None or isinstance(v, M.Node):
      return None
    return np.broadcast_to(np.asarray(v.cpu() if isinstance(v, torch.Tensor) else v,
                                      np.float64).reshape(-1) if np.ndim(v) else
                           np.asarray(v, np.float64), (n,)).astype(np.float64)
  if isinstance(org, M.Bernoulli):
    if org.probs is not None and const(org.probs) is not None:
      return ('bernoulli', const(org.probs), 1.0)
    if org.logits is not None and const(org.logits) is not None:
      return ('bernoulli', 1.0 / (1.0 + np.exp(-const(org.logits))), 1.0)
    return None
  if isinstance(org, M.Normal):
    loc, sc = const(org.loc), const(org.scale)
    if loc is not None and sc is not None:
      return ('normal', loc, sc)
  return None


def _evaluate_constant(metrics, const, y_true_np, n_samples, rng):
  """evaluate() for an output with constant parameters: the predictions need
  no posterior draws (Bernoulli: round the probabilities; Normal: the mean of
  n_samples draws, ``evaluate.py:146-150``)."""
  lik, par, sc = const
  dev = 'cuda' if torch.cuda.is_available() else 'cpu'
  y_true = torch.as_tensor(y_true_np, device=dev)
  p = torch.as_tensor(par, dtype=torch.float32, device=dev)
  if lik == 'bernoulli':
    u = torch.as_tensor(rng.random(p.shape), dtype=torch.float32, device=dev)
    y_pred = torch.round(torch.where(p == 0.5, u, p))
    ll_rows = y_true * torch.log(p) + (1 - y_true) * torch.log1p(-p)
  else:
    s = torch.as_tensor(sc, dtype=torch.float32, device=dev)
    z = torch.as_tensor(rng.normal(size=(n_samples,) + tuple(p.shape)), dtype=torch.float32,
                        device=dev)
    y_pred = (p + s * z).mean(dim=0)
    ll_rows = -0.5 * ((y_true - p) / s) ** 2 - torch.log(s) - 0.5 * np.log(2 * np.pi)
  out = []
  for m in metrics:
    name, params = (m if isinstance(m, tuple) else (m, {}))
    if name in ('accuracy', 'crossentropy'):
      name = 'binary_' + name
    if name in ('log_lik', 'log_likelihood'):
      out.append(float(ll_rows.mean()))
    elif callable(name):
      out.append(name(y_true, y_pred, **params))
    else:
      out.append(float(_metric(name, y_true, y_pred)))
  return out[0] if len(out) == 1 else out


# -- ppc -----------------------------------------------------------------------

def ppc(T, data, latent_vars=None, n_samples=100, seed=None):
  """``ed.ppc`` (``criticisms/ppc.py:13-121``): [T(x_rep_s, z_s)]_s and
  [T(x, z_s)]_s over n_samples posterior draws, as numpy arrays."""
  if not callable(T):
    raise TypeError('T must be a callable function.')
  _inf.check_data(data)
  if latent_vars is None:
    latent_vars = {}
  _inf.check_latent_vars(latent_vars)
  if not isinstance(n_samples, int):
    raise TypeError('n_samples must have type int.')
  feed = _placeholder_feed(data)
  outs = [k for k in data if isinstance(k, M.RandomVariable)]
  if len(outs) != 1:
    raise NotImplementedError('ppc supports one observed output')
  out = outs[0]
  design = _Design(out, feed)
  y_obs = np.asarray(data[out].cpu() if isinstance(data[out], torch.Tensor) else data[out],
                     np.float32).reshape(-1)
  rng = np.random.default_rng(seed)
  eng = _engine(design, y_obs)
  swap = dict(out.swap) if isinstance(out, PosteriorPredictive) else {}
  for k, v in latent_vars.items():
    swap.setdefault(k, v)
  W = _draws(design, swap, n_samples, rng, eng.device)
  _, _, rep = eng.predict(W, mean=False, loglik=False, replicate=True,
                          seed=int(rng.integers(0, 2 ** 63 - 1)))
  dev = eng.device
  xobs = {k: (torch.as_tensor(np.asarray(v), device=dev) if not isinstance(v, torch.Tensor)
              else v.to(dev)) for k, v in data.items()}
  offs = np.concatenate([[0], np.cumsum(design.sizes)])
  Treps, Ts = [], []
  for s in range(n_samples):
    zs = {}
    for i, rv in enumerate(design.latents):
      if rv in latent_vars:
        zs[rv] = W[s, offs[i]:offs[i + 1]].reshape(rv.shape if rv.shape else ())
    xrep = dict(xobs)
    xrep[out] = rep[s]
    Treps.append(np.asarray(_to_numpy(T(xrep, zs))))
    Ts.append(np.asarray(_to_numpy(T(xobs, zs))))
  return [np.stack(Treps), np.stack(Ts)]


def _to_numpy(v):
  if isinstance(v, torch.Tensor):
    return v.detach().cpu().numpy()
  return v
