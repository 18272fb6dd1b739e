"""Readout of Empirical sample stores (``edward/models/empirical.py:87-96``).

Device stores are reduced by ``edhmc_sample_moments`` (two-pass FP64 on the
GPU); a store the user keeps in host memory is reduced on the host in FP64,
since it never reaches the device."""

import numpy as np
import torch

from edward_amd import _lib as L

_ctx = [None]


def _context(device):
  from edward_amd.engine import Context
  if _ctx[0] is None or _ctx[0].device != device:
    _ctx[0] = Context(device.index or 0)
  return _ctx[0]


def empirical_moments(params, burn_in=0):
  """(mean, variance) over axis 0 of params [T, ...] for rows t >= burn_in,
  divisor T - burn_in; returned in fp64 with the event shape."""
  T = int(params.shape[0])
  if not 0 <= burn_in < T:
    raise ValueError('burn_in %d outside [0, %d)' % (burn_in, T))
  ev = tuple(params.shape[1:])
  if not params.is_cuda:
    a = params.detach().cpu().numpy().astype(np.float64)[burn_in:]
    return (torch.from_numpy(a.mean(0).reshape(ev)),
            torch.from_numpy(a.var(0).reshape(ev)))
  ctx = _context(params.device)
  x = params.contiguous()
  if x.dtype == torch.float32:
    dt = L.F32
  elif x.dtype == torch.float64:
    dt = 2
  else:
    x, dt = x.float(), L.F32
  K = int(np.prod(ev)) if ev else 1
  mean = torch.empty(K, dtype=torch.float64, device=x.device)
  var = torch.empty(K, dtype=torch.float64, device=x.device)
  ctx.check(ctx.lib.edhmc_sample_moments(ctx.h, L.ptr(x), dt, T, K, int(burn_in), L.ptr(mean),
                                         L.ptr(var), ctx.stream))
  return mean.reshape(ev), var.reshape(ev)
