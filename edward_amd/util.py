"""Session and seed helpers (``edward/util/graphs.py:13-73``)."""

import numpy as np
import torch

from edward_amd import models as M

_SEED = [None]
_SESSION = [None]


class Session(object):
  """Stand-in for the global ``InteractiveSession``: names the device the
  engines run on."""

  def __init__(self, device=0):
    self.device = torch.device('cuda', device) if torch.cuda.is_available() else None


def get_session():
  """``graphs.py:13-41``: one global session."""
  if _SESSION[0] is None:
    _SESSION[0] = Session()
  return _SESSION[0]


def set_seed(x):
  """``graphs.py:59-73``: seed numpy and the engine's device RNG; like the
  reference it refuses once model nodes exist."""
  if M._created[0] > 0:
    raise RuntimeError('Seeding is not supported after initializing part of the graph. '
                       'Please move set_seed to the beginning of your code.')
  np.random.seed(int(x) & 0xFFFFFFFF)
  _SEED[0] = int(x)


def get_seed():
  return 0 if _SEED[0] is None else _SEED[0]


def transform(x, *args, **kwargs):
  """``ed.util.transform`` (``util/random_variables.py:856-917``)."""
  from edward_amd.bijectors import transform as _transform
  return _transform(x, *args, **kwargs)
