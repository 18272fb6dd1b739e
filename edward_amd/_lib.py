"""ctypes binding of libedhmc.so, the C-ABI declared in include/edhmc.h.

torch is imported first: the HIP runtime it loads (SONAME libamdhip64.so.7)
is then the one the library resolves to, so torch streams and device
pointers are valid inside the library.  There is no fallback: if the library
is missing or a call fails, an exception is raised.
"""

import ctypes
import os

import torch  # noqa: F401  (must load the HIP runtime before the library)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libedhmc.so')

OK = 0
EINVAL = -1
ENOMEM = -2
EHIP = -3
EUNSUPPORTED = -4
ESTATE = -5
ENONFINITE = -6

F32 = 0
BF16 = 1
LIK_BERNOULLI = 0
LIK_NORMAL = 1

_c_p = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_u64 = ctypes.c_uint64
_f32 = ctypes.c_float

# name -> (restype, argtypes); mirrors include/edhmc.h
SIGNATURES = {
    'edhmc_version': (ctypes.c_char_p, []),
    'edhmc_create': (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(_c_p)]),
    'edhmc_destroy': (ctypes.c_int, [_c_p]),
    'edhmc_last_error': (ctypes.c_char_p, [_c_p]),
    'edhmc_bind_glm': (ctypes.c_int, [_c_p, _c_p, _i32, _c_p, _i64, _i32, _i32, _f32,
                                      _c_p, _c_p, _i32, _c_p]),
    'edhmc_set_state': (ctypes.c_int, [_c_p, _c_p, _i32, _c_p]),
    'edhmc_get_state': (ctypes.c_int, [_c_p, _c_p, _c_p, _c_p, _c_p]),
    'edhmc_logp_grad': (ctypes.c_int, [_c_p, _c_p, _i32, _c_p, _c_p, _c_p]),
    'edhmc_logp': (ctypes.c_int, [_c_p, _c_p, _i32, _c_p, _c_p]),
    'edhmc_predict': (ctypes.c_int, [_c_p, _c_p, _i32, _c_p, _c_p, _c_p, _u64, _c_p]),
    'edhmc_sgld_step': (ctypes.c_int, [_c_p, _f32, _u64, _u64, _c_p, _c_p, _c_p]),
    'edhmc_sghmc_step': (ctypes.c_int, [_c_p, _f32, _f32, _c_p, _u64, _u64, _c_p, _c_p, _c_p]),
    'edhmc_set_lik_weight': (ctypes.c_int, [_c_p, _f32]),
    'edhmc_set_prior_family': (ctypes.c_int, [_c_p, _c_p, _c_p, _c_p]),
    'edhmc_newton_terms': (ctypes.c_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    'edhmc_hier_set_scale_prior': (ctypes.c_int, [_c_p, _i32, ctypes.c_double, ctypes.c_double,
                                                   ctypes.c_double]),
    'edhmc_mh_step': (ctypes.c_int, [_c_p, _c_p, _c_p, _u64, _u64, _c_p, _c_p, _c_p, _c_p]),
    'edhmc_replica_exchange_step': (ctypes.c_int, [_c_p, _c_p, _c_p, _i32, _f32, _u64, _u64,
                                                   _c_p, _c_p, _c_p, _c_p, _c_p]),
    'edhmc_transition': (ctypes.c_int, [_c_p, _f32, _i32, _c_p, _c_p, _u64, _u64, _c_p,
                                        _c_p, _c_p, _c_p]),
    'edhmc_run': (ctypes.c_int, [_c_p, _f32, _i32, _c_p, _i64, _i64, _i64, _u64, _u64,
                                 _c_p, _c_p]),
    'edhmc_moments_enable': (ctypes.c_int, [_c_p, _i32]),
    'edhmc_moments_get': (ctypes.c_int, [_c_p, _c_p, _c_p, ctypes.POINTER(_i64), _c_p]),
    'edhmc_divergences': (ctypes.c_int, [_c_p, _c_p, _c_p]),
    'edhmc_last_ratio': (ctypes.c_int, [_c_p, _c_p, _c_p]),
    'edhmc_set_inner_gradient': (ctypes.c_int, [_c_p, _i32]),
    'edhmc_get_momentum': (ctypes.c_int, [_c_p, _c_p, _c_p]),
    'edhmc_sample_moments': (ctypes.c_int, [_c_p, _c_p, _i32, _i64, _i64, _i64, _c_p, _c_p,
                                            _c_p]),
    'edhmc_record_pairs': (ctypes.c_int, [_c_p, _c_p, _c_p, _i64, _i32]),
    'edhmc_record_count': (_i64, [_c_p]),
    'edhmc_col_stats': (ctypes.c_int, [_c_p, _c_p, _i64, _i32, _c_p, _c_p, _c_p]),
    'edhmc_col_stats_noalloc': (ctypes.c_int, [_c_p, _c_p, _i64, _i32, _c_p, _c_p, _c_p]),
    'edhmc_profile_enable': (ctypes.c_int, [_c_p, _i32]),
    'edhmc_profile_read': (ctypes.c_int, [_c_p, ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(_i64)]),
    'edhmc_profile_select': (ctypes.c_int, [_c_p, ctypes.c_uint32]),
    'edhmc_draw_momenta': (ctypes.c_int, [_c_p, _i32, _i32, _u64, _u64, _c_p, _c_p]),
    'edhmc_draw_uniforms': (ctypes.c_int, [_c_p, _i32, _u64, _u64, _c_p, _c_p]),
    'edhmc_autocov': (ctypes.c_int, [_c_p, _c_p, _i64, _i32, _i32, _i32, _c_p, _c_p, _c_p]),
    'edhmc_sur_create': (ctypes.c_int, [_c_p, _i32, _i32, _i32]),
    'edhmc_sur_num_params': (_i64, [_c_p]),
    'edhmc_sur_set_params': (ctypes.c_int, [_c_p, _c_p, _c_p]),
    'edhmc_sur_get_params': (ctypes.c_int, [_c_p, _c_p, _c_p]),
    'edhmc_sur_set_norm': (ctypes.c_int, [_c_p, _c_p, _c_p, _c_p, _c_p, _c_p]),
    'edhmc_sur_forward': (ctypes.c_int, [_c_p, _c_p, _i32, _c_p, _c_p]),
    'edhmc_sur_loss_grad': (ctypes.c_int, [_c_p, _c_p, _c_p, _i32, _c_p, _c_p, _c_p]),
    'edhmc_sur_train_step': (ctypes.c_int, [_c_p, _c_p, _c_p, _i32, _f32, _f32, _f32, _f32,
                                            _i32, _c_p, _c_p]),
    'edhmc_gather_rows': (ctypes.c_int, [_c_p, _c_p, _i64, _i32, _c_p, _i32, _c_p, _c_p]),
    'edhmc_sur_skip_gram': (ctypes.c_int, [_c_p, _c_p, _c_p, _i64, _c_p, _i32, _c_p, _c_p,
                                           _c_p]),
    'edhmc_affine_rows': (ctypes.c_int, [_c_p, _c_p, _i64, _i32, _c_p, _c_p, _c_p, _c_p]),
    'edhmc_sur_train_epoch': (ctypes.c_int, [_c_p, _c_p, _c_p, _i64, _c_p, _i32, _i32,
                                             ctypes.POINTER(_f32), _f32, _f32, _f32, _i32, _c_p,
                                             _c_p]),
    'edhmc_transition_surrogate': (ctypes.c_int, [_c_p, _f32, _i32, _c_p, _c_p, _u64, _u64,
                                                  _c_p, _c_p, _c_p, _c_p]),
    'edhmc_bind_hier': (ctypes.c_int, [_c_p, _c_p, _c_p, _c_p, _i32, _i64, _i32, _c_p]),
    'edhmc_hier_logp_grad': (ctypes.c_int, [_c_p, _c_p, _i32, _c_p, _c_p, _c_p]),
    'edhmc_hier_set_state': (ctypes.c_int, [_c_p, _c_p, _i32, _c_p]),
    'edhmc_hier_transition': (ctypes.c_int, [_c_p, _f32, _i32, _c_p, _c_p, _u64, _u64,
                                             _c_p, _c_p, _c_p, _c_p]),
}

_lib = None


class EdhmcError(RuntimeError):
  """A failed C-ABI call; `code` is the EDHMC_E* status."""

  def __init__(self, code, msg):
    super(EdhmcError, self).__init__('edhmc error %d: %s' % (code, msg))
    self.code = code


class NonFiniteError(EdhmcError, ValueError):
  """Non-finite data or positions (ed.dot's InvalidArgumentError)."""


def load(path=None):
  """Load and declare the library (once).  Raises OSError if it is absent."""
  global _lib
  if _lib is not None:
    return _lib
  path = path or LIB_PATH
  if not os.path.exists(path):
    raise OSError('libedhmc.so not built (%s); run __graft_entry__.build() or '
                  'make -C edward_amd/csrc' % path)
  lib = ctypes.CDLL(path)
  for name, (res, args) in SIGNATURES.items():
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args
  _lib = lib
  return lib


def check(rc, ctx=None):
  if rc == OK:
    return
  msg = load().edhmc_last_error(ctx).decode() if ctx else ''
  if rc == ENONFINITE:
    raise NonFiniteError(rc, msg)
  raise EdhmcError(rc, msg)


def ptr(t):
  """Device pointer of a tensor (None -> NULL)."""
  if t is None:
    return None
  return ctypes.c_void_p(t.data_ptr())


def stream_ptr(stream=None):
  s = stream if stream is not None else torch.cuda.current_stream()
  return ctypes.c_void_p(s.cuda_stream)
