"""CPU checks of the C-ABI boundary: the library loads, exports every function
include/edhmc.h declares, and the ctypes declarations match the header.  No
compute call is made (no GPU here)."""

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'edhmc.h')
LIB = os.path.join(ROOT, 'edward_amd', 'libedhmc.so')


def declared_functions():
  text = open(HEADER).read()
  text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
  return sorted(set(re.findall(r'\b(edhmc_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_the_boundary():
  names = declared_functions()
  for must in ('edhmc_create', 'edhmc_bind_glm', 'edhmc_logp_grad', 'edhmc_transition',
               'edhmc_run', 'edhmc_sur_train_step', 'edhmc_transition_surrogate',
               'edhmc_hier_transition'):
    assert must in names


def test_library_exports_every_declared_symbol():
  if not os.path.exists(LIB):
    pytest.skip('library not built')
  out = subprocess.run(['nm', '-D', '--defined-only', LIB], capture_output=True, text=True,
                       check=True).stdout
  exported = set(re.findall(r'\b(edhmc_[a-z0-9_]+)\b', out))
  missing = [n for n in declared_functions() if n not in exported]
  assert not missing, missing


def test_ctypes_declarations_cover_the_header():
  from edward_amd import _lib
  assert sorted(_lib.SIGNATURES) == declared_functions()


def test_library_loads_without_gpu():
  if not os.path.exists(LIB):
    pytest.skip('library not built')
  from edward_amd import _lib
  lib = _lib.load()
  assert lib.edhmc_version().startswith(b'edhmc')
  assert b'null' in lib.edhmc_last_error(None)


def test_single_hip_runtime_is_mapped():
  """torch's HIP runtime and the library's must be one and the same."""
  if not os.path.exists(LIB):
    pytest.skip('library not built')
  from edward_amd import _lib
  _lib.load()
  maps = open('/proc/self/maps').read()
  paths = {l.split()[-1] for l in maps.splitlines() if 'libamdhip64' in l}
  assert len(paths) == 1, paths
