"""CPU checks of the benchmark's contract: one GPU only (a request for more
exits with a message), and the CPU baseline's fields on a small workload."""

import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
  p = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                     capture_output=True, text=True, timeout=300)
  assert p.returncode == 2
  assert 'one GPU' in (p.stdout + p.stderr)


def test_cpu_baseline_fields():
  sys.path.insert(0, ROOT)
  import bench
  from edward_amd import workloads as W
  wl = W.logreg(3000, 16, 64, 0.01, 4, 10, seed=2)
  out = bench.cpu_baseline(wl, 0.2, 3.0)
  assert out['kind'] == 'port' and out['cores'] >= 1
  assert out['value'] > 0 and np.isfinite(out['value'])
  # ESS per transition / CPU seconds per transition
  np.testing.assert_allclose(out['value'], 3.0 / out['sec_per_transition'], rtol=1e-12)
  assert 'min-ESS/s' in out['unit'] and 'gradient evaluations' in out['sample']
