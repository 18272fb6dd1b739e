"""The allocation-free column statistics (edhmc_col_stats_noalloc) against the
allocating entry point (edhmc_col_stats), through the C ABI: bitwise equal
outputs on two calls in a row, the second with a smaller input (the shared
device scratch keeps the first call's partial sums, which the second must
not read); widths beyond the scratch are refused before any launch."""

import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _stats(ctx, fn, X):
  P, D = X.shape
  mean = torch.empty(D, device='cuda')
  sd = torch.empty(D, device='cuda')
  rc = getattr(ctx.lib, fn)(ctx.h, ctypes.c_void_p(X.data_ptr()), P, D,
                            ctypes.c_void_p(mean.data_ptr()), ctypes.c_void_p(sd.data_ptr()),
                            ctx.stream)
  return rc, mean, sd


def test_col_stats_noalloc_matches_the_allocating_entry_point():
  from edward_amd.engine import Context
  ctx = Context(0)
  g = torch.Generator(device='cuda')
  g.manual_seed(5)
  for P, D in ((70000, 4096), (1000, 37)):
    print('case P=%d D=%d' % (P, D), flush=True)
    X = (torch.randn(P, D, generator=g, device='cuda') * 3.0 + 1.5).contiguous()
    rc_a, m_a, s_a = _stats(ctx, 'edhmc_col_stats', X)
    rc_b, m_b, s_b = _stats(ctx, 'edhmc_col_stats_noalloc', X)
    torch.cuda.synchronize()
    assert rc_a == 0 and rc_b == 0
    np.testing.assert_array_equal(m_b.cpu().numpy(), m_a.cpu().numpy())
    np.testing.assert_array_equal(s_b.cpu().numpy(), s_a.cpu().numpy())
    # and both are the column moments (FP64 on the host, divisor P)
    x = X.double().cpu().numpy()
    np.testing.assert_allclose(m_b.cpu().numpy(), x.mean(0), rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(s_b.cpu().numpy(), x.std(0), rtol=1e-6)
  # beyond the scratch: refused with EDHMC_EINVAL, nothing launched
  X = torch.zeros(4, 4097, device='cuda')
  rc, _, _ = _stats(ctx, 'edhmc_col_stats_noalloc', X)
  assert rc == -1
  assert b'4096' in ctx.lib.edhmc_last_error(ctx.h)
  ctx.close()
