"""Writes tests/golden/posterior_large.npz: FP64 posterior moments of the large
logistic-regression workload (N = 1e6 rows, D = 512, X ~ N(0, 1) with column
scales U(0.5, 2) rounded to bf16, w* ~ N(0, 0.1^2 I), prior N(0, I); the
data of edward_amd.workloads.logreg_large, which the GPU samples).

The moments come from the oracle (oracle/hmc_oracle.py): Newton's method for
the mode and the inverse Hessian (laplace_logreg), then importance sampling
from that Gaussian with antithetic pairs (posterior_moments_is), all on the
bf16-rounded X in FP64.  The file also holds checksums of X and y, so a
reader can tell that its data are the ones generated now.

Run (about 30 minutes and 12 GB on 8 cores):
    python tests/golden/make_posterior_large.py [n_pairs]
"""

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from edward_amd import workloads  # noqa: E402
from oracle import hmc_oracle as O  # noqa: E402


def data_checksums(X, y):
  """Cheap fingerprints of the workload: FP64 sums of X, of X's first column
  weighted by the row index, of y, and X's first 4 x 8 block."""
  n = np.arange(X.shape[0], dtype=np.float64)
  return dict(x_sum=float(X.sum(dtype=np.float64)),
              x_col0_moment=float(n @ X[:, 0].astype(np.float64)),
              y_sum=float(y.sum(dtype=np.float64)),
              x_head=np.asarray(X[:4, :8], np.float32))


def main():
  n_pairs = int(sys.argv[1]) if len(sys.argv) > 1 else 65536
  t0 = time.time()
  wl = workloads.logreg_large()
  print('workload %.1f s' % (time.time() - t0), flush=True)
  m = O.GLM(wl.X.astype(np.float64), wl.y, wl.prior_loc, wl.prior_scale)
  mode, cov = O.laplace_logreg(m, tol=1e-12)
  print('laplace %.1f s' % (time.time() - t0), flush=True)
  est = O.posterior_moments_is(m, mode, cov, n_pairs, seed=1)
  print('importance sampling %.1f s, weight ESS %.0f of %d' % (time.time() - t0,
                                                              est['weight_ess'], 2 * n_pairs),
        flush=True)
  sd = np.sqrt(est['var'])
  print('mean - mode: max %.3e sd; var / laplace var: %.4f .. %.4f; se_mean max %.3e sd, '
        'se_var max %.3e relative'
        % (np.abs(est['mean'] - mode).max() / sd.min(), (est['var'] / np.diag(cov)).min(),
           (est['var'] / np.diag(cov)).max(), (est['se_mean'] / sd).max(),
           (est['se_var'] / est['var']).max()), flush=True)
  np.savez_compressed(os.path.join(HERE, 'posterior_large.npz'), mean=est['mean'],
                      var=est['var'], se_mean=est['se_mean'], se_var=est['se_var'],
                      mode=mode, laplace_var=np.diag(cov), weight_ess=est['weight_ess'],
                      n_pairs=n_pairs, **data_checksums(wl.X, wl.y))


if __name__ == '__main__':
  main()
