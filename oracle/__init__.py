"""CPU oracle for the HMC hot path -- TEST INFRASTRUCTURE ONLY.

This package restates, in FP64 NumPy, the algorithm of Edward 1.3.5's
``ed.HMC`` path (``edward/inferences/hmc.py``, ``monte_carlo.py``,
``inference.py``, ``models/empirical.py``) so that the HIP engine in
``edward_amd`` can be checked against it.  Only ``tests/``,
``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may
import it, and only as the checker / the timed CPU restatement.  The product
path (``edward_amd``) never imports it and fails loudly when its HIP library is
missing.

Modules:
  hmc_oracle  -- log-joints and gradients, leapfrog, Metropolis accept,
                 Empirical storage and moments, Philox4x32-10 RNG, ESS.
  mlp_oracle  -- FP64 forward/backward/Adam of the gradient-surrogate MLP.

Every function's docstring names the reference file:line it follows (reference
paths are relative to the Edward source tree) or, for build-only parts, the
published method it restates, and the test that pins it.
"""
