"""edward_amd -- MI355X-native engine behind Edward's ``ed.HMC`` surface.

Import as ``import edward_amd as ed``.  The public names mirror the subset of
Edward 1.3.5's API (``edward/__init__.py:32-90``) that an HMC program over a
generalised linear model uses: ``ed.HMC``, ``ed.models.{Normal, Bernoulli,
Empirical}``, ``ed.dot``, ``ed.set_seed``, ``ed.get_session``, plus
``ed.placeholder``/``ed.gather``/``ed.exp`` standing in for the TensorFlow
ops such programs call, the criticism of the fitted model
(``ed.copy``, ``ed.evaluate``, ``ed.ppc``), and ``ed.transform`` with the
bijectors of the auto-transform.  The arithmetic runs in
``libedhmc.so``.
"""

from edward_amd import bijectors  # noqa: F401
from edward_amd import models  # noqa: F401
from edward_amd.bijectors import transform  # noqa: F401
from edward_amd import util  # noqa: F401
from edward_amd.criticisms import copy, evaluate, ppc  # noqa: F401
from edward_amd.inferences import (HMC, SGHMC, SGLD, Inference, MetropolisHastings,  # noqa: F401
                                   MonteCarlo, ReplicaExchangeMC)
from edward_amd.models import dot, exp, gather, placeholder  # noqa: F401
from edward_amd.util import get_seed, get_session, set_seed  # noqa: F401
from edward_amd.version import __version__  # noqa: F401
