"""mocat_amd: mocat's data-parallel sampling core on MI355X (gfx950) HIP kernels.

Drop-in surface (names as in mocat/__init__.py:12-56): cdict, Scenario, run, Sampler,
MCMCSampler, Underdamped, Correction, Uncorrected, Metropolis, RMMetropolis, plus the fused
scenarios (LogisticRegression) and jax.random-compatible keys in mocat_amd.random.
Every computation runs in libmocat_amd.so; there is no CPU fallback.
"""
from . import _lib
from . import random
from .core import Scenario, cdict, static_cdict, save_cdict, load_cdict, is_implemented, impl_checkable
from .sample import Sampler, run
from .mcmc import MCMCSampler, Underdamped, Correction, Uncorrected, Metropolis, RMMetropolis
from .scenarios import LogisticRegression, DiagonalGaussian
from .transport import (TransportSampler, SMCSampler, TemperedSMCSampler, MetropolisedSMCSampler,
                        RMMetropolisedSMCSampler, SVGD)
from .scenarios import GaussianMixture
from . import kernels
from . import reduce, metrics, ssm
from .metrics import (ksd, autocorrelation, integrated_autocorrelation_time, ess_autocorrelation, log_ess_log_weight,
                      ess_log_weight, squared_jumping_distance)

__version__ = "0.1.0"
