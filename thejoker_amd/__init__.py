"""thejoker_amd -- The Joker's marginal-likelihood and rejection-sampling path
on AMD MI355X (gfx950).

Drop-in counterparts of the reference's public API (adrn/thejoker v1.3.0):
:class:`TheJoker`, :class:`RVData`, :class:`JokerPrior`,
:class:`JokerSamples`.  The likelihood, rejection and linear-posterior
arithmetic runs in hand-written HIP kernels (libjoker.so, C ABI in
include/joker.h) reached through ctypes; torch provides device memory and
streams.
"""
from . import distributions, units
from .data import RVData
from .prior import JokerPrior
from .samples import JokerSamples
from .thejoker import TheJoker

__all__ = ["TheJoker", "RVData", "JokerPrior", "JokerSamples", "units", "distributions"]
__version__ = "0.1.0"
