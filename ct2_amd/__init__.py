"""ct2_amd -- CosmicTiger 2's force path (direct-sum kick, drift, 3-D FFT,
particle-mesh force and its learned short-range correction) on MI355X.

The compute lives in ``libct2hip.so`` (HIP kernels for gfx950 behind the C
ABI of ``include/ct2hip.h``); this package is its ctypes face.
"""
from . import api  # noqa: F401
from ._lib import Ct2Error  # noqa: F401

__all__ = ["api", "Ct2Error"]
