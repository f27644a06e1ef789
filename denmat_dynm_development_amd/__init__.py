"""MI355X-native Lindblad scattering operator for DMD (denmat_dynm).

Drop-in for the reference's dRho/dt|scatt (electronphonon::evolve and the
initializer's compute_ddm) behind the C-ABI in include/dmd_scatt.h; the HIP
library is libdmd_scatt.so in this directory.
"""
from . import _lib  # noqa: F401
from ._lib import DMDError, SynthParams  # noqa: F401

__all__ = ["_lib", "DMDError", "SynthParams"]
