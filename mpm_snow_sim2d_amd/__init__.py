"""MI355X-native MLS-MPM snow solver (the hot path of michal-sitarz2/MPM_Snow_Sim2D:
P2G -> grid update -> G2P) behind the reference's solver interface.

The compute path is libmpm_hip.so (hand-written HIP for gfx950, C-ABI in
include/mpm/mpm.h); this package holds its ctypes host mirror, the parameter set
and seeded scene samplers.  Importing the solver requires the built library.
"""
from .params import (BOUNDARY_COULOMB_MARGIN, BOUNDARY_STICKY, Options, Params, reference_params,
                     snow_params)
from .scenes import ParticleSet

__all__ = ["Params", "Options", "ParticleSet", "reference_params", "snow_params",
           "BOUNDARY_COULOMB_MARGIN", "BOUNDARY_STICKY", "GPUSolver"]


def __getattr__(name):
    if name == "GPUSolver":
        from .solver import GPUSolver
        return GPUSolver
    raise AttributeError(name)
