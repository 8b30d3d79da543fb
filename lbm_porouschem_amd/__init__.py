"""MI355X-native coupled lattice-Boltzmann step (flow + thermal + species +
gray-LBM porous drag + pyrolysis source) behind the reference's solver-class API.

The compute path is liblbm_mi355x.so (hand-written gfx950 HIP kernels, C ABI in
include/lbm_porouschem.h).  Importing the solver classes loads it and fails if
it has not been built.
"""
from . import constant, flag  # noqa: F401
from .flag import *  # noqa: F401,F403


def __getattr__(name):
    if name in ("LBM2DSolver", "LBM3DSolver", "LBMSolverBase"):
        from . import solver
        return getattr(solver, name)
    raise AttributeError(name)


__all__ = ["LBM2DSolver", "LBM3DSolver", "constant", "flag"]
