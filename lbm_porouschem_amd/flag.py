"""Model and boundary-condition switches.

Same class names, member names and integer values as the reference's
src/LBM/util/flag.py:10-109, so a configuration written against the reference
selects the same model here.  The integers are what crosses the C ABI
(include/lbm_porouschem.h, LBM_* enum macros).
"""
from enum import Enum


class VISCOSITY_MODEL(Enum):
    NONE = 0
    CONSTANT = 1
    SUTHERLAND = 2
    MIXTURE = 3


class PORO_MODEL(Enum):
    SPHERICAL = 0
    DARCY = 1
    DARCYFORCHHEIMER = 2
    DARCY_HIGH = 3
    ERGUN = 4


class CONDUCTIVITY_MODEL(Enum):
    NONE = 0
    CONSTANT = 1
    POLYNOMIAL = 2
    MIXTURE = 3


class THERMO_MODEL(Enum):
    NONE = 0
    CONSTANT = 1
    POLYNOMIAL = 2
    NASA7 = 3
    MIXTURE = 4


class THERMAL_DIFF_MODEL(Enum):
    NONE = 0
    CONSTANT = 1
    DERIVED = 2
    PRANDTL = 3


class DIFF_MODEL(Enum):
    NONE = 0
    CONSANT = 1  # spelling of the reference (flag.py:51)
    SCHMIDT = 2


class SPECIE_UNIT(Enum):
    MASS = 1
    MOLE = 2


class REACTION_TYPE(Enum):
    ARREHNIUS = 0
    LMH = 1


class FLUID_STATE_EQUATION(Enum):
    IDEAL_GAS = 0
    INCOMPRESSIBLE = 1


class RADIATION_MODEL(Enum):
    NONE = 0
    SURFACE_UNIFORM = 1
    REAL_RADIATION = 2
    P1_APPROACH = 3


class SOURCE_TERM(Enum):
    NONE = 0
    MICRO = 1
    MACRO = 2


class FORCE_TERM(Enum):
    NONE = 0
    GUO = 1
    SHAN = 2
    MACRO = 3


class BC(Enum):
    periodic = 0
    zeroGradient = 1
    fixedValue = 2


class BC_FLOW(Enum):
    periodic = 0
    inlet = 1
    outlet = 2
    wall = 3
    symmetric = 4
    inlet_flow = 5


class BC_MODEL(Enum):
    NONE = 0
    NEE = 1
    NEBB = 2
    ES = 3


__all__ = [
    "VISCOSITY_MODEL", "PORO_MODEL", "CONDUCTIVITY_MODEL", "THERMO_MODEL", "THERMAL_DIFF_MODEL",
    "DIFF_MODEL", "SPECIE_UNIT", "REACTION_TYPE", "FLUID_STATE_EQUATION", "RADIATION_MODEL",
    "SOURCE_TERM", "FORCE_TERM", "BC", "BC_FLOW", "BC_MODEL",
]
