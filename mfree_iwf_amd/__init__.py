"""mfree_iwf_amd: MI355X (gfx950) implementation of mfree_iwf's SPH
particle-interaction hot path -- cell-list neighbour search, Gray-Monaghan
pair sums, Randles-Libersky corrected Total-Lagrangian SPH and RK4 -- behind
the C ABI of include/mfree_iwf_amd.h.
"""
from . import _capi  # noqa: F401
from ._capi import (METHOD_TL_COROTATIONAL, METHOD_TL_POTENTIAL, METHOD_TL_RL, METHOD_UL_GODUNOV,  # noqa: F401
                    METHOD_UL_MONAGHAN, PRECISION_FP64, PRECISION_MIXED, MfreeError,  # noqa: F401
                    Params)
from .scenes import ParticleSet, cubic_spline, make_plate, make_rings  # noqa: F401
from .solver import Solver, convex_hull, device_count  # noqa: F401
from .surrogate import Surrogate  # noqa: F401
