"""MI355X-native finite-element residual / Jacobian-action path with PyLith's integrator API.

The compute path is ``libpylith_amd.so`` (hand-written gfx950 HIP kernels behind the C ABI of
include/pylith_amd.h); this package only binds it.  Importing ``pylith_amd.dm`` loads the
library and fails loudly when it has not been built.
"""
from . import _lib  # noqa: F401
from ._lib import (ADD_VALUES, FP32, FP64, HEX_Q1, INSERT_VALUES, LHS, RHS, TET_P2,  # noqa: F401
                   F1V_ISOLINEAR_ELASTICITY_3D, JF3VU_ISOLINEAR_ELASTICITY_3D, F1V_GENMAXWELL_3D,
                   JF3VU_GENMAXWELL_3D, VISCOUS_STRAIN_GENMAXWELL_3D, TOTAL_STRAIN_3D)

__all__ = ["dm", "meshgen"]
