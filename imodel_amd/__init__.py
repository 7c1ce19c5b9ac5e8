"""imodel_amd: MI355X-native TRiSK/HTC shallow-water step (drop-in for the
iModel ``swm`` module path: swm_tests -> ode_rk4 -> tendency).

The compute path is libimx (hand-written HIP for gfx950 behind the C ABI in
include/imx_swm.h). This package is the Python host over that ABI.
"""
from .swm import Mesh, Swm, ode_rk4, tendency  # noqa: F401
