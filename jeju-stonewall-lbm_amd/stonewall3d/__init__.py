"""MI355X-native D3Q19 regularized-BGK + Smagorinsky-LES stone-wall LBM (host side).

Mirrors the solver surface of the reference's stonewall3d package (lbm3d.py, geometry.py)
over hand-written HIP kernels in libstonewall_lbm.so (C ABI: include/stonewall_lbm.h).
"""
from .lattice import C, W, OPP, MIRZ, MIRY, Q, FLUID, WALL, GROUND, flag_from_mask  # noqa: F401

__version__ = "1.0.0"
