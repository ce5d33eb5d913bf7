"""MI355X-native drop-in for LBDEM-Taichi's solver package (`src`).

Drivers written for the reference import `src.lbm3d.lbm_solver3d`, `src.lbdem.psclattice3d`,
`src.unlbdem.eqlattice`, `src.unlbdem.vanslattice` and `src.dem3d.*`; the same module paths
exist here.  State lives in torch tensors on the GPU; every time step runs hand-written HIP
kernels (gfx950) from liblbdem_hip.so through a C ABI.  `init()` replaces `ti.init()`.
"""
from .runtime import init, config

__all__ = ["init", "config"]
__version__ = "0.1.0"
