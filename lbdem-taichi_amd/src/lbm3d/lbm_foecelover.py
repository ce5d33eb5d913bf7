"""BGK with a constant body force (API of src/lbm3d/lbm_foecelover.py).

Same fused kernel as BasicLattice3D with Smagorinsky off (omega stays omega0) and the
forcing term rho * w_q * (c_q . F) / cs^2 added after relaxation (lbm_foecelover.py:37-68).
"""
from .. import _hip
from .lbm_solver3d import BasicLattice3D
from .lbmutils import CellType


class ForcedLattice3D(BasicLattice3D):
    # ForcedLattice3D.collide skips OBSTACLE | VEL_LADD | FREE_SLIP (lbm_foecelover.py:45)
    _skip_types = CellType.OBSTACLE | CellType.VEL_LADD | CellType.FREE_SLIP | CellType.VEL_INLET_LADD

    def __init__(self, Nx: int, Ny: int, Nz: int, omega: float, dx: float, dt: float, rho: float, F,
                 storage=None):
        self._F_phys = tuple(float(v) for v in F)
        super().__init__(Nx, Ny, Nz, omega, dx, dt, rho, storage=storage)
        self.F = tuple(self.unit.getLbAccel(v) for v in self._F_phys)  # lattice units (:31)

    def _collision(self):
        return _hip.COLL_BGK

    def _body_force(self):
        return getattr(self, "F", (0.0, 0.0, 0.0))
