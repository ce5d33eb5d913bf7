from .lbmutils import CellType, UnitConverter
from .lbm_solver3d import BasicLattice3D
from .lbm_foecelover import ForcedLattice3D

__all__ = ["CellType", "UnitConverter", "BasicLattice3D", "ForcedLattice3D"]
