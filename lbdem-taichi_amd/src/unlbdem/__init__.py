from .eqlattice import EqIMBlattice3D
from .vanslattice import Unresolvedlattice3D

__all__ = ["EqIMBlattice3D", "Unresolvedlattice3D"]
