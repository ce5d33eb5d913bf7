from .psclattice3d import PSCLattice3D

__all__ = ["PSCLattice3D"]
