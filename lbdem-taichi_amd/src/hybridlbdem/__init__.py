from .hybridlattice import HybridLattice3D

__all__ = ["HybridLattice3D"]
