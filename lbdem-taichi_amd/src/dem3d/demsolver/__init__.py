"""DEM driver."""
from .demsolver import DEMSolver
from .utils import quat2RotMatrix

__all__ = ["DEMSolver", "quat2RotMatrix"]
