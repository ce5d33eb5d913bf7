"""Domain and material property definitions (src/dem3d/demconfig/types.py)."""
from dataclasses import dataclass

import numpy as np


class Vector3(np.ndarray):
    """3-vector with .x/.y/.z access (stands in for ti.types.vector(3, float))."""

    def __new__(cls, x=0.0, y=0.0, z=0.0):
        obj = np.asarray([x, y, z], dtype=np.float64).view(cls)
        return obj

    x = property(lambda self: float(self[0]))
    y = property(lambda self: float(self[1]))
    z = property(lambda self: float(self[2]))


def _scalar(v):
    # drivers pass tuples like (value,) by accident (single_particle_settle.py:129-134)
    if isinstance(v, (tuple, list, np.ndarray)):
        v = np.asarray(v).reshape(-1)[0]
    return float(v)


@dataclass
class DomainBounds:
    """Simulation domain boundaries."""
    xmin: float
    xmax: float
    ymin: float
    ymax: float
    zmin: float
    zmax: float

    def __post_init__(self):
        for a in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
            setattr(self, a, _scalar(getattr(self, a)))
        for lo, hi in (("xmin", "xmax"), ("ymin", "ymax"), ("zmin", "zmax")):
            if getattr(self, lo) >= getattr(self, hi):
                raise ValueError(f"{lo} ({getattr(self, lo)}) must be < {hi} ({getattr(self, hi)})")

    def get_extended_bounds(self):
        return (Vector3(self.xmin, self.ymin, self.zmin), Vector3(self.xmax, self.ymax, self.zmax))


@dataclass
class ParticleProperties:
    """Material properties of particles."""
    elastic_modulus: float = 0
    poisson_ratio: float = 0
    density: float = 0
    max_coordinate_number: int = 0


@dataclass
class WallProperties:
    """Material properties of walls."""
    density: float = 0
    elastic_modulus: float = 0
    poisson_ratio: float = 0.0
