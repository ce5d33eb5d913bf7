"""Prototypes shared by the unresolved couplings."""
import ctypes

from .. import _hip
from ..lbdem.psclattice3d import Grains, Units

_vp = ctypes.c_void_p
_P = ctypes.POINTER
_hip.register({
    "lbdem_ibm_spread": (ctypes.c_int, [_P(_hip.Lattice), _P(_hip.StepArgs), _P(Grains), _P(Units), _vp, _vp, _vp,
                                        _vp, _vp, _vp, _vp, _vp]),
    "lbdem_ibm_drag": (ctypes.c_int, [_P(_hip.Lattice), _vp, _P(Grains), _P(Units), _vp, _vp, _vp]),
    "lbdem_ibm_weight": (ctypes.c_int, [_P(_hip.Lattice), _vp, _vp, _vp, _vp, _vp, _vp]),
    "lbdem_vans_spread": (ctypes.c_int, [_P(_hip.Lattice), _vp, _P(Grains), _P(Units), _vp, _vp, _vp]),
    "lbdem_vans_drag": (ctypes.c_int, [_P(_hip.Lattice), _vp, _P(Grains), _P(Units), _vp, _vp, _vp]),
})


class GrainStructs:
    """Cached ctypes views of the DEM particle tensors and the lattice units."""

    def __init__(self, lattice):
        from ..lbdem.psclattice3d import grains_struct, units_struct
        self.g = grains_struct(lattice.dem)
        self.u = units_struct(lattice.unit, lattice.dem)
        self.gp = ctypes.byref(self.g)
        self.up = ctypes.byref(self.u)
