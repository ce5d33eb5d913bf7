"""ctypes mirror of the DEM structs of include/lbdem_hip.h and their prototypes."""
import ctypes

from .. import _hip

_vp = ctypes.c_void_p
_d, _i = ctypes.c_double, ctypes.c_int32


class DemParams(ctypes.Structure):
    _fields_ = [("n", _i), ("K", _i), ("model", _i), ("T", _i), ("probe27", _i), ("periodic", _i * 3),
                ("dt", _d), ("g", _d * 3), ("dmin", _d * 3), ("dmax", _d * 3), ("cell", _d),
                ("mat", (_d * 9) * 2), ("wall_n", (_d * 3) * 6), ("wall_d", _d * 6), ("wall_type", _i * 6),
                ("fine_cell", _d), ("large_r", _d), ("nlarge", _i)]


class DemState(ctypes.Structure):
    _fields_ = [(name, _vp) for name in (
        "mass", "radius", "pos", "vel", "acc", "force", "quat", "omega", "omega_dot", "moment", "inertia",
        "force_fluid", "moment_fluid", "freeze", "fixvel", "cf_j", "cf_active", "cfn", "cf_pos", "cf_force",
        "cf_shear", "wcf_active", "wcf_pos", "wcf_force", "wcf_shear", "status")]


class DemWork(ctypes.Structure):
    _fields_ = [("count", _vp), ("offset", _vp), ("current", _vp), ("tile_sum", _vp), ("hash", _vp),
                ("pid", _vp), ("pairs", _vp), ("pair_cap", ctypes.c_int64), ("npairs", _vp),
                ("need_new", _vp), ("verlet", _vp), ("verlet_ref", _vp), ("spos", _vp),
                ("fhash", _vp), ("large", _vp)]


P = ctypes.POINTER
_hip.register({
    "lbdem_dem_broadphase": (ctypes.c_int, [P(DemParams), _vp, P(DemWork), _vp]),
    "lbdem_dem_run": (ctypes.c_int, [P(DemParams), P(DemState), P(DemWork), ctypes.c_int32, ctypes.c_int32, _vp]),
})
