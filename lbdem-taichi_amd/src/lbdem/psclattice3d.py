"""Resolved LBM-DEM coupling with partially saturated cells (API of src/lbdem/psclattice3d.py).

`collide()` runs, on the device and without host synchronisation:
  1. grain -> lattice mapping (lbdem_psc_map): the cells covered by each grain get a packed
     key (full cover / sub-cell count / grain index) by atomic max, grain-centric;
  2. the fused collide-stream step with the PSC weighted collision, which also writes the
     per-cell momentum exchange at covered cells;
  3. the per-grain FP64 force and torque reduction (lbdem_psc_forces) into the DEM
     solver's force_fluid / moment_fluid.
`id`, `volfrac`, `weight`, `velsolid` are decoded from the keys when read.
"""
import ctypes

import numpy as np
import torch

from .. import _hip
from .._field import Field
from ..lbm3d.lbm_solver3d import BasicLattice3D
from ..lbm3d.lbmutils import CellType

_vp = ctypes.c_void_p


class Grains(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("position", _vp), ("velocity", _vp), ("omega", _vp), ("radius", _vp),
                ("force_fluid", _vp), ("moment_fluid", _vp), ("order", _vp)]


class Units(ctypes.Structure):
    _fields_ = [("dx", ctypes.c_double), ("dt", ctypes.c_double), ("rho", ctypes.c_double),
                ("dmin", ctypes.c_double * 3)]


_P = ctypes.POINTER
_hip.register({
    "lbdem_psc_map": (ctypes.c_int, [_P(_hip.Lattice), _vp, _P(Grains), _P(Units), _vp, _vp, _vp, ctypes.c_int32,
                                     _vp]),

    "lbdem_psc_forces": (ctypes.c_int, [_P(_hip.Lattice), _P(Grains), _P(Units), _vp, _vp, _vp, _vp, _vp]),
    "lbdem_psc_decode": (ctypes.c_int, [_P(_hip.Lattice), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
})


def grains_struct(dem):
    t = dem.grain_tensors()
    g = Grains()
    g.n = dem.n
    for k in ("position", "velocity", "omega", "radius", "force_fluid", "moment_fluid"):
        setattr(g, k, t[k].data_ptr())
    return g


def units_struct(unit, dem):
    u = Units()
    u.dx, u.dt, u.rho = float(unit.dx), float(unit.dt), float(unit.rho)
    d = dem.config.domain
    u.dmin[0], u.dmin[1], u.dmin[2] = d.xmin, d.ymin, d.zmin
    return u


class PSCLattice3D(BasicLattice3D):
    """Partially saturated cell (Noble-Torczynski) coupling; BGK with constant omega."""
    _skip_types = CellType.OBSTACLE | CellType.VEL_LADD | CellType.FREE_SLIP | CellType.VEL_INLET_LADD
    _coupling = _hip.COUPLE_PSC

    def __init__(self, Nx: int, Ny: int, Nz: int, omega: float, dx: float, dt: float, rho: float, demslover,
                 storage=None, smagorinsky=None):
        super().__init__(Nx, Ny, Nz, omega, dx, dt, rho, storage=storage, smagorinsky=smagorinsky)
        self.dem = demslover
        dev = self.device
        ng = max(self.dem.n, 1)
        # Two key arrays (uint32 bit patterns) with their grain boxes, used by turns: the force
        # reduction after mapping n zeroes the keys of mapping n - 1 in the same launch, so
        # mapping n + 1 starts from clean keys without a clear pass.
        self._keys = [torch.zeros(self.N, dtype=torch.int32, device=dev) for _ in range(2)]
        self._bboxes = [torch.zeros(6 * ng, dtype=torch.int32, device=dev) for _ in range(2)]
        self._dirty = [False, False]   # keys of a mapping not yet zeroed
        self._cur = 1                  # buffer of the latest mapping
        self._snap = torch.zeros(12 * ng, dtype=torch.float32, device=dev)
        self._cell_force = torch.zeros(3 * self.N, dtype=torch.float32, device=dev)
        self._decoded_version = -1
        shp = (self.Nx, self.Ny, self.Nz)
        self.id = Field(shp, torch.int32, dev, refresh=self._decode)
        self.volfrac = Field(shp, torch.float32, dev, refresh=self._decode)
        self.weight = Field(shp, torch.float32, dev, refresh=self._decode)
        self.velsolid = Field(shp, torch.float32, dev, n=3, refresh=self._decode)
        self.hydroforce = Field(shp, torch.float32, dev, n=3, tensor=self._cell_force.view(*shp, 3))
        self._map_count = 0

    def _collision(self):
        return _hip.COLL_BGK

    @property
    def _key(self):
        return self._keys[self._cur]

    def _coupling_args(self, a):
        a.psc_key = self._key.data_ptr()
        a.psc_grain = self._snap.data_ptr()
        a.psc_force = self._cell_force.data_ptr()

    # ---------------------------------------------------------------------------------
    def initialize(self):
        """f <- feq, then map grains to the lattice (psclattice3d.py:60-78)."""
        super().initialize()
        self.grains2lattice()

    def _structs(self):
        if getattr(self, "_gu", None) is None:
            g = grains_struct(self.dem)
            u = units_struct(self.unit, self.dem)
            self._gu = (g, u, ctypes.byref(g), ctypes.byref(u))
        else:
            self.dem.gf.sync()
        return self._gu[2], self._gu[3]

    def grains2lattice(self):
        """Map grain id, solid fraction, weight and velocity to the lattice (:84-182)."""
        self._prepare()
        g, u = self._structs()
        b = 1 - self._cur
        _hip.check(_hip.lib().lbdem_psc_map(self._lat_c, _hip.ptr(self._flags), g, u,
                                            _hip.ptr(self._keys[b]), _hip.ptr(self._snap), _hip.ptr(self._bboxes[b]),
                                            1 if self._dirty[b] else 0, _hip.stream()), "lbdem_psc_map")
        self._cur = b
        self._dirty[b] = True
        self._map_count += 1

    def lattice2grains(self):
        """Hydrodynamic force and torque on the grains (:316-358), FP64, physical units."""
        g, u = self._structs()
        other = 1 - self._cur
        stale = (_hip.ptr(self._keys[other]), _hip.ptr(self._bboxes[other])) if self._dirty[other] else (None, None)
        _hip.check(_hip.lib().lbdem_psc_forces(self._lat_c, g, u, _hip.ptr(self._key), _hip.ptr(self._cell_force),
                                               stale[0], stale[1], _hip.stream()), "lbdem_psc_forces")
        self._dirty[other] = False
        self.dem.gf.force_fluid.touch()
        self.dem.gf.moment_fluid.touch()

    def collide(self):
        """Mapping, weighted collide-stream and force reduction (psclattice3d.py:187-218)."""
        self.grains2lattice()
        super().collide()
        self.hydroforce.touch()
        self.lattice2grains()

    def _after_load(self):
        # keys are remapped at the next collide; start both buffers clean
        for k, b in zip(self._keys, self._bboxes):
            k.zero_()
            b.zero_()
        self._dirty = [False, False]
        self._map_count += 1

    def _decode(self):
        if self._decoded_version == self._map_count:
            return
        _hip.check(_hip.lib().lbdem_psc_decode(self._lat_c, _hip.ptr(self._key), _hip.ptr(self._snap),
                                               _hip.ptr(self.id.dev), _hip.ptr(self.volfrac.dev),
                                               _hip.ptr(self.weight.dev), _hip.ptr(self.velsolid.dev),
                                               _hip.stream()), "lbdem_psc_decode")
        self._decoded_version = self._map_count
        for f in (self.id, self.volfrac, self.weight, self.velsolid):
            f._host = None
