"""Hybrid resolved / unresolved LBM-DEM coupling (API of src/hybridlbdem/hybridlattice.py).

Grains with 2R > dx are coarse: partially saturated cells, mapped to packed keys as in
PSCLattice3D.  Grains with 2R <= dx are fine: particle-equivalence IBM with the Tenneti weight,
as in EqIMBlattice3D, spread only into the room the coarse grains leave.  Per step, in the
order of the reference's driver (example/hybridlbdem/hybridexperiment/experiment.py):

  prepare_step()   coarse mapping (lbdem_hybrid_map_coarse), then the fine spreading with the
                   free-fraction correction (lbdem_hybrid_map_fine); the fine Tenneti weight and
                   the coarse weight B (with the step's own omega) are evaluated in the collision;
  collide()        the fused collide-stream with f + beta Os1 + B Os2 + (1 - B - beta) Of, which
                   also writes the coarse momentum exchange per cell, then the wet-node pass;
  stream()         nothing left;
  lattice2grains() coarse force and torque (FP64 reduction per grain) and the fine grains'
                   Tenneti drag from the post-stream fluid velocity (lbdem_hybrid_forces).
"""
import ctypes

import torch

from .. import _hip
from .._field import Field
from ..lbdem.psclattice3d import PSCLattice3D, Grains, Units
from ..lbm3d.lbm_solver3d import BasicLattice3D
from ..lbm3d.lbmutils import CellType

_vp = ctypes.c_void_p
_P = ctypes.POINTER
_i32 = ctypes.c_int32
_hip.register({
    "lbdem_hybrid_map_coarse": (ctypes.c_int, [_P(_hip.Lattice), _vp, _P(Grains), _P(Units), _vp, _i32, _vp, _vp,
                                               _vp, _i32, _vp]),
    "lbdem_hybrid_map_fine": (ctypes.c_int, [_P(_hip.Lattice), _vp, _P(Grains), _P(Units), _vp, _vp, _vp, _vp,
                                             _vp, _vp, _vp]),
    "lbdem_hybrid_forces": (ctypes.c_int, [_P(_hip.Lattice), _vp, _P(Grains), _P(Units), _vp, _i32, _vp, _vp, _vp,
                                           _vp, _vp, _vp, _vp]),
    "lbdem_hybrid_fine_weight": (ctypes.c_int, [_P(_hip.Lattice), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
})


class HybridLattice3D(PSCLattice3D):
    """Coarse grains by partially saturated cells, fine grains by the equivalent IBM, in one
    lattice (hybridlattice.py:88-845).  Collision with lagged (reference) or fixed-point
    Smagorinsky, as BasicLattice3D.computeOmega.  Shares the coarse-grain machinery of
    PSCLattice3D (keys, snapshot, force reduction); `coarse_id`, `coarse_volfrac` and
    `coarse_velsolid` are decoded from the keys on request, `coarse_weight` from them and the
    relaxation frequency field."""
    _skip_types = CellType.OBSTACLE | CellType.VEL_LADD | CellType.FREE_SLIP | CellType.VEL_INLET_LADD
    _coupling = _hip.COUPLE_HYBRID

    def __init__(self, Nx: int, Ny: int, Nz: int, omega: float, dx: float, dt: float, rho: float, dem_solver,
                 storage=None, smagorinsky=None):
        super().__init__(Nx, Ny, Nz, omega, dx, dt, rho, dem_solver, storage=storage, smagorinsky=smagorinsky)
        self.rho0 = rho
        self.nuLu = (1.0 / omega - 0.5) / 3.0
        self.nu = self.nuLu * dx ** 2 / dt
        self.mu = rho * self.nu
        shp = (self.Nx, self.Ny, self.Nz)
        dev = self.device
        self.coarse_id = self.id
        self.coarse_volfrac = self.volfrac
        self.coarse_velsolid = self.velsolid
        self.coarse_weight = Field(shp, torch.float32, dev, refresh=self._refresh_coarse_weight)
        self.fine_volfrac = Field(shp, torch.float32, dev)
        self.fine_velsolid = Field(shp, torch.float32, dev, n=3)
        self.fine_weight = Field(shp, torch.float32, dev, refresh=self._refresh_fine_weight)
        self._fine_velsum = Field(shp, torch.float32, dev, n=3)
        self._fine_weightsum = Field(shp, torch.float32, dev)
        self.hydrotorque = Field(shp, torch.float32, dev, n=3)  # declared, never written (:154)
        self._status = torch.zeros(2, dtype=torch.int32, device=dev)
        self._fine_weight_t = -1
        self._coarse_weight_t = -1
        self._coarse_list = None
        self._coarse_list_version = None

    def _collision(self):
        return BasicLattice3D._collision(self)

    def _coupling_args(self, a):
        super()._coupling_args(a)
        a.volfrac = self.fine_volfrac.dev.data_ptr()
        a.velsolid = self.fine_velsolid.dev.data_ptr()

    # ---------------------------------------------------------------------------------
    def initialize(self):
        """f <- feq at the collided cells (hybridlattice.py:158-171); no mapping."""
        BasicLattice3D.initialize(self)

    def _coarse(self):
        """Device list of the coarse grains (2R > dx), rebuilt when the radii change."""
        r = self.dem.gf.radius
        if self._coarse_list is None or self._coarse_list_version != r.version:
            rad = r.sync()
            lst = torch.nonzero(2.0 * rad > self.unit.dx).reshape(-1).to(torch.int32)
            if self._coarse_list is not None and not torch.equal(lst, self._coarse_list):
                # the keys are cleared through the listed grains' boxes: start both buffers clean
                for k, b in zip(self._keys, self._bboxes):
                    k.zero_()
                    b.zero_()
                self._dirty = [False, False]
            self._coarse_list = lst
            self._coarse_list_version = r.version
        return self._coarse_list

    def map_coarse_grains(self):
        """Coarse grains to packed keys (hybridlattice.py:173-234), one workgroup column per
        coarse grain; keys in the spare buffer of the two, as PSCLattice3D.grains2lattice."""
        self._prepare()
        g, u = self._structs()
        lst = self._coarse()
        b = 1 - self._cur
        _hip.check(_hip.lib().lbdem_hybrid_map_coarse(
            self._lat_c, _hip.ptr(self._flags), g, u, _hip.ptr(lst), int(lst.numel()), _hip.ptr(self._keys[b]),
            _hip.ptr(self._snap), _hip.ptr(self._bboxes[b]), 1 if self._dirty[b] else 0, _hip.stream()),
            "lbdem_hybrid_map_coarse")
        self._cur = b
        self._dirty[b] = True
        self._map_count += 1
        self._coarse_weight_t = -1

    grains2lattice = map_coarse_grains

    def map_fine_grains(self):
        """Fine-grain fractions and solid velocity (hybridlattice.py:260-376); call after
        map_coarse_grains, whose keys give the room left to the fine grains."""
        self._prepare()
        g, u = self._structs()
        self.fine_volfrac.sync()
        self.fine_velsolid.sync()
        _hip.check(_hip.lib().lbdem_hybrid_map_fine(
            self._lat_c, _hip.ptr(self._flags), g, u, _hip.ptr(self._key), _hip.ptr(self.fine_volfrac.dev),
            _hip.ptr(self._fine_velsum.dev), _hip.ptr(self._fine_weightsum.dev), _hip.ptr(self.fine_velsolid.dev),
            _hip.ptr(self._status), _hip.stream()), "lbdem_hybrid_map_fine")
        for f in (self.fine_volfrac, self.fine_velsolid, self._fine_velsum, self._fine_weightsum):
            f.touch()
        self._fine_weight_t = -1

    def compute_fine_weights(self):
        """The Tenneti weight beta (hybridlattice.py:378-431) is evaluated inside the collision
        from the current fields; `fine_weight` is computed on request for inspection."""

    def _refresh_fine_weight(self):
        if self._fine_weight_t == (self.t, self.fine_volfrac.version):
            return
        self._refresh_macros()
        _hip.check(_hip.lib().lbdem_hybrid_fine_weight(
            self._lat_c, _hip.ptr(self._flags), _hip.ptr(self._key), _hip.ptr(self.fine_volfrac.dev),
            _hip.ptr(self.fine_velsolid.dev), _hip.ptr(self.vel.dev), _hip.ptr(self.fine_weight.dev),
            _hip.stream()), "lbdem_hybrid_fine_weight")
        self._fine_weight_t = (self.t, self.fine_volfrac.version)
        self.fine_weight._host = None

    def _refresh_coarse_weight(self):
        """B = eps (tau - 1/2) / ((1 - eps) + (tau - 1/2)) of the last collision (_psc_compute_weight,
        :236-244): tau from the relaxation frequency field, which holds the step's omega with
        lagged Smagorinsky (omega0 otherwise)."""
        if self._coarse_weight_t == (self.t, self._map_count):
            return
        self._decode()
        eps = self.volfrac.dev.to(torch.float64)
        tm = 1.0 / self.omega.dev.to(torch.float64) - 0.5
        B = torch.where(eps > 0, eps * tm / ((1.0 - eps) + tm), torch.zeros_like(eps))
        self.coarse_weight.dev.copy_(B.to(torch.float32))
        self.coarse_weight._host = None
        self._coarse_weight_t = (self.t, self._map_count)

    def collide(self):
        """Fused collide-stream with the four-way weighted collision (hybridlattice.py:433-593)."""
        BasicLattice3D.collide(self)
        self.hydroforce.touch()

    def lattice2grains(self):
        """Coarse force and torque, and the fine grains' Tenneti drag (hybridlattice.py:595-740),
        into force_fluid / moment_fluid (physical units, FP64).  The fine drag reads the
        post-stream velocity field, refreshed here."""
        self._refresh_macros()
        g, u = self._structs()
        lst = self._coarse()
        other = 1 - self._cur
        stale = (_hip.ptr(self._keys[other]), _hip.ptr(self._bboxes[other])) if self._dirty[other] else (None, None)
        _hip.check(_hip.lib().lbdem_hybrid_forces(
            self._lat_c, _hip.ptr(self._flags), g, u, _hip.ptr(lst), int(lst.numel()), _hip.ptr(self._key),
            _hip.ptr(self._cell_force), stale[0], stale[1], _hip.ptr(self.fine_volfrac.dev), _hip.ptr(self.vel.dev),
            _hip.stream()), "lbdem_hybrid_forces")
        self._dirty[other] = False
        self.dem.gf.force_fluid.touch()
        self.dem.gf.moment_fluid.touch()

    # ---------------------------------------------------------------------------------
    def initialize_complete(self):
        """initialize, map_coarse_grains, map_fine_grains, compute_fine_weights (:808-816)."""
        self.initialize()
        self.map_coarse_grains()
        self.map_fine_grains()
        self.compute_fine_weights()

    def prepare_step(self):
        """Grain mappings before the collision (:818-831)."""
        self.map_coarse_grains()
        self.map_fine_grains()
        self.compute_fine_weights()

    def update_coupling(self):
        """Deprecated in the reference: prepare_step, then lattice2grains (:833-845)."""
        self.prepare_step()
        self.lattice2grains()

    def clamp_count(self):
        """Cells whose fine fraction overflowed the free room by more than 1e-6 (the reference's
        warnings in map_fine_grains)."""
        return int(self._status[0].item())
