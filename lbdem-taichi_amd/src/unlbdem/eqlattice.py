"""Unresolved LBM-DEM coupling, particle-equivalence IBM (API of src/unlbdem/eqlattice.py).

Per step, as in the reference driver: DEM substeps, `collide()` (fused collide-stream with
lagged Smagorinsky and the Tenneti-weighted solid collision evaluated inline from the solid
fraction, solid velocity and the cell's own fluid velocity), `stream()` (nothing left), then
`update_coupling()`: `grains2lattice()` (spreading + finalisation on the device),
`compute_weight()` (inline in the next collision; kept for the API) and `lattice2grains()`
(interpolation + Tenneti drag into force_fluid).
"""
import torch

from .. import _hip
from .._field import Field
from ..lbm3d.lbm_solver3d import BasicLattice3D
from ..lbm3d.lbmutils import CellType
from ._common import GrainStructs


class EqIMBlattice3D(BasicLattice3D):
    QINV_STATIC = (0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17)
    _skip_types = CellType.OBSTACLE | CellType.VEL_LADD | CellType.FREE_SLIP | CellType.VEL_INLET_LADD
    _coupling = _hip.COUPLE_EQIBM

    def __init__(self, Nx: int, Ny: int, Nz: int, omega: float, dx: float, dt: float, rho: float, demslover,
                 storage=None, smagorinsky=None):
        super().__init__(Nx, Ny, Nz, omega, dx, dt, rho, storage=storage, smagorinsky=smagorinsky)
        self.rho0 = rho
        self.nuLu = (1.0 / omega - 0.5) / 3.0
        self.nu = self.nuLu * self.unit.dx ** 2 / self.unit.dt
        self.mu = rho * self.nu
        self.omega0 = omega
        self.cssq = 1.0 / 3.0
        self.dem = demslover
        shp = (self.Nx, self.Ny, self.Nz)
        dev = self.device
        self.volfrac = Field(shp, torch.float32, dev)
        self.velsolid = Field(shp, torch.float32, dev, n=3)
        self.velsum = Field(shp, torch.float32, dev, n=3)
        self.weight_sum = Field(shp, torch.float32, dev)
        self.weight = Field(shp, torch.float32, dev, refresh=self._refresh_weight)
        self._status = torch.zeros(2, dtype=torch.int32, device=dev)
        self._gs = None
        self._weight_t = -1

    def _collision(self):
        return _hip.COLL_SMAG_LAGGED if self.smagorinsky == "lagged" else _hip.COLL_SMAG_FIXED

    def _coupling_args(self, a):
        a.volfrac = self.volfrac.dev.data_ptr()
        a.velsolid = self.velsolid.dev.data_ptr()

    def _structs(self):
        if self._gs is None:
            self._gs = GrainStructs(self)
        else:
            self.dem.gf.sync()
        # spreading and drag visit the grains in the DEM's bucket order when there is one
        order = self.dem.spatial_order()
        self._gs.g.order = order.data_ptr() if order is not None else None
        return self._gs

    def grains2lattice(self):
        """Spread solid fraction and momentum with the mirror 3-point kernel (eqlattice.py:104-194)."""
        self._prepare()
        gs = self._structs()
        self.volfrac.sync()
        _hip.check(_hip.lib().lbdem_ibm_spread(
            self._lat_c, self._args_c(), gs.gp, gs.up, _hip.ptr(self.volfrac.dev), _hip.ptr(self.velsum.dev),
            _hip.ptr(self.weight_sum.dev), _hip.ptr(self.velsolid.dev), _hip.ptr(self.rho.dev),
            _hip.ptr(self.vel.dev), _hip.ptr(self._status), _hip.stream()), "lbdem_ibm_spread")
        for f in (self.volfrac, self.velsum, self.weight_sum, self.velsolid):
            f.touch()
        self._weight_t = -1

    def compute_weight(self):
        """The Tenneti weight (eqlattice.py:200-238) is evaluated inside the collision kernel from
        the current fields; `weight` is computed on request for inspection."""

    def _refresh_weight(self):
        if self._weight_t == self.t:
            return
        self._refresh_macros()
        _hip.check(_hip.lib().lbdem_ibm_weight(self._lat_c, _hip.ptr(self._flags), _hip.ptr(self.volfrac.dev),
                                               _hip.ptr(self.velsolid.dev), _hip.ptr(self.vel.dev),
                                               _hip.ptr(self.weight.dev), _hip.stream()), "lbdem_ibm_weight")
        self._weight_t = self.t
        self.weight._host = None

    def lattice2grains(self):
        """Interpolated fluid velocity and Tenneti drag on each particle (eqlattice.py:309-444)."""
        gs = self._structs()
        _hip.check(_hip.lib().lbdem_ibm_drag(self._lat_c, _hip.ptr(self._flags), gs.gp, gs.up,
                                             _hip.ptr(self.volfrac.dev), _hip.ptr(self.vel.dev), _hip.stream()),
                   "lbdem_ibm_drag")
        self.dem.gf.force_fluid.touch()

    def initialize_complete(self):
        """initialize, grains2lattice, compute_weight (eqlattice.py:533-537)."""
        self.initialize()
        self.grains2lattice()
        self.compute_weight()

    def update_coupling(self):
        """grains2lattice, compute_weight, lattice2grains (eqlattice.py:539-543)."""
        self.grains2lattice()
        self.compute_weight()
        self.lattice2grains()

    def clamp_count(self):
        """Number of cells whose solid fraction reached 1 and was set to 0.99 (warnings in the reference)."""
        return int(self._status[0].item())
