"""Unresolved LBM-DEM coupling, volume-averaged Navier-Stokes (API of src/unlbdem/vanslattice.py).

Step order of the reference driver (one_particle.py:236-246): grains2lattice,
compute_hydrodynamic_force, DEM substeps, force2lattice, collide, stream.  The collision uses
the volume fraction and force of this step's force2lattice for the equilibrium, relaxation
and Guo-type force term, and those of the previous step for the macroscopic velocity (as the
reference, whose velocity field comes from the previous stream); both are kept on the device.
Deviations (both stated): the reference's shared `temp_weights` scratch, written by grains in
parallel, is replaced by per-grain weights; its wet-node boundary loop nested in the cell loop
runs once after streaming (there are no wet-node cells in the supported cell types).
"""
import ctypes

import torch

from .. import _hip
from .._field import Field
from ..lbm3d.lbm_solver3d import BasicLattice3D
from ..lbm3d.lbmutils import CellType
from ._common import GrainStructs


class Unresolvedlattice3D(BasicLattice3D):
    _skip_types = CellType.OBSTACLE | CellType.VEL_LADD | CellType.FREE_SLIP | CellType.VEL_INLET_LADD
    _coupling = _hip.COUPLE_VANS

    def __init__(self, Nx: int, Ny: int, Nz: int, omega: float, dx: float, dt: float, rho: float, demslover,
                 storage=None):
        super().__init__(Nx, Ny, Nz, omega, dx, dt, rho, storage=storage)
        self.rho0 = rho
        nuLu = (1.0 / omega - 0.5) / 3.0
        self.nu = nuLu * self.unit.dx ** 2 / self.unit.dt
        self.mu = rho * self.nu
        self.dem = demslover
        shp = (self.Nx, self.Ny, self.Nz)
        dev = self.device
        # force2lattice outputs, double-buffered: [cur] for this step, [prev] for the velocity
        self._vf = [torch.zeros(shp, dtype=torch.float32, device=dev) for _ in range(2)]
        self._hf = [torch.zeros(shp + (3,), dtype=torch.float32, device=dev) for _ in range(2)]
        self._cur = 0
        self.volfrac_g2l = Field(shp, torch.float32, dev)  # grains2lattice (used by the drag)
        self.volfrac = Field(shp, torch.float32, dev, tensor=self._vf[0])
        self.hydroforce = Field(shp, torch.float32, dev, n=3, tensor=self._hf[0])
        self._gs = None
        self._fvel_t = -1
        self._fvel = torch.zeros(shp + (3,), dtype=torch.float32, device=dev)
        self._frho = torch.zeros(shp, dtype=torch.float32, device=dev)

    def _collision(self):
        return _hip.COLL_BGK

    def _coupling_args(self, a):
        c, p = self._cur, 1 - self._cur
        a.volfrac = self._vf[c].data_ptr()
        a.hforce = self._hf[c].data_ptr()
        a.volfrac_prev = self._vf[p].data_ptr()
        a.hforce_prev = self._hf[p].data_ptr()

    def _structs(self):
        if self._gs is None:
            self._gs = GrainStructs(self)
        else:
            self.dem.gf.sync()
        return self._gs

    def _refresh_macros(self):
        """rho, vel of the last stream: computed with the force / fraction of the step that
        produced them (compute_rhof_vel, vanslattice.py:412-440)."""
        if self._macro_t == self.t:
            return
        self._prepare()
        args = self._args_c()
        a = self._args
        a.volfrac, a.hforce = self._vf[self._cur].data_ptr(), self._hf[self._cur].data_ptr()
        if self.t == 0:
            return
        _hip.check(_hip.lib().lbdem_lbm_macros(self._lat_c, args, _hip.ptr(self.rho.dev), _hip.ptr(self.vel.dev),
                                               None, _hip.stream()), "lbdem_lbm_macros")
        self._macro_t = self.t
        self.rho.touch()
        self.vel.touch()

    def grains2lattice(self):
        """Solid volume fraction by normalised 3-point spreading (vanslattice.py:135-196)."""
        self._prepare()
        gs = self._structs()
        _hip.check(_hip.lib().lbdem_vans_spread(self._lat_c, _hip.ptr(self._flags), gs.gp, gs.up,
                                                _hip.ptr(self.volfrac_g2l.dev), None, _hip.stream()),
                   "lbdem_vans_spread")
        self.volfrac_g2l.touch()

    def compute_hydrodynamic_force(self):
        """Kernel-weighted fluid velocity and drag on each particle (vanslattice.py:446-482)."""
        self._refresh_macros()
        gs = self._structs()
        _hip.check(_hip.lib().lbdem_vans_drag(self._lat_c, _hip.ptr(self._flags), gs.gp, gs.up,
                                              _hip.ptr(self.volfrac_g2l.dev), _hip.ptr(self.vel.dev),
                                              _hip.stream()), "lbdem_vans_drag")
        self.dem.gf.force_fluid.touch()

    def force2lattice(self):
        """Spread particle volume and fluid force to the lattice (vanslattice.py:489-557)."""
        self._prepare()
        self._refresh_macros()  # the velocity of the last stream uses the current buffers
        self._cur = 1 - self._cur
        gs = self._structs()
        _hip.check(_hip.lib().lbdem_vans_spread(self._lat_c, _hip.ptr(self._flags), gs.gp, gs.up,
                                                _hip.ptr(self._vf[self._cur]), _hip.ptr(self._hf[self._cur]),
                                                _hip.stream()), "lbdem_vans_spread")
        self.volfrac = Field(self.volfrac.shape, torch.float32, self.device, tensor=self._vf[self._cur])
        self.hydroforce = Field(self.volfrac.shape, torch.float32, self.device, n=3, tensor=self._hf[self._cur])

    def initialize(self):
        """f <- VANS equilibrium with the current volume fraction (vanslattice.py:106-127)."""
        self._prepare()
        a = self._args
        self._args_c()
        a.volfrac = self._vf[self._cur].data_ptr()
        self.t = 0
        _hip.check(_hip.lib().lbdem_lbm_init(self._lat_c, ctypes.byref(a), _hip.ptr(self.rho.dev),
                                             _hip.stream()), "lbdem_lbm_init")
        self._populated = True
        self._macro_t = 0
