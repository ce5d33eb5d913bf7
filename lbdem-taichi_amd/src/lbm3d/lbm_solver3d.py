"""D3Q19 lattice with BGK + Smagorinsky collision (API of src/lbm3d/lbm_solver3d.py).

`BasicLattice3D` keeps the reference's constructor, constants, fields and step methods
(`initialize`, `collide`, `stream`), but a time step is one fused HIP kernel: `collide()`
enqueues the collide-stream update of the populations (in place, AA access pattern) and
`stream()` has nothing left to do.  Macroscopic `rho` / `vel` are produced lazily from the
populations when they are read.  The reference's `feq` / `fpc` scratch fields do not exist;
`feq` is reconstructed on request.
"""
import ctypes

import numpy as np
import torch

from .. import _hip
from .._field import Field
from ..runtime import config
from .lbmutils import CellType, UnitConverter

_TYPE_BITS = (CellType.FLUID | CellType.INTERFACE | CellType.GAS | CellType.OBSTACLE | CellType.FREE_SLIP |
              CellType.VEL_LADD | CellType.VEL_ZOUHE | CellType.VEL_EXIT | CellType.Pre_ZOUHE |
              CellType.VEL_INLET_LADD)


def free_slip_pairs(ct, active, solid, moving):
    """Records (target cell, target direction, source cell, source direction), int32 (n, 4), of
    the free-slip pass (lbdem_lbm_free_slip) for the FREE_SLIP branch of stream
    (lbm_solver3d.py:230-236).  A target is population q of an active cell y whose upstream
    neighbour m = y - c_q is a free-slip cell; the reference pulls fpc[x'][qsy[q]], x' being m
    with its wall-normal index replaced by y's (the normal is the first of LEFT|RIGHT,
    BOTTOM|TOP, BACK|FRONT set on m).  The source is where the step put that post-collision
    value: the bounced population OPP[q'] of x' when x' + c_q' is a wall, else population q' of
    x' + c_q'; -1 (the value 0, the reference's never-written fpc) when x' does not collide.
    Indices wrap periodically like the reference's compute_neigh_index (:308-342)."""
    shape = ct.shape
    C = BasicLattice3D.c.astype(np.int64)
    qsy = (BasicLattice3D.qsyx, BasicLattice3D.qsyy, BasicLattice3D.qsyz)
    fs_cells = np.argwhere((ct & CellType.FREE_SLIP) != 0)
    axis = np.full(len(fs_cells), -1, dtype=np.int64)
    for a, bits in reversed(list(enumerate((CellType.LEFT | CellType.RIGHT, CellType.BOTTOM | CellType.TOP,
                                            CellType.BACK | CellType.FRONT)))):
        axis[(ct[tuple(fs_cells.T)] & bits) != 0] = a
    out = []
    for q in range(1, 19):
        y = (fs_cells + C[q]) % shape
        sel = active[tuple(y.T)]
        if not sel.any():
            continue
        m, y, ax = fs_cells[sel], y[sel], axis[sel]
        if (ax < 0).any():
            raise NotImplementedError("a FREE_SLIP cell next to a fluid cell needs a face flag (LEFT, RIGHT, BOTTOM, "
                                      "TOP, BACK or FRONT): without one the reference leaves the population as it was")
        xp = m.copy()
        xp[np.arange(len(m)), ax] = y[np.arange(len(m)), ax]
        qp = np.array([qsy[a][q] for a in ax], dtype=np.int64)
        coll = active[tuple(xp.T)]
        z = (xp + C[qp]) % shape
        bounced = solid[tuple(z.T)]
        if (coll & bounced & moving[tuple(z.T)]).any():
            raise NotImplementedError("a free-slip reflection whose source population bounced at a moving wall is not "
                                      "supported")
        lin = lambda v: (v[:, 0] * shape[1] + v[:, 1]) * shape[2] + v[:, 2]  # noqa: E731
        src_cell = np.where(coll, np.where(bounced, lin(xp), lin(z)), -1)
        src_dir = np.where(coll, np.where(bounced, BasicLattice3D.qinv[qp], qp), 0)
        out.append(np.column_stack([lin(y), np.full(len(y), q), src_cell, src_dir]))
    return np.concatenate(out).astype(np.int32) if out else np.zeros((0, 4), dtype=np.int32)


class _PopulationField:
    """`lattice.f`: unshifted populations of the post-stream state, shape (Nx, Ny, Nz, 19)."""

    def __init__(self, lattice):
        self._lat = lattice
        self.shape = (lattice.Nx, lattice.Ny, lattice.Nz)
        self.n = 19

    def to_torch(self, dtype=torch.float64):
        L = self._lat
        L._prepare()
        out = torch.empty((19, L.N), dtype=dtype, device=L.device)
        _hip.check(_hip.lib().lbdem_lbm_read_f(L._lat_c, L._args_c(), _hip.ptr(out), int(dtype == torch.float64),
                                               _hip.stream()), "lbdem_lbm_read_f")
        return out

    def to_numpy(self):
        L = self._lat
        return self.to_torch().T.reshape(L.Nx, L.Ny, L.Nz, 19).cpu().numpy()

    def from_numpy(self, arr):
        L = self._lat
        arr = np.asarray(arr, dtype=np.float64)
        if arr.shape != (L.Nx, L.Ny, L.Nz, 19):
            raise ValueError(f"expected shape {(L.Nx, L.Ny, L.Nz, 19)}, got {arr.shape}")
        L._prepare()
        src = torch.from_numpy(np.ascontiguousarray(arr.reshape(L.N, 19).T)).to(L.device)
        L._write_populations(src)
        L._macro_t = -1

    def __getitem__(self, idx):
        return self.to_numpy()[idx]


class BasicLattice3D:
    """LBM D3Q19, cs^2 = 1/3, w0 = 1/3, w1-6 = 1/18, w7-18 = 1/36 (reference lbm_solver3d.py:18-97)."""
    D = 3
    Q = 19
    qinv = np.array([0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17], dtype=np.int32)
    qsyx = np.array([0, 2, 1, 3, 4, 5, 6, 10, 9, 8, 7, 14, 13, 12, 11, 15, 16, 17, 18], dtype=np.int32)
    qsyy = np.array([0, 1, 2, 4, 3, 5, 6, 9, 10, 7, 8, 11, 12, 13, 14, 18, 17, 16, 15], dtype=np.int32)
    qsyz = np.array([0, 1, 2, 3, 4, 6, 5, 7, 8, 9, 10, 13, 14, 11, 12, 17, 18, 15, 16], dtype=np.int32)
    c = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                  [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
                  [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]], dtype=np.float64)
    w = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12, dtype=np.float64)
    velmax = (2.0 / 3.0) ** 0.5
    cssq = 1.0 / 3.0
    SmagorinskyConstant = 0.1  # read at every step, so drivers may change it (e.g. 0.17)

    # cells that collide() / stream() skip (lbm_solver3d.py:183)
    _skip_types = CellType.OBSTACLE | CellType.VEL_LADD | CellType.FREE_SLIP | CellType.VEL_INLET_LADD
    _coupling = _hip.COUPLE_NONE
    # collide-stream variant: 0 automatic, 1 one cell per thread, 2 row-tiled
    kernel = 0

    def __init__(self, Nx: int, Ny: int, Nz: int, omega: float, dx: float, dt: float, rho: float,
                 storage=None, smagorinsky=None):
        _hip.require_gpu()
        self.Nx, self.Ny, self.Nz = int(Nx), int(Ny), int(Nz)
        self.N = self.Nx * self.Ny * self.Nz
        self.omega0 = float(omega)
        self.storage = storage or config.storage
        if self.storage not in _hip.STORE:
            raise ValueError(f"unknown storage {self.storage!r}")
        self.smagorinsky = smagorinsky or config.smagorinsky
        self.device = config.torch_device
        shp = (self.Nx, self.Ny, self.Nz)
        dev = self.device

        self.CT = Field(shp, torch.int32, dev)
        self.CT.fill(CellType.FLUID)
        self.rho = Field(shp, torch.float32, dev, refresh=self._refresh_macros)
        self.rho.fill(1.0)
        self.vel = Field(shp, torch.float32, dev, n=3, refresh=self._refresh_macros)
        self.unit = UnitConverter(dx, dt, rho)

        self._f = torch.zeros(19 * self.N, dtype=_hip.STORE_DTYPE[self.storage], device=dev)
        self._flags = torch.zeros(self.N, dtype=torch.uint8, device=dev)
        self._nmask = torch.zeros(self.N, dtype=torch.int32, device=dev)  # wall-neighbour bits
        self._lag = None
        if self._collision() == _hip.COLL_SMAG_LAGGED:
            lag_dtype = torch.float64 if self.storage == "f64" else torch.float32
            self._lag = torch.zeros(5 * self.N, dtype=lag_dtype, device=dev)
            self.omega = Field(shp, lag_dtype, dev, tensor=self._lag[4 * self.N:].view(shp))
        else:
            self.omega = Field(shp, torch.float32, dev)
        self.omega.fill(self.omega0)
        self.f = _PopulationField(self)

        self.t = 0                 # steps taken (AA parity of the population array)
        self._macro_t = 0          # step at which rho/vel were last refreshed
        self._ct_version = -1
        self._lat_c = _hip.Lattice()
        self._args = _hip.StepArgs()
        self._sync_lattice_struct()

    # ----------------------------------------------------------------------------------
    # configuration of the fused kernel
    # ----------------------------------------------------------------------------------
    def _collision(self):
        return _hip.COLL_SMAG_LAGGED if self.smagorinsky == "lagged" else _hip.COLL_SMAG_FIXED

    def _body_force(self):
        return (0.0, 0.0, 0.0)

    def _sync_lattice_struct(self):
        L = self._lat_c
        L.nx, L.ny, L.nz = self.Nx, self.Ny, self.Nz
        L.storage = _hip.STORE[self.storage]
        L.collision = self._collision()
        L.coupling = self._coupling
        L.omega0 = self.omega0
        L.cs_smag = float(BasicLattice3D.SmagorinskyConstant)
        L.rho0 = 1.0
        fx, fy, fz = self._body_force()
        L.force[0], L.force[1], L.force[2] = float(fx), float(fy), float(fz)
        L.kernel = int(self.kernel)

    def _coupling_args(self, a):
        """Subclasses fill the coupling members of the step arguments."""

    def _args_c(self):
        a = self._args
        a.f = self._f.data_ptr()
        a.flags = self._flags.data_ptr()
        a.nmask = self._nmask.data_ptr()
        a.vel = self.vel.dev.data_ptr()
        a.lag = self._lag.data_ptr() if self._lag is not None else None
        a.t = self.t
        a.nu_lu = (1.0 / self.omega0 - 0.5) / 3.0
        a.rho = self.rho.dev.data_ptr()
        self._coupling_args(a)
        return ctypes.byref(a)

    # ----------------------------------------------------------------------------------
    # cell flags: CellType -> one byte per cell (active / near-wall / wall / moving / fluid)
    # ----------------------------------------------------------------------------------
    def _build_flags(self):
        ct = self.CT.to_numpy().astype(np.int64)
        kind = ct & _TYPE_BITS
        wet = {CellType.VEL_ZOUHE: _hip.BC_VEL_ZOUHE, CellType.Pre_ZOUHE: _hip.BC_PRE_ZOUHE,
               CellType.VEL_EXIT: _hip.BC_EXIT, CellType.VEL_INLET_LADD: _hip.BC_LADD_INLET}
        supported = ((kind == CellType.FLUID) | (kind == CellType.OBSTACLE) | (kind == CellType.VEL_LADD) |
                     (kind == CellType.FREE_SLIP))
        for t in wet:
            supported |= kind == t
        if not supported.all():
            bad = np.unique(kind[~supported])
            raise NotImplementedError(
                f"cell types {bad.tolist()} are not supported by the fused kernel (supported: FLUID, "
                "OBSTACLE, VEL_LADD, FREE_SLIP, VEL_ZOUHE, Pre_ZOUHE, VEL_EXIT, VEL_INLET_LADD, one each)")
        active = (ct & self._skip_types) == 0
        free_slip = kind == CellType.FREE_SLIP
        if (free_slip & active).any():
            raise NotImplementedError("FREE_SLIP cells must be among the skipped cell types")
        wall = (kind == CellType.OBSTACLE) | free_slip
        inlet = kind == CellType.VEL_INLET_LADD
        if (inlet & active).any():
            raise NotImplementedError("VEL_INLET_LADD cells must be among the skipped cell types")
        moving = (kind == CellType.VEL_LADD) | inlet
        fluid = kind == CellType.FLUID
        flags = (active * _hip.FLAG_ACTIVE | wall * _hip.FLAG_WALL | moving * _hip.FLAG_MOVING |
                 inlet * _hip.FLAG_INLET | fluid * _hip.FLAG_FLUID).astype(np.uint8)
        self._flags.copy_(torch.from_numpy(flags.reshape(-1)))
        pairs = free_slip_pairs(ct, active, wall | moving, moving) if free_slip.any() else None
        self._fs_pairs = None if pairs is None or not len(pairs) else \
            torch.from_numpy(pairs.reshape(-1)).to(self.device)
        self._fs_work = None if self._fs_pairs is None else \
            torch.empty(len(pairs), dtype=torch.float64, device=self.device)
        # wet-node boundary cells, applied after every step: type | face << 4, the face being the
        # first of LEFT, RIGHT, BOTTOM, TOP, BACK, FRONT set (7: none)
        bc_type = np.zeros(kind.shape, dtype=np.int64)
        for t, code in wet.items():
            bc_type[kind == t] = code
        face = np.full(kind.shape, 7, dtype=np.int64)
        for f, bit in reversed(list(enumerate((CellType.LEFT, CellType.RIGHT, CellType.BOTTOM, CellType.TOP,
                                                CellType.BACK, CellType.FRONT)))):
            face[(ct & bit) != 0] = f
        cells = np.nonzero(bc_type.reshape(-1))[0]
        self._bc_cells = torch.from_numpy(cells.astype(np.int32)).to(self.device)
        self._bc_kinds = torch.from_numpy((bc_type | face << 4).reshape(-1)[cells].astype(np.uint8)).to(self.device)
        dims = _hip.Lattice()
        dims.nx, dims.ny, dims.nz = self.Nx, self.Ny, self.Nz
        # NEAR bits and wall-neighbour masks are derived on the device
        _hip.check(_hip.lib().lbdem_lbm_prepare_flags(ctypes.byref(dims), _hip.ptr(self._flags),
                                                       _hip.ptr(self._nmask), _hip.stream()),
                   "lbdem_lbm_prepare_flags")
        self._ct_version = self.CT.version
        self._active_np = active

    def _write_populations(self, src):
        """Store unshifted FP64 populations (19, N) as the state after self.t steps."""
        _hip.check(_hip.lib().lbdem_lbm_write_f(self._lat_c, self._args_c(), _hip.ptr(src), 1, _hip.stream()),
                   "lbdem_lbm_write_f")
        self._populated = True

    def _prepare(self):
        """Bring device inputs up to date before a kernel launch."""
        if self.CT.version != self._ct_version:
            # The populations' layout depends on the walls (bounced populations are kept in the
            # wall cells): read them out with the old flags and store them back with the new ones,
            # so the next step starts from the same post-stream state, as in the reference.
            snap = None
            if getattr(self, "_populated", False):
                snap = torch.empty((19, self.N), dtype=torch.float64, device=self.device)
                _hip.check(_hip.lib().lbdem_lbm_read_f(self._lat_c, self._args_c(), _hip.ptr(snap), 1,
                                                       _hip.stream()), "lbdem_lbm_read_f")
            self._build_flags()
            if snap is not None:
                self._sync_lattice_struct()
                self._write_populations(snap)
        self.CT.sync()
        self.vel.sync()
        self.rho.sync()
        self.omega.sync()
        self._sync_lattice_struct()

    def _refresh_macros(self):
        if self._macro_t == self.t:
            return
        self._prepare()
        _hip.check(_hip.lib().lbdem_lbm_macros(self._lat_c, self._args_c(), _hip.ptr(self.rho.dev),
                                               _hip.ptr(self.vel.dev), None, _hip.stream()), "lbdem_lbm_macros")
        self._macro_t = self.t
        self.rho.touch()
        self.vel.touch()

    # ----------------------------------------------------------------------------------
    # reference step API
    # ----------------------------------------------------------------------------------
    def initialize(self):
        """f <- feq(rho, vel) at collided cells (lbm_solver3d.py:157-171)."""
        self._prepare()
        self.t = 0
        _hip.check(_hip.lib().lbdem_lbm_init(self._lat_c, self._args_c(), _hip.ptr(self.rho.dev), _hip.stream()),
                   "lbdem_lbm_init")
        self._populated = True
        self._macro_t = 0
        if self._lag is not None:
            self.omega.touch()

    # optional (start, end) event pairs recorded around every collide-stream launch
    step_events = None

    def collide(self):
        """One fused collide + stream step (lbm_solver3d.py:176-250)."""
        self._prepare()
        ev = self.step_events
        if ev is not None:
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
        _hip.check(_hip.lib().lbdem_lbm_step(self._lat_c, self._args_c(), _hip.stream()), "lbdem_lbm_step")
        if ev is not None:
            e1.record()
            ev.append((e0, e1))
        self.t += 1
        if self._fs_pairs is not None:  # free-slip reflection on the post-stream state (:230-236)
            _hip.check(_hip.lib().lbdem_lbm_free_slip(self._lat_c, self._args_c(), _hip.ptr(self._fs_pairs),
                                                      int(self._fs_work.numel()), _hip.ptr(self._fs_work),
                                                      _hip.stream()), "lbdem_lbm_free_slip")
        nbc = int(self._bc_cells.numel())
        if nbc:  # wet-node boundaries on the post-stream state (lbm_solver3d.py:241-250)
            _hip.check(_hip.lib().lbdem_lbm_boundaries(self._lat_c, self._args_c(), _hip.ptr(self._bc_cells),
                                                       _hip.ptr(self._bc_kinds), nbc, _hip.ptr(self.rho.dev),
                                                       _hip.ptr(self.vel.dev), _hip.stream()),
                       "lbdem_lbm_boundaries")
            self.rho.touch()
            self.vel.touch()
        if self._lag is not None:
            self.omega.touch()

    def stream(self):
        """Streaming and the wet-node boundary pass run in collide(); nothing is left to do
        here (lbm_solver3d.py:197-250)."""

    def step(self, n=1):
        for _ in range(n):
            self.collide()
            self.stream()

    # ----------------------------------------------------------------------------------
    # stress diagnostics (lbm_solver3d.py:347-456)
    # ----------------------------------------------------------------------------------
    _STRESS = ("stress_xx", "stress_yy", "stress_zz", "stress_xy", "stress_xz", "stress_yz", "shear_stress_mag")

    def _stress_fields(self):
        """Seven FP32 planes, allocated at the first use (28 B per cell)."""
        if getattr(self, "_stress", None) is None:
            shp = (self.Nx, self.Ny, self.Nz)
            self._stress = torch.zeros((7, self.N), dtype=torch.float32, device=self.device)
            for k, name in enumerate(self._STRESS):
                self.__dict__[name] = Field(shp, torch.float32, self.device, tensor=self._stress[k].view(shp))
        return self._stress

    def __getattr__(self, name):
        if name in BasicLattice3D._STRESS:
            self._stress_fields()
            return self.__dict__[name]
        raise AttributeError(name)

    def compute_stress(self):
        """sigma = -(1 - omega/2) sum_q c c (f - feq) per collided cell, with feq and omega of the
        last collision (exact with lagged Smagorinsky; otherwise those of the current
        populations), into stress_xx ... stress_yz and shear_stress_mag (:347-405)."""
        self._prepare()
        st = self._stress_fields()
        _hip.check(_hip.lib().lbdem_lbm_stress(self._lat_c, self._args_c(), _hip.ptr(st), _hip.stream()),
                   "lbdem_lbm_stress")
        for name in self._STRESS:
            self.__dict__[name].touch()

    def compute_plane_average_stress(self, axis: int, position: int):
        """(mean xy, mean xz, mean yz) shear stress over the FLUID cells of the plane
        {axis = position} (:407-456), from the last compute_stress."""
        st = self._stress_fields()
        out = torch.empty(3, dtype=torch.float64, device=self.device)
        _hip.check(_hip.lib().lbdem_lbm_plane_stress(self._lat_c, _hip.ptr(self._flags), _hip.ptr(st), int(axis),
                                                     int(position), _hip.ptr(out), _hip.stream()),
                   "lbdem_lbm_plane_stress")
        return tuple(float(v) for v in out.cpu())

    # ----------------------------------------------------------------------------------
    # restart snapshot (no counterpart in the reference)
    # ----------------------------------------------------------------------------------
    def _snapshot_fields(self):
        """Fields a run resumes from: inputs and state, not those derived on request."""
        seen, out = set(), {}
        for name, val in vars(self).items():
            if not isinstance(val, Field) or id(val) in seen or name in self._STRESS:
                continue
            seen.add(id(val))
            if val._refresh is not None and val is not self.rho and val is not self.vel:
                continue  # decoded / inspection fields, recomputed from the state
            out[name] = val
        return out

    def state_dict(self):
        """The populations (the stored array, and unshifted FP64 for a restore into another
        storage type), the step counter, the lagged-Smagorinsky state and every input or state
        field (CT, rho, vel, omega, coupling fields), as numpy arrays."""
        self._prepare()
        pops = torch.empty((19, self.N), dtype=torch.float64, device=self.device)
        _hip.check(_hip.lib().lbdem_lbm_read_f(self._lat_c, self._args_c(), _hip.ptr(pops), 1, _hip.stream()),
                   "lbdem_lbm_read_f")
        d = {"f": pops.cpu().numpy(), "t": np.array(self.t), "shape": np.array([self.Nx, self.Ny, self.Nz]),
             "storage": np.array(self.storage),
             # the stored array itself (its layout follows the parity and the walls): exact restore
             # into a lattice of the same storage type and cell types
             "f_raw": self._f.view(torch.uint8).cpu().numpy()}
        if self._lag is not None:
            d["lag"] = self._lag.cpu().numpy()
        for name, f in self._snapshot_fields().items():
            d["field." + name] = f.to_numpy()
        return d

    def load_state_dict(self, d):
        """Restore a state_dict() into a lattice of the same shape and class: the next step
        continues exactly where the snapshot was taken."""
        if tuple(int(v) for v in d["shape"]) != (self.Nx, self.Ny, self.Nz):
            raise ValueError(f"snapshot of shape {tuple(d['shape'])} does not fit {(self.Nx, self.Ny, self.Nz)}")
        fields = self._snapshot_fields()
        self.CT.from_numpy(d["field.CT"])
        self._populated = False  # the populations are written below, after the flags
        self.t = int(d["t"])
        self._prepare()
        if "f_raw" in d and str(d["storage"]) == self.storage:
            self._f.view(torch.uint8).copy_(torch.from_numpy(np.ascontiguousarray(d["f_raw"])))
            self._populated = True
        else:  # another storage type: through the unshifted FP64 populations
            self._write_populations(torch.from_numpy(np.ascontiguousarray(d["f"])).to(self.device))
        if self._lag is not None and "lag" in d:
            self._lag.copy_(torch.from_numpy(np.ascontiguousarray(d["lag"])))
        for name, f in fields.items():
            if name != "CT" and "field." + name in d:
                f.from_numpy(d["field." + name])
        self._macro_t = self.t
        self._after_load()

    def _after_load(self):
        """Subclasses reset state derived from what was restored."""

    def save_state(self, file_name):
        """Write state_dict() to an .npz file (no pickled objects)."""
        np.savez(file_name, **self.state_dict())

    def load_state(self, file_name):
        with np.load(file_name, allow_pickle=False) as z:
            self.load_state_dict({k: z[k] for k in z.files})

    # ----------------------------------------------------------------------------------
    # diagnostics
    # ----------------------------------------------------------------------------------
    def _reduce(self):
        self._prepare()
        work = torch.empty(int(_hip.lib().lbdem_reduce_work_doubles(self._lat_c)), dtype=torch.float64,
                           device=self.device)
        out = torch.empty(5, dtype=torch.float64, device=self.device)
        _hip.check(_hip.lib().lbdem_lbm_reduce(self._lat_c, self._args_c(), _hip.ptr(work), _hip.ptr(out),
                                               _hip.stream()), "lbdem_lbm_reduce")
        return out.cpu().numpy()

    def get_total_mass(self) -> float:
        """FP64 sum of rho over collided cells."""
        return float(self._reduce()[0])

    def get_total_momentum(self):
        """FP64 sum of rho*u over collided cells."""
        return self._reduce()[1:4].copy()

    def get_kinetic_energy(self) -> float:
        """FP64 sum of rho|u|^2/2 over collided cells."""
        return float(self._reduce()[4])

    @property
    def feq(self):
        """Equilibrium of the last collision, feq(rho, u), rebuilt on the host."""
        return _FeqView(self)

    @property
    def rate(self):
        """Strain-rate magnitude of the last collision (the `rate` field, lbm_solver3d.py:133,
        461-498), recovered from that collision's relaxation frequency, which the lagged
        Smagorinsky keeps: omega = 1 / (1/omega0 + Cs^2 rate / cs^2) (:500-513)."""
        return _RateView(self)

    @property
    def fpc(self):
        raise AttributeError("post-collision populations are not materialised: the collide-stream "
                             "kernel writes them straight into the streamed slots")


class _RateView:
    def __init__(self, lattice):
        self._lat = lattice

    def to_numpy(self):
        L = self._lat
        cs = float(type(L).SmagorinskyConstant)
        if L._lag is None or cs == 0.0:
            raise AttributeError("rate is kept only by the lagged Smagorinsky collision (smagorinsky='lagged', "
                                 "SmagorinskyConstant > 0): the other collisions store no relaxation frequency")
        om = L.omega.to_numpy().astype(np.float64)
        return (1.0 / om - 1.0 / L.omega0) * L.cssq / (cs * cs)


class _FeqView:
    def __init__(self, lattice):
        self._lat = lattice

    def to_numpy(self):
        L = self._lat
        if L._lag is not None:
            lag = L._lag.view(5, L.Nx, L.Ny, L.Nz).cpu().numpy().astype(np.float64)
            rho, u = lag[0], np.stack([lag[1], lag[2], lag[3]], axis=-1)
        else:
            rho, u = L.rho.to_numpy().astype(np.float64), L.vel.to_numpy().astype(np.float64)
        cu = u @ L.c.T
        usq = (u * u).sum(-1, keepdims=True)
        return L.w * rho[..., None] * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq)
