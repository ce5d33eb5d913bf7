"""ORACLE -- test infrastructure only (imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package).

ctypes wrapper of liblbdem_oracle.so, the C restatement of the reference's algorithms, plus
small drivers that sequence the restated kernels the way the reference's solver classes do.
Arrays follow the reference's layout: lattice fields are [i, j, k] (k fastest), vectors AoS.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblbdem_oracle.so")

FLUID, OBSTACLE, FREE_SLIP, VEL_LADD, VEL_INLET_LADD = 1, 8, 16, 32, 262144
VEL_ZOUHE, VEL_EXIT, PRE_ZOUHE = 64, 128, 256                      # lbmutils.py:69-97
RIGHT, LEFT, TOP, BOTTOM, FRONT, BACK = 4096, 8192, 16384, 32768, 65536, 131072
WET_NODE = VEL_ZOUHE | VEL_EXIT | PRE_ZOUHE | VEL_INLET_LADD
SKIP_BASIC = OBSTACLE | VEL_LADD | FREE_SLIP | VEL_INLET_LADD   # lbm_solver3d.py:183
SKIP_FORCED = OBSTACLE | VEL_LADD | FREE_SLIP                   # lbm_foecelover.py:45, psclattice3d.py:204

C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
              [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]], dtype=np.float64)
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)

_lib = None
_dp = ctypes.c_void_p


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            from . import build_oracle  # noqa: F401  (only when run as package)
        _lib = ctypes.CDLL(LIB_PATH)
        for sfx in ("_f64", "_f32"):
            getattr(_lib, "lbo_threedelta" + sfx).restype = ctypes.c_double
            getattr(_lib, "lbo_threedelta" + sfx).argtypes = [ctypes.c_double]
            getattr(_lib, "lbo_tenneti_cd" + sfx).restype = ctypes.c_double
            getattr(_lib, "lbo_tenneti_cd" + sfx).argtypes = [ctypes.c_double] * 3
            getattr(_lib, "lbo_kernel_mirror" + sfx).restype = ctypes.c_double
            getattr(_lib, "lbo_kernel_mirror" + sfx).argtypes = [ctypes.c_int] * 3 + [ctypes.c_double] * 3 + [ctypes.c_int] * 3
    return _lib


def P(a):
    """Pointer of a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "oracle arrays must be C-contiguous"
    return a.ctypes.data_as(_dp)


def _fn(name, real):
    return getattr(lib(), name + ("_f64" if real == np.float64 else "_f32"))


def _d(x):
    return ctypes.c_double(float(x))


class OracleLattice:
    """Reference-layout lattice state and its two-pass step (collide -> fpc, stream -> f).

    kind: "basic" (lagged Smagorinsky, BasicLattice3D), "forced" (ForcedLattice3D),
    "psc" (PSCLattice3D weighted collision), "eqibm" (EqIMBlattice3D), "vans",
    "hybrid" (HybridLattice3D: id / volfrac / velsolid / weight hold the coarse fields,
    fine_volfrac / fine_velsolid / fine_weight the fine ones).
    """

    def __init__(self, nx, ny, nz, omega, kind="basic", cs=0.1, F=None, real=np.float64):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz
        self.omega0 = omega
        self.kind = kind
        self.cs = cs
        self.real = real
        self.F = None if F is None else np.asarray(F, dtype=real)
        sh = (nx, ny, nz)
        self.CT = np.full(sh, FLUID, dtype=np.int32)
        self.f = np.zeros(sh + (19,), dtype=real)
        self.feq = np.zeros(sh + (19,), dtype=real)
        self.fpc = np.zeros(sh + (19,), dtype=real)
        self.rho = np.ones(sh, dtype=real)
        self.vel = np.zeros(sh + (3,), dtype=real)
        self.omega = np.full(sh, omega, dtype=real)
        self.rate = np.zeros(sh, dtype=real)
        # coupling fields
        self.volfrac = np.zeros(sh, dtype=real)
        self.weight = np.zeros(sh, dtype=real)
        self.velsolid = np.zeros(sh + (3,), dtype=real)
        self.hydroforce = np.zeros(sh + (3,), dtype=real)
        self.id = np.full(sh, -1, dtype=np.int32)
        if kind == "hybrid":
            self.fine_volfrac = np.zeros(sh, dtype=real)
            self.fine_weight = np.zeros(sh, dtype=real)
            self.fine_velsolid = np.zeros(sh + (3,), dtype=real)

    @property
    def skip(self):
        return SKIP_BASIC if self.kind in ("basic", "eqibm", "hybrid") else SKIP_FORCED

    def _g(self):
        return (ctypes.c_int(self.nx), ctypes.c_int(self.ny), ctypes.c_int(self.nz))

    def initialize(self):
        _fn("lbo_initialize", self.real)(*self._g(), P(self.CT), ctypes.c_int32(self.skip), P(self.rho),
                                         P(self.vel), P(self.feq), P(self.f))

    def collide(self):
        R = ctypes.c_double
        if self.kind in ("basic", "forced"):
            smag = 1 if self.kind == "basic" else 0
            _fn("lbo_collide_basic", self.real)(*self._g(), P(self.CT), ctypes.c_int32(self.skip),
                                                ctypes.c_int32(smag), _d(self.omega0), _d(self.cs),
                                                P(self.F), P(self.f), P(self.feq), P(self.rho), P(self.vel),
                                                P(self.omega), P(self.rate), P(self.fpc))
        elif self.kind in ("psc", "eqibm"):
            smag = 1 if self.kind == "eqibm" else 0
            _fn("lbo_collide_weighted", self.real)(*self._g(), P(self.CT), ctypes.c_int32(self.skip),
                                                   ctypes.c_int32(smag), _d(self.omega0), _d(self.cs),
                                                   P(self.F), P(self.f), P(self.feq), P(self.rho), P(self.vel),
                                                   P(self.omega), P(self.rate), P(self.volfrac), P(self.weight),
                                                   P(self.velsolid),
                                                   P(self.hydroforce) if self.kind == "psc" else None, P(self.fpc))
        elif self.kind == "hybrid":
            _fn("lbo_collide_hybrid", self.real)(*self._g(), P(self.CT), _d(self.omega0), _d(self.cs), P(self.f),
                                                 P(self.feq), P(self.rho), P(self.vel), P(self.omega), P(self.rate),
                                                 P(self.volfrac), P(self.velsolid), P(self.weight),
                                                 P(self.fine_volfrac), P(self.fine_weight), P(self.fine_velsolid),
                                                 P(self.hydroforce), P(self.fpc))
        elif self.kind == "vans":
            _fn("lbo_vans_collide", self.real)(*self._g(), P(self.CT), R(self.omega0), P(self.f), P(self.feq),
                                               P(self.rho), P(self.vel), P(self.omega), P(self.volfrac),
                                               P(self.hydroforce), P(self.fpc))
        else:
            raise ValueError(self.kind)

    def stream(self):
        vans = 1 if self.kind == "vans" else 0
        _fn("lbo_stream", self.real)(*self._g(), P(self.CT), ctypes.c_int32(self.skip), P(self.fpc), P(self.f),
                                     P(self.rho), P(self.vel), ctypes.c_int32(vans),
                                     P(self.volfrac) if vans else None, P(self.hydroforce) if vans else None)
        if (self.CT & WET_NODE).any():  # the boundary loop after streaming (lbm_solver3d.py:241-250)
            _fn("lbo_boundaries", self.real)(*self._g(), P(self.CT), P(self.f), P(self.rho), P(self.vel))

    def step(self, n=1):
        for _ in range(n):
            self.collide()
            self.stream()

    # conservation sums over collided cells, from the post-stream populations
    def sums(self):
        act = (self.CT & self.skip) == 0
        f = self.f[act].astype(np.float64)
        rho = f.sum(-1)
        j = f @ C
        return rho.sum(), j.sum(0), 0.5 * ((j * j).sum(-1) / rho).sum()


# ---------------------------------------------------------------------------------------
# coupling restatements
# ---------------------------------------------------------------------------------------
def psc_map(L, grains, dmin, dx, dt):
    """PSCLattice3D.grains2lattice (psclattice3d.py:84-182); grains (n, 10)."""
    g = np.ascontiguousarray(grains, dtype=np.float64)
    dm = np.ascontiguousarray(dmin, dtype=np.float64)
    _fn("lbo_psc_map", L.real)(*L._g(), P(L.CT), ctypes.c_int32(len(g)), P(g), P(dm), _d(dx), _d(dt),
                               P(L.omega), P(L.id), P(L.volfrac), P(L.weight), P(L.velsolid))


def psc_forces(L, grains, dmin, dx, dt, rho_phys):
    """PSCLattice3D.lattice2grains (psclattice3d.py:316-358) -> (n, 6) [F, T]."""
    g = np.ascontiguousarray(grains, dtype=np.float64)
    dm = np.ascontiguousarray(dmin, dtype=np.float64)
    out = np.zeros((len(g), 6))
    _fn("lbo_psc_forces", L.real)(*L._g(), P(L.CT), ctypes.c_int32(len(g)), P(g), P(dm), _d(dx), _d(dt),
                                  _d(rho_phys), P(L.id), P(L.volfrac), P(L.hydroforce), P(out))
    return out


def threedelta(r):
    return lib().lbo_threedelta_f64(float(r))


def kernel_mirror(nx, ny, nz, xc, yc, zc, i, j, k):
    return lib().lbo_kernel_mirror_f64(nx, ny, nz, xc, yc, zc, i, j, k)


def tenneti_cd(Re, svf, lim=1e-9):
    return lib().lbo_tenneti_cd_f64(float(Re), float(svf), float(lim))


def ibm_spread(L, grains, dmin, dx, dt):
    """EqIMBlattice3D.grains2lattice (eqlattice.py:104-194); grains (n, 7) [x, v, R]."""
    g = np.ascontiguousarray(grains, dtype=np.float64)
    dm = np.ascontiguousarray(dmin, dtype=np.float64)
    velsum = np.zeros_like(L.velsolid)
    wsum = np.zeros_like(L.volfrac)
    _fn("lbo_ibm_spread", L.real)(*L._g(), P(L.CT), ctypes.c_int32(len(g)), P(g), P(dm), _d(dx), _d(dt),
                                  P(L.volfrac), P(L.velsolid), P(velsum), P(wsum))
    return velsum, wsum


def ibm_weight(L, dx, dt, rho_phys):
    """EqIMBlattice3D.compute_weight (eqlattice.py:200-238)."""
    _fn("lbo_ibm_weight", L.real)(*L._g(), P(L.CT), _d(dx), _d(dt), _d(rho_phys), _d(L.omega0), P(L.volfrac),
                                  P(L.velsolid), P(L.vel), P(L.weight))


def ibm_drag(L, grains, dmin, dx, dt, rho_phys):
    """EqIMBlattice3D.lattice2grains (eqlattice.py:309-386) -> (n, 3) force_fluid."""
    g = np.ascontiguousarray(grains, dtype=np.float64)
    dm = np.ascontiguousarray(dmin, dtype=np.float64)
    out = np.zeros((len(g), 3))
    _fn("lbo_ibm_drag", L.real)(*L._g(), P(L.CT), ctypes.c_int32(len(g)), P(g), P(dm), _d(dx), _d(dt),
                                _d(rho_phys), _d(L.omega0), P(L.vel), P(L.volfrac), P(out))
    return out


def vans_spread(L, grains, dmin, dx, dt, rho_phys, with_force=True):
    """Unresolvedlattice3D.force2lattice / grains2lattice; grains (n, 7) [x, R, F_fluid]."""
    g = np.ascontiguousarray(grains, dtype=np.float64)
    dm = np.ascontiguousarray(dmin, dtype=np.float64)
    _fn("lbo_vans_spread", L.real)(*L._g(), P(L.CT), ctypes.c_int32(len(g)), P(g), P(dm), _d(dx), _d(dt),
                                   _d(rho_phys), ctypes.c_int32(1 if with_force else 0), P(L.volfrac),
                                   P(L.hydroforce))


def vans_drag(L, grains, dmin, dx, dt, rho_phys):
    """Unresolvedlattice3D.compute_hydrodynamic_force; grains (n, 7) [x, v, R] -> (n, 3)."""
    g = np.ascontiguousarray(grains, dtype=np.float64)
    dm = np.ascontiguousarray(dmin, dtype=np.float64)
    out = np.zeros((len(g), 3))
    _fn("lbo_vans_drag", L.real)(*L._g(), P(L.CT), ctypes.c_int32(len(g)), P(g), P(dm), _d(dx), _d(dt),
                                 _d(rho_phys), _d(L.omega0), P(L.volfrac), P(L.vel), P(out))
    return out


def hybrid_map_coarse(L, grains, dmin, dx, dt):
    """HybridLattice3D.map_coarse_grains (hybridlattice.py:173-257); grains (n, 10) [x, v, w, R]."""
    g = np.ascontiguousarray(grains, dtype=np.float64)
    dm = np.ascontiguousarray(dmin, dtype=np.float64)
    _fn("lbo_hybrid_map_coarse", L.real)(*L._g(), P(L.CT), ctypes.c_int32(len(g)), P(g), P(dm), _d(dx), _d(dt),
                                         P(L.id), P(L.volfrac), P(L.velsolid))
    L.weight[...] = 0  # coarse_weight.fill(0.0), :193


def hybrid_map_fine(L, grains, dmin, dx, dt):
    """HybridLattice3D.map_fine_grains (hybridlattice.py:260-376); grains (n, 7) [x, v, R].
    Returns (velsum, wsum, warnings)."""
    g = np.ascontiguousarray(grains, dtype=np.float64)
    dm = np.ascontiguousarray(dmin, dtype=np.float64)
    velsum = np.zeros_like(L.fine_velsolid)
    wsum = np.zeros_like(L.fine_volfrac)
    nwarn = np.zeros(1, dtype=np.int32)
    _fn("lbo_hybrid_map_fine", L.real)(*L._g(), P(L.CT), ctypes.c_int32(len(g)), P(g), P(dm), _d(dx), _d(dt),
                                       P(L.volfrac), P(L.fine_volfrac), P(L.fine_velsolid), P(velsum), P(wsum),
                                       P(nwarn))
    return velsum, wsum, int(nwarn[0])


def hybrid_fine_weights(L, dx, dt, rho_phys):
    """HybridLattice3D.compute_fine_weights (hybridlattice.py:378-431)."""
    _fn("lbo_hybrid_fine_weights", L.real)(*L._g(), P(L.CT), _d(dx), _d(dt), _d(rho_phys), _d(L.omega0),
                                           P(L.volfrac), P(L.fine_volfrac), P(L.fine_velsolid), P(L.vel),
                                           P(L.fine_weight))


def hybrid_forces(L, grains, dmin, dx, dt, rho_phys):
    """HybridLattice3D.lattice2grains (hybridlattice.py:595-740) -> (n, 6) [F, T]: the PSC
    exchange of the coarse grains (_transfer_coarse_force), the Tenneti drag of the fine ones."""
    g = np.ascontiguousarray(grains, dtype=np.float64)
    out = psc_forces(L, g, dmin, dx, dt, rho_phys)
    g7 = np.ascontiguousarray(np.concatenate([g[:, 0:6], g[:, 9:10]], axis=1))
    dm = np.ascontiguousarray(dmin, dtype=np.float64)
    drag = np.zeros((len(g), 3))
    _fn("lbo_hybrid_fine_drag", L.real)(*L._g(), P(L.CT), ctypes.c_int32(len(g)), P(g7), P(dm), _d(dx), _d(dt),
                                        _d(rho_phys), _d(L.omega0), P(L.vel), P(L.fine_volfrac), P(L.volfrac),
                                        P(drag))
    fine = 2.0 * g[:, 9] <= dx
    out[fine, :3] = drag[fine]
    out[fine, 3:] = 0.0
    return out


class HybridOracle:
    """HybridLattice3D driven as experiment.py's main loop (hybridexperiment/experiment.py:
    initialize_complete, then per step prepare_step, collide, stream, lattice2grains).
    grains (n, 10) [x, v, omega, R], physical units; held fixed (no DEM)."""

    def __init__(self, L, grains, dmin, dx, dt, rho_phys):
        assert L.kind == "hybrid"
        self.L, self.g = L, np.ascontiguousarray(grains, dtype=np.float64)
        self.dmin, self.dx, self.dt, self.rho_phys = np.asarray(dmin, dtype=np.float64), dx, dt, rho_phys
        self.warnings = 0

    def prepare_step(self):
        L, g = self.L, self.g
        hybrid_map_coarse(L, g, self.dmin, self.dx, self.dt)
        _, _, w = hybrid_map_fine(L, np.concatenate([g[:, 0:6], g[:, 9:10]], axis=1), self.dmin, self.dx, self.dt)
        self.warnings += w
        hybrid_fine_weights(L, self.dx, self.dt, self.rho_phys)

    def initialize_complete(self):
        self.L.initialize()
        self.prepare_step()

    def step(self):
        self.prepare_step()
        self.L.collide()
        self.L.stream()
        return hybrid_forces(self.L, self.g, self.dmin, self.dx, self.dt, self.rho_phys)


def stress(L):
    """BasicLattice3D.compute_stress (lbm_solver3d.py:347-405) -> (7, nx, ny, nz): xx, yy, zz, xy,
    xz, yz, shear magnitude; zeros at skipped cells."""
    out = np.zeros((7,) + L.CT.shape, dtype=L.real)
    _fn("lbo_stress", L.real)(*L._g(), P(L.CT), P(L.f), P(L.feq), P(L.omega), P(out))
    return out


def plane_stress(L, st, axis, position):
    """compute_plane_average_stress (lbm_solver3d.py:407-456) -> (xy, xz, yz) means."""
    out = np.zeros(3)
    _fn("lbo_plane_stress", L.real)(*L._g(), P(L.CT), P(np.ascontiguousarray(st, dtype=L.real)),
                                    ctypes.c_int(axis), ctypes.c_int(position), P(out))
    return tuple(out)


def cavity(L, lid_u, lid_axis=2):
    """Lid-driven cavity cell types (lid_driven_cavity.py:31-47 with the lid on the +lid_axis face)."""
    CT = L.CT
    CT[...] = FLUID
    for ax in range(3):
        sl_lo = [slice(None)] * 3
        sl_hi = [slice(None)] * 3
        sl_lo[ax] = 0
        sl_hi[ax] = -1
        CT[tuple(sl_lo)] = OBSTACLE
        if ax != lid_axis:
            CT[tuple(sl_hi)] = OBSTACLE
    lid = [slice(1, -1)] * 3
    lid[lid_axis] = -1
    CT[tuple(lid)] = VEL_LADD
    L.vel[tuple(lid)] = np.asarray(lid_u, dtype=L.real)
    return CT


# ---------------------------------------------------------------------------------------
# DEM restatement (dem_oracle.c)
# ---------------------------------------------------------------------------------------
_i32, _f64 = ctypes.c_int32, ctypes.c_double


class DemParams(ctypes.Structure):
    _fields_ = [("n", _i32), ("K", _i32), ("model", _i32), ("dt", _f64), ("g", _f64 * 3), ("dmin", _f64 * 3),
                ("dmax", _f64 * 3), ("periodic", _i32 * 3), ("cell", _f64), ("T", _i32), ("probe27", _i32),
                ("mat", (_f64 * 9) * 2), ("wall_n", (_f64 * 3) * 6), ("wall_d", _f64 * 6), ("wall_type", _i32 * 6)]


_STATE_FIELDS = ["mass", "radius", "pos", "vel", "acc", "force", "quat", "omega", "omega_dot", "inertia", "moment",
                 "force_fluid", "moment_fluid", "freeze", "fixvel", "cf_j", "cf_active", "cfn", "cf_pos", "cf_force",
                 "cf_shear", "wcf_active", "wcf_pos", "wcf_force", "wcf_shear", "status"]


class DemState(ctypes.Structure):
    _fields_ = [(k, _dp) for k in _STATE_FIELDS]


def morton3d32(x, y, z):
    L = lib()
    L.lbo_morton3d32.restype = ctypes.c_uint32
    return L.lbo_morton3d32(_i32(x), _i32(y), _i32(z))


def hash_table_size(dmin, dmax, cell):
    L = lib()
    L.lbo_hash_table_size.restype = ctypes.c_int64
    a = np.ascontiguousarray(dmin, dtype=np.float64)
    b = np.ascontiguousarray(dmax, dtype=np.float64)
    return int(L.lbo_hash_table_size(P(a), P(b), _f64(cell)))


def broadphase(pos, dmin, cell, T, probe27=False):
    """Candidate pairs (i < j) of the reference's broad phase, sorted (m, 2) int32."""
    L = lib()
    L.lbo_broadphase.restype = ctypes.c_int64
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    dm = np.ascontiguousarray(dmin, dtype=np.float64)
    n = len(pos)
    cap = max(64 * n, 16)
    while True:
        pairs = np.zeros((cap, 2), dtype=np.int32)
        m = L.lbo_broadphase(_i32(n), P(pos), P(dm), _f64(cell), _i32(T), _i32(1 if probe27 else 0), P(pairs),
                             ctypes.c_int64(cap))
        if m <= cap:
            break
        cap = m
    pr = pairs[:m]
    return pr[np.lexsort((pr[:, 1], pr[:, 0]))]


def material_rows(cfg, density):
    """mf[0] (particle) and mf[1] (wall) as DEMSolver.init_particle_fields sets them
    (demsolver.py:222-290); row layout density, E, nu, mu, e, kn, kt, dpn, dps."""
    cm = cfg.contact_model
    pp, wp = cfg.particle_props, cfg.wall_props
    m = np.zeros((2, 9))
    m[0, :3] = density, pp.elastic_modulus, pp.poisson_ratio
    m[1, :3] = wp.density, wp.elastic_modulus, wp.poisson_ratio
    if cm.get_model_name() == "linear":
        m[0, 5], m[0, 6], m[0, 7] = cm.stiffness_normal, cm.stiffness_tangential, cm.damping_normal
        m[0, 8] = 0.0  # dp_sratio is never assigned (demsolver.py:250 writes dp_tratio)
        m[0, 3], m[1, 3] = cm.pp_friction, cm.pw_friction
        m[0, 4] = getattr(cm, "pp_restitution", 0.0)
        m[1, 4] = getattr(cm, "pw_restitution", 0.0)
    else:
        m[0, 4], m[1, 4] = cm.pp_restitution, cm.pw_restitution
        m[0, 3], m[1, 3] = cm.pp_friction, cm.pw_friction
    return m


class OracleDEM:
    """DEMSolver state in reference semantics, stepped by the serial C restatement."""

    def __init__(self, cfg, radius, mass, position, velocity=None, K=None, cell=None, probe27=False,
                 dmin=None, dmax=None):
        n = len(radius)
        K = int(K if K is not None else cfg.particle_props.max_coordinate_number)
        self.n, self.K = n, K
        d = cfg.domain
        dmin = np.array([d.xmin, d.ymin, d.zmin]) if dmin is None else np.asarray(dmin, dtype=np.float64)
        dmax = np.array([d.xmax, d.ymax, d.zmax]) if dmax is None else np.asarray(dmax, dtype=np.float64)
        radius = np.asarray(radius, dtype=np.float64)
        mass = np.asarray(mass, dtype=np.float64)
        cell = 4.0 * radius.max() if cell is None else cell
        p = DemParams()
        p.n, p.K = n, K
        p.model = 1 if cfg.contact_model.get_model_name() == "linear" else 0
        p.dt = cfg.dt
        for a in range(3):
            p.g[a] = float(cfg.gravity[a])
            p.dmin[a] = float(dmin[a])
            p.dmax[a] = [d.xmax, d.ymax, d.zmax][a]
            p.periodic[a] = 1 if cfg.boundaryType[2 * a] == cfg.BOUNDARY_TYPE_PERIODIC else 0
        p.cell = cell
        p.T = hash_table_size(dmin, dmax, cell)
        p.probe27 = 1 if probe27 else 0
        density = 3 * mass[-1] / (4 * np.pi * radius[-1] ** 3)
        mat = material_rows(cfg, density)
        for k in range(2):
            for m in range(9):
                p.mat[k][m] = mat[k, m]
        for j in range(6):
            for a in range(3):
                p.wall_n[j][a] = float(cfg.wall_normals[j][a])
            p.wall_d[j] = float(cfg.wall_distances[j])
            p.wall_type[j] = int(cfg.boundaryType[j])
        self.params = p
        z = lambda *s: np.zeros(s)  # noqa: E731
        zi = lambda *s: np.zeros(s, dtype=np.int32)  # noqa: E731
        self.a = dict(mass=mass.copy(), radius=radius.copy(), pos=np.array(position, dtype=np.float64).reshape(n, 3),
                      vel=(z(n, 3) if velocity is None else np.array(velocity, dtype=np.float64).reshape(n, 3)),
                      acc=z(n, 3), force=z(n, 3), quat=np.tile([1.0, 0, 0, 0], (n, 1)), omega=z(n, 3),
                      omega_dot=z(n, 3), inertia=np.zeros((n, 3, 3)), moment=z(n, 3), force_fluid=z(n, 3),
                      moment_fluid=z(n, 3), freeze=zi(n), fixvel=zi(n), cf_j=zi(n * K), cf_active=zi(n * K),
                      cfn=zi(n), cf_pos=z(n * K, 3), cf_force=z(n * K, 3), cf_shear=z(n * K, 3),
                      wcf_active=zi(6 * n), wcf_pos=z(6 * n, 3), wcf_force=z(6 * n, 3), wcf_shear=z(6 * n, 3),
                      status=zi(8))
        for i in range(3):
            self.a["inertia"][:, i, i] = 2.0 / 5.0 * mass * radius * radius
        self._pairs = np.zeros((max(256 * n, 16), 2), dtype=np.int32)

    def step(self, n=1):
        L = lib()
        L.lbo_dem_step.restype = ctypes.c_int64
        s = DemState()
        for k in _STATE_FIELDS:
            setattr(s, k, P(self.a[k]))
        for _ in range(n):
            while True:
                r = L.lbo_dem_step(ctypes.byref(self.params), ctypes.byref(s), P(self._pairs),
                                   ctypes.c_int64(len(self._pairs)))
                if r >= 0:
                    break
                self._pairs = np.zeros((-r + 1024, 2), dtype=np.int32)  # more candidates than room

    def __getitem__(self, k):
        return self.a[k]


# ---------------------------------------------------------------------------------------
# single-sphere VANS settling, example/unlbdem/single_particle_settle/one_particle.py
# ---------------------------------------------------------------------------------------
# the four cases of one_particle.py:44-73: Re -> (mu [Pa s], rho [kg/m^3]); sphere of 1120 kg/m^3
SETTLE_CASES = {"31.9": (58e-3, 960.0), "11.6": (113e-3, 962.0), "4.1": (212e-3, 965.0), "1.5": (373e-3, 970.0)}
SETTLE_DX, SETTLE_TAU, SETTLE_R, SETTLE_DENS = 0.025, 0.5001, 0.0075, 1120.0
SETTLE_MASS = 0.00197920337176   # one_particle.p4p
SETTLE_X0 = (0.05, 0.1275, 0.05)


def settle_setup(case):
    """(mu, rho, dt, (nx, ny, nz), gravity) of one_particle.py:76-110 for a case."""
    mu, rho = SETTLE_CASES[case]
    nu_lu = (SETTLE_TAU - 0.5) / 3.0
    dt = SETTLE_DX ** 2 / ((mu / rho) / nu_lu)
    shape = (int(0.1 / SETTLE_DX) + 1, int(0.16 / SETTLE_DX) + 1, int(0.1 / SETTLE_DX) + 1)
    g = (0.0, -9.81 * (SETTLE_DENS - rho) / SETTLE_DENS, 0.0)
    return mu, rho, dt, shape, g


def settle_container(CT):
    """setContainer (one_particle.py:34-40): every face cell is an obstacle."""
    CT[0], CT[-1], CT[:, 0], CT[:, -1], CT[:, :, 0], CT[:, :, -1] = (OBSTACLE,) * 6


class VansSettleOracle:
    """one_particle.py's main loop (:236-246) on the restated VANS lattice and DEM: per lattice
    step grains2lattice, compute_hydrodynamic_force, 10 DEM substeps, force2lattice, collide,
    stream.  after_grains2lattice(L, pos) is called after the volume-fraction mapping, before the
    drag is evaluated."""

    def __init__(self, case, dem_cfg, after_grains2lattice=None):
        self.mu, self.rho, self.dt, shape, _ = settle_setup(case)
        self.L = OracleLattice(*shape, 1.0 / SETTLE_TAU, kind="vans")
        settle_container(self.L.CT)
        self.dem = OracleDEM(dem_cfg, [SETTLE_R], [SETTLE_MASS], [list(SETTLE_X0)])
        self.hook = after_grains2lattice
        self.L.initialize()
        self.t = 0

    def step(self):
        L, a, dx, R = self.L, self.dem.a, SETTLE_DX, SETTLE_R
        vans_spread(L, np.column_stack([a["pos"], [[R]], np.zeros((1, 3))]), [0, 0, 0], dx, self.dt, self.rho,
                    with_force=False)
        if self.hook is not None:
            self.hook(L, a["pos"][0].copy())
        a["force_fluid"][...] = vans_drag(L, np.column_stack([a["pos"], a["vel"], [[R]]]), [0, 0, 0], dx, self.dt,
                                          self.rho)
        self.dem.step(10)
        vans_spread(L, np.column_stack([a["pos"], [[R]], a["force_fluid"]]), [0, 0, 0], dx, self.dt, self.rho,
                    with_force=True)
        L.collide()
        L.stream()
        self.t += 1
