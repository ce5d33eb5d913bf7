"""DEM driver (API of src/dem3d/demsolver/demsolver.py).

Particle, contact and wall state live in FP64 device tensors (structure of arrays).  One call
of `run_simulation()` is one DEM substep of the reference (clear forces, broad phase, contact
resolution with history, walls, body and fluid forces, integration, slot compaction), executed
by `lbdem_dem_run`.  `run_substeps(k)` runs k substeps in one call; for up to 2048 particles
that is a single kernel launch.
"""
import ctypes
import math
import os
import time

import numpy as np
import torch

from ... import _hip
from ..._field import Field
from ...runtime import config
from .. import _abi

# lbdem_dem_run's automatic choice takes the single-workgroup path up to this many particles
# (kSmallN in csrc/dem.hip)
SMALL_PATH_MAX = 2048
from ..bpcd import BPCD
from ..contactmanager import HertzMindlinContactModel, LinearContactModel
from ..demconfig import DEMSolverConfig, HertzContactConfig, LinearContactConfig

DoublePrecisionTolerance = 1e-12

# material row layout of lbdem_dem_params.mat
_MAT = ("density", "elasticModulus", "poissonRatio", "coefficientFriction", "coefficientRestitution",
        "stiffness_normal", "stiffness_tangent", "dp_nratio", "dp_sratio")


class _Record:
    """Element proxy: `field[i].member` reads / writes element i of each member field."""

    def __init__(self, owner, i):
        object.__setattr__(self, "_owner", owner)
        object.__setattr__(self, "_i", i)

    def __getattr__(self, name):
        f = self._owner._fields[name]
        return f[self._i]

    def __setattr__(self, name, value):
        self._owner._fields[name][self._i] = value


# particle members the DEM kernels read but never write
_DEM_READ_ONLY = frozenset(("ID", "groupID", "materialType", "mass", "radius", "inertia", "freeze", "fixvel",
                            "force_fluid", "moment_fluid"))


class StructField:
    """Taichi StructField look-alike: member fields as attributes, records by index."""

    def __init__(self, n, fields):
        self.shape = (n,)
        self._fields = fields

    def __getattr__(self, name):
        fields = self.__dict__.get("_fields", {})
        if name in fields:
            return fields[name]
        raise AttributeError(name)

    def __getitem__(self, i):
        return _Record(self, i)

    def sync(self):
        for f in self._fields.values():
            f.sync()

    def touch(self, skip=()):
        """The device copies were written by a kernel (fields named in skip were only read)."""
        for name, f in self._fields.items():
            if name in skip:
                continue
            f._host = None
            f.version += 1


class _MaterialTable:
    """mf[0] (particles) and mf[1] (walls) with the reference's member names."""

    def __init__(self):
        self.rows = [dict.fromkeys(_MAT, 0.0), dict.fromkeys(_MAT, 0.0)]
        self.shape = (2,)

    def __getitem__(self, k):
        table = self

        class Row:
            def __getattr__(self, name):
                return table.rows[k][name]

            def __setattr__(self, name, value):
                if name not in table.rows[k]:
                    raise AttributeError(f"Material has no member {name!r}")
                table.rows[k][name] = float(value)

        return Row()


class DEMSolver:
    def __init__(self, config_: DEMSolverConfig):
        self.config = config_
        self.bpcd = None
        self.mf = _MaterialTable()
        self.gf = None
        self.cf = None
        self.cfn = None
        self.wf = None
        self.wcf = None
        self.contact_model = None
        self.device = config.torch_device
        self.path = 0  # 0 automatic, 1 single-workgroup, 2 grid-wide

    # ---------------------------------------------------------------------------------
    # output (p4p / p4c), demsolver.py:45-128
    # ---------------------------------------------------------------------------------
    def save(self, time_: float):
        out_dir = "output"
        os.makedirs(out_dir, exist_ok=True)
        idx = int(round(time_ / self.config.dt))
        name = os.path.join(out_dir, f"output_T{idx:06d}")
        with open(name + ".p4p", "w", encoding="UTF-8") as p4p, open(name + ".p4c", "w", encoding="UTF-8") as p4c:
            self.save_single(p4p, p4c, time_)

    def save_single(self, p4pfile, p4cfile, t: float):
        n = self.n
        p4pfile.write("TIMESTEP  PARTICLES\n")
        p4pfile.write(f"{t} {n}\n")
        p4pfile.write("ID  GROUP  RAD  MASS  PX  PY  PZ  VX  VY  VZ\n")
        ids = self.gf.ID.to_numpy()
        data = np.column_stack([ids, self.gf.groupID.to_numpy(), self.gf.radius.to_numpy(), self.gf.mass.to_numpy(),
                                self.gf.position.to_numpy(), self.gf.velocity.to_numpy()])
        np.savetxt(p4pfile, data, fmt="%d %d %.6e %.6e %.6e %.6e %.6e %.6e %.6e %.6e")
        K = self.K
        active = self._cf_active.cpu().numpy().astype(bool)
        i_idx = np.repeat(np.arange(n), K)[active]
        j_idx = self._cf_j.cpu().numpy()[active]
        pos = self._cf_pos.view(-1, 3).cpu().numpy()[active]
        frc = self._cf_force.view(-1, 3).cpu().numpy()[active]
        p4cfile.write("TIMESTEP  CONTACTS\n")
        p4cfile.write(f"{t} {int(active.sum())}\n")
        p4cfile.write("P1  P2  CX  CY  CZ  FX  FY  FZ\n")
        if active.any():
            np.savetxt(p4cfile, np.column_stack([ids[i_idx], ids[j_idx], pos, frc]),
                       fmt="%d %d %.6e %.6e %.6e %.6e %.6e %.6e")

    # ---------------------------------------------------------------------------------
    # restart snapshot (no counterpart in the reference, whose p4p / p4c output keeps neither
    # the rotation state nor the contact history a run resumes from)
    # ---------------------------------------------------------------------------------
    _HISTORY = ("_cf_j", "_cf_active", "_cfn", "_cf_pos", "_cf_force", "_cf_shear", "_wcf_active", "_wcf_pos",
                "_wcf_force", "_wcf_shear")

    def state_dict(self):
        """Every particle member (position, velocity, acceleration, force, quaternion, omega,
        omega_dot, inertia, moment, freeze, fixvel, fluid loads, ...) and the particle-particle
        and particle-wall contact history (partner, active flag, contact point, force, shear
        displacement), as numpy arrays."""
        self.gf.sync()
        d = {f"gf.{k}": f.dev.detach().cpu().numpy() for k, f in self.gf._fields.items()}
        for k in self._HISTORY:
            d[k] = getattr(self, k).detach().cpu().numpy()
        d["n"], d["K"] = np.array(self.n), np.array(self.K)
        return d

    def load_state_dict(self, d):
        """Restore a state_dict() into a solver set up with the same particle count and
        max_coordinate_number.  The single-workgroup path rebuilds its neighbour lists."""
        if int(d["n"]) != self.n or int(d["K"]) != self.K:
            raise ValueError(f"snapshot of n={int(d['n'])}, K={int(d['K'])} does not fit a solver of "
                             f"n={self.n}, K={self.K}")
        for k, f in self.gf._fields.items():
            f.from_numpy(d[f"gf.{k}"])
        for k in self._HISTORY:
            getattr(self, k).copy_(torch.from_numpy(np.ascontiguousarray(d[k])))
        for sf in (self.cf, self.wcf):
            sf.touch()
        self.cfn.touch()
        self._verlet[:1].zero_()
        self._radius_version = None

    def save_state(self, file_name):
        """Write state_dict() to an .npz file (no pickled objects)."""
        np.savez(file_name, **self.state_dict())

    def load_state(self, file_name):
        with np.load(file_name, allow_pickle=False) as z:
            self.load_state_dict({k: z[k] for k in z.files})

    # ---------------------------------------------------------------------------------
    # set-up
    # ---------------------------------------------------------------------------------
    @staticmethod
    def read_p4p(file_name):
        """Parse a p4p packing: ID GROUP RAD MASS PX PY PZ VX VY VZ (1-based IDs)."""
        with open(file_name, encoding="UTF-8") as fp:
            fp.readline()
            n = int(fp.readline().split()[1])
            fp.readline()
            rows = [fp.readline().split() for _ in range(n)]
        rows = [r for r in rows if r]
        a = np.array(rows, dtype=np.float64).reshape(-1, 10)
        order = np.argsort(a[:, 0].astype(np.int64), kind="stable")
        return a[order]

    def init_particle_fields(self, file_name: str, domain_min=None, domain_max=None):
        """Load a p4p packing (demsolver.py:130-309)."""
        a = self.read_p4p(file_name)
        self.init_particle_arrays(radius=a[:, 2], mass=a[:, 3], position=a[:, 4:7], velocity=a[:, 7:10],
                                  ids=a[:, 0].astype(np.int64), groups=a[:, 1].astype(np.int64),
                                  domain_min=domain_min, domain_max=domain_max)

    def init_particle_arrays(self, radius, mass, position, velocity=None, ids=None, groups=None, domain_min=None,
                             domain_max=None, cell_size=None, probe27=False):
        """Set up particles from arrays (the same state a p4p file gives)."""
        _hip.require_gpu()
        radius = np.asarray(radius, dtype=np.float64).reshape(-1)
        n = radius.shape[0]
        mass = np.asarray(mass, dtype=np.float64).reshape(n)
        position = np.asarray(position, dtype=np.float64).reshape(n, 3)
        velocity = np.zeros((n, 3)) if velocity is None else np.asarray(velocity, dtype=np.float64).reshape(n, 3)
        ids = np.arange(1, n + 1) if ids is None else np.asarray(ids).reshape(n)
        groups = np.zeros(n, dtype=np.int64) if groups is None else np.asarray(groups).reshape(n)
        self.n = n
        K = int(self.config.particle_props.max_coordinate_number)
        if K < 0:
            raise ValueError("set_particle_properties(max_coordinate_number=...) must not be negative")
        # K = 0 (the reference's default, types.py:42) is accepted as there: no contact slots,
        # so a particle-particle contact is counted as a slot overflow (dem.check() raises)
        self.K = K
        dev = self.device
        # density of the last particle is taken as the material density (demsolver.py:176-178, 228)
        material_density = 3 * mass[-1] / (4 * math.pi * radius[-1] ** 3) if n else 0.0
        inertia = np.zeros((n, 3, 3))
        inertia[:, [0, 1, 2], [0, 1, 2]] = (2.0 / 5.0 * mass * radius * radius)[:, None]

        def F(shape, dtype, vec=None, init=None):
            f = Field(shape, dtype, dev, n=vec)
            if init is not None:
                f.from_numpy(np.asarray(init).reshape(tuple(f.dev.shape)))
            return f

        f64, i32 = torch.float64, torch.int32
        quat = np.zeros((n, 4))
        quat[:, 0] = 1.0
        fields = {
            "ID": F((n,), torch.int64, init=ids), "groupID": F((n,), torch.int64, init=groups),
            "materialType": F((n,), i32), "mass": F((n,), f64, init=mass), "radius": F((n,), f64, init=radius),
            "position": F((n,), f64, 3, position), "velocity": F((n,), f64, 3, velocity),
            "acceleration": F((n,), f64, 3), "force": F((n,), f64, 3), "quaternion": F((n,), f64, 4, quat),
            "omega": F((n,), f64, 3), "omega_dot": F((n,), f64, 3),
            "inertia": Field((n,), f64, dev, n=9), "moment": F((n,), f64, 3),
            "freeze": F((n,), i32), "fixvel": F((n,), i32),
            "force_fluid": F((n,), f64, 3), "moment_fluid": F((n,), f64, 3),
        }
        fields["inertia"].from_numpy(inertia.reshape(n, 9))
        self.gf = StructField(n, fields)
        # contact history slots
        self._cf_j = torch.zeros(n * K, dtype=i32, device=dev)
        self._cf_active = torch.zeros(n * K, dtype=i32, device=dev)
        self._cfn = torch.zeros(n, dtype=i32, device=dev)
        self._cf_pos = torch.zeros(3 * n * K, dtype=f64, device=dev)
        self._cf_force = torch.zeros(3 * n * K, dtype=f64, device=dev)
        self._cf_shear = torch.zeros(3 * n * K, dtype=f64, device=dev)
        self._wcf_active = torch.zeros(6 * n, dtype=i32, device=dev)
        self._wcf_pos = torch.zeros(18 * n, dtype=f64, device=dev)
        self._wcf_force = torch.zeros(18 * n, dtype=f64, device=dev)
        self._wcf_shear = torch.zeros(18 * n, dtype=f64, device=dev)
        self._status = torch.zeros(8, dtype=i32, device=dev)
        self.cfn = Field((n,), i32, dev, tensor=self._cfn)
        self.cf = StructField(n * K, {
            "j": Field((n * K,), i32, dev, tensor=self._cf_j),
            "isActive": Field((n * K,), i32, dev, tensor=self._cf_active),
            "position": Field((n * K,), f64, dev, n=3, tensor=self._cf_pos.view(n * K, 3)),
            "force_a": Field((n * K,), f64, dev, n=3, tensor=self._cf_force.view(n * K, 3)),
            "shear_displacement": Field((n * K,), f64, dev, n=3, tensor=self._cf_shear.view(n * K, 3)),
        })
        self.wcf = StructField(6 * n, {
            "isActive": Field((6 * n,), i32, dev, tensor=self._wcf_active),
            "position": Field((6 * n,), f64, dev, n=3, tensor=self._wcf_pos.view(6 * n, 3)),
            "force_a": Field((6 * n,), f64, dev, n=3, tensor=self._wcf_force.view(6 * n, 3)),
            "shear_displacement": Field((6 * n,), f64, dev, n=3, tensor=self._wcf_shear.view(6 * n, 3)),
        })
        cfg = self.config
        self.wf = [dict(normal=np.asarray(cfg.wall_normals[j], dtype=np.float64), distance=float(cfg.wall_distances[j]),
                        materialType=1, boundaryType=int(cfg.boundaryType[j])) for j in range(6)]
        # materials (demsolver.py:222-290)
        cm = cfg.contact_model
        self.mf[0].density = material_density
        self.mf[0].elasticModulus = cfg.particle_props.elastic_modulus
        self.mf[0].poissonRatio = cfg.particle_props.poisson_ratio
        self.mf[1].density = cfg.wall_props.density
        self.mf[1].elasticModulus = cfg.wall_props.elastic_modulus
        self.mf[1].poissonRatio = cfg.wall_props.poisson_ratio
        if isinstance(cm, LinearContactConfig):
            self.mf[0].stiffness_normal = cm.stiffness_normal
            self.mf[0].stiffness_tangent = cm.stiffness_tangential
            self.mf[0].dp_nratio = cm.damping_normal
            # The reference assigns damping_tangential to a member `dp_tratio` that Material
            # does not have (demsolver.py:250, material.py:30) and the law reads dp_sratio,
            # which therefore stays 0: the tangential dashpot is inactive.  Kept as is.
            self.mf[0].coefficientFriction = cm.pp_friction
            self.mf[1].coefficientFriction = cm.pw_friction
            self.mf[0].coefficientRestitution = getattr(cm, "pp_restitution", 0.0)
            self.mf[1].coefficientRestitution = getattr(cm, "pw_restitution", 0.0)
        elif isinstance(cm, HertzContactConfig):
            self.mf[0].coefficientRestitution = cm.pp_restitution
            self.mf[1].coefficientRestitution = cm.pw_restitution
            self.mf[0].coefficientFriction = cm.pp_friction
            self.mf[1].coefficientFriction = cm.pw_friction
        else:
            raise ValueError(f"Unknown contact model type: {type(cm).__name__}. "
                             "Expected LinearContactConfig or HertzContactConfig.")
        max_radius = float(radius.max()) if n else 0.0
        dmin = np.asarray(domain_min if domain_min is not None else cfg.domain_min, dtype=np.float64).reshape(3)
        dmax = np.asarray(domain_max if domain_max is not None else cfg.domain_max, dtype=np.float64).reshape(3)
        self.bpcd = BPCD.create(n, max_radius, dmin, dmax, cell_size=cell_size, probe27=probe27)
        self._ws = None
        self._pair_cap = max(256 * n, 4096)
        # neighbour lists the single-workgroup kernel keeps between calls (first element: valid)
        self._verlet = torch.zeros(2 + 17 * n, dtype=torch.int32, device=self.device)
        self._verlet_ref = torch.zeros(1 + 3 * n, dtype=torch.float64, device=self.device)
        self._radius_version = None
        self._cached = None
        self._order_valid = False

    def set_contact_model(self, model_type: str = "hertz"):
        """Select the contact law; must match the config's contact model (demsolver.py:311-347)."""
        self._cached = None
        cm = self.config.contact_model
        if model_type.lower() == "hertz":
            if not isinstance(cm, HertzContactConfig):
                raise ValueError(f"Requested 'hertz' model but config uses {type(cm).__name__}. "
                                 "Please use HertzContactConfig in your configuration.")
            self.contact_model = HertzMindlinContactModel(self.mf)
        elif model_type.lower() == "linear":
            if not isinstance(cm, LinearContactConfig):
                raise ValueError(f"Requested 'linear' model but config uses {type(cm).__name__}. "
                                 "Please use LinearContactConfig in your configuration.")
            self.contact_model = LinearContactModel(self.mf)
        else:
            raise ValueError(f"Unknown contact model type: '{model_type}'. Valid options are 'hertz' or 'linear'.")

    # ---------------------------------------------------------------------------------
    # device call
    # ---------------------------------------------------------------------------------
    def _params(self):
        cfg = self.config
        P = self.bpcd.params(self.n)
        P.K = self.K
        if self.contact_model is None:
            self.set_contact_model(cfg.contact_model.get_model_name())
        P.model = self.contact_model.model_id
        per = cfg.periodic_axes()
        for a in range(3):
            P.periodic[a] = 1 if per[a] else 0
            P.g[a] = float(cfg.gravity[a])
        P.dt = float(cfg.dt)
        d = cfg.domain
        for a, (lo, hi) in enumerate(((d.xmin, d.xmax), (d.ymin, d.ymax), (d.zmin, d.zmax))):
            P.dmax[a] = hi
        for k in range(2):
            for m, name in enumerate(_MAT):
                P.mat[k][m] = float(self.mf.rows[k][name])
        for j in range(6):
            for a in range(3):
                P.wall_n[j][a] = float(self.wf[j]["normal"][a])
            P.wall_d[j] = float(self.wf[j]["distance"])
            P.wall_type[j] = int(self.wf[j]["boundaryType"])
        self._fine_grid(P)
        return P

    # at most this many particles are tested against all others on the second-grid search
    MAX_LARGE = 64

    def _fine_grid(self, P):
        """Second-grid touching search of the grid-wide path (lbdem_dem_params.fine_cell): used when
        all but at most MAX_LARGE particles have radii at most 1/8 of the broad-phase cell (a
        strongly polydisperse packing, whose 4 r_max cells hold most particles in a few buckets).
        The pairs are the same; only the search changes."""
        P.fine_cell, P.large_r, P.nlarge = 0.0, 0.0, 0
        self._large = None
        general = self.path == 2 or (self.path == 0 and self.n > SMALL_PATH_MAX)
        if not general or P.probe27 or any(P.periodic[a] for a in range(3)):
            return
        r = self.gf.radius.to_numpy()
        small = r[r <= P.cell / 8.0]
        if small.size == 0:
            return
        large = np.nonzero(r > small.max())[0]
        if large.size > self.MAX_LARGE:
            return
        P.large_r = float(small.max())
        P.fine_cell = 2.0 * P.large_r * (1.0 + 1e-6)
        P.nlarge = int(large.size)
        self._large = torch.as_tensor(large.astype(np.int32), device=self.device)

    def _state(self):
        g = self.gf
        g.sync()
        S = _abi.DemState()
        for name, key in (("mass", "mass"), ("radius", "radius"), ("pos", "position"), ("vel", "velocity"),
                          ("acc", "acceleration"), ("force", "force"), ("quat", "quaternion"), ("omega", "omega"),
                          ("omega_dot", "omega_dot"), ("moment", "moment"), ("inertia", "inertia"),
                          ("force_fluid", "force_fluid"), ("moment_fluid", "moment_fluid"), ("freeze", "freeze"),
                          ("fixvel", "fixvel")):
            setattr(S, name, g._fields[key].dev.data_ptr())
        for name in ("cf_j", "cf_active", "cfn", "cf_pos", "cf_force", "cf_shear", "wcf_active", "wcf_pos",
                     "wcf_force", "wcf_shear", "status"):
            setattr(S, name, getattr(self, "_" + name).data_ptr())
        return S

    def invalidate(self):
        """Rebuild the kernel parameter block at the next step (after changing the config)."""
        self._cached = None

    def run_substeps(self, nsub: int):
        """nsub DEM substeps (each one run_simulation of the reference) in one call."""
        if nsub <= 0 or self.n == 0:
            return
        cached = getattr(self, "_cached", None)
        if cached is not None and self.gf.radius.version != self._radius_version:
            cached = self._cached = None  # radii written: the second-grid parameters follow them
        if cached is None:
            P = self._params()
            S = self._state()
            # the workspace tensors are held here too, so a later broad-phase call that resizes the
            # BPCD workspace cannot free memory this solver's kernels still point at
            t, w = self.bpcd.workspace(self.n, self._pair_cap, self.device)
            w.verlet = self._verlet.data_ptr()
            w.verlet_ref = self._verlet_ref.data_ptr()
            w.large = self._large.data_ptr() if self._large is not None else None
            cached = self._cached = (P, S, w, ctypes.byref(P), ctypes.byref(S), ctypes.byref(w), t)
            self._verlet[:1].zero_()
        else:
            self.gf.sync()
        if self.gf.radius.version != self._radius_version:  # radii written since the last call
            self._verlet[:1].zero_()
        _hip.check(_hip.lib().lbdem_dem_run(cached[3], cached[4], cached[5], int(nsub), int(self.path),
                                            _hip.stream()), "lbdem_dem_run")
        # the grid-wide path leaves the broad phase's bucket order in the workspace's pid
        self._order_valid = self.path == 2 or (self.path == 0 and self.n > SMALL_PATH_MAX)
        self.gf.touch(skip=_DEM_READ_ONLY)
        self._radius_version = self.gf.radius.version

    def spatial_order(self):
        """The particle indices in broad-phase bucket order of the last grid-wide substep (an
        int32 permutation of 0..n-1, grains close in space close in the order), or None when no
        substep has run on the grid-wide path since the arrays were set up.  The unresolved
        couplings visit grains in this order for locality."""
        cached = getattr(self, "_cached", None)
        if not getattr(self, "_order_valid", False) or cached is None:
            return None
        return cached[6]["pid"]

    def run_simulation(self):
        """One DEM step (demsolver.py:848-862)."""
        self.run_substeps(1)

    def status(self):
        """Device anomaly counters: slot overflow, dt above critical, overlaps > 5%, pair-list overflow."""
        s = self._status.cpu().numpy()
        return dict(slot_overflow=int(s[0]), dt_warnings=int(s[1]), pp_overlap_warnings=int(s[2]),
                    pw_overlap_warnings=int(s[3]), pair_overflow=int(s[4]))

    def reset_status(self):
        """Clear the device anomaly counters (stream-ordered)."""
        self._status.zero_()

    def check(self):
        """Raise if contacts were lost to a full history slot array or candidate list."""
        st = self.status()
        if st["slot_overflow"]:
            raise RuntimeError(f"ERROR: coordinate number > set_max_coordinate_number({self.K}) "
                               f"({st['slot_overflow']} contacts dropped)")
        if st["pair_overflow"]:
            raise RuntimeError(f"candidate pair list overflow: {st['pair_overflow']} pairs needed")
        return st

    # helpers used by the couplings
    def grain_tensors(self):
        g = self.gf
        g.sync()
        return dict(position=g.position.dev, velocity=g.velocity.dev, omega=g.omega.dev, radius=g.radius.dev,
                    force_fluid=g.force_fluid.dev, moment_fluid=g.moment_fluid.dev)
