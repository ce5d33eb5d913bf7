"""Unresolved couplings (EqIBM, VANS) on the GPU vs the oracle's restatement."""
import numpy as np
import pytest

from dem_cases import gpu_solver, make_config
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX, DT, RHO = 1e-3, 5e-4, 1000.0


def _particles(n=70, N=20, seed=8, cluster=False):
    """Sub-grid spheres (R ~ 0.3 lu) within 1.5 lu of the walls (mirror images) and scattered
    ones.  cluster=True adds 25 larger spheres (R ~ 0.48 lu) packed round one point so that the
    spread solid fraction passes 1 and is clamped to 0.99.

    The cluster is used for one mapping only: the reference's Tenneti weight is not bounded by 1
    and a packed cluster drives the fluid unstable within ~20 steps (oracle and reference alike),
    and the VANS force is spread with the sign of the particle force, a positive feedback that
    the cluster also makes unstable; the trajectories use the scattered spheres."""
    rng = np.random.default_rng(seed)
    parts = [np.column_stack([rng.uniform(1.0, 2.4, 15), rng.uniform(2, N - 3, 15), rng.uniform(2, N - 3, 15)]),
             rng.uniform(2.0, N - 3.0, (n - 40, 3))]
    radii = [rng.uniform(0.25, 0.4, n - 25)]
    if cluster:
        parts.insert(0, rng.normal([6.0, 6.0, 6.0], 0.25, (25, 3)))
        radii.insert(0, rng.uniform(0.45, 0.5, 25))
    lat = np.concatenate(parts)
    pos = (lat - 0.5) * DX
    rad = np.concatenate(radii) * DX
    vel = rng.uniform(-0.02, 0.02, (len(pos), 3)) * DX / DT
    mass = 2500.0 * 4 / 3 * np.pi * rad ** 3
    return pos, vel, rad, mass


def _walls(CT):
    CT[0], CT[-1] = O.OBSTACLE, O.OBSTACLE
    CT[:, 0], CT[:, -1] = O.OBSTACLE, O.OBSTACLE
    CT[:, :, 0], CT[:, :, -1] = O.OBSTACLE, O.OBSTACLE


def _setup(cls, kind, N=20, tau=0.6, cs=0.17, cluster=False, particles=None, path=0):
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    pos, vel, rad, mass = _particles(N=N, cluster=cluster) if particles is None else particles
    box = (0.0, (N - 1) * DX, 0.0, (N - 1) * DX, 0.0, (N - 1) * DX)
    cfg = make_config(box=box, dt=DT / 10, g=(0, 0, 0))
    dem = gpu_solver(cfg, rad, mass, pos, vel, path=path)
    BasicLattice3D.SmagorinskyConstant = cs
    L = cls(N, N, N, 1.0 / tau, DX, DT, RHO, dem)
    orc = O.OracleLattice(N, N, N, 1.0 / tau, kind=kind, cs=cs)
    _walls(orc.CT)
    rng = np.random.default_rng(4)
    inner = (slice(1, -1),) * 3
    orc.vel[inner] = rng.uniform(-0.02, 0.02, orc.vel[inner].shape).astype(np.float32)
    L.CT.from_numpy(orc.CT)
    L.vel.from_numpy(orc.vel.astype(np.float32))
    return L, orc, dem, pos, vel, rad


def test_eqibm_matches_oracle():
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    from src.unlbdem.eqlattice import EqIMBlattice3D
    old = BasicLattice3D.SmagorinskyConstant
    try:
        L, orc, dem, pos, vel, rad = _setup(EqIMBlattice3D, "eqibm")
        g = np.column_stack([pos, vel, rad])
        L.initialize_complete()
        orc.initialize()
        O.ibm_spread(orc, g, [0, 0, 0], DX, DT)
        O.ibm_weight(orc, DX, DT, RHO)
        vf = L.volfrac.to_numpy()
        assert np.abs(vf - orc.volfrac).max() < 2e-6
        cov = orc.volfrac > 0
        assert np.abs(L.velsolid.to_numpy()[cov] - orc.velsolid[cov]).max() < 1e-5 * np.abs(orc.velsolid).max()
        assert np.abs(L.weight.to_numpy()[cov] - orc.weight[cov]).max() < 2e-4 * np.abs(orc.weight).max()
        F = None
        for _ in range(30):
            orc.collide()
            orc.stream()
            O.ibm_spread(orc, g, [0, 0, 0], DX, DT)
            O.ibm_weight(orc, DX, DT, RHO)
            F = O.ibm_drag(orc, g, [0, 0, 0], DX, DT, RHO)
            L.collide()
            L.stream()
            L.update_coupling()
        fluid = orc.CT == O.FLUID
        assert np.abs(L.vel.to_numpy()[fluid] - orc.vel[fluid]).max() < 2e-4 * np.abs(orc.vel[fluid]).max()
        Fg = dem.gf.force_fluid.to_numpy()
        assert np.abs(Fg - F).max() < 5e-4 * np.abs(F).max()
    finally:
        BasicLattice3D.SmagorinskyConstant = old


@pytest.mark.parametrize("layout", ["dense", "scattered"])
def test_eqibm_in_bucket_order_matches_oracle(layout):
    """After a grid-wide DEM substep the EqIBM spreading and drag visit the grains in the DEM's
    bucket order (DEMSolver.spatial_order).  'dense' puts 600 grains into 8^3 lattice cells, so
    most workgroups' unions of support boxes fit their LDS accumulator; 'scattered' spreads them
    over the lattice, so workgroups add to the fields directly.  Mapping, weights and drag at the
    DEM's positions, then the fluid velocity and the drag after 10 coupled steps, against the
    oracle."""
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    from src.unlbdem.eqlattice import EqIMBlattice3D
    N, n = 20, 600
    rng = np.random.default_rng(12)
    lo, hi = (5.0, 13.0) if layout == "dense" else (1.0, N - 2.0)
    lat, rl = [], []
    while len(lat) < n:  # no overlaps: the DEM substep must not turn contacts into large velocities
        x, r = rng.uniform(lo, hi, 3), rng.uniform(0.2, 0.35)
        if lat and (np.linalg.norm(np.array(lat) - x, axis=1) < np.array(rl) + r + 0.02).any():
            continue
        lat.append(x)
        rl.append(r)
    rad = np.array(rl) * DX
    parts = ((np.array(lat) - 0.5) * DX, rng.uniform(-0.02, 0.02, (n, 3)) * DX / DT, rad,
             2500.0 * 4 / 3 * np.pi * rad ** 3)
    old = BasicLattice3D.SmagorinskyConstant
    try:
        L, orc, dem, _, _, _ = _setup(EqIMBlattice3D, "eqibm", N=N, particles=parts, path=2)
        assert dem.spatial_order() is None
        dem.run_substeps(1)
        order = dem.spatial_order()
        assert order is not None and np.array_equal(np.sort(order.cpu().numpy()), np.arange(n))
        g = np.column_stack([dem.gf.position.to_numpy(), dem.gf.velocity.to_numpy(), rad])
        L.initialize_complete()
        orc.initialize()
        O.ibm_spread(orc, g, [0, 0, 0], DX, DT)
        O.ibm_weight(orc, DX, DT, RHO)
        assert np.abs(L.volfrac.to_numpy() - orc.volfrac).max() < 2e-6
        cov = orc.volfrac > 0
        assert np.abs(L.velsolid.to_numpy()[cov] - orc.velsolid[cov]).max() < 1e-5 * np.abs(orc.velsolid).max()
        assert np.abs(L.weight.to_numpy()[cov] - orc.weight[cov]).max() < 2e-4 * np.abs(orc.weight).max()
        F = O.ibm_drag(orc, g, [0, 0, 0], DX, DT, RHO)
        L.update_coupling()
        assert np.abs(dem.gf.force_fluid.to_numpy() - F).max() < 5e-4 * np.abs(F).max()
        for _ in range(10):
            orc.collide()
            orc.stream()
            O.ibm_spread(orc, g, [0, 0, 0], DX, DT)
            O.ibm_weight(orc, DX, DT, RHO)
            F = O.ibm_drag(orc, g, [0, 0, 0], DX, DT, RHO)
            L.collide()
            L.stream()
            L.update_coupling()
        fluid = orc.CT == O.FLUID
        assert np.abs(orc.vel[fluid]).max() < 0.05, "the oracle trajectory must stay stable"
        assert np.abs(L.vel.to_numpy()[fluid] - orc.vel[fluid]).max() < 2e-4 * np.abs(orc.vel[fluid]).max()
        assert np.abs(dem.gf.force_fluid.to_numpy() - F).max() < 5e-4 * np.abs(F).max()
    finally:
        BasicLattice3D.SmagorinskyConstant = old


def test_eqibm_clamped_cluster_mapping():
    """grains2lattice + compute_weight with a packed cluster: the clamp to 0.99, the solid
    velocity and the weights match the oracle, and the clamp counter equals the clamped cells."""
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    from src.unlbdem.eqlattice import EqIMBlattice3D
    old = BasicLattice3D.SmagorinskyConstant
    try:
        L, orc, dem, pos, vel, rad = _setup(EqIMBlattice3D, "eqibm", cluster=True)
        g = np.column_stack([pos, vel, rad])
        L.initialize_complete()
        orc.initialize()
        O.ibm_spread(orc, g, [0, 0, 0], DX, DT)
        O.ibm_weight(orc, DX, DT, RHO)
        clamped = int((orc.volfrac == 0.99).sum())
        assert clamped > 0, "the cluster must reach the solid-fraction clamp"
        assert L.clamp_count() == clamped
        vf = L.volfrac.to_numpy()
        assert np.abs(vf - orc.volfrac).max() < 2e-6
        cov = orc.volfrac > 0
        assert np.abs(L.velsolid.to_numpy()[cov] - orc.velsolid[cov]).max() < 1e-5 * np.abs(orc.velsolid).max()
        # per cell: the weight of a clamped cell is ~1e3 and sensitive to FP32 rounding of 0.99
        W, Wo = L.weight.to_numpy()[cov], orc.weight[cov]
        assert np.all(np.abs(W - Wo) <= 1e-3 * np.abs(Wo) + 1e-6)
        F = O.ibm_drag(orc, g, [0, 0, 0], DX, DT, RHO)
        L.update_coupling()
        Fg = dem.gf.force_fluid.to_numpy()
        assert np.abs(Fg - F).max() < 5e-4 * np.abs(F).max()
    finally:
        BasicLattice3D.SmagorinskyConstant = old


def test_vans_matches_oracle():
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    from src.unlbdem.vanslattice import Unresolvedlattice3D
    old = BasicLattice3D.SmagorinskyConstant
    try:
        L, orc, dem, pos, vel, rad = _setup(Unresolvedlattice3D, "vans")
        gv = np.column_stack([pos, vel, rad])
        L.initialize()
        orc.initialize()
        F = None
        for _ in range(25):
            # grains2lattice -> compute_hydrodynamic_force -> force2lattice -> collide -> stream
            O.vans_spread(orc, np.column_stack([pos, rad, np.zeros_like(pos)]), [0, 0, 0], DX, DT, RHO,
                          with_force=False)
            F = O.vans_drag(orc, gv, [0, 0, 0], DX, DT, RHO)
            O.vans_spread(orc, np.column_stack([pos, rad, F]), [0, 0, 0], DX, DT, RHO, with_force=True)
            orc.collide()
            orc.stream()
            L.grains2lattice()
            L.compute_hydrodynamic_force()
            L.force2lattice()
            L.collide()
            L.stream()
        fluid = orc.CT == O.FLUID
        assert np.abs(orc.vel[fluid]).max() < 0.05, "the oracle trajectory must stay stable"
        assert np.abs(L.volfrac.to_numpy() - orc.volfrac).max() < 2e-6
        assert np.abs(L.vel.to_numpy()[fluid] - orc.vel[fluid]).max() < 2e-4 * np.abs(orc.vel[fluid]).max()
        Fg = dem.gf.force_fluid.to_numpy()
        assert np.abs(Fg - F).max() < 5e-4 * np.abs(F).max()
    finally:
        BasicLattice3D.SmagorinskyConstant = old
