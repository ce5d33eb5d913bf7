"""PSC coupling at the resolved benchmark's own shape: 128 x 128 x 256 lattice with walls on every
face, 200 grains of radius 8 at seeded non-overlapping positions (seed 2026) plus one grain of
radius 12.5, all translating and spinning; FP32 populations.  At 201 grains the mapping runs 8
workgroups per grain, and the large grain's box (27^3 cells) takes several classification
chunks.  Three coupled steps against the oracle (psclattice3d.py:84-358): grain ids and sub-cell
counts exact, velocities, forces and torques within 2e-4."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

NX, NY, NZ = 128, 128, 256


def layout(seed=2026, n=200, R=8.0, big=12.5):
    rng = np.random.default_rng(seed)
    L = np.array([NX - 2, NY - 2, NZ - 2], dtype=np.float64)
    pos, rad = [], []
    while len(pos) < n + 1:
        r = R if len(pos) < n else big
        x = r + 0.5 + (L - 2 * r - 1.0) * rng.random(3)
        if pos and (np.linalg.norm(np.array(pos) - x, axis=1) < np.array(rad) + r + 1.0).any():
            continue
        pos.append(x)
        rad.append(r)
    pos, rad = np.array(pos), np.array(rad)
    vel = rng.uniform(-0.02, 0.02, pos.shape)
    omg = rng.uniform(-0.002, 0.002, pos.shape)
    return pos, vel, omg, rad


def test_psc_benchmark_shape_matches_oracle():
    from src.dem3d.demconfig import DEMSolverConfig, DomainBounds, HertzContactConfig
    from src.dem3d.demsolver import DEMSolver
    from src.lbdem.psclattice3d import PSCLattice3D
    pos, vel, omg, rad = layout()
    tau = 0.55
    cfg = DEMSolverConfig(domain=DomainBounds(0.0, NX - 2.0, 0.0, NY - 2.0, 0.0, NZ - 2.0), dt=0.1,
                          gravity=(0.0, 0.0, -1e-5),
                          contact_model=HertzContactConfig(pp_friction=0.3, pw_friction=0.3, pp_restitution=0.9,
                                                           pw_restitution=0.9))
    cfg.set_particle_properties(elastic_modulus=1e5, poisson_ratio=0.3, max_coordinate_number=16)
    cfg.set_wall_properties(elastic_modulus=1e5, poisson_ratio=0.3, density=2.5)
    dem = DEMSolver(cfg)
    dem.init_particle_arrays(radius=rad, mass=2.5 * 4.0 / 3.0 * np.pi * rad ** 3, position=pos, velocity=vel)
    dem.set_contact_model("hertz")
    dem.gf.omega.from_numpy(omg)
    lat = PSCLattice3D(NX, NY, NZ, 1.0 / tau, 1.0, 1.0, 1.0, dem, storage="f32")
    orc = O.OracleLattice(NX, NY, NZ, 1.0 / tau, kind="psc")
    orc.CT[0], orc.CT[-1], orc.CT[:, 0], orc.CT[:, -1], orc.CT[:, :, 0], orc.CT[:, :, -1] = (O.OBSTACLE,) * 6
    lat.CT.from_numpy(orc.CT)
    lat.initialize()
    orc.initialize()
    grains = np.column_stack([pos, vel, omg, rad])
    forces = None
    for step in range(3):
        O.psc_map(orc, grains, [0, 0, 0], 1.0, 1.0)
        orc.collide()
        forces = O.psc_forces(orc, grains, [0, 0, 0], 1.0, 1.0, 1.0)
        orc.stream()
        lat.collide()
        lat.stream()
        print(f"step {step} done", flush=True)
        if step == 0:
            gid = lat.id.to_numpy()
            assert np.array_equal(gid, orc.id), np.argwhere(gid != orc.id)[:5]
            cnt = np.rint(lat.volfrac.to_numpy().astype(np.float64) * 125).astype(int)
            assert np.array_equal(cnt, np.rint(orc.volfrac * 125).astype(int))
            assert (gid == len(rad) - 1).sum() > 7000   # the large grain is mapped
    fluid = orc.CT == O.FLUID
    u = orc.vel[fluid]
    assert np.abs(lat.vel.to_numpy()[fluid] - u).max() / np.abs(u).max() < 2e-4
    Ff, Tf = dem.gf.force_fluid.to_numpy(), dem.gf.moment_fluid.to_numpy()
    assert np.abs(Ff - forces[:, :3]).max() / np.abs(forces[:, :3]).max() < 2e-4
    assert np.abs(Tf - forces[:, 3:]).max() / np.abs(forces[:, 3:]).max() < 2e-4
