"""PSC coupling on the GPU vs the oracle: grain mapping, weighted collision, grain forces."""
import numpy as np
import pytest

from dem_cases import gpu_solver, make_config
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX, DT, RHO = 0.002, 0.001, 1000.0


def _case(nx=24, ny=22, nz=28, tau=0.55):
    """Five grains: 0/1 with overlapping partial shells, 2/3 interpenetrating (full-cover
    conflict), 4 touching a wall; all translating and rotating."""
    centres_lu = np.array([[7.3, 8.1, 9.0], [13.9, 8.4, 9.2], [11.0, 15.2, 18.6], [13.2, 15.9, 19.1],
                           [3.6, 17.0, 22.4]])
    R_lu = np.array([3.4, 3.1, 3.6, 3.3, 2.9])
    pos = (centres_lu - 0.5) * DX  # lattice coordinate = (x - xmin)/dx + 1/2, xmin = 0
    rng = np.random.default_rng(2)
    vel = rng.uniform(-0.02, 0.02, (5, 3)) * DX / DT
    omg = rng.uniform(-0.01, 0.01, (5, 3)) / DT
    rad = R_lu * DX
    mass = 2500.0 * 4 / 3 * np.pi * rad ** 3
    return nx, ny, nz, tau, pos, vel, omg, rad, mass


def _walls(CT):
    CT[0], CT[-1] = O.OBSTACLE, O.OBSTACLE
    CT[:, 0], CT[:, -1] = O.OBSTACLE, O.OBSTACLE
    CT[:, :, 0], CT[:, :, -1] = O.OBSTACLE, O.OBSTACLE


def _build(storage="f32"):
    from src.lbdem.psclattice3d import PSCLattice3D
    nx, ny, nz, tau, pos, vel, omg, rad, mass = _case()
    box = (0.0, (nx - 1) * DX, 0.0, (ny - 1) * DX, 0.0, (nz - 1) * DX)
    cfg = make_config(box=box, dt=DT / 10, g=(0, 0, 0))
    dem = gpu_solver(cfg, rad, mass, pos, vel)
    dem.gf.omega.from_numpy(omg)
    L = PSCLattice3D(nx, ny, nz, 1.0 / tau, DX, DT, RHO, dem, storage=storage)
    orc = O.OracleLattice(nx, ny, nz, 1.0 / tau, kind="psc")
    _walls(orc.CT)
    L.CT.from_numpy(orc.CT)
    grains = np.column_stack([pos, vel, omg, rad])
    return L, orc, dem, grains, cfg


def test_psc_mapping_matches_oracle_exactly():
    L, orc, dem, grains, cfg = _build()
    L.initialize()
    O.psc_map(orc, grains, [0, 0, 0], DX, DT)
    gid = L.id.to_numpy()
    assert np.array_equal(gid, orc.id), np.argwhere(gid != orc.id)[:5]
    cnt_gpu = np.rint(L.volfrac.to_numpy().astype(np.float64) * 125).astype(int)
    cnt_orc = np.rint(orc.volfrac * 125).astype(int)
    assert np.array_equal(cnt_gpu, cnt_orc)
    assert (cnt_orc == 125).sum() > 100 and ((cnt_orc > 0) & (cnt_orc < 125)).sum() > 100
    assert np.abs(L.weight.to_numpy() - orc.weight).max() < 1e-6
    assert np.abs(L.velsolid.to_numpy() - orc.velsolid).max() < 1e-6
    # both conflict rules are exercised: grain 3 loses full-cover cells to grain 2
    assert (gid == 2).sum() > 0 and (gid == 3).sum() > 0


def test_psc_coupled_steps_match_oracle():
    L, orc, dem, grains, cfg = _build()
    L.initialize()
    orc.initialize()
    steps = 40
    forces = None
    for _ in range(steps):
        O.psc_map(orc, grains, [0, 0, 0], DX, DT)
        orc.collide()
        forces = O.psc_forces(orc, grains, [0, 0, 0], DX, DT, RHO)
        orc.stream()
        L.collide()
        L.stream()
    fluid = orc.CT == O.FLUID
    u_orc = orc.vel[fluid]
    assert np.abs(L.vel.to_numpy()[fluid] - u_orc).max() / np.abs(u_orc).max() < 2e-4
    Ff = dem.gf.force_fluid.to_numpy()
    Tf = dem.gf.moment_fluid.to_numpy()
    fs = np.abs(forces[:, :3]).max()
    ts = np.abs(forces[:, 3:]).max()
    assert np.abs(Ff - forces[:, :3]).max() / fs < 2e-4
    assert np.abs(Tf - forces[:, 3:]).max() / ts < 2e-4


def test_psc_closed_forms():
    """A grain at rest in fluid at rest: exactly no momentum exchange.  Per covered cell the
    collision conserves mass and hands -hydroforce of momentum to the fluid."""
    from src.lbdem.psclattice3d import PSCLattice3D
    nx, ny, nz, tau, pos, vel, omg, rad, mass = _case()
    box = (0.0, (nx - 1) * DX, 0.0, (ny - 1) * DX, 0.0, (nz - 1) * DX)
    cfg = make_config(box=box, dt=DT / 10, g=(0, 0, 0))
    dem = gpu_solver(cfg, rad, mass, pos, 0 * vel)
    L = PSCLattice3D(nx, ny, nz, 1.0 / tau, DX, DT, RHO, dem, storage="f64")
    CT = L.CT.to_numpy()
    _walls(CT)
    L.CT.from_numpy(CT)
    L.initialize()
    L.collide()
    assert np.abs(dem.gf.force_fluid.to_numpy()).max() == 0.0
    assert np.abs(dem.gf.moment_fluid.to_numpy()).max() == 0.0
