"""Hybrid coupling (HybridLattice3D) on the GPU vs the oracle: coarse keys, fine spreading with
the free-fraction correction, the four-way collision with lagged Smagorinsky, coarse forces and
torques, fine Tenneti drag."""
import numpy as np
import pytest

from dem_cases import gpu_solver, make_config
from oracle import oracle as O

pytestmark = pytest.mark.gpu

DX, DT, RHO = 0.002, 0.001, 1000.0
NX, NY, NZ = 20, 18, 24
TAU = 0.56


def _grains():
    """Two coarse grains (translating, rotating), 40 fine grains scattered through the box,
    some inside the coarse grains' partial shells, one next to a wall (mirror kernel), and four
    fast fine grains on one spot of grain 0's shell, where B + beta > 1 (the clip)."""
    rng = np.random.default_rng(11)
    coarse_x = np.array([[7.2, 8.3, 9.6], [13.1, 9.4, 15.8]])
    coarse_R = np.array([3.3, 2.7])
    fine_x = rng.uniform([1.6, 1.6, 0.5], [NX - 2.6, NY - 2.6, NZ - 0.5], (38, 3))
    fine_x = np.vstack([fine_x, [[7.2, 8.3, 5.9], [1.3, 9.0, 12.0]]])
    fine_R = rng.uniform(0.25, 0.45, len(fine_x))
    x = np.vstack([coarse_x, fine_x, [[10.1, 8.3, 9.6]] * 4])
    R = np.concatenate([coarse_R, fine_R, [0.5] * 4])
    vel = rng.uniform(-0.02, 0.02, (len(x), 3)) * DX / DT
    vel[-4:] = np.array([0.0, 0.0, 0.05]) * DX / DT
    omg = np.zeros((len(x), 3))
    omg[:2] = rng.uniform(-0.01, 0.01, (2, 3)) / DT
    return (x - 0.5) * DX, vel, omg, R * DX


def _build(storage="f32", kernel=0):
    from src.hybridlbdem.hybridlattice import HybridLattice3D
    pos, vel, omg, rad = _grains()
    mass = 2500.0 * 4 / 3 * np.pi * rad ** 3
    box = (0.0, (NX - 1) * DX, 0.0, (NY - 1) * DX, 0.0, (NZ - 1) * DX)
    cfg = make_config(box=box, dt=DT / 10, g=(0, 0, 0))
    dem = gpu_solver(cfg, rad, mass, pos, vel)
    dem.gf.omega.from_numpy(omg)
    L = HybridLattice3D(NX, NY, NZ, 1.0 / TAU, DX, DT, RHO, dem, storage=storage, smagorinsky="lagged")
    L.kernel = kernel
    orc = O.OracleLattice(NX, NY, NZ, 1.0 / TAU, kind="hybrid", cs=L.SmagorinskyConstant)
    orc.CT[0], orc.CT[-1] = O.OBSTACLE, O.OBSTACLE
    orc.CT[:, 0], orc.CT[:, -1] = O.OBSTACLE, O.OBSTACLE
    x = np.arange(NX)
    orc.vel[..., 2] = 0.03 * np.sin(np.pi * (x - 0.5) / (NX - 2))[:, None, None]
    orc.vel[..., 0] = 0.01 * np.cos(2 * np.pi * np.arange(NZ) / NZ)[None, None, :]
    orc.vel[orc.CT == O.OBSTACLE] = 0
    L.CT.from_numpy(orc.CT)
    L.vel.from_numpy(orc.vel.astype(np.float32))
    grains = np.column_stack([pos, vel, omg, rad])
    H = O.HybridOracle(orc, grains, [0, 0, 0], DX, DT, RHO)
    return L, H, dem


def _rel(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


def test_hybrid_mapping_matches_oracle():
    L, H, dem = _build()
    L.initialize_complete()
    H.initialize_complete()
    orc = H.L
    assert np.array_equal(L.coarse_id.to_numpy(), orc.id)
    cnt = np.rint(L.coarse_volfrac.to_numpy().astype(np.float64) * 125).astype(int)
    assert np.array_equal(cnt, np.rint(orc.volfrac * 125).astype(int))
    assert _rel(L.coarse_velsolid.to_numpy(), orc.velsolid) < 1e-5
    fv = L.fine_volfrac.to_numpy()
    assert _rel(fv, orc.fine_volfrac) < 1e-5
    assert (orc.fine_volfrac > 0).sum() > 300 and ((orc.fine_volfrac > 0) & (orc.volfrac > 0)).sum() > 5
    assert (fv[orc.volfrac >= 1] == 0).all()
    cov = orc.fine_volfrac > 1e-6
    assert _rel(L.fine_velsolid.to_numpy()[cov], orc.fine_velsolid[cov]) < 1e-4
    assert L.clamp_count() == H.warnings == 0
    # inspection fields: fine Tenneti weight where defined, coarse weight B with omega0 at t = 0
    fw = L.fine_weight.to_numpy()
    d = (orc.fine_volfrac > 0) & (orc.CT == O.FLUID)
    assert _rel(fw[d], orc.fine_weight[d]) < 1e-4
    eps = orc.volfrac
    tm = TAU - 0.5
    B = np.where(eps > 0, eps * tm / ((1 - eps) + tm), 0)
    assert np.abs(L.coarse_weight.to_numpy() - B).max() < 1e-6


@pytest.mark.parametrize("storage,kernel,tol", [("f32", 0, 2e-4), ("f32", 1, 2e-4), ("f32", 3, 2e-4),
                                                ("f64", 0, 2e-5), ("bf16", 0, 2e-2)])
def test_hybrid_coupled_steps_match_oracle(storage, kernel, tol):
    L, H, dem = _build(storage, kernel)
    L.initialize_complete()
    H.initialize_complete()
    forces = None
    for _ in range(12):
        L.prepare_step()
        L.collide()
        L.stream()
        L.lattice2grains()
        forces = H.step()
    orc = H.L
    both = (orc.volfrac > 0) & (orc.fine_volfrac > 0)
    assert ((orc.weight + orc.fine_weight > 1) & both).any()  # the clip is exercised
    fluid = orc.CT == O.FLUID
    print(f"hybrid {storage} kernel {kernel}: vel err {_rel(L.vel.to_numpy()[fluid], orc.vel[fluid]):.2e}",
          flush=True)
    assert _rel(L.vel.to_numpy()[fluid], orc.vel[fluid]) < tol
    Ff = dem.gf.force_fluid.to_numpy()
    Tf = dem.gf.moment_fluid.to_numpy()
    coarse = slice(0, 2)
    assert _rel(Ff[coarse], forces[coarse, :3]) < 5 * tol
    assert _rel(Tf[coarse], forces[coarse, 3:]) < 5 * tol
    fine = slice(2, None)
    assert _rel(Ff[fine], forces[fine, :3]) < 5 * tol
    assert (Tf[fine] == 0).all()
    assert np.abs(forces[coarse, :3]).max() > 0 and np.abs(forces[fine, :3]).max() > 0


def test_hybrid_reduces_to_psc_without_fine_grains():
    """Only coarse grains, constant omega (Smagorinsky constant 0): the hybrid lattice steps like
    the PSC oracle (whose weight uses omega0)."""
    from src.hybridlbdem.hybridlattice import HybridLattice3D
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    pos, vel, omg, rad = _grains()
    pos, vel, omg, rad = pos[:2], vel[:2], omg[:2], rad[:2]
    mass = 2500.0 * 4 / 3 * np.pi * rad ** 3
    box = (0.0, (NX - 1) * DX, 0.0, (NY - 1) * DX, 0.0, (NZ - 1) * DX)
    dem = gpu_solver(make_config(box=box, dt=DT / 10, g=(0, 0, 0)), rad, mass, pos, vel)
    dem.gf.omega.from_numpy(omg)
    cs = BasicLattice3D.SmagorinskyConstant
    try:
        BasicLattice3D.SmagorinskyConstant = 0.0
        L = HybridLattice3D(NX, NY, NZ, 1.0 / TAU, DX, DT, RHO, dem, storage="f64", smagorinsky="lagged")
        orc = O.OracleLattice(NX, NY, NZ, 1.0 / TAU, kind="psc")
        orc.CT[0], orc.CT[-1] = O.OBSTACLE, O.OBSTACLE
        orc.vel[..., 2] = 0.02
        orc.vel[orc.CT == O.OBSTACLE] = 0
        L.CT.from_numpy(orc.CT)
        L.vel.from_numpy(orc.vel.astype(np.float32))
        L.initialize_complete()
        orc.initialize()
        grains = np.column_stack([pos, vel, omg, rad])
        for _ in range(8):
            L.prepare_step()
            L.collide()
            L.lattice2grains()
            O.psc_map(orc, grains, [0, 0, 0], DX, DT)
            orc.collide()
            F = O.psc_forces(orc, grains, [0, 0, 0], DX, DT, RHO)
            orc.stream()
    finally:
        BasicLattice3D.SmagorinskyConstant = cs
    fluid = orc.CT == O.FLUID
    assert _rel(L.vel.to_numpy()[fluid], orc.vel[fluid]) < 1e-5
    assert _rel(dem.gf.force_fluid.to_numpy(), F[:, :3]) < 1e-4
    assert _rel(dem.gf.moment_fluid.to_numpy(), F[:, 3:]) < 1e-4
