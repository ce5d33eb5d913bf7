"""Restart snapshots: a lattice (populations, step parity, lag state, fields) and a DEM solver
(rotation state and contact history) restored into fresh objects continue the run."""
import numpy as np
import pytest

from dem_cases import gpu_solver, make_config
from lbm_cases import gpu_from_oracle, oracle_cavity
from test_psc_gpu import _build as build_psc

pytestmark = pytest.mark.gpu


def _pops(L):
    return L.f.to_numpy()


@pytest.mark.parametrize("storage", ["f32", "bf16", "f64"])
@pytest.mark.parametrize("t0", [6, 7])
def test_lattice_restart_is_exact(tmp_path, storage, t0):
    """Cavity with lagged Smagorinsky: t0 steps, snapshot, 5 more steps; a fresh lattice loaded
    from the snapshot reproduces those 5 steps bit for bit (either parity of the AA state)."""
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    orc = oracle_cavity(12, tau=0.56, lid=0.08)
    A = gpu_from_oracle(BasicLattice3D, orc, storage=storage)
    A.initialize()
    A.step(t0)
    path = tmp_path / "lat.npz"
    A.save_state(path)
    A.step(5)
    B = gpu_from_oracle(BasicLattice3D, oracle_cavity(12, tau=0.56, lid=0.0), storage=storage)
    B.load_state(path)
    assert B.t == t0
    B.step(5)
    assert np.array_equal(_pops(A), _pops(B))
    assert np.array_equal(A.vel.to_numpy(), B.vel.to_numpy())
    assert np.array_equal(A.omega.to_numpy(), B.omega.to_numpy())


def test_lattice_restart_into_another_storage(tmp_path):
    """An FP32 snapshot restored into an FP64 lattice (through the FP64 populations) continues
    the run to FP32 accuracy."""
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    orc = oracle_cavity(12, tau=0.56, lid=0.08)
    A = gpu_from_oracle(BasicLattice3D, orc, storage="f32")
    A.initialize()
    A.step(9)
    A.save_state(tmp_path / "lat.npz")
    A.step(4)
    B = gpu_from_oracle(BasicLattice3D, orc, storage="f64")
    B.load_state(tmp_path / "lat.npz")
    B.step(4)
    fluid = orc.CT == 1
    va, vb = A.vel.to_numpy()[fluid], B.vel.to_numpy()[fluid]
    assert np.abs(va - vb).max() <= 1e-5 * np.abs(va).max()


def test_dem_restart_keeps_rotation_and_contact_history(tmp_path):
    """A settling bed with Hertz contacts (multi-kernel path): restored from the snapshot the
    run continues as the original; restored without the contact history it does not."""
    R = 0.004
    box = (0.0, 0.05, 0.0, 0.05, 0.0, 0.06)
    g = np.arange(6) * 1.99 * R + 1.01 * R
    X, Y, Z = np.meshgrid(g, g, np.arange(4) * 1.99 * R + 0.995 * R, indexing="ij")
    pos = np.column_stack([X.ravel(), Y.ravel(), Z.ravel()])  # slightly overlapping stack on the floor
    rng = np.random.default_rng(5)
    pos[:, :2] += rng.uniform(-0.05, 0.05, (len(pos), 2)) * R
    rad = np.full(len(pos), R)
    mass = 2500.0 * 4 / 3 * np.pi * rad ** 3
    vel = rng.normal(0, 0.02, (len(pos), 3))
    cfg = make_config(box=box, dt=2e-5, e=0.5, mu=0.4)
    A = gpu_solver(cfg, rad, mass, pos, vel, path=2)
    A.run_substeps(400)
    path = tmp_path / "dem.npz"
    A.save_state(path)
    st = A.state_dict()
    assert st["_cf_active"].sum() > 20 and st["_wcf_active"].sum() > 20 and np.abs(st["gf.omega"]).max() > 0
    A.run_substeps(40)
    B = gpu_solver(cfg, rad, mass, pos, vel, path=2)
    B.load_state(path)
    B.run_substeps(40)
    C = gpu_solver(cfg, rad, mass, pos, vel, path=2)
    d = dict(st)
    for k in ("_cf_active", "_cfn", "_wcf_active"):
        d[k] = np.zeros_like(d[k])
    C.load_state_dict(d)
    C.run_substeps(40)
    pa, pb, pc = (s.gf.position.to_numpy() for s in (A, B, C))
    wa, wb = A.gf.omega.to_numpy(), B.gf.omega.to_numpy()
    scale = np.abs(pa - st["gf.position"]).max()
    assert np.abs(pa - pb).max() <= 1e-9 * scale + 1e-15
    assert np.abs(wa - wb).max() <= 1e-9 * np.abs(wa).max()
    assert np.abs(pa - pc).max() > 1e3 * max(np.abs(pa - pb).max(), 1e-18)
    A.check()
    B.check()


def test_psc_coupled_restart(tmp_path):
    """PSC lattice + DEM: snapshot both after 6 coupled steps, continue 5; the restored pair
    follows the original (FP64 DEM forces are summed with atomics, so to round-off)."""
    L, orc, dem, grains, cfg = build_psc()
    L.initialize()
    for _ in range(6):
        L.collide()
        L.stream()
        dem.run_substeps(5)
    L.save_state(tmp_path / "lat.npz")
    dem.save_state(tmp_path / "dem.npz")
    for _ in range(5):
        L.collide()
        L.stream()
        dem.run_substeps(5)
    L2, _, dem2, _, _ = build_psc()
    L2.load_state(tmp_path / "lat.npz")
    dem2.load_state(tmp_path / "dem.npz")
    for _ in range(5):
        L2.collide()
        L2.stream()
        dem2.run_substeps(5)
    fluid = orc.CT == 1
    va, vb = L.vel.to_numpy()[fluid], L2.vel.to_numpy()[fluid]
    assert np.abs(va - vb).max() <= 1e-5 * np.abs(va).max()
    for k in ("position", "velocity", "omega"):
        a, b = getattr(dem.gf, k).to_numpy(), getattr(dem2.gf, k).to_numpy()
        assert np.abs(a - b).max() <= 1e-6 * max(np.abs(a).max(), 1e-30), k
    Fa, Fb = dem.gf.force_fluid.to_numpy(), dem2.gf.force_fluid.to_numpy()
    assert np.abs(Fa - Fb).max() <= 1e-4 * np.abs(Fa).max()
