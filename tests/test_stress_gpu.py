"""Stress diagnostics on the GPU vs the oracle: compute_stress and compute_plane_average_stress
after a lid-driven cavity with lagged Smagorinsky (exact: the lag state keeps the last
collision's equilibrium and omega) and in steady forced Poiseuille flow (BGK)."""
import numpy as np
import pytest

from lbm_cases import gpu_from_oracle, oracle_cavity
from oracle import oracle as O

pytestmark = pytest.mark.gpu

NAMES = ("stress_xx", "stress_yy", "stress_zz", "stress_xy", "stress_xz", "stress_yz", "shear_stress_mag")


def _gpu_stress(L):
    return np.stack([getattr(L, n).to_numpy() for n in NAMES])


# the stress fields are float32 (as the reference's default fields): FP64 storage agrees to
# their rounding
@pytest.mark.parametrize("storage,tol", [("f64", 3e-7), ("f32", 2e-4)])
def test_cavity_stress_matches_oracle(storage, tol):
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    orc = oracle_cavity(14, tau=0.56, lid=0.1, cs=0.17)
    L = gpu_from_oracle(BasicLattice3D, orc, storage=storage)
    BasicLattice3D.SmagorinskyConstant, cs = 0.17, BasicLattice3D.SmagorinskyConstant
    try:
        orc.initialize()
        L.initialize()
        orc.step(37)
        L.step(37)
        L.compute_stress()
    finally:
        BasicLattice3D.SmagorinskyConstant = cs
    ref = O.stress(orc)
    got = _gpu_stress(L)
    act = (orc.CT & O.SKIP_BASIC) == 0
    scale = np.abs(ref[:6][:, act]).max()
    assert scale > 1e-4
    assert np.abs(got[:, act] - ref[:, act]).max() <= tol * scale
    assert (got[:, ~act] == 0).all()
    for axis in range(3):
        for pos in (1, 6, 12):
            a = np.array(L.compute_plane_average_stress(axis, pos))
            b = np.array(O.plane_stress(orc, ref, axis, pos))
            assert np.abs(a - b).max() <= tol * scale, (axis, pos, a, b)
    assert L.compute_plane_average_stress(2, 0) == (0.0, 0.0, 0.0)


def test_poiseuille_stress_matches_oracle_and_the_force_balance():
    """BGK (ForcedLattice3D): at steady state the current populations' equilibrium is the last
    collision's, so the stress matches the oracle; sigma_xz = -rho F (z - z_c)."""
    from src.lbm3d.lbm_foecelover import ForcedLattice3D
    F, nz, tau = 1e-5, 18, 0.8
    orc = O.OracleLattice(4, 4, nz, 1.0 / tau, kind="forced", F=(F, 0.0, 0.0))
    orc.CT[:, :, 0] = O.OBSTACLE
    orc.CT[:, :, -1] = O.OBSTACLE
    L = gpu_from_oracle(ForcedLattice3D, orc, storage="f64", F=(F, 0.0, 0.0))
    orc.initialize()
    L.initialize()
    orc.step(4000)
    L.step(4000)
    L.compute_stress()
    ref = O.stress(orc)
    got = _gpu_stress(L)
    fl = orc.CT == O.FLUID
    assert np.abs(got[:, fl] - ref[:, fl]).max() <= 1e-6 * np.abs(ref[4][fl]).max()
    z = np.arange(1, nz - 1)
    expect = -F * (z - (nz - 1) / 2.0)
    assert np.abs(got[4, 2, 2, 1:-1] - expect).max() <= 2e-3 * np.abs(expect).max()
    xy, xz, yz = L.compute_plane_average_stress(2, 3)
    assert xz == pytest.approx(float(got[4, 0, 0, 3]), rel=1e-6)
