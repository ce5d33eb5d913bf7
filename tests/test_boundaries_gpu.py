"""Wet-node boundaries (Zou-He velocity and pressure, convective exit, Ladd velocity inlet) of the
HIP lattice against the oracle (lbm_solver3d.py:241-250, 516-957)."""
import numpy as np
import pytest

from lbm_cases import gpu_from_oracle, rel_err
from oracle import oracle as O

pytestmark = pytest.mark.gpu

LOW = {0: O.LEFT, 1: O.BOTTOM, 2: O.BACK}
HIGH = {0: O.RIGHT, 1: O.TOP, 2: O.FRONT}


@pytest.fixture
def smag017():
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    old = BasicLattice3D.SmagorinskyConstant
    BasicLattice3D.SmagorinskyConstant = 0.17
    yield
    BasicLattice3D.SmagorinskyConstant = old


def _channel(axis, inlet, outlet, shape=(14, 9, 8), real=np.float64, u0=0.03):
    """Flow along `axis` between walls on the next axis, periodic on the third; inlet cells at
    the low end (with a parabolic velocity), outlet cells at the high end."""
    L = O.OracleLattice(*shape, 1.0 / 0.8, kind="basic", cs=0.17, real=real)
    wall_axis = (axis + 1) % 3
    sl = [slice(None)] * 3
    sl[wall_axis] = 0
    L.CT[tuple(sl)] = O.OBSTACLE
    sl[wall_axis] = -1
    L.CT[tuple(sl)] = O.OBSTACLE
    n = shape[wall_axis]
    y = np.arange(n)
    prof = np.clip(4 * u0 * (y - 0.5) * (n - 1.5 - y) / (n - 2.0) ** 2, 0, None)
    lo = [slice(None)] * 3
    lo[axis] = 0
    hi = [slice(None)] * 3
    hi[axis] = -1
    lo_in, hi_in = list(lo), list(hi)
    lo_in[wall_axis] = hi_in[wall_axis] = slice(1, -1)
    L.CT[tuple(lo_in)] = inlet | LOW[axis]
    L.CT[tuple(hi_in)] = outlet | HIGH[axis]
    shp = [1, 1, 1]
    shp[wall_axis] = n
    v = np.zeros(shape + (3,))
    v[..., axis] = prof.reshape(shp)
    L.vel[tuple(lo)] = v[tuple(lo)].astype(np.float32)
    return L


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("kernel", [1, 2, 3])
def test_zouhe_channel_matches_oracle(smag017, axis, kernel):
    """Zou-He velocity inlet and pressure outlet on the two faces of each axis (FP64 storage,
    lagged Smagorinsky), per-cell and row-tiled kernels: populations, inlet density and outlet
    velocity after 151 steps."""
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    orc = _channel(axis, O.VEL_ZOUHE, O.PRE_ZOUHE)
    gpu = gpu_from_oracle(BasicLattice3D, orc, storage="f64")
    gpu.kernel = kernel
    orc.initialize()
    gpu.initialize()
    orc.step(151)
    gpu.step(151)
    act = (orc.CT & (O.OBSTACLE | O.VEL_LADD)) == 0
    assert rel_err(gpu.f.to_numpy()[act], orc.f[act]) < 1e-11
    wet = (orc.CT & (O.VEL_ZOUHE | O.PRE_ZOUHE)) != 0
    assert rel_err(gpu.rho.to_numpy()[wet], orc.rho[wet]) < 1e-6   # float32 fields
    assert rel_err(gpu.vel.to_numpy()[wet], orc.vel[wet]) < 1e-5
    fluid = orc.CT == O.FLUID
    assert rel_err(gpu.vel.to_numpy(), orc.vel, fluid) < 1e-5


@pytest.mark.parametrize("storage,tol", [("f64", 1e-8), ("f32", 2e-4)])
def test_ladd_inlet_and_exit_match_oracle(smag017, storage, tol):
    """Ladd velocity inlet (bounce-back with the inlet cell's density, which follows its inner
    neighbour) and convective exit (neighbour velocity, rho = 1, equilibrium populations).  The
    inlet density reaches the Ladd term through the float32 rho field (the reference's fields are
    float32 too), hence 1e-8 and not FP64 rounding at FP64 storage."""
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    orc = _channel(0, O.VEL_INLET_LADD, O.VEL_EXIT)
    gpu = gpu_from_oracle(BasicLattice3D, orc, storage=storage)
    orc.initialize()
    gpu.initialize()
    orc.step(120)
    gpu.step(120)
    act = (orc.CT & O.SKIP_BASIC) == 0
    assert rel_err(gpu.f.to_numpy()[act], orc.f[act]) < tol
    inl = (orc.CT & O.VEL_INLET_LADD) != 0
    assert rel_err(gpu.rho.to_numpy()[inl], orc.rho[inl]) < max(tol, 1e-6)
    ex = (orc.CT & O.VEL_EXIT) != 0
    assert rel_err(gpu.vel.to_numpy()[ex], orc.vel[ex]) < max(tol, 1e-5)
