"""Free-slip walls (the FREE_SLIP branch of BasicLattice3D.stream, src/lbm3d/lbm_solver3d.py:230-236)
on the HIP path against the oracle, whose branch is pinned by mirror symmetry
(tests/test_oracle_free_slip.py).  Cases: a channel between free-slip planes (periodic elsewhere);
free-slip planes meeting obstacle planes; free-slip planes meeting each other at edges, where the
reference reflects into a cell that never collides and takes 0 (reproduced).  Even and odd step
counts (both population layouts), FP64 and FP32 storage, lagged Smagorinsky."""
import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_free_slip import _fields

pytestmark = pytest.mark.gpu


def _case(kind):
    nx, ny, nz = 9, 11, 8
    v, r = _fields(nx, ny - 2, nz, np.arange(-1, ny - 1))
    ct = np.full((nx, ny, nz), O.FLUID, dtype=np.int32)
    if kind == "channel":
        ct[:, 0, :] = O.FREE_SLIP | O.BOTTOM
        ct[:, -1, :] = O.FREE_SLIP | O.TOP
    elif kind == "with_obstacles":
        ct[0], ct[-1] = O.FREE_SLIP | O.LEFT, O.FREE_SLIP | O.RIGHT
        ct[:, 0, :] = ct[:, -1, :] = O.OBSTACLE
        ct[3:5, 4:6, 2:4] = O.OBSTACLE          # a block inside the flow
    elif kind == "edges":
        ct[:, 0, :] = O.FREE_SLIP | O.BOTTOM
        ct[:, -1, :] = O.FREE_SLIP | O.TOP
        ct[0] = O.FREE_SLIP | O.LEFT | (ct[0] & (O.BOTTOM | O.TOP))
        ct[-1] = O.FREE_SLIP | O.RIGHT | (ct[-1] & (O.BOTTOM | O.TOP))
    v[(ct & O.FLUID) == 0] = 0.0
    return ct, v, r


@pytest.mark.parametrize("kind", ["channel", "with_obstacles", "edges"])
@pytest.mark.parametrize("storage,steps", [("f64", 40), ("f64", 41), ("f32", 41)])
def test_free_slip_matches_oracle(kind, storage, steps):
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    ct, v, r = _case(kind)
    tau, cs = 0.58, 0.1
    orc = O.OracleLattice(*ct.shape, 1.0 / tau, kind="basic", cs=cs)
    orc.CT[...] = ct
    orc.vel[...] = v.astype(np.float32)
    orc.rho[...] = r.astype(np.float32)
    old = BasicLattice3D.SmagorinskyConstant
    BasicLattice3D.SmagorinskyConstant = cs
    try:
        L = BasicLattice3D(*ct.shape, 1.0 / tau, 1.0, 1.0, 1.0, storage=storage, smagorinsky="lagged")
        L.CT.from_numpy(ct)
        L.vel.from_numpy(v.astype(np.float32))
        L.rho.from_numpy(r.astype(np.float32))
        orc.initialize()
        L.initialize()
        orc.step(steps)
        L.step(steps)
        act = (ct & O.SKIP_BASIC) == 0
        f = L.f.to_numpy()
        tol = 1e-12 if storage == "f64" else 2e-5
        assert np.abs(f[act] - orc.f[act]).max() / np.abs(orc.f[act]).max() < tol
        u, ru = L.vel.to_numpy()[act], orc.vel[act]
        assert np.abs(u - ru).max() / np.abs(ru).max() < (1e-6 if storage == "f64" else 2e-4)
        if kind == "edges":   # the reference's edge reflection drops populations: mass is lost
            m0 = r[act].sum()
            assert orc.sums()[0] < m0 - 1e-3
    finally:
        BasicLattice3D.SmagorinskyConstant = old

