"""The oracle's wet-node boundaries (vel_zouHe, pre_zouHe, bc_vel_exit, bc_ladd_inlet,
lbm_solver3d.py:516-957) against what the mathematics fixes: the rebuilt populations carry
exactly the prescribed density and velocity moments, the exit cell holds the equilibrium of its
neighbour's velocity, and a Zou-He channel develops the Poiseuille profile.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O

C = np.array([[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
              [1, 1, 0], [-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 0, 1], [-1, 0, -1], [1, 0, -1],
              [-1, 0, 1], [0, 1, 1], [0, -1, -1], [0, 1, -1], [0, -1, 1]])
W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)
FACES = [(O.LEFT, 0, 0), (O.RIGHT, 0, -1), (O.BOTTOM, 1, 0), (O.TOP, 1, -1), (O.BACK, 2, 0), (O.FRONT, 2, -1)]


def _feq(rho, u):
    cu = C @ u
    return W * rho * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * u @ u)


def _cell(face_axis, side, n=5):
    idx = [2, 2, 2]
    idx[face_axis] = side if side == 0 else n - 1
    return tuple(idx)


@pytest.mark.parametrize("face,axis,side", FACES)
def test_zouhe_velocity_moments(face, axis, side):
    """After the rebuild: sum f = rho (the Zou-He density) and sum f c = rho u_bc, every face."""
    rng = np.random.default_rng(axis * 2 + (side != 0))
    L = O.OracleLattice(5, 5, 5, 1.0, kind="basic")
    c = _cell(axis, side)
    L.CT[c] = O.VEL_ZOUHE | face
    u = np.array([0.03, -0.02, 0.015])
    L.vel[c] = u
    L.f[c] = _feq(1.0, np.array([0.01, 0.02, -0.01])) * (1 + 0.05 * rng.standard_normal(19))
    O.lib().lbo_boundaries_f64(*L._g(), O.P(L.CT), O.P(L.f), O.P(L.rho), O.P(L.vel))
    f = L.f[c]
    assert f.sum() == pytest.approx(L.rho[c], rel=1e-14)
    assert np.allclose(f @ C, L.rho[c] * u, rtol=0, atol=1e-15)


@pytest.mark.parametrize("face,axis,side", FACES)
def test_zouhe_pressure_moments(face, axis, side):
    """sum f = the prescribed rho, normal momentum rho u_n, transverse momentum 0."""
    rng = np.random.default_rng(10 + axis)
    L = O.OracleLattice(5, 5, 5, 1.0, kind="basic")
    c = _cell(axis, side)
    L.CT[c] = O.PRE_ZOUHE | face
    L.rho[c] = 1.02
    L.f[c] = _feq(1.0, np.array([0.02, 0.01, 0.0])) * (1 + 0.05 * rng.standard_normal(19))
    O.lib().lbo_boundaries_f64(*L._g(), O.P(L.CT), O.P(L.f), O.P(L.rho), O.P(L.vel))
    f = L.f[c]
    assert f.sum() == pytest.approx(1.02, rel=1e-14)
    mom = f @ C
    assert mom[axis] == pytest.approx(1.02 * L.vel[c][axis], rel=1e-13)
    for t in range(3):
        if t != axis:
            assert abs(mom[t]) < 1e-15


def test_exit_and_ladd_inlet():
    L = O.OracleLattice(5, 4, 3, 1.0, kind="basic")
    L.CT[4, 1, 1] = O.VEL_EXIT | O.RIGHT
    L.CT[0, 2, 1] = O.VEL_INLET_LADD | O.LEFT
    u = np.array([0.04, 0.01, -0.005])
    L.vel[3, 1, 1] = u
    L.rho[1, 2, 1] = 1.013
    O.lib().lbo_boundaries_f64(*L._g(), O.P(L.CT), O.P(L.f), O.P(L.rho), O.P(L.vel))
    assert np.array_equal(L.vel[4, 1, 1], u) and L.rho[4, 1, 1] == 1.0
    assert np.allclose(L.f[4, 1, 1], _feq(1.0, u), rtol=1e-15, atol=0)
    assert L.rho[0, 2, 1] == 1.013


def test_zouhe_channel_develops_poiseuille():
    """Walls at y = 0, ny-1, periodic z, a parabolic Zou-He velocity inlet at x = 0 and a Zou-He
    pressure outlet (rho = 1) at x = nx-1: the steady velocity is the inlet parabola at every x
    (fully developed flow, halfway walls at y = 0.5 and ny - 1.5)."""
    nx, ny, nz = 14, 11, 3
    L = O.OracleLattice(nx, ny, nz, 1.0 / 0.8, kind="basic", cs=0.0)
    L.CT[:, 0, :] = O.OBSTACLE
    L.CT[:, -1, :] = O.OBSTACLE
    y = np.arange(ny)
    h = ny - 2.0
    prof = np.clip(4 * 0.02 * (y - 0.5) * (ny - 1.5 - y) / h ** 2, 0, None)
    L.CT[0, 1:-1, :] = O.VEL_ZOUHE | O.LEFT
    L.CT[-1, 1:-1, :] = O.PRE_ZOUHE | O.RIGHT
    L.vel[0, :, :, 0] = prof[:, None]
    L.initialize()
    L.step(4000)
    ux = L.vel[:, 1:-1, 1, 0]
    rho = L.rho[:, 1:-1, 1]
    # steady: the same mass flux through every plane x = const (the density falls linearly from
    # inlet to outlet, so u grows by the same 0.8 %), a parabolic profile at every x
    flux = (rho * ux).sum(1)
    assert np.abs(flux / flux[0] - 1).max() < 1e-6
    shape = ux / ux[:, ny // 2 - 1: ny // 2]
    assert np.abs(shape - (prof[1:-1] / prof.max())[None, :]).max() < 2e-2
    assert rho[-1].max() == pytest.approx(1.0, abs=1e-12) and rho[0].min() > 1.005
    assert np.abs(L.vel[1:-1, 1:-1, :, 1]).max() < 5e-3 * prof.max()  # cross flow near the inlet
