"""The oracle's stress diagnostics (compute_stress, compute_plane_average_stress) against what
the mathematics fixes: no stress in fluid at rest, and in steady forced Poiseuille flow the shear
stress balances the body force, sigma_xz(z) = -rho F (z - z_c).  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O


def test_fluid_at_rest_has_no_stress():
    L = O.OracleLattice(6, 5, 7, 1.0 / 0.7, kind="basic")
    L.initialize()
    st = O.stress(L)
    assert np.abs(st).max() == 0.0


def _poiseuille(F=1e-5, nz=18, tau=0.8, steps=4000):
    L = O.OracleLattice(4, 3, nz, 1.0 / tau, kind="forced", F=(F, 0.0, 0.0))
    L.CT[:, :, 0] = O.OBSTACLE
    L.CT[:, :, -1] = O.OBSTACLE
    L.initialize()
    L.step(steps)
    return L


def test_poiseuille_shear_stress_balances_the_body_force():
    F, nz = 1e-5, 18
    L = _poiseuille(F, nz)
    st = O.stress(L)
    zc = (nz - 1) / 2.0  # bounce-back walls half way between z = 0, 1 and z = nz - 2, nz - 1
    z = np.arange(1, nz - 1)
    sxz = st[4, 1, 1, 1:-1]
    expect = -L.rho[1, 1, 1:-1] * F * (z - zc)
    np.testing.assert_allclose(sxz, expect, rtol=0, atol=2e-3 * np.abs(expect).max())
    # the other shear components vanish; the plane average is the cell value (uniform in x, y)
    assert np.abs(st[3]).max() < 1e-12 and np.abs(st[5]).max() < 1e-12
    for k in (1, 5, nz - 2):
        xy, xz, yz = O.plane_stress(L, st, 2, k)
        assert xz == pytest.approx(st[4, 0, 0, k], rel=1e-12)
        assert xy == pytest.approx(0.0, abs=1e-14) and yz == pytest.approx(0.0, abs=1e-14)
    assert O.plane_stress(L, st, 2, 0) == (0.0, 0.0, 0.0)  # a wall plane holds no fluid cell
    # shear magnitude sqrt((xy^2 + xz^2 + yz^2) / 2)
    np.testing.assert_allclose(st[6], np.sqrt(0.5 * (st[3] ** 2 + st[4] ** 2 + st[5] ** 2)), rtol=1e-14)
