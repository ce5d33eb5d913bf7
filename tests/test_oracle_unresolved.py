"""The oracle's unresolved-coupling functions (EqIBM, VANS) against what the mathematics fixes:
the Roma-Peskin kernel's closed-form values and its partition of unity, the Schiller-Naumann and
Stokes limits of the Tenneti drag, and conservation of spread volume, momentum and force.
CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O

DX, DT, RHO = 1e-3, 5e-4, 1000.0


def test_threedelta_closed_form_values():
    # phi(0) = 2/3, phi(1/2) = 1/2 from both branches, phi(1) = 1/6, phi(3/2) = 0
    assert O.threedelta(0.0) == pytest.approx(2.0 / 3.0, abs=1e-15)
    assert O.threedelta(0.5) == pytest.approx(0.5, abs=1e-15)
    assert O.threedelta(0.5 - 1e-12) == pytest.approx(0.5, abs=1e-9)
    assert O.threedelta(1.0) == pytest.approx(1.0 / 6.0, abs=1e-15)
    assert O.threedelta(1.5) == pytest.approx(0.0, abs=1e-15)
    assert O.threedelta(1.6) == 0.0


@pytest.mark.parametrize("x", [0.0, 0.13, 0.3, 0.5, 0.77, 0.999])
def test_threedelta_partition_of_unity_and_first_moment(x):
    """The 1-D three-point kernel sums to 1 over the integer grid and has zero first moment."""
    pts = np.arange(-3, 5)
    w = np.array([O.threedelta(abs(x - i)) for i in pts])
    assert w.sum() == pytest.approx(1.0, abs=1e-14)
    assert ((x - pts) * w).sum() == pytest.approx(0.0, abs=1e-14)


@pytest.mark.parametrize("Re", [0.0, 0.5, 3.0, 40.0])
def test_tenneti_dilute_limit_is_schiller_naumann(Re):
    assert O.tenneti_cd(Re, 0.0) == pytest.approx(1.0 + 0.15 * Re ** 0.687, rel=1e-14)


def test_tenneti_fully_solid_is_zero():
    assert O.tenneti_cd(1.0, 1.0) == 0.0


def _box(n=12, kind="eqibm", walls=False):
    L = O.OracleLattice(n, n, n, 1.0 / 0.6, kind=kind)
    if walls:
        L.CT[0] = O.OBSTACLE
    return L


@pytest.mark.parametrize("walls", [False, True])
def test_ibm_spread_conserves_volume_and_momentum(walls):
    """Normalised spreading: sum of eps = V_p / dx^3 and sum of eps * u_s = V_p / dx^3 * v dt/dx,
    also for a particle next to a wall (mirror images; the weights are renormalised)."""
    L = _box(walls=walls)
    x = np.array([1.7, 5.2, 6.4]) if walls else np.array([5.3, 6.1, 4.8])
    v = np.array([0.02, -0.01, 0.005]) * DX / DT
    R = 0.3 * DX
    g = np.concatenate([(x - 0.5) * DX, v, [R]])[None]
    O.ibm_spread(L, g, [0, 0, 0], DX, DT)
    Vp = 4.0 / 3.0 * math.pi * (R / DX) ** 3
    assert L.volfrac.sum() == pytest.approx(Vp, rel=1e-13)
    mom = (L.volfrac[..., None] * L.velsolid).reshape(-1, 3).sum(0)
    assert np.allclose(mom, Vp * v * DT / DX, rtol=1e-12, atol=0)
    if walls:
        assert L.volfrac[0].sum() == 0.0


def test_ibm_spread_clamps_at_one():
    L = _box()
    g = np.array([[5.5 * DX, 5.5 * DX, 5.5 * DX, 0, 0, 0, 0.9 * DX]])  # V_p = 3.05 lu^3
    O.ibm_spread(L, g, [0, 0, 0], DX, DT)
    assert L.volfrac.max() == 0.99


def test_ibm_drag_galilean_invariance_and_dilute_limit():
    """A particle moving with a uniform flow feels no drag (the kernel average of a constant
    field is that constant), and a tiny particle in fluid at rest feels the Schiller-Naumann drag
    -3 pi mu d Cd0(Re) v up to the dilute corrections 0.48 eps^(1/3) + O(eps)."""
    L = _box()
    nu = (1.0 / L.omega0 - 0.5) / 3.0 * DX * DX / DT
    mu = RHO * nu
    R = 0.05 * DX
    x = np.array([5.2, 6.7, 4.1]) * DX
    u_lu = np.array([0.005, 0.001, -0.003])
    g = np.concatenate([x, u_lu * DX / DT, [R]])[None]
    O.ibm_spread(L, g, [0, 0, 0], DX, DT)
    L.vel[...] = u_lu
    assert np.abs(O.ibm_drag(L, g, [0, 0, 0], DX, DT, RHO)).max() < 1e-18
    v = np.array([0.01, -0.02, 0.004]) * DX / DT
    g = np.concatenate([x, v, [R]])[None]
    L.vel[...] = 0.0
    F = O.ibm_drag(L, g, [0, 0, 0], DX, DT, RHO)[0]
    Re = RHO * 2 * R * np.linalg.norm(v) / mu
    stokes = -3 * math.pi * mu * 2 * R * (1 + 0.15 * Re ** 0.687) * v
    assert np.allclose(np.cross(F, v), 0.0, atol=1e-12 * np.linalg.norm(F) * np.linalg.norm(v))
    assert np.allclose(F, stokes, rtol=0.02, atol=0)


def test_ibm_weight_dilute_limit():
    """eps -> 0 with no slip: W -> 3 pi (d/dx) nu_lu, d the equivalent-sphere diameter."""
    L = _box(n=4)
    eps = 1e-9
    L.volfrac[...] = eps
    O.ibm_weight(L, DX, DT, RHO)
    nu_lu = (1.0 / L.omega0 - 0.5) / 3.0
    d = 2.0 * (3.0 * eps / (4.0 * math.pi)) ** (1.0 / 3.0)
    assert L.weight[1, 1, 1] == pytest.approx(3 * math.pi * d * nu_lu, rel=1e-3)  # + 0.48 eps^(1/3)


def test_vans_spread_conserves_volume_and_force():
    L = _box(kind="vans", walls=True)
    F = np.array([1e-6, -2e-6, 3e-7])
    R = 0.3 * DX
    g = np.concatenate([np.array([1.6, 5.2, 6.4]) * DX, [R], F])[None]
    O.vans_spread(L, g, [0, 0, 0], DX, DT, RHO, with_force=True)
    Vp = 4.0 / 3.0 * math.pi * (R / DX) ** 3
    assert L.volfrac.sum() == pytest.approx(Vp, rel=1e-13)
    assert np.allclose(L.hydroforce.reshape(-1, 3).sum(0), F * DT * DT / (RHO * DX ** 4), rtol=1e-12, atol=0)
    assert L.volfrac[0].sum() == 0.0


def test_vans_drag_stokes_limit():
    """No solid fraction in the box: the reference's fluid velocity sum runs over cells with
    eps > 0 only, so it is zero and the drag is the Schiller-Naumann drag on the particle
    velocity, -3 pi mu d Cd0(Re) v."""
    L = _box(kind="vans")
    nu = (1.0 / L.omega0 - 0.5) / 3.0 * DX * DX / DT
    mu = RHO * nu
    R = 0.2 * DX
    v = np.array([0.01, -0.02, 0.004]) * DX / DT
    g = np.concatenate([np.array([5.2, 6.7, 4.1]) * DX, v, [R]])[None]
    L.vel[...] = np.array([0.005, 0.001, -0.003])
    F = O.vans_drag(L, g, [0, 0, 0], DX, DT, RHO)[0]
    Re = RHO * 2 * R * np.linalg.norm(v) / mu
    assert np.allclose(F, -3 * math.pi * 2 * R * mu * (1 + 0.15 * Re ** 0.687) * v, rtol=1e-12, atol=0)
