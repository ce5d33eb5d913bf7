"""The oracle's hybrid-coupling functions (HybridLattice3D) against what the mathematics and the
already pinned PSC / EqIBM restatements fix: the coarse mapping is the PSC mapping of the coarse
grains; the fine spreading conserves volume and momentum with the free-fraction correction and
never enters fully covered cells; the four-way collision is the superposition of the two
single-solid weighted collisions and moves momentum by beta rho (u_f - u) + B rho (u_c - u);
the fine drag tends to Schiller-Naumann at vanishing solid fraction.  CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O

DX, DT, RHO = 1e-3, 5e-4, 1000.0
OMEGA = 1.0 / 0.62


def _grain(x, v=(0, 0, 0), w=(0, 0, 0), R=0.3):
    """[x, v, omega, R] row, x in lattice units (cell centres at integers), R in dx."""
    return np.concatenate([(np.asarray(x, float) - 0.5) * DX, np.asarray(v, float) * DX / DT,
                           np.asarray(w, float) / DT, [R * DX]])


def _lattice(n=(14, 12, 13), walls=True):
    L = O.OracleLattice(*n, OMEGA, kind="hybrid")
    if walls:
        L.CT[0] = O.OBSTACLE
        L.CT[-1] = O.OBSTACLE
    return L


def _g7(g):
    return np.ascontiguousarray(np.concatenate([g[:, 0:6], g[:, 9:10]], axis=1))


def test_coarse_map_is_the_psc_map_of_the_coarse_grains():
    coarse = np.array([_grain([6.3, 5.1, 6.7], v=(0.01, 0, -0.02), w=(0, 0.01, 0.02), R=3.2),
                       _grain([8.9, 7.0, 5.2], v=(0, 0.01, 0), R=2.4)])
    fine = np.array([_grain([3.2, 3.3, 9.1], R=0.4), _grain([7.7, 6.2, 6.0], R=0.45)])
    L = _lattice()
    O.hybrid_map_coarse(L, np.concatenate([coarse[:1], fine[:1], coarse[1:], fine[1:]]), [0, 0, 0], DX, DT)
    P = _lattice()
    O.psc_map(P, coarse, [0, 0, 0], DX, DT)
    ids = np.where(P.id >= 0, np.array([0, 2])[np.maximum(P.id, 0)], -1)  # grain order interleaved above
    assert (L.id == ids).all()
    np.testing.assert_array_equal(L.volfrac, P.volfrac)
    # the two references group the rotation term differently (cross(w, r dx) vs w r dx terms)
    np.testing.assert_allclose(L.velsolid, P.velsolid, rtol=1e-14, atol=1e-18)
    assert (L.volfrac == 1).any() and ((L.volfrac > 0) & (L.volfrac < 1)).any()


def test_coarse_map_skips_ladd_inlet_cells():
    L = _lattice()
    L.CT[:, :, 0] = O.VEL_INLET_LADD
    O.hybrid_map_coarse(L, _grain([6.0, 6.0, 1.5], R=3.0)[None], [0, 0, 0], DX, DT)
    assert (L.id[:, :, 0] == -1).all() and (L.volfrac[:, :, 1] > 0).any()


def _mixed():
    coarse = _grain([6.2, 5.7, 6.4], v=(0.01, 0.0, 0.0), R=3.1)
    rng = np.random.default_rng(7)
    fine = [_grain(rng.uniform([1.4, 1.0, 1.0], [12.6, 11.0, 12.0]), v=rng.normal(0, 0.01, 3), R=0.3)
            for _ in range(40)]
    fine.append(_grain([6.2, 5.7, 2.9], v=(0, 0, 0.02), R=0.35))  # at the coarse grain's surface
    fine.append(_grain([1.2, 6.0, 6.0], v=(0.01, 0, 0), R=0.3))   # next to the x = 0 wall (mirror)
    return np.array([coarse] + fine)


def test_fine_spread_conserves_volume_and_momentum_in_the_free_room():
    g = _mixed()
    L = _lattice()
    O.hybrid_map_coarse(L, g, [0, 0, 0], DX, DT)
    velsum, wsum, nwarn = O.hybrid_map_fine(L, _g7(g), [0, 0, 0], DX, DT)
    fine = 2 * g[:, 9] <= DX
    Vp = 4.0 / 3.0 * math.pi * (g[fine, 9] / DX) ** 3
    v = g[fine, 3:6] * DT / DX
    assert nwarn == 0
    assert L.fine_volfrac.sum() == pytest.approx(Vp.sum(), rel=1e-12)
    np.testing.assert_allclose(velsum.reshape(-1, 3).sum(0), (v * Vp[:, None]).sum(0), rtol=1e-11, atol=1e-15)
    assert (L.fine_volfrac[L.volfrac >= 1] == 0).all()
    assert (L.fine_volfrac <= 1 - L.volfrac + 1e-15).all()
    assert (L.fine_volfrac[L.CT == O.OBSTACLE] == 0).all()
    # each grain's normalised weights sum to one over the cells it reaches
    assert wsum.sum() == pytest.approx(fine.sum(), rel=1e-12)


def test_fine_spread_without_coarse_grains_is_the_eqibm_spread():
    g = _mixed()[1:]
    L = _lattice()
    O.hybrid_map_coarse(L, g, [0, 0, 0], DX, DT)
    assert (L.volfrac == 0).all()
    O.hybrid_map_fine(L, _g7(g), [0, 0, 0], DX, DT)
    E = O.OracleLattice(*L.CT.shape, OMEGA, kind="eqibm")
    E.CT[...] = L.CT
    O.ibm_spread(E, _g7(g), [0, 0, 0], DX, DT)
    np.testing.assert_allclose(L.fine_volfrac, E.volfrac, rtol=1e-13, atol=1e-18)
    cov = L.fine_volfrac > 1e-10
    np.testing.assert_allclose(L.fine_velsolid[cov], E.velsolid[cov], rtol=1e-12, atol=1e-16)


def test_fine_overflow_is_clamped_to_the_free_room():
    L = _lattice()
    g = np.array([_grain([6.0, 6.0, 6.0], v=(0.01, 0, 0), R=0.5)] * 8)
    velsum, _, nwarn = O.hybrid_map_fine(L, _g7(g), [0, 0, 0], DX, DT)
    c = (6, 6, 6)
    assert nwarn >= 1
    assert L.fine_volfrac[c] == pytest.approx(0.99, abs=1e-15)
    assert L.fine_velsolid[c][0] == pytest.approx(velsum[c][0] / 0.99, rel=1e-14)
    # fine_velsolid keeps its value where the fraction is 1e-10 or less
    L.fine_velsolid[0, 0, 0] = 7.0
    O.hybrid_map_fine(L, _g7(g), [0, 0, 0], DX, DT)
    assert L.fine_velsolid[0, 0, 0][0] == 7.0


def _stepped_state(kind_cs=0.0):
    """A lattice after a few steps of a sheared flow, with solid fields set by hand."""
    L = _lattice(walls=False)
    L.cs = kind_cs
    n = L.CT.shape
    z = np.arange(n[2])
    L.vel[..., 0] = 0.03 * np.sin(2 * np.pi * z / n[2])[None, None, :]
    L.vel[..., 1] = 0.01 * np.cos(2 * np.pi * np.arange(n[0]) / n[0])[:, None, None]
    L.initialize()
    for _ in range(3):
        L.kind = "basic"
        L.collide()
        L.stream()
    L.kind = "hybrid"
    rng = np.random.default_rng(3)
    sh = n
    cover = rng.uniform(size=sh)
    L.volfrac[...] = np.where(cover < 0.3, np.round(rng.uniform(0.05, 1.0, sh) * 125) / 125, 0.0)
    L.velsolid[...] = rng.normal(0, 0.02, sh + (3,))
    L.fine_volfrac[...] = np.where(rng.uniform(size=sh) < 0.4, rng.uniform(0.01, 0.3, sh), 0.0)
    L.fine_velsolid[...] = rng.normal(0, 0.02, sh + (3,))
    L.fine_weight[...] = rng.uniform(0.0, 1.4, sh)  # above 1 in places: clipped only with a coarse grain
    return L


def _copy(L, kind):
    M = O.OracleLattice(*L.CT.shape, OMEGA, kind=kind, cs=L.cs)
    for k in ("CT", "f", "feq", "rho", "vel", "omega", "rate"):
        getattr(M, k)[...] = getattr(L, k)
    return M


def _weighted(L, volfrac, weight, velsolid):
    M = _copy(L, "eqibm")
    M.volfrac[...] = volfrac
    M.weight[...] = weight
    M.velsolid[...] = velsolid
    M.collide()
    return M


def test_hybrid_collision_is_the_superposition_of_the_single_solid_collisions():
    """f + beta Os1 + B Os2 + (1 - B - beta) Of = g(beta, u_f) + g(B, u_c) - g(0), g being the
    weighted collision f + x Os(u_s) + (1 - x) Of of PSC / EqIBM (collide_weighted); B from the
    step's omega, beta clipped to 1 - B only where both solids are present."""
    L = _stepped_state()
    M = _copy(L, "hybrid")
    for k in ("volfrac", "velsolid", "fine_volfrac", "fine_velsolid", "fine_weight"):
        setattr(M, k, getattr(L, k).copy())
    M.collide()
    om = M.omega
    eps = L.volfrac
    tm = 1.0 / om - 0.5
    B = np.where(eps > 0, eps * tm / ((1 - eps) + tm), 0.0)
    np.testing.assert_allclose(M.weight[eps > 0], B[eps > 0], rtol=1e-15)
    has_c, has_f = eps > 0, L.fine_volfrac > 0
    beta = np.where(has_f, L.fine_weight, 0.0)
    beta = np.where(has_c & has_f & (B + beta > 1), 1 - B, beta)
    g0 = _weighted(L, np.zeros_like(eps), 0.0, 0.0)
    gc = _weighted(L, has_c * 1.0, B, L.velsolid)
    gf = _weighted(L, has_f * 1.0, beta, L.fine_velsolid)
    np.testing.assert_allclose(om, g0.omega, rtol=0, atol=0)
    expect = gc.fpc + gf.fpc - g0.fpc
    np.testing.assert_allclose(M.fpc, expect, rtol=0, atol=2e-15)
    assert ((B + L.fine_weight > 1) & has_c & has_f).any() and ((L.fine_weight > 1) & ~has_c & has_f).any()


def test_hybrid_collision_moves_momentum_toward_both_solids():
    """Per cell, sum_q (f* - f) = 0 and sum_q c_q (f* - f) = beta rho (u_f - u) + B rho (u_c - u),
    hydroforce = -B rho (u_c - u) (u = j / rho of the populations)."""
    L = _stepped_state(kind_cs=0.1)
    L.collide()
    act = np.ones(L.CT.shape, bool)
    d = L.fpc - L.f
    np.testing.assert_allclose(d.sum(-1), 0.0, atol=4e-15)
    rho = L.f.sum(-1)
    u = (L.f @ O.C) / rho[..., None]
    np.testing.assert_allclose(u, L.vel, atol=1e-15)
    eps = L.volfrac
    tm = 1.0 / L.omega - 0.5
    B = np.where(eps > 0, eps * tm / ((1 - eps) + tm), 0.0)
    has_c, has_f = eps > 0, L.fine_volfrac > 0
    beta = np.where(has_f, L.fine_weight, 0.0)
    beta = np.where(has_c & has_f & (B + beta > 1), 1 - B, beta)
    dj = d @ O.C
    expect = (beta[..., None] * rho[..., None] * (L.fine_velsolid - u) * has_f[..., None]
              + B[..., None] * rho[..., None] * (L.velsolid - u) * has_c[..., None])
    np.testing.assert_allclose(dj[act], expect[act], atol=4e-15)
    np.testing.assert_allclose(L.hydroforce, -B[..., None] * rho[..., None] * (L.velsolid - u) * has_c[..., None],
                               atol=4e-15)


def _node_mean_solid(L, x):
    """Mean of fine + coarse fraction over the nodes of positive kernel weight around x."""
    nx, ny, nz = L.CT.shape
    lo = [max(0, int(c - 1.5)) for c in x]
    hi = [min(n, int(c + 2.5)) for c, n in zip(x, (nx, ny, nz))]
    vals = [L.fine_volfrac[i, j, k] + L.volfrac[i, j, k]
            for i in range(lo[0], hi[0]) for j in range(lo[1], hi[1]) for k in range(lo[2], hi[2])
            if O.kernel_mirror(nx, ny, nz, *x, i, j, k) > 0]
    return float(np.mean(vals)), len(vals)


def test_fine_drag_interpolates_the_fluid_and_tends_to_schiller_naumann():
    """Uniform fluid U, one small fine grain with velocity V: the kernel interpolation returns U,
    the solid fraction is the mean over the nodes of positive weight, and the force is the
    Tenneti drag -3 pi d mu (1 - eps) Cd (V - U), within 2 % of Schiller-Naumann at eps ~ 1e-5."""
    L = _lattice(walls=False)
    U = np.array([0.012, -0.004, 0.007])
    L.vel[...] = U
    V = np.array([-0.01, 0.02, 0.0])
    g = np.array([_grain([6.3, 5.6, 6.8], v=V, R=0.05)])
    O.hybrid_map_coarse(L, g, [0, 0, 0], DX, DT)
    O.hybrid_map_fine(L, _g7(g), [0, 0, 0], DX, DT)
    out = O.hybrid_forces(L, g, [0, 0, 0], DX, DT, RHO)
    mu = RHO * (1 / OMEGA - 0.5) / 3 * DX * DX / DT
    slip = (V - U) * DX / DT
    dp = 2 * g[0, 9]
    Re = RHO * dp * np.linalg.norm(slip) / mu
    stokes = -3 * np.pi * dp * mu * (1 + 0.15 * Re ** 0.687) * slip
    ep, nodes = _node_mean_solid(L, [6.3, 5.6, 6.8])
    assert 7 <= nodes < 64 and 0 < ep < 1e-4
    Re_e = (1 - ep) * Re
    tenneti = -3 * np.pi * dp * mu * (1 - ep) * O.tenneti_cd(Re_e, ep) * slip
    np.testing.assert_allclose(out[0, :3], tenneti, rtol=1e-12)
    np.testing.assert_allclose(out[0, :3], stokes, rtol=2e-2)
    assert (out[0, 3:] == 0).all()


def test_fine_drag_counts_the_coarse_fraction():
    """Next to a coarse grain the Tenneti correction sees the total solid fraction (larger drag)."""
    U = np.array([0.0, 0.0, 0.01])
    V = np.zeros(3)
    far = _lattice(walls=False)
    near = _lattice(walls=False)
    for L in (far, near):
        L.vel[...] = U
    fine = _grain([6.0, 6.0, 2.6], v=V, R=0.3)
    coarse = _grain([6.0, 6.0, 6.6], R=3.4)
    gn = np.array([fine, coarse])
    gf = np.array([fine])
    for L, g in ((far, gf), (near, gn)):
        O.hybrid_map_coarse(L, g, [0, 0, 0], DX, DT)
        O.hybrid_map_fine(L, _g7(g), [0, 0, 0], DX, DT)
    Ff = O.hybrid_forces(far, gf, [0, 0, 0], DX, DT, RHO)[0, :3]
    Fn = O.hybrid_forces(near, gn, [0, 0, 0], DX, DT, RHO)[0, :3]
    assert (near.volfrac[5:8, 5:8, 2:5] > 0).any()
    assert Fn[2] > Ff[2] > 0


def test_closed_box_conserves_mass():
    L = O.OracleLattice(12, 11, 12, OMEGA, kind="hybrid")
    O.cavity(L, (0.0, 0.0, 0.0))
    L.CT[L.CT == O.VEL_LADD] = O.OBSTACLE
    g = np.array([_grain([5.8, 5.2, 6.1], v=(0.01, 0, 0), w=(0, 0, 0.02), R=2.8),
                  _grain([2.7, 3.1, 3.3], v=(0, 0.02, 0), R=0.4), _grain([8.4, 8.1, 8.9], v=(0, 0, -0.01), R=0.3)])
    H = O.HybridOracle(L, g, [0, 0, 0], DX, DT, RHO)
    H.initialize_complete()
    m0 = L.sums()[0]
    for _ in range(6):
        F = H.step()
    assert L.sums()[0] == pytest.approx(m0, rel=1e-13)
    assert np.isfinite(F).all() and np.abs(F[0, :3]).max() > 0 and np.abs(F[1, :3]).max() > 0
    assert (F[1:, 3:] == 0).all()


def test_fine_weights_are_the_eqibm_weights_and_respect_the_coarse_room():
    """Without coarse grains the fine Tenneti weight is the (pinned) EqIBM weight; where fine and
    coarse fractions leave no fluid the weight keeps its previous value."""
    g = _mixed()[1:]
    L = _lattice()
    L.vel[...] = np.array([0.01, -0.005, 0.002])
    O.hybrid_map_coarse(L, g, [0, 0, 0], DX, DT)
    O.hybrid_map_fine(L, _g7(g), [0, 0, 0], DX, DT)
    O.hybrid_fine_weights(L, DX, DT, RHO)
    E = O.OracleLattice(*L.CT.shape, OMEGA, kind="eqibm")
    E.CT[...] = L.CT
    E.volfrac[...] = L.fine_volfrac
    E.velsolid[...] = L.fine_velsolid
    E.vel[...] = L.vel
    O.ibm_weight(E, DX, DT, RHO)
    cov = L.fine_volfrac > 0
    assert cov.sum() > 100
    np.testing.assert_allclose(L.fine_weight[cov], E.weight[cov], rtol=1e-13)
    # a cell whose fine and coarse fractions fill it keeps the weight it held
    c = tuple(np.argwhere(cov)[0])
    L.volfrac[c] = 1.0 - L.fine_volfrac[c]
    L.fine_weight[c] = 123.0
    O.hybrid_fine_weights(L, DX, DT, RHO)
    assert L.fine_weight[c] == 123.0
