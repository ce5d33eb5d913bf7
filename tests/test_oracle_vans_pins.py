"""Pins of the oracle's VANS restatement (Unresolvedlattice3D, src/unlbdem/vanslattice.py):

- the forcing term F_q (vanslattice.py:278-318) by its exact D3Q19 moments: zero mass, momentum
  (1 - omega/2) F, second moment (1 - omega/2)(u F + F u);
- the fluid-fraction equilibrium (:238-276) by its moments rho, rho (1 - eps) u and
  rho / 3 delta + rho (1 - eps) u u, and the effective-viscosity omega by its eps = 0 value;
- the Tenneti drag at finite solid fraction and Reynolds number (:563-651) by the force balance
  of the reference's recorded settling runs at terminal velocity
  (example/unlbdem/single_particle_settle/Re{31.9,4.1}_un/output.p4p, copied under
  tests/golden/vans_settle/);
- the coupled settling loop of one_particle.py:236-246 (spreading, drag, DEM substeps, force
  spreading, collide, stream) end to end against the same recordings.

The recordings were not made with the volume-fraction rule of the current reference source: at
terminal velocity they balance the drag evaluated with un-normalised kernel weights
(volfrac_c = w_c V_p / V_c, so that svf = V_p / V_c * sum_c w_c^2), not with the weights divided
by their sum (vanslattice.py:167-196).  With the normalised weights the recorded velocities
carry 8-12 % more drag than the current code gives.  The tests state both facts and pin the
drag law and the loop under the recordings' rule.  The Re 11.6 and Re 1.5 recordings are not
used: their frame times are not whole multiples of the time step that one_particle.py's
parameters give (281.6 and 1843.2 steps per frame), so those runs used other parameters.
CPU only."""
import math
import os

import numpy as np
import pytest

from oracle import oracle as O

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vans_settle")
C = O.C


def read_frames(case):
    """(frames, 7) rows t, x, y, z, vx, vy, vz of a recorded one-particle run."""
    rows = []
    with open(os.path.join(GOLD, f"Re{case}_output.p4p"), encoding="UTF-8") as fh:
        blocks = fh.read().split("TIMESTEP  PARTICLES")[1:]
    for b in blocks:
        tok = b.split()
        vals = [float(v) for v in tok[12:22]]
        rows.append([float(tok[0])] + vals[4:10])
    return np.array(rows)


def dem_config(case):
    """DEMSolverConfig of one_particle.py:131-176 (linear contacts, walls at the domain)."""
    from src.dem3d.demconfig import DEMSolverConfig, DomainBounds, LinearContactConfig
    _, _, dt, _, g = O.settle_setup(case)
    cm = LinearContactConfig(stiffness_normal=1e6, stiffness_tangential=1e6, damping_normal=0.1,
                             damping_tangential=0.1, pp_friction=0.2, pw_friction=0.2)
    cfg = DEMSolverConfig(domain=DomainBounds(0.0, 0.1, 0.0, 0.16, 0.0, 0.1), dt=dt / 10, gravity=g, contact_model=cm)
    cfg.set_particle_properties(elastic_modulus=1e8, poisson_ratio=0.3)
    cfg.set_wall_properties(elastic_modulus=1e8, poisson_ratio=0.3)
    return cfg


def kernel_sums(pos, shape):
    """sum_c w_c and sum_c w_c^2 of the radial three-point kernel over the fluid cells of the
    settling box around a particle at pos [m]."""
    L = O.OracleLattice(*shape, 1.0, kind="vans")
    O.settle_container(L.CT)
    xc = (np.asarray(pos) + 0.5 * O.SETTLE_DX) / O.SETTLE_DX
    s1 = s2 = 0.0
    for i in range(shape[0]):
        for j in range(shape[1]):
            for k in range(shape[2]):
                if L.CT[i, j, k] & O.OBSTACLE:
                    continue
                w = O.threedelta(math.dist(xc, (i, j, k)))
                s1 += w
                s2 += w * w
    return s1, s2


def unnormalised(L, pos):
    """The recordings' volume-fraction rule: the mapped fraction times the kernel-weight sum."""
    L.volfrac *= kernel_sums(pos, L.CT.shape)[0]


def _vans_random(seed=3, n=5):
    rng = np.random.default_rng(seed)
    L = O.OracleLattice(n, n, n, 1.0 / 0.7, kind="vans")
    sh = (n, n, n)
    L.rho[...] = 1.0 + 0.05 * rng.standard_normal(sh)
    L.vel[...] = rng.uniform(-0.06, 0.06, sh + (3,))
    L.volfrac[...] = rng.uniform(0.0, 0.5, sh)
    L.hydroforce[...] = rng.uniform(-1e-3, 1e-3, sh + (3,))
    L.f[...] = rng.uniform(0.01, 0.1, sh + (19,))
    L.collide()
    return L


def test_vans_forcing_term_moments():
    L = _vans_random()
    om = L.omega[..., None]
    Fq = L.fpc - (1.0 - om) * L.f - om * L.feq
    F, u = L.hydroforce, L.vel
    h = 1.0 - 0.5 * L.omega
    assert np.abs(Fq.sum(-1)).max() < 1e-15
    np.testing.assert_allclose(Fq @ C, h[..., None] * F, rtol=0, atol=1e-15)
    M2 = np.einsum("...q,qa,qb->...ab", Fq, C, C)
    want = h[..., None, None] * (u[..., :, None] * F[..., None, :] + F[..., :, None] * u[..., None, :])
    np.testing.assert_allclose(M2, want, rtol=0, atol=1e-15)


def test_vans_equilibrium_moments_and_clear_fluid_omega():
    L = _vans_random(seed=5)
    rho, u, fvf = L.rho, L.vel, 1.0 - L.volfrac
    np.testing.assert_allclose(L.feq.sum(-1), rho, rtol=1e-14)
    np.testing.assert_allclose(L.feq @ C, (rho * fvf)[..., None] * u, rtol=0, atol=1e-15)
    M2 = np.einsum("...q,qa,qb->...ab", L.feq, C, C)
    want = (rho / 3.0)[..., None, None] * np.eye(3) + (rho * fvf)[..., None, None] * u[..., :, None] * u[..., None, :]
    np.testing.assert_allclose(M2, want, rtol=0, atol=1e-15)
    L.volfrac[...] = 0.0
    L.collide()
    np.testing.assert_allclose(L.omega, 1.0 / 0.7, rtol=1e-15)


def _terminal_frames(case):
    """Frames whose particle centre is at least 1.5 lattice spacings (the kernel support) from
    the floor and the top wall, with dv/dt from central differences."""
    fr = read_frames(case)
    out = []
    for k in range(1, len(fr) - 1):
        if 0.045 < fr[k, 2] < 0.09:
            acc = (fr[k + 1, 5] - fr[k - 1, 5]) / (fr[k + 1, 0] - fr[k - 1, 0])
            out.append((fr[k], acc))
    return out


@pytest.mark.parametrize("case", ["31.9", "4.1"])
def test_recorded_terminal_velocity_balances_tenneti_drag(case):
    """m (g' + dv/dt) = 3 pi d mu (1 - svf) Cd(Re, svf) |v| at every mid-box frame, with the
    recordings' solid fraction (svf ~ 0.054, Re ~ 28 and ~ 3)."""
    mu, rho, _, shape, g = O.settle_setup(case)
    d, m = 2 * O.SETTLE_R, O.SETTLE_MASS
    vp_vc = 4.0 / 3.0 * math.pi * O.SETTLE_R ** 3 / O.SETTLE_DX ** 3
    frames = _terminal_frames(case)
    assert len(frames) >= 3
    for fr, acc in frames:
        s1, s2 = kernel_sums(fr[1:4], shape)
        v = abs(fr[5])
        balance = []
        for svf in (vp_vc * s2, vp_vc * s2 / s1):   # recordings' rule, current source's rule
            Re = (1.0 - svf) * rho * d * v / mu
            drag = 3.0 * math.pi * d * mu * (1.0 - svf) * O.tenneti_cd(Re, svf, 1e-8) * v
            balance.append(drag / (m * (-g[1] + acc)))
        assert balance[0] == pytest.approx(1.0, abs=5e-3), (fr[0], balance)
        assert balance[1] < 0.93, (fr[0], balance)   # the current rule leaves drag unbalanced


@pytest.mark.parametrize("case,t_end", [("31.9", 0.85), ("4.1", 1.45)])
def test_settling_loop_reproduces_recorded_terminal_velocity(case, t_end):
    """The restated one_particle.py loop under the recordings' volume-fraction rule reproduces
    the recorded velocity at the mid-box frames within 3e-3 (relative; the recorded velocity
    wanders by about that much between frames) and the distance fallen between them within
    5e-3 (the start next to the top wall differs by a few per cent in
    velocity, so the heights themselves carry an offset)."""
    frames = read_frames(case)
    run = O.VansSettleOracle(case, dem_config(case), after_grains2lattice=unnormalised)
    per_frame = int(round(frames[0, 0] / run.dt))
    assert abs(per_frame * run.dt - frames[0, 0]) < 1e-9 * frames[0, 0]
    checked, prev = 0, None
    for k in range(len(frames)):
        if frames[k, 0] > t_end:
            break
        for _ in range(per_frame):
            run.step()
        y, vy = run.dem.a["pos"][0, 1], run.dem.a["vel"][0, 1]
        if 0.045 < frames[k, 2] < 0.09:
            assert vy == pytest.approx(frames[k, 5], rel=3e-3), frames[k, 0]
            if prev is not None:   # distance fallen since the previous mid-box frame
                assert y - prev[0] == pytest.approx(frames[k, 2] - prev[1], rel=5e-3), frames[k, 0]
            prev = (y, frames[k, 2])
            checked += 1
    assert checked >= 3


@pytest.mark.parametrize("case,steps_per_frame", [("11.6", 281.6), ("1.5", 1843.2)])
def test_unused_recordings_have_other_time_steps(case, steps_per_frame):
    frames = read_frames(case)
    _, _, dt, _, _ = O.settle_setup(case)
    assert frames[0, 0] / dt == pytest.approx(steps_per_frame, abs=1e-6)
