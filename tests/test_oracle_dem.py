"""The DEM oracle pinned by known answers, closed forms and brute force (CPU only)."""
import numpy as np
import pytest

from dem_cases import make_config, oracle_solver, random_packing, read_p4p
from oracle import oracle as O


def test_morton_known_answers():
    """bpcd.py:154-179: bit interleave x | y << 1 | z << 2 of 10-bit coordinates."""
    assert O.morton3d32(1, 0, 0) == 1
    assert O.morton3d32(0, 1, 0) == 2
    assert O.morton3d32(0, 0, 1) == 4
    assert O.morton3d32(1023, 1023, 1023) == 2 ** 30 - 1
    assert O.morton3d32(2, 0, 0) == 8
    assert O.morton3d32(1024, 0, 0) == 0  # coordinates wrap at 10 bits
    assert O.morton3d32(-1, 0, 0) == O.morton3d32(1023, 0, 0)


def test_hash_table_size_rule():
    """next_pow2 applied twice, clamped to [2^20, 2^22] (bpcd.py:55-60)."""
    assert O.hash_table_size([0, 0, 0], [1, 1, 1], 0.5) == 1 << 20
    # 200 x 200 x 200 cells = 8e6 -> next_pow2 -> 2^23 -> clamp 2^22
    assert O.hash_table_size([0, 0, 0], [200, 200, 200], 1.0) == 1 << 22


@pytest.mark.parametrize("fixture", ["vf0.35.p4p", "shear_5.p4p", "sill_5.p4p", "random"])
@pytest.mark.parametrize("probe27", [False, True])
def test_broadphase_contains_every_touching_pair(fixture, probe27):
    """Candidate set (i < j, unique) contains every pair closer than 2 r_max (brute force): half an
    octant-probe cell (4 r_max) or one 27-probe cell (2 r_max); every contact (gap < 0) is such a pair."""
    if fixture == "random":  # overlapping, polydisperse points: many close pairs
        rng = np.random.default_rng(11)
        pos = rng.uniform(0.0, 1.0, (3000, 3))
        rad = rng.uniform(0.005, 0.02, 3000)
    else:
        a = read_p4p(fixture)
        rad, pos = a[:, 2], a[:, 4:7]
    dmin = pos.min(0) - rad.max()
    cell = 2.0 * rad.max() if probe27 else 4.0 * rad.max()
    T = O.hash_table_size(dmin, pos.max(0) + rad.max(), cell)
    pairs = O.broadphase(pos, dmin, cell, T, probe27=probe27)
    assert (pairs[:, 0] < pairs[:, 1]).all()
    assert len(np.unique(pairs, axis=0)) == len(pairs)
    d = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=-1)
    close = np.argwhere(np.triu(d < 2.0 * rad.max() * (1 - 1e-9), 1))
    if fixture == "random":
        assert len(close) > 100
    have = set(map(tuple, pairs.tolist()))
    missing = [tuple(p) for p in close.tolist() if tuple(p) not in have]
    assert not missing


def test_integrator_closed_form_under_gravity():
    """demsolver.py:580-589: v += (a_prev + a) dt / 2, x += v dt + a dt^2 / 2 with a_prev = 0 at start:
    x_n = x0 + n v0 dt + g dt^2 n (n + 1) / 2."""
    cfg = make_config(box=(-10, 10, -10, 10, -10, 10), dt=1e-3, g=(0.0, 0.0, -9.81))
    x0 = np.array([[0.0, 0.0, 0.0]])
    v0 = np.array([[0.1, 0.0, 0.5]])
    o = oracle_solver(cfg, [0.01], [1.0], x0, v0)
    n = 250
    o.step(n)
    dt, g = 1e-3, np.array([0, 0, -9.81])
    expect = x0[0] + n * v0[0] * dt + g * dt * dt * n * (n + 1) / 2
    assert np.allclose(o["pos"][0], expect, rtol=0, atol=1e-13)
    assert np.allclose(o["vel"][0], v0[0] + (n - 0.5) * g * dt, rtol=0, atol=1e-13)


def test_hertz_static_force():
    """Two spheres at rest overlapping by delta, e = 1 (no damping): the first step applies
    F_n = 4/3 Y* sqrt(R*) delta^(3/2) (hertz.py:67-82)."""
    E, nu, R, m, delta = 1e6, 0.3, 0.1, 2.0, 1e-3
    cfg = make_config(box=(-1, 1, -1, 1, -1, 1), dt=1e-6, g=(0, 0, 0), E=E, nu=nu, e=1.0, mu=0.0)
    pos = np.array([[0.0, 0.0, 0.0], [2 * R - delta, 0.0, 0.0]])
    o = oracle_solver(cfg, [R, R], [m, m], pos, None)
    o.step(1)
    Ys = 1.0 / (2 * (1 - nu * nu) / E)
    Rs = R / 2
    F = 4.0 / 3.0 * Ys * np.sqrt(Rs) * delta ** 1.5
    assert np.allclose(o["acc"][0], [-F / m, 0, 0], rtol=1e-12, atol=1e-12)
    assert np.allclose(o["acc"][1], [F / m, 0, 0], rtol=1e-12, atol=1e-12)


def test_linear_head_on_restitution():
    """Linear spring-dashpot with C_n = 2 zeta sqrt(k m*): e = exp(-pi zeta / sqrt(1 - zeta^2))."""
    zeta, k, m, R, v = 0.3, 1e4, 1.0, 0.1, 0.5
    cfg = make_config(model="linear", box=(-1, 1, -1, 1, -1, 1), dt=2e-6, g=(0, 0, 0), kn=2 * k, kt=2 * k,
                      zeta=zeta, mu=0.0)
    pos = np.array([[-R - 1e-4, 0.0, 0.0], [R + 1e-4, 0.0, 0.0]])
    vel = np.array([[v, 0.0, 0.0], [-v, 0.0, 0.0]])
    o = oracle_solver(cfg, [R, R], [m, m], pos, vel)
    o.step(40000)
    e = (o["vel"][1, 0] - o["vel"][0, 0]) / (2 * v)
    assert abs(e - np.exp(-np.pi * zeta / np.sqrt(1 - zeta * zeta))) < 5e-3


def test_hertz_elastic_collision_conserves_energy():
    # E = 1e8: the contact lasts ~4 ms, well inside the 20 ms simulated
    cfg = make_config(box=(-1, 1, -1, 1, -1, 1), dt=1e-6, g=(0, 0, 0), E=1e8, e=1.0, mu=0.0)
    R, m, v = 0.1, 1.0, 0.3
    pos = np.array([[-R - 1e-4, 0.0, 0.0], [R + 1e-4, 0.02, 0.0]])
    vel = np.array([[v, 0.0, 0.0], [-v, 0.0, 0.0]])
    o = oracle_solver(cfg, [R, R], [m, m], pos, vel)
    ke0 = 0.5 * m * (vel ** 2).sum()
    o.step(20000)
    ke1 = 0.5 * m * (o["vel"] ** 2).sum()
    assert abs(ke1 / ke0 - 1.0) < 2e-3
    assert np.allclose(np.linalg.norm(o["quat"], axis=1), 1.0, atol=1e-12)


def test_oracle_packing_settles_without_slot_overflow():
    rng = np.random.default_rng(3)
    cfg = make_config(box=(0, 1, 0, 1, 0, 1), dt=5e-5)
    rad, mass, pos, vel = random_packing(60, 0.06, (0, 1, 0, 1, 0, 1), rng)
    o = oracle_solver(cfg, rad, mass, pos, vel)
    o.step(300)
    assert o["status"][0] == 0
    assert np.isfinite(o["pos"]).all()


def test_oracle_step_grows_its_candidate_scratch():
    """A step whose candidates outnumber the scratch is repeated with a larger one: the result
    is that of a scratch large enough from the start (no candidate is dropped)."""
    rng = np.random.default_rng(4)
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    rad, mass, pos, vel = random_packing(150, 0.04, box, rng, v=2.0, polydisperse=0.3)
    cfg = make_config(model="hertz", box=box, dt=5e-5)
    a = oracle_solver(cfg, rad, mass, pos, vel)
    b = oracle_solver(cfg, rad, mass, pos, vel)
    b._pairs = np.zeros((8, 2), dtype=np.int32)
    a.step(60)
    b.step(60)
    assert len(b._pairs) > 8
    for k in ("pos", "vel", "omega", "cf_active", "cf_j"):
        assert np.array_equal(a[k], b[k]), k
