"""Pin of the oracle's free-slip branch of stream (src/lbm3d/lbm_solver3d.py:230-236, mirror maps
:50-57) by mirror symmetry: a channel between two free-slip planes y = 1/2 and y = W + 1/2 (cells
j = 0 and W + 1) evolves exactly as the first half of a periodic channel of 2W cells whose initial
field is mirror symmetric about those planes (the D3Q19 collide-stream commutes with the
reflection y -> -y, q -> qsyy[q]).  The periodic stream is pinned on its own by the other lattice
tests.  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O


def _fields(nx, ny_half, nz, jp):
    """Velocity and density at periodic rows jp (theta = pi (jp + 1/2) / W): u_x, u_z, rho even
    and u_y odd about theta = 0 and theta = pi."""
    i = np.arange(nx)[:, None, None]
    k = np.arange(nz)[None, None, :]
    th = np.pi * (np.asarray(jp)[None, :, None] + 0.5) / ny_half
    U = 0.04
    ux = U * np.cos(th) * (1.0 + 0.3 * np.sin(2 * np.pi * i / nx)) * np.ones_like(k)
    uy = 0.4 * U * np.sin(th) * np.sin(2 * np.pi * i / nx + 0.3) * np.ones_like(k)
    uz = 0.5 * U * np.cos(th) * np.cos(2 * np.pi * k / nz) * np.ones_like(i)
    rho = 1.0 + 0.01 * np.cos(th) * np.cos(2 * np.pi * i / nx) * np.ones_like(k)
    return np.stack([ux, uy, uz], axis=-1), rho


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_free_slip_channel_is_the_mirrored_periodic_channel(axis):
    W, n1, n2, steps = 7, 6, 5, 60
    tau = 0.56
    # periodic channel of 2W cells along y, built in (x, y, z) and permuted so that y -> axis
    vp, rp = _fields(n1, W, n2, np.arange(2 * W))
    vc, rc = _fields(n1, W, n2, np.arange(-1, W + 1))   # rows j = 0 .. W + 1 of the channel
    perm = {0: (1, 0, 2), 1: (0, 1, 2), 2: (0, 2, 1)}[axis]
    comp = list(perm)

    def place(v, r):
        return np.transpose(v, perm + (3,))[..., comp].copy(), np.transpose(r, perm).copy()

    vp, rp = place(vp, rp)
    vc, rc = place(vc, rc)
    per = O.OracleLattice(*rp.shape, 1.0 / tau, kind="basic", cs=0.1)
    per.vel[...] = vp
    per.rho[...] = rp
    ch = O.OracleLattice(*rc.shape, 1.0 / tau, kind="basic", cs=0.1)
    lo = [slice(None)] * 3
    hi = [slice(None)] * 3
    lo[axis], hi[axis] = 0, -1
    face_lo, face_hi = [(O.LEFT, O.RIGHT), (O.BOTTOM, O.TOP), (O.BACK, O.FRONT)][axis]
    ch.CT[tuple(lo)] = O.FREE_SLIP | face_lo
    ch.CT[tuple(hi)] = O.FREE_SLIP | face_hi
    ch.vel[...] = vc
    ch.rho[...] = rc
    per.initialize()
    ch.initialize()
    per.step(steps)
    ch.step(steps)
    inner = [slice(None)] * 3
    inner[axis] = slice(1, W + 1)
    half = [slice(None)] * 3
    half[axis] = slice(0, W)
    a, b = ch.f[tuple(inner)], per.f[tuple(half)]
    assert np.abs(a - b).max() < 1e-13 * np.abs(b).max()
    np.testing.assert_allclose(ch.vel[tuple(inner)], per.vel[tuple(half)], rtol=0, atol=1e-14)
    # the flow is not trivial: the walls reflect a non-zero normal velocity
    assert np.abs(per.vel[..., axis]).max() > 1e-3


def test_free_slip_wall_conserves_mass():
    W, steps = 6, 40
    ch = O.OracleLattice(5, W + 2, 4, 1.0 / 0.6, kind="basic", cs=0.1)
    ch.CT[:, 0, :] = O.FREE_SLIP | O.BOTTOM
    ch.CT[:, -1, :] = O.FREE_SLIP | O.TOP
    v, r = _fields(5, W, 4, np.arange(-1, W + 1))
    ch.vel[...] = v
    ch.rho[...] = r
    ch.initialize()
    m0 = ch.sums()[0]
    ch.step(steps)
    assert ch.sums()[0] == pytest.approx(m0, rel=1e-13)
