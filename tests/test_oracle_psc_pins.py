"""Pins of the oracle's PSC restatement (PSCLattice3D, src/lbdem/psclattice3d.py) by what the
geometry and the symmetry fix, independently of the restated loops:

- grains2lattice (:84-182) samples every cell at the 5^3 points i - 0.4 + 0.2 a (a = 0..4), so
  over the lattice the samples form one cubic point lattice of spacing 1/5 and the counts add up
  to the number of its points inside the sphere of radius R - dx/2.  For a sphere centred on a
  node that is the integer count #{(a, b, c): a^2 + b^2 + c^2 < 25 r^2}, cell by cell;
- over uniformly random centre offsets the expected count of a shifted point lattice inside a
  ball is its volume over the cell volume, so the mean of sum(eps) converges to 4/3 pi r^3;
- lattice2grains (:316-358): in a box with the cubic symmetry of the grain, a spinning grain in
  fluid at rest feels a torque antiparallel to its spin and no force, and a translating grain a
  drag antiparallel to its velocity and no torque (signs and lever arm of the torque sum).
CPU only."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _grain(x, R, v=(0.0, 0.0, 0.0), w=(0.0, 0.0, 0.0)):
    return np.array([[*x, *v, *w, R]], dtype=np.float64)


def _lattice(n, walls=False, kind="psc"):
    L = O.OracleLattice(n, n, n, 1.0 / 0.6, kind=kind)
    if walls:
        L.CT[0], L.CT[-1], L.CT[:, 0], L.CT[:, -1], L.CT[:, :, 0], L.CT[:, :, -1] = (O.OBSTACLE,) * 6
    return L


@pytest.mark.parametrize("R,centre", [(3.37, 8), (6.37, 11), (8.21, 13), (12.44, 17)])
def test_node_centred_sphere_counts_are_lattice_point_counts(R, centre):
    """Grain centre on node `centre` (dx = 1, domain origin 0, so x = centre - 1/2)."""
    n = 2 * centre + 1
    L = _lattice(n)
    O.psc_map(L, _grain([centre - 0.5] * 3, R), [0, 0, 0], 1.0, 1.0)
    cnt = np.rint(L.volfrac * 125).astype(np.int64)
    r25 = 25.0 * (R - 0.5) ** 2
    # integer sub-sample coordinates of cell offset d: 5 d - 2 .. 5 d + 2
    a = np.arange(-2, 3)
    span = int(R) + 2
    for di in range(-span, span + 1):
        for dj in range(-span, span + 1):
            for dk in range(-span, span + 1):
                A, B, Cc = np.meshgrid(5 * di + a, 5 * dj + a, 5 * dk + a, indexing="ij")
                want = int((A * A + B * B + Cc * Cc < r25).sum())
                assert cnt[centre + di, centre + dj, centre + dk] == want, (di, dj, dk)
    # nothing outside the scanned block
    blk = cnt[centre - span:centre + span + 1, centre - span:centre + span + 1, centre - span:centre + span + 1]
    assert cnt.sum() == blk.sum()
    assert (L.id[cnt > 0] == 0).all() and (L.id[cnt == 0] == -1).all()


@pytest.mark.parametrize("R", [4.6, 8.0])
def test_mean_mapped_volume_over_random_offsets_is_the_sphere_volume(R):
    rng = np.random.default_rng(11)
    r = R - 0.5
    n = int(2 * R + 6)
    vols = []
    for _ in range(160):
        L = _lattice(n)
        x = n / 2 - 0.5 + rng.random(3)
        O.psc_map(L, _grain(x, R), [0, 0, 0], 1.0, 1.0)
        vols.append(L.volfrac.sum())
    vols = np.array(vols)
    want = 4.0 / 3.0 * math.pi * r ** 3
    se = vols.std(ddof=1) / math.sqrt(len(vols))
    assert abs(vols.mean() - want) < 4 * se + 1e-9 * want, (vols.mean(), want, se)
    # the sampling error of one mapping: within a few per cent of the volume
    assert np.abs(vols / want - 1).max() < 0.03


def _run_symmetric(n, R, v, w, steps=30):
    """Grain centred in a walled n^3 box (cubic symmetry about the grain centre).  25 (R - 1/2)^2
    must not be an integer: a sub-sample point exactly on the sphere would be classified by the
    rounding of its coordinates, which is not mirror symmetric."""
    L = _lattice(n, walls=True)
    c = (n - 1) / 2.0                     # lattice coordinate of the box centre
    g = _grain([c - 0.5] * 3, R, v, w)    # lattice coordinate = x + 1/2
    L.initialize()
    F = None
    for _ in range(steps):
        O.psc_map(L, g, [0, 0, 0], 1.0, 1.0)
        L.collide()
        F = O.psc_forces(L, g, [0, 0, 0], 1.0, 1.0, 1.0)
        L.stream()
    return F[0, :3], F[0, 3:]


@pytest.mark.parametrize("n", [20, 21])   # grain centred between nodes and on a node
def test_spinning_grain_feels_torque_against_its_spin_and_no_force(n):
    w = np.array([0.2, -0.35, 0.9]) * 1e-4   # small: the u^2 terms tilt T by O(u^2)
    F, T = _run_symmetric(n, 5.37, (0, 0, 0), w)
    tn = np.linalg.norm(T)
    assert tn > 0
    cos = T @ w / (tn * np.linalg.norm(w))
    assert cos == pytest.approx(-1.0, abs=1e-9)
    assert np.linalg.norm(F) < 1e-10 * tn / 5.37


@pytest.mark.parametrize("n", [20, 21])
def test_translating_grain_feels_drag_against_its_motion_and_no_torque(n):
    v = np.array([1e-3, 0.0, 0.0])
    F, T = _run_symmetric(n, 5.37, v, (0, 0, 0))
    fn = np.linalg.norm(F)
    assert F[0] < 0 and abs(F[1]) < 1e-10 * fn and abs(F[2]) < 1e-10 * fn
    assert np.linalg.norm(T) < 1e-10 * fn * 5.37
