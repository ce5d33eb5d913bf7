"""Host-side list of the free-slip pass (lbdem_lbm_free_slip) for the FREE_SLIP branch of
BasicLattice3D.stream (src/lbm3d/lbm_solver3d.py:230-236).  CPU only."""
import numpy as np
import pytest

from oracle import oracle as O
from test_free_slip_gpu import _case


def test_free_slip_pairs_count_and_sources():
    """Host-side pair list of a plain channel: every active cell next to a wall plane has five
    reflected directions, and each source is a bounced population of the same row."""
    from src.lbm3d.lbm_solver3d import BasicLattice3D, free_slip_pairs
    ct, _, _ = _case("channel")
    act = (ct & O.SKIP_BASIC) == 0
    solid = ~act
    pairs = free_slip_pairs(ct, act, solid, np.zeros_like(act))
    nx, ny, nz = ct.shape
    assert len(pairs) == 2 * 5 * nx * nz
    tj = (pairs[:, 0] // nz) % ny
    sj = (pairs[:, 2] // nz) % ny
    assert np.array_equal(tj, sj) and set(np.unique(tj)) == {1, ny - 2}
    # the source direction points back into the wall: opposite of the mirrored target direction
    for t, q, s, qs in pairs[:20]:
        assert BasicLattice3D.c[qs][1] == BasicLattice3D.c[q][1] != 0


def test_free_slip_without_face_is_rejected():
    from src.lbm3d.lbm_solver3d import free_slip_pairs
    ct, _, _ = _case("channel")
    ct[:, 0, :] = O.FREE_SLIP
    act = (ct & O.SKIP_BASIC) == 0
    with pytest.raises(NotImplementedError, match="face flag"):
        free_slip_pairs(ct, act, ~act, np.zeros_like(act))


def test_edge_reflection_into_a_free_slip_cell_reads_zero():
    """At an edge of two free-slip planes the reference's reflection source is itself a
    free-slip cell, whose fpc is never written: the record carries source -1."""
    from src.lbm3d.lbm_solver3d import free_slip_pairs
    ct, _, _ = _case("edges")
    act = (ct & O.SKIP_BASIC) == 0
    pairs = free_slip_pairs(ct, act, ~act, np.zeros_like(act))
    assert (pairs[:, 2] == -1).any()
    assert (pairs[:, 2] >= 0).sum() > (pairs[:, 2] == -1).sum()
