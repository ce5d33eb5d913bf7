"""lbdem_lbm_reduce_static (the allocation-free variant of lbdem_lbm_reduce, whose scratch is a
file-scope device array) called through the C ABI against lbdem_lbm_reduce with a caller-owned
work buffer: identical FP64 sums (same kernels, same fixed summation order) on two calls in a
row, the second on a smaller lattice, for three storage types."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _lattice(shape, storage, seed):
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    from src.lbm3d.lbmutils import CellType
    rng = np.random.default_rng(seed)
    L = BasicLattice3D(*shape, 1.0 / 0.6, 1.0, 1.0, 1.0, storage=storage, smagorinsky="fixed")
    ct = np.full(shape, CellType.FLUID, dtype=np.int32)
    ct[0] = ct[-1] = CellType.OBSTACLE          # inactive cells the sums must skip
    L.CT.from_numpy(ct)
    L.vel.from_numpy(rng.uniform(-0.05, 0.05, shape + (3,)).astype(np.float32))
    L.rho.from_numpy((1.0 + 0.01 * rng.standard_normal(shape)).astype(np.float32))
    L.initialize()
    L.step(3)
    return L


def _both(L):
    import torch
    from src import _hip
    lib = _hip.lib()
    L._prepare()
    nwork = int(lib.lbdem_reduce_work_doubles(ctypes.byref(L._lat_c)))
    work = torch.empty(nwork, dtype=torch.float64, device=L.device)
    ref = torch.empty(5, dtype=torch.float64, device=L.device)
    got = torch.full((5,), float("nan"), dtype=torch.float64, device=L.device)
    _hip.check(lib.lbdem_lbm_reduce(L._lat_c, L._args_c(), _hip.ptr(work), _hip.ptr(ref), _hip.stream()),
               "lbdem_lbm_reduce")
    _hip.check(lib.lbdem_lbm_reduce_static(L._lat_c, L._args_c(), _hip.ptr(got), _hip.stream()),
               "lbdem_lbm_reduce_static")
    torch.cuda.synchronize()
    return ref.cpu().numpy(), got.cpu().numpy()


@pytest.mark.parametrize("storage", ["f32", "f64", "bf16"])
def test_reduce_static_matches_reduce_on_two_calls(storage):
    big = _lattice((40, 36, 64), storage, 1)
    r1, s1 = _both(big)
    assert np.isfinite(s1).all() and np.array_equal(r1, s1), (r1, s1)
    small = _lattice((9, 7, 10), storage, 2)     # second call, smaller lattice: fewer partials
    r2, s2 = _both(small)
    assert np.isfinite(s2).all() and np.array_equal(r2, s2), (r2, s2)
    assert not np.array_equal(s1, s2)
    assert s2[0] == pytest.approx(small.get_total_mass(), rel=1e-15)
