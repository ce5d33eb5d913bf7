"""Precision of the storage types on the Taylor-Green vortex (Re 1600, u0 = 0.05 lu, periodic,
lagged Smagorinsky Cs = 0.1, the reference's collision, lbm_solver3d.py:176-192, 460-513).

The FP64-storage path (FP64 collision) is pinned to the FP64 oracle on a 32^3 vortex; it is then
the reference for the 256^3 vortex, where FP32 storage and the shifted 16-bit storages
(f - w rho0 in FP16 / BF16, FP32 collision) are compared after 1,000 steps by the relative L2
error of the velocity field.  Error bounds written here:
  FP32 <= 1e-4, FP16 <= 1e-3 (the 1e-3 target holds for FP16 through 1,000 steps of this
  vortex), BF16 <= 1e-2: BF16's 8-bit significand does not reach 1e-3, so no accuracy target
  is claimed for it beyond this bound."""
import numpy as np
import pytest
import torch

from oracle import oracle as O

pytestmark = pytest.mark.gpu

U0, RE, CS = 0.05, 1600.0, 0.1


def tgv_fields(n):
    k = 2 * np.pi / n
    x = np.arange(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    v = np.zeros((n, n, n, 3), np.float32)
    v[..., 0] = U0 * np.sin(k * X) * np.cos(k * Y) * np.cos(k * Z)
    v[..., 1] = -U0 * np.cos(k * X) * np.sin(k * Y) * np.cos(k * Z)
    rho = (1.0 + 3.0 * U0 * U0 / 16.0 * (np.cos(2 * k * X) + np.cos(2 * k * Y)) * (np.cos(2 * k * Z) + 2))
    return v, rho.astype(np.float32), 3.0 * U0 / (k * RE) + 0.5


@pytest.fixture
def smag():
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    old = BasicLattice3D.SmagorinskyConstant
    BasicLattice3D.SmagorinskyConstant = CS
    yield
    BasicLattice3D.SmagorinskyConstant = old


def _lattice(n, storage):
    from src.lbm3d.lbm_solver3d import BasicLattice3D
    v, rho, tau = tgv_fields(n)
    L = BasicLattice3D(n, n, n, 1.0 / tau, 1.0, 1.0, 1.0, storage=storage, smagorinsky="lagged")
    L.vel.from_numpy(v)
    L.rho.from_numpy(rho)
    L.initialize()
    return L


def test_tgv_f64_storage_matches_oracle(smag):
    n, steps = 32, 200
    v, rho, tau = tgv_fields(n)
    orc = O.OracleLattice(n, n, n, 1.0 / tau, kind="basic", cs=CS)
    orc.vel[...] = v
    orc.rho[...] = rho
    orc.initialize()
    orc.step(steps)
    L = _lattice(n, "f64")
    L.step(steps)
    f = L.f.to_numpy()
    assert np.abs(f - orc.f).max() / np.abs(orc.f).max() < 1e-12
    assert np.abs(L.vel.to_numpy() - orc.vel).max() / np.abs(orc.vel).max() < 1e-6


def test_tgv256_storage_errors_after_1000_steps(smag):
    n, steps = 256, 1000
    L = _lattice(n, "f64")
    L.step(steps)
    L._refresh_macros()   # rho, vel are derived from the populations when read
    ref = L.vel.dev.to(torch.float64).clone()
    ke_ref = L.get_kinetic_energy()
    del L
    torch.cuda.empty_cache()
    bounds = {"f32": 1e-4, "f16": 1e-3, "bf16": 1e-2}
    errs = {}
    for storage, bound in bounds.items():
        L = _lattice(n, storage)
        L.step(steps)
        L._refresh_macros()
        u = L.vel.dev.to(torch.float64)
        errs[storage] = float(torch.linalg.vector_norm(u - ref) / torch.linalg.vector_norm(ref))
        ke = L.get_kinetic_energy()
        print(f"{storage}: relative L2 velocity error {errs[storage]:.3e}, KE {ke / ke_ref - 1:+.3e}", flush=True)
        assert errs[storage] < bound, (storage, errs[storage])
        assert abs(ke / ke_ref - 1) < 10 * bound
        del L, u
        torch.cuda.empty_cache()
    assert errs["f32"] < errs["f16"] < errs["bf16"]
