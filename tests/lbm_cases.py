"""Shared set-ups for lattice parity tests: the same case built for the oracle and the GPU."""
import numpy as np

from oracle import oracle as O


def oracle_cavity(n, tau=0.6, lid=0.1, kind="basic", cs=0.17, real=np.float64):
    L = O.OracleLattice(n, n, n, 1.0 / tau, kind=kind, cs=cs, real=real)
    O.cavity(L, (lid, 0.0, 0.0), lid_axis=2)
    # the GPU fields are float32: give the oracle the same (rounded) wall velocity
    L.vel[...] = L.vel.astype(np.float32).astype(real)
    return L


def gpu_from_oracle(cls, orc, storage="f32", smagorinsky="lagged", **kw):
    """Build a GPU lattice with the oracle's cell types, rho and vel (dx = dt = rho = 1)."""
    if cls.__name__ == "ForcedLattice3D":
        L = cls(orc.nx, orc.ny, orc.nz, orc.omega0, 1.0, 1.0, 1.0, kw.pop("F"), storage=storage)
    else:
        L = cls(orc.nx, orc.ny, orc.nz, orc.omega0, 1.0, 1.0, 1.0, storage=storage, smagorinsky=smagorinsky, **kw)
    L.CT.from_numpy(orc.CT)
    L.vel.from_numpy(orc.vel.astype(np.float32))
    L.rho.from_numpy(orc.rho.astype(np.float32))
    return L


def rel_err(a, b, mask=None):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    if mask is not None:
        a, b = a[mask], b[mask]
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
