"""The ctypes binding shown in the integration guide, run as written, against BasicLattice3D."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT, PKG

pytestmark = pytest.mark.gpu


def _stub():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```python\n(import ctypes, math, torch\n.*?)```", text, flags=re.S)
    assert m, "the ctypes stub is missing from the integration guide"
    return m.group(1)


def test_ctypes_stub_matches_solver_class():
    import torch
    ns = {"LIB_PATH": os.path.join(PKG, "liblbdem_hip.so")}
    exec(compile(_stub(), "INTEGRATION.md", "exec"), ns)
    torch.cuda.synchronize()
    vel_stub = ns["vel"].cpu().numpy().reshape(32, 32, 32, 3)

    from src.lbm3d.lbm_solver3d import BasicLattice3D
    old = BasicLattice3D.SmagorinskyConstant
    try:
        BasicLattice3D.SmagorinskyConstant = 0.17
        L = BasicLattice3D(32, 32, 32, 1 / 0.6, 1.0, 1.0, 1.0, storage="f32", smagorinsky="lagged")
        v = np.zeros((32, 32, 32, 3), np.float32)
        v[..., 0] = (0.02 * np.sin(2 * np.pi * np.arange(32, dtype=np.float32) / 32)).astype(np.float32)
        L.vel.from_numpy(v)
        L.initialize()
        L.step(100)
        vel = L.vel.to_numpy()
    finally:
        BasicLattice3D.SmagorinskyConstant = old
    assert np.abs(vel_stub).max() > 1e-3
    assert np.allclose(vel_stub, vel, rtol=0, atol=1e-7)
