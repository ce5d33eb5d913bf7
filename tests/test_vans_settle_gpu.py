"""The reference's single-sphere VANS settling driver (example/unlbdem/single_particle_settle/
one_particle.py:44-246) on the HIP path, FP64 populations, against the oracle's restatement of the
same loop: 5 x 7 x 5 lattice, tau 0.5001, linear contacts, 10 DEM substeps per lattice step, the
particle read from one_particle.p4p.  The oracle loop itself is pinned to the reference's
recorded runs by tests/test_oracle_vans_pins.py."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from test_oracle_vans_pins import dem_config

pytestmark = pytest.mark.gpu

P4P = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "vans_settle", "one_particle.p4p")


@pytest.mark.parametrize("case,steps", [("31.9", 1500), ("4.1", 3000)])
def test_vans_settling_matches_oracle(case, steps):
    from src.dem3d.demsolver import DEMSolver
    from src.lbm3d.lbmutils import CellType
    from src.unlbdem.vanslattice import Unresolvedlattice3D
    mu, rho, dt, shape, g = O.settle_setup(case)
    dem = DEMSolver(dem_config(case))
    dem.init_particle_fields(P4P, [0.0, 0.0, 0.0], [0.1, 0.16, 0.1])
    dem.set_contact_model("linear")
    L = Unresolvedlattice3D(*shape, 1.0 / O.SETTLE_TAU, O.SETTLE_DX, dt, rho, dem, storage="f64")
    ct = np.full(shape, CellType.FLUID, dtype=np.int32)
    O.settle_container(ct)
    L.CT.from_numpy(ct)
    L.initialize()
    orc = O.VansSettleOracle(case, dem_config(case))
    assert dem.gf.position.to_numpy()[0] == pytest.approx(O.SETTLE_X0, abs=1e-15)
    checked = 0
    for s in range(1, steps + 1):
        L.grains2lattice()
        L.compute_hydrodynamic_force()
        for _ in range(10):
            dem.run_simulation()
        L.force2lattice()
        L.collide()
        L.stream()
        orc.step()
        if s % 250 == 0:
            y, vy = dem.gf.position.to_numpy()[0, 1], dem.gf.velocity.to_numpy()[0, 1]
            oy, ovy = orc.dem.a["pos"][0, 1], orc.dem.a["vel"][0, 1]
            assert vy == pytest.approx(ovy, rel=1e-6), s
            assert y == pytest.approx(oy, rel=1e-9), s
            ff = dem.gf.force_fluid.to_numpy()[0]
            assert ff == pytest.approx(orc.dem.a["force_fluid"][0], rel=1e-6, abs=1e-9 * abs(ff[1])), s
            checked += 1
    assert checked == steps // 250
    # the particle fell through the box interior and is still above the floor
    assert 0.02 < dem.gf.position.to_numpy()[0, 1] < 0.12
