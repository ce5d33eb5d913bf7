"""DEM on the GPU vs the oracle: bit-exact broad phase, trajectories within FP64 rounding."""
import numpy as np
import pytest
import torch

from dem_cases import gpu_solver, make_config, oracle_solver, random_packing, read_p4p
from oracle import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("fixture", ["vf0.35.p4p", "shear_5.p4p", "sill_5.p4p", "hybridcf.p4p", "random",
                                     "bed200k"])
@pytest.mark.parametrize("probe27", [False, True])
def test_broadphase_bit_exact(fixture, probe27):
    """Candidate pairs of the HIP broad phase == the restated reference set, as sorted sets.

    bed200k is the unresolved workload's particle set: 200,000 spheres, log-normal diameters
    (mean 0.5 lu, sigma_log 0.2, clipped below 1 lu) in a full-width 128 x 128 bed at porosity
    ~0.6 (2.25 lu tall), lattice units; 17.7 M (octant) / 9.3 M (27-cell) candidate pairs."""
    from src.dem3d.bpcd import BPCD
    if fixture == "random":
        rng = np.random.default_rng(5)
        pos = rng.uniform(0.0, 1.0, (20000, 3))
        rad = rng.uniform(0.002, 0.006, 20000)
    elif fixture == "bed200k":
        rng = np.random.default_rng(6)
        n = 200_000
        d = np.minimum(np.exp(np.log(0.5) - 0.02 + 0.2 * rng.standard_normal(n)), 0.999)
        rad = 0.5 * d
        pos = np.column_stack([rng.uniform(0, 128, n), rng.uniform(0, 128, n), rng.uniform(0, 2.25, n)])
    else:
        a = read_p4p(fixture)
        rad, pos = a[:, 2], a[:, 4:7]
    dmin = pos.min(0) - rad.max()
    dmax = pos.max(0) + rad.max()
    cell = 2.0 * rad.max() if probe27 else 4.0 * rad.max()
    bp = BPCD.create(len(pos), rad.max(), dmin, dmax, cell_size=cell, probe27=probe27)
    assert bp.table_size == O.hash_table_size(dmin, dmax, cell)
    got = bp.detect_collision(torch.tensor(pos, dtype=torch.float64, device="cuda")).cpu().numpy()
    got = got[np.lexsort((got[:, 1], got[:, 0]))]
    ref = O.broadphase(pos, dmin, cell, bp.table_size, probe27=probe27)
    print(f"{fixture} probe27={probe27}: {len(ref)} candidate pairs", flush=True)
    assert got.shape == ref.shape
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("probe27", [False, True])
def test_broadphase_without_spos_bit_exact(probe27):
    """With no spos buffer the scatter counts the bucket slots again (in `current`) instead of
    keeping those of the count kernel; same candidate set."""
    from src.dem3d.bpcd import BPCD
    a = read_p4p("sill_5.p4p")
    rad, pos = a[:, 2], a[:, 4:7]
    dmin = pos.min(0) - rad.max()
    dmax = pos.max(0) + rad.max()
    cell = 2.0 * rad.max() if probe27 else 4.0 * rad.max()
    bp = BPCD.create(len(pos), rad.max(), dmin, dmax, cell_size=cell, probe27=probe27)
    n = len(pos)
    _, w = bp.workspace(n, max(64 * n, 1024), torch.device("cuda"))
    w.spos = None
    ref = O.broadphase(pos, dmin, cell, bp.table_size, probe27=probe27)
    for _ in range(2):  # the second call finds the counters cleared
        got = bp.detect_collision(torch.tensor(pos, dtype=torch.float64, device="cuda")).cpu().numpy()
        got = got[np.lexsort((got[:, 1], got[:, 0]))]
        assert np.array_equal(got, ref)


def test_broadphase_large_table_bit_exact():
    """A table of 2^23 entries (4,096 scan tiles, more than the 2,048 whose sums every tile adds
    up itself): the tile sums are scanned by one workgroup first.  Same candidate set as the
    oracle on the same table, on two calls in a row."""
    from src.dem3d.bpcd import BPCD
    rng = np.random.default_rng(8)
    pos = rng.uniform(0.0, 1.0, (20000, 3))
    rad = rng.uniform(0.002, 0.006, 20000)
    dmin = pos.min(0) - rad.max()
    cell = 4.0 * rad.max()
    T = 1 << 23
    bp = BPCD(len(pos), T, rad.max(), dmin, cell_size=cell)
    ref = O.broadphase(pos, dmin, cell, T)
    for _ in range(2):
        got = bp.detect_collision(torch.tensor(pos, dtype=torch.float64, device="cuda")).cpu().numpy()
        got = got[np.lexsort((got[:, 1], got[:, 0]))]
        assert np.array_equal(got, ref) and len(ref) > 1000


def test_broadphase_empty_and_single():
    from src.dem3d.bpcd import BPCD
    bp = BPCD.create(1, 0.1, [0, 0, 0], [1, 1, 1])
    got = bp.detect_collision(torch.tensor([[0.5, 0.5, 0.5]], dtype=torch.float64, device="cuda"))
    assert got.shape == (0, 2)


def _run_pair(model, path, n=120, steps=400, periodic=(False, False, False), seed=1, r=0.045):
    rng = np.random.default_rng(seed)
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    cfg = make_config(model=model, box=box, dt=5e-5, periodic=periodic, kn=2e3, kt=2e3, zeta=0.3)
    rad, mass, pos, vel = random_packing(n, r, box, rng, v=0.5, polydisperse=0.3)
    gpu = gpu_solver(cfg, rad, mass, pos, vel, path=path)
    orc = oracle_solver(cfg, rad, mass, pos, vel)
    for _ in range(steps // 50):
        gpu.run_substeps(50)
        orc.step(50)
    torch.cuda.synchronize()
    return gpu, orc


@pytest.mark.parametrize("model", ["hertz", "linear"])
@pytest.mark.parametrize("path", [1, 2])
def test_dem_trajectories_match_oracle(model, path):
    gpu, orc = _run_pair(model, path)
    pos = gpu.gf.position.to_numpy()
    vel = gpu.gf.velocity.to_numpy()
    om = gpu.gf.omega.to_numpy()
    assert np.abs(pos - orc["pos"]).max() < 1e-9
    assert np.abs(vel - orc["vel"]).max() < 1e-7 * max(1.0, np.abs(orc["vel"]).max())
    assert np.abs(om - orc["omega"]).max() < 1e-6 * max(1.0, np.abs(orc["omega"]).max())
    assert np.allclose(np.linalg.norm(gpu.gf.quaternion.to_numpy(), axis=1), 1.0, atol=1e-12)
    # contact history: same active (i, j) set
    K = gpu.K
    act = gpu._cf_active.cpu().numpy().astype(bool)
    gi = np.repeat(np.arange(gpu.n), K)[act]
    gj = gpu._cf_j.cpu().numpy()[act]
    oact = orc["cf_active"].astype(bool)
    oi = np.repeat(np.arange(orc.n), K)[oact]
    oj = orc["cf_j"][oact]
    assert sorted(zip(gi.tolist(), gj.tolist())) == sorted(zip(oi.tolist(), oj.tolist()))
    assert gpu.status()["slot_overflow"] == 0


@pytest.mark.parametrize("search", ["second_grid", "tiled"])
@pytest.mark.parametrize("offset", [0.0, 1000.0])
def test_dem_coarse_grid_polydisperse_matches_oracle(offset, search):
    """Two grains of R 0.13 / 0.11 among ~2,900 of R <= 0.02: the broad-phase cell (4 r_max)
    makes a 2 x 2 x 2 grid, so each particle has ~1,000 candidates.  'second_grid': the
    grid-wide path's default here, the small grains on a 0.04-wide grid and the two large ones
    against all; 'tiled' (second grid switched off): the probe stages whole buckets in LDS.
    Trajectories and contact sets follow the oracle.  With the domain at [1000, 1001]^3 the
    staged probe's FP32 pre-test works on coordinates rounded by up to 3e-5, the size of the
    contact overlaps, so only its rounding slack (2^-20 of the coordinates' magnitude) keeps
    every touching pair."""
    rng = np.random.default_rng(3)
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    big = np.array([[0.3, 0.35, 0.4], [0.68, 0.62, 0.55]])
    Rb = np.array([0.13, 0.11])
    rad_s, _, pos_s, vel_s = random_packing(3000, 0.02, box, rng, v=2.0, polydisperse=0.3)
    keep = np.all(np.linalg.norm(pos_s[:, None, :] - big[None], axis=2) > Rb[None] + rad_s[:, None] + 0.005, axis=1)
    pos = np.vstack([big, pos_s[keep]])
    rad = np.concatenate([Rb, rad_s[keep]])
    mass = 2500.0 * 4 / 3 * np.pi * rad ** 3
    vel = np.vstack([np.zeros((2, 3)), vel_s[keep]])
    pos = pos + offset
    box = tuple(b + offset for b in box)
    cfg = make_config(model="hertz", box=box, dt=5e-5)
    gpu = gpu_solver(cfg, rad, mass, pos, vel, path=2)
    if search == "tiled":
        gpu.MAX_LARGE = 1
    orc = oracle_solver(cfg, rad, mass, pos, vel)
    for _ in range(4):
        gpu.run_substeps(50)
        orc.step(50)
    P = gpu._cached[0]
    assert (P.fine_cell > 0.0 and P.nlarge == 2) if search == "second_grid" else P.fine_cell == 0.0
    assert np.abs(gpu.gf.position.to_numpy() - orc["pos"]).max() < 1e-9
    assert np.abs(gpu.gf.velocity.to_numpy() - orc["vel"]).max() < 1e-7 * max(1.0, np.abs(orc["vel"]).max())
    K = gpu.K
    act = gpu._cf_active.cpu().numpy().astype(bool)
    got = sorted(zip(np.repeat(np.arange(gpu.n), K)[act].tolist(), gpu._cf_j.cpu().numpy()[act].tolist()))
    oact = orc["cf_active"].astype(bool)
    ref = sorted(zip(np.repeat(np.arange(orc.n), K)[oact].tolist(), orc["cf_j"][oact].tolist()))
    assert got == ref and len(ref) > 20
    gpu.check()


def test_dem_second_grid_without_large_particles_matches_oracle():
    """A broad-phase cell far above 4 r_max (cell_size = 10 r_max, set by the caller) takes the
    second-grid search with no large particle: every pair comes from the fine grid and the
    8-cell candidate filter.  Trajectories and contact sets against the oracle with the same cell."""
    rng = np.random.default_rng(17)
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    rad, mass, pos, vel = random_packing(2400, 0.018, box, rng, v=1.0, polydisperse=0.3)
    cell = float(10.0 * rad.max())
    cfg = make_config(model="hertz", box=box, dt=5e-5)
    gpu = gpu_solver(cfg, rad, mass, pos, vel, path=2, cell_size=cell)
    orc = oracle_solver(cfg, rad, mass, pos, vel, cell=cell)
    for _ in range(4):
        gpu.run_substeps(50)
        orc.step(50)
    P = gpu._cached[0]
    assert P.fine_cell > 0.0 and P.nlarge == 0
    assert np.abs(gpu.gf.position.to_numpy() - orc["pos"]).max() < 1e-9
    assert np.abs(gpu.gf.velocity.to_numpy() - orc["vel"]).max() < 1e-7 * max(1.0, np.abs(orc["vel"]).max())
    K = gpu.K
    act = gpu._cf_active.cpu().numpy().astype(bool)
    got = sorted(zip(np.repeat(np.arange(gpu.n), K)[act].tolist(), gpu._cf_j.cpu().numpy()[act].tolist()))
    oact = orc["cf_active"].astype(bool)
    ref = sorted(zip(np.repeat(np.arange(orc.n), K)[oact].tolist(), orc["cf_j"][oact].tolist()))
    assert got == ref and len(ref) > 20
    gpu.check()


def _dense_bed(n, ext, seed):
    """Grains of log-normal diameter (mean 0.5, sigma_log 0.2, below 1) seeded uniformly, with
    overlaps, in a box of extents ext (the unresolved bed's size law)."""
    rng = np.random.default_rng(seed)
    d = np.minimum(np.exp(np.log(0.5) - 0.02 + 0.2 * rng.standard_normal(n)), 0.999)
    rad = 0.5 * d
    pos = rad[:, None] + rng.random((n, 3)) * (np.array(ext) - 2 * rad[:, None])
    return rad, 2.5 * 4.0 / 3.0 * np.pi * rad ** 3, pos


@pytest.mark.parametrize("table", [None, 256, 32])
@pytest.mark.parametrize("periodic", [False, True])
def test_dem_fine_grid_probe27_matches_oracle(periodic, table):
    """6,000 grains seeded overlapping in a 12 x 12 x 6 box (about 7 per cell), broad-phase cells
    one maximum diameter wide with the 27-cell probe, as in the unresolved bed.  The touching
    probe lists each pair from one side only: with the table of the reference's size rule (2^20
    entries) and with 256 entries (cell coordinates kept modulo 8, 8 and 4; the smallest table
    the half shell is used for) it walks the half shell of 13 cells and its own bucket;
    with 32 entries (coordinates modulo 4, 4 and 2, so the 27 cells share 18 buckets) it walks
    the deduplicated buckets' entries after its own position.  The oracle runs with the same
    table (which cells alias decides which history slots the release keeps).  100, 40 and 10
    substeps of soft linear contacts (and a periodic x axis) against the oracle: positions,
    velocities and the contact sets."""
    rad, mass, pos = _dense_bed(6000, (12.0, 12.0, 6.0), seed=21)
    vel = np.zeros_like(pos)
    cell = float(2.0 * rad.max())
    cfg = make_config(model="linear", box=(0.0, 12.0, 0.0, 12.0, 0.0, 6.0), dt=0.1, g=(0.0, 0.0, -1e-5), K=64,
                      periodic=(periodic, False, False), kn=1e-2, kt=1e-2, zeta=0.3, mu=0.5)
    gpu = gpu_solver(cfg, rad, mass, pos, vel, path=2, cell_size=cell, probe27=True)
    if table is not None:
        gpu.bpcd.table_size = table
        gpu.bpcd._ws = None
        gpu._cached = None
    orc = oracle_solver(cfg, rad, mass, pos, vel, cell=cell, probe27=True)
    if table is not None:
        orc.params.T = table
    chunks = {None: (50, 50), 256: (20, 20), 32: (5, 5)}[table]
    for m in chunks:
        gpu.run_substeps(m)
        orc.step(m)
    torch.cuda.synchronize()
    assert np.abs(gpu.gf.position.to_numpy() - orc["pos"]).max() < 1e-9
    assert np.abs(gpu.gf.velocity.to_numpy() - orc["vel"]).max() < 1e-7 * max(1.0, np.abs(orc["vel"]).max())
    K = gpu.K
    act = gpu._cf_active.cpu().numpy().astype(bool)
    got = sorted(zip(np.repeat(np.arange(gpu.n), K)[act].tolist(), gpu._cf_j.cpu().numpy()[act].tolist()))
    oact = orc["cf_active"].astype(bool)
    ref = sorted(zip(np.repeat(np.arange(orc.n), K)[oact].tolist(), orc["cf_j"][oact].tolist()))
    assert got == ref and len(ref) > 10000
    assert gpu._cached[0].T == (table or 1 << 20)
    gpu.check()


@pytest.mark.parametrize("probe27", [False, True])
def test_dem_without_spos_matches_oracle(probe27):
    """The grid-wide path with no spos buffer: the scatter counts in `current`, the release
    kernel clears it with the counts, and the touching probe walks particles in index order
    (the 8-cell per-lane or staged walk, or the 27-cell per-lane walk).  The first substep runs
    with spos, the other 59 without; trajectories and contact sets against the oracle."""
    rad, mass, pos = _dense_bed(4000, (10.0, 10.0, 5.0), seed=23)
    vel = np.zeros_like(pos)
    cell = float((2.0 if probe27 else 4.0) * rad.max())
    cfg = make_config(model="linear", box=(0.0, 10.0, 0.0, 10.0, 0.0, 5.0), dt=0.1, g=(0.0, 0.0, -1e-5), K=64,
                      kn=1e-2, kt=1e-2, zeta=0.3, mu=0.5)
    gpu = gpu_solver(cfg, rad, mass, pos, vel, path=2, cell_size=cell, probe27=probe27)
    orc = oracle_solver(cfg, rad, mass, pos, vel, cell=cell, probe27=probe27)
    gpu.run_substeps(1)
    orc.step(1)
    gpu._cached[2].spos = None
    for m in (29, 30):
        gpu.run_substeps(m)
        orc.step(m)
    torch.cuda.synchronize()
    assert gpu._cached[2].spos is None and gpu._cached[0].fine_cell == 0.0
    assert np.abs(gpu.gf.position.to_numpy() - orc["pos"]).max() < 1e-9
    assert np.abs(gpu.gf.velocity.to_numpy() - orc["vel"]).max() < 1e-7 * max(1.0, np.abs(orc["vel"]).max())
    K = gpu.K
    act = gpu._cf_active.cpu().numpy().astype(bool)
    got = sorted(zip(np.repeat(np.arange(gpu.n), K)[act].tolist(), gpu._cf_j.cpu().numpy()[act].tolist()))
    oact = orc["cf_active"].astype(bool)
    ref = sorted(zip(np.repeat(np.arange(orc.n), K)[oact].tolist(), orc["cf_j"][oact].tolist()))
    assert got == ref and len(ref) > 5000
    gpu.check()


def test_dem_anisotropic_inertia_matches_oracle():
    """Full inertia tensors (principal moments 1 : 1.4 : 0.7, randomly rotated): the general
    rotation update (body frame, inverse inertia, Euler term), where spheres take the isotropic
    shortcut M / I."""
    rng = np.random.default_rng(21)
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    cfg = make_config(model="hertz", box=box, dt=5e-5, kn=2e3, kt=2e3, zeta=0.3)
    rad, mass, pos, vel = random_packing(120, 0.045, box, rng, v=0.5, polydisperse=0.3)
    gpu = gpu_solver(cfg, rad, mass, pos, vel, path=1)
    orc = oracle_solver(cfg, rad, mass, pos, vel)
    n = len(rad)
    inertia = np.zeros((n, 3, 3))
    for i in range(n):
        q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
        c = 0.4 * mass[i] * rad[i] ** 2
        inertia[i] = q @ np.diag([c, 1.4 * c, 0.7 * c]) @ q.T
    orc["inertia"][...] = inertia
    gpu.gf.inertia.from_numpy(inertia.reshape(n, 9))
    omega0 = rng.uniform(-20.0, 20.0, (n, 3))
    orc["omega"][...] = omega0
    gpu.gf.omega.from_numpy(omega0)
    for _ in range(8):
        gpu.run_substeps(50)
        orc.step(50)
    torch.cuda.synchronize()
    om = gpu.gf.omega.to_numpy()
    assert np.abs(gpu.gf.position.to_numpy() - orc["pos"]).max() < 1e-9
    assert np.abs(om - orc["omega"]).max() < 1e-6 * max(1.0, np.abs(orc["omega"]).max())
    assert np.abs(gpu.gf.quaternion.to_numpy() - orc["quat"]).max() < 1e-9


def test_dem_periodic_axis():
    gpu, orc = _run_pair("hertz", 0, periodic=(True, False, False), seed=4)
    assert np.abs(gpu.gf.position.to_numpy() - orc["pos"]).max() < 1e-9


def test_dem_small_path_state_in_global_memory():
    """600 particles: the single-workgroup path with the state in global memory (above the
    LDS-resident size) against the oracle."""
    gpu, orc = _run_pair("hertz", 1, n=600, seed=11, r=0.03)
    assert np.abs(gpu.gf.position.to_numpy() - orc["pos"]).max() < 1e-9
    assert np.abs(gpu.gf.velocity.to_numpy() - orc["vel"]).max() < 1e-7 * max(1.0, np.abs(orc["vel"]).max())
    assert gpu.status()["slot_overflow"] == 0


def test_dem_resident_path_neighbour_list_overflow():
    """One large sphere among 150 small ones: the large sphere has far more than the 16 Verlet
    neighbours the single-workgroup kernel keeps, so it falls back to the per-substep hash
    probe; the trajectories and the contact set still match the oracle."""
    rng = np.random.default_rng(12)
    box = (0.0, 1.0, 0.0, 1.0, 0.0, 1.0)
    cfg = make_config(model="hertz", box=box, dt=5e-5, g=(0.0, 0.0, 0.0))
    pos, rad = [np.array([0.5, 0.5, 0.5])], [0.12]
    while len(pos) < 151:
        x = rng.uniform(0.03, 0.97, 3)
        r = 0.02
        if (np.linalg.norm(np.array(pos) - x, axis=1) - np.array(rad) - r < 0.002).any():
            continue
        if np.linalg.norm(x - pos[0]) > 0.2:  # crowd the small spheres round the large one
            continue
        pos.append(x)
        rad.append(r)
    pos, rad = np.array(pos), np.array(rad)
    mass = 2500.0 * 4.0 / 3.0 * np.pi * rad ** 3
    vel = rng.uniform(-0.3, 0.3, pos.shape)
    vel[1:] -= 0.5 * (pos[1:] - pos[0])  # drift towards the large sphere: contacts form
    gpu = gpu_solver(cfg, rad, mass, pos, vel, path=1)
    orc = oracle_solver(cfg, rad, mass, pos, vel)
    for _ in range(6):
        gpu.run_substeps(50)
        orc.step(50)
    torch.cuda.synchronize()
    assert np.abs(gpu.gf.position.to_numpy() - orc["pos"]).max() < 1e-9
    assert np.abs(gpu.gf.velocity.to_numpy() - orc["vel"]).max() < 1e-7 * max(1.0, np.abs(orc["vel"]).max())
    K = gpu.K
    act = gpu._cf_active.cpu().numpy().astype(bool)
    gi, gj = np.repeat(np.arange(gpu.n), K)[act], gpu._cf_j.cpu().numpy()[act]
    oact = orc["cf_active"].astype(bool)
    oi, oj = np.repeat(np.arange(orc.n), K)[oact], orc["cf_j"][oact]
    assert len(gi) > 0, "the case must make contacts"
    assert sorted(zip(gi.tolist(), gj.tolist())) == sorted(zip(oi.tolist(), oj.tolist()))


def test_dem_small_and_general_paths_agree():
    a, _ = _run_pair("hertz", 1, seed=9)
    b, _ = _run_pair("hertz", 2, seed=9)
    assert np.abs(a.gf.position.to_numpy() - b.gf.position.to_numpy()).max() < 1e-10


def test_dem_p4p_loader_and_writer(tmp_path):
    """init_particle_fields reads a reference packing; save_single writes p4p / p4c."""
    from src.dem3d.demsolver import DEMSolver
    import os
    from dem_cases import GOLDEN
    a = read_p4p("shear_5.p4p")
    lo, hi = a[:, 4:7].min(0) - 0.01, a[:, 4:7].max(0) + 0.01
    cfg = make_config(box=(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2]), dt=1e-6)
    s = DEMSolver(cfg)
    s.init_particle_fields(os.path.join(GOLDEN, "shear_5.p4p"), lo, hi)
    s.set_contact_model("hertz")
    assert s.n == len(a)
    assert np.allclose(s.gf.position.to_numpy(), a[:, 4:7])
    s.run_simulation()
    with open(tmp_path / "o.p4p", "w") as p4p, open(tmp_path / "o.p4c", "w") as p4c:
        s.save_single(p4p, p4c, 0.0)
    txt = (tmp_path / "o.p4p").read_text().splitlines()
    assert txt[0].startswith("TIMESTEP") and int(txt[1].split()[1]) == len(a)
