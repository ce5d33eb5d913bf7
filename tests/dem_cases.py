"""Shared DEM set-ups: the same particles for the oracle and the GPU solver."""
import os

import numpy as np

from oracle import oracle as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def read_p4p(name):
    """(n, 10) rows ID GROUP RAD MASS PX PY PZ VX VY VZ of a reference packing."""
    with open(os.path.join(GOLDEN, name), encoding="UTF-8") as fp:
        fp.readline()
        n = int(fp.readline().split()[1])
        fp.readline()
        rows = [fp.readline().split() for _ in range(n)]
    return np.array([r for r in rows if r], dtype=np.float64)


def make_config(model="hertz", box=(0.0, 1.0, 0.0, 1.0, 0.0, 1.0), dt=1e-4, g=(0.0, 0.0, -9.81), K=32,
                periodic=(False, False, False), E=1e7, nu=0.3, e=0.8, mu=0.3, kn=1e4, kt=1e4, zeta=0.2):
    from src.dem3d.demconfig import DEMSolverConfig, DomainBounds, HertzContactConfig, LinearContactConfig
    if model == "hertz":
        cm = HertzContactConfig(pp_friction=mu, pw_friction=mu, pp_restitution=e, pw_restitution=e)
    else:
        cm = LinearContactConfig(stiffness_normal=kn, stiffness_tangential=kt, damping_normal=zeta,
                                 damping_tangential=zeta, pp_friction=mu, pw_friction=mu)
    cfg = DEMSolverConfig(domain=DomainBounds(*box), dt=dt, gravity=g, contact_model=cm)
    cfg.set_particle_properties(elastic_modulus=E, poisson_ratio=nu, max_coordinate_number=K)
    cfg.set_wall_properties(elastic_modulus=E, poisson_ratio=nu, density=2500.0)
    cfg.set_periodic_boundaries(*periodic)
    return cfg


def random_packing(n, R, box, rng, density=2500.0, v=0.1, polydisperse=0.0):
    """Non-overlapping spheres by rejection sampling inside box (xmin, xmax, ...)."""
    lo = np.array(box[0::2]) + 1.2 * R
    hi = np.array(box[1::2]) - 1.2 * R
    pos, rad = [], []
    tries = 0
    while len(pos) < n and tries < 200000:
        tries += 1
        r = R * (1.0 - polydisperse * rng.random())
        x = lo + (hi - lo) * rng.random(3)
        if pos:
            d = np.linalg.norm(np.array(pos) - x, axis=1) - np.array(rad) - r
            if (d < 0.02 * R).any():
                continue
        pos.append(x)
        rad.append(r)
    pos, rad = np.array(pos), np.array(rad)
    mass = density * 4.0 / 3.0 * np.pi * rad ** 3
    vel = v * (rng.random((len(pos), 3)) - 0.5)
    return rad, mass, pos, vel


def gpu_solver(cfg, rad, mass, pos, vel, path=0, **kw):
    from src.dem3d.demsolver import DEMSolver
    s = DEMSolver(cfg)
    s.init_particle_arrays(radius=rad, mass=mass, position=pos, velocity=vel, **kw)
    s.set_contact_model(cfg.contact_model.get_model_name())
    s.path = path
    return s


def oracle_solver(cfg, rad, mass, pos, vel, **kw):
    return O.OracleDEM(cfg, rad, mass, pos, vel, **kw)
