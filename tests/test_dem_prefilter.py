"""The FP32 pre-test of the staged broad-phase probe (dem.hip, k_bp_probe_touch) never drops a
pair that the FP64 contact test would call touching.  The same IEEE float32 operations, in the
same order and without fused multiply-adds (the library is built with -ffp-contract=off),
evaluated with numpy over pairs straddling contact at coordinate magnitudes from 1e-3 to 1e3.
CPU only."""
import numpy as np


def _f32_drops(xi, xj, ri, rj):
    f = np.float32
    xif, xjf = xi.astype(f), xj.astype(f)
    rif, rjf = ri.astype(f), rj.astype(f)
    mif = np.max(np.abs(xif), axis=1)
    mjf = np.max(np.abs(xjf), axis=1)
    d = xif - xjf
    x2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    T = (rif + rjf) * f(1.000001) + f(9.53674316e-7) * (mif + mjf)
    return x2 > T * T * f(1.000004)


def _f64_touching(xi, xj, ri, rj):
    d = xi - xj
    x2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
    a = ri + rj
    near = ~(x2 > a * a * (1.0 + 1e-9))
    return near & (np.sqrt(x2) - ri - rj < 0)


def test_fp32_pretest_is_conservative():
    rng = np.random.default_rng(0)
    n = 400_000
    for scale in (1e-3, 0.2, 1.0, 37.0, 1e3):
        ri = 10 ** rng.uniform(-4, -1, n) * max(scale, 1e-2)
        rj = ri * 10 ** rng.uniform(-1.5, 1.5, n)
        xi = rng.uniform(-scale, scale, (n, 3))
        u = rng.normal(size=(n, 3))
        u /= np.linalg.norm(u, axis=1, keepdims=True)
        # distances within a relative 1e-3 of contact, on both sides, down to round-off
        rel = np.concatenate([rng.uniform(-1e-3, 1e-3, n // 2), rng.uniform(-1e-12, 1e-12, n - n // 2)])
        xj = xi + u * ((ri + rj) * (1.0 + rel))[:, None]
        drop = _f32_drops(xi, xj, ri, rj)
        touch = _f64_touching(xi, xj, ri, rj)
        assert not (drop & touch).any(), scale
        # and the pre-test does drop pairs clearly apart (it is not vacuous)
        far = xi + u * ((ri + rj) * 1.01 + 1e-6 * scale)[:, None]
        assert _f32_drops(xi, far, ri, rj).mean() > 0.5, scale
