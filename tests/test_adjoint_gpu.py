"""Hand-written adjoint of the PSC + body-force channel vs central finite differences (FP64)
and the FP32 adjoint vs the FP64 adjoint."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

NX, NY, NZ, STEPS, K = 14, 12, 10, 60, 8


def _channel(precision):
    from src.adjoint import DiffPSCChannel, sphere_fraction
    from src.lbm3d.lbmutils import CellType
    CT = np.full((NX, NY, NZ), CellType.FLUID, dtype=np.int32)
    CT[:, 0, :] = CellType.OBSTACLE
    CT[:, -1, :] = CellType.OBSTACLE
    eps = sphere_fraction((NX, NY, NZ), [[4.3, 5.2, 4.1], [9.6, 6.4, 6.2]], 2.6)
    return DiffPSCChannel(CT, precision=precision, K=K), eps


def _loss(u, target):
    return ((u - target) ** 2).mean()


def _params(eps, dtype=torch.float64):
    F = torch.tensor([2e-5, 3e-6, -1e-6], dtype=dtype, device="cuda", requires_grad=True)
    tau = torch.tensor(0.6, dtype=dtype, device="cuda", requires_grad=True)
    e = torch.tensor(eps, dtype=dtype, device="cuda", requires_grad=True)
    return F, tau, e


def _target():
    rng = np.random.default_rng(0)
    y = np.arange(NY)
    prof = 1e-3 * np.clip((y - 0.5) * (NY - 1.5 - y), 0, None) / ((NY - 2) / 2) ** 2
    t = prof[:, None] * np.ones((1, NZ)) * (1 + 0.05 * rng.standard_normal((NY, NZ)))
    return torch.tensor(t, dtype=torch.float64, device="cuda")


def test_adjoint_matches_central_differences():
    from src.adjoint import PSCRollout
    ch, eps = _channel(64)
    target = _target()
    F, tau, e = _params(eps)
    loss = _loss(PSCRollout.apply(F, tau, e, ch, STEPS), target)
    loss.backward()

    def L(Fv, tv, ev):
        with torch.no_grad():
            return float(_loss(ch.forward(Fv, tv, torch.tensor(ev, dtype=torch.float64, device="cuda"), STEPS),
                               target))

    F0, t0 = F.detach().cpu().numpy(), float(tau)
    checks = []
    for a in range(3):
        h = 1e-4 * max(abs(F0[a]), 1e-6)
        Fp, Fm = F0.copy(), F0.copy()
        Fp[a] += h
        Fm[a] -= h
        checks.append((F.grad[a].item(), (L(Fp, t0, eps) - L(Fm, t0, eps)) / (2 * h)))
    h = 1e-4 * t0
    checks.append((tau.grad.item(), (L(F0, t0 + h, eps) - L(F0, t0 - h, eps)) / (2 * h)))
    rng = np.random.default_rng(1)
    fluid = np.argwhere((eps == 0))
    fluid = fluid[(fluid[:, 1] > 0) & (fluid[:, 1] < NY - 1)]
    partial = np.argwhere((eps > 0) & (eps < 1))
    full = np.argwhere(eps == 1)
    cells = [tuple(c) for c in (list(fluid[rng.choice(len(fluid), 3, replace=False)]) +
                                list(partial[rng.choice(len(partial), 3, replace=False)]) +
                                list(full[rng.choice(len(full), 2, replace=False)]))]
    for c in cells:
        h = 1e-4
        ep, em = eps.copy(), eps.copy()
        ep[c] += h
        em[c] -= h
        checks.append((e.grad[c].item(), (L(F0, t0, ep) - L(F0, t0, em)) / (2 * h)))
    scale = max(abs(fd) for _, fd in checks)
    for ad, fd in checks:
        assert abs(ad - fd) <= 1e-4 * max(abs(fd), 1e-3 * scale), (ad, fd)


def test_fp32_adjoint_close_to_fp64_adjoint():
    from src.adjoint import PSCRollout
    target = _target()
    grads = []
    for prec, dt in ((64, torch.float64), (32, torch.float32)):
        ch, eps = _channel(prec)
        F, tau, e = _params(eps, dt)
        loss = _loss(PSCRollout.apply(F, tau, e, ch, STEPS).double(), target)
        loss.backward()
        grads.append((F.grad.double().cpu().numpy(), tau.grad.double().item(), e.grad.double().cpu().numpy()))
    (F64, t64, e64), (F32, t32, e32) = grads
    assert np.abs(F32 - F64).max() <= 2e-3 * np.abs(F64).max()
    assert abs(t32 - t64) <= 2e-3 * abs(t64)
    assert np.abs(e32 - e64).max() <= 2e-3 * np.abs(e64).max()
