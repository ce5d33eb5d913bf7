"""Differentiable PSC channel flow with a hand-written adjoint, exposed to torch.autograd.

Forward model per step: PSC collision with fixed grains and a body force on the fluid part,
then bounce-back / periodic streaming (lbdem_diff_step).  The reference has no adjoint; this
module adds one.  `PSCRollout.apply(F, tau, eps, channel, steps)` returns u_x on the outlet
plane after `steps` steps from rest; its backward runs the adjoint steps
(lbdem_diff_adjoint_step) in reverse with checkpoint-every-K recomputation and returns
dL/dF (3), dL/dtau and dL/d eps (per cell).
"""
import ctypes

import numpy as np
import torch

from . import _hip
from .lbm3d.lbmutils import CellType

_vp = ctypes.c_void_p


class DiffArgs(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("precision", ctypes.c_int32), ("flags", _vp), ("eps", _vp), ("tau", ctypes.c_double),
                ("force", ctypes.c_double * 3), ("plane", ctypes.c_int32), ("reserved", ctypes.c_int32)]


_P = ctypes.POINTER
_hip.register({
    "lbdem_diff_init": (ctypes.c_int, [_P(DiffArgs), _vp, _vp]),
    "lbdem_diff_step": (ctypes.c_int, [_P(DiffArgs), _vp, _vp, _vp]),
    "lbdem_diff_adjoint_step": (ctypes.c_int, [_P(DiffArgs), _vp, _vp, _vp, _vp, _vp, _vp]),
    "lbdem_diff_outlet": (ctypes.c_int, [_P(DiffArgs), _vp, _vp, _vp]),
    "lbdem_diff_seed": (ctypes.c_int, [_P(DiffArgs), _vp, _vp, _vp, _vp]),
    "lbdem_diff_red_doubles": (ctypes.c_int64, [_P(DiffArgs)]),
    "lbdem_diff_reduce": (ctypes.c_int, [_P(DiffArgs), _vp, _vp, _vp]),
})


class DiffPSCChannel:
    """Lattice, walls and checkpoint storage of a differentiable PSC rollout.

    CT: (nx, ny, nz) CellType array (FLUID / OBSTACLE); plane: outlet index along x;
    precision: 32 or 64; K: checkpoint interval (states kept every K steps, and K states
    recomputed per segment during the adjoint sweep)."""

    def __init__(self, CT, plane=None, precision=64, K=100):
        _hip.require_gpu()
        CT = np.asarray(CT)
        self.nx, self.ny, self.nz = CT.shape
        self.N = CT.size
        self.plane = self.nx - 1 if plane is None else int(plane)
        self.precision = int(precision)
        self.dtype = torch.float64 if self.precision == 64 else torch.float32
        self.K = int(K)
        kind = CT & (CellType.FLUID | CellType.OBSTACLE)
        if not np.all((kind == CellType.FLUID) | (kind == CellType.OBSTACLE)):
            raise NotImplementedError("the differentiable channel supports FLUID and OBSTACLE cells")
        from . import _hip as H
        flags = np.where(kind == CellType.FLUID, H.FLAG_ACTIVE | H.FLAG_FLUID, H.FLAG_WALL).astype(np.uint8)
        self.device = torch.device("cuda")
        self.flags = torch.from_numpy(flags.reshape(-1)).to(self.device)
        # NEAR bits (cells with a wall neighbour take the bounce-back path) derived on the device
        dims = H.Lattice()
        dims.nx, dims.ny, dims.nz = self.nx, self.ny, self.nz
        nmask = torch.zeros(self.N, dtype=torch.int32, device=self.device)
        H.check(H.lib().lbdem_lbm_prepare_flags(ctypes.byref(dims), H.ptr(self.flags), H.ptr(nmask), H.stream()),
                "lbdem_lbm_prepare_flags")
        self._args = DiffArgs()
        a = self._args
        a.nx, a.ny, a.nz = self.nx, self.ny, self.nz
        a.precision = self.precision
        a.flags = self.flags.data_ptr()
        a.plane = self.plane
        self._ckpt = None
        self._fT = None

    def _set(self, F, tau, eps):
        a = self._args
        a.force[0], a.force[1], a.force[2] = (float(v) for v in F)
        a.tau = float(tau)
        a.eps = eps.data_ptr()
        return ctypes.byref(a)

    def _state(self):
        return torch.empty((19, self.N), dtype=self.dtype, device=self.device)

    def forward(self, F, tau, eps, steps):
        """Run `steps` steps from rest; keep every K-th state.  Returns u_x on the outlet plane."""
        eps = eps.to(self.device, self.dtype).contiguous().view(-1)
        self._eps = eps
        self._F, self._tau, self._steps = tuple(float(v) for v in F), float(tau), int(steps)
        a = self._set(self._F, self._tau, eps)
        L = _hip.lib()
        s = _hip.stream()
        cur, nxt = self._state(), self._state()
        _hip.check(L.lbdem_diff_init(a, _hip.ptr(cur), s), "lbdem_diff_init")
        self._ckpt = []
        for n in range(self._steps):
            if n % self.K == 0:
                self._ckpt.append(cur.clone())
            _hip.check(L.lbdem_diff_step(a, _hip.ptr(cur), _hip.ptr(nxt), s), "lbdem_diff_step")
            cur, nxt = nxt, cur
        self._fT = cur
        u = torch.empty((self.ny, self.nz), dtype=self.dtype, device=self.device)
        _hip.check(L.lbdem_diff_outlet(a, _hip.ptr(cur), _hip.ptr(u), s), "lbdem_diff_outlet")
        return u

    def adjoint(self, g_u):
        """Gradients of a loss L(u) given dL/du (ny, nz): (dL/dF [3], dL/dtau, dL/deps [nx, ny, nz])."""
        L = _hip.lib()
        s = _hip.stream()
        a = self._set(self._F, self._tau, self._eps)
        g_u = g_u.to(self.device, self.dtype).contiguous()
        lam, lam2 = self._state(), self._state()
        _hip.check(L.lbdem_diff_seed(a, _hip.ptr(self._fT), _hip.ptr(g_u), _hip.ptr(lam), s), "lbdem_diff_seed")
        g_eps = torch.zeros(self.N, dtype=self.dtype, device=self.device)
        g_red = torch.zeros(L.lbdem_diff_red_doubles(a), dtype=torch.float64, device=self.device)
        T, K = self._steps, self.K
        buf = torch.empty((K, 19, self.N), dtype=self.dtype, device=self.device)
        for seg in range((T + K - 1) // K - 1, -1, -1):
            n0 = seg * K
            m = min(K, T - n0)
            buf[0].copy_(self._ckpt[seg])
            for j in range(1, m):  # recompute the states of this segment
                _hip.check(L.lbdem_diff_step(a, _hip.ptr(buf[j - 1]), _hip.ptr(buf[j]), s), "lbdem_diff_step")
            for j in range(m - 1, -1, -1):
                _hip.check(L.lbdem_diff_adjoint_step(a, _hip.ptr(buf[j]), _hip.ptr(lam), _hip.ptr(lam2),
                                                     _hip.ptr(g_eps), _hip.ptr(g_red), s),
                           "lbdem_diff_adjoint_step")
                lam, lam2 = lam2, lam
        tot = torch.empty(4, dtype=torch.float64, device=self.device)
        _hip.check(L.lbdem_diff_reduce(a, _hip.ptr(g_red), _hip.ptr(tot), s), "lbdem_diff_reduce")
        gF = tot[1:4].clone()
        gtau = tot[0].clone()
        return gF, gtau, g_eps.view(self.nx, self.ny, self.nz)


class PSCRollout(torch.autograd.Function):
    """u_x(y, z) on the outlet plane after `steps` steps, differentiable in F, tau, eps."""

    @staticmethod
    def forward(ctx, F, tau, eps, channel, steps):
        Fh = F.detach().double().cpu().numpy().reshape(3)
        th = float(tau.detach().cpu())
        u = channel.forward(Fh, th, eps.detach(), steps)
        ctx.channel = channel
        ctx.dtypes = (F.dtype, tau.dtype, eps.dtype, F.device, tau.device, eps.device)
        return u

    @staticmethod
    def backward(ctx, g_u):
        gF, gtau, geps = ctx.channel.adjoint(g_u)
        dF, dtau, deps, devF, devt, deve = ctx.dtypes
        return gF.to(devF, dF), gtau.to(devt, dtau), geps.to(deve, deps), None, None


def sphere_fraction(shape, centres, radius, sub=5):
    """Solid fraction of cells covered by spheres, by sub x sub x sub point sampling per cell
    (the sampling rule of PSCLattice3D.grains2lattice, psclattice3d.py:119-131)."""
    nx, ny, nz = shape
    eps = np.zeros(shape)
    offs = -0.5 + (np.arange(sub) + 0.5) / sub
    for c in np.asarray(centres, dtype=np.float64):
        r = radius
        lo = np.maximum(np.floor(c - r - 1).astype(int), 0)
        hi = np.minimum(np.ceil(c + r + 1).astype(int) + 1, shape)
        ii, jj, kk = np.meshgrid(*(np.arange(lo[a], hi[a]) for a in range(3)), indexing="ij")
        cnt = np.zeros(ii.shape)
        for ox in offs:
            for oy in offs:
                for oz in offs:
                    d2 = (ii + ox - c[0]) ** 2 + (jj + oy - c[1]) ** 2 + (kk + oz - c[2]) ** 2
                    cnt += d2 < r * r
        sl = (slice(lo[0], hi[0]), slice(lo[1], hi[1]), slice(lo[2], hi[2]))
        eps[sl] = np.maximum(eps[sl], cnt / sub ** 3)
    return eps
