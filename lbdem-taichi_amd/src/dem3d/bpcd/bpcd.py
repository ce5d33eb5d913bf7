"""Broad-phase collision detection with Morton-coded spatial hashing (src/dem3d/bpcd/bpcd.py).

The counting sort, scan and probe run as HIP kernels (lbdem_dem_broadphase).  Cells are
`4 * max_radius` wide and probed by octant (8 cells), as in the reference; `cell_size` and
`probe27` select the variant with cells one diameter wide and a 27-cell probe.
"""
import ctypes

import numpy as np
import torch

from ... import _hip
from ...runtime import config
from .. import _abi
from .utils import next_pow2


class BPCD:
    def __init__(self, particle_count: int, hash_table_size: int, max_radius: float, domain_min, cell_size=None,
                 probe27=False):
        self.particle_count = int(particle_count)
        self.cell_size = float(cell_size) if cell_size is not None else max_radius * 4.0
        self.domain_min = np.asarray(domain_min, dtype=np.float64).reshape(3)
        self.table_size = int(hash_table_size)
        self.probe27 = bool(probe27)
        self._ws = None

    @staticmethod
    def table_size_for(domain_min, domain_max, cell):
        v = (np.asarray(domain_max, dtype=np.float64) - np.asarray(domain_min, dtype=np.float64)) / cell
        size = int(v[0] * v[1] * v[2])
        size = next_pow2(next_pow2(size))
        return min(max(size, 1 << 20), 1 << 22)

    @staticmethod
    def create(particle_count: int, max_radius: float, domain_min, domain_max, cell_size=None, probe27=False):
        """Table size from the domain volume in cells, rounded up to a power of two and clamped to
        [2^20, 2^22] (bpcd.py:48-65)."""
        cell = cell_size if cell_size is not None else 4.0 * max_radius
        size = BPCD.table_size_for(domain_min, domain_max, cell)
        return BPCD(particle_count, size, max_radius, domain_min, cell_size=cell, probe27=probe27)

    # ---------------------------------------------------------------------------------
    def params(self, n=None):
        P = _abi.DemParams()
        P.n = self.particle_count if n is None else int(n)
        P.K = 1
        P.model = 0
        P.T = self.table_size
        P.probe27 = 1 if self.probe27 else 0
        P.cell = self.cell_size
        for a in range(3):
            P.dmin[a] = float(self.domain_min[a])
        return P

    def workspace(self, n, pair_cap, device):
        key = (n, pair_cap)
        if self._ws is None or self._ws[0] != key:
            T = self.table_size
            z = lambda m, dt=torch.int32: torch.zeros(max(int(m), 1), dtype=dt, device=device)  # noqa: E731
            t = dict(count=z(T), offset=z(T), current=z(T), tile_sum=z((T + 1023) // 1024), hash=z(n), pid=z(n),
                     pairs=z(2 * pair_cap), npairs=z(1, torch.int64), need_new=z(pair_cap),
                     spos=z(4 * n, torch.float64), fhash=z(n))
            w = _abi.DemWork()
            for k, v in t.items():
                setattr(w, k, v.data_ptr())
            w.pair_cap = int(pair_cap)
            self._ws = (key, t, w)
        return self._ws[1], self._ws[2]

    def detect_collision(self, positions, collision_resolve_callback=None, pair_cap=None):
        """Candidate pairs (i < j) of the particles at `positions` ((n, 3) FP64 tensor or Field).

        Returns an (m, 2) int32 tensor on the device; when a callback is given it is called once
        with that tensor (the reference calls a Taichi function per pair)."""
        pos = positions.sync() if hasattr(positions, "sync") else positions
        pos = pos.to(device=config.torch_device, dtype=torch.float64).contiguous()
        n = pos.shape[0]
        cap = int(pair_cap) if pair_cap is not None else max(64 * n, 1024)
        for _ in range(2):
            t, w = self.workspace(n, cap, pos.device)
            t["npairs"].zero_()
            P = self.params(n)
            _hip.check(_hip.lib().lbdem_dem_broadphase(ctypes.byref(P), _hip.ptr(pos), ctypes.byref(w),
                                                       _hip.stream()), "lbdem_dem_broadphase")
            m = int(t["npairs"].item())
            if m <= cap:
                break
            cap = m
        pairs = t["pairs"][:2 * m].view(m, 2).clone()
        if collision_resolve_callback is not None:
            collision_resolve_callback(pairs)
        return pairs
