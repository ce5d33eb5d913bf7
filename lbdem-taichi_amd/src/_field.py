"""Taichi-field look-alike over a torch device tensor.

The reference's drivers touch fields element by element from Python (mask loops such as
`lattice.CT[i, j, k] = CellType.OBSTACLE`, `lattice.vel[i, j, k][0] = u`) and bulk-copy them
with `to_numpy()` / `from_numpy()` / `fill()`.  A Field keeps the authoritative copy on the
device and a host shadow for element access:

- element reads/writes go to the host shadow (downloaded once when it is stale);
- a write marks the shadow dirty; `sync()` (called before every kernel that reads the
  field) uploads it;
- kernels that write the device tensor call `touch()`, which invalidates the shadow;
- lazy fields (macroscopic rho, vel) carry a `refresh` hook that recomputes the device
  tensor from the populations before any host access.
"""
import numpy as np
import torch


class Field:
    def __init__(self, shape, dtype, device, n=None, refresh=None, tensor=None):
        self.shape = tuple(int(s) for s in shape)
        self.n = n
        full = self.shape + ((n,) if n is not None else ())
        self.dev = tensor if tensor is not None else torch.zeros(full, dtype=dtype, device=device)
        self._host = None
        self._host_dirty = False
        self._refresh = refresh
        self.version = 0  # bumped on every change (device or host)

    # --- state management -------------------------------------------------------------
    def touch(self):
        """The device tensor was written by a kernel."""
        if self._host_dirty:
            raise RuntimeError("field written on host and device without a sync in between")
        self._host = None
        self.version += 1

    def sync(self):
        """Upload pending host writes."""
        if self._host_dirty:
            self.dev.copy_(torch.from_numpy(self._host))
            self._host_dirty = False
        return self.dev

    def _pull(self):
        if self._refresh is not None and not self._host_dirty:
            self._refresh()
        if self._host is None:
            self._host = self.dev.detach().cpu().numpy().copy()
        return self._host

    # --- Taichi-like API --------------------------------------------------------------
    def __getitem__(self, idx):
        h = self._pull()
        if self.n is not None:
            # a vector element is returned as a writable view (lattice.vel[i, j, k][0] = u)
            self._host_dirty = True
            self.version += 1
            return h[idx]
        v = h[idx]
        return v.item() if isinstance(v, np.generic) else v

    def __setitem__(self, idx, value):
        h = self._pull()
        h[idx] = value
        self._host_dirty = True
        self.version += 1

    def to_numpy(self):
        return self._pull().copy()

    def to_torch(self, device=None):
        if self._refresh is not None and not self._host_dirty:
            self._refresh()
        t = self.sync()
        return t.clone() if device is None else t.to(device)

    def from_numpy(self, arr):
        arr = np.asarray(arr)
        if arr.shape != tuple(self.dev.shape):
            raise ValueError(f"shape {arr.shape} does not match field shape {tuple(self.dev.shape)}")
        self._host = None
        self._host_dirty = False
        self.dev.copy_(torch.from_numpy(np.ascontiguousarray(arr)).to(self.dev.dtype))
        self.version += 1

    def from_torch(self, t):
        self._host = None
        self._host_dirty = False
        self.dev.copy_(t)
        self.version += 1

    def fill(self, value):
        self._host = None
        self._host_dirty = False
        self.dev.fill_(value)
        self.version += 1
