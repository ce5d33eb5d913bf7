"""Process-wide settings that replace the reference's `ti.init(arch, default_fp, ...)`."""
import torch

_STORAGE_ALIASES = {
    "f32": "f32", "fp32": "f32", "float32": "f32",
    "f64": "f64", "fp64": "f64", "float64": "f64",
    "f16": "f16", "fp16": "f16", "float16": "f16", "half": "f16",
    "bf16": "bf16", "bfloat16": "bf16",
}


class _Config:
    def __init__(self):
        self.device = "cuda"
        self.storage = "f32"       # population storage of new lattices
        self.smagorinsky = "lagged"  # "lagged" (reference semantics) or "fixed" (lag-free)

    @property
    def torch_device(self):
        return torch.device(self.device)


config = _Config()


def init(arch=None, default_fp=None, device=None, storage=None, smagorinsky=None, **_ignored):
    """Select the device and the default population storage of new lattices.

    `default_fp` accepts the reference's precision switch (ti.f32 / ti.f64 in the
    reference; here a string or torch dtype).  Other keywords of ti.init (debug,
    default_ip, device_memory_fraction) are accepted and ignored.
    """
    if device is not None:
        config.device = device
    fp = storage if storage is not None else default_fp
    if fp is not None:
        if isinstance(fp, torch.dtype):
            fp = {torch.float32: "f32", torch.float64: "f64", torch.float16: "f16",
                  torch.bfloat16: "bf16"}[fp]
        key = str(fp).lower()
        if key not in _STORAGE_ALIASES:
            raise ValueError(f"unknown storage type {fp!r}")
        config.storage = _STORAGE_ALIASES[key]
    if smagorinsky is not None:
        if smagorinsky not in ("lagged", "fixed"):
            raise ValueError("smagorinsky must be 'lagged' or 'fixed'")
        config.smagorinsky = smagorinsky
    return config
