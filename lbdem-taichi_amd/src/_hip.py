"""ctypes binding of liblbdem_hip.so (declarations in include/lbdem_hip.h).

The library is the only compute path: if it is missing, or no GPU is present, the solver
classes raise instead of falling back to any CPU code.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("LBDEM_HIP_LIB") or os.path.normpath(os.path.join(_HERE, "..", "liblbdem_hip.so"))

STORE = {"f32": 0, "f16": 1, "bf16": 2, "f64": 3}
STORE_DTYPE = {"f32": torch.float32, "f16": torch.float16, "bf16": torch.bfloat16, "f64": torch.float64}
COLL_BGK, COLL_SMAG_LAGGED, COLL_SMAG_FIXED = 0, 1, 2
COUPLE_NONE, COUPLE_PSC, COUPLE_EQIBM, COUPLE_VANS, COUPLE_HYBRID = 0, 1, 2, 3, 4
FLAG_ACTIVE, FLAG_NEAR, FLAG_WALL, FLAG_MOVING, FLAG_FLUID, FLAG_INLET = 0x01, 0x02, 0x04, 0x08, 0x10, 0x20
FLAG_ROW_IDLE = 0x40
BC_VEL_ZOUHE, BC_PRE_ZOUHE, BC_EXIT, BC_LADD_INLET = 1, 2, 3, 4

_vp = ctypes.c_void_p


class Lattice(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("storage", ctypes.c_int32), ("collision", ctypes.c_int32), ("coupling", ctypes.c_int32),
                ("omega0", ctypes.c_double), ("cs_smag", ctypes.c_double), ("rho0", ctypes.c_double),
                ("force", ctypes.c_double * 3), ("kernel", ctypes.c_int32), ("reserved", ctypes.c_int32)]


ABI_VERSION = 9


class StepArgs(ctypes.Structure):
    _fields_ = [("f", _vp), ("flags", _vp), ("nmask", _vp), ("vel", _vp), ("lag", _vp),
                ("psc_key", _vp), ("psc_grain", _vp), ("psc_force", _vp),
                ("volfrac", _vp), ("velsolid", _vp), ("hforce", _vp),
                ("volfrac_prev", _vp), ("hforce_prev", _vp),
                ("t", ctypes.c_int64), ("nu_lu", ctypes.c_double), ("rho", _vp)]


_lib = None


def lib():
    """Load the library (once) and declare every prototype."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python __graft_entry__.py` "
                           "(or lbdem-taichi_amd/build_lib.py); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.POINTER
    sig = {
        "lbdem_abi_version": (ctypes.c_int, []),
        "lbdem_last_error": (ctypes.c_char_p, []),
        "lbdem_lbm_init": (ctypes.c_int, [P(Lattice), P(StepArgs), _vp, _vp]),
        "lbdem_lbm_step": (ctypes.c_int, [P(Lattice), P(StepArgs), _vp]),
        "lbdem_lbm_macros": (ctypes.c_int, [P(Lattice), P(StepArgs), _vp, _vp, _vp, _vp]),
        "lbdem_lbm_read_f": (ctypes.c_int, [P(Lattice), P(StepArgs), _vp, ctypes.c_int32, _vp]),
        "lbdem_lbm_write_f": (ctypes.c_int, [P(Lattice), P(StepArgs), _vp, ctypes.c_int32, _vp]),
        "lbdem_reduce_work_doubles": (ctypes.c_int64, [P(Lattice)]),
        "lbdem_lbm_reduce": (ctypes.c_int, [P(Lattice), P(StepArgs), _vp, _vp, _vp]),
        "lbdem_lbm_reduce_static": (ctypes.c_int, [P(Lattice), P(StepArgs), _vp, _vp]),
        "lbdem_lbm_prepare_flags": (ctypes.c_int, [P(Lattice), _vp, _vp, _vp]),
        "lbdem_lbm_boundaries": (ctypes.c_int, [P(Lattice), P(StepArgs), _vp, _vp, ctypes.c_int32, _vp, _vp, _vp]),
        "lbdem_lbm_free_slip": (ctypes.c_int, [P(Lattice), P(StepArgs), _vp, ctypes.c_int32, _vp, _vp]),
        "lbdem_lbm_stress": (ctypes.c_int, [P(Lattice), P(StepArgs), _vp, _vp]),
        "lbdem_lbm_plane_stress": (ctypes.c_int, [P(Lattice), _vp, _vp, ctypes.c_int32, ctypes.c_int32, _vp, _vp]),
    }
    sig.update(_EXTRA_SIGNATURES)
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.lbdem_abi_version() != ABI_VERSION:
        raise RuntimeError("liblbdem_hip.so ABI version mismatch")
    _lib = L
    return L


# prototypes registered by the coupling / DEM modules
_EXTRA_SIGNATURES = {}


def register(signatures):
    _EXTRA_SIGNATURES.update(signatures)
    if _lib is not None:
        for name, (res, args) in signatures.items():
            fn = getattr(_lib, name)
            fn.restype = res
            fn.argtypes = args


def check(rc, what):
    if rc != 0:
        msg = lib().lbdem_last_error()
        raise RuntimeError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def ptr(t):
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise RuntimeError("liblbdem_hip expects device tensors")
    return _vp(t.data_ptr())


def stream():
    return _vp(torch.cuda.current_stream().cuda_stream)


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the LBDEM solvers run only on the HIP path (gfx950)")
    lib()
