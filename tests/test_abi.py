"""The C-ABI library loads and exports every entry point its header declares (CPU only)."""
import ctypes
import os
import re

from conftest import ROOT, PKG

HEADER = os.path.join(ROOT, "include", "lbdem_hip.h")
LIB = os.path.join(PKG, "liblbdem_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(lbdem_[a-z0-9_]+)\s*\(", text)
    return sorted(set(names))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "lbdem_lbm_step" in names and "lbdem_abi_version" in names
    assert len(names) >= 9


def test_library_exports_every_declared_symbol():
    assert os.path.exists(LIB), "build liblbdem_hip.so first (python __graft_entry__.py)"
    # torch first: its HIP runtime (same soname) then serves the library too; a library loaded
    # before torch brings a second HIP runtime into the process, and GPU calls made later in
    # the same process see no device
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, f"not exported: {missing}"
    L.lbdem_abi_version.restype = ctypes.c_int
    import src._hip as H
    assert L.lbdem_abi_version() == H.ABI_VERSION


def test_python_binding_declares_every_header_function():
    import src._hip as H
    H.lib()
    import importlib
    for mod in ("src.lbm3d", "src.lbdem.psclattice3d", "src.unlbdem.eqlattice", "src.unlbdem.vanslattice",
                "src.hybridlbdem.hybridlattice", "src.dem3d.demsolver", "src.adjoint"):
        try:
            importlib.import_module(mod)  # registers that module's prototypes
        except ModuleNotFoundError:
            pass
    known = set(H._EXTRA_SIGNATURES) | {
        "lbdem_abi_version", "lbdem_last_error", "lbdem_lbm_init", "lbdem_lbm_step", "lbdem_lbm_macros",
        "lbdem_lbm_read_f", "lbdem_lbm_write_f", "lbdem_reduce_work_doubles", "lbdem_lbm_reduce",
        "lbdem_lbm_reduce_static",
        "lbdem_lbm_prepare_flags", "lbdem_lbm_boundaries", "lbdem_lbm_free_slip", "lbdem_lbm_stress",
        "lbdem_lbm_plane_stress"}
    missing = [n for n in declared_functions() if n not in known]
    assert not missing, f"no ctypes prototype for {missing}"


def test_errors_are_reported_without_a_gpu():
    """Argument validation happens on the host, before any HIP call."""
    import src._hip as H
    L = H.lib()
    lat = H.Lattice()
    lat.nx = 0
    args = H.StepArgs()
    rc = L.lbdem_lbm_step(ctypes.byref(lat), ctypes.byref(args), None)
    assert rc == -1
    assert b"bad extents" in L.lbdem_last_error()


def test_header_constants_match_python_binding():
    """Flag bits, storage / collision / coupling codes: header and ctypes binding agree."""
    import src._hip as H
    text = open(HEADER).read()
    consts = {m.group(1): int(m.group(2), 0) for m in re.finditer(r"#define\s+(LBDEM_[A-Z0-9_]+)\s+(0x[0-9a-fA-F]+|\d+)",
                                                                      text)}
    flags = ("ACTIVE", "NEAR", "WALL", "MOVING", "FLUID", "INLET", "ROW_IDLE")
    for name in flags:
        assert consts[f"LBDEM_FLAG_{name}"] == getattr(H, f"FLAG_{name}"), name
    assert len({consts[f"LBDEM_FLAG_{n}"] for n in flags}) == len(flags)
    for name in ("VEL_ZOUHE", "PRE_ZOUHE", "EXIT", "LADD_INLET"):
        assert consts[f"LBDEM_BC_{name}"] == getattr(H, f"BC_{name}"), name
    for key, code in H.STORE.items():
        assert consts[f"LBDEM_STORE_{key.upper()}"] == code, key
    for name in ("COLL_BGK", "COLL_SMAG_LAGGED", "COLL_SMAG_FIXED", "COUPLE_NONE", "COUPLE_PSC", "COUPLE_EQIBM",
                 "COUPLE_VANS", "COUPLE_HYBRID"):
        assert consts[f"LBDEM_{name}"] == getattr(H, name), name


def test_integration_stub_structs_match_the_binding():
    """The ctypes structs shown in the integration guide have the binding's (and the header's) layout."""
    import ctypes as C
    import src._hip as H
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    ns = {"ctypes": C}
    for name in ("Lattice", "StepArgs"):
        m = re.search(rf"^class {name}\(ctypes\.Structure\):\n(?:    .*\n|               .*\n)+", text, flags=re.M)
        assert m, name
        exec(m.group(0), ns)
        got = [(f, C.sizeof(t)) for f, t in ns[name]._fields_]
        want = [(f, C.sizeof(t)) for f, t in getattr(H, name)._fields_]
        assert got == want, name
        assert C.sizeof(ns[name]) == C.sizeof(getattr(H, name))


def test_reduce_static_rejects_a_lattice_beyond_its_scratch():
    """The allocation-free reduction checks the lattice size on the host, before any HIP call."""
    import src._hip as H
    L = H.lib()
    lat = H.Lattice()
    lat.nx, lat.ny, lat.nz = 512, 256, 257   # 2^25 + 2^17 cells
    lat.omega0 = 1.0
    args = H.StepArgs()
    dummy = ctypes.c_double(0.0)
    args.f = args.flags = args.nmask = ctypes.addressof(dummy)
    out = (ctypes.c_double * 5)()
    rc = L.lbdem_lbm_reduce_static(ctypes.byref(lat), ctypes.byref(args), ctypes.cast(out, ctypes.c_void_p), None)
    assert rc == -1
    assert b"exceed the static scratch" in L.lbdem_last_error()


def test_dem_run_rejects_a_second_grid_it_cannot_use():
    """The second-grid touching search (fine_cell > 0) is checked on the host, before any HIP
    call: it needs the 8-cell probe, no periodic axis, fine_cell > 2 large_r and its buffers."""
    import src._hip as H
    from src.dem3d import _abi
    L = H.lib()
    buf = (ctypes.c_double * 64)()
    addr = ctypes.addressof(buf)
    base = addr + (-addr) % 64

    def call(**kw):
        P = _abi.DemParams()
        P.n, P.K, P.T, P.cell, P.dt = 5000, 8, 1 << 20, 1.0, 1e-3
        P.fine_cell, P.large_r, P.nlarge = 0.3, 0.1, 1
        S = _abi.DemState()
        S.status = base
        w = _abi.DemWork()
        w.pairs = w.need_new = w.npairs = w.fhash = w.large = w.spos = base
        w.count = w.offset = w.current = w.hash = w.pid = w.tile_sum = base
        w.pair_cap = 16
        for k, v in kw.items():
            if k in ("fhash", "large", "spos"):
                setattr(w, k, v)
            elif k == "periodic":
                P.periodic[0] = v
            else:
                setattr(P, k, v)
        return L.lbdem_dem_run(ctypes.byref(P), ctypes.byref(S), ctypes.byref(w), 1, 2, None)

    for kw, msg in ((dict(probe27=1), b"8-cell probe"), (dict(periodic=1), b"8-cell probe"),
                    (dict(fine_cell=0.2), b"exceed 2 large_r"), (dict(fhash=None), b"large-particle list"),
                    (dict(large=None), b"large-particle list"), (dict(nlarge=70000), b"large-particle list")):
        assert call(**kw) == -1, kw
        assert msg in L.lbdem_last_error(), (kw, L.lbdem_last_error())
