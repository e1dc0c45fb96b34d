"""CPU-side checks of the C-ABI library: it loads, exports every entry point
declared in include/alpsg.h, and rejects bad input before touching a device."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import cases

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "alpsg.h")


def declared_symbols():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(alpsg_[a-z_0-9]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    from alps_amd import engine
    lib = engine.load_library()
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(lib, s), s
    assert sorted(engine.EXPORTS) == syms
    out = subprocess.run(["nm", "-D", "--defined-only", engine.LIB_PATH], capture_output=True,
                         text=True).stdout
    for s in syms:
        assert re.search(rf"\bT {s}\b", out), s


def test_library_has_no_unresolved_internal_symbols():
    """Every symbol the library needs resolves at load time (RTLD_NOW), so a
    launcher missing from one translation unit fails here, not on a GPU."""
    from alps_amd import engine
    C.CDLL(engine.LIB_PATH, mode=os.RTLD_NOW)
    out = subprocess.run(["nm", "-D", "--undefined-only", engine.LIB_PATH], capture_output=True,
                         text=True).stdout
    assert not re.search(r"\blaunch_\w+|_Z\d+launch_", out), out


def test_library_holds_gfx950_code():
    from alps_amd import engine
    data = open(engine.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def _desc_for(p, **over):
    from alps_amd.engine import _Desc
    fo = p.fortran()
    keep = dict(fo, ns=p.ns.copy(), qs=p.qs.copy(), ms=p.ms.copy(),
                ac=p.ACmethod.astype(np.int32), nf=p.n_fits.astype(np.int32),
                ft=np.asfortranarray(p.fit_type.astype(np.int32)))
    d = _Desc()
    d.nspec, d.nperp, d.npar = p.nspec, p.nperp, p.npar
    d.pp = keep["pp"].ctypes.data
    d.f0 = keep["f0"].ctypes.data
    d.ns, d.qs, d.ms = keep["ns"].ctypes.data, keep["qs"].ctypes.data, keep["ms"].ctypes.data
    d.vA, d.kperp_norm, d.positions_principal = p.vA, 1, p.positions_principal
    d.n_resonance_interval, d.Tlim, d.Bessel_zero = 100, 0.01, 1e-45
    d.ACmethod, d.n_fits, d.max_fits = keep["ac"].ctypes.data, keep["nf"].ctypes.data, 1
    d.fit_type = keep["ft"].ctypes.data
    d.perp_correction = keep["perp_correction"].ctypes.data
    d.param_fit = keep["param_fit"].ctypes.data
    d.param_fit_nperp = p.nperp
    for k, v in over.items():
        if isinstance(v, np.ndarray):
            keep[k] = v
            v = v.ctypes.data
        setattr(d, k, v)
    return d, keep


def _create(d):
    from alps_amd.engine import load_library
    L = load_library()
    h = C.c_void_p()
    rc = L.alpsg_create(C.byref(d), C.byref(h))
    return rc, L.alpsg_last_error().decode(), h


def test_create_rejects_bad_input_without_a_device():
    p = cases.small_bimax()
    d, keep = _desc_for(p, nspec=0)
    rc, msg, h = _create(d)
    assert rc == 1 and "nspec" in msg and not h.value
    # a use_bM species needs its &bM_spec parameters
    d, keep = _desc_for(p, usebM=np.array([1, 0], dtype=np.int32))
    rc, msg, _ = _create(d)
    assert rc == 1 and "bM_spec" in msg
    # cold use_bM species need kperp_norm (calc_chi_cold stops otherwise)
    bm = dict(usebM=np.array([1, 0], dtype=np.int32), bM_nmax=np.array([500, 500], dtype=np.int32),
              bM_Bessel_zero=np.array([1e-50, 1e-50]), bM_beta=np.array([0.0, 1.0]),
              bM_alpha=np.array([1.0, 1.0]), bM_pdrift=np.array([0.0, 0.0]))
    d, keep = _desc_for(p, kperp_norm=0, **bm)
    rc, msg, _ = _create(d)
    assert rc == 2 and "cold" in msg
    # a relativistic species needs its (Gamma, pparbar) tables and Juettner fits
    d, keep = _desc_for(p, relativistic=np.array([0, 1], dtype=np.int32))
    rc, msg, _ = _create(d)
    assert rc == 1 and "relativistic tables" in msg
    g = np.zeros((1, 11, 41))
    d, keep = _desc_for(p, relativistic=np.array([0, 1], dtype=np.int32), ngamma=10, npparbar=40,
                        gamma_rel=g, pparbar_rel=g.copy(), f0_rel=g.copy(),
                        df0_rel=np.zeros((1, 11, 41, 2)))
    rc, msg, _ = _create(d)
    assert rc == 2 and "fit_type 4 or 5" in msg
    # a grid that is not a tensor product of pperp and ppar
    pp = p.pp.copy()
    pp[0, 3, 5, 1] += 1e-3
    d, keep = _desc_for(p, pp=np.asfortranarray(pp))
    rc, msg, _ = _create(d)
    assert rc == 2 and "tensor product" in msg
    # positions_principal beyond what the principal-value window holds
    d, keep = _desc_for(p, positions_principal=30)
    rc, msg, _ = _create(d)
    assert rc == 2 and "positions_principal" in msg
    # positions_principal too large for the grid (npar < 2P + 8)
    p30 = cases.small_bimax(npar=30)
    d, keep = _desc_for(p30, positions_principal=12)
    rc, msg, _ = _create(d)
    assert rc == 1 and "grid too small" in msg


def test_null_handle_calls_fail_cleanly():
    from alps_amd.engine import load_library
    L = load_library()
    assert L.alpsg_set_k(None, 0.1, 0.1, None, None) == 1
    assert L.alpsg_disp_batch(None, 1, None, None, None, None, None) == 1
    assert L.alpsg_eval_count(None) == -1
    assert L.alpsg_autotune(None, 16, None, None) == 1
    assert L.alpsg_set_shape(None, 16, None) == 1
    assert L.alpsg_destroy(None) == 0
    assert L.alpsg_determine_param_fit(None, None, None, None, None, None, None) == 1


def test_fit_setup_rejects_bad_input_without_a_device():
    import ctypes as C
    from alps_amd.engine import load_library
    L = load_library()
    p = cases.small_bimax()
    pf = np.zeros(p.param_fit.size)
    ft = np.ones(p.nspec, dtype=np.int32)
    pc = np.ones(p.nspec)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))
    ip = lambda a: a.ctypes.data_as(C.POINTER(C.c_int))
    # f0 is needed (a description with df0 only cannot be fitted)
    d, keep = _desc_for(p, f0=None)
    assert L.alpsg_determine_param_fit(C.byref(d), None, dp(pf), ip(ft), dp(pc), None, None) == 1
    assert "f0" in L.alpsg_last_error().decode()
    # a fit type outside 1..6
    d, keep = _desc_for(p, fit_type=np.asfortranarray(np.array([[7], [1]], dtype=np.int32)))
    assert L.alpsg_determine_param_fit(C.byref(d), None, dp(pf), ip(ft), dp(pc), None, None) == 1
    assert "fit_type" in L.alpsg_last_error().decode()


def test_fortran_binding_compiles():
    from alps_amd import build
    exe = build.build_fortran()
    if exe is None:
        pytest.skip("no Fortran compiler in this image")
    assert os.path.exists(exe)
