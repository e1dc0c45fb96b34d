"""Python host of the gfx950 dispersion engine, mirroring ALPS's call surface.

The reference's public procedures (ALPS_fns.f90:26-29) map onto
:class:`Engine` methods:

=====================================  =====================================
reference                              this module
=====================================  =====================================
``disp(om)``            :252           ``Engine.disp(om)`` / ``disp_batch``
``determine_nmax`` +    :3971, :4214   ``Engine.set_k(kperp, kpar)``
``determine_bessel_array``
``secant_osc(om, in)``  :1919          ``Engine.secant_osc(om)`` (batched)
``refine_guess``        :3793          ``Engine.refine_guess(guesses)``
``map_search``          :3595          ``Engine.map_search(...)``
``om_scan(ik)``         :2198          ``Engine.om_scan(...)``
=====================================  =====================================

Every call goes through the C-ABI of ``libalpsg.so`` (include/alpsg.h); if the
library is missing the import fails: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os

import numpy as np

from .plasma import Plasma, RootParams

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libalpsg.so")

ALPSG_OK = 0
ALPSG_MAXITER = 6  # a served solve ended at numiter without |D| < D_threshold
_ERRNAMES = {1: "ALPSG_E_ARG", 2: "ALPSG_E_UNSUPPORTED", 3: "ALPSG_E_DEVICE",
             4: "ALPSG_E_STATE", 5: "ALPSG_E_NOTFOUND", 6: "ALPSG_MAXITER"}

# every symbol declared in include/alpsg.h
EXPORTS = ["alpsg_last_error", "alpsg_create", "alpsg_destroy", "alpsg_set_k",
           "alpsg_disp_batch", "alpsg_disp_plain", "alpsg_secant_batch", "alpsg_map_search",
           "alpsg_om_scan", "alpsg_get_nmax", "alpsg_eval_count", "alpsg_last_kernel_ms",
           "alpsg_last_root_times", "alpsg_autotune", "alpsg_set_shape", "alpsg_get_shape",
           "alpsg_last_kernel_split", "alpsg_calc_eigen", "alpsg_server_create",
           "alpsg_server_add_plasma", "alpsg_server_remove_plasma", "alpsg_server_submit_k",
           "alpsg_server_submit_batch", "alpsg_om_double_scan",
           "alpsg_server_submit", "alpsg_server_poll", "alpsg_server_stats",
           "alpsg_server_destroy", "alpsg_determine_param_fit", "alpsg_last_scan_steps",
           "alpsg_set_plain_roots", "alpsg_disp_batch_noalloc", "alpsg_server_work",
           "alpsg_set_sweep_series", "alpsg_sweep_work", "alpsg_map_search_refine"]


class AlpsgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRNAMES.get(code, code)}: {msg}")
        self.code = code


class _Desc(C.Structure):
    _fields_ = [
        ("nspec", C.c_int), ("nperp", C.c_int), ("npar", C.c_int),
        ("pp", C.c_void_p), ("f0", C.c_void_p), ("df0", C.c_void_p),
        ("ns", C.c_void_p), ("qs", C.c_void_p), ("ms", C.c_void_p),
        ("vA", C.c_double), ("kperp_norm", C.c_int), ("positions_principal", C.c_int),
        ("n_resonance_interval", C.c_int), ("Tlim", C.c_double), ("Bessel_zero", C.c_double),
        ("usebM", C.c_void_p), ("relativistic", C.c_void_p),
        ("ACmethod", C.c_void_p), ("n_fits", C.c_void_p), ("max_fits", C.c_int),
        ("fit_type", C.c_void_p), ("perp_correction", C.c_void_p), ("param_fit", C.c_void_p),
        ("param_fit_nperp", C.c_int),
        ("poly_kind", C.c_void_p), ("poly_order", C.c_void_p), ("logfit", C.c_void_p),
        ("poly_log_max", C.c_void_p), ("poly_fit_coeffs", C.c_void_p),
        ("poly_max_order", C.c_int),
        ("bM_nmax", C.c_void_p), ("bM_Bessel_zero", C.c_void_p), ("bM_beta", C.c_void_p),
        ("bM_alpha", C.c_void_p), ("bM_pdrift", C.c_void_p),
        ("ngamma", C.c_int), ("npparbar", C.c_int),
        ("gamma_rel", C.c_void_p), ("pparbar_rel", C.c_void_p), ("f0_rel", C.c_void_p),
        ("df0_rel", C.c_void_p),
        ("current_int", C.c_void_p),
    ]


class _RootParams(C.Structure):
    _fields_ = [("secant_method", C.c_int), ("numiter", C.c_int), ("D_threshold", C.c_double),
                ("D_prec", C.c_double), ("D_gap", C.c_double), ("D_tol", C.c_double)]


class _ScanParams(C.Structure):
    _fields_ = [("scan_type", C.c_int), ("log_scan", C.c_int), ("diff", C.c_double),
                ("diff2", C.c_double), ("n_out", C.c_int), ("n_res", C.c_int)]


class _FitParams(C.Structure):
    _fields_ = [("maxsteps_fit", C.c_int), ("lambda_initial_fit", C.c_double),
                ("lambdafac_fit", C.c_double), ("epsilon_fit", C.c_double)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libalpsg.so and declare the C-ABI prototypes (raises if absent)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is not built: run `python -m alps_amd.build`")
    L = C.CDLL(path)
    vp, ip, dp = C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_double)
    L.alpsg_last_error.restype = C.c_char_p
    L.alpsg_create.argtypes = [C.POINTER(_Desc), C.POINTER(vp)]
    L.alpsg_destroy.argtypes = [vp]
    L.alpsg_set_k.argtypes = [vp, C.c_double, C.c_double, ip, ip]
    L.alpsg_disp_batch.argtypes = [vp, C.c_int, dp, dp, dp, dp, dp]
    L.alpsg_disp_plain.argtypes = [vp, C.c_int, dp, dp]
    L.alpsg_disp_batch_noalloc.argtypes = [vp, C.c_int, dp, dp]
    L.alpsg_disp_batch_noalloc.restype = C.c_int
    L.alpsg_secant_batch.argtypes = [vp, C.POINTER(_RootParams), C.c_int, dp, dp, ip]
    L.alpsg_map_search.argtypes = [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                   C.c_int, C.c_int, C.c_int, dp, dp, dp, C.c_int, ip, ip]
    L.alpsg_map_search_refine.argtypes = [vp, C.POINTER(_RootParams), C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                          C.c_int, dp, dp, dp, C.c_int, ip, ip, dp, dp, ip]
    L.alpsg_map_search_refine.restype = C.c_int
    L.alpsg_om_scan.argtypes = [vp, C.POINTER(_RootParams), C.c_int, C.c_int, C.c_double,
                                C.c_double, C.c_int, C.c_int, C.c_int, dp, dp, dp, dp]
    L.alpsg_om_double_scan.argtypes = [vp, C.POINTER(_RootParams), C.POINTER(_ScanParams),
                                       C.POINTER(_ScanParams), C.c_int, dp, C.c_int, C.c_int, dp,
                                       dp, dp, ip, dp, dp, dp, dp, dp, dp, dp]
    L.alpsg_om_double_scan.restype = C.c_int
    L.alpsg_get_nmax.argtypes = [vp, ip]
    L.alpsg_eval_count.argtypes = [vp]
    L.alpsg_eval_count.restype = C.c_longlong
    L.alpsg_last_root_times.argtypes = [vp, C.c_int, dp]
    L.alpsg_last_root_times.restype = C.c_int
    L.alpsg_last_scan_steps.argtypes = [vp, C.c_int, C.c_int, dp, ip]
    L.alpsg_last_scan_steps.restype = C.c_int
    L.alpsg_set_plain_roots.argtypes = [vp, C.c_int]
    L.alpsg_set_plain_roots.restype = C.c_int
    L.alpsg_set_sweep_series.argtypes = [vp, C.c_int]
    L.alpsg_set_sweep_series.restype = C.c_int
    L.alpsg_sweep_work.argtypes = [vp, C.c_int, dp, dp]
    L.alpsg_sweep_work.restype = C.c_int
    L.alpsg_autotune.argtypes = [vp, C.c_int, dp, ip]
    L.alpsg_set_shape.argtypes = [vp, C.c_int, ip]
    L.alpsg_get_shape.argtypes = [vp, C.c_int, ip]
    for name in ["alpsg_autotune", "alpsg_set_shape", "alpsg_get_shape",
           "alpsg_last_kernel_split", "alpsg_calc_eigen"]:
        getattr(L, name).restype = C.c_int
    L.alpsg_server_create.argtypes = [C.c_double, C.c_int, C.POINTER(vp)]
    L.alpsg_server_submit.argtypes = [vp, C.POINTER(_Desc), C.c_double, C.c_double, dp,
                                      C.POINTER(_RootParams), C.c_longlong]
    L.alpsg_server_add_plasma.argtypes = [vp, C.POINTER(_Desc), ip]
    L.alpsg_server_remove_plasma.argtypes = [vp, C.c_int]
    L.alpsg_server_submit_k.argtypes = [vp, C.c_int, C.c_double, C.c_double, dp,
                                        C.POINTER(_RootParams), C.c_longlong]
    L.alpsg_server_submit_batch.argtypes = [vp, C.c_int, ip, dp, dp, dp, C.POINTER(_RootParams),
                                            C.POINTER(C.c_longlong)]
    L.alpsg_server_poll.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_longlong), dp, ip, ip,
                                    dp, ip]
    L.alpsg_server_stats.argtypes = [vp, C.POINTER(C.c_longlong), dp]
    L.alpsg_server_work.argtypes = [vp, dp]
    L.alpsg_server_work.restype = C.c_int
    L.alpsg_server_destroy.argtypes = [vp]
    for name in ["alpsg_server_create", "alpsg_server_submit", "alpsg_server_poll",
                 "alpsg_server_stats", "alpsg_server_destroy", "alpsg_server_add_plasma",
                 "alpsg_server_remove_plasma", "alpsg_server_submit_k",
                 "alpsg_server_submit_batch"]:
        getattr(L, name).restype = C.c_int
    L.alpsg_determine_param_fit.argtypes = [C.POINTER(_Desc), C.POINTER(_FitParams), dp, ip, dp,
                                            dp, dp]
    L.alpsg_determine_param_fit.restype = C.c_int
    L.alpsg_last_kernel_ms.argtypes = [vp]
    L.alpsg_last_kernel_ms.restype = C.c_double
    for name in ["alpsg_create", "alpsg_destroy", "alpsg_set_k", "alpsg_disp_batch",
                 "alpsg_disp_plain", "alpsg_secant_batch", "alpsg_map_search", "alpsg_om_scan",
                 "alpsg_get_nmax"]:
        getattr(L, name).restype = C.c_int
    _lib = L
    return L


def _dp(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_double))


def _ip(a):
    return None if a is None else a.ctypes.data_as(C.POINTER(C.c_int))


def _check(rc: int):
    if rc != ALPSG_OK:
        raise AlpsgError(rc, load_library().alpsg_last_error().decode())


def _rp(r: RootParams) -> _RootParams:
    return _RootParams(r.secant_method, r.numiter, r.D_threshold, r.D_prec, r.D_gap, r.D_tol)


def _make_desc(plasma: Plasma, use_df0: bool = False):
    """The alpsg_plasma_desc of a plasma and the arrays it points into (keep them alive)."""
    p = plasma
    fo = p.fortran()
    keep = dict(fo)
    keep["ns"], keep["qs"], keep["ms"] = p.ns.copy(), p.qs.copy(), p.ms.copy()
    keep["ACmethod"] = p.ACmethod.astype(np.int32)
    keep["n_fits"] = p.n_fits.astype(np.int32)
    keep["poly_kind"] = p.poly_kind.astype(np.int32)
    keep["poly_order"] = p.poly_order.astype(np.int32)
    keep["logfit"] = p.logfit.astype(np.int32)
    keep["poly_log_max"] = p.poly_log_max.astype(np.float64)
    keep["fit_type_f"] = np.asfortranarray(p.fit_type.astype(np.int32))
    keep["usebM"] = p.usebM.astype(np.int32)
    keep["bM_nmax"] = p.bM_nmax.astype(np.int32)
    for name in ("bM_Bessel_zero", "bM_beta", "bM_alpha", "bM_pdrift"):
        keep[name] = getattr(p, name).astype(np.float64)
    if use_df0:
        df0 = np.zeros((p.nspec, p.nperp - 1, p.npar - 1, 2))
        pp, f0 = p.pp, p.f0
        df0[..., 0] = ((f0[:, 2:, 1:-1] - f0[:, :-2, 1:-1]) /
                       (pp[:, 2:, 1:-1, 0] - pp[:, :-2, 1:-1, 0]))
        df0[..., 1] = ((f0[:, 1:-1, 2:] - f0[:, 1:-1, :-2]) /
                       (pp[:, 1:-1, 2:, 1] - pp[:, 1:-1, :-2, 1]))
        keep["df0"] = np.asfortranarray(df0)
    d = _Desc()
    d.nspec, d.nperp, d.npar = p.nspec, p.nperp, p.npar
    d.pp = keep["pp"].ctypes.data
    d.f0 = None if use_df0 else keep["f0"].ctypes.data
    d.df0 = keep["df0"].ctypes.data if use_df0 else None
    d.ns, d.qs, d.ms = keep["ns"].ctypes.data, keep["qs"].ctypes.data, keep["ms"].ctypes.data
    d.vA = p.vA
    d.kperp_norm = int(bool(p.kperp_norm))
    d.positions_principal = int(p.positions_principal)
    d.n_resonance_interval = int(p.n_resonance_interval)
    d.Tlim, d.Bessel_zero = p.Tlim, p.Bessel_zero
    d.usebM = keep["usebM"].ctypes.data
    keep["relativistic"] = p.relativistic.astype(np.int32)
    d.relativistic = keep["relativistic"].ctypes.data
    if p.nspec_rel:
        d.ngamma, d.npparbar = int(p.ngamma), int(p.npparbar)
        d.gamma_rel, d.pparbar_rel = fo["gamma_rel"].ctypes.data, fo["pparbar_rel"].ctypes.data
        d.f0_rel, d.df0_rel = fo["f0_rel"].ctypes.data, fo["df0_rel"].ctypes.data
    d.bM_nmax = keep["bM_nmax"].ctypes.data
    d.bM_Bessel_zero = keep["bM_Bessel_zero"].ctypes.data
    d.bM_beta, d.bM_alpha = keep["bM_beta"].ctypes.data, keep["bM_alpha"].ctypes.data
    d.bM_pdrift = keep["bM_pdrift"].ctypes.data
    d.ACmethod, d.n_fits = keep["ACmethod"].ctypes.data, keep["n_fits"].ctypes.data
    d.max_fits = p.max_fits
    d.fit_type = keep["fit_type_f"].ctypes.data
    d.perp_correction = keep["perp_correction"].ctypes.data
    d.param_fit = keep["param_fit"].ctypes.data
    d.param_fit_nperp = p.param_fit.shape[1] - 1
    d.poly_kind, d.poly_order = keep["poly_kind"].ctypes.data, keep["poly_order"].ctypes.data
    d.logfit, d.poly_log_max = keep["logfit"].ctypes.data, keep["poly_log_max"].ctypes.data
    d.poly_fit_coeffs = keep["poly_fit_coeffs"].ctypes.data
    d.poly_max_order = p.poly_fit_coeffs.shape[2] - 1
    return d, keep


def determine_param_fit(plasma: Plasma, maxsteps_fit: int = 500, lambda_initial_fit: float = 1.0,
                        lambdafac_fit: float = 10.0, epsilon_fit: float = 1e-8):
    """determine_param_fit (ALPS_analyt.f90:365-728) on the device: the
    Levenberg-Marquardt fits from the guesses in ``param_fit[:, 0]`` and the
    Chebyshev coefficients of ACmethod-2 species.  Returns the plasma with the
    fitted ``param_fit``, ``fit_type``, ``perp_correction`` and
    ``poly_fit_coeffs``, and the summed fit quality."""
    L = load_library()
    d, keep = _make_desc(plasma)
    p = plasma
    pf = np.array(p.param_fit, dtype=np.float64, order="F")
    ft = np.array(p.fit_type, dtype=np.int32, order="F")
    pc = np.array(p.perp_correction, dtype=np.float64, order="F")
    poly = np.array(p.poly_fit_coeffs, dtype=np.float64, order="F")
    q = C.c_double(0.0)
    fpar = _FitParams(int(maxsteps_fit), float(lambda_initial_fit), float(lambdafac_fit),
                      float(epsilon_fit))
    _check(L.alpsg_determine_param_fit(C.byref(d), C.byref(fpar), _dp(pf), _ip(ft), _dp(pc),
                                       _dp(poly), C.byref(q)))
    out = dataclasses.replace(p, param_fit=np.ascontiguousarray(pf),
                              fit_type=np.ascontiguousarray(ft),
                              perp_correction=np.ascontiguousarray(pc),
                              poly_fit_coeffs=np.ascontiguousarray(poly))
    return out, q.value


class Engine:
    """One device-resident plasma (a C-ABI handle)."""

    def __init__(self, plasma: Plasma, use_df0: bool = False):
        L = load_library()
        self.plasma = plasma
        d, self._keep = _make_desc(plasma, use_df0)
        h = C.c_void_p()
        _check(L.alpsg_create(C.byref(d), C.byref(h)))
        self._h = h
        self.kperp = self.kpar = None

    def close(self):
        if getattr(self, "_h", None):
            load_library().alpsg_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- wavevector -------------------------------------------------------
    def set_k(self, kperp: float, kpar: float, nmax=None):
        ns = self.plasma.nspec
        out = np.zeros(ns, dtype=np.int32)
        nin = None if nmax is None else np.asarray(nmax, dtype=np.int32)
        _check(load_library().alpsg_set_k(self._h, kperp, kpar, _ip(nin), _ip(out)))
        self.kperp, self.kpar = kperp, kpar
        return [int(x) for x in out]

    @property
    def nmax(self):
        out = np.zeros(self.plasma.nspec, dtype=np.int32)
        _check(load_library().alpsg_get_nmax(self._h, _ip(out)))
        return [int(x) for x in out]

    @property
    def eval_count(self) -> int:
        return int(load_library().alpsg_eval_count(self._h))

    @property
    def last_kernel_ms(self) -> float:
        return float(load_library().alpsg_last_kernel_ms(self._h))

    @property
    def last_kernel_split_ms(self):
        """(k_sweep, k_resonant, k_assemble) device times of the last batch, ms."""
        out = np.zeros(3)
        _check(load_library().alpsg_last_kernel_split(self._h, _dp(out)))
        return tuple(float(x) for x in out)

    # -- disp ---------------------------------------------------------------
    def disp_batch(self, om, want_tensors: bool = False):
        om = np.ascontiguousarray(np.asarray(om, dtype=np.complex128).ravel())
        n = om.size
        det = np.zeros(n, dtype=np.complex128)
        ns = self.plasma.nspec
        chi = low = wave = None
        if want_tensors:
            chi = np.zeros(n * ns * 9, dtype=np.complex128)
            low = np.zeros(n * ns * 27, dtype=np.complex128)
            wave = np.zeros(n * 9, dtype=np.complex128)
        _check(load_library().alpsg_disp_batch(self._h, n, _dp(om), _dp(det), _dp(chi), _dp(low),
                                              _dp(wave)))
        if not want_tensors:
            return det
        # Fortran layouts -> [n][s][i][j], [n][s][i][j][q], [n][i][j]
        chi = chi.reshape(n, 3, 3, ns).transpose(0, 3, 2, 1)
        low = low.reshape(n, 3, 3, 3, ns).transpose(0, 4, 3, 2, 1)
        wave = wave.reshape(n, 3, 3).transpose(0, 2, 1)
        return det, chi, low, wave

    def disp(self, om):
        """det D at one complex frequency (ALPS_fns.f90:252)."""
        return complex(self.disp_batch([om])[0])

    def disp_batch_noalloc(self, om):
        """det D at every omega through alpsg_disp_batch_noalloc (no device
        allocation: the library's fixed per-GPU scratch array)."""
        om = np.ascontiguousarray(np.asarray(om, dtype=np.complex128).ravel())
        det = np.zeros(om.size, dtype=np.complex128)
        _check(load_library().alpsg_disp_batch_noalloc(self._h, om.size, _dp(om), _dp(det)))
        return det

    def disp_plain(self, om):
        om = np.ascontiguousarray(np.asarray(om, dtype=np.complex128).ravel())
        det = np.zeros(om.size, dtype=np.complex128)
        _check(load_library().alpsg_disp_plain(self._h, om.size, _dp(om), _dp(det)))
        return det

    # -- root searches --------------------------------------------------------
    def solve_batch(self, guesses, roots: RootParams | None = None):
        """The root solver roots.secant_method selects (0 secant, 1 rtsec,
        2 secant_osc) on every guess, batched; returns (roots, D at the roots,
        iterations)."""
        r = roots or self.plasma.roots
        om = np.ascontiguousarray(np.asarray(guesses, dtype=np.complex128).ravel()).copy()
        dfin = np.zeros_like(om)
        it = np.zeros(om.size, dtype=np.int32)
        rp = _rp(r)
        _check(load_library().alpsg_secant_batch(self._h, C.byref(rp), om.size, _dp(om), _dp(dfin),
                                                _ip(it)))
        return om, dfin, it

    def secant_osc_batch(self, guesses, roots: RootParams | None = None):
        """secant_osc (ALPS_fns.f90:1919-2101) on every guess, batched."""
        import dataclasses
        r = roots or self.plasma.roots
        return self.solve_batch(guesses, dataclasses.replace(r, secant_method=2))

    def secant_osc(self, om, roots: RootParams | None = None):
        o, d, it = self.secant_osc_batch([om], roots)
        return complex(o[0]), int(it[0])

    # -- launch shapes ----------------------------------------------------------
    SWEEP_VARIANTS = [(1, 1), (1, 2), (1, 4), (2, 4), (4, 4), (8, 4), (16, 4), (16, 8), (32, 8),
                      (64, 4), (64, 8)]
    RESONANT_LANES = [4, 8, 16, 32, 64]

    def autotune(self, n: int, om=None):
        """Time the kernel variants on a batch of n omegas; returns the chosen shape
        (sweep lanes/omega, waves/workgroup, row chunks, resonant lanes/omega, harmonic blocks)."""
        out = np.zeros(5, dtype=np.int32)
        o = None if om is None else np.ascontiguousarray(np.asarray(om, dtype=np.complex128).ravel())
        _check(load_library().alpsg_autotune(self._h, int(n), _dp(o), _ip(out)))
        return tuple(int(x) for x in out)

    def set_shape(self, n: int, shape=None):
        sh = None if shape is None else np.asarray(shape, dtype=np.int32)
        _check(load_library().alpsg_set_shape(self._h, int(n), _ip(sh)))

    def get_shape(self, n: int):
        out = np.zeros(5, dtype=np.int32)
        _check(load_library().alpsg_get_shape(self._h, int(n), _ip(out)))
        return tuple(int(x) for x in out)

    def last_root_times_ms(self, nroots: int):
        """Completion time of each root of the last secant batch (ms from its
        start; after map_search_refine, from the start of that call)."""
        out = np.zeros(nroots)
        _check(load_library().alpsg_last_root_times(self._h, nroots, _dp(out)))
        return out

    def last_scan_steps(self, nsteps: int, nroots: int):
        """The last om_scan per k step: end time (ms from its start) [nsteps]
        and each root's secant iterations [nsteps, nroots] (-1: not followed)."""
        ms = np.zeros(nsteps)
        it = np.zeros(nsteps * nroots, dtype=np.int32)
        _check(load_library().alpsg_last_scan_steps(self._h, nsteps, nroots, _dp(ms), _ip(it)))
        return ms, it.reshape(nsteps, nroots)

    def set_plain_roots(self, plain: bool):
        """Root searches step on the host with one disp launch per omega (the
        plain form) instead of device-stepped graph rounds."""
        _check(load_library().alpsg_set_plain_roots(self._h, int(bool(plain))))

    def set_sweep_series(self, on: bool):
        """Far rows of the one-lane node sweep take their per-k series in m w
        (on, the default) or pass over their nodes (off); both give the
        trapezoid node sums to double rounding."""
        _check(load_library().alpsg_set_sweep_series(self._h, int(bool(on))))

    def sweep_work(self, om):
        """Work of the node sweep over these omegas in the batch's launch shape,
        counted on the device: (node units summed node by node, far-row signs
        summed by series, sum of those series' degrees), each summed over the
        omegas."""
        om = np.ascontiguousarray(np.asarray(om, dtype=np.complex128).ravel())
        out = np.zeros(3)
        _check(load_library().alpsg_sweep_work(self._h, om.size, _dp(om), _dp(out)))
        return tuple(float(x) for x in out)

    def refine_guess(self, guesses, roots: RootParams | None = None):
        """refine_guess (ALPS_fns.f90:3793-3856): secant on each guess, then det D there."""
        om, dfin, _ = self.solve_batch(guesses, roots)
        return om, dfin

    def calc_eigen(self, roots, eigen=True, heat=True):
        """calc_eigen (ALPS_fns.f90:2605-2899) at every root, batched: ef, bf [n, 3],
        Us [n, nspec, 3], ds [n, nspec], Ps [n, nspec], Ps_split [n, nspec, 4], W_EM [n]."""
        om = np.ascontiguousarray(np.asarray(roots, dtype=np.complex128).ravel())
        n, ns = om.size, self.plasma.nspec
        ef = np.zeros((n, 3), dtype=np.complex128)
        bf = np.zeros((n, 3), dtype=np.complex128)
        us = np.zeros((n, ns, 3), dtype=np.complex128)
        ds = np.zeros((n, ns), dtype=np.complex128)
        ps = np.zeros((n, ns))
        psp = np.zeros((n, ns, 4))
        ew = np.zeros(n)
        _check(load_library().alpsg_calc_eigen(self._h, n, _dp(om), int(eigen), int(heat), _dp(ef),
                                               _dp(bf), _dp(us), _dp(ds), _dp(ps), _dp(psp),
                                               _dp(ew)))
        return dict(ef=ef, bf=bf, Us=us, ds=ds, Ps=ps, Ps_split=psp, W_EM=ew)

    def map_search(self, omi, omf, gami, gamf, nr, ni, loggridw=False, loggridg=False,
                   numroots=100, out=None):
        """det D on the nr x ni grid plus the local minima (ALPS_fns.f90:3595-3788).
        Returns om[nr, ni], cal[nr, ni], val[nr, ni], iroots [(ir, ii)], nfound.
        ``out`` = (om, cal, val), flat contiguous arrays of nr*ni complex128,
        complex128 and float64, receives the grids instead of new arrays (a
        caller that maps repeatedly reuses them; the returned grids are views)."""
        N = nr * ni
        if out is None:
            om = np.zeros(N, dtype=np.complex128)
            cal = np.zeros(N, dtype=np.complex128)
            val = np.zeros(N)
        else:
            om, cal, val = out
            for a, dt in ((om, np.complex128), (cal, np.complex128), (val, np.float64)):
                if a.dtype != dt or a.size != N or not a.flags.c_contiguous:
                    raise ValueError("map_search: out arrays must be contiguous with nr*ni "
                                     "elements of complex128, complex128, float64")
        iroots = np.zeros(2 * numroots, dtype=np.int32)
        nf = np.zeros(1, dtype=np.int32)
        _check(load_library().alpsg_map_search(self._h, omi, omf, gami, gamf, nr, ni,
                                              int(loggridw), int(loggridg), _dp(om), _dp(cal),
                                              _dp(val), numroots, _ip(iroots), _ip(nf)))
        n = int(nf[0])
        f = lambda a: a.reshape(ni, nr).T  # Fortran order (ir fastest)
        return f(om), f(cal), f(val), iroots.reshape(numroots, 2)[: min(n, numroots)], n

    def map_search_refine(self, omi, omf, gami, gamf, nr, ni, loggridw=False, loggridg=False,
                          numroots=100, out=None, roots: RootParams | None = None):
        """map_search with refine_guess on its minima (ALPS_fns.f90:3595-3788,
        3793-3856) in one call: as map_search, plus the refined roots of the
        recorded minima, D at them and their iteration counts.  Returns om,
        cal, val, iroots, nfound, roots, D_final, iters."""
        r = roots or self.plasma.roots
        N = nr * ni
        if out is None:
            om = np.zeros(N, dtype=np.complex128)
            cal = np.zeros(N, dtype=np.complex128)
            val = np.zeros(N)
        else:
            om, cal, val = out
            for a, dt in ((om, np.complex128), (cal, np.complex128), (val, np.float64)):
                if a.dtype != dt or a.size != N or not a.flags.c_contiguous:
                    raise ValueError("map_search_refine: out arrays must be contiguous with "
                                     "nr*ni elements of complex128, complex128, float64")
        iroots = np.zeros(2 * numroots, dtype=np.int32)
        nf = np.zeros(1, dtype=np.int32)
        w = np.zeros(numroots, dtype=np.complex128)
        dfin = np.zeros(numroots, dtype=np.complex128)
        it = np.zeros(numroots, dtype=np.int32)
        rp = _rp(r)
        _check(load_library().alpsg_map_search_refine(
            self._h, C.byref(rp), omi, omf, gami, gamf, nr, ni, int(loggridw), int(loggridg),
            _dp(om), _dp(cal), _dp(val), numroots, _ip(iroots), _ip(nf), _dp(w), _dp(dfin),
            _ip(it)))
        n = int(nf[0])
        k = min(n, numroots)
        f = lambda a: a.reshape(ni, nr).T  # Fortran order (ir fastest)
        return (f(om), f(cal), f(val), iroots.reshape(numroots, 2)[:k], n, w[:k], dfin[:k],
                it[:k])

    def om_double_scan(self, s1, s2, roots, eigen=False, heat=False,
                       rparams: RootParams | None = None):
        """om_double_scan (ALPS_fns.f90:2900-3591) from the current k.  s1, s2:
        (scan_type, log_scan, diff, diff2, n_out, n_res) of the outer and inner
        scans.  Returns a dict: kperp, kpar [n1+1, n2+1]; roots and written
        [n1+1, n2+1, nroots]; last_roots (along scan 1 at its end); with
        eigen/heat the calc_eigen outputs per (point, root)."""
        r = rparams or self.plasma.roots
        roots = np.ascontiguousarray(np.asarray(roots, dtype=np.complex128).ravel()).copy()
        nr, ns = roots.size, self.plasma.nspec
        n1, n2 = int(s1[4]) + 1, int(s2[4]) + 1
        rows = n1 * n2
        kp, kz = np.zeros(rows), np.zeros(rows)
        out = np.zeros(rows * nr, dtype=np.complex128)
        wr = np.zeros(rows * nr, dtype=np.int32)
        want = eigen or heat
        ef = np.zeros(rows * nr * 3, dtype=np.complex128) if want else None
        bf = np.zeros(rows * nr * 3, dtype=np.complex128) if want else None
        us = np.zeros(rows * nr * 3 * ns, dtype=np.complex128) if eigen else None
        ds = np.zeros(rows * nr * ns, dtype=np.complex128) if eigen else None
        ps = np.zeros(rows * nr * ns) if heat else None
        psp = np.zeros(rows * nr * 4 * ns) if heat else None
        ew = np.zeros(rows * nr) if heat else None
        a, b = _ScanParams(*s1), _ScanParams(*s2)
        _check(load_library().alpsg_om_double_scan(
            self._h, C.byref(_rp(r)), C.byref(a), C.byref(b), nr, _dp(roots), int(bool(eigen)),
            int(bool(heat)), _dp(kp), _dp(kz), _dp(out), _ip(wr), _dp(ef), _dp(bf), _dp(us),
            _dp(ds), _dp(ps), _dp(psp), _dp(ew)))
        sh = (n1, n2, nr)
        res = dict(kperp=kp.reshape(n1, n2), kpar=kz.reshape(n1, n2), roots=out.reshape(sh),
                   written=wr.reshape(sh).astype(bool), last_roots=roots)
        if want:
            res.update(ef=ef.reshape(sh + (3,)), bf=bf.reshape(sh + (3,)))
        if eigen:
            res.update(Us=us.reshape(sh + (ns, 3)), ds=ds.reshape(sh + (ns,)))
        if heat:
            res.update(Ps=ps.reshape(sh + (ns,)), Ps_split=psp.reshape(sh + (ns, 4)),
                       W_EM=ew.reshape(sh))
        return res

    def om_scan(self, scan_type, log_scan, diff, diff2, n_out, n_res, roots,
                rparams: RootParams | None = None):
        r = rparams or self.plasma.roots
        roots = np.ascontiguousarray(np.asarray(roots, dtype=np.complex128).ravel()).copy()
        nr = roots.size
        kp = np.zeros(n_out + 1)
        kz = np.zeros(n_out + 1)
        out = np.zeros((n_out + 1) * nr, dtype=np.complex128)
        rp = _rp(r)
        _check(load_library().alpsg_om_scan(self._h, C.byref(rp), scan_type, int(log_scan), diff,
                                           diff2, n_out, n_res, nr, _dp(roots), _dp(kp), _dp(kz),
                                           _dp(out)))
        return kp, kz, out.reshape(n_out + 1, nr)


class RootServer:
    """Continuous-batching root server (alpsg_server_* in include/alpsg.h).

    Each request is one secant_osc solve (ALPS_fns.f90:1919-2101) of a plasma
    at its own wavevector.  ``add_plasma`` registers a plasma once (its device
    tables are then shared by every request naming it); ``submit`` queues a
    solve on a registered plasma id, or on a :class:`Plasma` given inline, and
    returns at once; ``poll`` returns finished solves as ``(tag, root, iters,
    status, latency_ms)`` tuples.  The service thread inside the library
    batches every admitted solve into shared device launches.
    """

    def __init__(self, hbm_bytes: float = 8e9, max_slots: int = 4096):
        L = load_library()
        h = C.c_void_p()
        _check(L.alpsg_server_create(float(hbm_bytes), int(max_slots), C.byref(h)))
        self._h = h
        self._descs = {}

    def _desc(self, plasma: Plasma):
        key = id(plasma)
        hit = self._descs.get(key)
        if hit is None or hit[0] is not plasma:
            d, keep = _make_desc(plasma)
            hit = (plasma, d, keep)
            self._descs[key] = hit
        return hit[1]

    def add_plasma(self, plasma: Plasma) -> int:
        d, keep = _make_desc(plasma)
        pid = C.c_int(0)
        _check(load_library().alpsg_server_add_plasma(self._h, C.byref(d), C.byref(pid)))
        return pid.value

    def remove_plasma(self, plasma_id: int):
        _check(load_library().alpsg_server_remove_plasma(self._h, int(plasma_id)))

    def submit(self, plasma, kperp: float, kpar: float, guess: complex,
               roots: RootParams | None = None, tag: int = 0):
        """Queue one solve; ``plasma`` is a registered id or a :class:`Plasma`."""
        rp = roots or (plasma.roots if isinstance(plasma, Plasma) else RootParams(secant_method=2))
        om = np.array([complex(guess).real, complex(guess).imag], dtype=np.float64)
        L = load_library()
        if isinstance(plasma, (int, np.integer)):
            _check(L.alpsg_server_submit_k(self._h, int(plasma), float(kperp), float(kpar),
                                           _dp(om), C.byref(_rp(rp)), int(tag)))
        else:
            _check(L.alpsg_server_submit(self._h, C.byref(self._desc(plasma)), float(kperp),
                                         float(kpar), _dp(om), C.byref(_rp(rp)), int(tag)))

    def submit_batch(self, plasma_ids, kperp, kpar, guesses, roots: RootParams, tags):
        """Queue len(tags) solves on registered plasmas in one call."""
        ids = np.ascontiguousarray(plasma_ids, dtype=np.int32)
        kp = np.ascontiguousarray(kperp, dtype=np.float64)
        kz = np.ascontiguousarray(kpar, dtype=np.float64)
        g = np.asarray(guesses, dtype=np.complex128)
        om = np.ascontiguousarray(np.stack([g.real, g.imag], axis=-1).ravel())
        tg = np.ascontiguousarray(tags, dtype=np.int64)
        n = tg.size
        if not (ids.size == kp.size == kz.size == g.size == n):
            raise ValueError("submit_batch: arrays of different lengths")
        _check(load_library().alpsg_server_submit_batch(
            self._h, int(n), _ip(ids), _dp(kp), _dp(kz), _dp(om), C.byref(_rp(roots)),
            tg.ctypes.data_as(C.POINTER(C.c_longlong))))

    def poll(self, max_results: int = 1024, wait: bool = True):
        n = int(max_results)
        tags = np.zeros(n, dtype=np.int64)
        roots = np.zeros(2 * n)
        iters = np.zeros(n, dtype=np.int32)
        status = np.zeros(n, dtype=np.int32)
        lat = np.zeros(n)
        got = C.c_int(0)
        _check(load_library().alpsg_server_poll(
            self._h, n, int(bool(wait)), tags.ctypes.data_as(C.POINTER(C.c_longlong)), _dp(roots),
            _ip(iters), _ip(status), _dp(lat), C.byref(got)))
        k = got.value
        return [(int(tags[i]), complex(roots[2 * i], roots[2 * i + 1]), int(iters[i]),
                 int(status[i]), float(lat[i])) for i in range(k)]

    def drain(self, count: int):
        """Poll until ``count`` results have arrived; returns them keyed by tag."""
        out = {}
        while len(out) < count:
            got = self.poll(max(count - len(out), 1), wait=True)
            if not got:
                st = self.stats()
                if st["finished"] >= st["submitted"]:
                    raise RuntimeError(f"drain({count}): only {len(out)} results outstanding")
            for r in got:
                out[r[0]] = r
        return out

    def stats(self):
        c = np.zeros(10, dtype=np.int64)
        t = np.zeros(4)
        _check(load_library().alpsg_server_stats(
            self._h, c.ctypes.data_as(C.POINTER(C.c_longlong)), _dp(t)))
        keys = ("submitted", "finished", "passes", "evals", "admitted", "waiting",
                "arena_bytes", "arena_used_bytes", "plasma_uploads", "plasma_evictions")
        out = {k: int(v) for k, v in zip(keys, c)}
        out.update(pass_ms=t[0], admit_ms=t[1], launch_ms=t[2], wait_ms=t[3])
        w = np.zeros(7)
        _check(load_library().alpsg_server_work(self._h, _dp(w)))
        out.update(work_nodes=float(w[0]), work_rel_points=float(w[1]), work_rel_rows=float(w[2]),
                   sweep_ms=float(w[3]), resonant_ms=float(w[4]), rel_ms=float(w[5]),
                   assemble_ms=float(w[6]))
        return out

    def close(self):
        if getattr(self, "_h", None):
            load_library().alpsg_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
