"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py to check and time the HIP path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")

MAXSPEC = 8
MAXFITS = 4


class _Plasma(C.Structure):
    _fields_ = [
        ("nspec", C.c_int), ("nperp", C.c_int), ("npar", C.c_int),
        ("pp", C.POINTER(C.c_double)), ("f0", C.POINTER(C.c_double)),
        ("df0", C.POINTER(C.c_double)),
        ("ns", C.c_double * MAXSPEC), ("qs", C.c_double * MAXSPEC), ("ms", C.c_double * MAXSPEC),
        ("vA", C.c_double),
        ("kperp_norm", C.c_int), ("positions_principal", C.c_int),
        ("n_resonance_interval", C.c_int), ("Tlim", C.c_double), ("Bessel_zero", C.c_double),
        ("ACmethod", C.c_int * MAXSPEC), ("n_fits", C.c_int * MAXSPEC),
        ("fit_type", (C.c_int * MAXFITS) * MAXSPEC),
        ("perp_correction", (C.c_double * MAXFITS) * MAXSPEC),
        ("param_fit", C.POINTER(C.c_double)),
        ("poly_kind", C.c_int * MAXSPEC), ("poly_order", C.c_int * MAXSPEC),
        ("logfit", C.c_int * MAXSPEC), ("poly_log_max", C.c_double * MAXSPEC),
        ("poly_fit_coeffs", C.POINTER(C.c_double)), ("poly_stride", C.c_int),
        ("kperp", C.c_double), ("kpar", C.c_double),
        ("nmax", C.c_int * MAXSPEC),
        ("bessel", C.POINTER(C.c_double) * MAXSPEC),
        ("numiter", C.c_int), ("D_threshold", C.c_double), ("D_prec", C.c_double),
        ("D_gap", C.c_double), ("D_tol", C.c_double), ("secant_method", C.c_int),
        ("nproc", C.c_int),
        ("n_disp_calls", C.c_longlong), ("n_threads", C.c_int),
        ("usebM", C.c_int * MAXSPEC), ("bM_nmax", C.c_int * MAXSPEC),
        ("bM_Bessel_zero", C.c_double * MAXSPEC), ("bM_beta", C.c_double * MAXSPEC),
        ("bM_alpha", C.c_double * MAXSPEC), ("bM_pdrift", C.c_double * MAXSPEC),
        ("relativistic", C.c_int * MAXSPEC), ("ngamma", C.c_int), ("npparbar", C.c_int),
        ("param_np", C.c_int),
        ("gamma_rel", C.POINTER(C.c_double) * MAXSPEC), ("pparbar_rel", C.POINTER(C.c_double) * MAXSPEC),
        ("f0_rel", C.POINTER(C.c_double) * MAXSPEC), ("df0_rel", C.POINTER(C.c_double) * MAXSPEC),
        ("current_int", C.c_double * MAXSPEC),
    ]


class _C128(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


class _RootCtl(C.Structure):
    _fields_ = [("numiter", C.c_int), ("D_threshold", C.c_double), ("D_prec", C.c_double),
                ("D_tol", C.c_double)]


class _TestCtx(C.Structure):
    _fields_ = [("kind", C.c_int), ("a", C.c_double * 2), ("r1", C.c_double * 2),
                ("r2", C.c_double * 2), ("pole", C.c_double * 2), ("n", C.c_int),
                ("cap", C.c_int), ("traj", C.POINTER(C.c_double))]


TEST_KINDS = {"linear": 0, "quadratic": 1, "rational": 2, "exp": 3}
SOLVERS = {"secant": 0, "rtsec": 1, "secant_osc": 2}


def solve_test(method, kind, om, a=1.0, r1=0.0, r2=0.0, pole=0.0, numiter=50,
               D_threshold=1e-12, D_prec=1e-5, D_tol=1e-12, cap=4096):
    """One root solve of a closed-form test function by the restated secant,
    rtsec or secant_osc (orc_solve_test): kind "linear" a (w - r1),
    "quadratic" a (w^2 - r2), "rational" a (w - r1)(w - r2)/(w - pole), "exp"
    a exp(w).  Returns (root, iterations, the arguments of every function
    call in order)."""
    t = _TestCtx()
    t.kind = TEST_KINDS[kind]
    for name, v in (("a", a), ("r1", r1), ("r2", r2), ("pole", pole)):
        getattr(t, name)[0], getattr(t, name)[1] = complex(v).real, complex(v).imag
    buf = np.zeros(2 * cap)
    t.cap, t.traj = cap, _ptr(buf)
    c = _RootCtl(numiter, D_threshold, D_prec, D_tol)
    w = np.array([complex(om).real, complex(om).imag])
    it = lib().orc_solve_test(SOLVERS[method], C.byref(t), C.byref(c), _ptr(w))
    n = min(t.n, cap)
    return complex(w[0], w[1]), it, buf[: 2 * n:2] + 1j * buf[1: 2 * n:2]


def build(force: bool = False) -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True, capture_output=True)
    return _LIB


_lib = None


class OrcScan(C.Structure):
    """orc_scan: one &scan_input scan (type, log, diff, diff2, n_out, n_res)."""
    _fields_ = [("type", C.c_int), ("log_scan", C.c_int), ("diff", C.c_double),
                ("diff2", C.c_double), ("n_out", C.c_int), ("n_res", C.c_int)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        P = C.POINTER(_Plasma)
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.orc_bessj.restype = C.c_double
        L.orc_bessj.argtypes = [C.c_int, C.c_double]
        L.orc_bessj0.restype = C.c_double
        L.orc_bessj0.argtypes = [C.c_double]
        L.orc_bessj1.restype = C.c_double
        L.orc_bessj1.argtypes = [C.c_double]
        L.orc_derivative_f0.argtypes = [P]
        L.orc_determine_param_fit.restype = C.c_int
        L.orc_determine_param_fit.argtypes = [P, C.c_int, C.c_double, C.c_double, C.c_double, dp, ip]
        L.orc_set_k.argtypes = [P, C.c_double, C.c_double]
        L.orc_set_k.restype = C.c_int
        L.orc_free_k.argtypes = [P]
        L.orc_disp.restype = _C128
        L.orc_disp.argtypes = [P, _C128, dp, dp, dp]
        L.orc_full_integrate.restype = _C128
        L.orc_full_integrate.argtypes = [P, C.c_int, _C128, C.c_int, C.c_int, C.c_int]
        L.orc_landau_integrate.restype = _C128
        L.orc_landau_integrate.argtypes = [P, C.c_int, _C128, C.c_int, C.c_int]
        L.orc_integrate_res.restype = _C128
        L.orc_integrate_res.argtypes = [P, C.c_int, _C128, C.c_int, C.c_int]
        L.orc_int_ee.restype = _C128
        L.orc_int_ee.argtypes = [P, C.c_int]
        L.orc_eval_fit.restype = _C128
        L.orc_eval_fit.argtypes = [P, C.c_int, C.c_int, _C128]
        L.orc_determine_resonances.argtypes = [P, C.c_int, _C128, C.c_int, ip, ip]
        L.orc_secant_osc.restype = C.c_int
        L.orc_secant_osc.argtypes = [P, dp]
        L.orc_secant.restype = C.c_int
        L.orc_secant.argtypes = [P, dp]
        L.orc_bessi.restype = C.c_double
        L.orc_bessi.argtypes = [C.c_int, C.c_double]
        L.orc_wofz.restype = None
        L.orc_wofz.argtypes = [C.c_double, C.c_double, dp, dp, C.POINTER(C.c_int)]
        L.orc_nhds_nmaxrun.restype = C.c_int
        L.orc_nhds_nmaxrun.argtypes = [P, C.c_int, C.c_double]
        L.orc_calc_chi.restype = C.c_int
        L.orc_calc_chi.argtypes = [P, C.c_int, C.c_double, C.c_double, _C128, dp, dp]
        L.orc_gamma_fn.restype = C.c_double
        L.orc_gamma_fn.argtypes = [C.c_double]
        L.orc_cbessj.restype = _C128
        L.orc_cbessj.argtypes = [_C128, C.c_int]
        L.orc_int_ee_rel.restype = _C128
        L.orc_int_ee_rel.argtypes = [P, C.c_int]
        L.orc_current_int.restype = None
        L.orc_current_int.argtypes = [P]
        L.orc_calc_eigen.restype = None
        L.orc_calc_eigen.argtypes = [P, _C128, dp, dp, C.c_int, C.c_int, dp, dp, dp, dp, dp, dp, dp]
        L.orc_solve_test.argtypes = [C.c_int, C.POINTER(_TestCtx), C.POINTER(_RootCtl), dp]
        L.orc_solve_test.restype = C.c_int
        L.orc_rtsec.restype = _C128
        L.orc_rtsec.argtypes = [P, _C128, C.POINTER(C.c_int)]
        L.orc_map_search.restype = C.c_int
        L.orc_map_search.argtypes = [P, C.c_double, C.c_double, C.c_double, C.c_double,
                                     C.c_int, C.c_int, C.c_int, C.c_int, dp, dp, dp, C.c_int, ip]
        L.orc_find_minima.restype = C.c_int
        L.orc_find_minima.argtypes = [dp, C.c_int, C.c_int, C.c_int, ip]
        L.orc_om_scan.restype = C.c_int
        L.orc_om_scan.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int,
                                  C.c_int, C.c_int, dp, dp, dp, dp]
        L.orc_om_double_scan.restype = C.c_int
        L.orc_om_double_scan.argtypes = [P, C.POINTER(OrcScan), C.POINTER(OrcScan), C.c_int, dp,
                                         C.c_int, C.c_int, dp, dp, dp, ip, dp, dp, dp, dp, dp,
                                         dp, dp]
        _lib = L
    return _lib


BRANCHES = ["nonres", "res_pos", "res_neg", "res_real", "edge_low", "edge_high", "pv_tlim",
            "pv_analytic", "tiny_rest", "landau_zero", "rel_res", "rel_pv_tlim", "rel_pv_analytic",
            "rel_tiny", "rel_edge", "rel_landau"]


def bessi(n: int, x: float) -> float:
    """BESSI of ALPS_NHDS.f90: I_n(x) exp(-x)."""
    return lib().orc_bessi(n, x)


def wofz(z) -> complex:
    """WOFZ of ALPS_NHDS.f90: the Faddeeva function w(z)."""
    u, v, f = C.c_double(), C.c_double(), C.c_int()
    lib().orc_wofz(complex(z).real, complex(z).imag, C.byref(u), C.byref(v), C.byref(f))
    if f.value:
        raise OverflowError("WOFZ overflow flag")
    return complex(u.value, v.value)


def gamma_fn(x: float) -> float:
    """Gamma of ALPS_fns_rel.f90 (Lanczos series)."""
    return lib().orc_gamma_fn(x)


def cbessj(z, nu: int) -> complex:
    """CBESSJ of ALPS_fns_rel.f90: J_nu(z) from 21 terms of its power series."""
    return _z(lib().orc_cbessj(_c(z), nu))


def rel_error() -> int:
    """8 once a relativistic row hit the reference's alps_error(8) (sticky)."""
    return C.c_int.in_dll(lib(), "orc_rel_error").value


def reset_rel_error() -> None:
    C.c_int.in_dll(lib(), "orc_rel_error").value = 0


def reset_branch_counts():
    lib().orc_reset_branch_counts()


def branch_counts() -> dict:
    arr = (C.c_longlong * len(BRANCHES)).in_dll(lib(), "orc_branch_counts")
    return {name: int(arr[i]) for i, name in enumerate(BRANCHES)}


def _c(z) -> _C128:
    z = complex(z)
    return _C128(z.real, z.imag)


def _z(c: _C128) -> complex:
    return complex(c.re, c.im)


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """CPU restatement of ALPS's disp()/secant/map_search/om_scan for one plasma."""

    def __init__(self, plasma, nproc: int = 1, n_threads: int = 0):
        L = lib()
        p = plasma
        if p.nspec > MAXSPEC or p.max_fits > MAXFITS:
            raise ValueError("too many species or fits for the oracle")
        self.plasma = p
        self._pp = np.ascontiguousarray(p.pp, dtype=np.float64)
        self._f0 = np.ascontiguousarray(p.f0, dtype=np.float64)
        self._df0 = np.zeros(self._pp.shape)
        pf = np.zeros((p.nspec, p.param_fit.shape[1], 5, MAXFITS))
        pf[..., : p.max_fits] = p.param_fit
        self._pf = pf
        self._poly = np.ascontiguousarray(p.poly_fit_coeffs, dtype=np.float64)
        s = _Plasma()
        s.nspec, s.nperp, s.npar = p.nspec, p.nperp, p.npar
        s.pp, s.f0, s.df0 = _ptr(self._pp), _ptr(self._f0), _ptr(self._df0)
        for i in range(p.nspec):
            s.ns[i], s.qs[i], s.ms[i] = p.ns[i], p.qs[i], p.ms[i]
            s.ACmethod[i] = int(p.ACmethod[i])
            s.n_fits[i] = int(p.n_fits[i])
            for f in range(p.max_fits):
                s.fit_type[i][f] = int(p.fit_type[i, f])
                s.perp_correction[i][f] = float(p.perp_correction[i, f])
            s.poly_kind[i] = int(p.poly_kind[i])
            s.poly_order[i] = int(p.poly_order[i])
            s.logfit[i] = int(p.logfit[i])
            s.poly_log_max[i] = float(p.poly_log_max[i])
            s.usebM[i] = int(p.usebM[i])
            s.bM_nmax[i] = int(p.bM_nmax[i])
            s.bM_Bessel_zero[i] = float(p.bM_Bessel_zero[i])
            s.bM_beta[i] = float(p.bM_beta[i])
            s.bM_alpha[i] = float(p.bM_alpha[i])
            s.bM_pdrift[i] = float(p.bM_pdrift[i])
        s.vA = p.vA
        s.kperp_norm = int(bool(p.kperp_norm))
        s.positions_principal = int(p.positions_principal)
        s.n_resonance_interval = int(p.n_resonance_interval)
        s.Tlim = p.Tlim
        s.Bessel_zero = p.Bessel_zero
        s.param_fit = _ptr(self._pf)
        s.poly_fit_coeffs = _ptr(self._poly)
        s.poly_stride = self._poly.shape[2]
        r = p.roots
        s.numiter, s.D_threshold, s.D_prec = r.numiter, r.D_threshold, r.D_prec
        s.D_gap, s.D_tol, s.secant_method = r.D_gap, r.D_tol, r.secant_method
        s.nproc = nproc
        s.param_np = p.param_fit.shape[1] - 1
        s.ngamma, s.npparbar = int(p.ngamma), int(p.npparbar)
        self._rel = []
        r = 0
        for i in range(p.nspec):
            s.relativistic[i] = int(p.relativistic[i])
            if p.relativistic[i]:
                arrs = [np.ascontiguousarray(a[r]) for a in
                        (p.gamma_rel, p.pparbar_rel, p.f0_rel, p.df0_rel)]
                self._rel.append(arrs)
                s.gamma_rel[i], s.pparbar_rel[i], s.f0_rel[i], s.df0_rel[i] = [_ptr(a) for a in arrs]
                r += 1
        s.n_threads = n_threads
        self.s = s
        L.orc_derivative_f0(C.byref(s))
        L.orc_current_int(C.byref(s))

    def __del__(self):
        try:
            lib().orc_free_k(C.byref(self.s))
        except Exception:
            pass

    @property
    def df0(self) -> np.ndarray:
        return self._df0

    @property
    def nmax(self):
        return [self.s.nmax[i] for i in range(self.plasma.nspec)]

    @property
    def n_disp_calls(self) -> int:
        return self.s.n_disp_calls

    def set_k(self, kperp: float, kpar: float):
        rc = lib().orc_set_k(C.byref(self.s), kperp, kpar)
        if rc != 0:
            raise RuntimeError(f"orc_set_k failed ({rc})")
        return self.nmax

    def bessel_table(self, s: int) -> np.ndarray:
        """bessel_array(-1:nmax+1, 0:nperp) of species s (0-based) as [nmax+3, nperp+1]."""
        n = (self.s.nmax[s] + 3) * (self.plasma.nperp + 1)
        a = np.ctypeslib.as_array(self.s.bessel[s], shape=(n,)).copy()
        return a.reshape(self.s.nmax[s] + 3, self.plasma.nperp + 1)

    def disp(self, om, want_tensors: bool = False):
        ns = self.plasma.nspec
        if want_tensors:
            chi0 = np.zeros(ns * 9, dtype=np.complex128)
            low = np.zeros(ns * 27, dtype=np.complex128)
            wave = np.zeros(9, dtype=np.complex128)
            d = lib().orc_disp(C.byref(self.s), _c(om), _ptr(chi0), _ptr(low), _ptr(wave))
            return _z(d), chi0.reshape(ns, 3, 3), low.reshape(ns, 3, 3, 3), wave.reshape(3, 3)
        return _z(lib().orc_disp(C.byref(self.s), _c(om), None, None, None))

    def full_integrate(self, sproc: int, om, nn: int, mode: int, found: bool) -> complex:
        return _z(lib().orc_full_integrate(C.byref(self.s), sproc, _c(om), nn, mode, int(found)))

    def landau(self, sproc: int, om, nn: int, mode: int) -> complex:
        return _z(lib().orc_landau_integrate(C.byref(self.s), sproc, _c(om), nn, mode))

    def int_ee(self, sproc: int) -> complex:
        return _z(lib().orc_int_ee(C.byref(self.s), sproc))

    def eval_fit(self, sproc: int, iperp: int, p) -> complex:
        return _z(lib().orc_eval_fit(C.byref(self.s), sproc, iperp, _c(p)))

    def resonances(self, sproc: int, om, nn: int):
        a, b = C.c_int(), C.c_int()
        lib().orc_determine_resonances(C.byref(self.s), sproc, _c(om), nn, C.byref(a), C.byref(b))
        return bool(a.value), bool(b.value)

    def secant_osc(self, om):
        z = np.array([complex(om)], dtype=np.complex128)
        it = lib().orc_secant_osc(C.byref(self.s), _ptr(z))
        return complex(z[0]), it

    def secant(self, om):
        z = np.array([complex(om)], dtype=np.complex128)
        it = lib().orc_secant(C.byref(self.s), _ptr(z))
        return complex(z[0]), it

    def calc_chi(self, sproc: int, om, kpar=None, kperp=None):
        """NHDS chi [3,3] and chi_low [3,3,3] (n = -1, 0, 1) of species sproc (1-based)."""
        chi = np.zeros(9, dtype=np.complex128)
        low = np.zeros(27, dtype=np.complex128)
        kz = self.s.kpar if kpar is None else kpar
        kp = self.s.kperp if kperp is None else kperp
        rc = lib().orc_calc_chi(C.byref(self.s), sproc, kz, kp, _c(om), _ptr(chi), _ptr(low))
        if rc != 0:
            raise ValueError("calc_chi stops here (cold species with kperp_norm = F)")
        return chi.reshape(3, 3), low.reshape(3, 3, 3)

    def nhds_nmaxrun(self, sproc: int, kperp: float) -> int:
        return lib().orc_nhds_nmaxrun(C.byref(self.s), sproc, kperp)

    def int_ee_rel(self, sproc: int) -> complex:
        return _z(lib().orc_int_ee_rel(C.byref(self.s), sproc))

    @property
    def current_int(self):
        return [self.s.current_int[i] for i in range(self.plasma.nspec)]

    def calc_eigen(self, om, eigen=True, heat=True):
        """calc_eigen (ALPS_fns.f90:2605-2899) at om after disp(om): electric and
        magnetic field ef, bf [3], velocities Us [nspec, 3], densities ds
        [nspec], heating Ps [nspec], Ps_split [nspec, 4], W_EM."""
        ns = self.plasma.nspec
        chi = np.zeros(ns * 9, dtype=np.complex128)
        low = np.zeros(ns * 27, dtype=np.complex128)
        wave = np.zeros(9, dtype=np.complex128)
        lib().orc_disp(C.byref(self.s), _c(om), _ptr(chi), _ptr(low), _ptr(wave))
        ef = np.zeros(3, dtype=np.complex128)
        bf = np.zeros(3, dtype=np.complex128)
        us = np.zeros(3 * ns, dtype=np.complex128)
        ds = np.zeros(ns, dtype=np.complex128)
        ps = np.zeros(ns)
        psp = np.zeros(4 * ns)
        ew = np.zeros(1)
        lib().orc_calc_eigen(C.byref(self.s), _c(om), _ptr(chi), _ptr(wave), int(eigen), int(heat),
                             _ptr(ef), _ptr(bf), _ptr(us), _ptr(ds), _ptr(ps), _ptr(psp), _ptr(ew))
        return dict(ef=ef, bf=bf, Us=us.reshape(ns, 3), ds=ds, Ps=ps, Ps_split=psp.reshape(ns, 4),
                    W_EM=float(ew[0]))

    def rtsec(self, om):
        flag = C.c_int(0)
        r = lib().orc_rtsec(C.byref(self.s), _c(om), C.byref(flag))
        return _z(r), flag.value

    def map_search(self, omi, omf, gami, gamf, nr, ni, loggridw=False, loggridg=False,
                   numroots=100):
        om = np.zeros(nr * ni, dtype=np.complex128)
        cal = np.zeros(nr * ni, dtype=np.complex128)
        val = np.zeros(nr * ni)
        iroots = np.zeros(2 * numroots, dtype=np.int32)
        n = lib().orc_map_search(C.byref(self.s), omi, omf, gami, gamf, nr, ni, int(loggridw),
                                 int(loggridg), _ptr(om), _ptr(cal), _ptr(val), numroots,
                                 iroots.ctypes.data_as(C.POINTER(C.c_int)))
        return (om.reshape(nr, ni), cal.reshape(nr, ni), val.reshape(nr, ni),
                iroots.reshape(numroots, 2)[: min(n, numroots)], n)

    def om_double_scan(self, s1, s2, roots, eigen=False, heat=False):
        """om_double_scan (ALPS_fns.f90:2900-3591).  s1, s2: (type, log_scan,
        diff, diff2, n_out, n_res).  Returns rc, kperp and kpar [n1+1, n2+1],
        roots [n1+1, n2+1, nroots], written [same] and, with eigen or heat,
        the calc_eigen outputs per written (point, root)."""
        roots = np.array(roots, dtype=np.complex128)
        nr, ns = roots.size, self.plasma.nspec
        a, b = OrcScan(*s1), OrcScan(*s2)
        n1, n2 = s1[4] + 1, s2[4] + 1
        rows = n1 * n2
        kp, kz = np.zeros(rows), np.zeros(rows)
        out = np.zeros(rows * nr, dtype=np.complex128)
        wr = np.zeros(rows * nr, dtype=np.int32)
        ef = np.zeros(rows * nr * 3, dtype=np.complex128)
        bf = np.zeros(rows * nr * 3, dtype=np.complex128)
        us = np.zeros(rows * nr * 3 * ns, dtype=np.complex128)
        ds = np.zeros(rows * nr * ns, dtype=np.complex128)
        ps = np.zeros(rows * nr * ns)
        psp = np.zeros(rows * nr * 4 * ns)
        ew = np.zeros(rows * nr)
        rc = lib().orc_om_double_scan(C.byref(self.s), C.byref(a), C.byref(b), nr, _ptr(roots),
                                      int(eigen), int(heat), _ptr(kp), _ptr(kz), _ptr(out),
                                      wr.ctypes.data_as(C.POINTER(C.c_int)), _ptr(ef), _ptr(bf),
                                      _ptr(us), _ptr(ds), _ptr(ps), _ptr(psp), _ptr(ew))
        sh = (n1, n2, nr)
        return dict(rc=rc, kperp=kp.reshape(n1, n2), kpar=kz.reshape(n1, n2),
                    roots=out.reshape(sh), written=wr.reshape(sh).astype(bool),
                    last_roots=roots, ef=ef.reshape(sh + (3,)), bf=bf.reshape(sh + (3,)),
                    Us=us.reshape(sh + (ns, 3)), ds=ds.reshape(sh + (ns,)),
                    Ps=ps.reshape(sh + (ns,)), Ps_split=psp.reshape(sh + (ns, 4)),
                    W_EM=ew.reshape(sh))

    def determine_param_fit(self, maxsteps=500, lambda0=1.0, lambdafac=10.0, eps=1e-8):
        """determine_param_fit (ALPS_analyt.f90:365-728) from the initial
        guesses in param_fit[:, 0]; updates this oracle's fit state (as the
        reference updates its globals) and returns param_fit [nspec, np+1, 5,
        max_fits], fit_type, perp_correction [nspec, max_fits],
        poly_fit_coeffs, the summed quality and LM step count."""
        q, st = C.c_double(0.0), C.c_int(0)
        rc = lib().orc_determine_param_fit(C.byref(self.s), int(maxsteps), float(lambda0),
                                           float(lambdafac), float(eps), C.byref(q), C.byref(st))
        p = self.plasma
        mf = p.max_fits
        ft = np.array([[self.s.fit_type[i][f] for f in range(mf)] for i in range(p.nspec)])
        pc = np.array([[self.s.perp_correction[i][f] for f in range(mf)] for i in range(p.nspec)])
        return dict(rc=rc, param_fit=self._pf[..., :mf].copy(), fit_type=ft, perp_correction=pc,
                    poly_fit_coeffs=self._poly.copy(), quality=q.value, steps=st.value)

    def om_scan(self, scan_type, log_scan, diff, diff2, n_out, n_res, roots):
        roots = np.array(roots, dtype=np.complex128)
        nr = roots.size
        kp = np.zeros(n_out + 1)
        kz = np.zeros(n_out + 1)
        out = np.zeros((n_out + 1) * nr, dtype=np.complex128)
        rc = lib().orc_om_scan(C.byref(self.s), scan_type, int(log_scan), diff, diff2, n_out,
                               n_res, nr, _ptr(roots), _ptr(kp), _ptr(kz), _ptr(out))
        return rc, kp, kz, out.reshape(n_out + 1, nr)
