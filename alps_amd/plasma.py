"""Plasma description: the inputs ``disp()`` reads from ALPS's global state.

The reference keeps these as module variables of ``alps_var``
(ALPS_var.f90:19-427) filled by the namelist reader (ALPS_io.f90:46-277),
``read_f0`` (ALPS_io.f90:643-692) and the fit set-up
(ALPS_analyt.f90:365-727).  :class:`Plasma` holds the same quantities as
numpy arrays indexed like the Fortran arrays but with the species first:

* ``pp[s, iperp, ipar, 0/1]``    -> ``pp(s+1, iperp, ipar, 1/2)``
* ``f0[s, iperp, ipar]``         -> ``f0(s+1, iperp, ipar)``
* ``param_fit[s, iperp, k, f]``  -> ``param_fit(s+1, iperp, k+1, f+1)``
* ``poly_fit_coeffs[s, iperp, n]`` -> ``poly_fit_coeffs(s+1, iperp, n)``

:meth:`Plasma.fortran` returns the column-major copies the C-ABI takes,
which are byte-identical to the Fortran arrays.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from . import distribution as _dist

__all__ = ["Plasma", "RootParams", "bimaxwellian_plasma", "kappa_plasma", "juttner_plasma"]


@dataclass
class RootParams:
    """Root-search controls (&system namelist, ALPS_var.f90:147-165)."""
    secant_method: int = 2
    numiter: int = 50
    D_threshold: float = 1.0e-5
    D_prec: float = 1.0e-5
    D_gap: float = 1.0e-5
    D_tol: float = 1.0e-7


@dataclass
class Plasma:
    pp: np.ndarray
    f0: np.ndarray
    ns: np.ndarray
    qs: np.ndarray
    ms: np.ndarray
    vA: float
    kperp_norm: bool = True
    positions_principal: int = 5
    n_resonance_interval: int = 100
    Tlim: float = 0.01
    Bessel_zero: float = 1.0e-45
    ACmethod: np.ndarray = None          # [nspec] 1 = fits, 2 = Chebyshev
    n_fits: np.ndarray = None            # [nspec]
    fit_type: np.ndarray = None          # [nspec, maxfits]
    perp_correction: np.ndarray = None   # [nspec, maxfits]
    param_fit: np.ndarray = None         # [nspec, nperp+1, 5, maxfits]
    poly_kind: np.ndarray = None         # [nspec]
    poly_order: np.ndarray = None        # [nspec]
    logfit: np.ndarray = None            # [nspec]
    poly_log_max: np.ndarray = None      # [nspec]
    poly_fit_coeffs: np.ndarray = None   # [nspec, nperp+1, maxorder+1]
    # closed-form species (use_bM) and their &bM_spec parameters
    # (ALPS_var.f90:332-350, defaults of ALPS_io.f90:145-150)
    usebM: np.ndarray = None             # [nspec] 0/1
    bM_nmax: np.ndarray = None           # [nspec], 500
    bM_Bessel_zero: np.ndarray = None    # [nspec], 1e-50
    bM_beta: np.ndarray = None           # [nspec], 1 (0 = cold plasma)
    bM_alpha: np.ndarray = None          # [nspec], 1
    bM_pdrift: np.ndarray = None         # [nspec], 0
    # relativistic species and their (Gamma, pparbar) tables as derivative_f0_rel
    # leaves them (ALPS_fns_rel.f90:19-408); the *_rel arrays hold the
    # relativistic species only, in species order: [nspec_rel, ngamma+1, npparbar+1]
    relativistic: np.ndarray = None      # [nspec] 0/1
    ngamma: int = 0
    npparbar: int = 0
    gamma_rel: np.ndarray = None
    pparbar_rel: np.ndarray = None
    f0_rel: np.ndarray = None
    df0_rel: np.ndarray = None           # [nspec_rel, ngamma+1, npparbar+1, 2]
    roots: RootParams = field(default_factory=RootParams)

    def __post_init__(self):
        self.pp = np.ascontiguousarray(self.pp, dtype=np.float64)
        self.f0 = np.ascontiguousarray(self.f0, dtype=np.float64)
        self.ns = np.asarray(self.ns, dtype=np.float64).ravel()
        self.qs = np.asarray(self.qs, dtype=np.float64).ravel()
        self.ms = np.asarray(self.ms, dtype=np.float64).ravel()
        n = self.nspec
        if self.ACmethod is None:
            self.ACmethod = np.ones(n, dtype=np.int32)
        if self.n_fits is None:
            self.n_fits = np.ones(n, dtype=np.int32)
        if self.fit_type is None:
            self.fit_type = np.ones((n, 1), dtype=np.int32)
        if self.perp_correction is None:
            self.perp_correction = np.ones((n, self.fit_type.shape[1]))
        if self.param_fit is None:
            self.param_fit = np.zeros((n, self.nperp + 1, 5, self.fit_type.shape[1]))
        if self.poly_kind is None:
            self.poly_kind = np.zeros(n, dtype=np.int32)
        if self.poly_order is None:
            self.poly_order = np.zeros(n, dtype=np.int32)
        if self.logfit is None:
            self.logfit = np.ones(n, dtype=np.int32)
        if self.poly_log_max is None:
            self.poly_log_max = np.zeros(n)
        if self.poly_fit_coeffs is None:
            self.poly_fit_coeffs = np.zeros((n, self.nperp + 1, 1))
        for name, val, dt in (("usebM", 0, np.int32), ("bM_nmax", 500, np.int32),
                              ("bM_Bessel_zero", 1e-50, np.float64), ("bM_beta", 1.0, np.float64),
                              ("bM_alpha", 1.0, np.float64), ("bM_pdrift", 0.0, np.float64)):
            v = getattr(self, name)
            setattr(self, name, np.full(n, val, dtype=dt) if v is None
                    else np.asarray(v, dtype=dt).ravel().copy())
        self.relativistic = (np.zeros(n, dtype=np.int32) if self.relativistic is None
                             else np.asarray(self.relativistic, dtype=np.int32).ravel().copy())
        nrel = int(self.relativistic.sum())
        if nrel:
            shp = (nrel, self.ngamma + 1, self.npparbar + 1)
            for name in ("gamma_rel", "pparbar_rel", "f0_rel"):
                a = np.ascontiguousarray(getattr(self, name), dtype=np.float64)
                if a.shape != shp:
                    raise ValueError(f"{name} must have shape {shp}")
                setattr(self, name, a)
            self.df0_rel = np.ascontiguousarray(self.df0_rel, dtype=np.float64)
            if self.df0_rel.shape != shp + (2,):
                raise ValueError(f"df0_rel must have shape {shp + (2,)}")
        self.ACmethod = np.asarray(self.ACmethod, dtype=np.int32)
        self.n_fits = np.asarray(self.n_fits, dtype=np.int32)
        self.fit_type = np.ascontiguousarray(self.fit_type, dtype=np.int32)
        self.perp_correction = np.ascontiguousarray(self.perp_correction, dtype=np.float64)
        self.param_fit = np.ascontiguousarray(self.param_fit, dtype=np.float64)
        self.poly_kind = np.asarray(self.poly_kind, dtype=np.int32)
        self.poly_order = np.asarray(self.poly_order, dtype=np.int32)
        self.logfit = np.asarray(self.logfit, dtype=np.int32)
        self.poly_log_max = np.asarray(self.poly_log_max, dtype=np.float64)
        self.poly_fit_coeffs = np.ascontiguousarray(self.poly_fit_coeffs, dtype=np.float64)

    @property
    def nspec(self) -> int:
        return self.pp.shape[0]

    @property
    def nperp(self) -> int:
        return self.pp.shape[1] - 1

    @property
    def npar(self) -> int:
        return self.pp.shape[2] - 1

    @property
    def max_fits(self) -> int:
        return self.fit_type.shape[1]

    @property
    def nspec_rel(self) -> int:
        return int(self.relativistic.sum())

    def fortran(self) -> dict:
        """Column-major copies with the Fortran dimension order (species fastest)."""
        f = np.asfortranarray
        out = dict(
            pp=f(self.pp), f0=f(self.f0), fit_type=f(self.fit_type),
            perp_correction=f(self.perp_correction), param_fit=f(self.param_fit),
            poly_fit_coeffs=f(self.poly_fit_coeffs),
        )
        if self.nspec_rel:
            out.update(gamma_rel=f(self.gamma_rel), pparbar_rel=f(self.pparbar_rel),
                       f0_rel=f(self.f0_rel), df0_rel=f(self.df0_rel))
        return out


def bimaxwellian_plasma(ms, qs, ns, taus, alphs, ps, beta, vA, nperp, npar, maxP,
                        perpcorr=None, **numerics) -> Plasma:
    """A plasma of bi-Maxwellian species tabulated like generate_distribution
    (distributions=1), with fit_type-1 fits at their least-squares optimum."""
    sp = [_dist.SpeciesDist(ms=m, taus=t, alphs=a, ps=p, distribution=1)
          for m, t, a, p in zip(ms, taus, alphs, ps)]
    pp, f0, ideal = _dist.generate(sp, beta, vA, nperp, npar, maxP)
    nspec = len(sp)
    param_fit = np.zeros((nspec, nperp + 1, 5, 1))
    perp_corr = np.ones((nspec, 1))
    for s in range(nspec):
        pc = ideal[s]["perpcorr"] if perpcorr is None or perpcorr[s] is None else perpcorr[s]
        perp_corr[s, 0] = pc
        param_fit[s, :, :, 0] = _dist.maxwell_fit_params(
            pp, s, ideal[s]["fit_1"], ideal[s]["fit_2"], ideal[s]["fit_3"], pc,
            ideal[s]["perpcorr"])
    roots = numerics.pop("roots", RootParams())
    return Plasma(pp=pp, f0=f0, ns=ns, qs=qs, ms=ms, vA=vA, fit_type=np.ones((nspec, 1)),
                  perp_correction=perp_corr, param_fit=param_fit, roots=roots, **numerics)


def kappa_plasma(ms, qs, ns, taus, alphs, ps, kappas, beta, vA, nperp, npar, maxP,
                 **numerics) -> Plasma:
    """A plasma of bi-kappa species (distributions=2) with fit_type-2 fits at the
    generator's ideal parameters."""
    sp = [_dist.SpeciesDist(ms=m, taus=t, alphs=a, ps=p, kappas=k, distribution=2)
          for m, t, a, p, k in zip(ms, taus, alphs, ps, kappas)]
    pp, f0, ideal = _dist.generate(sp, beta, vA, nperp, npar, maxP)
    nspec = len(sp)
    param_fit = np.zeros((nspec, nperp + 1, 5, 1))
    perp_corr = np.ones((nspec, 1))
    for s in range(nspec):
        perp_corr[s, 0] = ideal[s]["perpcorr"]
        param_fit[s, :, :, 0] = _dist.kappa_fit_params(
            pp, s, ideal[s]["fit_1"], ideal[s]["fit_2"], ideal[s]["fit_3"],
            ideal[s]["fit_4"], ideal[s]["fit_5"])
    roots = numerics.pop("roots", RootParams())
    return Plasma(pp=pp, f0=f0, ns=ns, qs=qs, ms=ms, vA=vA, fit_type=np.full((nspec, 1), 2),
                  perp_correction=perp_corr, param_fit=param_fit, roots=roots, **numerics)


def juttner_plasma(ms, qs, ns, taus, beta, vA, nperp, npar, maxP, ngamma, npparbar,
                   analytic_rel=False, **numerics) -> Plasma:
    """A plasma of relativistic Juettner species (distributions=3, alphs = 1, no
    drift) with fit_type-4 fits (Gamma only) at the generator's ideal values,
    and the (Gamma, pparbar) tables of derivative_f0_rel.  ``analytic_rel``
    evaluates f0 on the new grid from the Juettner formula instead of the
    reference's spline of the table (needed for large tables)."""
    import math
    sp = [_dist.SpeciesDist(ms=m, taus=t, alphs=1.0, ps=0.0, distribution=3)
          for m, t in zip(ms, taus)]
    pp, f0, ideal = _dist.generate(sp, beta, vA, nperp, npar, maxP)
    nspec = len(sp)
    npf = max(nperp, ngamma)
    param_fit = np.zeros((nspec, npf + 1, 5, 1))
    perp_corr = np.ones((nspec, 1))
    rel = [_dist.relativistic_tables(
        pp[s], f0[s], ms[s], vA, ngamma, npparbar,
        f_of=(lambda g, q, c=ideal[s]["perpcorr"]: np.exp(-c * g)) if analytic_rel else None)
        for s in range(nspec)]
    for s in range(nspec):
        perp_corr[s, 0] = ideal[s]["perpcorr"]
        param_fit[s, :, 0, 0] = ideal[s]["fit_1"]
    roots = numerics.pop("roots", RootParams())
    return Plasma(pp=pp, f0=f0, ns=ns, qs=qs, ms=ms, vA=vA, fit_type=np.full((nspec, 1), 4),
                  perp_correction=perp_corr, param_fit=param_fit, roots=roots,
                  relativistic=np.ones(nspec, dtype=np.int32), ngamma=ngamma, npparbar=npparbar,
                  gamma_rel=np.stack([r[0] for r in rel]), pparbar_rel=np.stack([r[1] for r in rel]),
                  f0_rel=np.stack([r[2] for r in rel]), df0_rel=np.stack([r[3] for r in rel]),
                  **numerics)
