"""Background distribution tables f0(pperp, ppar) in the ALPS table format.

ALPS reads one table per species, ``distribution/<name>.<is>.array``, with
``(nperp+1)*(npar+1)`` lines ``pperp ppar f0`` in the order
``do iperp=0,nperp; do ipar=0,npar`` (ALPS_io.f90:643-692).  The tables are
made by the reference's generator (distribution/generate_distribution.f90);
:func:`generate` follows that generator for the distributions the solver
path consumes, and returns the "ideal" fit parameters the generator prints
(generate_distribution.f90:176-265) so that a plasma can be set up without a
Levenberg-Marquardt fit.

Momenta are in units of m_ref v_A,ref.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

__all__ = ["SpeciesDist", "generate", "write_array", "read_array", "maxwell_fit_params",
           "kappa_fit_params", "relativistic_tables"]


@dataclass
class SpeciesDist:
    """Input of one species, as in generate_distribution's &spec_j namelist."""
    ms: float = 1.0          # m_j / m_ref
    taus: float = 1.0        # T_par,j / T_par,ref
    alphs: float = 1.0       # T_perp,j / T_par,j
    ps: float = 0.0          # drift momentum m_j v_drift / (m_ref v_A,ref)
    kappas: float = 8.0      # kappa index (distribution 2)
    distribution: int = 1    # 1 bi-Maxwellian, 2 bi-kappa, 3 Juettner
    autoscale: bool = True
    maxPperp: float = 1.0
    maxPpar: float = 1.0


def _grid(nperp: int, npar: int, pperp_max: float, ppar_max: float, p_drift: float):
    """pperp = real(iperp)*dpperp, ppar = real(ipar)*dppar - ppar_max + p_drift
    (generate_distribution.f90:267-273)."""
    dpperp = pperp_max / float(nperp)
    dppar = 2.0 * ppar_max / float(npar)
    iperp = np.arange(nperp + 1, dtype=np.float64)
    ipar = np.arange(npar + 1, dtype=np.float64)
    pperp = iperp * dpperp
    ppar = ipar * dppar - ppar_max + p_drift
    return pperp, ppar


def generate(species: list[SpeciesDist], beta: float, vA: float, nperp: int, npar: int,
             maxP: float):
    """Tabulate f0 for every species (generate_distribution.f90:164-316).

    Returns ``(pp, f0, ideal)`` with ``pp[s, iperp, ipar, 0:2]`` = (pperp, ppar),
    ``f0[s, iperp, ipar]`` and ``ideal[s]`` = dict of the ideal fit values
    (``fit_1..fit_5``, ``perpcorr``).
    """
    nspec = len(species)
    pp = np.zeros((nspec, nperp + 1, npar + 1, 2))
    f0 = np.zeros((nspec, nperp + 1, npar + 1))
    ideal = []
    pi = math.atan(1.0) * 4.0
    for s, sp in enumerate(species):
        ms, tau, alph, pd = sp.ms, sp.taus, sp.alphs, sp.ps
        fit = dict(fit_1=0.0, fit_2=0.0, fit_3=0.0, fit_4=0.0, fit_5=0.0, perpcorr=1.0)
        if sp.distribution == 1:
            norm = pi ** (-1.5) / ((ms * beta * tau) ** (3.0 / 2.0) * alph)
            if sp.autoscale:
                pperp_max = maxP * math.sqrt(ms * tau * alph)
                ppar_max = maxP * math.sqrt(ms * tau)
            else:
                pperp_max, ppar_max = sp.maxPperp, sp.maxPpar
            fit.update(fit_1=norm, fit_2=1.0 / (beta * ms * tau), fit_3=pd, fit_4=0.0,
                       perpcorr=1.0 / (tau * beta * ms * alph))
        elif sp.distribution == 2:
            kap = sp.kappas
            a = math.sqrt((2.0 * kap - 3.0) / (2.0 * kap))
            norm = 1.0 / ((ms * beta * tau * pi * kap) ** (3.0 / 2.0) * alph)
            norm = norm * math.gamma(kap + 1.0) / (math.gamma(kap - 0.5) * a ** 3)
            if sp.autoscale:
                k1 = species[0].kappas
                sc = math.sqrt((kap - 1.5) / (k1 - 1.5))
                pperp_max = maxP * math.sqrt(ms * tau * alph) * sc
                ppar_max = maxP * math.sqrt(ms * tau) * sc
            else:
                pperp_max, ppar_max = sp.maxPperp, sp.maxPpar
            fit.update(fit_1=norm, fit_2=1.0 / (beta * ms * kap * a * a * tau), fit_3=pd,
                       fit_4=(-1.0 - kap), fit_5=1.0,
                       perpcorr=1.0 / (tau * beta * ms * kap * a * a * alph))
        elif sp.distribution == 3:
            from scipy.special import kn as _bessk
            norm = vA / (2.0 * pi * math.sqrt(alph) * ms ** 2 * beta * tau)
            norm = norm / _bessk(2, 2.0 * ms / (vA * vA * alph * beta * tau))
            if sp.autoscale:
                pm = math.sqrt(maxP * maxP * tau + (tau - ms * ms) / (vA * vA))
                pperp_max, ppar_max = pm * math.sqrt(alph), pm
            else:
                pperp_max, ppar_max = sp.maxPperp, sp.maxPpar
            # fit type 4 (Gamma only): fit_1 and perpcorr
            fit.update(fit_1=norm, fit_2=vA * vA * alph / (ms * ms), fit_3=pd, fit_4=0.0,
                       fit_5=0.0, perpcorr=(2.0 * ms / (vA * vA * beta * tau * alph)))
        else:
            raise ValueError(f"distribution {sp.distribution} is not tabulated here")

        pperp, ppar = _grid(nperp, npar, pperp_max, ppar_max, pd)
        P2 = pperp[:, None]
        Q = ppar[None, :] - pd
        if sp.distribution == 1:
            f = np.exp(-((Q ** 2.0) / (beta * ms * tau) + (P2 ** 2.0) / (tau * beta * ms * alph)))
        elif sp.distribution == 3:
            f = np.exp(-(2.0 * ms / (vA * vA * beta * tau * alph))
                       * np.sqrt(1.0 + P2 ** 2 * vA * vA / (ms * ms)
                                 + Q ** 2 * vA * vA * alph / (ms * ms)))
        else:
            f = (1.0 + (Q ** 2.0 / (beta * ms * kap * a * a * tau))
                 + (P2 ** 2.0) / (tau * beta * ms * kap * a * a * alph)) ** (-1.0 - kap)
        f0[s] = f * norm
        pp[s, :, :, 0] = pperp[:, None]
        pp[s, :, :, 1] = ppar[None, :]
        ideal.append(fit)
    return pp, f0, ideal


def maxwell_fit_params(pp: np.ndarray, s: int, fit_1: float, fit_2: float, fit_3: float,
                       perpcorr: float, perp_exact: float) -> np.ndarray:
    """param_fit(s, 0:nperp, 1:3) of a fit_type-1 (Maxwell) fit to a bi-Maxwellian table.

    The Maxwell fit form is u1*exp(-perpcorr*pperp**2)*exp(-u2*(ppar-u3)**2)
    (ALPS_analyt.f90:225-229).  For a bi-Maxwellian table the logarithmic
    least-squares fit is exact, so its minimum is u2 = fit_2, u3 = fit_3 and
    u1(pperp) = fit_1*exp((perpcorr - perp_exact)*pperp**2), where perp_exact is
    the table's own 1/(tau*beta*ms*alph).  Returns an array [nperp+1, 5].
    """
    pperp = pp[s, :, 1, 0]
    out = np.zeros((pperp.size, 5))
    out[:, 0] = fit_1 * np.exp((perpcorr - perp_exact) * pperp ** 2)
    out[:, 1] = fit_2
    out[:, 2] = fit_3
    return out


def kappa_fit_params(pp: np.ndarray, s: int, fit_1: float, fit_2: float, fit_3: float,
                     fit_4: float, fit_5: float) -> np.ndarray:
    """param_fit(s, 0:nperp, 1:5) of a fit_type-2 (kappa) fit to a bi-kappa table with
    perp_correction equal to the generator's ideal value: the fit form
    u1*(1 + u2*(ppar-u3)**2 + perpcorr*u5*pperp**2)**u4 (ALPS_analyt.f90:231-236)
    reproduces the table exactly with the ideal parameters at every pperp."""
    n = pp.shape[1]
    out = np.zeros((n, 5))
    out[:] = [fit_1, fit_2, fit_3, fit_4, fit_5]
    return out


def write_array(path: str, pp_s: np.ndarray, f0_s: np.ndarray) -> None:
    """Write one species table in the ``.array`` layout read by ALPS_io.f90:674-679."""
    with open(path, "w") as fh:
        for iperp in range(f0_s.shape[0]):
            for ipar in range(f0_s.shape[1]):
                fh.write(f"{pp_s[iperp, ipar, 0]!r} {pp_s[iperp, ipar, 1]!r} {f0_s[iperp, ipar]!r}\n")


def read_array(path: str, nperp: int, npar: int):
    """Read one species table (``pperp ppar f0`` per line, iperp outer)."""
    data = np.loadtxt(path).reshape(nperp + 1, npar + 1, 3)
    return data[:, :, :2].copy(), data[:, :, 2].copy()


def relativistic_tables(pp_s: np.ndarray, f0_s: np.ndarray, ms: float, vA: float, ngamma: int,
                        npparbar: int, f_of=None):
    """The (Gamma, pparbar) grid of one relativistic species as derivative_f0_rel
    leaves it (ALPS_fns_rel.f90:19-408): ``gamma_rel``, ``pparbar_rel``,
    ``f0_rel`` (-1 outside the subluminal cone Gamma^2 - 1 < pparbar^2,
    normalised to unit density) and ``df0_rel[..., 0:2]`` (d/dGamma, d/dpparbar
    on the interior, one-sided at the cone's edge, zero on the frame).

    f0 on the new grid comes from the reference's thin-plate (polyharmonic)
    spline of log f0 over the table points, solved densely -- feasible for
    small tables only (the system has (nperp+1)(npar+1)+3 unknowns).  With
    ``f_of(gamma, pparbar)`` the spline is replaced by that function (an
    analytic f0 for large grids)."""
    pperp, ppar = pp_s[..., 0], pp_s[..., 1]
    gam = np.sqrt((pperp ** 2 + ppar ** 2) * vA ** 2 / ms ** 2 + 1.0)
    gamma_min, gamma_max = gam.min(), gam.max()
    pb_min, pb_max = ppar.min() * vA / ms, ppar.max() * vA / ms
    gamma_max_use = math.sqrt(1.0 + pp_s[-1, 1, 0] ** 2 * vA ** 2 / ms ** 2)
    ig = np.arange(ngamma + 1, dtype=np.float64)[:, None]
    ip = np.arange(npparbar + 1, dtype=np.float64)[None, :]
    gamma_rel = np.broadcast_to(gamma_min + ((gamma_max_use - gamma_min) * ig) / (1.0 * ngamma),
                                (ngamma + 1, npparbar + 1)).copy()
    pparbar_rel = np.broadcast_to(pb_min + ((pb_max - pb_min) * ip) / (1.0 * npparbar),
                                  (ngamma + 1, npparbar + 1)).copy()
    if f_of is None:
        gc = np.sqrt(1.0 + (pperp ** 2 + ppar ** 2) * vA * vA / (ms * ms)).ravel()
        pc = (ppar * vA / ms).ravel()
        y = np.log(f0_s).ravel()
        n = gc.size

        def kern(r):
            out = np.zeros_like(r)
            big = r >= 1.0
            small = (r > 0.0) & ~big
            out[big] = r[big] * r[big] * np.log(r[big])
            out[small] = r[small] * np.log(r[small] ** r[small])
            return out

        A = np.zeros((n + 3, n + 3))
        A[:n, :n] = kern(np.sqrt((gc[:, None] - gc[None, :]) ** 2 + (pc[:, None] - pc[None, :]) ** 2))
        A[:n, n], A[:n, n + 1], A[:n, n + 2] = 1.0, gc, pc
        A[n, :n], A[n + 1, :n], A[n + 2, :n] = 1.0, gc, pc
        rhs = np.zeros(n + 3)
        rhs[:n] = y
        w = np.linalg.solve(A, rhs)
        g, q = gamma_rel.ravel(), pparbar_rel.ravel()
        f = np.zeros(g.size)
        for k0 in range(0, g.size, 4096):
            sl = slice(k0, k0 + 4096)
            r = np.sqrt((g[sl, None] - gc[None, :]) ** 2 + (q[sl, None] - pc[None, :]) ** 2)
            f[sl] = kern(r) @ w[:n]
        f = f + w[n] + w[n + 1] * g + w[n + 2] * q
        f0_rel = np.exp(f.reshape(gamma_rel.shape))
    else:
        f0_rel = np.asarray(f_of(gamma_rel, pparbar_rel), dtype=np.float64).copy()
    f0_rel[(gamma_rel ** 2 - 1.0) < pparbar_rel ** 2] = -1.0
    pi = math.atan(1.0) * 4.0
    dg = gamma_rel[2, 2] - gamma_rel[1, 2]
    dpb = pparbar_rel[2, 2] - pparbar_rel[2, 1]
    valid = f0_rel > -1.0
    integ = np.sum(gamma_rel[valid] * f0_rel[valid]) * 2.0 * pi * dg * dpb * (ms / vA) ** 3
    f0_rel[f0_rel != -1.0] /= integ
    f = f0_rel
    df = np.zeros(f.shape + (2,))
    for i in range(1, ngamma):
        for j in range(1, npparbar):
            if f[i - 1, j] > 0.0 and f[i + 1, j] > 0.0:
                df[i, j, 0] = (f[i + 1, j] - f[i - 1, j]) / (gamma_rel[i + 1, j] - gamma_rel[i - 1, j])
            if f[i, j + 1] > 0.0 and f[i, j - 1] > 0.0:
                df[i, j, 1] = (f[i, j + 1] - f[i, j - 1]) / (pparbar_rel[i, j + 1] - pparbar_rel[i, j - 1])
            if f[i, j] >= 0.0:
                if f[i, j + 1] <= 0.0 and f[i, j - 1] > 0.0:
                    df[i, j, 1] = (f[i, j] - f[i, j - 1]) / (pparbar_rel[i, j] - pparbar_rel[i, j - 1])
                if f[i, j + 1] > 0.0 and f[i, j - 1] <= 0.0:
                    df[i, j, 1] = (f[i, j + 1] - f[i, j]) / (pparbar_rel[i, j + 1] - pparbar_rel[i, j])
    return gamma_rel, pparbar_rel, f0_rel, df
