"""Plasma set-ups shared by the CPU and GPU tests (synthetic f0 tables made with
the generator formulas of distribution/generate_distribution.f90)."""
import math

import numpy as np

from alps_amd.plasma import Plasma, RootParams, bimaxwellian_plasma, kappa_plasma

M_E = 5.44662e-4  # m_e/m_p used by the reference's test inputs


def small_bimax(kperp_norm=True, nperp=24, npar=48, drift=0.0, alph=1.0, P=3):
    """Protons + electrons, 24 x 48, small enough for the literal CPU oracle."""
    return bimaxwellian_plasma(ms=[1.0, M_E], qs=[1.0, -1.0], ns=[1.0, 1.0], taus=[1.0, 1.0],
                               alphs=[alph, 1.0], ps=[drift, 0.0], beta=1.0, vA=1e-4,
                               nperp=nperp, npar=npar, maxP=6.0, perpcorr=[None, 1.836e3],
                               kperp_norm=kperp_norm, positions_principal=P,
                               n_resonance_interval=100, Tlim=0.01, Bessel_zero=1e-45)


def three_species(nperp=24, npar=48):
    """Protons (anisotropic), electrons, drifting alphas (5 %)."""
    return bimaxwellian_plasma(ms=[1.0, M_E, 4.0], qs=[1.0, -1.0, 2.0], ns=[1.0, 1.1, 0.05],
                               taus=[1.0, 1.0, 1.0], alphs=[2.0, 1.0, 1.0], ps=[0.0, 0.0, 4.0],
                               beta=1.0, vA=1e-4, nperp=nperp, npar=npar, maxP=6.0,
                               positions_principal=3, n_resonance_interval=100, Tlim=0.01,
                               Bessel_zero=1e-45)


def kappa_case(nperp=24, npar=48):
    return kappa_plasma(ms=[1.0, M_E], qs=[1.0, -1.0], ns=[1.0, 1.0], taus=[1.0, 1.0],
                        alphs=[2.0, 1.0], ps=[0.0, 0.0], kappas=[4.0, 4.0], beta=1.0, vA=1e-4,
                        nperp=nperp, npar=npar, maxP=6.0, positions_principal=3,
                        n_resonance_interval=100, Tlim=0.01, Bessel_zero=1e-45)


def poly_case(nperp=24, npar=48, order=8):
    """Species 1 continued analytically by a Chebyshev fit of log10 f0 (ACmethod 2)."""
    p = small_bimax(nperp=nperp, npar=npar)
    ns = p.nspec
    coeffs = np.zeros((ns, nperp + 1, order + 1))
    s = 0
    pa = p.pp[s, 0, :, 1]
    n1 = 0.5 * (pa[-1] + pa[0])
    n2 = 0.5 * (pa[-1] - pa[0])
    t = (pa - n1) / n2
    for i in range(nperp + 1):
        coeffs[s, i] = np.polynomial.chebyshev.chebfit(t, np.log10(p.f0[s, i]), order)
    ac = np.array([2, 1], dtype=np.int32)
    return Plasma(pp=p.pp, f0=p.f0, ns=p.ns, qs=p.qs, ms=p.ms, vA=p.vA, kperp_norm=True,
                  positions_principal=3, n_resonance_interval=100, Tlim=0.01,
                  Bessel_zero=1e-45, ACmethod=ac, n_fits=p.n_fits, fit_type=p.fit_type,
                  perp_correction=p.perp_correction, param_fit=p.param_fit,
                  poly_kind=np.array([1, 0]), poly_order=np.array([order, 0]),
                  logfit=np.array([1, 1]), poly_log_max=np.array([18.0, 0.0]),
                  poly_fit_coeffs=coeffs)


def bm_mix(kperp_norm=True, cold=True):
    """Closed-form (use_bM) species beside a tabulated one: an anisotropic
    drifting bi-Maxwellian proton population and, with ``cold``, a cold
    drifting alpha population (test_cold_plasma.in style), both use_bM, plus
    tabulated electrons."""
    import dataclasses
    p = three_species()
    p = dataclasses.replace(p, kperp_norm=kperp_norm)
    return dataclasses.replace(
        p, usebM=[1, 0, 1 if cold else 0], bM_nmax=[500, 500, 500],
        bM_Bessel_zero=[1e-50, 1e-50, 1e-50], bM_beta=[1.0, 1.0, 0.0 if cold else 1.0],
        bM_alpha=[1.5, 1.0, 1.0], bM_pdrift=[0.2, 0.0, 1.0])


REL_OMEGAS = [0.5 + 0.05j, 0.5 - 0.05j, 0.3 - 0.002j, 0.8 + 0.0j, 1.0 - 1.655e-6j,
              0.0627 - 4.662e-8j, 1.5 - 0.3j, 0.12 + 0.3j]


def rel_pair(kperp_norm=True, nperp=12, npar=24, ngamma=20, npparbar=200):
    """Relativistic pair plasma in the style of test_relativistic.in (Juettner
    e+/e-, vA = 1, fit type 4), on a small table."""
    from alps_amd.plasma import juttner_plasma
    return juttner_plasma(ms=[1.0, 1.0], qs=[1.0, -1.0], ns=[1.0, 1.0], taus=[1.0, 1.0], beta=1.0,
                          vA=1.0, nperp=nperp, npar=npar, maxP=5.0, ngamma=ngamma,
                          npparbar=npparbar, positions_principal=5, n_resonance_interval=100,
                          Tlim=0.01, Bessel_zero=1e-45, kperp_norm=kperp_norm)


def rel_mixed():
    """A relativistic Juettner species (vA = 0.3) beside a non-relativistic
    bi-Maxwellian one on the same (pperp, ppar) grid size."""
    import dataclasses
    from alps_amd.plasma import juttner_plasma
    r = juttner_plasma(ms=[1.0], qs=[-1.0], ns=[1.0], taus=[1.0], beta=1.0, vA=0.3, nperp=12,
                       npar=24, maxP=5.0, ngamma=20, npparbar=200, positions_principal=3,
                       n_resonance_interval=100, Tlim=0.01, Bessel_zero=1e-45)
    b = bimaxwellian_plasma(ms=[1.0], qs=[1.0], ns=[1.0], taus=[1.0], alphs=[1.5], ps=[0.0],
                            beta=1.0, vA=0.3, nperp=12, npar=24, maxP=5.0)
    npf = r.param_fit.shape[1]
    pf = np.zeros((2, npf, 5, 1))
    pf[0, : b.param_fit.shape[1]] = b.param_fit[0]
    pf[1] = r.param_fit[0]
    return dataclasses.replace(
        r, pp=np.concatenate([b.pp, r.pp]), f0=np.concatenate([b.f0, r.f0]), ns=[1.0, 1.0],
        qs=[1.0, -1.0], ms=[1.0, 1.0], fit_type=np.array([[1], [4]]),
        perp_correction=np.array([[b.perp_correction[0, 0]], [r.perp_correction[0, 0]]]),
        param_fit=pf, relativistic=[0, 1], ACmethod=None, n_fits=None, poly_kind=None,
        poly_order=None, logfit=None, poly_log_max=None, poly_fit_coeffs=None, usebM=None,
        bM_nmax=None, bM_Bessel_zero=None, bM_beta=None, bM_alpha=None, bM_pdrift=None)


def kpar_fast(nperp=120, npar=240):
    """The reference's tests/test_kpar_fast.in (with test_kpar_fast_dist.in)."""
    rp = RootParams(secant_method=2, numiter=50, D_threshold=1e-15, D_prec=1e-5, D_gap=1e-5)
    return bimaxwellian_plasma(ms=[1.0, M_E], qs=[1.0, -1.0], ns=[1.0, 1.0], taus=[1, 1],
                               alphs=[1, 1], ps=[0, 0], beta=1.0, vA=1e-4, nperp=nperp,
                               npar=npar, maxP=6.0, perpcorr=[1.0, 1.836e3],
                               positions_principal=3, n_resonance_interval=100, Tlim=0.01,
                               Bessel_zero=1e-50, roots=rp)


def config0(nperp=50, npar=100):
    """2 species, 50 x 100, kperp d_p = 0.01, kpar d_p = 0.1 (CPU-runnable case)."""
    rp = RootParams(secant_method=2, numiter=50, D_threshold=1e-12, D_prec=1e-5, D_gap=1e-5)
    return bimaxwellian_plasma(ms=[1.0, M_E], qs=[1.0, -1.0], ns=[1.0, 1.0], taus=[1, 1],
                               alphs=[1, 1], ps=[0, 0], beta=1.0, vA=1e-4, nperp=nperp,
                               npar=npar, maxP=8.0, positions_principal=3,
                               n_resonance_interval=100, Tlim=0.01, Bessel_zero=1e-45, roots=rp)


def config1(nperp=128, npar=256):
    """Batched map case: 2 species at kperp d_p = kpar d_p = 0.5 on 128 x 256."""
    return config0(nperp, npar)


# complex frequencies that exercise every branch of full_integrate on the
# small 24 x 48 plasmas at kperp = kpar = 0.3 (protons: p_res = (w - n)/0.3)
BRANCH_OMEGAS = [
    0.5 + 0.05j,     # resonant, Im > 0, |Im p_res| > Tlim
    0.5 - 0.05j,     # Landau x2, Tlim branch
    0.5 - 0.001j,    # analytic (|Im p_res| <= Tlim) branch, Landau x2
    0.5 + 0.001j,    # analytic branch, Im > 0
    0.5 + 0.0j,      # Im = 0: Landau x1
    1.62 - 0.02j,    # resonance near the upper ppar edge (edge fall-back)
    1.80 - 0.02j,    # resonance just outside the grid
    -1.7 + 0.02j,    # lower edge
    3.3 - 0.2j,      # mostly non-resonant rows
    0.05 - 0.3j,     # strongly damped
]


def det_scale(wave):
    """Magnitude of the terms of the determinant formula (ALPS_fns.f90:622-624)."""
    w = np.abs(wave)
    return (w[0, 0] * (w[1, 1] * w[2, 2] + w[1, 2] ** 2) + 2 * w[0, 1] * w[1, 2] * w[0, 2]
            + w[0, 2] ** 2 * w[1, 1] + w[0, 1] ** 2 * w[2, 2])
