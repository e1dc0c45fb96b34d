"""Pins of the CPU restatement of ALPS's relativistic path (relativistic = T
species, ALPS_fns_rel.f90): the series Bessel function and Gamma function
against scipy/math, the resonant (Gamma, pparbar) integration against an
independent trapezoid of the same integrand off the real axis, its convergence
to the closed-form susceptibility of a cool plasma as the Gamma grid is
refined, and the Landau term through the continuity of chi across the real
omega axis (a wrong sign or factor of the residue leaves an O(1) jump)."""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.special import jv

from alps_amd.plasma import juttner_plasma
from nhds_closed_form import chi_bimax
from oracle import Oracle, cbessj, gamma_fn, rel_error


@pytest.mark.parametrize("z", [0.3 + 0.1j, 1.5 - 0.4j, -0.8 + 0.6j, 2.5 + 0.05j, 0.01 - 0.02j])
@pytest.mark.parametrize("nu", [0, 1, 2, 5])
def test_cbessj_matches_scipy(z, nu):
    ref = jv(nu, z)
    assert abs(cbessj(z, nu) - ref) <= 1e-9 * max(abs(ref), 1e-12), (z, nu)


def test_gamma_matches_math():
    for x in [1.0, 1.5, 2.0, 3.7, 7.0, 12.0, 21.0, 30.0]:
        assert abs(gamma_fn(x) - math.gamma(x)) <= 1e-9 * math.gamma(x), x


def _cool(ngamma, npparbar, nperp=24, npar=48):
    """A proton Juettner plasma with v_th/c = 0.06 (vA = 0.06, beta = 1)."""
    return juttner_plasma(ms=[1.0], qs=[1.0], ns=[1.0], taus=[1.0], beta=1.0, vA=0.06,
                          nperp=nperp, npar=npar, maxP=5.0, ngamma=ngamma, npparbar=npparbar,
                          positions_principal=3, n_resonance_interval=100, Tlim=0.01,
                          Bessel_zero=1e-45, analytic_rel=True)


def _brute(p, om, n, mode, kperp, kpar):
    """Trapezoid of resU_rel * int_T_rel over the subluminal part of every Gamma
    row (weights of integrate_res_rel), with scipy's J_n."""
    g, pb, f, df = p.gamma_rel[0], p.pparbar_rel[0], p.f0_rel[0], p.df0_rel[0]
    m, q, vA = p.ms[0], p.qs[0], p.vA
    dg, dpb = g[2, 2] - g[1, 2], pb[2, 2] - pb[2, 1]
    zbar = kperp * m / (vA * q)
    tot = 0.0
    for ig in range(1, p.ngamma):
        row = np.nonzero(f[ig, 1:p.npparbar] > -1.0)[0] + 1
        w = np.full(row.size, 2.0)
        w[0] = w[-1] = 1.0
        G, PB = g[ig, row], pb[ig, row]
        U = (-2 * np.pi * (m / vA) ** 3 * (q * vA / (kpar * m))
             * (om * df[ig, row, 0] + (kpar / vA) * df[ig, row, 1])
             / (PB - G * om * vA / kpar + n * q * vA / (kpar * m)))
        J = jv(n, zbar * np.sqrt(G * G - 1 - PB * PB))
        T = J * J * PB ** 2 if mode == 3 else n * n * J * J / zbar ** 2
        tot += (2.0 if ig <= p.ngamma - 2 else 1.0) * np.sum(w * U * T) * dpb
    return tot * dg * 0.25


@pytest.mark.parametrize("om", [0.5 + 0.05j, 0.5 + 0.2j])
@pytest.mark.parametrize("n,mode", [(0, 3), (1, 1)])
def test_resonant_rel_integral_matches_trapezoid(om, n, mode):
    p = _cool(40, 200)
    O = Oracle(p)
    O.set_k(0.3, 0.3)
    got = O.full_integrate(1, om, n, mode, True)
    ref = _brute(p, om, n, mode, 0.3, 0.3)
    assert abs(got - ref) <= 2e-3 * abs(ref), (got, ref)
    assert rel_error() == 0


@pytest.mark.slow
def test_rel_chi_converges_to_closed_form_above_the_real_axis():
    om = 0.5 + 0.05j
    errs = []
    for ng, npb in ((100, 500), (200, 1000)):
        p = _cool(ng, npb, 48, 96)
        O = Oracle(p, n_threads=8)
        O.set_k(0.3, 0.3)
        _, chi, _, _ = O.disp(om, want_tensors=True)
        got = chi[0] * (om * om * p.vA * p.vA)
        ref = chi_bimax(om, 0.3, 0.3, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0)
        errs.append(np.max(np.abs(got - ref)) / np.max(np.abs(ref)))
    assert errs[1] < 0.15 and errs[1] < 0.7 * errs[0], errs


def test_landau_term_continues_chi_across_the_real_axis():
    p = _cool(100, 500, 48, 96)
    O = Oracle(p, n_threads=8)
    O.set_k(0.3, 0.3)
    jumps = []
    for e in (0.004, 0.001):
        a = O.full_integrate(1, 0.5 + e * 1j, 1, 1, True)
        b = O.full_integrate(1, 0.5 - e * 1j, 1, 1, True)
        jumps.append(abs(a - b) / abs(a))
    assert jumps[1] < 0.05 and jumps[1] < 0.5 * jumps[0], jumps
