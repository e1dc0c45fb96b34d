"""Pins of the analytic continuations that eval_fit takes beyond the fitted
Maxwellian/kappa/bi-Moyal forms: the pre-coded distribution_analyt of
ACmethod 0 (ALPS_analyt.f90:77-79, distribution/distribution_analyt.f90:30-96)
and the Juettner fit in (pperp, ppar), fit type 3 (ALPS_analyt.f90:238-242).

Each is checked against its closed form written with numpy at complex ppar
(principal square root), against the table it continues at real ppar, and
through disp(): with ACmethod 0 the tabulated chi of a Maxwellian converges
to the closed-form hot-plasma chi as the grid is refined, also below the real
axis where the Landau term is the continuation's; with fit type 3 chi is
continuous across the real omega axis (the Landau term switches on there),
to the grid's own discretisation level, and the jump shrinks with the grid."""
import dataclasses
import math

import numpy as np
import pytest

import cases
from nhds_closed_form import chi_bimax
from oracle import Oracle
from test_oracle_physics import KP, KZ, _proton

P_COMPLEX = [0.3 + 0.0j, -1.2 + 0.4j, 2.5 - 0.7j, 0.05 - 1.5j, -3.0 - 0.2j]


def _maxwell_closed(is_, pperp, ppar):
    """The shipped distribution_analyt: beta = 1, m = 1 (species 1) or 1/1836 (2)."""
    m = 1.0 if is_ == 1 else 1.0 / 1836.0
    return math.pi ** -1.5 / m ** 1.5 * np.exp(-(ppar * ppar / m + pperp * pperp / m))


def test_distribution_analyt_is_the_shipped_maxwellian():
    p = cases.analytical(ms2=1.0 / 1836.0)
    O = Oracle(p)
    for s in (1, 2):
        pe = p.pp[s - 1, :, 0, 0]
        scale = 1.0 if s == 1 else math.sqrt(1.0 / 1836.0)
        for i in (0, 5, p.nperp // 2, p.nperp):
            for z in P_COMPLEX:
                v = O.eval_fit(s, i, z * scale)
                want = _maxwell_closed(s, pe[i], z * scale)
                assert abs(v - want) <= 1e-13 * abs(want) + 1e-300, (s, i, z, v, want)
            # at the grid's real ppar the continuation is the table itself
            for j in (0, 7, p.npar // 2, p.npar):
                v = O.eval_fit(s, i, p.pp[s - 1, i, j, 1])
                assert abs(v.real / p.f0[s - 1, i, j] - 1) < 1e-12 and v.imag == 0.0


def test_juettner_fit_type3_closed_form_and_table():
    p = cases.juettner3()
    O = Oracle(p)
    u1, u2, u3 = cases.JUETTNER3
    pe = p.pp[0, :, 0, 0]
    for i in (0, 3, p.nperp // 2, p.nperp):
        for z in P_COMPLEX:
            want = u1 * np.exp(-u2 * np.sqrt(1 + (pe[i] ** 2 + (z - u3) ** 2) * p.vA ** 2 / p.ms[0] ** 2))
            v = O.eval_fit(1, i, z)
            assert abs(v - want) <= 1e-12 * abs(want), (i, z, v, want)
        for j in (0, 9, p.npar // 2, p.npar):
            v = O.eval_fit(1, i, p.pp[0, i, j, 1])
            assert abs(v.real / p.f0[0, i, j] - 1) < 1e-12 and abs(v.imag) <= 1e-14 * abs(v.real)
    # the branch: a point where 1 + (pperp^2 + (p - u3)^2) vA^2 crosses the
    # negative real axis takes the principal root (Re sqrt >= 0), as Fortran's sqrt
    z = u3 + 1j * math.sqrt(1 / p.vA ** 2 + 1.0) * 1.0000001
    w = np.sqrt(1 + (pe[0] ** 2 + (z - u3) ** 2) * p.vA ** 2)
    assert w.real >= 0 and abs(O.eval_fit(1, 0, z) - u1 * np.exp(-u2 * w)) <= 1e-12 * abs(u1 * np.exp(-u2 * w))


def _chi_err(p, om):
    O = Oracle(p)
    O.set_k(KP, KZ)
    d, chi0, low, wave = O.disp(om, want_tensors=True)
    got = chi0[0] * (om * om * p.vA * p.vA)
    ref = chi_bimax(om, KP, KZ, 1.0, 1.0, 1.0, 1.0, 1.0, 0.0)
    return np.max(np.abs(got - ref)) / np.max(np.abs(ref))


@pytest.mark.parametrize("om", [0.5 - 0.05j, 0.7 - 0.05j, 1.2 - 0.1j])
def test_acmethod0_chi_converges_to_closed_form(om):
    """Below the real axis the Landau term is built from distribution_analyt:
    the protons' chi converges to the closed form as the grid is refined."""
    errs = []
    for n in (64, 128):
        p = dataclasses.replace(_proton(n, 2 * n), ACmethod=np.array([0]))
        errs.append(_chi_err(p, om))
    assert errs[1] < 5e-3 and errs[1] < 0.7 * errs[0], errs


def test_juettner_fit_type3_chi_is_continuous_across_the_real_axis():
    """chi(w + i eps) - chi(w - i eps) is the grid's discretisation error only
    (the continuation is the table's own closed form), and it falls with the grid."""
    jumps = []
    for n in (24, 48):
        p = cases.juettner3(nperp=n, npar=2 * n)
        O = Oracle(p)
        O.set_k(KP, KZ)
        worst = 0.0
        for w in (0.5, 0.8, 1.3):
            a = O.disp(w + 1e-8j, want_tensors=True)[1][0]
            b = O.disp(w - 1e-8j, want_tensors=True)[1][0]
            worst = max(worst, np.max(np.abs(a - b)) / np.max(np.abs(a)))
        jumps.append(worst)
    assert jumps[1] < 2e-2 and jumps[1] < 0.7 * jumps[0], jumps
    # a continuation that is not the table's (u2 off by 10 %) breaks continuity
    p = cases.juettner3(nperp=48, npar=96)
    pf = p.param_fit.copy()
    pf[0, :, 1, 0] *= 1.1
    O = Oracle(dataclasses.replace(p, param_fit=pf))
    O.set_k(KP, KZ)
    a = O.disp(0.8 + 1e-8j, want_tensors=True)[1][0]
    b = O.disp(0.8 - 1e-8j, want_tensors=True)[1][0]
    assert np.max(np.abs(a - b)) / np.max(np.abs(a)) > 5 * jumps[1]
