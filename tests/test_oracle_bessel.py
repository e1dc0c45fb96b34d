"""Pins of the Bessel restatement (BESSJ/BESSJ0/BESSJ1, ALPS_fns_rel.f90:1580-1735)
and of determine_nmax / determine_bessel_array."""
import numpy as np
import pytest
from scipy.special import jv

import cases
from oracle import Oracle, lib


def test_j0_j1_against_scipy():
    L = lib()
    xs = np.concatenate([np.linspace(-30, 30, 601), [1e-8, -1e-8, 7.999999, 8.0, 8.000001]])
    big = []
    for x in xs:
        assert abs(L.orc_bessj0(x) - jv(0, x)) < 2e-8, x
        e1 = abs(L.orc_bessj1(x) - jv(1, x))
        if abs(x) < 8.0:
            assert e1 < 2e-8, x
        else:
            # the asymptotic branch shifts the phase by the single-precision
            # literal 2.35619491 (ALPS_fns_rel.f90:1729): ~5e-7 rad
            assert e1 < 2.5e-7, x
            big.append(e1)
    assert max(big) > 5e-8   # the float literal is reproduced, not 3*pi/4
    assert L.orc_bessj0(0.0) == 1.0
    assert L.orc_bessj1(0.0) == 0.0


@pytest.mark.parametrize("n", [2, 3, 5, 10, 20, 40, 80])
def test_jn_against_scipy(n):
    L = lib()
    for x in np.concatenate([np.linspace(0.0, 25, 126)[1:], [0.003, 1e-3 * n]]):
        ref = jv(n, x)
        got = L.orc_bessj(n, x)
        if x > n:
            # upward recurrence carries the absolute error of J0/J1 (up to ~1e-7)
            assert abs(got - ref) <= 1e-6, (n, x, got, ref)
        else:
            # Miller's algorithm is accurate relative to |J_n| away from zeros
            assert abs(got - ref) <= 1e-6 * abs(ref) + 1e-9, (n, x, got, ref)
        if x <= n:
            # negative arguments (negative charge) take the same Miller branch:
            # exactly (-1)^n times the positive-argument value
            assert L.orc_bessj(n, -x) == (-1) ** n * got
    assert L.orc_bessj(n, 0.0) == 0.0


def test_negative_argument_beyond_miller_range_is_the_reference_value():
    """For x < -n the reference still runs Miller's recurrence from m(n) (the test
    "X > FLOAT(N)" is false), which is inaccurate once |x| exceeds m(n): the
    restatement keeps that behaviour (values finite, not J_n(x))."""
    L = lib()
    v = L.orc_bessj(2, -25.0)
    assert np.isfinite(v) and abs(v - jv(2, -25.0)) > 0.1


def test_jn_tiny_values_relative():
    L = lib()
    for n in [10, 20, 30]:
        for x in [1e-3, 0.05, 0.5]:
            ref = jv(n, x)
            assert abs(L.orc_bessj(n, x) / ref - 1) < 1e-6, (n, x)


def test_nmax_is_first_order_below_cutoff():
    """determine_nmax: smallest n >= 1 with max_iperp |J_n(kperp pperp/q)| <= Bessel_zero."""
    L = lib()
    p = cases.small_bimax()
    O = Oracle(p)
    kperp = 0.3
    nmax = O.set_k(kperp, 0.3)
    for s in range(p.nspec):
        xs = kperp * p.pp[s, :, 1, 0] / p.qs[s]
        bm = [max(abs(L.orc_bessj(n, x)) for x in xs) for n in range(1, nmax[s] + 1)]
        assert bm[-1] <= p.Bessel_zero
        assert all(b > p.Bessel_zero for b in bm[:-1])


def test_bessel_table_rows():
    """bessel_array(n, iperp) = J_n(kperp pperp/q), row -1 = -J_1 (ALPS_fns.f90:4241-4253)."""
    p = cases.small_bimax()
    O = Oracle(p)
    O.set_k(0.7, 0.3)
    for s in range(p.nspec):
        tab = O.bessel_table(s)
        x = 0.7 * p.pp[s, :, 1, 0] / p.qs[s]
        np.testing.assert_allclose(tab[0], -jv(1, x), atol=1e-7)
        np.testing.assert_allclose(tab[1], jv(0, x), atol=1e-7)
        np.testing.assert_allclose(tab[3], jv(2, x), atol=1e-7)
        assert np.array_equal(tab[0], -tab[2])
