"""The CPU restatement reproduces the reference's committed test_kpar_fast scan
(tests/test_kpar_fast.py of ALPS checks it at rtol 1e-2; here every printed
digit must agree)."""
import math
import os

import numpy as np
import pytest

import cases
from oracle import Oracle

GOLD = os.path.join(os.path.dirname(__file__), "golden", "test_kpar_fast.scan_kpara_1.root_1")
GOLD_EIGEN = os.path.join(os.path.dirname(__file__), "golden", "test_kpar_fast.eigen_kpara_1.root_1")
GOLD_HEAT = os.path.join(os.path.dirname(__file__), "golden", "test_kpar_fast.heat_kpara_1.root_1")


def eigen_rows(kperp, kpar, w, r):
    """The .eigen and .heat rows om_scan writes (ALPS_fns.f90:2567-2583):
    kperp, kpar, omega, ef, bf, Us, ds / kperp, kpar, omega, Ps, W_EM."""
    e = [kperp, kpar, w.real, w.imag]
    for a in list(r["ef"]) + list(r["bf"]) + list(np.ravel(r["Us"])) + list(r["ds"]):
        e += [a.real, a.imag]
    h = [kperp, kpar, w.real, w.imag] + list(r["Ps"]) + [float(r["W_EM"])]
    return e, h


def printed_equal(v, g):
    """v rounds to the es14.4 value g (half a unit of the last printed digit)."""
    if g == 0.0:
        return abs(v) < 1e-300
    e = math.floor(math.log10(abs(g)))
    return abs(v - g) <= 0.5 * 10.0 ** (e - 4) * (1 + 1e-9)


def kpar_fast_scan(solver):
    """refine_guess at the starting k, then om_scan(1): scan_type 4, swf = 0.1,
    32 log steps (test_kpar_fast.in)."""
    om, _ = solver.secant_osc(9.9e-3 - 5.5e-6j)
    diff = (math.log10(1e-1) - math.log10(1e-2)) / (32 * 1.0)
    return om, solver.om_scan(4, True, diff, 0.0, 32, 1, [om])


@pytest.mark.slow
def test_oracle_reproduces_kpar_fast_golden():
    gold = np.loadtxt(GOLD)
    p = cases.kpar_fast()
    O = Oracle(p)
    assert O.set_k(1e-2, 1e-2) == [21, 13]
    om, (rc, kp, kz, roots) = kpar_fast_scan(O)
    assert rc == 0
    assert gold.shape == (33, 4)
    for i in range(33):
        row = [kp[i], kz[i], roots[i, 0].real, roots[i, 0].imag]
        for v, g in zip(row, gold[i]):
            assert printed_equal(v, g), (i, row, gold[i])
    # calc_eigen at every output row: every printed digit of the .eigen and
    # .heat goldens (32 and 7 columns)
    ge, gh = np.loadtxt(GOLD_EIGEN), np.loadtxt(GOLD_HEAT)
    assert ge.shape == (33, 32) and gh.shape == (33, 7)
    for i in range(33):
        O.set_k(kp[i], kz[i])
        e, h = eigen_rows(kp[i], kz[i], roots[i, 0], O.calc_eigen(roots[i, 0]))
        for j, (v, g) in enumerate(zip(e, ge[i])):
            assert printed_equal(v, g), ("eigen", i, j, v, g)
        for j, (v, g) in enumerate(zip(h, gh[i])):
            assert printed_equal(v, g), ("heat", i, j, v, g)


def test_oracle_secant_and_rtsec_reach_the_golden_root():
    """secant (method 0) and rtsec (method 1) from the test_kpar_fast guess land on
    the root the reference printed (first row of the scan golden).  rtsec moves
    half a secant step per iteration (ALPS_fns.f90:2170), so near the root its
    error halves per step: tightening D_tol by 1e3 costs about log2(1e3) = 10
    more iterations."""
    import dataclasses
    gold = np.loadtxt(GOLD)
    p = cases.kpar_fast()
    O = Oracle(p)
    O.set_k(1e-2, 1e-2)
    w0 = 9.9e-3 - 5.5e-6j
    r, it = O.secant(w0)
    assert printed_equal(r.real, gold[0][2]) and printed_equal(r.imag, gold[0][3]), r
    assert 1 <= it < p.roots.numiter
    its = []
    for tol in (1e-10, 1e-13):
        q = dataclasses.replace(p, roots=dataclasses.replace(p.roots, D_tol=tol))
        Q = Oracle(q)
        Q.set_k(1e-2, 1e-2)
        rr, itr = Q.rtsec(w0)
        assert printed_equal(rr.real, gold[0][2]) and printed_equal(rr.imag, gold[0][3]), rr
        its.append(itr)
    assert 8 <= its[1] - its[0] <= 12, its
