"""The oracle's restatement of determine_param_fit (ALPS_analyt.f90:365-728)
pinned by what fixes its result independently of the restatement:

* tables of the fit functions' own form, whose least-squares optimum is known
  in closed form (Maxwellian with and without an exact perp_correction, a
  two-component core + beam, bi-Moyal, bi-kappa up to the row-wise degeneracy
  of its parameters, and a relativistic fit_type-5 table on Gamma rows with
  the few-point fallback), reached from guesses away from it;
* the reference's own fit report for test_kpar_fast (tests/test_kpar_fast.out
  lines 70-77);
* the Chebyshev coefficients against numpy's least squares on the same basis;
* the reset of closed-form species and the reference's error exits.
"""
import dataclasses

import numpy as np

import cases
from oracle import Oracle

TIGHT = dict(eps=0.0, maxsteps=60)  # run every row to the optimum


def _guess(p, mult, add):
    ns, mf = p.nspec, p.max_fits
    m = np.ones((ns, 5, mf))
    a = np.zeros((ns, 5, mf))
    m[:, :, :] = np.asarray(mult, dtype=float)[None, :, None] if np.ndim(mult) == 1 else mult
    a[:, :, :] = np.asarray(add, dtype=float)[None, :, None] if np.ndim(add) == 1 else add
    return cases.with_guess(p, m, a)


def _rel_err(a, b, floor=1.0):
    return np.abs(a - b) / np.maximum(np.abs(b), floor)


def test_maxwell_fit_reaches_the_closed_form_optimum():
    p = cases.fit_maxwell()
    g = _guess(p, [1.05, 0.97, 1.0, 1.0, 1.0], [0.0, 0.0, 0.02, 0.0, 0.0])
    r = Oracle(g).determine_param_fit(**TIGHT)
    assert r["rc"] == 0
    pf, ref = r["param_fit"], p.param_fit
    # u1 (amplitude) varies with pperp for the electrons, whose perp_correction
    # is not the table's: maxwell_fit_params gives the exact optimum
    assert np.ptp(ref[1, :, 0, 0]) > 0.5 * ref[1, 0, 0, 0]
    for s in range(2):
        assert np.max(_rel_err(pf[s, :, 0, 0], ref[s, :, 0, 0], 0)) < 1e-12
        assert np.max(_rel_err(pf[s, :, 1, 0], ref[s, :, 1, 0], 0)) < 1e-12
        assert np.max(np.abs(pf[s, :, 2, 0] - ref[s, :, 2, 0])) < 1e-12
    assert r["quality"] < 1e-20


def test_default_controls_stop_at_the_first_small_change():
    """epsilon_fit = 1e-8 on |LSQnew - LSQ| ends a row's fit early: from a
    guess 5 % off the optimum the parameters land near, not on, it."""
    p = cases.fit_maxwell()
    g = _guess(p, [1.05, 0.97, 1.0, 1.0, 1.0], [0.0, 0.0, 0.02, 0.0, 0.0])
    r = Oracle(g).determine_param_fit()
    e = _rel_err(r["param_fit"][:, :, 0, 0], p.param_fit[:, :, 0, 0], 0)
    assert 1e-12 < np.max(e) < 1e-2
    assert r["steps"] < 5 * (p.nperp + 1) * p.nspec


def test_core_beam_two_fits():
    p = cases.fit_core_beam()
    mult = np.ones((2, 5, 2))
    add = np.zeros((2, 5, 2))
    mult[0, :2, 0] = [1.05, 0.97]
    mult[0, :2, 1] = [0.95, 1.03]
    add[0, 2, :] = [0.02, -0.03]
    r = Oracle(cases.with_guess(p, mult, add)).determine_param_fit(**TIGHT)
    pf, ref = r["param_fit"][0], p.param_fit[0]
    for f in range(2):
        assert np.max(_rel_err(pf[:, :3, f], ref[:, :3, f])) < 1e-11, f


def test_bi_moyal_fit():
    p = cases.fit_moyal()
    g = _guess(p, [1.05, 0.97, 1.0, 1.03, 1.0], [0.0, 0.0, 0.02, 0.0, 0.0])
    r = Oracle(g).determine_param_fit(**TIGHT)
    pf, ref = r["param_fit"][0, :, :, 0], p.param_fit[0, :, :, 0]
    # row 0 (pperp = 0) does not fit the pperp parameter u4: it keeps the guess
    assert pf[0, 3] == g.param_fit[0, 0, 3, 0]
    assert np.max(_rel_err(pf[1:, :4], ref[1:, :4])) < 1e-11
    assert np.max(_rel_err(pf[0, :3], ref[0, :3])) < 1e-11


def test_kappa_fit_up_to_its_rowwise_degeneracy():
    """Within one pperp row f = u1 (c + u2 d^2)^u4 with c = 1 + pc u5 pperp^2:
    only u1 c^u4, u2 / c, u3 and u4 are fixed by the row, and eval_fit depends
    on nothing else; those reach the optimum."""
    p = cases.fit_kappa()
    g = _guess(p, [1.05, 0.97, 1.0, 1.02, 1.05], [0.0, 0.0, 0.02, 0.0, 0.0])
    r = Oracle(g).determine_param_fit(**TIGHT)
    for s in range(2):
        pperp = p.pp[s, :, 0, 0]
        pc = p.perp_correction[s, 0]

        def combos(u):
            c = 1.0 + pc * u[:, 4] * pperp ** 2
            return np.stack([u[:, 0] * c ** u[:, 3], u[:, 1] / c, u[:, 2], u[:, 3]], 1)

        a = combos(r["param_fit"][s, :, :, 0])
        b = combos(p.param_fit[s, :, :, 0])
        assert np.max(_rel_err(a, b)) < 1e-12, s
    # u5 is not fitted in row 0
    assert r["param_fit"][0, 0, 4, 0] == g.param_fit[0, 0, 4, 0]


def test_relativistic_rows_and_the_few_point_fallback():
    p = cases.fit_rel5()
    g = _guess(p, [1.05, 0.97, 1.0, 1.0, 1.0], [0.0, 0.0, 0.0, 0.0, 0.0])
    r = Oracle(g).determine_param_fit(**TIGHT)
    assert r["rc"] == 0
    pf = r["param_fit"][1, : p.ngamma + 1, :3, 0]
    # rows with fewer than 4 points inside f0_rel > -1: amplitude from the middle
    # point over exp(-pc Gamma(row, 1)), u2 = 1e-12, u3 = 0 (ALPS_analyt.f90:617-628)
    inside = (p.f0_rel[0] > -1.0).sum(axis=1)
    few = np.nonzero(inside <= 3)[0]
    assert few.size >= 1 and few[0] == 0
    for ig in few:
        j = np.nonzero(p.f0_rel[0, ig] > -1.0)[0]
        mid = (j[0] + j[-1]) // 2
        pc = p.perp_correction[1, 0]
        expect = p.f0_rel[0, ig, mid] / np.exp(-pc * p.gamma_rel[0, ig, 1])
        assert abs(pf[ig, 0] - expect) <= 1e-14 * abs(expect)
        assert pf[ig, 1] == 1e-12 and pf[ig, 2] == 0.0
    rest = np.setdiff1d(np.arange(p.ngamma + 1), few)
    u = p.param_fit[1, 0, :3, 0]
    assert np.max(_rel_err(pf[rest], np.broadcast_to(u, pf[rest].shape))) < 1e-10


def test_kpar_fast_fit_report():
    """tests/test_kpar_fast.in guesses (fit_1..3 of &ffit_1_1 and &ffit_2_1) on
    the test's bi-Maxwellian tables with the default controls: the reference
    reports param_fit at iperp = 1 as 1.7959E-001, 1.0000E+000, ~0 and
    1.4128E+004, 1.8360E+003, ~0 (test_kpar_fast.out:70-77)."""
    p = cases.kpar_fast()
    pf = np.zeros_like(p.param_fit)
    pf[0, 0, :3, 0] = [1.7959e-1, 1.0, 0.0]
    pf[1, 0, :3, 0] = [1.4128e4, 1.836e3, 0.0]
    r = Oracle(dataclasses.replace(p, param_fit=pf)).determine_param_fit()
    out = r["param_fit"][:, 1, :3, 0]
    fmt = lambda v: f"{v:.4E}"
    assert [fmt(v) for v in out[0, :2]] == ["1.7959E-01", "1.0000E+00"]
    assert [fmt(v) for v in out[1, :2]] == ["1.4128E+04", "1.8360E+03"]
    assert abs(out[0, 2]) < 1e-15 and abs(out[1, 2]) < 1e-15


def _cheb_basis(npar, order):
    yy = -1.0 + np.arange(npar + 1) * (2.0 / npar)
    return np.polynomial.chebyshev.chebvander(yy, order)


def test_chebyshev_coefficients_match_least_squares():
    p = cases.poly_case()
    f0 = p.f0.copy()
    f0[0, 5, :3] = 0.0  # zeros in a row of a logarithmic fit: 1 % of the smallest value
    p = dataclasses.replace(p, f0=f0)
    O = Oracle(p)
    r = O.determine_param_fit()
    assert r["rc"] == 0
    order = int(p.poly_order[0])
    B = _cheb_basis(p.npar, order)
    for i in range(p.nperp + 1):
        row = f0[0, i].copy()
        if np.min(row) <= 0:
            row[row == 0] = 0.01 * np.min(row[row > 0])
        c, *_ = np.linalg.lstsq(B, np.log10(row), rcond=None)
        assert np.max(np.abs(r["poly_fit_coeffs"][0, i, : order + 1] - c)) < 1e-9 * np.max(np.abs(c))
    # the Chebyshev species is reset: param_fit 0, fit_type 1, perp_correction 1
    assert np.all(r["param_fit"][0] == 0.0)
    assert np.all(r["fit_type"][0] == 1) and np.all(r["perp_correction"][0] == 1.0)


def test_closed_form_species_are_reset():
    p = cases.bm_mix()
    r = Oracle(p).determine_param_fit()
    assert r["rc"] == 0
    for s in (0, 2):
        assert np.all(r["param_fit"][s] == 0.0)
        assert np.all(r["fit_type"][s] == 1) and np.all(r["perp_correction"][s] == 1.0)


def test_error_exits():
    # alps_error(2): more parameters than points (4 kappa fits, 20 parameters, npar = 10)
    p = cases.fit_kappa(nperp=6, npar=10)
    q = cases._replace_species0(p, p.f0[0], [2, 2, 2, 2], [1.0] * 4,
                                [p.param_fit[0, 0, :, 0]] * 4)
    assert Oracle(q).determine_param_fit()["rc"] == 2
    # a singular fit matrix: with perp_correction 0 the kappa u5 derivative
    # vanishes in every row past the first
    q = cases._replace_species0(p, p.f0[0], [2], [0.0], [p.param_fit[0, 0, :, 0]])
    assert Oracle(q).determine_param_fit()["rc"] == 3
