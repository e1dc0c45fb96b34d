"""Pins of the restated root finders (secant ALPS_fns.f90:1815-1917, rtsec
:2105-2195, secant_osc :1919-2101) on closed-form functions whose iterates
are known in closed form: the secant step of a linear function lands on its
root; on w^2 - c the secant iterates obey q(k+1) = q(k) q(k-1) with
q = (w - sqrt c)/(w + sqrt c); rtsec's half step halves the error of a linear
function each step; secant_osc's jump cap moves a far start by
float(0.1) |w| per step.  The solvers take the function as an argument (as
the reference's rtsec(func, ...) does), so these are the same code that
solves disp."""
import cmath

import numpy as np
import pytest

from oracle import solve_test

EPS = np.finfo(float).eps
F32_TENTH = float(np.float32(0.1))   # the default-real literal 0.1 of the source


def _distinct(traj):
    """The successive evaluation points, without immediate repeats."""
    out = [traj[0]]
    for w in traj[1:]:
        if w != out[-1]:
            out.append(w)
    return np.array(out)


def test_secant_linear_lands_on_the_root():
    a, r, w0, prec = 1.3 - 0.4j, 0.7 - 0.05j, 0.5 - 0.01j, 1e-5
    root, it, traj = solve_test("secant", "linear", w0, a=a, r1=r, D_prec=prec)
    # D(prevom = w0 (1 - D_prec)), D(w0) twice (before and in the loop), then the root
    assert traj[0] == w0 * (1 - prec) and traj[1] == w0 and traj[2] == w0
    assert it == 2 and len(traj) == 4
    # the step divides differences of relative size D_prec: rounding grows by 1/D_prec
    bound = 8 * EPS / prec * abs(w0 - r)
    assert abs(traj[3] - r) <= bound and abs(root - r) <= bound, (root, r)


@pytest.mark.parametrize("c,w0", [(0.49 - 0.1j, 0.6 - 0.05j), (-0.3 + 0.8j, 0.5 + 0.7j)])
def test_secant_quadratic_iterates_follow_the_product_recurrence(c, w0):
    sc = cmath.sqrt(c)
    root, it, traj = solve_test("secant", "quadratic", w0, r2=c, D_threshold=1e-13)
    x = _distinct(traj)
    q = (x - sc) / (x + sc)
    checked = 0
    for k in range(1, len(x) - 1):
        if abs(q[k]) < 1e-9:
            break   # D is at rounding level once the iterate sits on the root
        # how far x[k+1] lies from the point the recurrence gives, in units of
        # x: a few ulps; the first step differences two points D_prec apart
        # and carries rounding / D_prec
        dev = abs(q[k + 1] / (q[k] * q[k - 1]) - 1) * abs(x[k + 1] - sc)
        assert dev <= 4 * EPS * (1 / 1e-5 if k == 1 else 1.0), (k, dev / EPS)
        checked += 1
    assert checked >= 4
    assert abs(root - sc) <= 1e-12 * abs(sc) and abs(root * root - c) < 1e-13


def test_secant_max_iterations_returns_the_update_after_the_best_point():
    """With numiter spent, secant returns minom, recorded as the updated om
    after the smallest |D| (ALPS_fns.f90:1902-1905, 1910-1911): on w^2 - c it is
    the next secant iterate q = q(last) q(last - 1), a point never evaluated."""
    c = 0.49 - 0.1j
    sc = cmath.sqrt(c)
    root, it, traj = solve_test("secant", "quadratic", 0.9 - 0.2j, r2=c, numiter=3,
                                D_threshold=0.0)
    assert it == 3
    x = _distinct(traj)
    assert not np.any(traj == root)
    q = lambda w: (w - sc) / (w + sc)
    want = q(x[-1]) * q(x[-2])
    assert abs(q(root) / want - 1) * abs(root - sc) <= 4 * EPS
    # |D| fell at every step, so the best point is the last one evaluated
    d = np.abs(x * x - c)
    assert np.all(np.diff(d[1:]) < 0)


def test_rtsec_linear_halves_the_error_every_step():
    a, r, w0, tol = 1.3 - 0.4j, 0.7 - 0.05j, 0.5 - 0.01j, 1e-9
    root, iflag, traj = solve_test("rtsec", "linear", w0, a=a, r1=r, D_tol=tol)
    # x1 = w0, x2 = w0 (1 + D_prec); the one with the smaller |f| starts
    assert traj[0] == w0 and traj[1] == w0 * (1 + 1e-5)
    start = traj[1] if abs(traj[1] - r) < abs(traj[0] - r) else traj[0]
    e = np.abs(traj[2:] - r)
    e0 = abs(start - r)
    for k, ek in enumerate(e):
        want = e0 / 2.0 ** (k + 1)
        assert abs(ek - want) <= 1e-6 * want + 64 * EPS, (k, ek, want)
    # the loop stops at the first full step |dx| = e(j-1) below D_tol
    j = 1
    while e0 / 2.0 ** (j - 1) >= tol:
        j += 1
    assert iflag == j and len(e) == j
    assert abs(root - r) <= tol


def test_rtsec_rational_reaches_the_nearer_root():
    r1, r2, pole = 0.4 - 0.02j, 1.5 + 0.1j, 0.9 - 0.3j
    root, iflag, traj = solve_test("rtsec", "rational", 0.45 - 0.03j, a=2.0, r1=r1, r2=r2,
                                   pole=pole, D_tol=1e-12)
    assert iflag < 49 and abs(root - r1) <= 1e-11


def test_secant_osc_jump_cap_is_a_tenth_of_om_in_single_precision():
    """Far from the root the secant jump of a linear function (the exact
    distance to the root) exceeds the cap, so each step moves exactly
    float(0.1) |w| towards the root (ALPS_fns.f90:2062-2064), halved where |D|
    grew over the previous point (:2066-2068: the first step, whose previous
    point w0 (1 - D_prec) lies nearer the root); the last step lands on it."""
    a, r, w0 = 0.8 + 0.3j, 0.2 - 0.01j, 1.0 - 0.2j
    root, it, traj = solve_test("secant_osc", "linear", w0, a=a, r1=r, D_threshold=1e-12)
    # D(w0), D(prevom = w0 (1 - D_prec)), then D(om) once per iteration
    assert traj[0] == w0 and traj[1] == w0 * (1 - 1e-5)
    x = _distinct(traj[2:])
    assert x[0] == w0
    capped = halved = 0
    for k in range(len(x) - 1):
        full = x[k] - r
        prev = traj[1] if k == 0 else x[k - 1]
        half = 0.5 if abs(x[k] - r) > abs(prev - r) else 1.0
        if abs(full) <= F32_TENTH * abs(x[k]):
            assert abs(x[k + 1] - (x[k] - half * full)) <= 1e-12
            break
        halved += half < 1
        want = x[k] - half * F32_TENTH * abs(x[k]) * full / abs(full)
        assert abs(x[k + 1] - want) <= 1e-13 * abs(want), (k, x[k + 1], want)
        # the double literal 0.1 would move it by a different amount
        other = x[k] - half * 0.1 * abs(x[k]) * full / abs(full)
        assert abs(x[k + 1] - other) > 100 * abs(x[k + 1] - want)
        capped += 1
    assert capped >= 5 and halved == 1 and abs(root - r) <= 1e-12
