"""GPU root searches against the CPU restatement and the reference golden:
secant_osc roots to |dw/w| <= 1e-10, map_search values and minima, the
test_kpar_fast k-scan, and the Fortran host binding."""
import math
import os
import subprocess

import numpy as np
import pytest

import cases
from oracle import Oracle
from test_oracle_golden import GOLD, GOLD_EIGEN, GOLD_HEAT, eigen_rows, kpar_fast_scan, printed_equal

pytestmark = pytest.mark.gpu

ROOT_TOL = 1e-10


def _engine(p):
    from alps_amd.engine import Engine
    return Engine(p)


def test_secant_osc_roots_match_oracle():
    p = cases.small_bimax()
    p.roots.D_threshold = 1e-14
    O, E = Oracle(p), _engine(p)
    O.set_k(0.3, 0.3)
    E.set_k(0.3, 0.3)
    guesses = [0.29 - 0.001j, 0.5 - 0.05j, 1.1 - 0.1j, 0.05 - 0.02j]
    g_roots, g_dfin, g_it = E.secant_osc_batch(guesses)
    for k, w0 in enumerate(guesses):
        o_root, o_it = O.secant_osc(w0)
        assert abs(g_roots[k] - o_root) <= ROOT_TOL * abs(o_root), (w0, g_roots[k], o_root)
        # a root in the batch equals the same root solved alone
        alone, _ = E.secant_osc(w0)
        assert abs(alone - g_roots[k]) <= 1e-13 * abs(alone)


@pytest.mark.parametrize("method", [0, 1])
def test_secant_and_rtsec_batches_match_oracle(method):
    """secant (ALPS_fns.f90:1815) and rtsec (:2105) as batched host state
    machines on the device disp agree with the restatement root by root."""
    import dataclasses
    p = cases.small_bimax()
    p = dataclasses.replace(p, roots=dataclasses.replace(p.roots, secant_method=method,
                                                         D_threshold=1e-14, D_tol=1e-13))
    O, E = Oracle(p), _engine(p)
    O.set_k(0.3, 0.3)
    E.set_k(0.3, 0.3)
    guesses = [0.29 - 0.001j, 0.5 - 0.05j, 1.1 - 0.1j]
    g_roots, g_dfin, g_it = E.solve_batch(guesses)
    for k, w0 in enumerate(guesses):
        o_root, o_it = O.secant(w0) if method == 0 else O.rtsec(w0)
        assert abs(g_roots[k] - o_root) <= ROOT_TOL * abs(o_root), (w0, g_roots[k], o_root)
        assert abs(g_dfin[k] - E.disp(g_roots[k])) == 0.0


def test_map_search_matches_oracle():
    p = cases.small_bimax()
    O, E = Oracle(p), _engine(p)
    O.set_k(0.3, 0.3)
    E.set_k(0.3, 0.3)
    args = (0.05, 1.5, -0.4, 0.05, 18, 12)
    om_o, cal_o, val_o, ir_o, n_o = O.map_search(*args)
    om_g, cal_g, val_g, ir_g, n_g = E.map_search(*args)
    assert np.array_equal(om_o, om_g)
    assert np.max(np.abs(cal_g - cal_o) / np.abs(cal_o)) <= 1e-10
    assert np.max(np.abs(val_g - val_o)) <= 1e-10
    assert n_o == n_g and np.array_equal(ir_o, ir_g)


@pytest.mark.parametrize("loggrid", [(0, 0), (1, 0), (0, 1)])
def test_map_grid_on_the_device_is_the_returned_grid(loggrid):
    """det D of the map equals det D of a batch at the returned grid, bit for
    bit (the same launch shape): the device-built linear grid rounds like the
    host's, and log grids are built on the host."""
    p = cases.small_bimax()
    E = _engine(p)
    E.set_k(0.3, 0.3)
    gam = (0.02, 0.4) if loggrid[1] else (-0.4, 0.05)
    om, cal, val, _, _ = E.map_search(0.05, 1.5, gam[0], gam[1], 18, 12,
                                      loggridw=bool(loggrid[0]), loggridg=bool(loggrid[1]))
    det = E.disp_batch(om.T.ravel())
    assert np.array_equal(det, cal.T.ravel())
    E.close()


def test_refine_after_map_matches_oracle():
    p = cases.small_bimax()
    p.roots.D_threshold = 1e-14
    O, E = Oracle(p), _engine(p)
    O.set_k(0.3, 0.3)
    E.set_k(0.3, 0.3)
    om, cal, val, iroots, n = E.map_search(0.05, 1.5, -0.4, 0.05, 18, 12)
    guesses = [om[i - 1, j - 1] for i, j in iroots]
    assert len(guesses) >= 1
    roots, dfin = E.refine_guess(guesses)
    for g, r in zip(guesses, roots):
        ro, _ = O.secant_osc(g)
        assert abs(r - ro) <= ROOT_TOL * abs(ro)


@pytest.mark.parametrize("method", [2, 0, 1])
def test_map_search_refine_equals_map_then_batch(method):
    """alpsg_map_search_refine (the reference's map_search with its
    refine_guess call) equals alpsg_map_search followed by the batched solver
    on the minima's grid frequencies, bit for bit: grids, minima, roots, D at
    the roots and iteration counts; roots to the oracle's at 1e-10.  Also with
    fewer recorded minima than found, and twice in a row into the same arrays
    (the device-stepped solver copies the grids while its rounds run)."""
    import dataclasses
    p = cases.config1(64, 128)
    p = dataclasses.replace(p, roots=dataclasses.replace(p.roots, secant_method=method,
                                                         D_threshold=1e-14))
    E = _engine(p)
    E.set_k(0.5, 0.5)
    args = (0.01, 2.0, -0.5, 0.05, 48, 40)
    om, cal, val, ir, n = E.map_search(*args, numroots=6)
    assert n >= 2
    k = min(n, 6)
    g = np.array([om[i - 1, j - 1] for i, j in ir])
    r_b, d_b, it_b = E.solve_batch(g)
    N = args[4] * args[5]
    out = (np.zeros(N, dtype=np.complex128), np.zeros(N, dtype=np.complex128), np.zeros(N))
    for rep in range(2):
        om2, cal2, val2, ir2, n2, r2, d2, it2 = E.map_search_refine(*args, numroots=6, out=out)
        assert n2 == n and np.array_equal(ir2, ir)
        assert np.array_equal(om2, om) and np.array_equal(cal2, cal) and np.array_equal(val2, val)
        assert np.array_equal(r2, r_b) and np.array_equal(d2, d_b) and np.array_equal(it2, it_b)
        t = E.last_root_times_ms(k)
        assert np.all(t > 0.0) and np.all(np.isfinite(t))
    # fewer recorded minima than found
    _, _, _, ir3, n3, r3, _, _ = E.map_search_refine(*args, numroots=1)
    assert n3 == n and ir3.shape == (1, 2) and r3[0] == r_b[0]
    E.close()
    if method == 2:
        # against the oracle on two minima away from the map's Re w = 0.01
        # edge (a minimum there walks into the flat |D| region near w = 0,
        # where the secant's end point is ill-conditioned)
        O = Oracle(p)
        O.set_k(0.5, 0.5)
        pick = [q for q in range(k) if g[q].real > 0.05][:2]
        assert pick
        for q in pick:
            # (iteration counts are not compared: at D_threshold 1e-14 |D|
            # reaches the threshold within the rounding of det D)
            ro, _ = O.secant_osc(g[q])
            assert abs(r_b[q] - ro) <= ROOT_TOL * abs(ro), (q, r_b[q], ro)


@pytest.mark.slow
def test_kpar_fast_scan_on_gpu_matches_golden_and_oracle():
    gold = np.loadtxt(GOLD)
    p = cases.kpar_fast()
    E = _engine(p)
    assert E.set_k(1e-2, 1e-2) == [21, 13]
    om, (kp, kz, roots) = kpar_fast_scan(E)
    for i in range(33):
        row = [kp[i], kz[i], roots[i, 0].real, roots[i, 0].imag]
        for v, g in zip(row, gold[i]):
            assert printed_equal(v, g), (i, row, gold[i])
    O = Oracle(p)
    O.set_k(1e-2, 1e-2)
    _, (rc, okp, okz, oroots) = kpar_fast_scan(O)
    rel = np.abs(roots[:, 0] - oroots[:, 0]) / np.abs(oroots[:, 0])
    assert np.max(rel) <= ROOT_TOL, rel
    # eigenfunctions and heating rates on the device path, against the goldens
    ge, gh = np.loadtxt(GOLD_EIGEN), np.loadtxt(GOLD_HEAT)
    for i in range(33):
        E.set_k(kp[i], kz[i])
        r = E.calc_eigen([roots[i, 0]])
        e, h = eigen_rows(kp[i], kz[i], roots[i, 0], {k: v[0] for k, v in r.items()})
        for j, (v, g) in enumerate(zip(e, ge[i])):
            assert printed_equal(v, g), ("eigen", i, j, v, g)
        for j, (v, g) in enumerate(zip(h, gh[i])):
            # Ps and W_EM difference chi over a 1e-6 step in omega, which lifts
            # chi's 1e-14 device/CPU differences to ~1e-8: a value on a rounding
            # boundary of the fifth printed digit may land one unit away
            ok = printed_equal(v, g) if j < 4 else abs(v - g) <= 1.0 * 10.0 ** (
                math.floor(math.log10(abs(g))) - 4) * (1 + 1e-9)
            assert ok, ("heat", i, j, v, g)


@pytest.mark.parametrize("case", ["drift", "three"])
def test_calc_eigen_matches_oracle(case):
    """Batched calc_eigen against the restatement: eigenfunctions to 1e-9;
    heating rates to 1e-5 (their wave energy differentiates chi over a 1e-6
    step in omega, which scales chi's 1e-14 differences by 1e6)."""
    p = cases.small_bimax(drift=0.7, alph=2.5) if case == "drift" else cases.three_species()
    O, E = Oracle(p), _engine(p)
    O.set_k(0.3, 0.3)
    E.set_k(0.3, 0.3)
    oms = [0.5 - 0.05j, 0.8 - 0.02j, 1.3 - 0.1j]
    g = E.calc_eigen(oms)
    for r, om in enumerate(oms):
        o = O.calc_eigen(om)
        for key, tol in (("ef", 1e-9), ("bf", 1e-9), ("Us", 1e-9), ("ds", 1e-9),
                         ("Ps", 1e-5), ("Ps_split", 1e-5)):
            a, b = np.asarray(g[key][r]), np.asarray(o[key])
            assert np.max(np.abs(a - b)) <= tol * max(np.max(np.abs(b)), 1e-300), (key, om, a, b)
        assert abs(g["W_EM"][r] - o["W_EM"]) <= 1e-5 * abs(o["W_EM"])


def test_fortran_host_matches_python_host():
    from alps_amd import build
    exe = os.path.join(os.path.dirname(build.__file__), "fortran", "alpsg_demo")
    if not os.path.exists(exe):
        pytest.skip("Fortran demo not built")
    out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    lines = out.stdout.split("\n")
    nmax = [int(x) for x in next(l for l in lines if l.startswith("NMAX")).split()[1:]]
    disp = [[float(x) for x in l.split()[1:]] for l in lines if l.startswith("DISP")]
    root = [float(x) for x in next(l for l in lines if l.startswith("ROOT")).split()[1:]]
    p = cases.small_bimax()
    E = _engine(p)
    assert E.set_k(0.3, 0.3) == nmax
    for r in disp:
        d = E.disp(complex(r[0], r[1]))
        assert abs(d - complex(r[2], r[3])) <= 1e-13 * abs(d)
    p.roots.D_threshold = 1e-15
    w, _ = E.secant_osc(0.5 - 0.05j)
    assert abs(w - complex(*root)) <= 1e-12 * abs(w)
    # om_double_scan through the Fortran interface equals the Python host's
    from test_oracle_double_scan import K0, S1, S2
    rows = [l.split() for l in lines if l.startswith("DSCAN")]
    assert len(rows) == (S1[4] + 1) * (S2[4] + 1)
    p.roots.D_threshold = 1e-14
    E.set_k(*K0)
    om0, _ = E.secant_osc(0.048 - 0.0001j)
    r = E.om_double_scan(S1, S2, [om0])
    for k, row in enumerate(rows):
        i1, i2 = divmod(k, S2[4] + 1)
        kp, kz, re, im = (float(x) for x in row[2:6])
        assert int(row[6]) == int(r["written"][i1, i2, 0])
        assert abs(kp - r["kperp"][i1, i2]) <= 1e-15 * kp and abs(kz - r["kpar"][i1, i2]) <= 1e-15 * kz
        assert abs(complex(re, im) - r["roots"][i1, i2, 0]) <= 1e-14 * abs(complex(re, im))
    # the root server through the Fortran interface: two solves and the counters
    served = {int(l.split()[1]): l.split()[2:] for l in lines if l.startswith("SERVED")}
    E.set_k(0.3, 0.3)
    for tag, g in ((11, 0.29 - 0.001j), (12, 0.5 - 0.05j)):
        w, it = E.secant_osc(g)
        re, im, its, st = served[tag]
        assert abs(complex(float(re), float(im)) - w) <= 1e-12 * abs(w)
        assert int(its) == it and int(st) == (0 if it < p.roots.numiter else 6)
    stats = [int(x) for x in next(l for l in lines if l.startswith("STATS")).split()[1:]]
    assert stats[0] == 2 and stats[1] == 2 and stats[2] >= 1 and stats[3] >= stats[2]
