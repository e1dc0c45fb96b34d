"""GPU parity at the full grid sizes of the benchmark workloads, against the
oracle on the same tables (a few frequencies each, so the oracle finishes in
seconds):

* the batched map workload: 2 species, 128 x 256, kperp = kpar = 0.5
  (nmax 47/20), frequencies spread over the map's range, and a map_search on
  a 6 x 5 sub-grid;
* the non-Maxwellian workload: kappa proton core (kappa 4, T_perp/T_par = 2),
  a 10 % kappa beam drifting at 1.5 v_A (kappa 5) and electrons, 256 x 512;
* the relativistic workload: a Juettner species (kT = 0.3 m c^2) on a
  200 x 400 (pperp, ppar) grid and a 100 x 200 (Gamma, pparbar) grid, beside a
  non-relativistic species;
* the served workload at its largest grid (3 species, 256 x 512): served
  roots equal the one-handle solver's.

Bars as in test_gpu_disp.py: chi per species 1e-11 relative (norm-wise), det D
1e-11 of its terms' magnitude; map values 1e-10 with identical minima; roots
1e-10.

The reference's algorithm is discontinuous in omega where a resonance falls on
a p_par grid node or a principal-value ladder point on a cell midpoint (the
interpolation cell of funct_g changes there): with kpar = 0.3 on a 0.025 grid,
omega = 0.3 + 0.05i (p_res = 1.0, a node) moves chi by 5e-3 in the oracle when
omega moves by 1e-15, and at such points two evaluations that round
differently agree to about 1e-10 only.  The relativistic case therefore uses
frequencies away from such points and checks first that the oracle is
continuous there.
"""
import dataclasses

import numpy as np
import pytest

import cases
from oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _disp_parity(p, kperp, kpar, oms, continuity=False):
    from alps_amd.engine import Engine
    O = Oracle(p)
    E = Engine(p)
    assert O.set_k(kperp, kpar) == E.set_k(kperp, kpar)
    det, chi, low, wave = E.disp_batch(np.array(oms), want_tensors=True)
    for k, om in enumerate(oms):
        d_o, chi_o, low_o, wave_o = O.disp(om, want_tensors=True)
        if continuity:
            # the oracle itself is continuous at om (no cell or ladder tie)
            chi_p = O.disp(om * (1.0 + 1e-13), want_tensors=True)[1]
            for s in range(p.nspec):
                assert np.max(np.abs(chi_p[s] - chi_o[s])) <= 1e-11 * np.max(np.abs(chi_o[s]))
        for s in range(p.nspec):
            e = np.max(np.abs(chi[k, s] - chi_o[s])) / np.max(np.abs(chi_o[s]))
            assert e <= TOL, (om, s, e)
        assert abs(det[k] - d_o) / cases.det_scale(wave_o) <= TOL, om
    E.close()


def test_map_workload_grid():
    p = cases.config1()
    oms = [0.01 - 0.5j, 0.6 - 0.05j, 1.2 - 0.2j, 2.0 + 0.05j, 0.3 + 0.0j, 1.7 - 0.45j]
    _disp_parity(p, 0.5, 0.5, oms)


def test_map_workload_sub_grid_search():
    from alps_amd.engine import Engine
    p = cases.config1()
    O = Oracle(p)
    E = Engine(p)
    O.set_k(0.5, 0.5)
    E.set_k(0.5, 0.5)
    args = (0.3, 0.9, -0.3, 0.0, 6, 5)
    om_g, cal_g, val_g, roots_g, n_g = E.map_search(*args, numroots=10)
    r = O.map_search(*args, numroots=10)
    om_o, cal_o, val_o, roots_o, n_o = r[:5]
    assert np.array_equal(om_g, om_o)
    assert np.max(np.abs(val_g - val_o) / np.maximum(np.abs(val_o), 1.0)) <= 1e-10
    assert n_g == n_o and np.array_equal(np.asarray(roots_g), np.asarray(roots_o))
    E.close()


def _kappa_core_beam(kappas=(4.0, 5.0, 4.0)):
    from alps_amd.plasma import kappa_plasma
    return kappa_plasma(ms=[1.0, 1.0, cases.M_E], qs=[1.0, 1.0, -1.0], ns=[0.9, 0.1, 1.0],
                        taus=[1.0, 1.0, 1.0], alphs=[2.0, 2.0, 1.0], ps=[0.0, 1.5, 0.0],
                        kappas=list(kappas), beta=1.0, vA=1e-4, nperp=256, npar=512, maxP=6.0,
                        positions_principal=3, n_resonance_interval=100, Tlim=0.01,
                        Bessel_zero=1e-45)


def test_kappa_core_beam_workload_grid():
    _disp_parity(_kappa_core_beam(), 0.3, 0.3, [0.6 - 0.05j, 0.2 - 0.001j, 1.1 + 0.02j])


@pytest.mark.parametrize("kappa", [3.0, 6.0])
def test_kappa_core_beam_extreme_kappas(kappa):
    """The ends of the workload's kappa range, every species at that kappa."""
    _disp_parity(_kappa_core_beam((kappa, kappa, kappa)), 0.3, 0.3,
                 [0.6 - 0.05j, 0.2 - 0.001j, 1.1 + 0.02j])


@pytest.mark.parametrize("tke,oms", [
    (0.1, [0.1313 - 0.0041j, 0.3071 + 0.0213j]),
    (0.3, [0.1313 - 0.0041j, 0.3071 + 0.0213j, 0.8133 - 0.0517j]),
    (1.0, [0.4519 + 0.0311j, 0.5923 - 0.0087j, 0.0973 - 0.0011j])])
def test_relativistic_workload_at_each_temperature(tke, oms):
    """Juettner electrons at kT/mc^2 = 0.1, 0.3 and 1.0 beside cold protons,
    200 x 400 (pperp, ppar) and 100 x 200 (Gamma, pparbar): the served
    relativistic workload's plasmas at kpar = 0.5, kperp = 0.05.  At kT = m c^2
    the oracle itself moves by 1e-9 .. 1e-7 under a 1e-13 relative change of
    omega at most frequencies (a principal-value window of some Gamma row
    sits on a cell boundary); the frequencies here are the quietest found, and
    the bar is 1e-11 or ten times the oracle's own jump there, whichever is
    larger."""
    from alps_amd.engine import Engine
    from alps_amd.plasma import RootParams
    from bench import serve_rel_variant
    p = serve_rel_variant(tke, RootParams(secant_method=2, numiter=50, D_threshold=1e-12))
    O, E = Oracle(p), Engine(p)
    assert O.set_k(0.05, 0.5) == E.set_k(0.05, 0.5)
    det, chi, low, wave = E.disp_batch(np.array(oms), want_tensors=True)
    for k, om in enumerate(oms):
        d_o, chi_o, low_o, wave_o = O.disp(om, want_tensors=True)
        chi_p = O.disp(om * (1.0 + 1e-13), want_tensors=True)[1]
        for s in range(p.nspec):
            sc = np.max(np.abs(chi_o[s]))
            jump = np.max(np.abs(chi_p[s] - chi_o[s])) / sc
            e = np.max(np.abs(chi[k, s] - chi_o[s])) / sc
            print(tke, om, s, "err", e, "oracle jump", jump, flush=True)
            assert e <= max(TOL, 10 * jump), (tke, om, s, e, jump)
    E.close()


def test_relativistic_workload_grid():
    from alps_amd.plasma import bimaxwellian_plasma, juttner_plasma
    r = juttner_plasma(ms=[1.0], qs=[-1.0], ns=[1.0], taus=[0.3], beta=1.0, vA=1.0, nperp=200,
                       npar=400, maxP=5.0, ngamma=100, npparbar=200, analytic_rel=True,
                       positions_principal=3, n_resonance_interval=100, Tlim=0.01,
                       Bessel_zero=1e-45)
    b = bimaxwellian_plasma(ms=[1.0], qs=[1.0], ns=[1.0], taus=[1.0], alphs=[1.0], ps=[0.0],
                            beta=1.0, vA=1.0, nperp=200, npar=400, maxP=5.0)
    npf = r.param_fit.shape[1]
    pf = np.zeros((2, npf, 5, 1))
    pf[0, : b.param_fit.shape[1]] = b.param_fit[0]
    pf[1] = r.param_fit[0]
    p = dataclasses.replace(
        r, pp=np.concatenate([b.pp, r.pp]), f0=np.concatenate([b.f0, r.f0]), ns=[1.0, 1.0],
        qs=[1.0, -1.0], ms=[1.0, 1.0], fit_type=np.array([[1], [4]]),
        perp_correction=np.array([[b.perp_correction[0, 0]], [r.perp_correction[0, 0]]]),
        param_fit=pf, relativistic=[0, 1], ACmethod=None, n_fits=None, poly_kind=None,
        poly_order=None, logfit=None, poly_log_max=None, poly_fit_coeffs=None, usebM=None,
        bM_nmax=None, bM_Bessel_zero=None, bM_beta=None, bM_alpha=None, bM_pdrift=None)
    _disp_parity(p, 0.3, 0.3, [0.5137 - 0.0471j, 0.3137 + 0.0471j, 0.8071 + 0.0j, 0.6211 + 0.0123j],
                 continuity=True)


def test_served_workload_largest_grid():
    """Served solves at 256 x 512 (protons, electrons, drifting alphas) equal
    the one-handle solver's roots."""
    from alps_amd.engine import Engine, RootServer
    p = cases.three_species(nperp=256, npar=512)
    p = dataclasses.replace(p, roots=dataclasses.replace(p.roots, secant_method=2,
                                                         D_threshold=1e-14))
    reqs = [(0.3, 0.3, 0.45 - 0.01j), (0.05, 1.0, 0.9 - 0.05j), (1.0, 0.2, 0.25 - 0.001j)]
    S = RootServer(hbm_bytes=8e9, max_slots=16)
    for i, (kp, kz, g) in enumerate(reqs):
        S.submit(p, kp, kz, g, p.roots, tag=i)
    got = S.drain(len(reqs))
    S.close()
    E = Engine(p)
    for i, (kp, kz, g) in enumerate(reqs):
        E.set_k(kp, kz)
        w, _ = E.secant_osc(g, p.roots)
        assert got[i][3] == 0
        assert abs(got[i][1] - w) <= 1e-10 * abs(w), (i, got[i][1], w)
    E.close()
