"""GPU parity of the batched disp() kernel against the CPU restatement.

Bar: chi per species agrees norm-wise to <= 1e-11 relative
(max|chi_gpu - chi_cpu| / max|chi_cpu|), chi_low likewise, and det D to
<= 1e-11 of the magnitude of the determinant's terms (det_scale), for
omegas covering every branch of full_integrate (ALPS_fns.f90:750-792).
"""
import numpy as np
import pytest

import cases
from oracle import Oracle

pytestmark = pytest.mark.gpu

TOL = 1e-11


def _engine(p):
    from alps_amd.engine import Engine
    return Engine(p)


def _check(p, kperp, kpar, oms, tol=TOL):
    O = Oracle(p)
    E = _engine(p)
    n_o = O.set_k(kperp, kpar)
    n_g = E.set_k(kperp, kpar)
    assert n_o == n_g, (n_o, n_g)
    det_g, chi_g, low_g, wave_g = E.disp_batch(oms, want_tensors=True)
    worst = 0.0
    for k, om in enumerate(oms):
        d_o, chi_o, low_o, wave_o = O.disp(om, want_tensors=True)
        for s in range(p.nspec):
            sc = np.max(np.abs(chi_o[s]))
            e = np.max(np.abs(chi_g[k, s] - chi_o[s])) / sc
            el = np.max(np.abs(low_g[k, s] - low_o[s])) / max(np.max(np.abs(low_o[s])), 1e-300)
            worst = max(worst, e, el)
            assert e <= tol, (om, s, e)
            assert el <= tol, (om, s, el)
        ew = np.max(np.abs(wave_g[k] - wave_o)) / np.max(np.abs(wave_o))
        ed = abs(det_g[k] - d_o) / cases.det_scale(wave_o)
        assert ew <= tol, (om, ew)
        assert ed <= tol, (om, ed)
        worst = max(worst, ew, ed)
    E.close()
    return worst


@pytest.mark.parametrize("kperp_norm", [True, False])
def test_branches_bimax(kperp_norm):
    p = cases.small_bimax(kperp_norm=kperp_norm)
    _check(p, 0.3, 0.3, cases.BRANCH_OMEGAS)


def test_branches_drift_aniso():
    p = cases.small_bimax(drift=0.7, alph=2.5)
    _check(p, 0.3, 0.3, cases.BRANCH_OMEGAS)


def test_three_species():
    p = cases.three_species()
    _check(p, 0.5, 0.2, cases.BRANCH_OMEGAS)


def test_kappa_fits():
    p = cases.kappa_case()
    _check(p, 0.3, 0.3, cases.BRANCH_OMEGAS)


def test_bi_moyal_fit():
    """A bi-Moyal species (fit type 6, ALPS_analyt.f90:256-260): its Landau
    term takes the general continuation."""
    _check(cases.fit_moyal(), 0.3, 0.3, cases.BRANCH_OMEGAS)


@pytest.mark.parametrize("case", ["kappa_poly_mix", "kappa_two_fits"])
def test_kappa_landau_form_beside_the_general_form(case):
    """A species whose every fit is a kappa fit takes the Landau form with the
    kappa evaluation inlined (a kernel of its own), the others the general
    form (a call of eval_fit): beside a Chebyshev species both run in one
    pass.  Against the oracle to the disp bar; every resonant lane count, the
    separate and the fused (small-batch) resonant kernels agree to rounding."""
    from alps_amd.engine import Engine
    p = getattr(cases, case)()
    _check(p, 0.3, 0.3, cases.BRANCH_OMEGAS)
    E = Engine(p)
    E.set_k(0.3, 0.3)
    oms = np.array(cases.BRANCH_OMEGAS * 3)
    n = oms.size
    ref, chi_r, _, wave = E.disp_batch(oms, want_tensors=True)
    scale = np.max(np.abs(chi_r), axis=(2, 3))
    for i, nb in enumerate(Engine.RESONANT_LANES):
        nch, w = Engine.SWEEP_VARIANTS[i % len(Engine.SWEEP_VARIANTS)]
        E.set_shape(n, (nch, w, 1, nb, 1 + i % 3))
        _, chi, _, _ = E.disp_batch(oms, want_tensors=True)
        e = np.max(np.abs(chi - chi_r), axis=(2, 3)) / scale
        assert np.all(e <= 1e-12), (nb, e.max())
    dsc = np.array([cases.det_scale(x) for x in wave])
    for m in (1, 3, 16):
        d = E.disp_batch(oms[:m])
        assert np.all(np.abs(d - ref[:m]) / dsc[:m] <= 1e-12), m
    E.close()


@pytest.mark.parametrize("case", ["poly_case", "poly_row0_clipped"])
def test_chebyshev_continuation(case):
    """Chebyshev continuation, including the Landau early return
    (ALPS_fns.f90:1404-1407; the oracle's counter shows it is taken) and, with
    the pperp = 0 row clipped to zero, the zero at row iperp - 1 that the
    source does not test."""
    from oracle import branch_counts, reset_branch_counts
    p = getattr(cases, case)()
    reset_branch_counts()
    _check(p, 0.3, 0.3, cases.BRANCH_OMEGAS)
    assert branch_counts()["landau_zero"] > 0


@pytest.mark.parametrize("kperp_norm,cold", [(True, True), (False, False)])
def test_closed_form_species(kperp_norm, cold):
    """use_bM species (NHDS calc_chi, hot and cold) beside a tabulated one."""
    p = cases.bm_mix(kperp_norm=kperp_norm, cold=cold)
    _check(p, 0.5, 0.2, cases.BRANCH_OMEGAS)
    _check(p, 0.5, -0.2, cases.BRANCH_OMEGAS[:4])


def test_single_omega_paths_agree():
    """The plain one-omega-per-launch path gives the batched results."""
    from alps_amd.engine import Engine
    p = cases.small_bimax()
    E = Engine(p)
    E.set_k(0.3, 0.3)
    oms = np.array(cases.BRANCH_OMEGAS)
    a = E.disp_batch(oms)
    b = E.disp_plain(oms)
    assert np.all(np.abs(a - b) <= 1e-12 * np.abs(a) + 1e-300)


def test_replayed_small_batches_match_the_large_batch():
    """Batches of up to 16 omegas (det only) replay a captured graph of their
    disp chain: every replay, interleaved with large eager batches and with a
    change of k, equals the large-batch evaluation of the same omegas."""
    from alps_amd.engine import Engine
    p = cases.three_species()
    E = Engine(p)
    oms = np.array(cases.BRANCH_OMEGAS * 3)
    assert oms.size > 16
    for k in ((0.5, 0.2), (0.3, 0.3)):
        E.set_k(*k)
        ref, chi, low, wave = E.disp_batch(oms, want_tensors=True)
        scale = np.array([cases.det_scale(w) for w in wave])
        for rep in range(3):
            for n in (1, 3, 16):
                for i0 in range(0, 16, n):
                    d = E.disp_batch(oms[i0:i0 + n])
                    e = np.abs(d - ref[i0:i0 + n]) / scale[i0:i0 + n]
                    assert np.all(e <= 1e-12), (k, rep, n, i0, e.max())
            again = E.disp_batch(oms)  # eager work between the replays
            assert np.all(np.abs(again - ref) / scale <= 1e-13)


def test_df0_input_matches_f0_input():
    """Passing df0 (as the Fortran host holds it) equals deriving it on the device."""
    from alps_amd.engine import Engine
    p = cases.small_bimax()
    a = Engine(p)
    b = Engine(p, use_df0=True)
    a.set_k(0.3, 0.3)
    b.set_k(0.3, 0.3)
    oms = np.array(cases.BRANCH_OMEGAS)
    assert np.array_equal(a.disp_batch(oms), b.disp_batch(oms))


def test_every_launch_shape_gives_the_same_results():
    """Each instantiated kernel variant (lanes per omega, wavefronts, row chunks,
    harmonic blocks) sums the same terms: results agree to rounding."""
    from alps_amd.engine import Engine
    p = cases.three_species()
    E = Engine(p)
    E.set_k(0.5, 0.2)
    oms = np.array(cases.BRANCH_OMEGAS * 3)
    n = oms.size
    ref, chi_r, low_r, _ = E.disp_batch(oms, want_tensors=True)
    scale = np.max(np.abs(chi_r), axis=(2, 3))
    shapes = []
    for i, (nch, w) in enumerate(Engine.SWEEP_VARIANTS):
        nb = Engine.RESONANT_LANES[i % len(Engine.RESONANT_LANES)]
        shapes.append((nch, w, 1 + i % 3, nb, 1 + (i % 4)))
    for sh in shapes:
        E.set_shape(n, sh)
        assert E.get_shape(n) == sh
        d, chi, low, _ = E.disp_batch(oms, want_tensors=True)
        e = np.max(np.abs(chi - chi_r), axis=(2, 3)) / scale
        assert np.all(e <= 1e-12), (sh, e.max())
    E.set_shape(n, None)
    sh = E.autotune(n, oms)
    assert E.get_shape(n) == sh
    d, chi, low, _ = E.disp_batch(oms, want_tensors=True)
    assert np.max(np.abs(chi - chi_r) / scale[:, :, None, None]) <= 1e-12
    E.close()


@pytest.mark.parametrize("kperp_norm", [True, False])
def test_relativistic_pair(kperp_norm):
    """Relativistic species: resonant harmonics on the (Gamma, pparbar) grid
    (integrate_res_rel, landau_integrate_rel), the others on the (pperp, ppar)
    grid with Gamma in resU (ALPS_fns.f90:1584-1585)."""
    from oracle import rel_error, reset_rel_error
    reset_rel_error()
    p = cases.rel_pair(kperp_norm=kperp_norm)
    _check(p, 0.3, 0.1, cases.REL_OMEGAS)
    assert rel_error() == 0


def test_relativistic_beside_nonrelativistic():
    _check(cases.rel_mixed(), 0.3, 0.2, cases.REL_OMEGAS)


def test_relativistic_window_error_is_reported():
    """A principal-value window wider than a Gamma row's subluminal part is
    alps_error(8) in the reference; the engine returns an error for it."""
    from alps_amd.engine import AlpsgError
    p = cases.rel_pair(ngamma=30, npparbar=30)
    E = _engine(p)
    E.set_k(1e-3, 0.1)
    with pytest.raises(AlpsgError) as ei:
        E.disp_batch([0.0627 - 4.662e-8j])
    assert "alps_error(8)" in str(ei.value)
    E.close()
