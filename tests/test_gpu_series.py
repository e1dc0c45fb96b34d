"""GPU parity of the far-row series of the one-lane node sweep: a row whose
every ppar node is far from resonance for every omega of a wavefront takes its
trapezoid node sum (integrate(), ALPS_fns.f90:799-864) from the per-k
polynomial in m w instead of a pass over the nodes.

* det D with the series equals det D with every row swept node by node, to
  1e-12 of the determinant's term scale, on batches that mix far rows, split
  rows and resonant rows (the map region of the 128 x 256 two-species case, the
  three-species and kappa plasmas, both signs of Im w);
* det D equals the oracle on a strided sample, to 1e-11;
* the device work counts add up: with the series off every (species, harmonic,
  node) unit is summed node by node, n * sum_s (2 nmax_s + 1)(npar - 1), and
  with it on the far-row signs account for the rest, (npar - 1) units each."""
import numpy as np
import pytest

import cases
from oracle import Oracle

pytestmark = pytest.mark.gpu

ONE_LANE = (1, 2, 4, 16, 1)   # the one-lane-per-omega sweep (the map batches' shape)


def _map_omegas(n, re=(0.01, 2.0), im=(-0.5, 0.05), seed=5):
    rng = np.random.default_rng(seed)
    return rng.uniform(*re, n) + 1j * rng.uniform(*im, n)


def _series_vs_nodes(p, kp, kz, om, tol=1e-12, shape=ONE_LANE):
    from alps_amd.engine import Engine
    E = Engine(p)
    nmax = E.set_k(kp, kz)
    if shape is not None:
        E.set_shape(om.size, shape)
    on, _, _, wave = E.disp_batch(om, want_tensors=True)
    w_on = E.sweep_work(om)
    E.set_sweep_series(False)
    off = E.disp_batch(om)
    w_off = E.sweep_work(om)
    E.set_sweep_series(True)
    again = E.disp_batch(om)
    E.close()
    sc = np.array([cases.det_scale(wave[i]) for i in range(om.size)])
    err = np.max(np.abs(on - off) / sc)
    print("series vs nodes: max err", err, "work on", w_on, "off", w_off, flush=True)
    assert err <= tol
    assert np.array_equal(on, again)   # the switch restores the same arithmetic
    if shape is not None:
        units = om.size * sum((2 * n + 1) * (p.npar - 1) for n in nmax)
        assert w_off[0] == units and w_off[1] == 0 and w_off[2] == 0
        assert w_on[0] + w_on[1] * (p.npar - 1) == units
    return on, wave, w_on


def test_series_matches_node_sums_on_the_map_region():
    p = cases.config1()
    om = _map_omegas(4096)
    det, wave, w = _series_vs_nodes(p, 0.5, 0.5, om)
    # most of the work of this case lies in far rows
    assert w[1] > 0 and w[2] >= w[1]
    O = Oracle(p)
    O.set_k(0.5, 0.5)
    for i in range(0, om.size, 512):
        d = O.disp(om[i])
        assert abs(det[i] - d) <= 1e-11 * cases.det_scale(wave[i]), (i, om[i])


@pytest.mark.parametrize("case,k", [("three_species", (0.3, 0.3)), ("kappa_case", (0.3, 0.3)),
                                    ("small_bimax", (0.05, 0.05))])
def test_series_matches_node_sums_small_plasmas(case, k):
    p = getattr(cases, case)()
    om = np.concatenate([_map_omegas(1024, (0.01, 3.5), (-0.4, 0.1), seed=7),
                         _map_omegas(1024, (0.01, 0.3), (-0.05, 0.05), seed=8)])
    _series_vs_nodes(p, *k, om)


def test_series_with_positive_growth_and_far_omegas():
    """Im w > 0 throughout and omegas far above every resonance (every row far)."""
    p = cases.config1(64, 128)
    om = np.concatenate([_map_omegas(512, (0.01, 2.0), (0.01, 0.3), seed=3),
                         _map_omegas(512, (40.0, 60.0), (-1.0, 1.0), seed=4)])
    det, wave, w = _series_vs_nodes(p, 0.5, 0.5, om)
    O = Oracle(p)
    O.set_k(0.5, 0.5)
    for i in (0, 100, 511, 512, 700, 1023):
        assert abs(det[i] - O.disp(om[i])) <= 1e-11 * cases.det_scale(wave[i]), (i, om[i])


def test_series_in_small_batches_and_secant_rounds():
    """Batches of at most 16 omegas (the request sweep, captured in graphs) and
    the device-stepped secant rounds sum every row node by node, also after a
    large batch has built the k's series tables: the same det D and roots with
    the series switched on and off."""
    from alps_amd.engine import Engine
    p = cases.config1(64, 128)
    om = _map_omegas(16, seed=12)
    E = Engine(p)
    E.set_k(0.5, 0.5)
    # guesses: the minima of a 64 x 64 map of the bench's frequency window
    gom, _, _, iroots, nf = E.map_search(0.01, 2.0, -0.5, 0.05, 64, 64)
    guesses = np.array([gom[i - 1, j - 1] for i, j in iroots[:8]])
    assert guesses.size >= 2
    on = [E.disp_batch(om[:n]) for n in (1, 5, 16)]
    r_on = E.secant_osc_batch(guesses)
    E.set_sweep_series(False)
    off = [E.disp_batch(om[:n]) for n in (1, 5, 16)]
    r_off = E.secant_osc_batch(guesses)
    _, _, _, wave = E.disp_batch(om, want_tensors=True)
    E.close()
    for a, b in zip(on, off):
        assert np.array_equal(a, b)
    assert np.array_equal(r_on[0], r_off[0]) and np.array_equal(r_on[2], r_off[2])


@pytest.mark.parametrize("case,k", [("config1", (0.5, 0.5)), ("three_species", (0.3, 0.3)),
                                    ("kappa_case", (0.3, 0.3))])
def test_partial_range_rows_in_the_split_form(case, k):
    """Rows resonant for some omega of a wavefront: the one-lane sweep sums
    their lanes' interior ranges in the split form (C/den node by node, G/m
    from cumulative sums); two lanes per omega take the U/den form.  Both
    agree to 1e-12 of the determinant's term scale, series off in both."""
    from alps_amd.engine import Engine
    p = getattr(cases, case)() if case != "config1" else cases.config1(64, 128)
    rng = np.random.default_rng(21)
    om = rng.uniform(0.05, 2.5, 2048) + 1j * rng.choice([-0.2, -0.02, -0.001, 0.0, 0.001, 0.05], 2048)
    E = Engine(p)
    E.set_k(*k)
    E.set_sweep_series(False)
    E.set_shape(om.size, ONE_LANE)
    one, _, _, wave = E.disp_batch(om, want_tensors=True)
    E.set_shape(om.size, (2, 4, 1, 16, 1))
    two = E.disp_batch(om)
    E.close()
    err = max(abs(one[i] - two[i]) / cases.det_scale(wave[i]) for i in range(om.size))
    print("split vs U/den form:", err, flush=True)
    assert err <= 1e-12
