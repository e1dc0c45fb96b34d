"""The restatement of om_double_scan (ALPS_fns.f90:2900-3591), pinned by what
fixes it without a reference run (the reference ships test_double_scan.in but
no output for it):

* the wavevector of every output point, from the closed forms of the scan
  types and the reference's reset of the second scan's coordinate between
  outer steps;
* every inner scan equals om_scan (already pinned by the test_kpar_fast
  golden) run from that outer point with the point's first root;
* the outer path equals om_scan of the first scan from the starting roots;
* the reference's error cases (alps_error 5 and 6)."""
import dataclasses
import math

import numpy as np

import cases
from oracle import Oracle


def _tight(p):
    return dataclasses.replace(p, roots=dataclasses.replace(p.roots, secant_method=2,
                                                            D_threshold=1e-14))


K0 = (0.05, 0.05)
# scan_read (ALPS_io.f90:482-548): log steps from the starting k
S1 = (4, 1, (math.log10(0.2) - math.log10(K0[1])) / 3, 0.0, 3, 1)   # kpar 0.05 -> 0.2
S2 = (3, 1, (math.log10(0.2) - math.log10(K0[0])) / 4, 0.0, 2, 2)   # kperp 0.05 -> 0.2


def _start():
    p = _tight(cases.small_bimax())
    O = Oracle(p)
    O.set_k(*K0)
    om, _ = O.secant_osc(0.048 - 0.0001j)
    return p, O, om


def test_double_scan_k_grid_is_the_closed_form():
    p, O, om = _start()
    r = O.om_double_scan(S1, S2, [om])
    assert r["rc"] == 0 and r["written"].all()
    for i1 in range(S1[4] + 1):
        for i2 in range(S2[4] + 1):
            kpar = 10.0 ** (math.log10(K0[1]) + S1[2] * i1 * S1[5])
            kperp = 10.0 ** (math.log10(K0[0]) + S2[2] * i2 * S2[5])
            assert math.isclose(r["kpar"][i1, i2], kpar, rel_tol=1e-14)
            assert math.isclose(r["kperp"][i1, i2], kperp, rel_tol=1e-14)


def test_inner_scans_are_om_scans_and_the_outer_path_is_one():
    p, O, om = _start()
    r = O.om_double_scan(S1, S2, [om])
    # the outer path: om_scan of scan 1 from the start
    O2 = Oracle(p)
    O2.set_k(*K0)
    rc, kp, kz, outer = O2.om_scan(S1[0], S1[1], S1[2], S1[3], S1[4], S1[5], [om])
    assert rc == 0
    for i1 in range(S1[4] + 1):
        # row (i1, 0) re-solves the outer root once more at the same k
        assert abs(r["roots"][i1, 0, 0] - outer[i1, 0]) <= 1e-10 * abs(outer[i1, 0])
        # the inner scan: om_scan of scan 2 from (i1, 0) with that root
        O3 = Oracle(p)
        O3.set_k(r["kperp"][i1, 0], r["kpar"][i1, 0])
        rc, kp2, kz2, inner = O3.om_scan(S2[0], S2[1], S2[2], S2[3], S2[4], S2[5],
                                         [r["roots"][i1, 0, 0]])
        assert rc == 0
        for i2 in range(S2[4] + 1):
            assert kp2[i2] == r["kperp"][i1, i2] and kz2[i2] == r["kpar"][i1, i2]
            assert inner[i2, 0] == r["roots"][i1, i2, 0], (i1, i2)


def test_double_scan_error_cases():
    p, O, om = _start()
    assert O.om_double_scan(S1, S1, [om])["rc"] == 5
    assert O.om_double_scan((0,) + S1[1:], S2, [om])["rc"] == 6
