"""GPU parity of the continuations of ACmethod 0 (the pre-coded
distribution_analyt, ALPS_analyt.f90:77-79) and of fit type 3 (Juettner in
pperp and ppar, ALPS_analyt.f90:238-242): disp() over the branch omegas at
1e-11, the reference's own tests/test_analytical.in set-up (root refined at
kperp = 1e-2, kpar = 1e-3 from 9.9e-4 - 5.5e-10i on 120 x 240, then the
32-step log kpar scan to 1e-2) at 1e-10 per root, and served solves."""
import dataclasses
import math

import numpy as np
import pytest

import cases
from oracle import Oracle
from test_gpu_disp import _check

pytestmark = pytest.mark.gpu


def test_acmethod0_branches():
    _check(cases.analytical(), 0.3, 0.3, cases.BRANCH_OMEGAS)


def test_juettner_fit_type3_branches():
    _check(cases.juettner3(), 0.3, 0.3, cases.BRANCH_OMEGAS)


def test_analytical_input_root_and_scan():
    from alps_amd.engine import Engine
    p = cases.analytical(120, 240)
    O, E = Oracle(p), Engine(p)
    assert E.set_k(1e-2, 1e-3) == O.set_k(1e-2, 1e-3)
    g, g_it = E.secant_osc(9.9e-4 - 5.5e-10j)
    o, o_it = O.secant_osc(9.9e-4 - 5.5e-10j)
    print("root", g, o, g_it, o_it, flush=True)
    assert abs(g - o) <= 1e-10 * abs(o) and g_it == o_it < p.roots.numiter
    diff = (math.log10(1e-2) - math.log10(1e-3)) / 32
    _, _, groots = E.om_scan(4, True, diff, 0.0, 32, 1, [g])
    rc, _, _, oroots = O.om_scan(4, True, diff, 0.0, 32, 1, [o])
    assert rc == 0
    rel = np.abs(groots[:, 0] - oroots[:, 0]) / np.abs(oroots[:, 0])
    print("max |dw/w| over the scan:", float(rel.max()), flush=True)
    assert rel.max() <= 1e-10


def test_served_continuations_match_oracle():
    from alps_amd.engine import RootServer
    tight = lambda p: dataclasses.replace(p, roots=dataclasses.replace(
        p.roots, secant_method=2, numiter=50, D_threshold=1e-14))
    reqs = [(tight(cases.analytical()), 0.3, 0.3, 0.29 - 0.001j),
            (tight(cases.juettner3()), 0.3, 0.3, 0.5 - 0.05j),
            (tight(cases.juettner3()), 0.2, 0.4, 0.4 - 0.01j)]
    S = RootServer(hbm_bytes=2e9, max_slots=16)
    for i, (p, kp, kz, g) in enumerate(reqs):
        S.submit(p, kp, kz, g, p.roots, tag=i)
    got = S.drain(len(reqs))
    S.close()
    for i, (p, kp, kz, g) in enumerate(reqs):
        O = Oracle(p)
        O.set_k(kp, kz)
        w, it = O.secant_osc(g)
        assert got[i][2] == it and (got[i][3] == 6) == (it >= p.roots.numiter), (got[i], it)
        assert abs(got[i][1] - w) <= 1e-10 * abs(w), (i, got[i][1], w)
