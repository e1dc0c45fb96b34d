"""The CPU-runnable workload on the GPU: 2 species on 50 x 100, kperp d_p =
0.01, kpar d_p = 0.1.  secant_osc refines the Alfven/ion-cyclotron root from
0.1 - 0.001i to |det D| < 1e-12 (ALPS_fns.f90:1919-2101), then om_scan type 4
follows it through 64 logarithmic kpar steps from 0.1 to 1.0
(ALPS_fns.f90:2198-2600).  Every step's root must agree with the oracle to
|dw/w| <= 1e-10."""
import math

import numpy as np
import pytest

import cases
from oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT_TOL = 1e-10
GUESS = 0.1 - 0.001j
NSTEPS = 64
DIFF = (math.log10(1.0) - math.log10(0.1)) / NSTEPS


def test_config0_refine_and_log_kpar_scan_match_oracle():
    from alps_amd.engine import Engine
    p = cases.config0()
    O, E = Oracle(p), Engine(p)
    assert E.set_k(0.01, 0.1) == O.set_k(0.01, 0.1) == [20, 12]
    g_om, g_it = E.secant_osc(GUESS)
    o_om, o_it = O.secant_osc(GUESS)
    print("refined", g_om, o_om, g_it, o_it, flush=True)
    assert abs(g_om - o_om) <= ROOT_TOL * abs(o_om)
    assert g_it < p.roots.numiter
    assert abs(E.disp(g_om)) < p.roots.D_threshold
    gk, gz, groots = E.om_scan(4, True, DIFF, 0.0, NSTEPS, 1, [g_om])
    rc, ok, oz, oroots = O.om_scan(4, True, DIFF, 0.0, NSTEPS, 1, [o_om])
    assert rc == 0
    assert np.allclose(gz, oz, rtol=1e-14, atol=0) and np.allclose(gk, ok, rtol=1e-14, atol=0)
    assert math.isclose(gz[-1], 1.0, rel_tol=1e-12) and math.isclose(gz[0], 0.1, rel_tol=1e-12)
    rel = np.abs(groots[:, 0] - oroots[:, 0]) / np.abs(oroots[:, 0])
    print("max |dw/w| over the scan:", float(rel.max()), flush=True)
    assert rel.max() <= ROOT_TOL, (int(np.argmax(rel)), float(rel.max()))
    # the followed root stays on the branch: Re w grows monotonically with kpar
    assert np.all(np.diff(groots[:, 0].real) > 0)
