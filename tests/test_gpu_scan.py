"""om_double_scan on the device against the restatement, and the solution
files of the device path against the reference's own output files."""
import math
import os

import numpy as np
import pytest

import cases
from oracle import Oracle
from test_oracle_double_scan import K0, S1, S2, _tight
from test_oracle_golden import GOLD, GOLD_EIGEN, GOLD_HEAT, kpar_fast_scan

pytestmark = pytest.mark.gpu

ROOT_TOL = 1e-10


def _engine(p):
    from alps_amd.engine import Engine
    return Engine(p)


@pytest.mark.parametrize("want", [False, True])
def test_double_scan_matches_oracle(want):
    p = _tight(cases.small_bimax(drift=0.3))
    O, E = Oracle(p), _engine(p)
    O.set_k(*K0)
    E.set_k(*K0)
    om, _ = O.secant_osc(0.048 - 0.0001j)
    guesses = [om, 0.3 - 0.05j]
    o = O.om_double_scan(S1, S2, guesses, eigen=want, heat=want)
    g = E.om_double_scan(S1, S2, guesses, eigen=want, heat=want)
    assert o["rc"] == 0
    assert np.array_equal(o["written"], g["written"])
    assert np.allclose(g["kperp"], o["kperp"], rtol=1e-14, atol=0)
    assert np.allclose(g["kpar"], o["kpar"], rtol=1e-14, atol=0)
    w = o["written"]
    rel = np.abs(g["roots"][w] - o["roots"][w]) / np.abs(o["roots"][w])
    assert np.max(rel) <= ROOT_TOL, rel
    if want:
        for key, tol in (("ef", 1e-9), ("bf", 1e-9), ("Us", 1e-9), ("ds", 1e-9), ("Ps", 1e-5),
                         ("Ps_split", 1e-5), ("W_EM", 1e-5)):
            a, b = g[key][w], o[key][w]
            scale = np.max(np.abs(b.reshape(len(b), -1)), axis=1)
            err = np.max(np.abs((a - b).reshape(len(b), -1)), axis=1)
            assert np.all(err <= tol * np.maximum(scale, 1e-300)), (key, err / scale)


def test_double_scan_files(tmp_path):
    from alps_amd import output
    p = _tight(cases.small_bimax())
    E = _engine(p)
    E.set_k(*K0)
    om, _ = E.secant_osc(0.048 - 0.0001j)
    r = E.om_double_scan(S1, S2, [om], eigen=True, heat=True)
    output.write_om_double_scan("run", S1[0], S2[0], r, eigen=True, heat=True, outdir=tmp_path)
    text = open(tmp_path / "run.scan_kpara_kperp.root_1").read()
    blocks = text.split("\n\n")
    assert len(blocks) == S1[4] + 2 and blocks[-1] == ""  # a block per outer point
    for i1, blk in enumerate(blocks[:-1]):
        rows = [l for l in blk.split("\n") if l]
        assert len(rows) == S2[4] + 1
        for i2, l in enumerate(rows):
            assert l == output.scan_row(r["kperp"][i1, i2], r["kpar"][i1, i2],
                                        r["roots"][i1, i2, 0]).rstrip("\n")
    for kind, ncol in (("eigen", 4 + 12 + 8 * p.nspec), ("heat", 4 + p.nspec + 1),
                       ("heat_mech", 4 + 4 * p.nspec)):
        rows = [l for l in open(tmp_path / f"run.{kind}_kpara_kperp.root_1").read().split("\n") if l]
        assert len(rows) == (S1[4] + 1) * (S2[4] + 1) and all(len(l) == 14 * ncol for l in rows)


def test_device_path_writes_the_kpar_fast_files(tmp_path):
    """om_scan and calc_eigen on the device, written with alps_amd.output: the
    .scan and .eigen files equal the reference's byte for byte; the .heat file
    to its printed digits (see test_kpar_fast_scan_on_gpu_matches_golden_and_oracle
    for the one-unit allowance of the heating columns)."""
    from alps_amd import output
    p = cases.kpar_fast()
    E = _engine(p)
    E.set_k(1e-2, 1e-2)
    om, (kp, kz, roots) = kpar_fast_scan(E)
    eig = {"ef": [], "bf": [], "Us": [], "ds": [], "Ps": [], "Ps_split": [], "W_EM": []}
    for i in range(len(kp)):
        E.set_k(kp[i], kz[i])
        r = E.calc_eigen([roots[i, 0]])
        for k in eig:
            eig[k].append([r[k][0]])
    output.write_om_scan("test_kpar_fast", 4, 1, kp, kz, roots, eigen=eig, heat=eig,
                         outdir=tmp_path)
    assert open(tmp_path / "test_kpar_fast.scan_kpara_1.root_1").read() == open(GOLD).read()
    assert open(tmp_path / "test_kpar_fast.eigen_kpara_1.root_1").read() == open(GOLD_EIGEN).read()
    got = np.loadtxt(tmp_path / "test_kpar_fast.heat_kpara_1.root_1")
    gold = np.loadtxt(GOLD_HEAT)
    unit = 10.0 ** (np.floor(np.log10(np.abs(gold) + 1e-300)) - 4)
    assert np.all(np.abs(got - gold) <= unit * (1 + 1e-9))


@pytest.mark.parametrize("scan_type,log_scan", [(0, True), (0, False), (1, True), (1, False),
                                                (2, True), (2, False)])
def test_om_scan_types_0_to_2_match_oracle(scan_type, log_scan):
    """om_scan (ALPS_fns.f90:2198-2600) along k1 -> k2 (type 0), in theta
    (type 1) and in |k| at constant theta (type 2), log and linear steps, n_res
    sub-steps: every output root against the oracle to 1e-10."""
    p = _tight(cases.small_bimax())
    O, E = Oracle(p), _engine(p)
    O.set_k(*K0)
    E.set_k(*K0)
    om, _ = O.secant_osc(0.048 - 0.0001j)
    theta = math.atan(K0[0] / K0[1])
    if scan_type == 0:   # kperp 0.05 -> 0.1 and kpar 0.05 -> 0.15
        d1 = (math.log10(0.1) - math.log10(K0[0])) / 6 if log_scan else (0.1 - K0[0]) / 6
        d2 = (math.log10(0.15) - math.log10(K0[1])) / 6 if log_scan else (0.15 - K0[1]) / 6
    elif scan_type == 1:  # theta to 1.2 theta
        d1 = (math.log10(1.2 * theta) - math.log10(theta)) / 6 if log_scan else 0.2 * theta / 6
        d2 = 0.0
    else:                 # |k| to 2 |k|
        k0 = math.hypot(*K0)
        d1 = (math.log10(2 * k0) - math.log10(k0)) / 6 if log_scan else k0 / 6
        d2 = 0.0
    rc, ok, oz, oroots = O.om_scan(scan_type, log_scan, d1, d2, 3, 2, [om])
    gk, gz, groots = E.om_scan(scan_type, log_scan, d1, d2, 3, 2, [om])
    assert rc == 0
    assert np.allclose(gk, ok, rtol=1e-14, atol=0) and np.allclose(gz, oz, rtol=1e-14, atol=0)
    rel = np.abs(groots[:, 0] - oroots[:, 0]) / np.abs(oroots[:, 0])
    assert rel.max() <= ROOT_TOL, rel
