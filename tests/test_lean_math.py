"""The register-lean complex power of the kappa Landau pass
(alps_amd/csrc/alpsg_lean.h), compiled for the host from the same header the
kernels include, against mpmath at 40 digits:

* z^c to within 4e-16 times its condition number |c| (|log|z|| + pi) + 1
  (the error any double evaluation of exp(c log z) carries), over |z| from
  1e-200 to 1e200, both half planes, the negative real axis, and |c| <= 20;
* exp to 3e-16 relative, the reduced log to 2e-16, atan2 to 1e-15, sin and
  cos to 3e-16 absolute (a few ulp each);
* z = 0 gives 0, as the library complex power of eval_fit does."""
import os
import shutil
import subprocess

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")
HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    exe = str(tmp_path_factory.mktemp("lean") / "lean_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-o", exe, os.path.join(HERE, "lean_math_check.cpp")],
                   check=True)
    return exe


def _run(exe, xs, ys, cs):
    inp = "\n".join(f"{float(x)!r} {float(y)!r} {float(c)!r}" for x, y, c in zip(xs, ys, cs))
    out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout
    rows = [l.split() for l in out.strip().split("\n")]
    assert len(rows) == len(xs)
    return rows


def test_lean_power_and_its_parts_against_mpmath(checker):
    mp.mp.dps = 40
    rng = np.random.default_rng(11)
    n = 1500
    xs = np.concatenate([rng.uniform(-50, 200, n), 10 ** rng.uniform(-200, 200, 300) * rng.choice([-1, 1], 300),
                         rng.uniform(-1, 1, 200), np.full(50, -3.0), np.array([0.0, 1.0, -2.0])])
    ys = np.concatenate([rng.uniform(-100, 100, n), 10 ** rng.uniform(-200, 200, 300) * rng.choice([-1, 1], 300),
                         rng.uniform(-1e-8, 1e-8, 200), np.zeros(50), np.array([0.0, 0.0, 0.0])])
    cs = np.concatenate([rng.uniform(-20, 20, n), rng.uniform(-2, 2, 300), rng.uniform(-10, 10, 200),
                         rng.uniform(-8, 8, 50), np.array([-4.5, 3.0, -5.0])])
    rows = _run(checker, xs, ys, cs)
    worst = dict(cpow=0.0, exp=0.0, logm=0.0, atan2=0.0, sincos=0.0)
    for r, x, y, c in zip(rows, xs, ys, cs):
        re, im, ex, E, lm, at, s, co = (float(r[0]), float(r[1]), float(r[2]), int(r[3]), float(r[4]),
                                        float(r[5]), float(r[6]), float(r[7]))
        z = mp.mpc(x, y)
        if z == 0:
            assert re == 0.0 and im == 0.0
        else:
            ref = z ** mp.mpf(c)
            if 1e-300 < abs(ref) < 1e300:
                cond = abs(c) * (abs(mp.log(abs(z))) + mp.pi) + 1
                worst["cpow"] = max(worst["cpow"], float(abs(mp.mpc(re, im) - ref) / abs(ref) / cond))
            worst["atan2"] = max(worst["atan2"], float(abs(at - mp.atan2(y, x))))
        worst["exp"] = max(worst["exp"], float(abs(ex - mp.exp(c)) / mp.exp(c)))
        xx = x if x > 0 else 1.0
        worst["logm"] = max(worst["logm"], float(abs(lm - (mp.log(xx) - E * mp.log(2)))))
        worst["sincos"] = max(worst["sincos"], float(max(abs(s - mp.sin(c)), abs(co - mp.cos(c)))))
    print(worst)
    assert worst["cpow"] <= 4e-16
    assert worst["exp"] <= 3e-16
    assert worst["logm"] <= 2e-16
    assert worst["atan2"] <= 1e-15
    assert worst["sincos"] <= 3e-16


def test_lean_sincos_over_the_reduction_range(checker):
    """Arguments up to the largest |c arg z| the kappa pass admits (|c| <= 1e4)."""
    mp.mp.dps = 40
    rng = np.random.default_rng(12)
    cs = np.concatenate([rng.uniform(-3.2e4, 3.2e4, 400), np.arange(-20, 21) * 0.5 * np.pi])
    rows = _run(checker, np.ones_like(cs), np.zeros_like(cs), cs)
    err = max(max(abs(float(r[6]) - mp.sin(c)), abs(float(r[7]) - mp.cos(c))) for r, c in zip(rows, cs))
    assert err <= 3e-16, err
