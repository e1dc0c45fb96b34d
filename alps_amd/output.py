"""Writers of ALPS's solution files, byte-compatible with the reference's
formatted writes (ALPS_fns.f90):

* ``.scan_*`` rows ``(4es14.4e3)``: kperp, kpar, Re w, Im w (om_scan :2352,
  :2562; om_double_scan :3460);
* ``.eigen_*`` rows ``(4es14.4e3,12es14.4e3,<8 nspec>es14.4e3)``: kperp, kpar,
  w, ef(1:3), bf(1:3), Us(1:3,1:nspec), ds(1:nspec) (:2334, :2379);
* ``.heat_*`` rows ``(4es14.4e3,<nspec>es14.4e3,es14.4e3)``: kperp, kpar, w,
  Ps(1:nspec), W_EM (:2339, :2390);
* ``.heat_mech_*`` rows ``(4es14.4e3,<4 nspec>es14.4e3)``: kperp, kpar, w,
  Ps_split(1:4,1:nspec) (:2340, :2399);
* ``.map`` rows ``(5es16.6e3)``: w, log10|D|, D, with an empty record after
  each real-frequency row (:3745, :3754);
* ``.roots`` rows ``(i4,5es14.4e3)``: index, w, log10|D|, D (:3849).

File names follow the reference: ``solution/<runname>.scan_<id><ik>.root_<n>``
for om_scan and ``solution/<runname>.scan_<id1><id2>.root_<n>`` for
om_double_scan, where the ids are k1_k2, theta, kcstq, kperp, kpara (with a
trailing underscore for the first scan of a double scan, :3063-3090).
"""
from __future__ import annotations

import math
import os

import numpy as np

SCAN_IDS = ["k1_k2", "theta", "kcstq", "kperp", "kpara"]


def es(v: float, w: int = 14, d: int = 4, e: int = 3) -> str:
    """Fortran ``ESw.dEe`` edit descriptor: d.dddd mantissa, e-digit exponent,
    right-justified in w columns."""
    v = float(v)
    if math.isnan(v):
        s = "NaN"
    elif math.isinf(v):
        s = "Infinity" if v > 0 else "-Infinity"
    else:
        m, x = f"{v:.{d}e}".split("e")
        xi = int(x)
        s = f"{m}E{'-' if xi < 0 else '+'}{abs(xi):0{e}d}"
    if len(s) > w:
        return "*" * w
    return s.rjust(w)


def row(values, w=14, d=4, e=3) -> str:
    return "".join(es(v, w, d, e) for v in values) + "\n"


def _c(z):
    z = complex(z)
    return [z.real, z.imag]


def scan_row(kperp, kpar, root) -> str:
    return row([kperp, kpar] + _c(root))


def eigen_row(kperp, kpar, root, ef, bf, Us, ds) -> str:
    """Us as [nspec][3] (Us(1:3,1:nspec) in Fortran order)."""
    vals = [kperp, kpar] + _c(root)
    for z in list(ef) + list(bf) + list(np.asarray(Us).ravel()) + list(ds):
        vals += _c(z)
    return row(vals)


def heat_row(kperp, kpar, root, Ps, W_EM) -> str:
    return row([kperp, kpar] + _c(root) + list(Ps) + [W_EM])


def heat_mech_row(kperp, kpar, root, Ps_split) -> str:
    """Ps_split as [nspec][4] (Ps_split(1:4,1:nspec) in Fortran order)."""
    return row([kperp, kpar] + _c(root) + list(np.asarray(Ps_split).ravel()))


def map_text(om, val, cal) -> str:
    """om, cal complex and val real, each [nr][ni] (the reference's (nr, ni))."""
    out = []
    nr, ni = np.shape(val)
    for ir in range(nr):
        for ii in range(ni):
            out.append(row(_c(om[ir][ii]) + [val[ir][ii]] + _c(cal[ir][ii]), 16, 6, 3))
        out.append("\n")
    return "".join(out)


def roots_text(roots, dfin) -> str:
    """The .roots file of refine_guess: rows only for roots with D != 0."""
    out = []
    for i, (w, D) in enumerate(zip(roots, dfin), start=1):
        if abs(D) != 0.0:
            out.append(f"{i:4d}" + row(_c(w) + [math.log10(abs(D))] + _c(D)))
    return "".join(out)


def scan_files(runname, kind, ids, n, outdir="solution"):
    """Path of the ``kind`` file (scan, eigen, heat, heat_mech) of root n."""
    return os.path.join(outdir, f"{runname}.{kind}_{ids}.root_{n}")


def write_om_scan(runname, scan_type, ik, kperp, kpar, roots, eigen=None, heat=None,
                  outdir="solution"):
    """The files om_scan(ik) writes for each root: kperp, kpar [rows], roots
    [rows][nroots]; eigen/heat: dicts of calc_eigen outputs [rows][nroots][...]."""
    ids = f"{SCAN_IDS[scan_type]}_{ik}"
    os.makedirs(outdir, exist_ok=True)
    nroots = np.shape(roots)[1]
    for n in range(nroots):
        with open(scan_files(runname, "scan", ids, n + 1, outdir), "w") as f:
            for i in range(len(kperp)):
                f.write(scan_row(kperp[i], kpar[i], roots[i][n]))
        _write_eigen_heat(runname, ids, n, kperp, kpar, [r[n] for r in roots], eigen, heat, outdir,
                          lambda i: i)


def write_om_double_scan(runname, type1, type2, res, eigen=False, heat=False, outdir="solution"):
    """The files om_double_scan writes for each root from the dict of
    ``Engine.om_double_scan``: every inner scan as a block of rows (the points
    where the root was still followed), an empty record after each block."""
    ids = f"{SCAN_IDS[type1]}_{SCAN_IDS[type2]}"
    os.makedirs(outdir, exist_ok=True)
    n1, n2, nroots = res["roots"].shape
    for n in range(nroots):
        kinds = ["scan"] + (["eigen"] if eigen else []) + (["heat", "heat_mech"] if heat else [])
        files = {k: open(scan_files(runname, k, ids, n + 1, outdir), "w") for k in kinds}
        for i1 in range(n1):
            for i2 in range(n2):
                if not res["written"][i1, i2, n]:
                    continue
                kp, kz, w = res["kperp"][i1, i2], res["kpar"][i1, i2], res["roots"][i1, i2, n]
                files["scan"].write(scan_row(kp, kz, w))
                if eigen:
                    files["eigen"].write(eigen_row(kp, kz, w, res["ef"][i1, i2, n],
                                                   res["bf"][i1, i2, n], res["Us"][i1, i2, n],
                                                   res["ds"][i1, i2, n]))
                if heat:
                    files["heat"].write(heat_row(kp, kz, w, res["Ps"][i1, i2, n],
                                                 res["W_EM"][i1, i2, n]))
                    files["heat_mech"].write(heat_mech_row(kp, kz, w, res["Ps_split"][i1, i2, n]))
            for f in files.values():
                f.write("\n")
        for f in files.values():
            f.close()


def _write_eigen_heat(runname, ids, n, kperp, kpar, roots, eigen, heat, outdir, at):
    if eigen is not None:
        with open(scan_files(runname, "eigen", ids, n + 1, outdir), "w") as f:
            for i in range(len(kperp)):
                e = eigen
                f.write(eigen_row(kperp[i], kpar[i], roots[i], e["ef"][at(i)][n], e["bf"][at(i)][n],
                                  e["Us"][at(i)][n], e["ds"][at(i)][n]))
    if heat is not None:
        with open(scan_files(runname, "heat", ids, n + 1, outdir), "w") as f:
            for i in range(len(kperp)):
                f.write(heat_row(kperp[i], kpar[i], roots[i], heat["Ps"][at(i)][n],
                                 heat["W_EM"][at(i)][n]))
        with open(scan_files(runname, "heat_mech", ids, n + 1, outdir), "w") as f:
            for i in range(len(kperp)):
                f.write(heat_mech_row(kperp[i], kpar[i], roots[i], heat["Ps_split"][at(i)][n]))
