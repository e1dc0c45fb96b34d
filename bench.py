"""Benchmark of the batched map search + root refinement on one MI355X.

Workload (the configuration the headline metric is quoted on): 2-species
bi-Maxwellian plasma (protons, electrons with m_p/m_e = 1836, T_e = T_p,
beta_par = 1, v_A/c = 1e-4), f0 tabulated in FP64 on nperp x npar = 128 x 256,
kperp d_p = kpar d_p = 0.5.  One step = det D on the 256 x 256 complex-omega
map (Re w in [0.01, 2], Im w in [-0.5, 0.05]) in one fused launch, local
minima (at most 100), batched secant_osc refinement of every minimum
(|det D| < 1e-12) and det D at every refined root.

Prints one JSON line.  value = chi-tensor (det D) evaluations per second over
the timed steps; roots/s and p50/p99 root latency (from the start of the step
to the root's convergence) are reported beside it.

Beside the map line, "serve" reports served root solves (--workload serve runs
them alone): 4096 independent secant_osc requests arriving as a Poisson stream
through the continuous-batching root server.  Each request draws kperp, kpar
log-uniform in [0.01, 3], Re w log-uniform in [0.01, 3] and Im w uniform in
[-0.1, 0.01], and one of 64 plasma variants: protons with T_perp/T_par uniform
in [0.3, 3], electrons (n = 1.1), alphas (5 %, drifting at 1 v_A), on grids
32x64, 64x128, 128x256 or 256x512.  It reports roots/s, p50/p99 latency
(submit to result), the same requests on the one-root-per-launch path and on
the CPU restatement.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("roots/sec and p50/p99 root latency at |Δω/ω|≤1e-10; "
          "chi-tensor evals/sec")
FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 (vector and matrix) dense peak
HBM_PEAK_GBS = 8000.0

M_E = 5.44662e-4


def make_plasma(nperp=128, npar=256):
    from alps_amd.plasma import RootParams, bimaxwellian_plasma
    rp = RootParams(secant_method=2, numiter=50, D_threshold=1e-12, D_prec=1e-5, D_gap=1e-5)
    return bimaxwellian_plasma(ms=[1.0, M_E], qs=[1.0, -1.0], ns=[1.0, 1.0], taus=[1.0, 1.0],
                               alphs=[1.0, 1.0], ps=[0.0, 0.0], beta=1.0, vA=1e-4,
                               nperp=nperp, npar=npar, maxP=8.0, positions_principal=3,
                               n_resonance_interval=100, Tlim=0.01, Bessel_zero=1e-45, roots=rp)


MAP = dict(omi=0.01, omf=2.0, gami=-0.5, gamf=0.05)
KPERP = KPAR = 0.5


# HBM traffic of the map-batch k_sweep launch by its shape.
# FETCH_SIZE and WRITE_SIZE (KiB per dispatch) of k_sweep on the 256 x 256 map
# batch for every one-lane launch shape (lanes, wavefronts, workgroups per omega
# block), one counter pass each over `bench.py --pmc-shapes` (far rows summed
# by their series, after one warm-up batch); a 262144 KiB copy reads
# FETCH_SIZE 131084-131089 and WRITE_SIZE 262144, so FETCH_SIZE counts half
# the bytes fetched and WRITE_SIZE all bytes written.  The writes are the
# partial-chi slots, 12288 KiB per slot (W x workgroups slots per omega).
_SWEEP_PMC_KIB = {
    (1, 1, 1): (3352.8, 12288.0), (1, 1, 2): (3352.9, 24576.0), (1, 1, 3): (3501.9, 36864.0),
    (1, 1, 4): (3505.1, 49152.0), (1, 1, 6): (3471.1, 73728.0), (1, 1, 8): (3447.4, 98304.0),
    (1, 1, 12): (3420.6, 147456.0), (1, 1, 16): (3457.1, 196608.0), (1, 2, 1): (3353.2, 24576.0),
    (1, 2, 2): (3351.1, 49152.0), (1, 2, 3): (3410.5, 73728.0), (1, 2, 4): (3528.6, 98304.0),
    (1, 2, 6): (3788.7, 147456.0), (1, 2, 8): (3852.8, 196608.0), (1, 2, 12): (4067.6, 294912.0),
    (1, 2, 16): (4487.9, 393216.0), (1, 4, 1): (4415.7, 49152.0), (1, 4, 2): (4501.9, 98304.0),
    (1, 4, 3): (4982.6, 147456.0), (1, 4, 4): (5384.8, 196608.0), (1, 4, 6): (6266.7, 294912.0),
    (1, 4, 8): (7017.7, 393216.0), (1, 4, 12): (8868.6, 589824.0), (1, 4, 16): (10304.1, 786432.0),
}
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0
SWEEP_TRAFFIC_BYTES = {k: (f * FETCH_SCALE + w * WRITE_SCALE) * 1024.0
                       for k, (f, w) in _SWEEP_PMC_KIB.items()}


def flops_per_eval(nmax, npar):
    """Algorithmic FP64 flops of one non-relativistic det D evaluation:
    54 per (species, harmonic, ppar node) of the resonant-denominator sweep."""
    return 54.0 * sum((2 * n + 1) * (npar - 1) for n in nmax)


def sweep_flops(work):
    """Algorithmic FP64 flops of one k_sweep launch from its device work
    counts (alpsg_sweep_work): 54 per node-summed (species, harmonic, ppar
    node) unit, and 24 J + 12 per far-row sign summed by its degree-J series
    (two FMAs per coefficient and node product, 6 products)."""
    return 54.0 * work[0] + 24.0 * work[2] + 12.0 * work[1]


def resonant_rows(plasma, kpar, nmax, om):
    """Rows (species, harmonic) that k_resonant takes at each omega, counted on
    the host the way the device classifies them (determine_resonances,
    ALPS_fns.f90:870-950): (principal-value rows, Landau rows) summed over
    the omegas."""
    om = np.asarray(om).ravel()
    P = plasma.positions_principal
    pv = landau = 0
    for s, nm in enumerate(nmax):
        pa = plasma.pp[s, 1, :, 1]
        npar = pa.size - 1
        d = pa[2] - pa[1]
        m, q = plasma.ms[s], plasma.qs[s]
        for n in range(-nm, nm + 1):
            pr = (m * om.real - n * q) / kpar
            found = ((pa[1] <= pr) & (pr < pa[npar])) | ((pr < pa[1]) & (pr >= pa[1] - P * d)) | \
                    ((pr >= pa[npar - 1]) & (pr < pa[npar - 1] + P * d))
            inner = (pa[1] <= pr) & (pr < pa[npar - 1])
            ipr = np.where(inner, np.clip(np.searchsorted(pa, pr, side="right") - 1, 1, npar - 2), 0)
            ipr = np.where(inner, ipr, np.where(pr < pa[1], 0, npar - 1))
            is_pv = found & inner & (ipr - P > 2) & (ipr + P < npar - 2)
            pv += int(np.sum(is_pv))
            landau += int(np.sum(found & (om.imag <= 0.0)))
    return pv, landau


def resonant_flops(pv_rows, landau_rows, nri, P):
    """Algorithmic FP64 flops of k_resonant: per principal-value row the
    ladder's n_resonance_interval + 1 mirror pairs at 24 flops (the point, its
    complex reciprocal counted with one division, and the two cells' c0 and
    f-weighted sums) and the window contraction of 2P + 8 nodes at 72 flops
    (six node products of U = w G + H and six complex multiply-adds); per
    Landau row 50 flops (the per-k sums' O(1) form; a lower bound for other
    continuations)."""
    return pv_rows * (24.0 * (nri + 1) + 72.0 * (2 * P + 8)) + 50.0 * landau_rows


def progress(msg):
    """One line per finished leg of the run, on stderr (stdout carries the JSON line)."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    """The host CPU's model name (from /proc/cpuinfo)."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return (line.split(":", 1)[1].strip() +
                        f" ({os.cpu_count()} logical CPUs visible)")
    except OSError:
        pass
    return "unknown"


def run_step(E, nr, ni, numroots=100, out=None):
    """One step of the map workload: map_search with refine_guess on every
    minimum (ALPS_fns.f90:3595-3856) in one call; root latencies count from
    the step's start."""
    t0 = time.perf_counter()
    om, cal, val, iroots, nfound, roots, dfin, iters = E.map_search_refine(
        MAP["omi"], MAP["omf"], MAP["gami"], MAP["gamf"], nr, ni, numroots=numroots, out=out)
    guesses = np.array([om[i - 1, j - 1] for i, j in iroots])
    lat = (E.last_root_times_ms(guesses.size) if guesses.size else np.zeros(0))
    return dict(roots=roots, dfin=dfin, iters=iters, guesses=guesses, latency_ms=lat,
                map_kernel_ms=E.last_kernel_ms, split_ms=E.last_kernel_split_ms, cal=cal, om=om,
                nfound=nfound, wall_ms=(time.perf_counter() - t0) * 1e3)


def oracle_roots(plasma, kperp, kpar, guesses, threads):
    """secant_osc of the CPU restatement from each guess: (roots, iterations)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    O = Oracle(plasma, n_threads=threads)
    O.set_k(kperp, kpar)
    out = [O.secant_osc(g) for g in guesses]
    return np.array([w for w, _ in out]), np.array([it for _, it in out])


def cpu_baseline(plasma, om_grid, cal_gpu, n_sample, threads):
    """The CPU restatement (oracle/) on a strided sample of the same map."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    O = Oracle(plasma, n_threads=threads)
    O.set_k(KPERP, KPAR)
    flat_om = om_grid.T.ravel()
    flat_cal = cal_gpu.T.ravel()
    idx = np.linspace(0, flat_om.size - 1, n_sample).astype(int)
    t0 = time.perf_counter()
    vals = [O.disp(flat_om[k], want_tensors=True) for k in idx]
    dt = time.perf_counter() - t0
    err = 0.0
    from_scale = []
    for (d, chi, low, wave), k in zip(vals, idx):
        w = np.abs(wave)
        sc = (w[0, 0] * (w[1, 1] * w[2, 2] + w[1, 2] ** 2) + 2 * w[0, 1] * w[1, 2] * w[0, 2]
              + w[0, 2] ** 2 * w[1, 1] + w[0, 1] ** 2 * w[2, 2])
        err = max(err, abs(flat_cal[k] - d) / sc)
    return dict(value=n_sample / dt, unit="chi-tensor evals/s", cores=threads, kind="port",
                sample=f"{n_sample} strided points of the 256x256 map, full 128x256 grid, "
                       f"oracle/liboracle.so with {threads} OpenMP threads",
                cpu_model=cpu_model(), wall_s=dt), err


# -- served root solves ------------------------------------------------------
SERVE_GRIDS = [(32, 64), (64, 128), (128, 256), (256, 512)]


def serve_variant(alph, nperp, npar, rp):
    """Protons (T_perp/T_par = alph), electrons, alphas at 5 % density drifting
    at 1 v_A along B (p = m v = 4), quasi-neutral, beta_par = 1, v_A/c = 1e-4."""
    from alps_amd.plasma import bimaxwellian_plasma
    return bimaxwellian_plasma(ms=[1.0, M_E, 4.0], qs=[1.0, -1.0, 2.0], ns=[1.0, 1.1, 0.05],
                               taus=[1.0, 1.0, 1.0], alphs=[alph, 1.0, 1.0], ps=[0.0, 0.0, 4.0],
                               beta=1.0, vA=1e-4, nperp=nperp, npar=npar, maxP=6.0,
                               positions_principal=3, n_resonance_interval=100, Tlim=0.01,
                               Bessel_zero=1e-45, roots=rp)


def serve_requests(n, n_variants, seed=7):
    """n requests over a pool of plasma variants (anisotropy uniform in [0.3, 3],
    grids cycling 32x64 .. 256x512): (variant, kperp, kpar, guess)."""
    from alps_amd.plasma import RootParams
    rng = np.random.default_rng(seed)
    rp = RootParams(secant_method=2, numiter=50, D_threshold=1e-12, D_prec=1e-5, D_gap=1e-5)
    pool = []
    for v in range(n_variants):
        nperp, npar = SERVE_GRIDS[v % len(SERVE_GRIDS)]
        pool.append(serve_variant(float(rng.uniform(0.3, 3.0)), nperp, npar, rp))
    lu = lambda a, b, size: np.exp(rng.uniform(np.log(a), np.log(b), size))
    kperp, kpar = lu(0.01, 3.0, n), lu(0.01, 3.0, n)
    guess = lu(0.01, 3.0, n) + 1j * rng.uniform(-0.1, 0.01, n)
    var = rng.integers(0, n_variants, n)
    return pool, [(int(var[i]), float(kperp[i]), float(kpar[i]), complex(guess[i]))
                  for i in range(n)]


def serve_rel_variant(tke, rp, nperp=200, npar=400, ngamma=100, npparbar=200):
    """Relativistic electrons (Juettner, kT_e / m_e c^2 = tke, the (Gamma,
    pparbar) tables of derivative_f0_rel) and cold-ish protons (T_p = 0.01
    T_e), on a 200 x 400 (pperp, ppar) grid, electron-normalised: masses m_e = 1,
    m_p = 1836, v_A / c = 0.3."""
    import dataclasses
    from alps_amd.plasma import bimaxwellian_plasma, juttner_plasma
    vA = 0.3
    tau = 2.0 * tke / (vA * vA)  # kT/mc^2 = vA^2 beta tau / (2 m) with beta = 1, m = 1
    r = juttner_plasma(ms=[1.0], qs=[-1.0], ns=[1.0], taus=[tau], beta=1.0, vA=vA, nperp=nperp,
                       npar=npar, maxP=5.0, ngamma=ngamma, npparbar=npparbar, analytic_rel=True,
                       positions_principal=3, n_resonance_interval=100, Tlim=0.01,
                       Bessel_zero=1e-45)
    b = bimaxwellian_plasma(ms=[1836.0], qs=[1.0], ns=[1.0], taus=[0.01 * tau], alphs=[1.0],
                            ps=[0.0], beta=1.0, vA=vA, nperp=nperp, npar=npar, maxP=5.0)
    npf = r.param_fit.shape[1]
    pf = np.zeros((2, npf, 5, 1))
    pf[0] = r.param_fit[0]
    pf[1, : b.param_fit.shape[1]] = b.param_fit[0]
    return dataclasses.replace(
        r, pp=np.concatenate([r.pp, b.pp]), f0=np.concatenate([r.f0, b.f0]), ns=[1.0, 1.0],
        qs=[-1.0, 1.0], ms=[1.0, 1836.0], fit_type=np.array([[4], [1]]),
        perp_correction=np.array([[r.perp_correction[0, 0]], [b.perp_correction[0, 0]]]),
        param_fit=pf, relativistic=[1, 0], ACmethod=None, n_fits=None, poly_kind=None,
        poly_order=None, logfit=None, poly_log_max=None, poly_fit_coeffs=None, usebM=None,
        bM_nmax=None, bM_Bessel_zero=None, bM_beta=None, bM_alpha=None, bM_pdrift=None, roots=rp)


def run_serve_rel(args, seed=5):
    """Served kpar-scan root requests on relativistic plasmas (kT_e / m_e c^2 in
    {0.1, 0.3, 1.0}): kpar log-uniform in [0.05, 2], kperp = 0.1 kpar, guesses
    near the whistler branch; roots/s and p50/p99 latency, and the plain path on
    a sample."""
    from alps_amd.engine import ALPSG_OK, Engine, RootServer
    from alps_amd.plasma import RootParams
    n_req = args.rel_requests
    rp = RootParams(secant_method=2, numiter=50, D_threshold=1e-12, D_prec=1e-5, D_gap=1e-5)
    pool = [serve_rel_variant(t, rp) for t in (0.1, 0.3, 1.0)]
    rng = np.random.default_rng(seed)
    kz = np.exp(rng.uniform(np.log(0.05), np.log(2.0), n_req))
    var = rng.integers(0, len(pool), n_req)
    guess = (0.5 * kz * kz) * (1.0 + rng.uniform(-0.05, 0.05, n_req)) - 1e-3j * kz
    S = RootServer(hbm_bytes=args.hbm_gb * 1e9, max_slots=min(args.slots, 1024))
    ids = [S.add_plasma(p) for p in pool]
    for v in range(len(pool)):  # warm-up: tables and code paths
        S.submit(ids[v], 0.05, 0.5, 0.125 - 5e-4j, rp, tag=-1 - v)
    S.drain(len(pool))
    st0 = S.stats()
    t0 = time.perf_counter()
    S.submit_batch(np.array([ids[v] for v in var], dtype=np.int32), 0.1 * kz, kz, guess, rp,
                   np.arange(n_req))
    got = S.drain(n_req)
    wall = time.perf_counter() - t0
    st = S.stats()
    S.close()
    # the relativistic kernels: 55 flops per (point, harmonic) of the 2-D
    # trapezoid of the non-resonant harmonics (k_rel_nonres) over the time of
    # both relativistic kernels -- a lower bound: the resonant harmonics'
    # Gamma-row integrals and the resonance scans are not counted
    rel_pts = st["work_rel_points"] - st0["work_rel_points"]
    rel_ms = st["rel_ms"] - st0["rel_ms"]
    rel_ach = 55.0 * rel_pts / (rel_ms * 1e-3) / 1e12 if rel_ms > 0 else 0.0
    lat = np.array([got[i][4] for i in range(n_req)])
    status = np.array([got[i][3] for i in range(n_req)])
    iters = np.array([got[i][2] for i in range(n_req)])
    served = np.array([got[i][1] for i in range(n_req)])
    idx = np.linspace(0, n_req - 1, 8).astype(int)
    engines, plain_t, dev, n_cmp = {}, 0.0, 0.0, 0
    for i in idx:
        v = int(var[i])
        E = engines.setdefault(v, Engine(pool[v]))
        tp = time.perf_counter()
        try:
            E.set_k(0.1 * kz[i], kz[i])
            r, it = E.secant_osc(guess[i])
        except Exception:
            r, it = None, rp.numiter
        plain_t += time.perf_counter() - tp
        # roots that converged on both paths (|D| < D_threshold fixes them only
        # to |D| / |dD/dw|, so the two paths' rounding shows at that level)
        if r is not None and status[i] == ALPSG_OK and it < rp.numiter:
            dev = max(dev, abs(r - served[i]) / max(abs(r), 1e-300))
            n_cmp += 1
    conv = status == ALPSG_OK
    res = {
        "workload": "relativistic kpar-scan requests: Juettner electrons (kT/mc^2 0.1, 0.3, 1.0) "
                    "and cold protons, 200 x 400 (pperp, ppar), 100 x 200 (Gamma, pparbar), "
                    "kpar in [0.05, 2], all submitted at once",
        "requests": n_req,
        "wall_s": wall,
        "roots_per_sec": int(np.sum(conv)) / wall,
        "converged": int(np.sum(conv)),
        "max_iterations_reached": int(np.sum(status == 6)),
        "failed": int(np.sum((status != ALPSG_OK) & (status != 6))),
        "latency_p50_ms": float(np.percentile(lat[conv], 50)) if conv.any() else None,
        "latency_p99_ms": float(np.percentile(lat[conv], 99)) if conv.any() else None,
        "mean_iterations": float(iters.mean()),
        "roofline_rel": {"kernel": "k_rel_nonres + k_rel_res (all passes)", "bound": "fp64 valu",
                         "flops_lower_bound": 55.0 * rel_pts, "kernel_ms": rel_ms,
                         "achieved_lower_bound": rel_ach, "peak": FP64_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac_lower_bound": rel_ach / FP64_PEAK_TFLOPS,
                         "kernel_ms_split": {k: st[k] - st0[k] for k in
                                             ("sweep_ms", "resonant_ms", "rel_ms", "assemble_ms")}},
        "plain_roots_per_sec": len(idx) / plain_t,
        "plain_sample": int(len(idx)),
        "max_rel_dev_served_vs_plain_converged": dev,
        "compared": n_cmp,
    }
    res["speedup_vs_plain"] = res["roots_per_sec"] / res["plain_roots_per_sec"]
    if not args.no_cpu:
        # two converged served roots against the oracle's secant_osc (about
        # 3 s each on the host at this grid size)
        ci = [i for i in idx if conv[i]][:2]
        err = 0.0
        thr = min(args.cpu_threads, os.cpu_count() or 1)
        tc = time.perf_counter()
        for i in ci:
            w, _ = oracle_roots(pool[int(var[i])], 0.1 * kz[i], kz[i], [guess[i]], thr)
            err = max(err, abs(served[i] - w[0]) / abs(w[0]))
        t_cpu = time.perf_counter() - tc
        res["roots_vs_oracle"] = {"compared": len(ci), "max_rel_err": err, "tolerance": 1e-10}
        if ci:
            res["cpu_baseline"] = {"value": len(ci) / t_cpu, "unit": "roots/s", "cores": thr,
                                   "kind": "port", "cpu_model": cpu_model(),
                                   "sample": f"{len(ci)} requests of the stream, oracle/"
                                             f"liboracle.so with {thr} OpenMP threads",
                                   "wall_s": t_cpu}
    return res


def served_roofline(st0, st):
    """The served node sweep (k_sweep_req) against the FP64 peak: 54 flops per
    node record (species, harmonic, ppar node) of every evaluated omega, over
    its device time from the server's events; the resonant parts beside it."""
    nodes = st["work_nodes"] - st0["work_nodes"]
    ms = st["sweep_ms"] - st0["sweep_ms"]
    flops = 54.0 * nodes
    ach = flops / (ms * 1e-3) / 1e12 if ms > 0 else 0.0
    return {"kernel": "k_sweep_req (all passes)", "bound": "fp64 valu", "flops": flops,
            "node_records": nodes, "kernel_ms": ms, "achieved": ach, "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s", "frac": ach / FP64_PEAK_TFLOPS,
            "resonant_ms": st["resonant_ms"] - st0["resonant_ms"],
            "rel_ms": st["rel_ms"] - st0["rel_ms"],
            "assemble_ms": st["assemble_ms"] - st0["assemble_ms"]}


def run_serve(args):
    """Poisson stream of root requests through the continuous-batching server;
    the same requests (a sample) on the plain one-root-per-launch path and on
    the CPU restatement."""
    from alps_amd.engine import ALPSG_OK, Engine, RootServer
    pool, reqs = serve_requests(args.requests, args.variants)
    rng = np.random.default_rng(11)
    S = RootServer(hbm_bytes=args.hbm_gb * 1e9, max_slots=args.slots)
    # the plasma variants are registered once; requests name them (inline
    # plasmas with --inline: each request carries and uploads its own tables)
    ref = list(pool) if args.inline else [S.add_plasma(p) for p in pool]
    # warm-up: one request of each grid size
    for v in range(min(len(pool), len(SERVE_GRIDS))):
        S.submit(ref[v], 0.5, 0.5, 0.5 - 0.01j, pool[v].roots, tag=-1 - v)
    S.drain(min(len(pool), len(SERVE_GRIDS)))
    st0 = S.stats()
    gaps = rng.exponential(1.0 / args.rate, len(reqs))
    t_arr = np.cumsum(gaps)
    rp = pool[0].roots
    t0 = time.perf_counter()
    if args.inline:
        for i, (v, kp, kz, g) in enumerate(reqs):
            dt = t_arr[i] - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
            S.submit(ref[v], kp, kz, g, rp, tag=i)
    else:
        # every request whose arrival time has passed goes in one call
        ids = np.array([ref[v] for v, *_ in reqs], dtype=np.int32)
        kps = np.array([r[1] for r in reqs])
        kzs = np.array([r[2] for r in reqs])
        gs = np.array([r[3] for r in reqs])
        i = 0
        while i < len(reqs):
            now = time.perf_counter() - t0
            if t_arr[i] > now:
                time.sleep(t_arr[i] - now)
                now = time.perf_counter() - t0
            j = int(np.searchsorted(t_arr, now, side="right"))
            j = max(j, i + 1)
            S.submit_batch(ids[i:j], kps[i:j], kzs[i:j], gs[i:j], rp, np.arange(i, j))
            i = j
    t_sub = time.perf_counter() - t0
    got = S.drain(len(reqs))
    wall = time.perf_counter() - t0
    st = S.stats()
    S.close()
    lat = np.array([got[i][4] for i in range(len(reqs))])
    status = np.array([got[i][3] for i in range(len(reqs))])
    iters = np.array([got[i][2] for i in range(len(reqs))])
    served = np.array([got[i][1] for i in range(len(reqs))])
    evals = st["evals"] - st0["evals"]
    passes = st["passes"] - st0["passes"]
    # plain path: one Engine per variant, set_k and secant_osc with one root per launch
    engines = {}
    idx = np.linspace(0, len(reqs) - 1, args.plain_requests).astype(int)
    for i in idx[:2]:
        v = reqs[i][0]
        engines.setdefault(v, Engine(pool[v]))
    plain_t = 0.0
    dev = 0.0
    conv_plain = 0
    for i in idx:
        v, kp, kz, g = reqs[i]
        E = engines.get(v)
        if E is None:
            E = engines[v] = Engine(pool[v])
        tp = time.perf_counter()
        E.set_k(kp, kz)
        r, it = E.secant_osc(g)
        plain_t += time.perf_counter() - tp
        if status[i] == ALPSG_OK:
            dev = max(dev, abs(r - served[i]) / max(abs(r), 1e-300))
        conv_plain += 1
    conv = status == ALPSG_OK
    out = {
        "requests": len(reqs),
        "plasma_variants": len(pool),
        "arrival_rate_per_s": args.rate,
        "submit_s": t_sub,
        "wall_s": wall,
        # converged solves only: a solve that used up numiter iterations
        # (ALPSG_MAXITER) or failed is not a root
        "roots_per_sec": int(np.sum(conv)) / wall,
        "converged": int(np.sum(conv)),
        "max_iterations_reached": int(np.sum(status == 6)),
        "failed": int(np.sum((status != ALPSG_OK) & (status != 6))),
        "finished_per_sec": len(reqs) / wall,
        "latency_p50_ms": float(np.percentile(lat[conv], 50)),
        "latency_p99_ms": float(np.percentile(lat[conv], 99)),
        "latency_all_p50_ms": float(np.percentile(lat, 50)),
        "latency_all_p99_ms": float(np.percentile(lat, 99)),
        "mean_iterations": float(iters.mean()),
        "passes": passes,
        "disp_evals": evals,
        "evals_per_pass": evals / max(passes, 1),
        "service_ms": {k: st[k] - st0[k] for k in ("pass_ms", "admit_ms", "launch_ms", "wait_ms")},
        "hbm_arena_gb": args.hbm_gb,
        "plasmas": "inline per request" if args.inline else "registered once",
        "plasma_uploads": st["plasma_uploads"] - st0["plasma_uploads"],
        "plasma_evictions": st["plasma_evictions"] - st0["plasma_evictions"],
        "slots": args.slots,
        "roofline_sweep": served_roofline(st0, st),
        "plain_roots_per_sec": len(idx) / plain_t,
        "plain_sample": int(len(idx)),
        "max_rel_dev_served_vs_plain": dev,
    }
    out["speedup_vs_plain"] = out["roots_per_sec"] / out["plain_roots_per_sec"]
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        thr = min(args.cpu_threads, os.cpu_count() or 1)
        n_cpu, t_cpu, err, n_conv_cpu, n_cmp, err_mi, n_mi, st_agree = 0, 0.0, 0.0, 0, 0, 0.0, 0, 0
        for i in idx:
            v, kp, kz, g = reqs[i]
            tc = time.perf_counter()
            O = Oracle(pool[v], n_threads=thr)
            O.set_k(kp, kz)
            w, it = O.secant_osc(g)
            t_cpu += time.perf_counter() - tc
            n_cpu += 1
            n_conv_cpu += int(it < rp.numiter)
            st_agree += int((status[i] == ALPSG_OK) == (it < rp.numiter))
            rel = abs(served[i] - w) / max(abs(w), 1e-300)
            if status[i] == ALPSG_OK and it < rp.numiter:
                err = max(err, rel)   # converged roots: the 1e-10 bar
                n_cmp += 1
            elif status[i] == 6:
                # both spent numiter: 50 steps of a non-converging secant
                # amplify rounding, so these best points agree more loosely
                err_mi = max(err_mi, rel)
                n_mi += 1
            if t_cpu > args.cpu_seconds:
                break
        out["cpu_baseline"] = {"value": n_conv_cpu / t_cpu, "unit": "roots/s", "cores": thr,
                               "kind": "port", "cpu_model": cpu_model(),
                               "sample": f"{n_cpu} requests of the stream (strided; "
                                         f"{n_conv_cpu} converged), oracle/liboracle.so with "
                                         f"{thr} OpenMP threads",
                               "wall_s": t_cpu}
        out["roots_vs_oracle"] = {"compared": n_cmp, "max_rel_err": err, "tolerance": 1e-10,
                                  "same_convergence_status": f"{st_agree}/{n_cpu}",
                                  "max_iteration_solves_compared": n_mi,
                                  "max_iteration_max_rel_dev": err_mi}
        out["speedup_vs_cpu"] = out["roots_per_sec"] / out["cpu_baseline"]["value"]
    return out


# the non-Maxwellian workload: proton core + 10 % beam drifting at 1.5 v_A,
# bi-kappa (T_perp/T_par = 2) with kappa in {3, 4, 5, 6}, beside electrons,
# f0 on 256 x 512; per kappa a 512 x 512 omega map at kperp = kpar = 0.3, and
# a 32 x 32 (kperp, kpar) double scan of one root
KAPPAS = (3.0, 4.0, 5.0, 6.0)
KMAP = dict(kperp=0.3, kpar=0.3, omi=0.01, omf=2.0, gami=-0.5, gamf=0.05)


def kappa_plasma_cb(kappa, nperp=256, npar=512):
    from alps_amd.plasma import RootParams, kappa_plasma
    rp = RootParams(secant_method=2, numiter=50, D_threshold=1e-12, D_prec=1e-5, D_gap=1e-5)
    return kappa_plasma(ms=[1.0, 1.0, M_E], qs=[1.0, 1.0, -1.0], ns=[0.9, 0.1, 1.0],
                        taus=[1.0, 1.0, 1.0], alphs=[2.0, 2.0, 1.0], ps=[0.0, 1.5, 0.0],
                        kappas=[kappa, kappa, kappa], beta=1.0, vA=1e-4, nperp=nperp, npar=npar,
                        maxP=6.0, positions_principal=3, n_resonance_interval=100, Tlim=0.01,
                        Bessel_zero=1e-45, roots=rp)


def run_kappa(args):
    """512 x 512 maps of det D, one per kappa, and a 1024-point double scan
    (om_double_scan, ALPS_fns.f90:2900-3591: kperp 0.3 -> 1 outer, kpar
    0.3 -> 1 inner, 31 log steps each) of the refined deepest minimum of the
    kappa = 4 map; chi-tensor evals/s, roots/s, the sweep's roofline, and the
    oracle on a sample of map points and scan roots."""
    import math
    from alps_amd.engine import Engine
    nr = ni = args.kmap
    maps, t_map, evals, split, nmaxs, grids, works = [], 0.0, 0, [], [], {}, []
    engines = {}
    for kap in KAPPAS:
        p = kappa_plasma_cb(kap)
        E = engines[kap] = Engine(p)
        nmaxs.append(E.set_k(KMAP["kperp"], KMAP["kpar"]))
        E.map_search(KMAP["omi"], KMAP["omf"], KMAP["gami"], KMAP["gamf"], nr, ni)  # warm-up
        ev0 = E.eval_count
        t0 = time.perf_counter()
        om, cal, val, iroots, nf = E.map_search(KMAP["omi"], KMAP["omf"], KMAP["gami"],
                                                KMAP["gamf"], nr, ni)
        t_map += time.perf_counter() - t0
        evals += E.eval_count - ev0
        split.append(E.last_kernel_split_ms)
        grids[kap] = (om, cal, val, iroots, p)
        works.append(E.sweep_work(om.T.ravel()))
    flops = sum(sweep_flops(w) for w in works)
    sweep_ms = sum(sp[0] for sp in split)
    # the double scan from the deepest minimum of the kappa = 4 map
    om4, cal4, val4, ir4, p4 = grids[4.0]
    E = engines[4.0]
    E.set_k(KMAP["kperp"], KMAP["kpar"])
    # refine the 16 deepest minima in one batch; the scan starts from the
    # deepest one whose solve converged
    cands = sorted(ir4, key=lambda ij: val4[ij[0] - 1, ij[1] - 1])[:16] or [(nr // 2, ni // 2)]
    gs = [om4[i - 1, j - 1] for i, j in cands]
    wb, _, itb = E.solve_batch(gs)
    q0 = next((q for q in range(len(gs)) if itb[q] < p4.roots.numiter), 0)
    guess, w0 = gs[q0], wb[q0:q0 + 1]
    n_s = args.kscan
    r, t_scan, n_pts = None, 0.0, 0
    if n_s >= 2:   # --kscan 0 or 1: the maps only (a profiler run of the maps)
        s1 = (3, 1, (math.log10(1.0) - math.log10(KMAP["kperp"])) / (n_s - 1), 0.0, n_s - 1, 1)
        s2 = (4, 1, (math.log10(1.0) - math.log10(KMAP["kpar"])) / (n_s - 1), 0.0, n_s - 1, 1)
        t0 = time.perf_counter()
        r = E.om_double_scan(s1, s2, [w0[0]])
        t_scan = time.perf_counter() - t0
        n_pts = int(np.sum(r["written"]))
    out = {
        "workload": "kappa core (90 %) + beam (10 %, 1.5 v_A) bi-kappa T_perp/T_par = 2, "
                    "electrons, 256 x 512; a 512 x 512 omega map per kappa in {3,4,5,6} at "
                    "kperp = kpar = 0.3; a 32 x 32 (kperp, kpar) om_double_scan 0.3 -> 1",
        "map": [nr, ni],
        "nmax": nmaxs,
        "map_evals_per_sec": evals / t_map,
        "map_ms": 1e3 * t_map / len(KAPPAS),
        "map_kernel_split_ms": [list(sp) for sp in split],
        "roofline_sweep": {"bound": "fp64 valu", "achieved": flops / (sweep_ms * 1e-3) / 1e12,
                           "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                           "frac": flops / (sweep_ms * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                           "flops": flops, "kernel_ms": sweep_ms,
                           "flops_if_all_rows_node_summed":
                               sum(flops_per_eval(nm, 512) for nm in nmaxs) * nr * ni,
                           "kernel": "k_sweep, the four %dx%d map batches" % (nr, ni)},
        "scan_points": n_pts,
        "scan_s": t_scan,
        "scan_roots_per_sec": n_pts / t_scan if t_scan > 0 else None,
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        thr = min(args.cpu_threads, os.cpu_count() or 1)
        # 16 strided points of the kappa = 4 map, then the first 2 scan points
        O = Oracle(p4, n_threads=thr)
        O.set_k(KMAP["kperp"], KMAP["kpar"])
        flat_om, flat_cal = om4.T.ravel(), cal4.T.ravel()
        idx = np.linspace(0, flat_om.size - 1, 16).astype(int)
        tc = time.perf_counter()
        err = 0.0
        for k in idx:
            d, _, _, wave = O.disp(flat_om[k], want_tensors=True)
            w = np.abs(wave)
            sc = (w[0, 0] * (w[1, 1] * w[2, 2] + w[1, 2] ** 2) + 2 * w[0, 1] * w[1, 2] * w[0, 2]
                  + w[0, 2] ** 2 * w[1, 1] + w[0, 1] ** 2 * w[2, 2])
            err = max(err, abs(flat_cal[k] - d) / sc)
        t_eval = time.perf_counter() - tc
        # the refined start, then the scan's points (0, 0) (the start re-solved)
        # and (0, 1), (0, 2) (inner steps), each against the oracle's
        # secant_osc from the same start at the same wavevector.  The test of
        # a GPU root is the oracle's |D| there (below D_threshold: it is a root
        # of the oracle's D); the roots themselves agree to 1e-10 only where
        # the two solves take the same steps (a strongly damped mode sits where
        # |D| is flat, and a solve that stops one step apart stops elsewhere
        # in the |D| < D_threshold set).
        numiter, thr_d = p4.roots.numiter, p4.roots.D_threshold
        pts = [(KMAP["kperp"], KMAP["kpar"], guess, w0[0], int(itb[q0]))]
        prev = w0[0]
        for j in ((0, 1, 2) if r is not None else ()):
            pts.append((r["kperp"][0, j], r["kpar"][0, j], prev, r["roots"][0, j, 0], None))
            prev = r["roots"][0, j, 0]
        dmax, drel, agree, rows = 0.0, 0.0, 0, []
        for kp, kz, start, g, git in pts:
            O.set_k(kp, kz)
            dgc, _, _, wave = O.disp(g, want_tensors=True)
            w = np.abs(wave)
            sc = (w[0, 0] * (w[1, 1] * w[2, 2] + w[1, 2] ** 2) + 2 * w[0, 1] * w[1, 2] * w[0, 2]
                  + w[0, 2] ** 2 * w[1, 1] + w[0, 1] ** 2 * w[2, 2])
            dg = abs(dgc)
            dmax = max(dmax, dg / thr_d)
            drel = max(drel, dg / sc)
            o, oit = O.secant_osc(start)
            agree += int(abs(g - o) <= 1e-10 * abs(o))
            rows.append([g.real, g.imag, o.real, o.imag, git, oit, dg])
        out["cpu_baseline"] = {"value": len(idx) / t_eval, "unit": "chi-tensor evals/s",
                               "cores": thr, "kind": "port", "cpu_model": cpu_model(),
                               "sample": "16 strided points of the kappa = 4 map, oracle/"
                                         f"liboracle.so with {thr} OpenMP threads",
                               "wall_s": t_eval}
        out["map_vs_oracle_on_sample"] = err
        out["roots_vs_oracle"] = {"roots": len(pts),
                                  "max_oracle_abs_D_over_threshold": dmax,
                                  "max_oracle_abs_D_over_term_scale": drel,
                                  "agree_to_1e-10": agree,
                                  "rows": "GPU root re, im; oracle root from the same start re, im; "
                                          "GPU steps (start solve only); oracle steps; "
                                          "oracle |D| at the GPU root",
                                  "pairs": rows}
        out["speedup_vs_cpu"] = out["map_evals_per_sec"] / out["cpu_baseline"]["value"]
    return out


# the CPU-runnable scan: 2 species on 50 x 100, the root refined at
# kperp d_p = 0.01, kpar d_p = 0.1, then followed over 64 log steps to kpar d_p = 1
SCAN = dict(nperp=50, npar=100, kperp=0.01, kpar=0.1, kpar_end=1.0, guess=0.1 - 0.001j, steps=64)


def run_scan(args):
    """secant_osc from 0.1 - 0.001i to |det D| < 1e-12, then om_scan type 4
    over 64 logarithmic kpar steps 0.1 -> 1.0 (ALPS_fns.f90:1919-2101,
    2198-2600): roots/s (converged), per-step latency, the plain path (one
    disp launch per frequency, stepped on the host), the oracle's own run of
    the same scan as the CPU baseline, and every root against it."""
    import math
    from alps_amd.engine import Engine
    p = make_plasma(SCAN["nperp"], SCAN["npar"])
    n = SCAN["steps"]
    diff = (math.log10(SCAN["kpar_end"]) - math.log10(SCAN["kpar"])) / n
    E = Engine(p)

    def one():
        E.set_k(SCAN["kperp"], SCAN["kpar"])
        t0 = time.perf_counter()
        om, dfin, it0 = E.solve_batch([SCAN["guess"]])
        t_ref = (time.perf_counter() - t0) * 1e3
        kp, kz, roots = E.om_scan(4, True, diff, 0.0, n, 1, [om[0]])
        wall = time.perf_counter() - t0
        ms, its = E.last_scan_steps(n, 1)
        steps_ms = np.diff(np.concatenate([[0.0], ms]))
        conv = int(it0[0] < p.roots.numiter) + int(np.sum((its >= 0) & (its < p.roots.numiter)))
        return dict(wall=wall, refine_ms=t_ref, steps_ms=steps_ms, roots=roots[:, 0], kpar=kz,
                    converged=conv)

    reps = max(1, args.scan_repeats)
    one()  # warm-up: graphs and launch shapes
    runs = [one() for _ in range(reps)]
    wall = sum(r["wall"] for r in runs)
    n_conv = sum(r["converged"] for r in runs)
    step_ms = np.concatenate([r["steps_ms"] for r in runs])
    E.set_plain_roots(True)
    one()
    plain = [one() for _ in range(max(1, reps // 2))]
    E.set_plain_roots(False)
    plain_rate = sum(r["converged"] for r in plain) / sum(r["wall"] for r in plain)
    plain_dev = max(float(np.max(np.abs(r["roots"] - runs[-1]["roots"]) / np.abs(runs[-1]["roots"])))
                    for r in plain)
    out = {
        "workload": "refine one root at kperp d_p = 0.01, kpar d_p = 0.1 from 0.1-0.001i, then "
                    "om_scan type 4 over 64 log kpar steps to 1.0; 2 species on 50 x 100",
        "scans": reps,
        "roots_per_scan": n + 1,
        "roots_per_sec": n_conv / wall,
        "converged": n_conv,
        "scan_ms": wall * 1e3 / reps,
        "refine_ms": float(np.mean([r["refine_ms"] for r in runs])),
        "step_latency_p50_ms": float(np.percentile(step_ms, 50)),
        "step_latency_p99_ms": float(np.percentile(step_ms, 99)),
        "plain_roots_per_sec": plain_rate,
        "plain_scan_ms": 1e3 * sum(r["wall"] for r in plain) / len(plain),
        "max_rel_dev_plain_vs_batched": plain_dev,
        "final_root": [runs[-1]["roots"][-1].real, runs[-1]["roots"][-1].imag],
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        from oracle import Oracle
        thr = min(args.cpu_threads, os.cpu_count() or 1)
        O = Oracle(p, n_threads=thr)
        tc = time.perf_counter()
        O.set_k(SCAN["kperp"], SCAN["kpar"])
        w0, it0 = O.secant_osc(SCAN["guess"])
        rc, _, _, oroots = O.om_scan(4, True, diff, 0.0, n, 1, [w0])
        t_cpu = time.perf_counter() - tc
        g = runs[-1]["roots"]
        out["cpu_baseline"] = {"value": (n + 1) / t_cpu, "unit": "roots/s", "cores": thr,
                               "kind": "port", "cpu_model": cpu_model(),
                               "sample": f"the whole scan ({n + 1} roots), oracle/liboracle.so "
                                         f"with {thr} OpenMP threads",
                               "wall_s": t_cpu}
        out["roots_vs_oracle"] = {
            "compared": n + 1,
            "max_rel_err": float(np.max(np.abs(g - oroots[:, 0]) / np.abs(oroots[:, 0]))),
            "tolerance": 1e-10}
        out["speedup_vs_cpu"] = out["roots_per_sec"] / out["cpu_baseline"]["value"]
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nr", type=int, default=256)
    ap.add_argument("--ni", type=int, default=256)
    ap.add_argument("--cpu-sample", type=int, default=64)
    ap.add_argument("--cpu-threads", type=int, default=8)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain", type=int, default=32, help="omegas timed on the plain path")
    ap.add_argument("--shapes", default=None,
                    help="JSON {batch size: [5 ints]} pinning launch shapes instead of autotuning "
                         "(e.g. the launch_shapes of an earlier run, for a profiler pass)")
    ap.add_argument("--pmc-shapes", action="store_true",
                    help="run one map batch per one-lane sweep shape (for counter passes)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 3 device-to-device copies of MIB MiB (torch) and exit: the "
                         "known-size kernel that scales the profiler's FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--workload", choices=["map", "serve", "serve_rel", "scan", "kappa"],
                    default="map",
                    help="map: the batched map search (default); serve: the Poisson stream "
                         "of independent root requests through the root server; serve_rel: "
                         "the relativistic kpar-scan requests through the root server; scan: "
                         "one root refined and followed over 64 log kpar steps (50 x 100); "
                         "kappa: the non-Maxwellian core-beam maps and double scan")
    ap.add_argument("--requests", type=int, default=4096)
    ap.add_argument("--rel-requests", type=int, default=1024,
                    help="serve_rel: relativistic kpar-scan requests")
    ap.add_argument("--scan-repeats", type=int, default=10, help="scan: timed repeats")
    ap.add_argument("--kmap", type=int, default=512, help="kappa: omega map side")
    ap.add_argument("--kscan", type=int, default=32, help="kappa: double-scan side")
    ap.add_argument("--variants", type=int, default=64, help="plasma variants in the request pool")
    ap.add_argument("--rate", type=float, default=20000.0, help="request arrivals per second")
    ap.add_argument("--slots", type=int, default=4096)
    ap.add_argument("--hbm-gb", type=float, default=32.0)
    ap.add_argument("--plain-requests", type=int, default=32)
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--no-serve", action="store_true",
                    help="map: skip the served-requests leg reported beside the map line")
    ap.add_argument("--no-series", action="store_true",
                    help="map: every sweep row summed node by node (no far-row series), "
                         "for comparison")
    ap.add_argument("--inline", action="store_true",
                    help="serve: every request carries its plasma (no registry)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(the omega/request axis would shard across GPUs with no communication: "
              "one process and one handle or server per GPU)", file=sys.stderr)
        sys.exit(2)

    if args.copy_probe:
        import torch
        nb = args.copy_probe << 20
        a = torch.ones(nb // 8, dtype=torch.float64, device="cuda")
        c = torch.empty_like(a)
        for _ in range(3):
            c.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": nb, "copies": 3}))
        return

    if args.workload == "serve":
        sv = run_serve(args)
        print(json.dumps({"workload": "serve", **sv}))
        return
    if args.workload == "serve_rel":
        print(json.dumps(run_serve_rel(args)))
        return
    if args.workload == "scan":
        print(json.dumps({"workload": "scan", **run_scan(args)}))
        return
    if args.workload == "kappa":
        print(json.dumps(run_kappa(args)))
        return

    if args.pmc_shapes:
        # one map batch per one-lane sweep shape the autotuner can pick, for the
        # FETCH_SIZE / WRITE_SIZE counter passes (dispatches tell the shape by
        # their workgroup size and grid y = workgroups per omega block)
        from alps_amd.engine import Engine
        E = Engine(make_plasma())
        E.set_k(KPERP, KPAR)
        # one batch in the built-in shape first: the table builds' dirty lines
        # leave L2 during it, so none of the counted dispatches pays for them
        E.map_search(MAP["omi"], MAP["omf"], MAP["gami"], MAP["gamf"], args.nr, args.ni)
        shapes = []
        for w in (1, 2, 4):
            for rc in (1, 2, 3, 4, 6, 8, 12, 16):
                E.set_shape(args.nr * args.ni, (1, w, rc, 4, 4))
                E.map_search(MAP["omi"], MAP["omf"], MAP["gami"], MAP["gamf"], args.nr, args.ni)
                shapes.append([1, w, rc])
        print(json.dumps({"pmc_shapes": shapes}))
        return

    from alps_amd.engine import Engine
    p = make_plasma()
    E = Engine(p)
    if args.no_series:
        E.set_sweep_series(False)
    nmax = E.set_k(KPERP, KPAR)
    # warm-up, then tune the launch shapes on this workload's own batches: the
    # map grid, and the secant request batches (size classes 1, 2, 4, ... up to
    # the number of minima)
    first = run_step(E, args.nr, args.ni)
    if args.shapes:
        shapes = {int(k): tuple(v) for k, v in json.loads(args.shapes).items()}
        for k, v in shapes.items():
            E.set_shape(k, v)
    else:
        shapes = {args.nr * args.ni: E.autotune(args.nr * args.ni, first["om"].T.ravel())}
        g = first["guesses"]
        k = 1
        while g.size and k <= 2 * g.size:
            shapes[k] = E.autotune(k, np.resize(g, k))
            k *= 2
    # the map grids of every later step go to the same host arrays (a caller
    # that maps repeatedly reuses its buffers); the last step's stay readable
    N = args.nr * args.ni
    grids = (np.zeros(N, dtype=np.complex128), np.zeros(N, dtype=np.complex128), np.zeros(N))
    for _ in range(max(0, args.warmup - 1)):
        run_step(E, args.nr, args.ni, out=grids)

    ev0 = E.eval_count
    res = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        res.append(run_step(E, args.nr, args.ni, out=grids))
    t1 = time.perf_counter()
    evals = E.eval_count - ev0
    wall = t1 - t0
    ms_step = wall * 1e3 / args.steps
    n_roots = sum(r["roots"].size for r in res)
    # a root counts when its solve converged (|D| < D_threshold before numiter)
    n_conv = sum(int(np.sum(r["iters"] < p.roots.numiter)) for r in res)
    lat = np.concatenate([r["latency_ms"][r["iters"] < p.roots.numiter] for r in res]) \
        if n_conv else np.zeros(1)
    kern = np.array([r["map_kernel_ms"] for r in res])
    split = np.array([r["split_ms"] for r in res]).mean(axis=0)
    last = res[-1]
    conv = int(np.sum(last["iters"] < p.roots.numiter))

    # roofline of the dominant kernel: k_sweep of the map batch (HIP events
    # around it on the engine's stream, averaged over the timed steps)
    # algorithmic work of the map batch, from the device's own tally of the
    # rows each omega summed node by node (54 flops per (species, harmonic,
    # node)) and the far-row signs it summed by series (24 J + 12 flops for
    # degree J)
    f_eval = flops_per_eval(nmax, p.npar)
    work = E.sweep_work(first["om"].T.ravel())
    flops_launch = sweep_flops(work)
    achieved = flops_launch / (split[0] * 1e-3) / 1e12
    pv_rows, landau_rows = resonant_rows(p, KPAR, nmax, first["om"].T.ravel())
    res_flops = resonant_flops(pv_rows, landau_rows, p.n_resonance_interval, p.positions_principal)

    # plain path: one omega per launch
    plain_om = last["om"].ravel()[:: max(1, last["om"].size // args.plain)][: args.plain]
    E.disp_plain(plain_om[:2])
    tp = time.perf_counter()
    E.disp_plain(plain_om)
    plain_rate = plain_om.size / (time.perf_counter() - tp)

    out = {
        "metric": METRIC,
        "value": evals / wall,
        "unit": "chi-tensor evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (bi-Maxwellian f0 tables from the generator formulas)",
        "config": {"workload": "map_search 256x256 omega map + secant refinement of every "
                               "minimum, 2 species, nperp x npar = 128 x 256, "
                               "kperp d_p = kpar d_p = 0.5",
                   "nmax": nmax, "map": [args.nr, args.ni], "omega_re": [MAP["omi"], MAP["omf"]],
                   "omega_im": [MAP["gami"], MAP["gamf"]], "D_threshold": p.roots.D_threshold},
        "roots_per_sec": n_conv / wall,
        "roots_per_step": n_roots / args.steps,
        "roots_converged_per_step": n_conv / args.steps,
        "roots_converged_last_step": conv,
        "root_latency_p50_ms": float(np.percentile(lat, 50)),
        "root_latency_p99_ms": float(np.percentile(lat, 99)),
        "plain_evals_per_sec": plain_rate,
        "map_kernel_ms": float(kern.mean()),
        "map_kernel_split_ms": {"k_sweep": float(split[0]), "k_resonant": float(split[1]),
                                "k_assemble": float(split[2])},
        "map_evals_per_sec_device": args.nr * args.ni / (kern.mean() * 1e-3),
        "launch_shapes": {str(k): list(v) for k, v in shapes.items()},
        "far_row_series": not args.no_series,
        "roofline": {"bound": "fp64 valu", "compute": "fp64 (VALU)", "achieved": achieved,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": (SWEEP_TRAFFIC_BYTES.get(tuple(shapes.get(args.nr * args.ni, (0,) * 5)[:3]))
                                 if (args.nr, args.ni) == (256, 256) else None),
                     "flops_per_launch": flops_launch,
                     "kernel": "k_sweep (map batch, %d omegas)" % (args.nr * args.ni),
                     "kernel_ms": float(split[0]),
                     "work": {"node_units": work[0], "series_signs": work[1],
                              "series_degree_sum": work[2],
                              "node_units_if_all_rows_node_summed": f_eval / 54.0 * args.nr * args.ni,
                              "flops_if_all_rows_node_summed": f_eval * args.nr * args.ni}},
        "roofline_resonant": {"bound": "fp64 valu", "kernel": "k_resonant, both parts (map batch)",
                              "flops": res_flops, "kernel_ms": float(split[1]),
                              "achieved": res_flops / (split[1] * 1e-3) / 1e12,
                              "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                              "frac": res_flops / (split[1] * 1e-3) / 1e12 / FP64_PEAK_TFLOPS,
                              "pv_rows": pv_rows, "landau_rows": landau_rows},
        "cpu_baseline": None,
    }
    progress(f"map: {out['value']:.4g} evals/s, {out['roots_per_sec']:.4g} roots/s")
    if not args.no_cpu:
        thr = min(args.cpu_threads, os.cpu_count() or 1)
        cb, err = cpu_baseline(p, last["om"], last["cal"], args.cpu_sample, thr)
        out["cpu_baseline"] = cb
        out["parity_vs_oracle_on_sample"] = err
        # refined roots of the last step against the oracle's secant_osc from
        # the same minima (two of them: the oracle takes ~1 s per root here)
        pick = np.linspace(0, last["roots"].size - 1, min(2, last["roots"].size)).astype(int)
        if pick.size:
            w, it = oracle_roots(p, KPERP, KPAR, last["guesses"][pick], thr)
            g = last["roots"][pick]
            out["roots_vs_oracle"] = {
                "compared": int(pick.size),
                "max_rel_err": float(np.max(np.abs(g - w) / np.abs(w))),
                "same_iterations": bool(np.all(it == last["iters"][pick])),
                "tolerance": 1e-10}
    if not args.no_serve:
        # served root solves beside the map line: 4096 independent requests
        progress("map: oracle checks done")
        out["serve"] = run_serve(args)
        progress(f"serve: {out['serve']['roots_per_sec']:.4g} roots/s")
        # and relativistic kpar-scan requests through the same server
        out["serve_relativistic"] = run_serve_rel(args)
        progress(f"serve_relativistic: {out['serve_relativistic']['roots_per_sec']:.4g} roots/s")
        # the CPU-runnable scan: one root refined, then 64 log kpar steps
        out["scan"] = run_scan(args)
        progress(f"scan: {out['scan']['roots_per_sec']:.4g} roots/s")
        # the non-Maxwellian kappa core-beam maps and double scan
        out["kappa"] = run_kappa(args)
        progress(f"kappa: {out['kappa']['map_evals_per_sec']:.4g} evals/s")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
