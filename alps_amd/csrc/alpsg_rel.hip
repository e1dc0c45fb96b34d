// Relativistic species (relativistic = T) of disp(om), for a batch of omegas.
//
// The reference treats a relativistic species in two ways per harmonic n,
// chosen by determine_resonances (ALPS_fns.f90:680-701), which asks whether
// the resonance p_res = (Gamma m om - n q)/kpar of some point of the
// (pperp, ppar) grid falls into that point's own ppar cell:
//
//  * no such point: integrate() over the (pperp, ppar) grid with the point's
//    Gamma in resU (ALPS_fns.f90:799-864, 1560-1596).  Gamma couples pperp and
//    ppar, so the pperp moments of the non-relativistic sweep do not apply;
//    k_rel_nonres sums the 2-D trapezoid directly, one workgroup per
//    (omega, species, harmonic block), after deciding the flag of each n with
//    a workgroup-wide OR over the grid.
//  * otherwise: integrate_res_rel and landau_integrate_rel along the
//    resonance curve on the (Gamma, pparbar) grid (ALPS_fns_rel.f90:460-1092):
//    k_rel_res, one workgroup per (omega, species, n), lanes over the Gamma
//    rows.  The row integral is the trapezoid of resU_rel * int_T_rel around
//    the principal-value window, plus principal_integral_rel; int_T_rel reads
//    J_n from a per-k table on the same grid (k_rel_bessel) and the Landau
//    term evaluates int_T_res_rel with the 21-term series CBESSJ.
//
// Both kernels write their partial chi into slots of chiR (non-resonant:
// slot = harmonic block; resonant: slot = NB + n + nmaxR) and the n = -1, 0, 1
// terms into lowR, read by k_assemble for relativistic species only.
// Decision-critical arithmetic (resonance flags, window limits, the
// principal-value ladder positions) is evaluated without FMA contraction in
// the reference's operation order.
#include "alpsg_device.h"
#include "alpsg_fit.h"

namespace {

constexpr double PI_R = 3.141592653589793238462643383279502884;

__device__ __forceinline__ cd cdivr(cd a, double d) { return cd{a.re / d, a.im / d}; }
__device__ __forceinline__ double m1p(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }

// mode order k: 0 xx, 1 yy, 2 zz, 3 xy, 4 xz, 5 yz
__device__ constexpr int ACC_OF_MODE_R[6] = {0, 3, 2, 4, 1, 5};

// T-tensor constant C_k(n) of int_T (ALPS_fns.f90:1659-1703) for the
// (pperp, ppar) grid; z = kperp/q or 1/q.
__device__ __forceinline__ cd tconst_r(const ProbDev &Pd, const SpecDev &S, int n, int k) {
  const double kp = Pd.kperp, z = S.z;
  const bool kn = Pd.kperp_norm != 0;
  switch (k) {
    case 0: return mk(1.0 * (n * n) / (z * z), 0.0);
    case 1: return mk(kn ? 1.0 : kp * kp, 0.0);
    case 2: return mk(kn ? 1.0 : kp * kp, 0.0);
    case 3: return mk(0.0, (kn ? 1.0 : kp) * (1.0 * n) / z);
    case 4: return mk((kn ? 1.0 : kp) * (1.0 * n) / z, 0.0);
    default: return mk(0.0, kn ? -1.0 : -(kp * kp));
  }
}

// workgroup sum of a complex value (blockDim.x = 256 or 64), result to all
template <int NT>
__device__ __forceinline__ cd block_sum(cd v, double *lds) {
  for (int o = 32; o > 0; o >>= 1) {
    v.re += __shfl_xor(v.re, o, 64);
    v.im += __shfl_xor(v.im, o, 64);
  }
  if constexpr (NT == 64) return v;
  const int w = threadIdx.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    lds[2 * w] = v.re;
    lds[2 * w + 1] = v.im;
  }
  __syncthreads();
  cd t = mk(0.0, 0.0);
  for (int i = 0; i < NT / 64; ++i) t = t + mk(lds[2 * i], lds[2 * i + 1]);
  return t;
}

// ---------------------------------------------------------------------------
// k_rel_nonres: flags of every harmonic of the block and the integrate()
// terms of the non-resonant ones.  grid (omega, species, NB), 256 threads.
// ---------------------------------------------------------------------------
// three wavefronts per SIMD: with two harmonics per pass the natural
// register count allows two, and three (a few spills) ran the relativistic
// served stream 7 % faster
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_rel_nonres(const ProbDev *__restrict__ probp, int nom,
                                                    const double2 *__restrict__ om_in,
                                                    const int *__restrict__ active,
                                                    double2 *__restrict__ chiR,
                                                    double2 *__restrict__ lowR,
                                                    int *__restrict__ flags,
                                                    const int *__restrict__ pidx, int nstride) {
  __shared__ double lds[16];
  const int iom = blockIdx.x, s = blockIdx.y, NB = gridDim.z, nb = blockIdx.z;
  const ProbDev &Pd = pidx != nullptr ? probp[pidx[iom]] : *probp;
  if (s >= Pd.nspec) return;
  const SpecDev &S = Pd.sp[s];
  if (!S.rel) return;
  if (!(active == nullptr || active[iom] != 0)) return;
  const int nspec = Pd.nspec, nperp = S.nperp, npar = S.npar, nmax = S.nmax_rel;
  const double2 ov = om_in[iom];
  const cd om = mk(ov.x, ov.y);
  const double m = S.m, q = S.q, kpar = Pd.kpar;
  const double scale = 2.0 * PI_R * S.dpperp * S.dppar_abs * 0.25;
  cd res[6];
  for (int k = 0; k < 6; ++k) res[k] = mk(0.0, 0.0);
  // the Gamma range of the (pperp, ppar) points (from the host's table):
  // p_res(Gamma) = (Gamma m Re w - n q)/kpar is linear in Gamma, so a harmonic
  // whose p_res over [gmin, gmax] stays clear of [ppar(0), ppar(npar)] (with a
  // margin far above rounding) has no point in a resonant cell and skips the
  // scan below
  const double gmin = S.rel_gmin, gmax = S.rel_gmax;
  // whether harmonic n's p_res stays clear of the ppar grid over [gmin, gmax]
  auto clear_of = [&](int n) {
#pragma clang fp contract(off)
    const double nq = 1.0 * n * q;
    const double pa = (gmin * m * om.re - nq) / kpar, pb = (gmax * m * om.re - nq) / kpar;
    const double lo = S.ppar[0], hi = S.ppar[npar];
    const double tol = 1e-9 * (fabs(pa) + fabs(pb) + fabs(lo) + fabs(hi));
    return fmax(pa, pb) < lo - tol || fmin(pa, pb) > hi + tol;
  };
  // the block's harmonics n = -nmax + nb + j NB two at a time: both share each
  // point's loads in the resonance scan and in the trapezoid (the same terms,
  // summed in the same order, as one harmonic at a time)
  for (int n0 = -nmax + nb; n0 <= nmax; n0 += 2 * NB) {
    const int n1 = n0 + NB;
    const bool has1 = n1 <= nmax;
    const bool scan0 = !clear_of(n0), scan1 = has1 && !clear_of(n1);
    int hit0 = 0, hit1 = 0;
    if ((scan0 || scan1) && S.rel_gmono) {
      // Along a ppar column, Gamma does not decrease with iperp, and the
      // reference's pr = (Gamma m Re w - n q)/kpar (each operation rounded,
      // hence monotone in Gamma) is monotone along it: the column has a point
      // with ppar(ia) <= pr < ppar(ia+1) iff the first point past the cell's
      // near end (found by bisection on that same pr) is short of its far end
      for (int ia = threadIdx.x; ia < npar; ia += 256) {
        const double lo = S.ppar[ia], hi = S.ppar[ia + 1];
        const double *gc = S.gam_pp + ia;
        for (int h = 0; h < 2; ++h) {
          if (h == 0 ? !scan0 : !scan1) continue;
          const double nq = 1.0 * (h == 0 ? n0 : n1) * q;
          auto pr = [&](int ip) {
#pragma clang fp contract(off)
            return (gc[(size_t)ip * (npar + 1)] * m * om.re - nq) / kpar;
          };
          const bool up = pr(0) <= pr(nperp);
          // the first ip whose pr is past the near end: pr >= lo (rising) or
          // pr < hi (falling); nperp + 1 if none
          int a0 = 0, a1 = nperp + 1;
          while (a0 < a1) {
            const int mid = (a0 + a1) >> 1;
            const double v = pr(mid);
            if (up ? (v >= lo) : (v < hi))
              a1 = mid;
            else
              a0 = mid + 1;
          }
          if (a0 <= nperp) {
            const double v = pr(a0);
            if (lo <= v && hi > v) (h == 0 ? hit0 : hit1) = 1;
          }
        }
      }
    } else if (scan0 || scan1) {
#pragma clang fp contract(off)
      const double nq0 = 1.0 * n0 * q, nq1 = 1.0 * n1 * q;
      const double ik = 1.0 / kpar;
      // point pt = ip npar + ia, (ip, ia) stepped without a division per point.
      // The test takes the reference's quotient pr = (g m Re w - n q)/kpar;
      // the product by 1/kpar is within 4 ulp of it, so where that product is
      // farther than 1e-14 (relative) from both cell ends the test's answer
      // is the same, and only the other points take the division
      int ip = (int)threadIdx.x / npar, ia = (int)threadIdx.x % npar;
      for (int pt = threadIdx.x; pt < (nperp + 1) * npar; pt += 256) {
        const double g = S.gam_pp[ip * (npar + 1) + ia];
        const double lo = S.ppar[ia], hi = S.ppar[ia + 1];
        const double gm = g * m * om.re;
        if (scan0) {
          const double num = gm - nq0;
          double pr = num * ik;
          const double tol = 1e-14 * fabs(pr);
          if (fabs(pr - lo) <= tol || fabs(pr - hi) <= tol) pr = num / kpar;
          if (lo <= pr && hi > pr) hit0 = 1;
        }
        if (scan1) {
          const double num = gm - nq1;
          double pr = num * ik;
          const double tol = 1e-14 * fabs(pr);
          if (fabs(pr - lo) <= tol || fabs(pr - hi) <= tol) pr = num / kpar;
          if (lo <= pr && hi > pr) hit1 = 1;
        }
        for (ia += 256; ia >= npar; ia -= npar) ++ip;
      }
    }
    const int found0 = __syncthreads_or(hit0);
    const int found1 = __syncthreads_or(hit1);
    if (threadIdx.x == 0) {
      flags[((size_t)iom * nspec + s) * nstride + n0 + Pd.nmaxR] = found0;
      if (has1) flags[((size_t)iom * nspec + s) * nstride + n1 + Pd.nmaxR] = found1;
    }
    const bool do0 = !found0, do1 = has1 && !found1;
    if (!do0 && !do1) continue;
    const double *at0 = S.atab + (size_t)(n0 < 0 ? -n0 : n0) * (nperp + 1) * ATAB_STRIDE;
    const double *at1 = S.atab + (size_t)(n1 < 0 ? -n1 : n1) * (nperp + 1) * ATAB_STRIDE;
    cd acc0[6], acc1[6];
    for (int k = 0; k < 6; ++k) acc0[k] = acc1[k] = mk(0.0, 0.0);
    // one harmonic's terms of the point: resU (ALPS_fns.f90:1587-1592) with the
    // point's Gamma, num / den as num conj(den) / |den|^2 with one reciprocal
    // (|den| is O(1) here: no scaling needed), times the int_T moments
    auto add = [&](const double *at, int ip, double w, cd num, double dre, double dim, double p2,
                   cd acc[6]) {
      const double inv = w * rcp1_d(dre * dre + dim * dim);
      const cd r = mk((num.re * dre + num.im * dim) * inv, (num.im * dre - num.re * dim) * inv);
      const double a1 = at[ip * ATAB_STRIDE + 0], a2 = at[ip * ATAB_STRIDE + 1],
                   a3 = at[ip * ATAB_STRIDE + 2];
      const cd u1 = a1 * r, u3 = a3 * r;
      acc[0] = acc[0] + u1;
      acc[1] = acc[1] + p2 * u1;
      acc[2] = acc[2] + (p2 * p2) * u1;
      acc[3] = acc[3] + a2 * r;
      acc[4] = acc[4] + u3;
      acc[5] = acc[5] + p2 * u3;
    };
    // point pt = (ip - 1)(npar - 1) + ia - 1, stepped as above
    int ip = 1 + (int)threadIdx.x / (npar - 1), ia = 1 + (int)threadIdx.x % (npar - 1);
    for (int pt = threadIdx.x; pt < (nperp - 1) * (npar - 1); pt += 256) {
      const double w = (ip == nperp - 1 ? 1.0 : 2.0) * ((ia == 1 || ia == npar - 1) ? 1.0 : 2.0);
      const size_t o = (size_t)ip * (npar + 1) + ia;
      const double g = S.gam_pp[o], d1 = S.df1[o], d2 = S.df2[o];
      const double p1 = S.pperp[ip], p2 = S.ppar[ia];
      const double A = (kpar * rcp_d(g * m)) * (p1 * d2 - p2 * d1);
      const cd num = mk(q * (om.re * d1 + A), q * (om.im * d1));
      const double base = (g * m) * om.re - kpar * p2, dim = (g * m) * om.im;
      if (do0) add(at0, ip, w, num, base - (1.0 * n0) * q, dim, p2, acc0);
      if (do1) add(at1, ip, w, num, base - (1.0 * n1) * q, dim, p2, acc1);
      for (ia += 256; ia > npar - 1; ia -= npar - 1) ++ip;
    }
    for (int h = 0; h < 2; ++h) {
      const int n = h == 0 ? n0 : n1;
      if (h == 0 ? !do0 : !do1) continue;
      const cd *acc = h == 0 ? acc0 : acc1;
      const int an = n < 0 ? -n : n;
      cd v[6];
      for (int k = 0; k < 6; ++k) {
        const cd t = block_sum<256>(acc[ACC_OF_MODE_R[k]], lds);
        v[k] = (scale * t) * tconst_r(Pd, S, n, k);
      }
      if (n == 0) v[0] = v[3] = v[4] = mk(0.0, 0.0);
      for (int k = 0; k < 6; ++k) res[k] = res[k] + v[k];
      if (threadIdx.x == 0 && lowR != nullptr && an <= 1) {
        double2 *lo = lowR + ((size_t)iom * nspec + s) * 18 + 6 * (n + 1);
        for (int k = 0; k < 6; ++k) lo[k] = make_double2(v[k].re, v[k].im);
      }
    }
  }
  if (threadIdx.x == 0) {
    double2 *o = chiR + ((size_t)(nb * nspec + s) * 6) * nom + iom;
    for (int k = 0; k < 6; ++k) o[(size_t)k * nom] = make_double2(res[k].re, res[k].im);
  }
}

// ---------------------------------------------------------------------------
// the (Gamma, pparbar) grid: helpers of integrate_res_rel
// ---------------------------------------------------------------------------
struct RelCtx {
  const SpecDev *S;
  cd om;
  int n, npb;
  double vA, kpar, kperp;
  bool kn;
  // per-call factors of the integrand (not of any decision), so that a grid
  // point costs one reciprocal: vA/kpar, n q vA/(kpar m), 1/zbar, 1/dpparbar
  double vak, nterm, izbar, idpb;
};

// a / b for complex b of moderate size: a conj(b) / |b|^2 with one reciprocal
__device__ __forceinline__ cd cdiv_r(cd a, cd b) {
  const double inv = rcp_d(b.re * b.re + b.im * b.im);
  return mk((a.re * b.re + a.im * b.im) * inv, (a.im * b.re - a.re * b.im) * inv);
}

// bessel and besselP of int_T_rel (ALPS_fns_rel.f90:1324-1338) from the J table
__device__ __forceinline__ void rel_bessel(const RelCtx &C, int pt, double &b, double &bP) {
  const SpecDev &S = *C.S;
  const int n = C.n;
  const size_t np = (size_t)(S.ng + 1) * (S.npb + 1);
  auto J = [&](int k) { return S.relbes[(size_t)k * np + pt]; };
  b = n < 0 ? m1p(n) * J(-n) : J(n);
  if (n >= 1)
    bP = 0.5 * (J(n - 1) - J(n + 1));
  else if (n < -1)
    bP = 0.5 * ((m1p(n - 1) * J(-(n - 1))) - (m1p(n + 1) * J(-(n + 1))));
  else if (n == 0)
    bP = -J(1);
  else
    bP = 0.5 * (J(2) - J(0));
}

// int_T_rel for the six modes at grid point pt (ALPS_fns_rel.f90:1340-1366)
__device__ __forceinline__ void int_T_rel6(const RelCtx &C, int pt, cd T[6]) {
  const SpecDev &S = *C.S;
  const double g = S.grel[pt], pb = S.pbrel[pt];
  double pperpbar;
  {
#pragma clang fp contract(off)
    pperpbar = sqrt(g * g - 1.0 - pb * pb);
  }
  double b, bP;
  rel_bessel(C, pt, b, bP);
  const double iz = C.izbar, kp = C.kperp, nn = 1.0 * C.n;
  if (C.kn) {
    T[0] = mk(1.0 * (C.n * C.n) * b * b * (iz * iz), 0.0);
    T[1] = mk(bP * bP * pperpbar * pperpbar, 0.0);
    T[2] = mk(b * b * (pb * pb), 0.0);
    T[3] = mk(0.0, nn * b * bP * pperpbar * iz);
    T[4] = mk(nn * b * b * pb * iz, 0.0);
    T[5] = mk(0.0, -(b * bP * pb * pperpbar));
  } else {
    T[0] = mk(1.0 * (C.n * C.n) * b * b * (iz * iz), 0.0);
    T[1] = mk(kp * kp * bP * bP * pperpbar * pperpbar, 0.0);
    T[2] = mk(kp * kp * b * b * (pb * pb), 0.0);
    T[3] = mk(0.0, nn * kp * b * bP * pperpbar * iz);
    T[4] = mk(nn * kp * b * b * pb * iz, 0.0);
    T[5] = mk(0.0, -(kp * kp * b * bP * pb * pperpbar));
  }
}

// the numerator of resU_rel and funct_g_rel: pref * (om df/dGamma + (kpar/vA) df/dpparbar)
__device__ __forceinline__ cd rel_num(const RelCtx &C, int pt) {
  const double2 d = C.S->dfrel[pt];
  const cd t = mk(C.om.re * d.x + (C.kpar / C.vA) * d.y, C.om.im * d.x);
  return C.S->rel_pref * t;
}

// funct_g_rel's interpolation cell (ALPS_fns_rel.f90:918-999): the index
// close of row ig whose neighbours close-1, close+1 it interpolates at real x
// (*lo, *hi): the bounds pb[k] <= x < pb[k+1] of the cell the search found,
// or (+inf, -inf) if none: any x' in [lo, hi) gives the same index
__device__ int rel_close(const RelCtx &C, int ig, double x, double *lo = nullptr,
                         double *hi = nullptr) {
  const SpecDev &S = *C.S;
  const int npb = C.npb, row = ig * (npb + 1);
  const double *pb = S.pbrel + row;
  int close = -2;
  double cl = 1.0 / 0.0, ch = -1.0 / 0.0;
  {
#pragma clang fp contract(off)
    // the last ip in [0, npb-1] with pb[ip] <= x < pb[ip+1] (increasing row)
    int k = (int)floor((x - pb[0]) / S.rel_dpb);
    if (k >= -1 && k <= npb) {
      k = max(0, min(k, npb - 1));
      while (k > 0 && pb[k] > x) --k;
      while (k < npb - 1 && pb[k + 1] <= x) ++k;
      if (pb[k + 1] > x && pb[k] <= x) {
        close = k;
        cl = pb[k];
        ch = pb[k + 1];
      }
    }
  }
  if (lo) {
    *lo = cl;
    *hi = ch;
  }
  const double *f0 = S.f0rel + row;
  if (close + 1 >= 0 && close + 1 <= npb && f0[close + 1] <= -1.0) close = close - 1;
  if (close - 1 >= 0 && close - 1 <= npb && f0[close - 1] <= -1.0) close = close + 1;
  if (x == pb[npb]) close = npb - 2;
  if (close >= npb - 1) close = npb - 2;
  if (close <= 1) close = 2;
  return close;
}

// funct_g_rel (ALPS_fns_rel.f90:918-999) for the six modes at real pparbar x
// of row ig: g(close) + (x - pparbar(close)) (g(close+1) - g(close-1)) /
// (2 dpparbar), g = -(rel_num T), with one node's T values live at a time
__device__ void funct_g_rel6(const RelCtx &C, int ig, double x, cd g[6]) {
  const SpecDev &S = *C.S;
  const int npb = C.npb, row = ig * (npb + 1);
  const double *pb = S.pbrel + row;
  const int close = rel_close(C, ig, x);
  const double f = 0.5 * ((x - pb[close]) * C.idpb);
  cd T[6];
  int_T_rel6(C, row + close, T);
  const cd nc = rel_num(C, row + close);
  for (int k = 0; k < 6; ++k) g[k] = -1.0 * (nc * T[k]);
  int_T_rel6(C, row + close + 1, T);
  const cd np_ = -f * rel_num(C, row + close + 1);
  for (int k = 0; k < 6; ++k) g[k] = g[k] + np_ * T[k];
  int_T_rel6(C, row + close - 1, T);
  const cd nm = f * rel_num(C, row + close - 1);
  for (int k = 0; k < 6; ++k) g[k] = g[k] + nm * T[k];
}

// the resU_rel * int_T_rel terms of point ip of row ig, weight w, into acc
__device__ __forceinline__ void rel_point(const RelCtx &C, int ig, int ip, double w, cd acc[6]) {
  const SpecDev &S = *C.S;
  const int pt = ig * (C.npb + 1) + ip;
  const double G = S.grel[pt], PB = S.pbrel[pt];
  const cd gom = mk(G * C.om.re * C.vak, G * C.om.im * C.vak);
  const cd den = mk((PB - gom.re) + C.nterm, -gom.im);
  const cd U = -1.0 * cdiv_r(rel_num(C, pt), den);
  cd T[6];
  int_T_rel6(C, pt, T);
  for (int k = 0; k < 6; ++k) acc[k] = acc[k] + w * (U * T[k]);
}

// funct_g_rel is linear in the values g_k(node) = -(rel_num T_k) at the three
// nodes of its cell (close - 1, close, close + 1): g(x) = g(close) +
// f (g(close+1) - g(close-1)), f = 0.5 (x - pparbar(close)) / dpparbar.  A
// weighted sum of g over a monotone ladder of x is therefore the sum over the
// cells it passes of three coefficients times the node values: RelCell holds
// the open cell's coefficients, and the node values are evaluated once per
// cell, when the ladder leaves it.
struct RelCell {
  int close;  // -1: empty
  cd cm, c0, cp;
  // the last cell search of this ladder: any x in [klo, khi) has index kc,
  // whose pparbar is kp (klo > khi: none yet)
  double klo, khi, kp;
  int kc;
};

__device__ __forceinline__ void rel_node_g(const RelCtx &C, int pt, cd g[6]) {
  cd T[6];
  int_T_rel6(C, pt, T);
  const cd nm = rel_num(C, pt);
  for (int k = 0; k < 6; ++k) g[k] = -1.0 * (nm * T[k]);
}

// the open cell's three coefficients, times f, into acc
__device__ void rel_cell_flush(const RelCtx &C, int ig, RelCell &cc, cd acc[6], double f) {
  if (cc.close < 0) return;
  const int pt = ig * (C.npb + 1) + cc.close;
  cd g[6];
  rel_node_g(C, pt, g);
  const cd c0 = f * cc.c0, cp = f * cc.cp, cm = f * cc.cm;
  for (int k = 0; k < 6; ++k) acc[k] = acc[k] + c0 * g[k];
  rel_node_g(C, pt + 1, g);
  for (int k = 0; k < 6; ++k) acc[k] = acc[k] + cp * g[k];
  rel_node_g(C, pt - 1, g);
  for (int k = 0; k < 6; ++k) acc[k] = acc[k] + cm * g[k];
  cc.close = -1;
  cc.cm = cc.c0 = cc.cp = mk(0.0, 0.0);
}

// add wt g(x) through the cell cache (a cell leaving it goes into acc, times
// fac).  The ladder's x rise (or fall) monotonically: while x stays in the
// raw cell of the last search, the index is that search's.
__device__ __forceinline__ void rel_cell_add(const RelCtx &C, int ig, RelCell &cc, double x, cd wt,
                                             cd acc[6], double fac) {
  if (!(cc.klo <= x && x < cc.khi)) {
    cc.kc = rel_close(C, ig, x, &cc.klo, &cc.khi);
    cc.kp = C.S->pbrel[ig * (C.npb + 1) + cc.kc];
  }
  const int close = cc.kc;
  if (close != cc.close) {
    rel_cell_flush(C, ig, cc, acc, fac);
    cc.close = close;
  }
  const double f = 0.5 * ((x - cc.kp) * C.idpb);
  cc.c0 = cc.c0 + wt;
  cc.cp = cc.cp + f * wt;
  cc.cm = cc.cm - f * wt;
}

// principal_integral_rel (ALPS_fns_rel.f90:724-913), six modes, times smdelta
// and the row's trapezoid weight wr, added to acc (each term carries the
// factor, so that no partial sum of the row is held beside acc).  The ladder
// sums go through cell coefficients (RelCell), with the ladder's points
// visited in increasing x (the far end of each ladder after its inner
// points), so that each cell is opened once.
__device__ void principal_rel6(const RelCtx &C, int ig, int ipres, int upperlimit, double Tlim,
                               int pos, int nri, double wr, cd acc[6]) {
  const SpecDev &S = *C.S;
  const int row = ig * (C.npb + 1);
  double denomR, denomI, capDelta, smdelta;
  {
#pragma clang fp contract(off)
    const double G = S.grel[row + ipres];
    const double gr = G * C.om.re * C.vA / C.kpar, gi = G * C.om.im * C.vA / C.kpar;
    denomR = gr - (1.0 * C.n) * (S.q / S.m) * C.vA / C.kpar;
    denomI = gi;
    capDelta = denomR - S.pbrel[1 * (C.npb + 1) + ipres - pos];
    smdelta = capDelta / (1.0 * nri);
  }
  const double f = wr * smdelta;
  cd g[6], g2[6];
  RelCell up, dn;  // x = denomR + t (increasing) and its mirror 2 denomR - x (decreasing)
  up.close = dn.close = -1;
  up.cm = up.c0 = up.cp = dn.cm = dn.c0 = dn.cp = mk(0.0, 0.0);
  up.klo = dn.klo = 1.0 / 0.0;
  up.khi = dn.khi = -1.0 / 0.0;
  up.kc = dn.kc = 0;
  up.kp = dn.kp = 0.0;
  auto pv_pair = [&](double x, double wgt) {
    // wgt / (t - i denomI) and wgt / (t + i denomI), t = x - denomR: one
    // reciprocal, the second the conjugate of the first
    const double t = x - denomR;
    const double inv = wgt * rcp_d(t * t + denomI * denomI);
    const cd r = mk(t * inv, denomI * inv);
    rel_cell_add(C, ig, up, x, r, acc, f);
    rel_cell_add(C, ig, dn, 2.0 * denomR - x, mk(-r.re, r.im), acc, f);
  };
  if (fabs(denomI) > Tlim) {
    pv_pair(denomR, 1.0);
    for (int ip = 1; ip <= nri - 1; ++ip) {
      double x;
      {
#pragma clang fp contract(off)
        x = denomR + smdelta * ip;
      }
      pv_pair(x, 2.0);
    }
    pv_pair(denomR + capDelta, 1.0);
    rel_cell_flush(C, ig, up, acc, f);
    rel_cell_flush(C, ig, dn, acc, f);
  } else {
    funct_g_rel6(C, ig, denomR + S.rel_dpb, g);
    funct_g_rel6(C, ig, denomR - S.rel_dpb, g2);
    cd gp[6];
    for (int k = 0; k < 6; ++k) gp[k] = cdivr(g[k] - g2[k], 2.0 * S.rel_dpb);
    auto lorentz = [&](double x) {
      const double d = x - denomR;
      return (d * d) / (d * d + denomI * denomI);
    };
    {
      const double l = lorentz(denomR + capDelta);
      for (int k = 0; k < 6; ++k) acc[k] = acc[k] + (f * (2.0 * l)) * gp[k];
    }
    for (int ip = 1; ip <= nri - 1; ++ip) {
      double x;
      {
#pragma clang fp contract(off)
        x = denomR + smdelta * ip;
      }
      const double l = lorentz(x);
      for (int k = 0; k < 6; ++k) acc[k] = acc[k] + (f * ((2.0 * 2.0) * l)) * gp[k];
    }
    if (denomI != 0.0) {
      funct_g_rel6(C, ig, denomR, g);
      const double sg = denomI > 0.0 ? 1.0 : -1.0;
      for (int k = 0; k < 6; ++k)
        acc[k] = acc[k] + wr * timesI((sg * 2.0 * PI_R) * g[k]);  // (2 pi i g / smdelta) f
    }
  }
  int ntiny;
  double base;
  {
#pragma clang fp contract(off)
    base = S.pbrel[row + upperlimit] - denomR - capDelta;
    ntiny = (int)(base / smdelta);
  }
  if (ntiny > 0) {
    double correction;
    {
#pragma clang fp contract(off)
      correction = (base / (1.0 * ntiny)) / smdelta;
    }
    auto tiny = [&](double x, double wgt) {
      const double t = x - denomR;
      const double inv = (wgt * correction) * rcp_d(t * t + denomI * denomI);
      rel_cell_add(C, ig, up, x, mk(t * inv, denomI * inv), acc, f);
    };
    double x0, x1;
    {
#pragma clang fp contract(off)
      x0 = denomR + capDelta;
      x1 = denomR + capDelta + correction * smdelta * ntiny;
    }
    tiny(x0, 1.0);
    for (int ip = 1; ip <= ntiny - 1; ++ip) {
      double x;
      {
#pragma clang fp contract(off)
        x = denomR + capDelta + correction * smdelta * ip;
      }
      tiny(x, 2.0);
    }
    tiny(x1, 1.0);
    rel_cell_flush(C, ig, up, acc, f);
  }
}

// integrate_resU_rel (ALPS_fns_rel.f90:516-719) of row ig, six modes, times
// the row's trapezoid weight wr, added to acc (term by term); returns false,
// before adding anything, where the reference calls alps_error(8)
__device__ bool resU_row6(const RelCtx &C, int ig, int pos, int nri, double Tlim, double wr,
                          cd acc[6]) {
  const SpecDev &S = *C.S;
  const int npb = C.npb;
  const double dpb = S.rel_dpb;
  const double *pb2 = S.pbrel + 2 * (npb + 1);  // row 2 of pparbar_rel
  int ipres = 0;
  bool found = false;
  double pr;
  {
#pragma clang fp contract(off)
    const double g1 = S.grel[ig * (npb + 1) + 1];
    const double tr = g1 * C.om.re - (1.0 * C.n) * S.q / S.m;
    pr = tr * C.vA / C.kpar;
    if ((pr * pr) <= (g1 * g1 - 1.0)) {
      // the unique ip in [1, npb-2] with pb2[ip] <= pr < pb2[ip+1]
      if (pb2[1] <= pr && pr < pb2[npb - 1]) {
        int k = 1 + (int)((pr - pb2[1]) / dpb);
        k = max(1, min(k, npb - 2));
        while (k > 1 && pb2[k] > pr) --k;
        while (k < npb - 2 && pb2[k + 1] <= pr) ++k;
        ipres = k;
        found = true;
      }
    }
    for (int ip = 0; ip <= pos; ++ip) {
      if (pr >= (pb2[0] - dpb * ip) && pr < (pb2[0] - dpb * (ip - 1))) {
        ipres = -ip;
        found = true;
      }
      if (pr >= (pb2[npb - 1] + dpb * ip) && pr < (pb2[npb - 1] + dpb * (ip + 1))) {
        ipres = npb - 1 + ip;
        found = true;
      }
    }
  }
  const int2 lim = S.rowlim[ig];
  const int lower = lim.x, upper = lim.y;
  int int_start, int_end, lowerlimit, upperlimit;
  if (found) {
    int_start = lower;
    int_end = upper;
    lowerlimit = ipres - pos;
    upperlimit = ipres + pos + 1;
    if (ipres >= 0 && ipres <= npb) {
#pragma clang fp contract(off)
      if (fabs(pr - pb2[ipres]) > (0.5 * dpb)) upperlimit = upperlimit + 1;
    }
    if (lowerlimit < lower && upperlimit > upper) return false;
    if (lowerlimit <= lower) {
      int_start = 1;
      lowerlimit = 0;
      upperlimit = lower;
    } else if (upperlimit >= upper) {
      lowerlimit = upper;
      upperlimit = npb;
      int_end = npb - 1;
    }
  } else {
    int_start = lower;
    lowerlimit = upper;
    int_end = npb - 1;
    upperlimit = npb;
  }
  const double w1 = wr * dpb, w2 = 2.0 * w1;  // trapezoid weights times dpparbar and wr
  if (int_start <= lowerlimit) rel_point(C, ig, int_start, w1, acc);
  for (int ip = int_start + 1; ip <= lowerlimit - 1; ++ip) rel_point(C, ig, ip, w2, acc);
  if (int_start < lowerlimit) rel_point(C, ig, lowerlimit, w1, acc);
  if (upperlimit <= int_end) rel_point(C, ig, upperlimit, w1, acc);
  for (int ip = upperlimit + 1; ip <= int_end - 1; ++ip) rel_point(C, ig, ip, w2, acc);
  if (upperlimit < int_end) rel_point(C, ig, int_end, w1, acc);
  if (found && lowerlimit >= int_start && upperlimit <= int_end)
    principal_rel6(C, ig, ipres, upperlimit, Tlim, pos, nri, wr, acc);
  return true;
}

// complex ** integer by repeated squaring
__device__ __forceinline__ cd cpowi_d(cd a, int n) {
  cd pw = mk(1.0, 0.0), x = a;
  unsigned u = (unsigned)n;
  if (n == 0) return pw;
  for (;;) {
    if (u & 1u) pw = pw * x;
    u >>= 1;
    if (u)
      x = x * x;
    else
      break;
  }
  return pw;
}

// CBESSJ (ALPS_fns_rel.f90:1502-1553), 21 terms of the power series, for
// J_nu0, J_nu0+1, J_nu0+2 of the same z: the three
// orders share the powers (-z^2/4)^k, built by one product per term, and
// take the coefficients 1/(k! Gamma(nu+k+1)) from the table (requires
// nu0 + 2 < RFG_NU)
__device__ void cbessj3_d(const double *gtab, cd z, int nu0, cd J[3]) {
  const double *rf = gtab + GTAB_N + 21 + (size_t)nu0 * 21;
  const cd w = cdivr(-1.0 * (z * z), 4.0);
  cd t = mk(1.0, 0.0), s0 = mk(0.0, 0.0), s1 = s0, s2 = s0;
  for (int k = 0; k <= 20; ++k) {
    if (k > 0) t = t * w;
    s0 = s0 + t * rf[k];
    s1 = s1 + t * rf[21 + k];
    s2 = s2 + t * rf[42 + k];
  }
  const cd h = cdivr(z, 2.0);
  cd pw = cpowi_d(h, nu0);
  J[0] = pw * s0;
  pw = pw * h;
  J[1] = pw * s1;
  pw = pw * h;
  J[2] = pw * s2;
}

// int_T_res_rel (ALPS_fns_rel.f90:1374-1495), six modes, at complex pparbar p of row ig
__device__ void int_T_res_rel6(const RelCtx &C, const double *gtab, int ig, cd p, cd T[6]) {
  const SpecDev &S = *C.S;
  const double g = S.grel[ig * (C.npb + 1) + 1];
  const cd pp2 = p * p;
  const cd pperpbar = csqrtd(mk((g * g - 1.0) - pp2.re, -pp2.im));
  const cd z = S.rel_zfac * pperpbar;
  const int n = C.n;
  const int an = n < 0 ? -n : n;
  // the orders |n| - 1, |n|, |n| + 1 (0, 1, 2 for |n| <= 1)
  cd J[3];
  cbessj3_d(gtab, z, an >= 1 ? an - 1 : 0, J);
  cd b, bP;
  if (n >= 1) {
    b = J[1];
    bP = 0.5 * (J[0] - J[2]);
  } else if (n < -1) {
    b = m1p(n) * J[1];
    bP = 0.5 * ((m1p(n - 1) * J[2]) - (m1p(n + 1) * J[0]));
  } else if (n == 0) {
    b = J[0];
    bP = -1.0 * J[1];
  } else {  // n == -1
    b = m1p(n) * J[1];
    bP = 0.5 * (J[2] - J[0]);
  }
  const double zbar = S.rel_zbar, kp = C.kperp, nn = 1.0 * n;
  const double f = C.kn ? 1.0 : kp, f2 = C.kn ? 1.0 : kp * kp;
  T[0] = cdivr(1.0 * (n * n) * (b * b), zbar * zbar);
  T[1] = f2 * (((bP * bP) * pperpbar) * pperpbar);
  T[2] = f2 * ((b * b) * pp2);
  T[3] = cdivr(timesI((nn * f) * ((b * bP) * pperpbar)), zbar);
  T[4] = cdivr((nn * f) * ((b * b) * p), zbar);
  T[5] = -1.0 * timesI(f2 * (((b * bP) * p) * pperpbar));
}

// ---------------------------------------------------------------------------
// k_rel_res / k_rel_landau: one resonant harmonic, grid (omega, species,
// 2 nmaxR + 1), 64 threads over the Gamma rows.  k_rel_res (LANDAU = false):
// integrate_res_rel (the trapezoid and the principal-value part), into slot
// NB + z of chiR; k_rel_landau (LANDAU = true): landau_integrate_rel for
// Im om <= 0, into slot NB + (2 nmaxR + 1) + z, and added to the harmonic's
// lowR terms (k_rel_res wrote them).  Two launches rather than one body: the
// Landau rows' Bessel series and continuation need registers that would hold
// the whole pass to one wavefront per SIMD.
// ---------------------------------------------------------------------------
template <bool LANDAU>
__device__ __forceinline__ void rel_res_body(const ProbDev *__restrict__ probp, int nom,
                                             const double2 *__restrict__ om_in,
                                             const int *__restrict__ active,
                                             double2 *__restrict__ chiR,
                                             double2 *__restrict__ lowR,
                                             const int *__restrict__ flags, int NB,
                                             int *__restrict__ err,
                                             const int *__restrict__ pidx, int nstride) {
  const int iom = blockIdx.x, s = blockIdx.y;
  const ProbDev &Pd = pidx != nullptr ? probp[pidx[iom]] : *probp;
  if (s >= Pd.nspec) return;
  const SpecDev &S = Pd.sp[s];
  if (!S.rel) return;
  if (!(active == nullptr || active[iom] != 0)) return;
  const int nspec = Pd.nspec;
  const int n = (int)blockIdx.z - Pd.nmaxR;
  const int slot = NB + (LANDAU ? (int)gridDim.z : 0) + (int)blockIdx.z;
  double2 *o = chiR + ((size_t)(slot * nspec + s) * 6) * nom + iom;
  const bool inr = (n >= -S.nmax_rel && n <= S.nmax_rel);
  const bool found = inr && flags[((size_t)iom * nspec + s) * nstride + blockIdx.z] != 0;
  if (!found) {
    if (threadIdx.x == 0)
      for (int k = 0; k < 6; ++k) o[(size_t)k * nom] = make_double2(0.0, 0.0);
    return;
  }
  const double2 ov = om_in[iom];
  RelCtx C;
  C.S = &S;
  C.om = mk(ov.x, ov.y);
  C.n = n;
  C.npb = S.npb;
  C.vA = Pd.vA;
  C.kpar = Pd.kpar;
  C.kperp = Pd.kperp;
  C.kn = Pd.kperp_norm != 0;
  C.vak = C.vA / C.kpar;
  C.nterm = (1.0 * C.n) * S.q * C.vA / (C.kpar * S.m);
  C.izbar = 1.0 / S.rel_zbar;
  C.idpb = 1.0 / S.rel_dpb;
  const int ng = S.ng;
  cd acc[6];
  for (int k = 0; k < 6; ++k) acc[k] = mk(0.0, 0.0);
  int bad = 0;
  for (int ig = 1 + threadIdx.x; ig <= ng - 1; ig += 64) {
    if constexpr (!LANDAU) {
      const double wr = ig <= ng - 2 ? 2.0 : 1.0;
      if (!resU_row6(C, ig, Pd.P, Pd.NRI, Pd.Tlim, wr, acc)) bad = 1;
    } else if (C.om.im <= 0.0) {
      // landau_integrate_rel row (ALPS_fns_rel.f90:1065-1083)
      const double g1 = S.grel[ig * (S.npb + 1) + 1];
      cd pres;
      bool in;
      {
#pragma clang fp contract(off)
        const double a = g1 * C.om.re * C.vA / C.kpar, b = g1 * C.om.im * C.vA / C.kpar;
        pres = mk(a - (1.0 * n) * S.q * C.vA / (C.kpar * S.m), b);
        in = (pres.re * pres.re) <= (g1 * g1 - 1.0);
      }
      if (in) {
        const double h = (ig == ng - 1) ? 0.5 : 1.0;
        cd dfg;
        // (a relativistic species' fits are stored in the Maxwell form)
        if (ig == 1)
          dfg = cdivr(eval_fit_maxwell(S, ig + 1, pres) - eval_fit_maxwell(S, ig, pres), S.rel_dg);
        else
          dfg = cdivr(eval_fit_maxwell(S, ig + 1, pres) - eval_fit_maxwell(S, ig - 1, pres),
                      2.0 * S.rel_dg);
        const cd dfp = cdivr(eval_fit_maxwell(S, ig, mk(pres.re + S.rel_dpb, pres.im)) -
                                 eval_fit_maxwell(S, ig, mk(pres.re - S.rel_dpb, pres.im)),
                             2.0 * S.rel_dpb);
        const cd fac = h * (C.om * dfg + (C.kpar / C.vA) * dfp);
        cd T[6];
        int_T_res_rel6(C, Pd.gtab, ig, pres, T);
        for (int k = 0; k < 6; ++k) acc[k] = acc[k] - fac * T[k];
      }
    }
  }
  if (bad) err[pidx != nullptr ? iom : 0] = 8;  // with problem indices: one flag per omega
  cd v[6];
  for (int k = 0; k < 6; ++k) {
    const cd r = block_sum<64>(acc[k], nullptr);
    if constexpr (!LANDAU) {
      v[k] = (r * S.rel_dg) * 0.25;
    } else if (C.om.im <= 0.0) {
      const cd Lf = (((timesI(r) * S.rel_dg) * PI_R) * 2.0) * PI_R *
                    ((S.q * C.vA / (C.kpar * S.m)) * ((S.m / C.vA) * (S.m / C.vA) * (S.m / C.vA)));
      v[k] = (C.om.im < 0.0 ? 2.0 : 1.0) * Lf;
    } else {
      v[k] = mk(0.0, 0.0);
    }
  }
  if (n == 0) v[0] = v[3] = v[4] = mk(0.0, 0.0);
  if (threadIdx.x == 0) {
    for (int k = 0; k < 6; ++k) o[(size_t)k * nom] = make_double2(v[k].re, v[k].im);
    if (lowR != nullptr && n >= -1 && n <= 1) {
      // the harmonic's whole term: the Landau launch adds its part to the
      // integrate_res_rel part the first launch wrote
      double2 *lo = lowR + ((size_t)iom * nspec + s) * 18 + 6 * (n + 1);
      for (int k = 0; k < 6; ++k) {
        const cd t = LANDAU ? mk(lo[k].x, lo[k].y) + v[k] : v[k];
        lo[k] = make_double2(t.re, t.im);
      }
    }
  }
}

// the integrate_res_rel part at three wavefronts per SIMD (a few spilled
// registers; two by its register count)
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_rel_res(
    const ProbDev *__restrict__ probp, int nom, const double2 *__restrict__ om_in,
    const int *__restrict__ active, double2 *__restrict__ chiR, double2 *__restrict__ lowR,
    const int *__restrict__ flags, int NB, int *__restrict__ err, const int *__restrict__ pidx,
    int nstride) {
  rel_res_body<false>(probp, nom, om_in, active, chiR, lowR, flags, NB, err, pidx, nstride);
}

__global__ __launch_bounds__(64) void k_rel_landau(
    const ProbDev *__restrict__ probp, int nom, const double2 *__restrict__ om_in,
    const int *__restrict__ active, double2 *__restrict__ chiR, double2 *__restrict__ lowR,
    const int *__restrict__ flags, int NB, int *__restrict__ err, const int *__restrict__ pidx,
    int nstride) {
  rel_res_body<true>(probp, nom, om_in, active, chiR, lowR, flags, NB, err, pidx, nstride);
}

}  // namespace

// partial slots of chiR: NB harmonic blocks of k_rel_nonres, then 2 nmaxR + 1
// of each k_rel_res launch
int rel_slots(int NB, int nmaxR) { return NB + 2 * (2 * nmaxR + 1); }

// nmaxR: the largest relativistic nmax of the batch (of any omega's problem
// with pidx); flags [nom][nspec][2 nmaxR + 1]; err [1], or [nom] with pidx
hipError_t launch_rel(hipStream_t st, const ProbDev *prob, int nspec, int nmaxR, int NB, int nom,
                      const double2 *om, const int *active, double2 *chiR, double2 *lowR,
                      int *flags, int *err, const int *pidx) {
  if (nom <= 0) return hipSuccess;
  const int nstride = 2 * nmaxR + 1;
  hipLaunchKernelGGL(k_rel_nonres, dim3(nom, nspec, NB), dim3(256), 0, st, prob, nom, om, active,
                     chiR, lowR, flags, pidx, nstride);
  hipLaunchKernelGGL(k_rel_res, dim3(nom, nspec, 2 * nmaxR + 1), dim3(64), 0, st, prob, nom, om,
                     active, chiR, lowR, flags, NB, err, pidx, nstride);
  hipLaunchKernelGGL(k_rel_landau, dim3(nom, nspec, 2 * nmaxR + 1), dim3(64), 0, st, prob, nom,
                     om, active, chiR, lowR, flags, NB, err, pidx, nstride);
  return hipGetLastError();
}
