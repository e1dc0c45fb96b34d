// Fused batched evaluation of disp(om) (ALPS_fns.f90:252-636).
//
// One workgroup is one 64-lane wavefront.  Its lanes form OMB = 64/NCH groups
// of NCH lanes; each group owns one omega and walks every harmonic row of
// every species, so a wavefront evaluates OMB omegas side by side:
//   * node sweep  (integrate, ALPS_fns.f90:799-864): the NCH lanes of a group
//     split the ppar nodes of the row; the +n and -n rows of a harmonic share
//     one pass over the pperp-reduced node table (alpsg_tables.hip);
//   * principal value (integrate_res, ALPS_fns.f90:870-1238): the ladder of
//     n_resonance_interval sub-steps is split over the group's lanes, with
//     funct_g's linear interpolation applied to the pperp-summed node values;
//   * Landau term (landau_integrate, ALPS_fns.f90:1327-1452): the pperp
//     points are split over the group's lanes; the analytic continuation
//     (eval_fit) is evaluated once per (omega, n) and shared by all 6 modes;
//   * group sums use cross-lane shuffles; the species sum, the dielectric
//     tensor, the wave tensor and det D are formed in registers.
// Every branch decision (resonance detection, ipar_res, ipar_close, ladder
// positions) is computed without FMA contraction in the reference's
// operation order, so the GPU takes the same branches as the CPU restatement.
#include "alpsg_device.h"

namespace {

constexpr double PI_D = 3.141592653589793238462643383279502884;

struct RowClass {
  int a1, b1, a2, b2;  // ppar index ranges of the trapezoid (integrate calls)
  int pv;              // principal-value part present
  int landau;          // Landau multiplier 0, 1, 2 (full_integrate, ALPS_fns.f90:782-790)
  int ipar_res, upper;
  double pr, pim;      // resonant momentum
};

// determine_resonances (ALPS_fns.f90:704-742) and the ipar_res / range logic of
// integrate_res (ALPS_fns.f90:950-1007) for one (n, omega).
__device__ __noinline__ void classify(const SpecDev &S, int P, double kpar, double omr,
                                      double omi, int n, RowClass &rc) {
#pragma clang fp contract(off)
  const double *pa = S.ppar;
  const int npar = S.npar;
  const double d = S.dppar;
  const double pr = (S.m * omr - (1.0 * n) * S.q) / kpar;
  rc.pr = pr;
  rc.pim = (S.m * omi) / kpar;
  rc.a2 = rc.b2 = -1000;
  rc.pv = 0;
  rc.landau = 0;
  rc.ipar_res = 0;
  rc.upper = 0;
  bool found = (pa[1] <= pr && pr < pa[npar]);
  const double pos = (double)P;
  if (pr < pa[1] && pr >= (pa[1] - pos * d)) found = true;
  if (pr >= pa[npar - 1] && pr < (pa[npar - 1] + pos * d)) found = true;
  if (!found) {
    rc.a1 = 1;
    rc.b1 = npar - 1;
    return;
  }
  int ipr = 0;
  if (pa[1] <= pr && pr < pa[npar - 1]) {
    int lo = 1, hi = npar - 2;  // largest i in [1, npar-2] with pa[i] <= pr
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (pa[mid] <= pr)
        lo = mid;
      else
        hi = mid - 1;
    }
    ipr = lo;
  }
  for (int ip = 0; ip <= P; ++ip) {
    if (pr >= (pa[0] - d * ip) && pr < (pa[0] - d * (ip - 1))) ipr = -ip;
    if (pr >= (pa[npar - 1] + d * ip) && pr < (pa[npar - 1] + d * (ip + 1))) ipr = npar - 1 + ip;
  }
  rc.ipar_res = ipr;
  rc.landau = (omi < 0.0) ? 2 : ((omi == 0.0) ? 1 : 0);
  if ((ipr - P) <= 2) {
    rc.a1 = ipr + P;
    rc.b1 = npar - 1;
    return;
  }
  if ((ipr + P) >= (npar - 2)) {
    rc.a1 = 1;
    rc.b1 = ipr - P;
    return;
  }
  rc.a1 = 1;
  rc.b1 = ipr - P;
  rc.upper = (fabs(pr - pa[ipr]) < (0.5 * d)) ? ipr + P + 1 : ipr + P + 2;
  rc.a2 = rc.upper;
  rc.b2 = npar - 1;
  rc.pv = 1;
}

// trapezoid weight of node i for integrate(a1,b1) + integrate(a2,b2)
__device__ __forceinline__ double node_weight(const RowClass &rc, int i) {
  int w = (i == rc.a1) + (i == rc.b1) + 2 * (i > rc.a1 && i < rc.b1) + (i == rc.a2) +
          (i == rc.b2) + 2 * (i > rc.a2 && i < rc.b2);
  return (double)w;
}

// funct_g's ipar_close (ALPS_fns.f90:1285-1294): the last ipar in [1,npar-1]
// with |ppar(ipar) - x| <= dppar/2, clamped to [2, npar-2].
__device__ __noinline__ int ipar_close(const SpecDev &S, double x) {
#pragma clang fp contract(off)
  const double *pa = S.ppar;
  const int npar = S.npar;
  const double half = 0.5 * S.dppar_abs;
  double t = (x - pa[1]) / S.dppar_abs;
  int k0;
  if (!(t > -4.0))
    k0 = -10;
  else if (!(t < (double)npar + 4.0))
    k0 = npar + 10;
  else
    k0 = 1 + (int)floor(t + 0.5);
  int best = 0;
  int khi = min(k0 + 2, npar - 1), klo = max(k0 - 2, 1);
  for (int k = khi; k >= klo; --k)
    if (fabs(pa[k] - x) <= half) {
      best = k;
      break;
    }
  if (best >= npar - 1) best = npar - 2;
  if (best <= 1) best = 2;
  return best;
}

// pperp-summed funct_g at real ppar x for the 6 node products
// [U1, p U1, p^2 U1, U2, U3, p U3] with U_a = om*G_a + H_a (no -q/kpar factor).
__device__ __forceinline__ void interp6(const SpecDev &S, const double *__restrict__ rows, double x,
                                        double omr, double omi, cd out[6]) {
  const int ic = ipar_close(S, x);
  cd u[3][6];
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    const double *r = rows + (size_t)(ic - 1 + t - 1) * NODE_STRIDE;
    double G1 = r[0], H1 = r[1], G2 = r[2], H2 = r[3], G3 = r[4], H3 = r[5], p = r[6];
    cd U1 = mk(omr * G1 + H1, omi * G1), U2 = mk(omr * G2 + H2, omi * G2),
       U3 = mk(omr * G3 + H3, omi * G3);
    u[t][0] = U1;
    u[t][1] = p * U1;
    u[t][2] = (p * p) * U1;
    u[t][3] = U2;
    u[t][4] = U3;
    u[t][5] = p * U3;
  }
  const double dx = x - S.ppar[ic];
  const double inv = 1.0 / S.dppar_abs;
#pragma unroll
  for (int k = 0; k < 6; ++k) out[k] = u[1][k] + (0.5 * ((u[2][k] - u[0][k]) * inv)) * dx;
}

__device__ __forceinline__ bool czero(cd a) { return a.re == 0.0 && a.im == 0.0; }

// eval_fit (ALPS_analyt.f90:32-363) at (iperp, complex p).
__device__ cd eval_fit(const SpecDev &S, int iperp, cd p) {
  if (S.acmethod == 2) {
    // fit_function_poly, Chebyshev kind 1 (ALPS_analyt.f90:307-358)
    cd t = mk((p.re - S.poly_norm1) / S.poly_norm2, p.im / S.poly_norm2);
    if (cabsd(t) > 1.0) return mk(0.0, 0.0);
    const double *c = S.polyc + (size_t)iperp * (S.poly_order + 1);
    cd bm2 = mk(1.0, 0.0), bm1 = t, f = c[0] * bm2;
    f = f + c[1] * bm1;
    for (int n = 2; n <= S.poly_order; ++n) {
      cd b = (2.0 * t) * bm1 - bm2;
      f = f + c[n] * b;
      bm2 = bm1;
      bm1 = b;
    }
    if (S.poly_logfit) {
      double lm = S.poly_log_max;
      if (f.re < -lm || f.im < -lm || f.re > lm || f.im > lm) return mk(0.0, 0.0);
      const double ln10 = 2.302585092994045684017991454684364208;
      return cexpd(mk(f.re * ln10, f.im * ln10));
    }
    return f;
  }
  cd f = mk(0.0, 0.0);
  for (int fi = 0; fi < S.nfits; ++fi) {
    const double *c = S.fitc + ((size_t)iperp * S.nfits + fi) * FIT_STRIDE;
    cd dq = mk(p.re - c[2], p.im);
    cd sq = dq * dq;
    switch (S.fit_type[fi]) {
      case 1:  // Maxwell: c0 = u1*exp(-pc*pperp^2), c1 = u2, c2 = u3
        f = f + c[0] * cexpd(mk(-c[1] * sq.re, -c[1] * sq.im));
        break;
      case 2: {  // kappa: c3 = u4, c4 = pc*u5*pperp^2
        cd kp = mk(1.0 + c[1] * sq.re + c[4], c[1] * sq.im);
        f = f + c[0] * cpowr(kp, c[3]);
      } break;
      case 6: {  // bi-Moyal: c3 = u4*pc*pperp^2
        cd arg = mk(c[3] + c[1] * sq.re, c[1] * sq.im);
        cd e = cexpd(arg);
        f = f + c[0] * cexpd(mk(0.5 * (arg.re - e.re), 0.5 * (arg.im - e.im)));
      } break;
      default:
        break;
    }
  }
  return f;
}

template <int NCH>
__device__ __forceinline__ double gsum(double v) {
#pragma unroll
  for (int o = NCH / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <int NCH>
__device__ __forceinline__ cd gsumc(cd v) {
  return mk(gsum<NCH>(v.re), gsum<NCH>(v.im));
}
template <int NCH>
__device__ __forceinline__ int gor(int v) {
#pragma unroll
  for (int o = NCH / 2; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
  return v;
}

// mode order k: 0 xx, 1 yy, 2 zz, 3 xy, 4 xz, 5 yz; node-product index per mode
__device__ constexpr int ACC_OF_MODE[6] = {0, 3, 2, 4, 1, 5};
// (i,j) of mode k (0-based)
__device__ constexpr int MI[6] = {0, 1, 2, 0, 0, 1};
__device__ constexpr int MJ[6] = {0, 1, 2, 1, 2, 2};

// T-tensor constant C_k(n) (int_T, ALPS_fns.f90:1659-1703), real and imaginary
// parts: modes xy and yz are purely imaginary.
__device__ __forceinline__ cd tconst(const ProbDev &Pd, const SpecDev &S, int n, int k) {
  const double kp = Pd.kperp, z = S.z;
  const bool kn = Pd.kperp_norm != 0;
  switch (k) {
    case 0:
      return mk(1.0 * (n * n) / (z * z), 0.0);
    case 1:
      return mk(kn ? 1.0 : kp * kp, 0.0);
    case 2:
      return mk(kn ? 1.0 : kp * kp, 0.0);
    case 3:
      return mk(0.0, kn ? (1.0 * n) / z : kp * (1.0 * n) / z);
    case 4:
      return mk(kn ? (1.0 * n) / z : (1.0 * n) * kp / z, 0.0);
    default:
      return mk(0.0, kn ? -1.0 : -(kp * kp));
  }
}

// Principal-value and tiny-rest contributions of one resonant row
// (integrate_res after the two integrate() calls), partial sums of this lane.
template <int NCH>
__device__ void pv_part(const SpecDev &S, const ProbDev &Pd, const double *__restrict__ rows,
                        const RowClass &rc, double omr, double omi, int c, cd pv[6]) {
  const int NRI = Pd.NRI, P = Pd.P;
  const double pr = rc.pr, pim = rc.pim;
  double capDelta, smdelta;
  {
#pragma clang fp contract(off)
    capDelta = pr - S.ppar[rc.ipar_res - P];
    smdelta = capDelta / (1.0 * NRI);
  }
  cd a[6], b[6];
  if (fabs(pim) > Pd.Tlim) {
    for (int j = c; j <= NRI; j += NCH) {
      double x, xm;
      {
#pragma clang fp contract(off)
        x = (j == 0) ? pr : ((j == NRI) ? pr + capDelta : pr + smdelta * j);
        xm = 2.0 * pr - x;
      }
      const double w = (j == 0 || j == NRI) ? 1.0 : 2.0;
      cd r1 = cscale_inv(w, mk(x - pr, -pim));
      cd r2 = conjc(r1);
      interp6(S, rows, x, omr, omi, a);
      interp6(S, rows, xm, omr, omi, b);
#pragma unroll
      for (int k = 0; k < 6; ++k) pv[k] = pv[k] + (a[k] * r1 - b[k] * r2);
    }
  } else {
    double sphi = 0.0;
    for (int j = 1 + c; j <= NRI; j += NCH) {
      double x;
      {
#pragma clang fp contract(off)
        x = (j == NRI) ? pr + capDelta : pr + smdelta * j;
      }
      const double t = x - pr;
      const double w = (j == NRI) ? 1.0 : 2.0;
      sphi += w * ((t * t) / ((t * t) + pim * pim));
    }
    sphi = gsum<NCH>(sphi);
    if (c == 0) {
      double xp, xq;
      {
#pragma clang fp contract(off)
        xp = pr + S.dppar;
        xq = pr - S.dppar;
      }
      interp6(S, rows, xp, omr, omi, a);
      interp6(S, rows, xq, omr, omi, b);
      const double inv2d = 1.0 / (2.0 * S.dppar);
#pragma unroll
      for (int k = 0; k < 6; ++k) pv[k] = pv[k] + (2.0 * sphi) * ((a[k] - b[k]) * inv2d);
      if (pim != 0.0) {
        interp6(S, rows, pr, omr, omi, a);
        const double sg = (pim > 0.0) ? 1.0 : -1.0;
        const double f = sg * 2.0 * PI_D / smdelta;
#pragma unroll
        for (int k = 0; k < 6; ++k) pv[k] = pv[k] + timesI(f * a[k]);
      }
    }
  }
  // tiny rest (ALPS_fns.f90:1169-1230)
  int ntiny;
  double corr = 0.0, x0 = 0.0, cs = 0.0;
  {
#pragma clang fp contract(off)
    const double rest = S.ppar[rc.upper] - pr - capDelta;
    ntiny = (int)(rest / smdelta);
    if (ntiny > 0) {
      corr = (rest / (1.0 * ntiny)) / smdelta;
      x0 = pr + capDelta;
      cs = corr * smdelta;
    }
  }
  for (int kk = c; kk <= ntiny && ntiny > 0; kk += NCH) {
    double x;
    {
#pragma clang fp contract(off)
      x = (kk == 0) ? x0 : x0 + cs * kk;
    }
    const double w = (kk == 0 || kk == ntiny) ? 1.0 : 2.0;
    cd r1 = cscale_inv(corr * w, mk(x - pr, -pim));
    interp6(S, rows, x, omr, omi, a);
#pragma unroll
    for (int k = 0; k < 6; ++k) pv[k] = pv[k] + a[k] * r1;
  }
}

// Landau sums L_a = sum_iperp h a(n,iperp) K(iperp) of this lane (a = J^2,
// pperp^2 J'^2, pperp J J') and the early-return flag of landau_integrate.
template <int NCH>
__device__ void landau_part(const SpecDev &S, const ProbDev &Pd, const double *__restrict__ arow,
                            const RowClass &rc, double omr, double omi, int c, cd L[3],
                            int &zero) {
  const int nperp = S.nperp;
  const double d = S.dppar_abs, dpp = S.dpperp;
  const cd p = mk(rc.pr, rc.pim);
  const cd pp_ = mk(rc.pr + d, rc.pim), pm_ = mk(rc.pr - d, rc.pim);
  const double qk = S.q / fabs(Pd.kpar), km = Pd.kpar / S.m;
  for (int ip = c; ip <= nperp; ip += NCH) {
    double h;
    cd dfperp, dfpar;
    if (ip >= 1 && ip <= nperp - 1) {
      h = (ip == nperp - 1) ? 0.5 : 1.0;
      cd fpi = eval_fit(S, ip, pp_), fpf = eval_fit(S, ip, pm_);
      cd fperpi = eval_fit(S, ip + 1, p), fperpf = eval_fit(S, ip - 1, p);
      if (czero(fpi) || czero(fpf) || czero(fperpi)) zero = 1;
      dfperp = (fperpi - fperpf) * (1.0 / (2.0 * dpp));
      dfpar = (fpi - fpf) * (1.0 / (2.0 * d));
    } else if (ip == 0) {
      h = 0.5;
      dfperp = (eval_fit(S, 1, p) - eval_fit(S, 0, p)) * (1.0 / dpp);
      dfpar = (eval_fit(S, 0, pp_) - eval_fit(S, 0, pm_)) * (1.0 / (2.0 * d));
    } else {
      h = 0.5;
      dfperp = (eval_fit(S, nperp, p) - eval_fit(S, nperp - 1, p)) * (1.0 / dpp);
      dfpar = (eval_fit(S, nperp, pp_) - eval_fit(S, nperp, pm_)) * (1.0 / (2.0 * d));
    }
    cd K = qk * ((S.pperp[ip] * dfpar - p * dfperp) * km + mk(omr, omi) * dfperp);
    const double *a = arow + (size_t)ip * ATAB_STRIDE;
    L[0] = L[0] + (h * a[0]) * K;
    L[1] = L[1] + (h * a[1]) * K;
    L[2] = L[2] + (h * a[2]) * K;
  }
}

}  // namespace

// chi/chi_low/wave are written only when the pointers are non-null.
template <int NCH>
__global__ __launch_bounds__(64) void k_disp(const ProbDev *__restrict__ probp, int nom,
                                             const double2 *__restrict__ om_in,
                                             const int *__restrict__ active,
                                             double2 *__restrict__ det_out,
                                             double2 *__restrict__ chi_out,
                                             double2 *__restrict__ low_out,
                                             double2 *__restrict__ wave_out) {
  constexpr int OMB = 64 / NCH;
  const ProbDev &Pd = *probp;
  const int lane = threadIdx.x;
  const int g = lane / NCH, c = lane % NCH;
  const int iom = blockIdx.x * OMB + g;
  const bool valid = (iom < nom) && (active == nullptr || active[iom] != 0);
  const double2 omv = valid ? om_in[iom] : make_double2(1.0, 0.5);
  const double omr = omv.x, omi = omv.y;
  const int nspec = Pd.nspec;
  const bool want_low = (low_out != nullptr) && valid && c == 0;

  cd eps[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) eps[k] = mk(0.0, 0.0);

  for (int s = 0; s < nspec; ++s) {
    const SpecDev &S = Pd.sp[s];
    const int npar = S.npar, nmax = S.nmax;
    const double scaleN = 2.0 * PI_D * S.dpperp * S.dppar_abs * 0.25 * S.q;
    cd chi[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) chi[k] = mk(0.0, 0.0);

    for (int nn = 0; nn <= nmax; ++nn) {
      const double *rows = S.nodes + (size_t)nn * (npar - 1) * NODE_STRIDE;
      RowClass rp, rm;
      classify(S, Pd.P, Pd.kpar, omr, omi, nn, rp);
      if (nn > 0) {
        classify(S, Pd.P, Pd.kpar, omr, omi, -nn, rm);
      } else {
        rm.a1 = rm.b1 = rm.a2 = rm.b2 = -1000;
        rm.pv = 0;
        rm.landau = 0;
      }
      const double nq = (1.0 * nn) * S.q;
      cd ap[6], am[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) ap[k] = am[k] = mk(0.0, 0.0);

      // ---- node sweep: integrate() over the ppar nodes of this row pair ----
      for (int i = 1 + c; i <= npar - 1; i += NCH) {
        const double *r = rows + (size_t)(i - 1) * NODE_STRIDE;
        const double G1 = r[0], H1 = r[1], G2 = r[2], H2 = r[3], G3 = r[4], H3 = r[5],
                     pa = r[6];
        const cd n1 = mk(omr * G1 + H1, omi * G1), n2 = mk(omr * G2 + H2, omi * G2),
                 n3 = mk(omr * G3 + H3, omi * G3);
        const double base = S.m * omr - Pd.kpar * pa;
        const double dimag = S.m * omi;
        const double p2 = pa * pa;
        const double wp = node_weight(rp, i);
        if (wp != 0.0) {
          cd rr = cscale_inv(wp, mk(base - nq, dimag));
          cd u1 = n1 * rr, u2 = n2 * rr, u3 = n3 * rr;
          ap[0] = ap[0] + u1;
          ap[1] = ap[1] + pa * u1;
          ap[2] = ap[2] + p2 * u1;
          ap[3] = ap[3] + u2;
          ap[4] = ap[4] + u3;
          ap[5] = ap[5] + pa * u3;
        }
        const double wm = node_weight(rm, i);
        if (wm != 0.0) {
          cd rr = cscale_inv(wm, mk(base + nq, dimag));
          cd u1 = n1 * rr, u2 = n2 * rr, u3 = n3 * rr;
          am[0] = am[0] + u1;
          am[1] = am[1] + pa * u1;
          am[2] = am[2] + p2 * u1;
          am[3] = am[3] + u2;
          am[4] = am[4] + u3;
          am[5] = am[5] + pa * u3;
        }
      }
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        ap[k] = gsumc<NCH>(ap[k]);
        am[k] = gsumc<NCH>(am[k]);
      }

      // ---- resonant extras for the + and - rows ----
      cd rowv[2][6];
      for (int sg = 0; sg < 2; ++sg) {
        const RowClass &rc = sg == 0 ? rp : rm;
        const int n = sg == 0 ? nn : -nn;
        cd *acc = sg == 0 ? ap : am;
        cd Ck[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          Ck[k] = tconst(Pd, S, n, k);
          rowv[sg][k] = (scaleN * acc[ACC_OF_MODE[k]]) * Ck[k];
        }
        if (sg == 1 && nn == 0) continue;
        if (rc.pv) {
          cd pv[6];
#pragma unroll
          for (int k = 0; k < 6; ++k) pv[k] = mk(0.0, 0.0);
          pv_part<NCH>(S, Pd, rows, rc, omr, omi, c, pv);
          double capDelta, smdelta;
          {
#pragma clang fp contract(off)
            capDelta = rc.pr - S.ppar[rc.ipar_res - Pd.P];
            smdelta = capDelta / (1.0 * Pd.NRI);
          }
          const double scalePV = 2.0 * PI_D * smdelta * S.dpperp * 0.25 * (-S.q / Pd.kpar);
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            cd v = gsumc<NCH>(pv[ACC_OF_MODE[k]]);
            rowv[sg][k] = rowv[sg][k] + (scalePV * v) * Ck[k];
          }
        }
        if (rc.landau) {
          cd L[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
          int zero = 0;
          landau_part<NCH>(S, Pd, S.atab + (size_t)nn * (S.nperp + 1) * ATAB_STRIDE, rc, omr, omi,
                           c, L, zero);
          zero = gor<NCH>(zero);
          if (!zero) {
            const cd p = mk(rc.pr, rc.pim);
            cd Ls[3] = {gsumc<NCH>(L[0]), gsumc<NCH>(L[1]), gsumc<NCH>(L[2])};
            // T_res (int_T_res, ALPS_fns.f90:1768-1808): zz, xz, yz take p_res
            cd Lm[6] = {Ls[0] * Ck[0], Ls[1] * Ck[1], (Ls[0] * (p * p)) * Ck[2],
                        Ls[2] * Ck[3], (Ls[0] * p) * Ck[4], (Ls[2] * p) * Ck[5]};
            const double f = -(double)rc.landau * S.dpperp * PI_D * 2.0 * PI_D;
#pragma unroll
            for (int k = 0; k < 6; ++k) rowv[sg][k] = rowv[sg][k] + timesI(f * Lm[k]);
          }
        }
      }

      // ---- species accumulation in the reference's order ----
      if (nn == 0) {
        chi[1] = chi[1] + rowv[0][1];
        chi[2] = chi[2] + rowv[0][2];
        chi[5] = chi[5] + rowv[0][5];
      } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) chi[k] = (chi[k] + rowv[0][k]) + rowv[1][k];
      }
      if (want_low && nn <= 1) {
        // chi_low(s,i,j,-1:1) in Fortran order: s + nspec*(i + 3*(j + 3*q))
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          if (nn == 0 && (k == 0 || k == 3 || k == 4)) continue;
          const int i = MI[k], j = MJ[k];
          double2 *o = low_out + (size_t)iom * nspec * 27;
          if (nn == 0) {
            o[s + nspec * (i + 3 * (j + 3 * 1))] = make_double2(rowv[0][k].re, rowv[0][k].im);
          } else {
            o[s + nspec * (i + 3 * (j + 3 * 2))] = make_double2(rowv[0][k].re, rowv[0][k].im);
            o[s + nspec * (i + 3 * (j + 3 * 0))] = make_double2(rowv[1][k].re, rowv[1][k].im);
          }
        }
      }
    }

    // ---- ee term, density*charge normalisation (ALPS_fns.f90:479-514) ----
    const double ee = Pd.kperp_norm ? S.int_ee : Pd.kperp * Pd.kperp * S.int_ee;
    chi[2] = chi[2] + mk(ee, 0.0);
    const double norm = S.dens * S.q;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      chi[k] = chi[k] * norm;
      eps[k] = eps[k] + chi[k];
    }
    const cd om2 = mk(omr, omi) * mk(omr, omi);
    cd dn = om2 * (Pd.vA * Pd.vA);
    if (!Pd.kperp_norm) dn = dn * (Pd.kperp * Pd.kperp);
    if (chi_out != nullptr && valid && c == 0) {
      double2 *o = chi_out + (size_t)iom * nspec * 9;
      cd v[3][3];
#pragma unroll
      for (int k = 0; k < 6; ++k) v[MI[k]][MJ[k]] = cdiv(chi[k], dn);
      v[1][0] = mk(-v[0][1].re, -v[0][1].im);
      v[2][0] = v[0][2];
      v[2][1] = mk(-v[1][2].re, -v[1][2].im);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o[s + nspec * (i + 3 * j)] = make_double2(v[i][j].re, v[i][j].im);
    }
    if (want_low) {
      double2 *o = low_out + (size_t)iom * nspec * 27;
      // add the ee term to the n = 0 zz entry, normalise, symmetrise
      {
        double2 t = o[s + nspec * (2 + 3 * (2 + 3 * 1))];
        o[s + nspec * (2 + 3 * (2 + 3 * 1))] = make_double2(t.x + ee, t.y);
      }
      for (int q = 0; q < 3; ++q) {
        cd v[3][3];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) v[i][j] = mk(0.0, 0.0);
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          if (q == 1 && (k == 0 || k == 3 || k == 4)) continue;
          double2 t = o[s + nspec * (MI[k] + 3 * (MJ[k] + 3 * q))];
          v[MI[k]][MJ[k]] = cdiv(mk(t.x, t.y) * norm, dn);
        }
        v[1][0] = mk(-v[0][1].re, -v[0][1].im);
        v[2][0] = v[0][2];
        v[2][1] = mk(-v[1][2].re, -v[1][2].im);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            o[s + nspec * (i + 3 * (j + 3 * q))] = make_double2(v[i][j].re, v[i][j].im);
      }
    }
  }

  if (!valid || c != 0) return;
  // ---- dielectric tensor, wave tensor, determinant (ALPS_fns.f90:566-624) ----
  const cd om = mk(omr, omi);
  cd unit = Pd.kperp_norm ? (om * Pd.vA) * (om * Pd.vA)
                          : ((Pd.kperp * om) * Pd.vA) * ((Pd.kperp * om) * Pd.vA);
  double enx2, enz2, enxnz;
  if (Pd.kperp_norm) {
    enx2 = Pd.kperp * Pd.kperp;
    enz2 = Pd.kpar * Pd.kpar;
    enxnz = Pd.kpar * Pd.kperp;
  } else {
    const double k2 = Pd.kperp * Pd.kperp;
    enx2 = k2 * k2;
    enz2 = (Pd.kpar * Pd.kpar) * k2;
    enxnz = Pd.kpar * (k2 * Pd.kperp);
  }
  cd w11 = (eps[0] + unit) - mk(enz2, 0.0);
  cd w22 = ((eps[1] + unit) - mk(enz2, 0.0)) - mk(enx2, 0.0);
  cd w33 = (eps[2] + unit) - mk(enx2, 0.0);
  cd w13 = eps[4] + mk(enxnz, 0.0);
  cd w12 = eps[3];
  cd w23 = eps[5];
  cd det = w11 * (w22 * w33 + w23 * w23) + 2.0 * w12 * w23 * w13 - (w13 * w13) * w22 +
           (w12 * w12) * w33;
  det_out[iom] = make_double2(det.re, det.im);
  if (wave_out != nullptr) {
    double2 *o = wave_out + (size_t)iom * 9;
    cd W[3][3] = {{w11, w12, w13},
                  {mk(-w12.re, -w12.im), w22, w23},
                  {w13, mk(-w23.re, -w23.im), w33}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) o[i + 3 * j] = make_double2(W[i][j].re, W[i][j].im);
  }
}

hipError_t launch_disp(hipStream_t st, int nch, const ProbDev *prob, int nom, const double2 *om,
                       const int *active, double2 *det, double2 *chi, double2 *low,
                       double2 *wave) {
  if (nom <= 0) return hipSuccess;
  switch (nch) {
#define CASE(N)                                                                              \
  case N: {                                                                                  \
    int grid = (nom + (64 / N) - 1) / (64 / N);                                              \
    hipLaunchKernelGGL(k_disp<N>, dim3(grid), dim3(64), 0, st, prob, nom, om, active, det, chi, \
                       low, wave);                                                           \
  } break;
    CASE(1)
    CASE(2)
    CASE(4)
    CASE(8)
    CASE(16)
    CASE(32)
    CASE(64)
#undef CASE
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}
