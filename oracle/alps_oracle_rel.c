/*
 * alps_oracle_rel.c -- CPU restatement of ALPS's relativistic susceptibility
 * path (relativistic = T species), ALPS_fns_rel.f90:460-1791 and the
 * relativistic branch of determine_resonances, ALPS_fns.f90:680-701.
 *
 * TEST INFRASTRUCTURE ONLY: part of the checker for the HIP path (see
 * alps_oracle.c).
 *
 * The resonant harmonics of a relativistic species are integrated on the
 * (Gamma, pparbar) grid row by row along the resonance curve
 * pparbar_res = (Gamma om - n q/m) vA/kpar; the non-resonant ones use
 * integrate() on the (pperp, ppar) grid with Gamma in resU (alps_oracle.c).
 * Loops and expressions follow the reference in order.  The inputs
 * gamma_rel, pparbar_rel, f0_rel and df0_rel are taken as the reference's
 * set-up (derivative_f0_rel, ALPS_fns_rel.f90:19-408) leaves them.
 *
 * Parity pins (tests/test_oracle_rel.py): CBESSJ against scipy's jv at complex
 * argument, the Gamma function against math.gamma, int_ee_rel and the whole
 * relativistic chi of a cool Maxwellian against the non-relativistic closed
 * form as vA -> 0.
 */
#include <math.h>
#include <string.h>

#include "alps_oracle.h"

static const double PI_R = 3.141592653589793238462643383279502884;

#define COUNT(b)            \
  do {                      \
    _Pragma("omp atomic")  \
    orc_branch_counts[b]++; \
  } while (0)

#define GR(P, s, ig, ip) ((P)->gamma_rel[(s)-1][(size_t)(ig) * ((P)->npparbar + 1) + (ip)])
#define PB(P, s, ig, ip) ((P)->pparbar_rel[(s)-1][(size_t)(ig) * ((P)->npparbar + 1) + (ip)])
#define F0R(P, s, ig, ip) ((P)->f0_rel[(s)-1][(size_t)(ig) * ((P)->npparbar + 1) + (ip)])
#define DF0R(P, s, ig, ip, k) \
  ((P)->df0_rel[(s)-1][((size_t)(ig) * ((P)->npparbar + 1) + (ip)) * 2 + ((k)-1)])
/* pp(1:nspec,0:nperp,0:npar,1:2) */
#define PPR(P, is, iperp, ipar, k) \
  ((P)->pp[((((size_t)((is)-1) * ((P)->nperp + 1) + (iperp)) * ((P)->npar + 1) + (ipar)) * 2) + ((k)-1)])

static inline double cube(double x) { return x * x * x; }

/* complex ** integer by repeated squaring (the Fortran runtime's pow_c8_i4) */
static double complex cpowi(double complex a, int n) {
  double complex pw = 1.0, x = a;
  unsigned u = (unsigned)n;
  if (n == 0) return 1.0;
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

/* Fact, ALPS_fns_rel.f90:1557-1574 */
static double fact(int k) {
  double f = 1.0;
  for (int i = 2; i <= k; ++i) f = f * (1.0 * i);
  return f;
}

/* Gamma, ALPS_fns_rel.f90:1738-1791 (Lanczos series of Numerical Recipes' gammln) */
double orc_gamma_fn(double xx) {
  const double ONE = 1.0, FPF = 5.5, HALF = 0.5;
  const double cof[6] = {76.18009173, -86.50532033, 24.01409822, -1.231739516, 0.120858003e-2,
                         -0.536382e-5};
  const double stp = 2.50662827465;
  double x = xx - ONE, tmp = x + FPF, ser = ONE;
  tmp = (x + HALF) * log(tmp) - tmp;
  for (int j = 0; j < 6; ++j) {
    x = x + ONE;
    ser = ser + cof[j] / x;
  }
  return exp(tmp + log(stp * ser));
}

/* CBESSJ, ALPS_fns_rel.f90:1502-1553: J_nu(z) by 21 terms of the power series */
double complex orc_cbessj(double complex z, int nu) {
  const int MAXK = 20;
  double complex sum = 0.0, tmp;
  for (int k = 0; k <= MAXK; ++k) {
    tmp = cpowi(-z * z / 4.0, k);
    tmp = tmp / fact(k);
    tmp = tmp / orc_gamma_fn(1.0 * (nu + k + 1));
    sum = sum + tmp;
  }
  tmp = cpowi(z / 2.0, nu);
  return tmp * sum;
}

static inline double m1p(int n) { return (n % 2 == 0) ? 1.0 : -1.0; }

/* int_T_rel, ALPS_fns_rel.f90:1256-1369 */
static double complex int_T_rel(const orc_plasma *P, int s, int nn, int ig, int ip, int mode) {
  const double complex ii = I;
  const double kperp = P->kperp, ms = P->ms[s - 1], qs = P->qs[s - 1], vA = P->vA;
  const double g = GR(P, s, ig, ip), pb = PB(P, s, ig, ip);
  const double pperpbar = sqrt(g * g - 1.0 - pb * pb);
  const double z = (kperp * ms / (vA * qs)) * pperpbar;
  const double zbar = P->kperp_norm ? kperp * ms / (vA * qs) : ms / (vA * qs);
  double bessel, besselP;
  if (nn < 0)
    bessel = m1p(nn) * orc_bessj(-nn, z);
  else
    bessel = orc_bessj(nn, z);
  if (nn >= 1)
    besselP = 0.5 * (orc_bessj(nn - 1, z) - orc_bessj(nn + 1, z));
  else if (nn < -1)
    besselP = 0.5 * ((m1p(nn - 1) * orc_bessj(-(nn - 1), z)) - (m1p(nn + 1) * orc_bessj(-(nn + 1), z)));
  else if (nn == 0)
    besselP = -orc_bessj(1, z);
  else
    besselP = 0.5 * (orc_bessj(2, z) - orc_bessj(0, z));
  const int kn = P->kperp_norm;
  switch (mode) {
    case 1:
      return 1.0 * (nn * nn) * bessel * bessel / (zbar * zbar);
    case 2:
      return kn ? besselP * besselP * pperpbar * pperpbar
                : kperp * kperp * besselP * besselP * pperpbar * pperpbar;
    case 3:
      return kn ? bessel * bessel * (pb * pb) : kperp * kperp * bessel * bessel * (pb * pb);
    case 4:
      return kn ? ii * (1.0 * nn) * bessel * besselP * pperpbar / zbar
                : ii * (1.0 * nn) * kperp * bessel * besselP * pperpbar / zbar;
    case 5:
      return kn ? (1.0 * nn) * bessel * bessel * pb / zbar
                : (1.0 * nn) * kperp * bessel * bessel * pb / zbar;
    case 6:
      return kn ? (-1.0 * ii) * bessel * besselP * pb * pperpbar
                : (-1.0 * ii) * kperp * kperp * bessel * besselP * pb * pperpbar;
  }
  return 0.0;
}

/* int_T_res_rel, ALPS_fns_rel.f90:1374-1495 (T at complex pparbar) */
static double complex int_T_res_rel(const orc_plasma *P, int s, int nn, int ig, double complex pb,
                                    int mode) {
  const double complex ii = I;
  const double kperp = P->kperp, ms = P->ms[s - 1], qs = P->qs[s - 1], vA = P->vA;
  const double g = GR(P, s, ig, 1);
  const double complex pperpbar = csqrt(g * g - 1.0 - cpowi(pb, 2));
  const double complex z = (kperp * ms / (vA * qs)) * pperpbar;
  const double zbar = P->kperp_norm ? kperp * ms / (vA * qs) : ms / (vA * qs);
  double complex bessel, besselP, besselH;
  if (nn < 0) {
    bessel = orc_cbessj(z, -nn);
    bessel = bessel * m1p(nn);
  } else {
    bessel = orc_cbessj(z, nn);
  }
  if (nn >= 1) {
    besselP = orc_cbessj(z, nn - 1);
    besselH = orc_cbessj(z, nn + 1);
    besselP = 0.5 * (besselP - besselH);
  } else if (nn < -1) {
    besselP = orc_cbessj(z, -(nn - 1));
    besselH = orc_cbessj(z, -(nn + 1));
    besselP = 0.5 * ((m1p(nn - 1) * besselP) - (m1p(nn + 1) * besselH));
  } else if (nn == 0) {
    besselP = orc_cbessj(z, 1);
    besselP = -besselP;
  } else {
    besselP = orc_cbessj(z, 2);
    besselH = orc_cbessj(z, 0);
    besselP = 0.5 * (besselP - besselH);
  }
  const int kn = P->kperp_norm;
  switch (mode) {
    case 1:
      return 1.0 * (nn * nn) * bessel * bessel / (zbar * zbar);
    case 2:
      return kn ? besselP * besselP * pperpbar * pperpbar
                : kperp * kperp * besselP * besselP * pperpbar * pperpbar;
    case 3:
      return kn ? bessel * bessel * cpowi(pb, 2) : kperp * kperp * bessel * bessel * cpowi(pb, 2);
    case 4:
      return kn ? ii * (1.0 * nn) * bessel * besselP * pperpbar / zbar
                : ii * (1.0 * nn) * kperp * bessel * besselP * pperpbar / zbar;
    case 5:
      return kn ? (1.0 * nn) * bessel * bessel * pb / zbar
                : (1.0 * nn) * kperp * bessel * bessel * pb / zbar;
    case 6:
      return kn ? (-1.0 * ii) * bessel * besselP * pb * pperpbar
                : (-1.0 * ii) * kperp * kperp * bessel * besselP * pb * pperpbar;
  }
  return 0.0;
}

/* the common prefactor -2 pi (m/vA)^3 (q vA/(kpar m)) of resU_rel and funct_g_rel */
static inline double rel_pref(const orc_plasma *P, int s) {
  const double ms = P->ms[s - 1], qs = P->qs[s - 1], vA = P->vA, kpar = P->kpar;
  return ((2.0 * PI_R) * cube(ms / vA)) * (qs * vA / (kpar * ms));
}

/* resU_rel, ALPS_fns_rel.f90:1220-1249 */
static double complex resU_rel(const orc_plasma *P, int s, double complex om, int nn, int ig, int ip) {
  const double ms = P->ms[s - 1], qs = P->qs[s - 1], vA = P->vA, kpar = P->kpar;
  const double complex num = om * DF0R(P, s, ig, ip, 1) + (kpar / (vA)) * DF0R(P, s, ig, ip, 2);
  const double complex den =
      (PB(P, s, ig, ip) - GR(P, s, ig, ip) * om * vA / kpar) + (1.0 * nn) * qs * vA / (kpar * ms);
  return -((rel_pref(P, s) * num) / den);
}

/* funct_g_rel, ALPS_fns_rel.f90:918-999.  The search leaves ipparbar_close = -2
   when pparbar is off the row; the clamp below then sets it to 2 whatever
   f0_rel(-1) / f0_rel(-3) hold, so those out-of-range reads are skipped. */
static double complex funct_g_rel(const orc_plasma *P, int s, double x, int ig, double complex om,
                                  int nn, int mode) {
  const int npb = P->npparbar;
  const double vA = P->vA, kpar = P->kpar;
  const double dpb = PB(P, s, 2, 2) - PB(P, s, 2, 1);
  int close = -2;
  for (int ip = 0; ip <= npb - 1; ++ip)
    if (PB(P, s, ig, ip + 1) > x && PB(P, s, ig, ip) <= x) close = ip;
  if (close + 1 >= 0 && close + 1 <= npb && F0R(P, s, ig, close + 1) <= -1.0) close = close - 1;
  if (close - 1 >= 0 && close - 1 <= npb && F0R(P, s, ig, close - 1) <= -1.0) close = close + 1;
  if (x == PB(P, s, ig, npb)) close = npb - 2;
  if (close >= (npb - 1)) close = npb - 2;
  if (close <= 1) close = 2;
#define GI(ip)                                                                                 \
  (-((rel_pref(P, s) * (om * DF0R(P, s, ig, (ip), 1) + (kpar / (vA)) * DF0R(P, s, ig, (ip), 2))) * \
     int_T_rel(P, s, nn, ig, (ip), mode)))
  const double complex plus = GI(close + 1);
  const double complex mid = GI(close);
  const double complex minus = GI(close - 1);
#undef GI
  return mid + 0.5 * ((plus - minus) / dpb) * (x - PB(P, s, ig, close));
}

/* principal_integral_rel, ALPS_fns_rel.f90:724-913 */
static double complex principal_integral_rel(const orc_plasma *P, int s, double complex om, int nn,
                                             int mode, int ig, int ipres, int upperlimit) {
  const double complex ii = I;
  const double ms = P->ms[s - 1], qs = P->qs[s - 1], vA = P->vA, kpar = P->kpar;
  const int pos = P->positions_principal, nri = P->n_resonance_interval;
  const double dpb = PB(P, s, 2, 2) - PB(P, s, 2, 1);
  double complex res = 0.0;
  const double complex gom = GR(P, s, ig, ipres) * om * vA / kpar;
  const double denomR = creal(gom - (1.0 * nn) * (qs / ms) * vA / kpar);
  const double denomI = cimag(gom);
  const double complex pres = denomR + denomI * ii;
  const double capDelta = creal(pres) - PB(P, s, 1, ipres - pos);
  const double smdelta = capDelta / (1.0 * nri);
  double x;
#define G(v) funct_g_rel(P, s, (v), ig, om, nn, mode)
  if (fabs(denomI) > P->Tlim) {
    COUNT(ORC_BR_REL_PV_TLIM);
    x = creal(pres);
    res = res + 1.0 * G(x) / (x - denomR - ii * denomI);
    res = res - 1.0 * G(2.0 * denomR - x) / (x - denomR + ii * denomI);
    x = creal(pres) + capDelta;
    res = res + 1.0 * G(x) / (x - denomR - ii * denomI);
    res = res - 1.0 * G(2.0 * denomR - x) / (x - denomR + ii * denomI);
    for (int ip = 1; ip <= nri - 1; ++ip) {
      x = creal(pres) + smdelta * ip;
      res = res + 2.0 * G(x) / (x - denomR - ii * denomI);
      res = res - 2.0 * G(2.0 * denomR - x) / (x - denomR + ii * denomI);
    }
  } else {
    COUNT(ORC_BR_REL_PV_ANALYTIC);
    const double complex gprimetr = (G(denomR + dpb) - G(denomR - dpb)) / (2.0 * dpb);
    x = creal(pres) + capDelta;
    res = res + 2.0 * gprimetr * ((x - denomR) * (x - denomR)) /
                    ((x - denomR) * (x - denomR) + denomI * denomI);
    for (int ip = 1; ip <= nri - 1; ++ip) {
      x = creal(pres) + smdelta * ip;
      res = res + 2.0 * 2.0 * gprimetr * ((x - denomR) * (x - denomR)) /
                      ((x - denomR) * (x - denomR) + denomI * denomI);
    }
    if (denomI > 0.0)
      res = res + 2.0 * ii * PI_R * G(denomR) / smdelta;
    else if (denomI < 0.0)
      res = res - 2.0 * ii * PI_R * G(denomR) / smdelta;
  }
  const int ntiny = (int)((PB(P, s, ig, upperlimit) - creal(pres) - capDelta) / smdelta);
  if (ntiny > 0) {
    COUNT(ORC_BR_REL_TINY);
    const double correction =
        ((PB(P, s, ig, upperlimit) - creal(pres) - capDelta) / (1.0 * ntiny)) / smdelta;
    x = creal(pres) + capDelta;
    res = res + 1.0 * correction * (G(x) / (x - denomR - ii * denomI));
    x = creal(pres) + capDelta + correction * smdelta * ntiny;
    res = res + 1.0 * correction * (G(x) / (x - denomR - ii * denomI));
    for (int ip = 1; ip <= ntiny - 1; ++ip) {
      x = creal(pres) + capDelta + correction * smdelta * ip;
      res = res + 2.0 * correction * (G(x) / (x - denomR - ii * denomI));
    }
  }
#undef G
  return res * smdelta;
}

/* the subluminal limits of row ig: the first index entering f0_rel > -1 and the
   last one before leaving it (ALPS_fns_rel.f90:593-606, 1128-1141) */
static void row_limits(const orc_plasma *P, int s, int ig, int *lower, int *upper) {
  const int npb = P->npparbar;
  int fl = 0, fu = 0;
  *lower = 1;
  *upper = npb - 1;
  for (int ip = 1; ip <= npb - 1; ++ip) {
    if (!fl && F0R(P, s, ig, ip - 1) <= -1.0 && F0R(P, s, ig, ip) > -1.0) {
      *lower = ip;
      fl = 1;
    }
    if (!fu && F0R(P, s, ig, ip) > -1.0 && F0R(P, s, ig, ip + 1) <= -1.0) {
      *upper = ip;
      fu = 1;
    }
  }
}

/* integrate_resU_rel, ALPS_fns_rel.f90:516-719.  *err = 8 where the reference
   calls alps_error(8). */
static double complex integrate_resU_rel(const orc_plasma *P, int s, double complex om, int nn,
                                         int mode, int ig, int *err) {
  const int npb = P->npparbar, pos = P->positions_principal;
  const double ms = P->ms[s - 1], qs = P->qs[s - 1], vA = P->vA, kpar = P->kpar;
  const double dpb = PB(P, s, 2, 2) - PB(P, s, 2, 1);
  double complex res = 0.0;
  int ip = 0, ipres = 0, found = 0;
  const double g1 = GR(P, s, ig, 1);
  const double complex pres = (g1 * om - (1.0 * nn) * qs / ms) * vA / kpar;
  const double pr = creal(pres);
  if ((pr * pr) <= (g1 * g1 - 1.0)) {
    while ((ip < (npb - 2)) && !found) {
      ip = ip + 1;
      if (PB(P, s, 2, ip + 1) > pr && PB(P, s, 2, ip) <= pr) {
        ipres = ip;
        found = 1;
      }
    }
  }
  for (ip = 0; ip <= pos; ++ip) {
    if (pr >= (PB(P, s, 2, 0) - dpb * ip) && pr < (PB(P, s, 2, 0) - dpb * (ip - 1))) {
      ipres = -ip;
      found = 1;
    }
    if (pr >= (PB(P, s, 2, npb - 1) + dpb * ip) && pr < (PB(P, s, 2, npb - 1) + dpb * (ip + 1))) {
      ipres = npb - 1 + ip;
      found = 1;
    }
  }
  int lower, upper;
  row_limits(P, s, ig, &lower, &upper);
  int int_start, int_end, lowerlimit, upperlimit;
  if (found) {
    int_start = lower;
    int_end = upper;
    lowerlimit = ipres - pos;
    upperlimit = ipres + pos + 1;
    if (ipres >= 0 && ipres <= npb)
      if (fabs(pr - PB(P, s, 2, ipres)) > (0.5 * dpb)) upperlimit = upperlimit + 1;
    if (lowerlimit < lower && upperlimit > upper) {
      *err = 8;
      return 0.0;
    } else if (lowerlimit <= lower) {
      COUNT(ORC_BR_REL_EDGE);
      int_start = 1;
      lowerlimit = 0;
      upperlimit = lower;
    } else if (upperlimit >= upper) {
      COUNT(ORC_BR_REL_EDGE);
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
#define RT(i) (resU_rel(P, s, om, nn, ig, (i)) * int_T_rel(P, s, nn, ig, (i), mode))
  if (int_start <= lowerlimit) res = res + RT(int_start);
  for (ip = int_start + 1; ip <= lowerlimit - 1; ++ip) res = res + 2.0 * RT(ip);
  if (int_start < lowerlimit) res = res + RT(lowerlimit);
  if (upperlimit <= int_end) res = res + RT(upperlimit);
  for (ip = upperlimit + 1; ip <= int_end - 1; ++ip) res = res + 2.0 * RT(ip);
  if (upperlimit < int_end) res = res + RT(int_end);
#undef RT
  res = res * dpb;
  if (found && lowerlimit >= int_start && upperlimit <= int_end)
    res = res + principal_integral_rel(P, s, om, nn, mode, ig, ipres, upperlimit);
  return res;
}

/* integrate_res_rel, ALPS_fns_rel.f90:460-511 */
int orc_rel_error = 0;

double complex orc_integrate_res_rel(const orc_plasma *P, int s, double complex om, int nn, int mode) {
  const int ng = P->ngamma;
  const double dg = GR(P, s, 2, 2) - GR(P, s, 1, 2);
  double complex res = 0.0;
  int err = 0;
  COUNT(ORC_BR_REL_RES);
  for (int ig = 1; ig <= ng - 2; ++ig) res = res + 2.0 * integrate_resU_rel(P, s, om, nn, mode, ig, &err);
  res = res + integrate_resU_rel(P, s, om, nn, mode, ng - 1, &err);
  if (err) orc_rel_error = err;
  return res * dg * 0.25;
}

/* landau_integrate_rel, ALPS_fns_rel.f90:1005-1092 */
double complex orc_landau_integrate_rel(const orc_plasma *P, int s, double complex om, int nn, int mode) {
  const double complex ii = I;
  const int ng = P->ngamma;
  const double ms = P->ms[s - 1], qs = P->qs[s - 1], vA = P->vA, kpar = P->kpar;
  const double dg = GR(P, s, 2, 2) - GR(P, s, 1, 2);
  const double dpb = PB(P, s, 2, 2) - PB(P, s, 2, 1);
  double complex res = 0.0;
  for (int ig = 1; ig <= ng - 1; ++ig) {
    const double g1 = GR(P, s, ig, 1);
    const double complex pres = g1 * om * vA / kpar - (1.0 * nn) * qs * vA / (kpar * ms);
    if ((creal(pres) * creal(pres)) <= (g1 * g1 - 1.0)) {
      COUNT(ORC_BR_REL_LANDAU);
      double h = 1.0;
      if (ig == ng - 1) h = 0.5;
      double complex dfg, dfp;
      if (ig == 1)
        dfg = (orc_eval_fit(P, s, ig + 1, pres) - orc_eval_fit(P, s, ig, pres)) / dg;
      else
        dfg = (orc_eval_fit(P, s, ig + 1, pres) - orc_eval_fit(P, s, ig - 1, pres)) / (2.0 * dg);
      dfp = (orc_eval_fit(P, s, ig, pres + dpb) - orc_eval_fit(P, s, ig, pres - dpb)) / (2.0 * dpb);
      res = res - h * (om * dfg + (kpar / (vA)) * dfp) * int_T_res_rel(P, s, nn, ig, pres, mode);
    }
  }
  return res * ii * dg * PI_R * 2.0 * PI_R * (qs * vA / (kpar * ms)) * cube(ms / vA);
}

/* int_ee_rel, ALPS_fns_rel.f90:1097-1215 (independent of om) */
double complex orc_int_ee_rel(const orc_plasma *P, int s) {
  const int ng = P->ngamma;
  const double ms = P->ms[s - 1], qs = P->qs[s - 1], vA = P->vA;
  const double dg = GR(P, s, 2, 2) - GR(P, s, 1, 2);
  const double dpb = PB(P, s, 2, 2) - PB(P, s, 2, 1);
  double complex res = 0.0;
  int lower, upper, ig = ng - 1;
  row_limits(P, s, ig, &lower, &upper);
  res = res + PB(P, s, ig, lower) * DF0R(P, s, ig, lower, 2);
  res = res + PB(P, s, ig, upper) * DF0R(P, s, ig, upper, 2);
  for (int ip = lower + 1; ip <= upper - 1; ++ip)
    res = res + 2.0 * PB(P, s, ig, ip) * DF0R(P, s, ig, ip, 2);
  for (ig = 1; ig <= ng - 2; ++ig) {
    row_limits(P, s, ig, &lower, &upper);
    for (int ip = lower + 1; ip <= upper - 1; ++ip)
      res = res + 4.0 * PB(P, s, ig, ip) * DF0R(P, s, ig, ip, 2);
    res = res + 2.0 * PB(P, s, ig, lower) * DF0R(P, s, ig, lower, 2);
    res = res + 2.0 * PB(P, s, ig, upper) * DF0R(P, s, ig, upper, 2);
  }
  res = res * 2.0 * PI_R * qs / (ms);
  return res * dg * dpb * 0.25 * cube(ms / vA);
}

/* determine_resonances, relativistic branch (ALPS_fns.f90:680-701): over the
   whole (pperp, ppar) grid with each point's Gamma. */
void orc_determine_resonances_rel(const orc_plasma *P, int s, double complex om, int nn, int *fp,
                                  int *fm) {
  const double ms = P->ms[s - 1], qs = P->qs[s - 1], vA = P->vA, kpar = P->kpar;
  *fp = *fm = 0;
  for (int iperp = 0; iperp <= P->nperp; ++iperp)
    for (int ipar = 0; ipar <= P->npar - 1; ++ipar) {
      const double p1 = PPR(P, s, iperp, ipar, 1), p2 = PPR(P, s, iperp, ipar, 2);
      const double gamma = sqrt((p1 * p1 + p2 * p2) * (vA * vA) / (ms * ms) + 1.0);
      double complex p_res = (gamma * ms * om - 1.0 * nn * qs) / kpar;
      if (PPR(P, s, 2, ipar, 2) <= creal(p_res) && PPR(P, s, 2, ipar + 1, 2) > creal(p_res)) *fp = 1;
      p_res = (gamma * ms * om + 1.0 * nn * qs) / kpar;
      if (PPR(P, s, 2, ipar, 2) <= creal(p_res) && PPR(P, s, 2, ipar + 1, 2) > creal(p_res)) *fm = 1;
    }
}
