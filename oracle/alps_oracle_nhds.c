/*
 * alps_oracle_nhds.c -- CPU restatement of ALPS's closed-form susceptibility
 * for bi-Maxwellian and cold species (use_bM = T), ALPS_NHDS.f90:59-866.
 *
 * TEST INFRASTRUCTURE ONLY: part of the checker for the HIP path (see
 * alps_oracle.c).
 *
 * Arithmetic follows the reference expression by expression, including the
 * default-real (single precision) literals of WOFZ (6.3, 4.4, 0.85, 1.88 and
 * the single-precision 1.0/J of its power series, ALPS_NHDS.f90:634-652) and
 * the single-precision sqrt in BESSI's Miller start index (:801).
 *
 * Parity pins (tests/test_oracle_nhds.py):
 *   orc_wofz        scipy.special.wofz (1e-12 relative off the power-series
 *                   region, 1e-7 in it, where the single 1/J terms limit it)
 *   orc_bessi       scipy.special.ive (the A&S 9.8.1-9.8.4 polynomials, 2e-7)
 *   orc_calc_chi    closed-form Stix 10-57 with scipy's wofz/ive (1e-6), the
 *                   cold-plasma chi as the beta -> 0 limit of the hot chi, and
 *                   the tabulated-f0 integration of the same bi-Maxwellian.
 */
#include <math.h>
#include <string.h>

#include "alps_oracle.h"

/* BESSI0: I_0(x) e^{-|x|}, ALPS_NHDS.f90:822-841 */
static double bessi0(double x) {
  const double P1 = 1.0, P2 = 3.5156229, P3 = 3.0899424, P4 = 1.2067492, P5 = 0.2659732,
               P6 = 0.360768e-1, P7 = 0.45813e-2;
  const double Q1 = 0.39894228, Q2 = 0.1328592e-1, Q3 = 0.225319e-2, Q4 = -0.157565e-2,
               Q5 = 0.916281e-2, Q6 = -0.2057706e-1, Q7 = 0.2635537e-1, Q8 = -0.1647633e-1,
               Q9 = 0.392377e-2;
  double ax = fabs(x), y;
  if (ax < 3.75) {
    y = (x / 3.75) * (x / 3.75);
    return (P1 + y * (P2 + y * (P3 + y * (P4 + y * (P5 + y * (P6 + y * P7)))))) * exp(-ax);
  }
  y = 3.75 / ax;
  double bx = 1.0 / sqrt(ax);
  ax = Q1 + y * (Q2 + y * (Q3 + y * (Q4 + y * (Q5 + y * (Q6 + y * (Q7 + y * (Q8 + y * Q9)))))));
  return ax * bx;
}

/* BESSI1: I_1(x) e^{-|x|}, ALPS_NHDS.f90:846-866 */
static double bessi1(double x) {
  const double P1 = 0.5, P2 = 0.87890594, P3 = 0.51498869, P4 = 0.15084934, P5 = 0.2658733e-1,
               P6 = 0.301532e-2, P7 = 0.32411e-3;
  const double Q1 = 0.39894228, Q2 = -0.3988024e-1, Q3 = -0.362018e-2, Q4 = 0.163801e-2,
               Q5 = -0.1031555e-1, Q6 = 0.2282967e-1, Q7 = -0.2895312e-1, Q8 = 0.1787654e-1,
               Q9 = -0.420059e-2;
  double ax = fabs(x), y;
  if (ax < 3.75) {
    y = (x / 3.75) * (x / 3.75);
    return x * (P1 + y * (P2 + y * (P3 + y * (P4 + y * (P5 + y * (P6 + y * P7)))))) * exp(-ax);
  }
  y = 3.75 / ax;
  double bx = 1.0 / sqrt(ax);
  ax = Q1 + y * (Q2 + y * (Q3 + y * (Q4 + y * (Q5 + y * (Q6 + y * (Q7 + y * (Q8 + y * Q9)))))));
  return ax * bx;
}

/* BESSI: I_n(x) e^{-x} for n >= 0 by Miller's backward recurrence with
   rescaling by 2^-512 (IBIGNO = maxexponent/2), ALPS_NHDS.f90:761-817 */
double orc_bessi(int n, double x) {
  const int IACC = 40, IBIGNO = 512;
  if (n == 0) return bessi0(x);
  if (n == 1) return bessi1(x);
  if (x == 0.0) return 0.0;
  double tox = 2.0 / x, bip = 0.0, bi = 1.0, res = 0.0;
  int m = 2 * (n + (int)sqrtf((float)(IACC * n)));
  for (int j = m; j >= 1; --j) {
    double bim = bip + ((double)j * tox) * bi;
    bip = bi;
    bi = bim;
    int e;
    frexp(bi, &e);
    if (e > IBIGNO) {
      bi = ldexp(bi, -IBIGNO);
      bip = ldexp(bip, -IBIGNO);
      res = ldexp(res, -IBIGNO);
    }
    if (j == n) res = bip;
  }
  return bessi0(x) * (res / bi);
}

/* besselI, ALPS_NHDS.f90:493-508 */
double orc_besselI(int n, double x) { return n < 0 ? orc_bessi(-n, x) : orc_bessi(n, x); }

/* integer power by repeated squaring (Fortran real**integer) */
static double ipow(double b, int e) {
  double r = 1.0;
  while (e > 0) {
    if (e & 1) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

/* WOFZ: Faddeeva function w(z) = exp(-z^2) erfc(-iz) by the algorithm of
   Poppe & Wijers (ACM TOMS 680, 1990), ALPS_NHDS.f90:541-744. */
void orc_wofz(double xi, double yi, double *u_out, double *v_out, int *flag) {
  const double FACTOR = 1.12837916709551257388, RMAXREAL = 0.5e+154,
               RMAXEXP = 708.503061461606, RMAXGONI = 3.53711887601422e+15;
  double u = 0.0, v = 0.0, u2 = 0.0, v2 = 0.0;
  *flag = 0;
  const double xabs = fabs(xi), yabs = fabs(yi);
  const double x = xabs / (double)6.3f, y = yabs / (double)4.4f;
  if (xabs > RMAXREAL || yabs > RMAXREAL) goto overflow;
  double qrho = x * x + y * y;
  const double xabsq = xabs * xabs;
  double xquad = xabsq - yabs * yabs;
  const double yquad = 2 * xabs * yabs;
  const int a = qrho < 0.085264;
  if (a) {
    qrho = (1 - (double)0.85f * y) * sqrt(qrho);
    const int nn = (int)lround(6 + 72 * qrho);
    int j = 2 * nn + 1;
    double xsum = (double)(1.0f / (float)j), ysum = 0.0;
    for (int i = nn; i >= 1; --i) {
      j = j - 2;
      const double xaux = (xsum * xquad - ysum * yquad) / i;
      ysum = (xsum * yquad + ysum * xquad) / i;
      xsum = xaux + (double)(1.0f / (float)j);
    }
    const double u1 = -FACTOR * (xsum * yabs + ysum * xabs) + 1.0;
    const double v1 = FACTOR * (xsum * xabs - ysum * yabs);
    const double daux = exp(-xquad);
    u2 = daux * cos(yquad);
    v2 = -daux * sin(yquad);
    u = u1 * u2 - v1 * v2;
    v = u1 * v2 + v1 * u2;
  } else {
    double h, h2 = 0.0, qlambda = 0.0;
    int kapn, nu;
    if (qrho > 1.0) {
      h = 0.0;
      kapn = 0;
      qrho = sqrt(qrho);
      nu = (int)(3 + (1442 / (26 * qrho + 77)));
    } else {
      qrho = (1 - y) * sqrt(1 - qrho);
      h = (double)1.88f * qrho;
      h2 = 2 * h;
      kapn = (int)lround(7 + 34 * qrho);
      nu = (int)lround(16 + 26 * qrho);
    }
    const int b = h > 0.0;
    if (b) qlambda = ipow(h2, kapn);
    double rx = 0.0, ry = 0.0, sx = 0.0, sy = 0.0;
    for (int n = nu; n >= 0; --n) {
      const int np1 = n + 1;
      double tx = yabs + h + np1 * rx;
      const double ty = xabs - np1 * ry;
      const double c = 0.5 / (tx * tx + ty * ty);
      rx = c * tx;
      ry = c * ty;
      if (b && n <= kapn) {
        tx = qlambda + sx;
        sx = rx * tx - ry * sy;
        sy = ry * tx + rx * sy;
        qlambda = qlambda / h2;
      }
    }
    if (h == 0.0) {
      u = FACTOR * rx;
      v = FACTOR * ry;
    } else {
      u = FACTOR * sx;
      v = FACTOR * sy;
    }
    if (yabs == 0.0) u = exp(-xabs * xabs);
  }
  if (yi < 0.0) {
    if (a) {
      u2 = 2 * u2;
      v2 = 2 * v2;
    } else {
      xquad = -xquad;
      if (yquad > RMAXGONI || xquad > RMAXEXP) goto overflow;
      const double w1 = 2 * exp(xquad);
      u2 = w1 * cos(yquad);
      v2 = -w1 * sin(yquad);
    }
    u = u2 - u;
    v = v2 - v;
    if (xi > 0.0) v = -v;
  } else {
    if (xi < 0.0) v = -v;
  }
  *u_out = u;
  *v_out = v;
  return;
overflow:
  *flag = 1;
  *u_out = 0.0;
  *v_out = 0.0;
}

/* dispfunct: the plasma dispersion function Z(zeta) = i sqrt(pi) w(zeta),
   mirrored for kpar < 0, ALPS_NHDS.f90:511-537 */
static double complex dispfunct(double complex zeta, int kpos) {
  const double sqpi = sqrt(4.0 * atan(1.0));
  double u, v;
  int flag;
  if (kpos) {
    orc_wofz(creal(zeta), cimag(zeta), &u, &v, &flag);
    return (I * sqpi) * (u + I * v);
  }
  orc_wofz(-creal(zeta), -cimag(zeta), &u, &v, &flag);
  return (-I * sqpi) * (u + I * v);
}

/* species constants of the NHDS routines, ALPS_NHDS.f90:163-166 */
static void nhds_consts(const orc_plasma *P, int j, double *Omega, double *ell, double *vtherm,
                        double *vdrift) {
  const int s = j - 1;
  *Omega = P->qs[s] / P->ms[s];
  *ell = sqrt(P->ms[s] / (P->ns[s] * P->qs[s] * P->qs[s]));
  *vtherm = sqrt(P->bM_beta[s] / (P->ns[s] * P->ms[s]));
  *vdrift = P->bM_pdrift[s] / P->ms[s];
}

/* calc_ypsilon: the n-th term of Stix's Y tensor (10-57), ALPS_NHDS.f90:250-375.
   Y is [3][3] row-major. */
static void calc_ypsilon(const orc_plasma *P, double complex Y[3][3], int j, int n, double kz,
                         double kperp, double complex x) {
  const int s = j - 1;
  const double alpha = P->bM_alpha[s];
  double Omega, ell, vtherm, vdrift;
  nhds_consts(P, j, &Omega, &ell, &vtherm, &vdrift);
  const int kpos = !(kz < 0.0);
  const double nO = (1.0 * n) * Omega;
  const double complex zeta = ((x - kz * vdrift) - nO) / (kz * vtherm);
  const double complex resfac = (x - kz * vdrift) - nO;
  const double z = ((0.5 * (kperp * vtherm / Omega)) * (kperp * vtherm / Omega)) * alpha;
  const double zp = ((0.5 * (vtherm / Omega)) * (vtherm / Omega)) * alpha;
  const double complex Z = dispfunct(zeta, kpos);
  double complex An = alpha - 1.0;
  An = An + ((1.0 / (kz * vtherm)) * (alpha * resfac + nO)) * Z;
  double complex Bn = (alpha * (x - nO) - (kz * vdrift - nO)) / kz;
  Bn = Bn + (((x - nO) * (alpha * resfac + nO)) / ((kz * kz) * vtherm)) * Z;
  double BInz, dB;
  if (n >= 0)
    BInz = 1.0 * orc_besselI(n, z);
  else
    BInz = 1.0 * orc_besselI(-n, z);
  dB = 5.e-1 * (orc_besselI(n + 1, z) + orc_besselI(n - 1, z));
  const double nn2 = 1.0 * (n * n);
  if (P->kperp_norm) {
    Y[0][0] = ((nn2 * BInz) * An) / z;
    Y[0][1] = -(((I * n) * (BInz - dB)) * An);
    Y[0][2] = (((kperp * n) * BInz) * Bn) / (Omega * z);
    Y[1][0] = ((I * n) * (BInz - dB)) * An;
    Y[1][1] = ((nn2 * BInz) / z + (2.0 * z) * BInz - (2.0 * z) * dB) * An;
    Y[1][2] = (((I * kperp) * (BInz - dB)) * Bn) / Omega;
    Y[2][0] = (((kperp * BInz) * n) * Bn) / (Omega * z);
    Y[2][1] = -((((I * kperp) * (BInz - dB)) * Bn) / Omega);
    Y[2][2] = (((2.0 * (x - nO)) * BInz) * Bn) / (((kz * vtherm) * vtherm) * alpha);
  } else {
    const double k2 = kperp * kperp;
    Y[0][0] = ((nn2 * BInz) * An) / zp;
    Y[0][1] = -(((((I * n) * (BInz - dB)) * An) * kperp) * kperp);
    Y[0][2] = (((kperp * n) * BInz) * Bn) / (Omega * zp);
    Y[1][0] = ((((I * n) * (BInz - dB)) * An) * kperp) * kperp;
    Y[1][1] = ((nn2 * BInz) / zp + ((k2 * 2.0) * z) * BInz - ((k2 * 2.0) * z) * dB) * An;
    Y[1][2] = (((((I * kperp) * kperp) * kperp) * (BInz - dB)) * Bn) / Omega;
    Y[2][0] = (((kperp * BInz) * n) * Bn) / (Omega * zp);
    Y[2][1] = -((((((I * kperp) * kperp) * kperp) * (BInz - dB)) * Bn) / Omega);
    Y[2][2] = ((((k2 * 2.0) * (x - nO)) * BInz) * Bn) / (((kz * vtherm) * vtherm) * alpha);
  }
}

/* calc_chi_cold, ALPS_NHDS.f90:379-464 (kperp_norm = T only; the reference
   stops otherwise).  Returns 0, or -1 for kperp_norm = F. */
static int calc_chi_cold(const orc_plasma *P, double complex chi[3][3], int j, double kz,
                         double kperp, double complex x) {
  double Omega, ell, vtherm, vdrift;
  nhds_consts(P, j, &Omega, &ell, &vtherm, &vdrift);
  const double il2 = 1.0 / (ell * ell);
  const double complex xd = x - kz * vdrift;
  const double complex den = xd * xd - Omega * Omega;
  const double complex R = -((il2 * xd) / (xd + Omega));
  const double complex L = -((il2 * xd) / (xd - Omega));
  double complex Pp = (x * x) / (xd * xd) + ((kperp * vdrift) * (kperp * vdrift)) / den;
  Pp = -(il2 * Pp);
  const double complex J = -((((il2 * kperp) * vdrift) * xd) / den);
  const double complex M = (((((I * il2) * kperp) * vdrift) * Omega) / den);
  if (!P->kperp_norm) return -1;
  chi[0][0] = (R + L) / 2.0;
  chi[0][1] = -((I * (R - L)) / 2.0);
  chi[0][2] = J;
  chi[1][0] = (I * (R - L)) / 2.0;
  chi[1][1] = (R + L) / 2.0;
  chi[1][2] = M;
  chi[2][0] = J;
  chi[2][1] = -M;
  chi[2][2] = Pp;
  return 0;
}

/* nmaxrun of calc_chi, ALPS_NHDS.f90:176-185: the first n >= 0 with
   I_n(z) e^{-z} < bM_Bessel_zero, capped at bM_nmax. */
int orc_nhds_nmaxrun(const orc_plasma *P, int j, double kperp) {
  const int s = j - 1;
  double Omega, ell, vtherm, vdrift;
  nhds_consts(P, j, &Omega, &ell, &vtherm, &vdrift);
  const double z = ((0.5 * (kperp * vtherm / Omega)) * (kperp * vtherm / Omega)) * P->bM_alpha[s];
  int n = 0;
  for (;;) {
    if (n >= P->bM_nmax[s] || orc_besselI(n, z) < P->bM_Bessel_zero[s]) return n;
    n = n + 1;
  }
}

/* calc_chi, ALPS_NHDS.f90:59-242: chi [3][3] and chi_low [3][3][3] (n = -1, 0, 1
   in the last index) of species j.  Returns 0, or -1 where the reference stops. */
int orc_calc_chi(const orc_plasma *P, int j, double kz, double kperp, double complex x,
                 double complex chi[3][3], double complex chi_low[3][3][3]) {
  const int s = j - 1;
  memset(chi_low, 0, sizeof(double complex) * 27);
  if (P->bM_beta[s] == 0.0) return calc_chi_cold(P, chi, j, kz, kperp, x);
  double Omega, ell, vtherm, vdrift;
  nhds_consts(P, j, &Omega, &ell, &vtherm, &vdrift);
  const double alpha = P->bM_alpha[s];
  double complex Y[3][3], Y0[3][3], Y1[3][3], Yn1[3][3], Ynew[3][3];
  memset(Y, 0, sizeof(Y));
  memset(Y0, 0, sizeof(Y0));
  memset(Y1, 0, sizeof(Y1));
  memset(Yn1, 0, sizeof(Yn1));
  const int nmaxrun = orc_nhds_nmaxrun(P, j, kperp);
  for (int n = -nmaxrun; n <= nmaxrun; ++n) {
    calc_ypsilon(P, Ynew, j, n, kz, kperp, x);
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) {
        Y[i][k] = Y[i][k] + Ynew[i][k];
        if (n == 0) Y0[i][k] = Ynew[i][k];
        if (n == 1) Y1[i][k] = Y1[i][k] + Ynew[i][k];
        if (n == -1) Yn1[i][k] = Yn1[i][k] + Ynew[i][k];
      }
  }
  const double l2 = ell * ell;
  const double complex drift =
      P->kperp_norm ? ((2.0 * x) * vdrift) / ((((l2 * kz) * vtherm) * vtherm) * alpha)
                    : ((((kperp * kperp) * 2.0) * x) * vdrift) / ((((l2 * kz) * vtherm) * vtherm) * alpha);
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      chi[i][k] = Y[i][k] / l2;
      chi_low[i][k][1] = Y0[i][k] / l2;
      chi_low[i][k][2] = Y1[i][k] / l2;
      chi_low[i][k][0] = Yn1[i][k] / l2;
    }
  chi[2][2] = drift + Y[2][2] / l2;
  chi_low[2][2][1] = Y0[2][2] / l2 + drift;
  return 0;
}
