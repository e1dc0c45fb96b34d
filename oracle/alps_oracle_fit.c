/*
 * alps_oracle_fit.c -- CPU restatement of ALPS's fit setup for the hybrid
 * analytic continuation: determine_param_fit (ALPS_analyt.f90:365-728) with
 * its Levenberg-Marquardt fit (LM_nonlinear_fit :1395-1636, determine_JT
 * :1183-1389, fit_function :185-266) and the Chebyshev least squares of
 * ACmethod 2 (set_polynomial_basis :850-889, determine_GLLS :731-793,
 * least_squares_fit :795-848).
 *
 * TEST INFRASTRUCTURE ONLY (see alps_oracle.h).
 *
 * The reference inverts JTJ + lambda diag(JTJ) with LAPACK dgetrf/dgetri and
 * multiplies the inverse by JT r; here the same system is solved by Gaussian
 * elimination with partial pivoting (the same LU factorisation, without
 * forming the inverse), which differs only in rounding.  The normal equations
 * of the Chebyshev fit (dgemm, dgemv, dgesv) are formed and solved the same
 * way.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "alps_oracle.h"

#define PF(P, is, ip, k, f) \
  ((P)->param_fit[((((size_t)(is) * ((P)->param_np + 1) + (ip)) * 5 + (k)) * ORC_MAXFITS) + (f)])

/* number of parameters of one fit of the given type (ALPS_analyt.f90:466-473) */
static int nparams_of(int type) {
  switch (type) {
    case 1: return 3;
    case 2: return 5;
    case 3: return 3;
    case 4: return 1;
    case 5: return 3;
    case 6: return 4;
    default: return 0;
  }
}

/* fit_function (ALPS_analyt.f90:185-266) at real (pperp, ppar): the
   reference evaluates it in complex arithmetic with a zero imaginary part;
   complex**real of a positive base is exp(p log x) there. */
static double fit_value(const orc_plasma *P, int is, const double *par, double pperp, double ppar) {
  double f = 0.0;
  int b = 0;
  for (int fi = 0; fi < P->n_fits[is]; ++fi) {
    const int t = P->fit_type[is][fi];
    const double pc = P->perp_correction[is][fi];
    const double *u = par + b;
    switch (t) {
      case 1: {
        const double d = ppar - u[2];
        f += u[0] * exp(-pc * (pperp * pperp)) * exp(-u[1] * (d * d));
      } break;
      case 2: {
        const double d = ppar - u[2];
        const double k = 1.0 + u[1] * (d * d) + pc * u[4] * (pperp * pperp);
        f += u[0] * exp(u[3] * log(k));
      } break;
      case 3: {
        const double d = ppar - u[2];
        const double s = sqrt(1.0 + ((pperp * pperp) + (d * d)) * P->vA * P->vA / (P->ms[is] * P->ms[is]));
        f += u[0] * exp(-u[1] * s);
      } break;
      case 4:
        f += u[0] * exp(-pc * pperp);
        break;
      case 5: {
        const double d = ppar - u[2];
        f += u[0] * exp(-pc * pperp) * exp(-u[1] * (d * d));
      } break;
      case 6: {
        const double d = ppar - u[2];
        const double a = u[3] * pc * (pperp * pperp) + u[1] * (d * d);
        f += u[0] * exp(0.5 * (a - exp(a)));
      } break;
    }
    b += nparams_of(t);
  }
  return f;
}

/* determine_JT (ALPS_analyt.f90:1183-1389) at one point: the rows of the
   free parameters (the pperp parameter of kappa and bi-Moyal fits is fixed
   at iperp = 0). */
static void jt_column(const orc_plasma *P, int is, const double *par, int iperp, double pperp,
                      double ppar, double *jt) {
  int b = 0, r = 0;
  for (int fi = 0; fi < P->n_fits[is]; ++fi) {
    const int t = P->fit_type[is][fi];
    const double pc = P->perp_correction[is][fi];
    const double *u = par + b;
    switch (t) {
      case 1: {
        const double d = ppar - u[2];
        const double e = exp(-u[1] * (d * d) - pc * (pperp * pperp));
        jt[r + 0] = e;
        jt[r + 1] = -(d * d) * u[0] * e;
        jt[r + 2] = 2.0 * u[1] * d * u[0] * e;
        r += 3;
      } break;
      case 2: {
        const double d = ppar - u[2];
        const double k = 1.0 + u[1] * (d * d) + pc * u[4] * (pperp * pperp);
        const double km1 = pow(k, u[3] - 1.0);
        jt[r + 0] = pow(k, u[3]);
        jt[r + 1] = u[0] * u[3] * km1 * (d * d);
        jt[r + 2] = u[0] * u[3] * km1 * 2.0 * u[1] * (u[2] - ppar);
        jt[r + 3] = log(k) * u[0] * pow(k, u[3]);
        if (iperp == 0) {
          r += 4;
        } else {
          jt[r + 4] = u[0] * u[3] * km1 * pc * (pperp * pperp);
          r += 5;
        }
      } break;
      case 3: {
        const double d = ppar - u[2];
        const double va2 = P->vA * P->vA / (P->ms[is] * P->ms[is]);
        const double s = sqrt(1.0 + ((pperp * pperp) + (d * d)) * va2);
        jt[r + 0] = exp(-u[1] * s);
        jt[r + 1] = -u[0] * exp(-u[1] * s) * s;
        jt[r + 2] = u[0] * exp(-u[1] * s) * (u[1] / s) * d * P->vA * P->vA / (P->ms[is] * P->ms[is]);
        r += 3;
      } break;
      case 4:
        jt[r] = exp(-pc * pperp);
        r += 1;
        break;
      case 5: {
        const double d = ppar - u[2];
        const double e = exp(-u[1] * (d * d)) * exp(-pc * pperp);
        jt[r + 0] = e;
        jt[r + 1] = -u[0] * e * (d * d);
        jt[r + 2] = 2.0 * u[0] * d * u[1] * e;
        r += 3;
      } break;
      case 6: {
        const double d = ppar - u[2];
        const double a = u[3] * pc * (pperp * pperp) + u[1] * (d * d);
        const double ea = exp(a);
        const double e = exp(0.5 * (a - ea));
        jt[r + 0] = e;
        jt[r + 1] = u[0] * e * 0.5 * (d * d) * (1.0 - ea);
        jt[r + 2] = u[0] * e * u[1] * (u[2] - ppar) * (1.0 - ea);
        if (iperp == 0) {
          r += 3;
        } else {
          jt[r + 3] = u[0] * e * 0.5 * pc * (pperp * pperp) * (1.0 - ea);
          r += 4;
        }
      } break;
    }
    b += nparams_of(t);
  }
}

/* Solve A x = b (n x n, row-major, overwritten) by Gaussian elimination with
   partial pivoting; returns 1 for an exactly singular pivot (dgetrf info > 0). */
static int lu_solve(int n, double *A, double *b) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + k]) > amax) {
        amax = fabs(A[i * n + k]);
        p = i;
      }
    if (A[p * n + k] == 0.0) return 1;
    if (p != k) {
      for (int j = 0; j < n; ++j) {
        const double t = A[k * n + j];
        A[k * n + j] = A[p * n + j];
        A[p * n + j] = t;
      }
      const double t = b[k];
      b[k] = b[p];
      b[p] = t;
    }
    const double rp = 1.0 / A[k * n + k];
    for (int i = k + 1; i < n; ++i) {
      const double l = A[i * n + k] * rp;
      A[i * n + k] = l;
      for (int j = k + 1; j < n; ++j) A[i * n + j] -= l * A[k * n + j];
      b[i] -= l * b[k];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int j = i + 1; j < n; ++j) s -= A[i * n + j] * b[j];
    b[i] = s / A[i * n + i];
  }
  return 0;
}

#define LM_MAXP (5 * ORC_MAXFITS)

/* LM_nonlinear_fit (ALPS_analyt.f90:1395-1636) on the points 0..np of one
   row: coordinates xs (pperp or Gamma) and ys (ppar or pparbar), targets g.
   Returns 0, or 3 for a singular matrix (the reference stops). */
static int lm_fit(const orc_plasma *P, int is, int iperp, int np, const double *xs, const double *ys,
                  const double *g, int n_params, const int *mask, double *par, int maxsteps,
                  double lambda0, double lambdafac, double eps, double *quality, int *steps) {
  int nJT = 0;
  for (int k = 0; k < n_params; ++k) nJT += mask[k];
  double *JT = (double *)malloc(sizeof(double) * (size_t)nJT * (np + 1));
  double *res = (double *)malloc(sizeof(double) * (np + 1));
  double A[LM_MAXP * LM_MAXP], rhs[LM_MAXP];
  double lambda = lambda0, LSQ = 0.0;
  int counter = 0, converged = 0, rc = 0;
  const int logf = P->logfit[is];
  while (!converged) {
    ++counter;
    LSQ = 0.0;
    for (int i = 0; i <= np; ++i) {
      double col[LM_MAXP];
      jt_column(P, is, par, iperp, xs[i], ys[i], col);
      const double f = fit_value(P, is, par, xs[i], ys[i]);
      for (int k = 0; k < nJT; ++k) JT[(size_t)k * (np + 1) + i] = logf ? col[k] / f : col[k];
      res[i] = g[i] - (logf ? log(f) : f);
      LSQ = LSQ + res[i] * res[i];
    }
    for (int a = 0; a < nJT; ++a) {
      for (int b = 0; b < nJT; ++b) {
        double s = 0.0;
        for (int i = 0; i <= np; ++i) s += JT[(size_t)a * (np + 1) + i] * JT[(size_t)b * (np + 1) + i];
        A[a * nJT + b] = s;
      }
      double s = 0.0;
      for (int i = 0; i <= np; ++i) s += JT[(size_t)a * (np + 1) + i] * res[i];
      rhs[a] = s;
    }
    for (int a = 0; a < nJT; ++a) A[a * nJT + a] = A[a * nJT + a] + lambda * A[a * nJT + a];
    if (lu_solve(nJT, A, rhs)) {
      rc = 3;
      break;
    }
    for (int k = 0, l = 0; k < n_params; ++k)
      if (mask[k]) par[k] = par[k] + rhs[l++];
    double LSQnew = 0.0;
    for (int i = 0; i <= np; ++i) {
      const double f = fit_value(P, is, par, xs[i], ys[i]);
      const double r = g[i] - (logf ? log(f) : f);
      LSQnew = LSQnew + r * r;
    }
    if (LSQnew > LSQ) {
      for (int k = 0, l = 0; k < n_params; ++k)
        if (mask[k]) par[k] = par[k] - rhs[l++];
      lambda = lambda * lambdafac;
    } else {
      lambda = lambda / lambdafac;
    }
    if (fabs(LSQnew - LSQ) < eps) converged = 1;
    if (counter == maxsteps) converged = 1;
  }
  *quality = LSQ;
  *steps = counter;
  free(JT);
  free(res);
  return rc;
}

/* set_polynomial_basis (kind 1, Chebyshev on yy = -1 + 2 ipar/npar) and
   determine_GLLS: per iperp the normal equations of the fit of log10 f0 (or
   f0) to poly_order+1 basis functions. */
static int glls(orc_plasma *P, int is) {
  const int npar = P->npar, no = P->poly_order[is];
  double *B = (double *)malloc(sizeof(double) * (size_t)(npar + 1) * (no + 1));
  double *y = (double *)malloc(sizeof(double) * (npar + 1));
  double *A = (double *)malloc(sizeof(double) * (size_t)(no + 1) * (no + 1));
  double *rhs = (double *)malloc(sizeof(double) * (no + 1));
  int rc = 0;
  for (int i = 0; i <= npar; ++i) {
    const double yy = -1.0 + i * (2.0 / npar);
    B[(size_t)i * (no + 1) + 0] = 1.0;
    if (no >= 1) B[(size_t)i * (no + 1) + 1] = yy;
    for (int n = 2; n <= no; ++n)
      B[(size_t)i * (no + 1) + n] = 2.0 * yy * B[(size_t)i * (no + 1) + n - 1] - B[(size_t)i * (no + 1) + n - 2];
  }
  for (int ip = 0; ip <= P->nperp && !rc; ++ip) {
    const double *f = P->f0 + ((size_t)is * (P->nperp + 1) + ip) * (npar + 1);
    if (P->logfit[is]) {
      double mn = f[0];
      for (int i = 1; i <= npar; ++i) mn = f[i] < mn ? f[i] : mn;
      if (mn > 0.0) {
        for (int i = 0; i <= npar; ++i) y[i] = log10(f[i]);
      } else {
        double minpos = 0.0;
        int any = 0;
        for (int i = 0; i <= npar; ++i)
          if (f[i] > 0.0 && (!any || f[i] < minpos)) {
            minpos = f[i];
            any = 1;
          }
        if (!any) minpos = 1.0e-40;
        for (int i = 0; i <= npar; ++i) y[i] = log10(f[i] == 0.0 ? 0.01 * minpos : f[i]);
      }
    } else {
      for (int i = 0; i <= npar; ++i) y[i] = f[i];
    }
    for (int a = 0; a <= no; ++a) {
      for (int b = 0; b <= no; ++b) {
        double s = 0.0;
        for (int i = 0; i <= npar; ++i) s += B[(size_t)i * (no + 1) + a] * B[(size_t)i * (no + 1) + b];
        A[a * (no + 1) + b] = s;
      }
      double s = 0.0;
      for (int i = 0; i <= npar; ++i) s += B[(size_t)i * (no + 1) + a] * y[i];
      rhs[a] = s;
    }
    if (lu_solve(no + 1, A, rhs)) {
      rc = 3;
      break;
    }
    double *c = P->poly_fit_coeffs + ((size_t)is * (P->nperp + 1) + ip) * P->poly_stride;
    for (int n = 0; n <= no; ++n) c[n] = rhs[n];
  }
  free(B);
  free(y);
  free(A);
  free(rhs);
  return rc;
}

int orc_determine_param_fit(orc_plasma *P, int maxsteps, double lambda0, double lambdafac, double eps,
                            double *qualitytotal, int *steps_total) {
  double qt = 0.0, quality = 0.0;
  long long st = 0;
  for (int is = 0; is < P->nspec; ++is) {
    if (P->usebM[is] || P->ACmethod[is] == 0 || P->ACmethod[is] == 2) {
      for (int ip = 0; ip <= P->param_np; ++ip)
        for (int k = 0; k < 5; ++k)
          for (int f = 0; f < ORC_MAXFITS; ++f) PF(P, is, ip, k, f) = 0.0;
      for (int f = 0; f < ORC_MAXFITS; ++f) {
        P->fit_type[is][f] = 1;
        P->perp_correction[is][f] = 1.0;
      }
      if (!P->usebM[is] && P->ACmethod[is] == 2) {
        if (P->poly_kind[is] != 1) return 10; /* alps_error(10) */
        if (glls(P, is)) return 3;
      }
      continue;
    }
    int n_params = 0;
    for (int f = 0; f < P->n_fits[is]; ++f) n_params += nparams_of(P->fit_type[is][f]);
    const int rel = P->relativistic[is];
    const int upper = rel ? P->ngamma : P->nperp;
    /* a relativistic row with too few points counts the quality of the
       previous row again; the reference's variable also carries over from
       the previous species (uninitialised before the first fit), here it
       starts at 0 for each species */
    quality = 0.0;
    for (int ip = 0; ip <= upper; ++ip) {
      if (ip != 0)
        for (int k = 0; k < 5; ++k)
          for (int f = 0; f < ORC_MAXFITS; ++f) PF(P, is, ip, k, f) = PF(P, is, ip - 1, k, f);
      double par[LM_MAXP];
      int mask[LM_MAXP], b = 0;
      for (int f = 0; f < P->n_fits[is]; ++f) {
        const int t = P->fit_type[is][f], np = nparams_of(t);
        for (int k = 0; k < np; ++k) {
          par[b + k] = PF(P, is, ip, k, f);
          mask[b + k] = 1;
        }
        if (ip == 0 && t == 2) mask[b + 4] = 0;
        if (ip == 0 && t == 6) mask[b + 3] = 0;
        b += np;
      }
      if (rel) {
        const int ngp = P->npparbar;
        const double *fr = P->f0_rel[is] + (size_t)ip * (ngp + 1);
        const double *gr = P->gamma_rel[is] + (size_t)ip * (ngp + 1);
        const double *pr = P->pparbar_rel[is] + (size_t)ip * (ngp + 1);
        int lo = 1, hi = ngp - 1, flo = 0, fhi = 0;
        for (int j = 1; j <= ngp - 1; ++j) {
          if (!flo && fr[j - 1] <= -1.0 && fr[j] > -1.0) {
            lo = j;
            flo = 1;
          }
          if (!fhi && fr[j] > -1.0 && fr[j + 1] <= -1.0) {
            hi = j;
            fhi = 1;
          }
        }
        if (hi - lo > 2) {
          const int np = hi - lo;
          double *g = (double *)malloc(sizeof(double) * (np + 1));
          for (int j = 0; j <= np; ++j) g[j] = P->logfit[is] ? log(fr[lo + j]) : fr[lo + j];
          int steps = 0;
          if (1 + np < n_params) {
            free(g);
            return 2; /* alps_error(2) */
          }
          const int rc = lm_fit(P, is, ip, np, gr + lo, pr + lo, g, n_params, mask, par, maxsteps,
                                lambda0, lambdafac, eps, &quality, &steps);
          free(g);
          st += steps;
          if (rc) return rc;
        } else {
          /* too few points inside the subluminal range (ALPS_analyt.f90:617-628):
             the amplitude from the middle point; quality keeps its last value */
          int pb = 0;
          for (int f = 0; f < P->n_fits[is]; ++f) {
            par[f + pb] = fr[(hi + lo) / 2] / exp(-P->perp_correction[is][f] * gr[1]);
            if (P->fit_type[is][f] == 5) {
              par[f + pb + 1] = 1.0e-12;
              par[f + pb + 2] = 0.0;
              pb += 2;
            }
          }
        }
      } else {
        const int npar = P->npar;
        const double *f0r = P->f0 + ((size_t)is * (P->nperp + 1) + ip) * (npar + 1);
        double *xs = (double *)malloc(sizeof(double) * (npar + 1));
        double *ys = (double *)malloc(sizeof(double) * (npar + 1));
        double *g = (double *)malloc(sizeof(double) * (npar + 1));
        for (int j = 0; j <= npar; ++j) {
          const double *pp = P->pp + (((size_t)is * (P->nperp + 1) + ip) * (npar + 1) + j) * 2;
          xs[j] = pp[0];
          ys[j] = pp[1];
          g[j] = P->logfit[is] ? log(f0r[j]) : f0r[j];
        }
        int steps = 0, rc = 0;
        if (1 + npar < n_params) rc = 2;
        if (!rc)
          rc = lm_fit(P, is, ip, npar, xs, ys, g, n_params, mask, par, maxsteps, lambda0, lambdafac, eps,
                      &quality, &steps);
        free(xs);
        free(ys);
        free(g);
        st += steps;
        if (rc) return rc;
      }
      qt += quality;
      b = 0;
      for (int f = 0; f < P->n_fits[is]; ++f) {
        const int np = nparams_of(P->fit_type[is][f]);
        for (int k = 0; k < np; ++k) PF(P, is, ip, k, f) = par[b + k];
        b += np;
      }
    }
  }
  if (qualitytotal) *qualitytotal = qt;
  if (steps_total) *steps_total = (int)st;
  return 0;
}
