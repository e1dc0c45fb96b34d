// Fit setup of the hybrid analytic continuation on the device:
// determine_param_fit (ALPS_analyt.f90:365-728).
//
// * ACmethod-1 species: the Levenberg-Marquardt fit of LM_nonlinear_fit
//   (:1395-1636) with the Jacobian of determine_JT (:1183-1389), row by row
//   in pperp (or Gamma for a relativistic species), each row starting from
//   the previous row's result.  The rows form a chain, so one workgroup runs
//   one species: its 256 threads evaluate the fit and the Jacobian at the
//   row's points, form JT JT^T and JT r (one matrix entry per thread, summed
//   over the points), and one lane solves the small (<= 20) system.
// * ACmethod-2 species: the Chebyshev basis of set_polynomial_basis
//   (:850-889) and the normal equations of determine_GLLS /
//   least_squares_fit (:731-848), one workgroup per (species, pperp row).
//
// The reference inverts JTJ + lambda diag(JTJ) (dgetrf, dgetri) and
// multiplies by JT r; the same system is solved here by Gaussian elimination
// with partial pivoting, which differs from it only in rounding.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/alpsg.h"
#include "alpsg_guard.h"

int alpsg_internal_fail(int code, const std::string &msg);

namespace {

constexpr int FIT_NT = 256;      // threads of one fit workgroup
constexpr int FIT_MAXP = 20;     // 4 fits x 5 parameters
constexpr int FIT_MAXPT = 2048;  // points of one row (held in LDS)
constexpr int POLY_MAXN = 32;    // Chebyshev basis functions

// One species of the Levenberg-Marquardt fit.  Point j of row r of the
// coordinate arrays x (pperp or Gamma), y (ppar or pparbar) and of f0 is
// element off + r*srow + j*spt of the Fortran arrays as the caller holds them.
struct FitJob {
  const double *x, *y, *f;
  long long srow, spt;
  int npts;    // last point index of a full row (npar or npparbar)
  int nrows;   // rows 0..nrows-1 (nperp+1 or ngamma+1)
  int rel, logfit, nfits, nparams;
  int type[4];
  double pc[4];     // perp_correction of each fit
  double va2;       // vA^2 / ms^2 (fit type 3)
  const double *par0;  // [nparams] initial guess, parameters of fit 1 first
  double *par_out;     // [nrows][FIT_MAXP]
  double *jt;          // workspace [FIT_MAXP][npts+1]
  double *quality;     // [1]: sum over rows of the least-squares quality
  int *steps;          // [1]: LM steps over all rows
  int *err;            // [1]: 0, 2 (alps_error(2)) or 3 (singular matrix)
};

// One ACmethod-2 species: f0 row r point j at f[off + r*srow + j*spt].
struct PolyJob {
  const double *f;
  long long srow, spt;
  int npar, order, logfit, nrows;
  double *basis;  // workspace [(npar+1)][order+1]
  double *coef;   // [nrows][order+1]
  int *err;
};

__device__ __forceinline__ int nparams_of(int t) {
  return t == 1 ? 3 : t == 2 ? 5 : t == 3 ? 3 : t == 4 ? 1 : t == 5 ? 3 : t == 6 ? 4 : 0;
}

// fit_function (ALPS_analyt.f90:185-266) at a real point.  The reference
// works in complex arithmetic with zero imaginary parts, where complex**real
// of a positive base is exp(p log x).
__device__ double fit_value(const FitJob &J, const double *par, double x, double y) {
  double f = 0.0;
  int b = 0;
  for (int fi = 0; fi < J.nfits; ++fi) {
    const int t = J.type[fi];
    const double pc = J.pc[fi];
    const double *u = par + b;
    const double d = y - u[2 < nparams_of(t) ? 2 : 0];
    switch (t) {
      case 1:
        f += u[0] * exp(-pc * (x * x)) * exp(-u[1] * (d * d));
        break;
      case 2: {
        const double k = 1.0 + u[1] * (d * d) + pc * u[4] * (x * x);
        f += u[0] * exp(u[3] * log(k));
      } break;
      case 3: {
        const double s = sqrt(1.0 + ((x * x) + (d * d)) * J.va2);
        f += u[0] * exp(-u[1] * s);
      } break;
      case 4:
        f += u[0] * exp(-pc * x);
        break;
      case 5:
        f += u[0] * exp(-pc * x) * exp(-u[1] * (d * d));
        break;
      case 6: {
        const double a = u[3] * pc * (x * x) + u[1] * (d * d);
        f += u[0] * exp(0.5 * (a - exp(a)));
      } break;
      default:
        break;
    }
    b += nparams_of(t);
  }
  return f;
}

// determine_JT (ALPS_analyt.f90:1183-1389) at one point: the derivatives by
// the free parameters (the pperp parameter of kappa and bi-Moyal fits is not
// fitted in row 0), divided by f for a logarithmic fit (:1507-1511); row k of
// the Jacobian goes to out[k * stride].
__device__ void jt_point(const FitJob &J, const double *par, int row, double x, double y, double f,
                         double *out, int stride) {
  int b = 0, r = 0;
  const bool lg = J.logfit != 0;
  auto put = [&](double v) {
    out[(size_t)r * stride] = lg ? v / f : v;
    ++r;
  };
  for (int fi = 0; fi < J.nfits; ++fi) {
    const int t = J.type[fi];
    const double pc = J.pc[fi];
    const double *u = par + b;
    switch (t) {
      case 1: {
        const double d = y - u[2];
        const double e = exp(-u[1] * (d * d) - pc * (x * x));
        put(e);
        put(-(d * d) * u[0] * e);
        put(2.0 * u[1] * d * u[0] * e);
      } break;
      case 2: {
        const double d = y - u[2];
        const double k = 1.0 + u[1] * (d * d) + pc * u[4] * (x * x);
        const double km1 = pow(k, u[3] - 1.0), kp = pow(k, u[3]);
        put(kp);
        put(u[0] * u[3] * km1 * (d * d));
        put(u[0] * u[3] * km1 * 2.0 * u[1] * (u[2] - y));
        put(log(k) * u[0] * kp);
        if (row != 0) put(u[0] * u[3] * km1 * pc * (x * x));
      } break;
      case 3: {
        const double d = y - u[2];
        const double s = sqrt(1.0 + ((x * x) + (d * d)) * J.va2);
        const double e = exp(-u[1] * s);
        put(e);
        put(-u[0] * e * s);
        put(u[0] * e * (u[1] / s) * d * J.va2);
      } break;
      case 4:
        put(exp(-pc * x));
        break;
      case 5: {
        const double d = y - u[2];
        const double e = exp(-u[1] * (d * d)) * exp(-pc * x);
        put(e);
        put(-u[0] * e * (d * d));
        put(2.0 * u[0] * d * u[1] * e);
      } break;
      case 6: {
        const double d = y - u[2];
        const double a = u[3] * pc * (x * x) + u[1] * (d * d);
        const double ea = exp(a);
        const double e = exp(0.5 * (a - ea));
        put(e);
        put(u[0] * e * 0.5 * (d * d) * (1.0 - ea));
        put(u[0] * e * u[1] * (u[2] - y) * (1.0 - ea));
        if (row != 0) put(u[0] * e * 0.5 * pc * (x * x) * (1.0 - ea));
      } break;
      default:
        break;
    }
    b += nparams_of(t);
  }
}

// A x = b for a small n x n system (row-major, overwritten) by Gaussian
// elimination with partial pivoting; false for an exactly singular pivot
// (dgetrf's info > 0).
__device__ bool lu_solve(int n, double *A, double *b) {
  for (int k = 0; k < n; ++k) {
    int p = k;
    double amax = fabs(A[k * n + k]);
    for (int i = k + 1; i < n; ++i)
      if (fabs(A[i * n + k]) > amax) {
        amax = fabs(A[i * n + k]);
        p = i;
      }
    if (A[p * n + k] == 0.0) return false;
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
  return true;
}

// sum over i < n of a[i] * b[i] in four interleaved partial sums
__device__ __forceinline__ double dot4(const double *a, const double *b, int n) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = 0;
  for (; i + 3 < n; i += 4) {
    s0 += a[i] * b[i];
    s1 += a[i + 1] * b[i + 1];
    s2 += a[i + 2] * b[i + 2];
    s3 += a[i + 3] * b[i + 3];
  }
  for (; i < n; ++i) s0 += a[i] * b[i];
  return (s0 + s1) + (s2 + s3);
}

__device__ __forceinline__ double sum4(const double *a, int n) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int i = 0;
  for (; i + 3 < n; i += 4) {
    s0 += a[i];
    s1 += a[i + 1];
    s2 += a[i + 2];
    s3 += a[i + 3];
  }
  for (; i < n; ++i) s0 += a[i];
  return (s0 + s1) + (s2 + s3);
}

__global__ __launch_bounds__(FIT_NT) void k_lm_fit(const FitJob *__restrict__ jobs, int maxsteps,
                                                   double lambda0, double lambdafac, double eps) {
  const FitJob &J = jobs[blockIdx.x];
  __shared__ double par[FIT_MAXP], old[FIT_MAXP], A[FIT_MAXP * FIT_MAXP], rhs[FIT_MAXP];
  __shared__ double res[FIT_MAXPT], rn[FIT_MAXPT];
  __shared__ int mask[FIT_MAXP];
  __shared__ int s_lo, s_np, s_nJT, s_state, s_fallback;
  __shared__ double s_lsq, s_lambda;
  const int tid = threadIdx.x;
  if (tid < FIT_MAXP) par[tid] = tid < J.nparams ? J.par0[tid] : 0.0;
  double quality = 0.0, qtotal = 0.0;  // thread 0's copies
  int steps = 0;
  __syncthreads();
  for (int row = 0; row < J.nrows; ++row) {
    const double *xr = J.x + row * J.srow, *yr = J.y + row * J.srow, *fr = J.f + row * J.srow;
    if (tid == 0) {
      int b = 0, n = 0;
      for (int fi = 0; fi < J.nfits; ++fi) {
        const int t = J.type[fi], np = nparams_of(t);
        for (int k = 0; k < np; ++k) mask[b + k] = 1;
        if (row == 0 && t == 2) mask[b + 4] = 0;
        if (row == 0 && t == 6) mask[b + 3] = 0;
        b += np;
      }
      for (int k = 0; k < J.nparams; ++k) n += mask[k];
      s_nJT = n;
      s_fallback = 0;
      s_state = 0;
      if (J.rel) {
        // the subluminal part of the row: f0_rel > -1 (ALPS_analyt.f90:588-600)
        int lo = 1, hi = J.npts - 1;
        bool flo = false, fhi = false;
        for (int j = 1; j <= J.npts - 1; ++j) {
          const double fm = fr[(j - 1) * J.spt], f0 = fr[j * J.spt], fp = fr[(j + 1) * J.spt];
          if (!flo && fm <= -1.0 && f0 > -1.0) {
            lo = j;
            flo = true;
          }
          if (!fhi && f0 > -1.0 && fp <= -1.0) {
            hi = j;
            fhi = true;
          }
        }
        if (hi - lo > 2) {
          s_lo = lo;
          s_np = hi - lo;
        } else {
          // too few points (ALPS_analyt.f90:617-628): the amplitude from the
          // middle point; the quality of the previous row is counted again
          s_fallback = 1;
          int pb = 0;
          for (int fi = 0; fi < J.nfits; ++fi) {
            par[fi + pb] = fr[((hi + lo) / 2) * J.spt] / exp(-J.pc[fi] * xr[1 * J.spt]);
            if (J.type[fi] == 5) {
              par[fi + pb + 1] = 1.0e-12;
              par[fi + pb + 2] = 0.0;
              pb += 2;
            }
          }
        }
      } else {
        s_lo = 0;
        s_np = J.npts;
      }
      if (!s_fallback && 1 + s_np < J.nparams) {
        *J.err = 2;
        s_state = -1;
      }
      s_lambda = lambda0;
    }
    __syncthreads();
    if (s_state < 0) return;
    if (!s_fallback) {
      const int lo = s_lo, np = s_np, nJT = s_nJT, ld = np + 1;
      int counter = 0;
      for (;;) {
        ++counter;
        // residuals and the Jacobian at the current parameters
        for (int j = tid; j <= np; j += FIT_NT) {
          const double x = xr[(lo + j) * J.spt], y = yr[(lo + j) * J.spt], fv = fr[(lo + j) * J.spt];
          const double g = J.logfit ? log(fv) : fv;
          const double f = fit_value(J, par, x, y);
          jt_point(J, par, row, x, y, f, J.jt + j, ld);
          res[j] = g - (J.logfit ? log(f) : f);
          rn[j] = res[j] * res[j];
        }
        __syncthreads();
        // JT JT^T, JT r and the least-squares sum
        for (int e = tid; e < nJT * nJT + nJT + 1; e += FIT_NT) {
          if (e < nJT * nJT) {
            const int a = e / nJT, b = e % nJT;
            if (b >= a) {
              const double s = dot4(J.jt + (size_t)a * ld, J.jt + (size_t)b * ld, ld);
              A[a * nJT + b] = s;
              A[b * nJT + a] = s;
            }
          } else if (e < nJT * nJT + nJT) {
            const int a = e - nJT * nJT;
            rhs[a] = dot4(J.jt + (size_t)a * ld, res, ld);
          } else {
            s_lsq = sum4(rn, ld);
          }
        }
        __syncthreads();
        if (tid == 0) {
          for (int a = 0; a < nJT; ++a) A[a * nJT + a] = A[a * nJT + a] + s_lambda * A[a * nJT + a];
          if (!lu_solve(nJT, A, rhs)) {
            *J.err = 3;
            s_state = -1;
          } else {
            for (int k = 0, l = 0; k < J.nparams; ++k) {
              old[k] = par[k];
              if (mask[k]) par[k] = par[k] + rhs[l++];
            }
          }
        }
        __syncthreads();
        if (s_state < 0) return;
        // the least-squares sum at the new parameters
        for (int j = tid; j <= np; j += FIT_NT) {
          const double x = xr[(lo + j) * J.spt], y = yr[(lo + j) * J.spt], fv = fr[(lo + j) * J.spt];
          const double g = J.logfit ? log(fv) : fv;
          const double f = fit_value(J, par, x, y);
          const double r = g - (J.logfit ? log(f) : f);
          rn[j] = r * r;
        }
        __syncthreads();
        if (tid == 0) {
          const double lsqnew = sum4(rn, ld), lsq = s_lsq;
          if (lsqnew > lsq) {
            for (int k = 0; k < J.nparams; ++k) par[k] = old[k];
            s_lambda = s_lambda * lambdafac;
          } else {
            s_lambda = s_lambda / lambdafac;
          }
          if (fabs(lsqnew - lsq) < eps || counter == maxsteps) s_state = 1;
          quality = lsq;
        }
        __syncthreads();
        if (s_state != 0) break;
      }
      if (tid == 0) steps += counter;
    }
    if (tid == 0) {
      qtotal += quality;
      for (int k = 0; k < FIT_MAXP; ++k) J.par_out[(size_t)row * FIT_MAXP + k] = par[k];
    }
    __syncthreads();
  }
  if (tid == 0) {
    *J.quality = qtotal;
    *J.steps = steps;
  }
}

// Chebyshev basis (set_polynomial_basis, kind 1): T_n at yy = -1 + 2 j / npar.
__global__ __launch_bounds__(256) void k_poly_basis(const PolyJob *__restrict__ jobs) {
  const PolyJob &J = jobs[blockIdx.x];
  const int no = J.order;
  for (int j = threadIdx.x; j <= J.npar; j += blockDim.x) {
    const double yy = -1.0 + j * (2.0 / J.npar);
    double *B = J.basis + (size_t)j * (no + 1);
    B[0] = 1.0;
    if (no >= 1) B[1] = yy;
    for (int n = 2; n <= no; ++n) B[n] = 2.0 * yy * B[n - 1] - B[n - 2];
  }
}

// determine_GLLS for one pperp row: log10 f0 (zeros replaced by 1 % of the
// smallest positive value, ALPS_analyt.f90:752-776) fitted by the normal
// equations B^T B c = B^T y.  grid (rows, jobs).
__global__ __launch_bounds__(256) void k_glls(const PolyJob *__restrict__ jobs) {
  const PolyJob &J = jobs[blockIdx.y];
  const int row = blockIdx.x, no = J.order, nb = no + 1, np1 = J.npar + 1;
  if (row >= J.nrows) return;
  __shared__ double y[FIT_MAXPT], A[POLY_MAXN * POLY_MAXN], rhs[POLY_MAXN];
  __shared__ double s_min, s_minpos;
  __shared__ int s_any;
  const double *fr = J.f + row * J.srow;
  const int tid = threadIdx.x;
  for (int j = tid; j < np1; j += blockDim.x) y[j] = fr[j * J.spt];
  __syncthreads();
  if (tid == 0 && J.logfit) {
    double mn = y[0], mp = 0.0;
    int any = 0;
    for (int j = 0; j < np1; ++j) {
      mn = y[j] < mn ? y[j] : mn;
      if (y[j] > 0.0 && (!any || y[j] < mp)) {
        mp = y[j];
        any = 1;
      }
    }
    s_min = mn;
    s_minpos = any ? mp : 1.0e-40;
  }
  __syncthreads();
  if (J.logfit) {
    const bool repl = !(s_min > 0.0);
    for (int j = tid; j < np1; j += blockDim.x) {
      const double v = y[j];
      y[j] = log10((repl && v == 0.0) ? 0.01 * s_minpos : v);
    }
  }
  __syncthreads();
  for (int e = tid; e < nb * nb + nb; e += blockDim.x) {
    if (e < nb * nb) {
      const int a = e / nb, b = e % nb;
      double s = 0.0;
      for (int j = 0; j < np1; ++j) s += J.basis[(size_t)j * nb + a] * J.basis[(size_t)j * nb + b];
      A[a * nb + b] = s;
    } else {
      const int a = e - nb * nb;
      double s = 0.0;
      for (int j = 0; j < np1; ++j) s += J.basis[(size_t)j * nb + a] * y[j];
      rhs[a] = s;
    }
  }
  __syncthreads();
  if (tid == 0) {
    if (!lu_solve(nb, A, rhs)) {
      *J.err = 3;
    } else {
      for (int n = 0; n < nb; ++n) J.coef[(size_t)row * nb + n] = rhs[n];
    }
  }
}

// device allocations of one call, freed after the call's stream has drained
struct Dev {
  hipStream_t st = nullptr;
  std::vector<void *> ptrs;
  ~Dev() {
    if (st) (void)hipStreamSynchronize(st);
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <class T>
  hipError_t alloc(T **p, size_t n) {
    hipError_t e = hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T));
    if (e == hipSuccess) ptrs.push_back(*p);
    return e;
  }
};

struct StreamGuard {
  hipStream_t st = nullptr;
  ~StreamGuard() {
    if (st) (void)hipStreamDestroy(st);
  }
};

int nparams_host(int t) {
  return t == 1 ? 3 : t == 2 ? 5 : t == 3 ? 3 : t == 4 ? 1 : t == 5 ? 3 : t == 6 ? 4 : 0;
}

#define FCHK(expr)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return alpsg_internal_fail(ALPSG_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

}  // namespace

extern "C" int alpsg_determine_param_fit(const alpsg_plasma_desc *d, const alpsg_fit_params *fpar,
                                         double *param_fit, int *fit_type, double *perp_correction,
                                         double *poly_fit_coeffs, double *quality) {
  ALPSG_TRY
  if (!d || !param_fit || !fit_type || !perp_correction)
    return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: null argument");
  const int nspec = d->nspec, nperp = d->nperp, npar = d->npar, mf = d->max_fits;
  if (nspec < 1 || nspec > ALPSG_MAXSPEC || nperp < 2 || npar < 2 || mf < 1 || mf > ALPSG_MAXFITS)
    return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: bad nspec/nperp/npar/max_fits");
  if (!d->pp || !d->f0 || !d->ACmethod || !d->n_fits || !d->fit_type || !d->perp_correction ||
      !d->param_fit)
    return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: pp, f0, ACmethod, n_fits, fit_type, "
                                            "perp_correction and param_fit are needed");
  const int np = d->param_fit_nperp;
  alpsg_fit_params fp = {500, 1.0, 10.0, 1.0e-8};  // ALPS_var.f90:288-297
  if (fpar) fp = *fpar;
  if (fp.maxsteps_fit < 1)
    return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: maxsteps_fit < 1");
  auto is_bm = [&](int s) { return d->usebM && d->usebM[s]; };
  auto is_rel = [&](int s) { return d->relativistic && d->relativistic[s]; };
  auto logfit = [&](int s) { return d->logfit ? d->logfit[s] != 0 : true; };  // ALPS_io.f90:142
  const size_t nrel_tot = [&] {
    size_t n = 0;
    for (int s = 0; s < nspec; ++s) n += is_rel(s) ? 1 : 0;
    return n;
  }();
  // Fortran layouts
  auto PF = [&](const double *a, int s, int ip, int k, int f) {
    return a[s + (size_t)nspec * (ip + (size_t)(np + 1) * (k + 5 * (size_t)f))];
  };
  auto PFw = [&](int s, int ip, int k, int f) -> double & {
    return param_fit[s + (size_t)nspec * (ip + (size_t)(np + 1) * (k + 5 * (size_t)f))];
  };
  // read what the outputs may overwrite (they may alias the description's arrays)
  std::vector<int> ft_in((size_t)nspec * mf);
  std::vector<double> pc_in((size_t)nspec * mf);
  for (size_t i = 0; i < ft_in.size(); ++i) {
    ft_in[i] = d->fit_type[i];
    pc_in[i] = d->perp_correction[i];
  }
  auto FT = [&](int s, int f) { return ft_in[s + (size_t)nspec * f]; };
  auto PC = [&](int s, int f) { return pc_in[s + (size_t)nspec * f]; };
  struct SpecPlan {
    int kind = 0;  // 0 reset only, 1 LM fit, 2 Chebyshev
    int nparams = 0, nrows = 0, srel = -1;
    std::vector<double> par0;
  };
  std::vector<SpecPlan> plan(nspec);
  int nrel = 0, nfit = 0, npoly = 0, maxorder = 0;
  for (int s = 0; s < nspec; ++s) {
    SpecPlan &P = plan[s];
    if (is_rel(s)) P.srel = nrel++;
    if (is_bm(s) || d->ACmethod[s] == 0) continue;
    if (d->ACmethod[s] == 2) {
      if (!d->poly_kind || !d->poly_order || !poly_fit_coeffs)
        return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: ACmethod 2 needs poly_kind, "
                                                "poly_order and poly_fit_coeffs");
      if (d->poly_kind[s] != 1)
        return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: poly_kind must be 1 "
                                                "(alps_error(10))");
      const int no = d->poly_order[s];
      if (no < 0 || no + 1 > POLY_MAXN || no > d->poly_max_order)
        return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "determine_param_fit: poly_order out of range");
      if (npar + 1 > FIT_MAXPT)
        return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "determine_param_fit: npar too large");
      P.kind = 2;
      P.nrows = nperp + 1;
      maxorder = std::max(maxorder, no);
      ++npoly;
      continue;
    }
    if (d->ACmethod[s] != 1)
      return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: ACmethod must be 0, 1 or 2");
    const int nf = d->n_fits[s];
    if (nf < 1 || nf > mf) return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: bad n_fits");
    P.kind = 1;
    for (int f = 0; f < nf; ++f) {
      const int t = FT(s, f);
      if (t < 1 || t > 6) return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: fit_type must be 1..6");
      for (int k = 0; k < nparams_host(t); ++k) P.par0.push_back(PF(d->param_fit, s, 0, k, f));
    }
    P.nparams = (int)P.par0.size();
    if (is_rel(s)) {
      if (!d->gamma_rel || !d->pparbar_rel || !d->f0_rel || d->ngamma < 2 || d->npparbar < 2)
        return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: relativistic species need "
                                                "gamma_rel, pparbar_rel and f0_rel");
      if (d->npparbar + 1 > FIT_MAXPT)
        return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "determine_param_fit: npparbar too large");
      P.nrows = d->ngamma + 1;
    } else {
      if (npar + 1 > FIT_MAXPT)
        return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "determine_param_fit: npar too large");
      P.nrows = nperp + 1;
    }
    if (P.nrows - 1 > np)
      return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: param_fit_nperp below the fitted rows");
    ++nfit;
  }

  // device copies of the grids, the jobs and their outputs
  StreamGuard sg;
  FCHK(hipStreamCreateWithFlags(&sg.st, hipStreamNonBlocking));
  hipStream_t st = sg.st;
  Dev dev;
  dev.st = st;
  const size_t npp = (size_t)nspec * (nperp + 1) * (npar + 1);
  double *pp_d = nullptr, *f0_d = nullptr, *g_d = nullptr, *pb_d = nullptr, *fr_d = nullptr;
  std::vector<FitJob> fj;
  std::vector<PolyJob> pj;
  std::vector<int> fj_spec, pj_spec;
  double *par_out = nullptr, *jt = nullptr, *q_d = nullptr, *basis = nullptr, *coef = nullptr;
  int *steps_d = nullptr, *err_d = nullptr;
  std::vector<size_t> par_off, coef_off;
  if (nfit + npoly > 0) {
    FCHK(dev.alloc(&pp_d, 2 * npp));
    FCHK(dev.alloc(&f0_d, npp));
    FCHK(hipMemcpyAsync(pp_d, d->pp, sizeof(double) * 2 * npp, hipMemcpyHostToDevice, st));
    FCHK(hipMemcpyAsync(f0_d, d->f0, sizeof(double) * npp, hipMemcpyHostToDevice, st));
    FCHK(dev.alloc(&err_d, nfit + npoly));
    FCHK(hipMemsetAsync(err_d, 0, sizeof(int) * (nfit + npoly), st));
  }
  const size_t nrp = nrel_tot * (size_t)(d->ngamma + 1) * (d->npparbar + 1);
  if (nrel > 0 && nfit > 0) {
    bool need = false;
    for (int s = 0; s < nspec; ++s) need = need || (plan[s].kind == 1 && is_rel(s));
    if (need) {
      FCHK(dev.alloc(&g_d, nrp));
      FCHK(dev.alloc(&pb_d, nrp));
      FCHK(dev.alloc(&fr_d, nrp));
      FCHK(hipMemcpyAsync(g_d, d->gamma_rel, sizeof(double) * nrp, hipMemcpyHostToDevice, st));
      FCHK(hipMemcpyAsync(pb_d, d->pparbar_rel, sizeof(double) * nrp, hipMemcpyHostToDevice, st));
      FCHK(hipMemcpyAsync(fr_d, d->f0_rel, sizeof(double) * nrp, hipMemcpyHostToDevice, st));
    }
  }
  std::vector<double> par0_all;
  if (nfit > 0) {
    size_t rows = 0, ptsmax = 0;
    for (int s = 0; s < nspec; ++s)
      if (plan[s].kind == 1) {
        par_off.push_back(rows);
        rows += plan[s].nrows;
        ptsmax = std::max<size_t>(ptsmax, is_rel(s) ? d->npparbar + 1 : npar + 1);
      }
    FCHK(dev.alloc(&par_out, rows * FIT_MAXP));
    FCHK(dev.alloc(&jt, (size_t)nfit * FIT_MAXP * ptsmax));
    FCHK(dev.alloc(&q_d, nfit));
    FCHK(dev.alloc(&steps_d, nfit));
    double *par0_d = nullptr;
    FCHK(dev.alloc(&par0_d, (size_t)nfit * FIT_MAXP));
    par0_all.assign((size_t)nfit * FIT_MAXP, 0.0);
    int i = 0;
    for (int s = 0; s < nspec; ++s) {
      const SpecPlan &P = plan[s];
      if (P.kind != 1) continue;
      FitJob J{};
      if (is_rel(s)) {
        const long long nr = (long long)nrel_tot;
        J.x = g_d + P.srel;
        J.y = pb_d + P.srel;
        J.f = fr_d + P.srel;
        J.srow = nr;
        J.spt = nr * (d->ngamma + 1);
        J.npts = d->npparbar;
        J.rel = 1;
      } else {
        J.x = pp_d + s;
        J.y = pp_d + s + npp;
        J.f = f0_d + s;
        J.srow = nspec;
        J.spt = (long long)nspec * (nperp + 1);
        J.npts = npar;
        J.rel = 0;
      }
      J.nrows = P.nrows;
      J.logfit = logfit(s) ? 1 : 0;
      J.nfits = d->n_fits[s];
      J.nparams = P.nparams;
      for (int f = 0; f < 4; ++f) {
        J.type[f] = f < J.nfits ? FT(s, f) : 0;
        J.pc[f] = f < J.nfits ? PC(s, f) : 0.0;
      }
      J.va2 = d->vA * d->vA / (d->ms[s] * d->ms[s]);
      for (int k = 0; k < P.nparams; ++k) par0_all[(size_t)i * FIT_MAXP + k] = P.par0[k];
      J.par0 = par0_d + (size_t)i * FIT_MAXP;
      J.par_out = par_out + par_off[i] * FIT_MAXP;
      J.jt = jt + (size_t)i * FIT_MAXP * ptsmax;
      J.quality = q_d + i;
      J.steps = steps_d + i;
      J.err = err_d + i;
      fj.push_back(J);
      fj_spec.push_back(s);
      ++i;
    }
    FCHK(hipMemcpyAsync(par0_d, par0_all.data(), sizeof(double) * par0_all.size(), hipMemcpyHostToDevice, st));
    FitJob *fj_d = nullptr;
    FCHK(dev.alloc(&fj_d, fj.size()));
    FCHK(hipMemcpyAsync(fj_d, fj.data(), sizeof(FitJob) * fj.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_lm_fit, dim3(nfit), dim3(FIT_NT), 0, st, fj_d, fp.maxsteps_fit,
                       fp.lambda_initial_fit, fp.lambdafac_fit, fp.epsilon_fit);
    FCHK(hipGetLastError());
  }
  if (npoly > 0) {
    size_t ctot = 0;
    const int nb = maxorder + 1;
    FCHK(dev.alloc(&basis, (size_t)npoly * (npar + 1) * nb));
    for (int s = 0; s < nspec; ++s)
      if (plan[s].kind == 2) {
        coef_off.push_back(ctot);
        ctot += (size_t)(nperp + 1) * (d->poly_order[s] + 1);
      }
    FCHK(dev.alloc(&coef, ctot));
    int i = 0;
    for (int s = 0; s < nspec; ++s) {
      if (plan[s].kind != 2) continue;
      PolyJob J{};
      J.f = f0_d + s;
      J.srow = nspec;
      J.spt = (long long)nspec * (nperp + 1);
      J.npar = npar;
      J.order = d->poly_order[s];
      J.logfit = logfit(s) ? 1 : 0;
      J.nrows = nperp + 1;
      J.basis = basis + (size_t)i * (npar + 1) * nb;
      J.coef = coef + coef_off[i];
      J.err = err_d + nfit + i;
      pj.push_back(J);
      pj_spec.push_back(s);
      ++i;
    }
    PolyJob *pj_d = nullptr;
    FCHK(dev.alloc(&pj_d, pj.size()));
    FCHK(hipMemcpyAsync(pj_d, pj.data(), sizeof(PolyJob) * pj.size(), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(k_poly_basis, dim3(npoly), dim3(256), 0, st, pj_d);
    hipLaunchKernelGGL(k_glls, dim3(nperp + 1, npoly), dim3(256), 0, st, pj_d);
    FCHK(hipGetLastError());
  }
  // results
  std::vector<int> err_h(nfit + npoly, 0);
  std::vector<double> q_h(nfit, 0.0), par_h, coef_h;
  if (nfit + npoly > 0)
    FCHK(hipMemcpyAsync(err_h.data(), err_d, sizeof(int) * err_h.size(), hipMemcpyDeviceToHost, st));
  if (nfit > 0) {
    size_t rows = 0;
    for (const FitJob &J : fj) rows += J.nrows;
    par_h.resize(rows * FIT_MAXP);
    FCHK(hipMemcpyAsync(q_h.data(), q_d, sizeof(double) * nfit, hipMemcpyDeviceToHost, st));
    FCHK(hipMemcpyAsync(par_h.data(), par_out, sizeof(double) * par_h.size(), hipMemcpyDeviceToHost, st));
  }
  if (npoly > 0) {
    size_t ctot = 0;
    for (int i = 0; i < npoly; ++i) ctot += (size_t)pj[i].nrows * (pj[i].order + 1);
    coef_h.resize(ctot);
    FCHK(hipMemcpyAsync(coef_h.data(), coef, sizeof(double) * ctot, hipMemcpyDeviceToHost, st));
  }
  FCHK(hipStreamSynchronize(st));
  for (size_t i = 0; i < err_h.size(); ++i) {
    if (err_h[i] == 2)
      return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: fewer points than fit parameters "
                                              "(alps_error(2))");
    if (err_h[i] == 3)
      return alpsg_internal_fail(ALPSG_E_ARG, "determine_param_fit: fit matrix is numerically singular");
  }
  // write back in the Fortran layouts
  double qt = 0.0;
  int fi_job = 0, pi_job = 0;
  for (int s = 0; s < nspec; ++s) {
    const SpecPlan &P = plan[s];
    if (P.kind == 1) {
      const double *pr = par_h.data() + par_off[fi_job] * FIT_MAXP;
      qt += q_h[fi_job];
      for (int ip = 0; ip < P.nrows; ++ip) {
        int b = 0;
        for (int f = 0; f < d->n_fits[s]; ++f) {
          const int npf = nparams_host(FT(s, f));
          for (int k = 0; k < npf; ++k) PFw(s, ip, k, f) = pr[(size_t)ip * FIT_MAXP + b + k];
          b += npf;
        }
      }
      for (int f = 0; f < mf; ++f) {
        fit_type[s + (size_t)nspec * f] = FT(s, f);
        perp_correction[s + (size_t)nspec * f] = PC(s, f);
      }
      ++fi_job;
      continue;
    }
    // use_bM, ACmethod 0 and 2 (ALPS_analyt.f90:430-461)
    for (int f = 0; f < mf; ++f) {
      for (int ip = 0; ip <= np; ++ip)
        for (int k = 0; k < 5; ++k) PFw(s, ip, k, f) = 0.0;
      fit_type[s + (size_t)nspec * f] = 1;
      perp_correction[s + (size_t)nspec * f] = 1.0;
    }
    if (P.kind == 2) {
      const int no = d->poly_order[s], pm = d->poly_max_order;
      const double *c = coef_h.data() + coef_off[pi_job];
      for (int ip = 0; ip <= nperp; ++ip)
        for (int n = 0; n <= no; ++n)
          poly_fit_coeffs[s + (size_t)nspec * (ip + (size_t)(nperp + 1) * n)] = c[(size_t)ip * (no + 1) + n];
      (void)pm;
      ++pi_job;
    }
  }
  if (quality) *quality = qt;
  return ALPSG_OK;
  ALPSG_CATCH
}
