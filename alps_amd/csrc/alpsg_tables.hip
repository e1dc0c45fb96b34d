// Per-handle and per-k tables of the dispersion engine, built on the device.
//
//   k_derivative   df0 from f0 (derivative_f0, ALPS_fns.f90:103-115)
//   k_int_ee       omega-independent ee term (int_ee, ALPS_fns.f90:1457-1555)
//   k_bessel_max   max_iperp |J_n| for n = 1..ntry (determine_nmax, ALPS_fns.f90:4017-4035)
//   k_bessel       J_n table, n = -1..nmax+1 (determine_bessel_array, ALPS_fns.f90:4214-4255)
//   k_atab         J^2, pperp^2 J'^2, pperp J J' per (n, iperp) (int_T, ALPS_fns.f90:1639-1707)
//   k_moments      pperp-reduced node tables G_a(n,ipar), H_a(n,ipar)
//
// The pperp moments rest on the factorisation of resU*int_T
// (ALPS_fns.f90:1587-1592, 1659-1703): the numerator of resU is
// q*(om*df1 + A) with A = (kpar/m)*(pperp*df2 - ppar*df1), the denominator
// depends on (n, ipar, om) only, and T = C_mode(n) * b_mode(ipar) * a(n, iperp)
// with a in {J^2, pperp^2 J'^2, pperp J J'} and b in {1, ppar, ppar^2}.  The
// trapezoid weights of integrate() (ALPS_fns.f90:838-860) are separable,
// w = wperp(iperp) * wpar(ipar) with wperp = 2,...,2,1, so
//   G_a(n,ipar) = sum_iperp wperp a(n,iperp) df1(iperp,ipar)
//   H_a(n,ipar) = sum_iperp wperp a(n,iperp) A(iperp,ipar)
// are omega-independent and every later omega needs only (n, ipar) data.
#include "alpsg_device.h"

// ---------------------------------------------------------------------------
// Bessel functions (ALPS_fns_rel.f90:1580-1735), evaluated with the same
// operation order as the reference and without FMA contraction, so the table
// matches the CPU restatement up to the last bit of sin/cos/sqrt.
// ---------------------------------------------------------------------------
__device__ static double dev_bessj0(double x) {
#pragma clang fp contract(off)
  const double P1 = 1.e0, P2 = -.1098628627e-2, P3 = .2734510407e-4, P4 = -.2073370639e-5,
               P5 = .2093887211e-6;
  const double Q1 = -.1562499995e-1, Q2 = .1430488765e-3, Q3 = -.6911147651e-5,
               Q4 = .7621095161e-6, Q5 = -.9349451520e-7;
  const double R1 = 57568490574.e0, R2 = -13362590354.e0, R3 = 651619640.7e0,
               R4 = -11214424.18e0, R5 = 77392.33017e0, R6 = -184.9052456e0;
  const double S1 = 57568490411.e0, S2 = 1029532985.e0, S3 = 9494680.718e0, S4 = 59272.64853e0,
               S5 = 267.8532712e0, S6 = 1.e0;
  if (x == 0.0) return 1.0;
  double ax = fabs(x);
  if (ax < 8.0) {
    double y = x * x;
    double fr = R1 + y * (R2 + y * (R3 + y * (R4 + y * (R5 + y * R6))));
    double fs = S1 + y * (S2 + y * (S3 + y * (S4 + y * (S5 + y * S6))));
    return fr / fs;
  }
  double z = 8.0 / ax, y = z * z, xx = ax - .785398164e0;
  double fp = P1 + y * (P2 + y * (P3 + y * (P4 + y * P5)));
  double fq = Q1 + y * (Q2 + y * (Q3 + y * (Q4 + y * Q5)));
  return sqrt(.636619772e0 / ax) * (fp * cos(xx) - z * fq * sin(xx));
}

__device__ static double dev_bessj1(double x) {
#pragma clang fp contract(off)
  const double P1 = 1.e0, P2 = .183105e-2, P3 = -.3516396496e-4, P4 = .2457520174e-5,
               P5 = -.240337019e-6, P6 = .636619772e0;
  const double Q1 = .04687499995e0, Q2 = -.2002690873e-3, Q3 = .8449199096e-5,
               Q4 = -.88228987e-6, Q5 = .105787412e-6;
  const double R1 = 72362614232.e0, R2 = -7895059235.e0, R3 = 242396853.1e0,
               R4 = -2972611.439e0, R5 = 15704.48260e0, R6 = -30.16036606e0;
  const double S1 = 144725228442.e0, S2 = 2300535178.e0, S3 = 18583304.74e0,
               S4 = 99447.43394e0, S5 = 376.9991397e0, S6 = 1.e0;
  double ax = fabs(x);
  if (ax < 8.0) {
    double y = x * x;
    double fr = R1 + y * (R2 + y * (R3 + y * (R4 + y * (R5 + y * R6))));
    double fs = S1 + y * (S2 + y * (S3 + y * (S4 + y * (S5 + y * S6))));
    return x * (fr / fs);
  }
  double z = 8.0 / ax, y = z * z;
  double xx = ax - (double)2.35619491f;  // default-real literal in the reference
  double fp = P1 + y * (P2 + y * (P3 + y * (P4 + y * P5)));
  double fq = Q1 + y * (Q2 + y * (Q3 + y * (Q4 + y * Q5)));
  double r = sqrt(P6 / ax) * (cos(xx) * fp - z * sin(xx) * fq);
  return r * copysign(S6, x);
}

// BESSJ with the Miller start index m = 2*((n+int(sqrt(float(40n))))/2)
// supplied by the host (mtab[n]) so that it is the same integer as the CPU's
// single-precision sqrt gives.
__device__ static double dev_bessj(int n, double x, const int *__restrict__ mtab) {
#pragma clang fp contract(off)
  const double BIGNO = 1.e10, BIGNI = 1.e-10;
  if (n == 0) return dev_bessj0(x);
  if (n == 1) return dev_bessj1(x);
  if (x == 0.0) return 0.0;
  double tox = 2.0 / x;
  if (x > (double)(float)n) {
    double bjm = dev_bessj0(x), bj = dev_bessj1(x), bjp;
    for (int j = 1; j <= n - 1; ++j) {
      bjp = j * tox * bj - bjm;
      bjm = bj;
      bj = bjp;
    }
    return bj;
  }
  int m = mtab[n];
  double res = 0.0, sum = 0.0, bjp = 0.0, bj = 1.0, bjm;
  int jsum = 0;
  for (int j = m; j >= 1; --j) {
    bjm = j * tox * bj - bjp;
    bjp = bj;
    bj = bjm;
    if (fabs(bj) > BIGNO) {
      bj = bj * BIGNI;
      bjp = bjp * BIGNI;
      res = res * BIGNI;
      sum = sum * BIGNI;
    }
    if (jsum != 0) sum = sum + bj;
    jsum = 1 - jsum;
    if (j == n) res = bjp;
  }
  sum = 2.0 * sum - bj;
  return res / sum;
}

// df0 on the interior (derivative_f0).  f0, pp in the handle's C layout:
// f0[s][iperp][ipar], pperp[s][iperp], ppar[s][ipar]; df1/df2 [s][iperp][ipar].
__global__ void k_derivative(int nspec, int nperp, int npar, const double *__restrict__ f0,
                             const double *__restrict__ pperp, const double *__restrict__ ppar,
                             double *__restrict__ df1, double *__restrict__ df2) {
  int ipar = blockIdx.x * blockDim.x + threadIdx.x;
  int iperp = blockIdx.y;
  int s = blockIdx.z;
  if (ipar < 1 || ipar > npar - 1 || iperp < 1 || iperp > nperp - 1) return;
  size_t row = (size_t)s * (nperp + 1) * (npar + 1);
  size_t o = row + (size_t)iperp * (npar + 1) + ipar;
  const double *pe = pperp + (size_t)s * (nperp + 1);
  const double *pa = ppar + (size_t)s * (npar + 1);
  df1[o] = (f0[o + (npar + 1)] - f0[o - (npar + 1)]) / (pe[iperp + 1] - pe[iperp - 1]);
  df2[o] = (f0[o + 1] - f0[o - 1]) / (pa[ipar + 1] - pa[ipar - 1]);
}

// Copy a user-supplied df0 (already interior-only) is done on the host side.

// int_ee for one species per block (blockDim = 256).  Weights follow the
// reference loops: corners 2,2,1,1 and the (1,1) corner's quirk
// pp(..,1)*df0(..,1) in place of pp(..,2)*df0(..,1).
__global__ void k_int_ee(int nperp, int npar, const double *__restrict__ pperp,
                         const double *__restrict__ ppar, const double *__restrict__ df1,
                         const double *__restrict__ df2, const double *__restrict__ qs,
                         const double *__restrict__ ms, double *__restrict__ out) {
  __shared__ double red[256];
  int s = blockIdx.x;
  const double *pe = pperp + (size_t)s * (nperp + 1);
  const double *pa = ppar + (size_t)s * (npar + 1);
  const double *d1 = df1 + (size_t)s * (nperp + 1) * (npar + 1);
  const double *d2 = df2 + (size_t)s * (nperp + 1) * (npar + 1);
  double acc = 0.0;
  int n_int = (nperp - 1) * (npar - 1);
  for (int t = threadIdx.x; t < n_int; t += blockDim.x) {
    int iperp = 1 + t / (npar - 1), ipar = 1 + t % (npar - 1);
    size_t o = (size_t)iperp * (npar + 1) + ipar;
    double wperp = (iperp == nperp - 1) ? 1.0 : 2.0;
    double wpar = (ipar == 1 || ipar == npar - 1) ? 1.0 : 2.0;
    double term;
    if (iperp == 1 && ipar == 1)
      term = pa[ipar] * (d2[o] * pe[iperp] - pe[iperp] * d1[o]);
    else
      term = pa[ipar] * (d2[o] * pe[iperp] - pa[ipar] * d1[o]);
    acc += (wperp * wpar) * term;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    double dpperp = pe[2] - pe[1];
    double dppar = fabs(pa[2] - pa[1]);
    double v = red[0];
    v = v * 2.0 * 3.141592653589793238462643383279502884 * qs[s] / ms[s];
    v = v * dpperp * dppar * 0.25;
    out[s] = v;
  }
}

// max_iperp |J_n(kperp*pperp/q)| for one n (one 64-lane block)
// One block of BMAX_T threads per n: every thread takes at most a few pperp
// points, so a block's time is that of one or two Miller recurrences.
constexpr int BMAX_T = 256;

__device__ __forceinline__ void bessel_max_block(int n, int nperp, double kperp,
                                                 const double *__restrict__ pe, double q,
                                                 const int *__restrict__ mtab, double *out,
                                                 double *red) {
  double m = 0.0;
  for (int ip = threadIdx.x; ip <= nperp; ip += blockDim.x) {
    double z = kperp * pe[ip] / q;
    m = fmax(m, fabs(dev_bessj(n, z, mtab)));
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// max_iperp |J_n(kperp*pperp/q)| for n = 1..ntry: one block per (n, species).
__global__ void k_bessel_max(int nperp, int ntry, double kperp, const double *__restrict__ pperp,
                             const double *__restrict__ qs, const int *__restrict__ mtab,
                             double *__restrict__ out) {
  __shared__ double red[BMAX_T];
  int n = blockIdx.x + 1;
  int s = blockIdx.y;
  bessel_max_block(n, nperp, kperp, pperp + (size_t)s * (nperp + 1), qs[s], mtab,
                   out + (size_t)s * ntry + (n - 1), red);
}

// bessel_array(n, iperp) for n = -1..nmax+1 -> bes[(n+1)*(nperp+1) + iperp]
__device__ __forceinline__ void bessel_entry(int ip, int n, int nperp, int nmax, double kperp,
                                             double q, const double *__restrict__ pperp,
                                             const int *__restrict__ mtab,
                                             double *__restrict__ bes) {
  if (ip > nperp || n > nmax + 1) return;
  double z = kperp * pperp[ip] / q;
  double v = (n == -1) ? -dev_bessj(1, z, mtab) : dev_bessj(n, z, mtab);
  bes[(size_t)(n + 1) * (nperp + 1) + ip] = v;
}

__global__ void k_bessel(int nperp, int nmax, double kperp, double q,
                         const double *__restrict__ pperp, const int *__restrict__ mtab,
                         double *__restrict__ bes) {
  bessel_entry(blockIdx.x * blockDim.x + threadIdx.x, (int)blockIdx.y - 1, nperp, nmax, kperp, q,
               pperp, mtab, bes);
}

// a-types per (n >= 0, iperp): J^2, pperp^2 J'^2, pperp J J', J.
// J'_n = (J_{n-1} - J_{n+1})/2 for n >= 1, -J_1 for n = 0 (int_T, ALPS_fns.f90:1647-1656).
// The negative harmonics use the same values: J_{-n}^2 = J_n^2 and
// J_{-n} J'_{-n} = J_n J'_n exactly in floating point.
__device__ __forceinline__ void atab_entry(int ip, int n, int nperp, int nmax,
                                           const double *__restrict__ pperp,
                                           const double *__restrict__ bes,
                                           double *__restrict__ atab) {
  if (ip > nperp || n > nmax) return;
  const int W = nperp + 1;
  double J = bes[(size_t)(n + 1) * W + ip];
  double Jp = (n >= 1) ? 0.5 * (bes[(size_t)n * W + ip] - bes[(size_t)(n + 2) * W + ip])
                       : -bes[(size_t)2 * W + ip];
  double p = pperp[ip];
  double *o = atab + ((size_t)n * W + ip) * ATAB_STRIDE;
  o[0] = J * J;
  o[1] = (p * p) * Jp * Jp;
  o[2] = p * J * Jp;
  o[3] = J;
}

__global__ void k_atab(int nperp, int nmax, const double *__restrict__ pperp,
                       const double *__restrict__ bes, double *__restrict__ atab) {
  atab_entry(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y, nperp, nmax, pperp, bes, atab);
}

// Landau sums of a single-Maxwellian species (the omega-independent part of
// landau_integrate, ALPS_fns.f90:1382-1444, when eval_fit = c0(iperp) E(p)):
// with dfpar = c0 (E(p+d) - E(p-d))/2d and dfperp = beta E(p), the sum over
// iperp of h a K is qk [km A sum(h a pperp c0) + (om - km p) E(p) sum(h a beta)],
// A = (E(p+d) - E(p-d))/2d.  One thread per (n, sum), iperp in order.
__global__ void k_lsum(int nperp, int nmax, double dpp, const double *__restrict__ pperp,
                       const double *__restrict__ fitc, const double *__restrict__ atab,
                       double *__restrict__ lsum) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * (nmax + 1)) return;
  const int n = t / 6, w = t % 6, k = w % 3;
  const int W = nperp + 1;
  auto c0 = [&](int ip) { return fitc[(size_t)ip * FIT_STRIDE]; };
  double acc = 0.0;
  for (int ip = 0; ip <= nperp; ++ip) {
    const double h = (ip == 0 || ip >= nperp - 1) ? 0.5 : 1.0;
    const double a = atab[((size_t)n * W + ip) * ATAB_STRIDE + k];
    double v;
    if (w < 3) {
      v = pperp[ip] * c0(ip);
    } else if (ip == 0) {
      v = (c0(1) - c0(0)) * (1.0 / dpp);
    } else if (ip == nperp) {
      v = (c0(nperp) - c0(nperp - 1)) * (1.0 / dpp);
    } else {
      v = (c0(ip + 1) - c0(ip - 1)) * (1.0 / (2.0 * dpp));
    }
    acc += (h * a) * v;
  }
  lsum[t] = acc;
}

hipError_t launch_lsum(hipStream_t st, int nperp, int nmax, double dpp, const double *pperp,
                       const double *fitc, const double *atab, double *lsum) {
  const int n = 6 * (nmax + 1);
  hipLaunchKernelGGL(k_lsum, dim3((n + 63) / 64), dim3(64), 0, st, nperp, nmax, dpp, pperp, fitc,
                     atab, lsum);
  return hipGetLastError();
}

// Node tables: one thread per (n, ipar), summing over iperp.  The a-values of a
// given n are shared by the whole block (staged in LDS); df1/df2 rows are read
// coalesced along ipar.
__device__ __forceinline__ void moments_block(int bx, int n, int nperp, int npar, double kpar,
                                              double m, const double *__restrict__ pperp,
                                              const double *__restrict__ ppar,
                                              const double *__restrict__ df1,
                                              const double *__restrict__ df2,
                                              const double *__restrict__ atab,
                                              double *__restrict__ nodes, double *lds_a) {
  int ipar = 1 + bx * blockDim.x + threadIdx.x;
  const int W = nperp + 1;
  for (int t = threadIdx.x; t < (nperp - 1); t += blockDim.x) {
    const double *a = atab + ((size_t)n * W + (t + 1)) * ATAB_STRIDE;
    double wperp = (t + 1 == nperp - 1) ? 1.0 : 2.0;
    lds_a[3 * t + 0] = wperp * a[0];
    lds_a[3 * t + 1] = wperp * a[1];
    lds_a[3 * t + 2] = wperp * a[2];
  }
  __syncthreads();
  if (ipar > npar - 1) return;
  double pa = ppar[ipar];
  double kom = kpar / m;
  double g1 = 0, h1 = 0, g2 = 0, h2 = 0, g3 = 0, h3 = 0;
  for (int iperp = 1; iperp <= nperp - 1; ++iperp) {
    size_t o = (size_t)iperp * (npar + 1) + ipar;
    double d1 = df1[o], d2 = df2[o];
    double A = kom * (pperp[iperp] * d2 - pa * d1);
    double a1 = lds_a[3 * (iperp - 1) + 0], a2 = lds_a[3 * (iperp - 1) + 1],
           a3 = lds_a[3 * (iperp - 1) + 2];
    g1 += a1 * d1;
    h1 += a1 * A;
    g2 += a2 * d1;
    h2 += a2 * A;
    g3 += a3 * d1;
    h3 += a3 * A;
  }
  double *o = nodes + ((size_t)n * (npar - 1) + (ipar - 1)) * NODE_STRIDE;
  o[0] = g1;
  o[1] = h1;
  o[2] = g2;
  o[3] = h2;
  o[4] = g3;
  o[5] = h3;
}

__global__ void k_moments(int nperp, int npar, int nmax, double kpar, double m,
                          const double *__restrict__ pperp, const double *__restrict__ ppar,
                          const double *__restrict__ df1, const double *__restrict__ df2,
                          const double *__restrict__ atab, double *__restrict__ nodes) {
  extern __shared__ double lds_a[];  // (nperp-1) * 3
  moments_block(blockIdx.x, blockIdx.y, nperp, npar, kpar, m, pperp, ppar, df1, df2, atab, nodes,
                lds_a);
}

// Split-form records of the one-lane-per-omega sweep.  With den = m w - kpar p
// -+ n q and U_a = w G_a + H_a,
//   U_a / den = G_a / m + C_a / den,  C_a = G_a (kpar p +- n q) / m + H_a,
// where C_a is real and omega-independent.  Record (n, ipar):
//   [0] kpar p, [1..6] C1, p C1, p^2 C1, C2, C3, p C3 for +n, [7..12] for -n.
__global__ void k_crec(int npar, int nmax, double kpar, double q, double m,
                       const double *__restrict__ ppar, const double *__restrict__ nodes,
                       double *__restrict__ crec) {
  const int ipar = 1 + blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (ipar > npar - 1 || n > nmax) return;
  const double *r = nodes + ((size_t)n * (npar - 1) + (ipar - 1)) * NODE_STRIDE;
  const double p = ppar[ipar], p2 = p * p;
  const double kp = kpar * p, nq = (1.0 * n) * q;
  double *o = crec + ((size_t)n * (npar - 1) + (ipar - 1)) * CREC_STRIDE;
  o[0] = kp;
  for (int sg = 0; sg < 2; ++sg) {
    const double t = (sg == 0 ? kp + nq : kp - nq) / m;
    const double c1 = r[0] * t + r[1], c2 = r[2] * t + r[3], c3 = r[4] * t + r[5];
    double *c = o + 1 + 6 * sg;
    c[0] = c1;
    c[1] = p * c1;
    c[2] = p2 * c1;
    c[3] = c2;
    c[4] = c3;
    c[5] = p * c3;
  }
}

// rsum[n] = sum over the interior nodes ipar = 2..npar-2 of
// {G1, p G1, p^2 G1, G2, G3, p G3} / m: one wavefront per harmonic, lanes
// strided over the nodes, then a fixed-order butterfly over the lanes
__global__ __launch_bounds__(64) void k_rsum(int npar, int nmax, double m,
                                             const double *__restrict__ ppar,
                                             const double *__restrict__ nodes,
                                             double *__restrict__ rsum) {
  const int n = blockIdx.x, lane = threadIdx.x;
  double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int ipar = 2 + lane; ipar <= npar - 2; ipar += 64) {
    const double *r = nodes + ((size_t)n * (npar - 1) + (ipar - 1)) * NODE_STRIDE;
    const double p = ppar[ipar];
    a[0] += r[0] / m;
    a[1] += (p * r[0]) / m;
    a[2] += ((p * p) * r[0]) / m;
    a[3] += r[2] / m;
    a[4] += r[4] / m;
    a[5] += (p * r[4]) / m;
  }
#pragma unroll
  for (int k = 0; k < 6; ++k)
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a[k] += __shfl_xor(a[k], o, 64);
  if (lane < RSUM_STRIDE) rsum[(size_t)n * RSUM_STRIDE + lane] = (lane < 6) ? a[lane] : 0.0;
}

// Far-row series tables.  For harmonic n, sign sg (d_i = kpar p_i + n q for
// +n, kpar p_i - n q for -n) and D = min |d_i| over ipar = 1..npar-1, the
// full-range node sum of node product k (the same products and trapezoid
// weights as the sweep: 1/2 at ipar = 1 and npar-1, 1 between) is
//   sum_i w_i pk_i (w G_i + H_i) / (m w - d_i) = sum_q c_q z^q,  z = m w / D,
// from 1/(m w - d) = -(1/d) sum_j (m w / d)^j and w = z D / m:
//   c_q = -sum_i w_i pk_i rho_i^q ((G_i / m) [q >= 1] + (H_i / D) rho_i),
//   rho_i = D / d_i, |rho_i| <= 1.
// One workgroup per (n, sg): lane q of each of the 4 wavefronts sums c_q over
// a quarter of the nodes; the quarters are added in wavefront order.
static_assert(SER_Q == 64, "one lane per series coefficient");
__device__ __forceinline__ void series_block(int n, int sg, int npar, double kpar, double q,
                                             double m, const double *__restrict__ ppar,
                                             const double *__restrict__ nodes,
                                             double *__restrict__ ser, double *__restrict__ serD) {
  __shared__ double part[4][6][64];
  __shared__ double dpart[4];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const double nq = (sg == 0 ? 1.0 : -1.0) * ((1.0 * n) * q);
  double dmin = INFINITY;
  for (int i = 1 + threadIdx.x; i <= npar - 1; i += 256) dmin = fmin(dmin, fabs(kpar * ppar[i] + nq));
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) dmin = fmin(dmin, __shfl_xor(dmin, o, 64));
  if (lane == 0) dpart[wv] = dmin;
  __syncthreads();
  dmin = fmin(fmin(dpart[0], dpart[1]), fmin(dpart[2], dpart[3]));
  // the -0 row does not exist; a node on d = 0 leaves the row to the sweep
  const bool ok = dmin > 0.0 && dmin < INFINITY && !(n == 0 && sg == 1);
  double a[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  if (ok) {
    const double gq = lane >= 1 ? 1.0 / m : 0.0, hD = 1.0 / dmin;
    for (int i = 1 + wv; i <= npar - 1; i += 4) {
      const double *r = nodes + ((size_t)n * (npar - 1) + (i - 1)) * NODE_STRIDE;
      const double p = ppar[i];
      const double rho = dmin / (kpar * p + nq);
      double pw = 1.0, b = rho;
      int e = lane;
#pragma unroll
      for (int t = 0; t < 6; ++t) {
        pw = (e & 1) ? pw * b : pw;
        b = b * b;
        e >>= 1;
      }
      const double f = ((i == 1 || i == npar - 1) ? -0.5 : -1.0) * pw;
      const double hr = hD * rho;
      const double e1 = f * (r[0] * gq + r[1] * hr);
      const double e2 = f * (r[2] * gq + r[3] * hr);
      const double e3 = f * (r[4] * gq + r[5] * hr);
      a[0] += e1;
      a[1] += p * e1;
      a[2] += (p * p) * e1;
      a[3] += e2;
      a[4] += e3;
      a[5] += p * e3;
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) part[wv][k][lane] = a[k];
  __syncthreads();
  if (wv == 0) {
    double *o = ser + ((size_t)(2 * n + sg) * SER_Q + lane) * SER_STRIDE;
#pragma unroll
    for (int k = 0; k < 6; ++k)
      o[k] = ((part[0][k][lane] + part[1][k][lane]) + part[2][k][lane]) + part[3][k][lane];
    o[6] = o[7] = 0.0;
    if (lane == 0) serD[2 * n + sg] = ok ? dmin : 0.0;
  }
}

__global__ __launch_bounds__(256) void k_series(int npar, int nmax, double kpar, double q, double m,
                                                const double *__restrict__ ppar,
                                                const double *__restrict__ nodes,
                                                double *__restrict__ ser, double *__restrict__ serD) {
  series_block(blockIdx.x, blockIdx.y, npar, kpar, q, m, ppar, nodes, ser, serD);
}

// gcum[n][i] = sum over ipar = 1..i of {G1, p G1, p^2 G1, G2, G3, p G3} / m,
// gcum[n][0] = 0: one wavefront per harmonic, an inclusive lane scan per
// chunk of 64 nodes plus the running carry
__global__ __launch_bounds__(64) void k_gcum(int npar, int nmax, double m,
                                             const double *__restrict__ ppar,
                                             const double *__restrict__ nodes,
                                             double *__restrict__ gcum) {
  const int n = blockIdx.x, lane = threadIdx.x;
  double *g = gcum + (size_t)n * npar * GCUM_STRIDE;
  if (lane < GCUM_STRIDE) g[lane] = 0.0;
  double carry[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int base = 1; base <= npar - 1; base += 64) {
    const int i = base + lane;
    double t[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (i <= npar - 1) {
      const double *r = nodes + ((size_t)n * (npar - 1) + (i - 1)) * NODE_STRIDE;
      const double p = ppar[i];
      t[0] = r[0] / m;
      t[1] = (p * r[0]) / m;
      t[2] = ((p * p) * r[0]) / m;
      t[3] = r[2] / m;
      t[4] = r[4] / m;
      t[5] = (p * r[4]) / m;
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const double v = __shfl_up(t[k], o, 64);
        if (lane >= o) t[k] += v;
      }
      t[k] += carry[k];
      carry[k] = __shfl(t[k], 63, 64);
    }
    if (i <= npar - 1) {
#pragma unroll
      for (int k = 0; k < 6; ++k) g[(size_t)i * GCUM_STRIDE + k] = t[k];
      g[(size_t)i * GCUM_STRIDE + 6] = g[(size_t)i * GCUM_STRIDE + 7] = 0.0;
    }
  }
}

hipError_t launch_split_tables(hipStream_t st, int npar, int nmax, double kpar, double q, double m,
                               const double *ppar, const double *nodes, double *crec,
                               double *rsum, double *gcum) {
  hipLaunchKernelGGL(k_crec, dim3((npar - 1 + 63) / 64, nmax + 1), dim3(64), 0, st, npar, nmax,
                     kpar, q, m, ppar, nodes, crec);
  hipLaunchKernelGGL(k_rsum, dim3(nmax + 1), dim3(64), 0, st, npar, nmax, m, ppar, nodes, rsum);
  hipLaunchKernelGGL(k_gcum, dim3(nmax + 1), dim3(64), 0, st, npar, nmax, m, ppar, nodes, gcum);
  return hipGetLastError();
}

hipError_t launch_series_tables(hipStream_t st, int npar, int nmax, double kpar, double q, double m,
                                const double *ppar, const double *nodes, double *ser, double *serD) {
  hipLaunchKernelGGL(k_series, dim3(nmax + 1, 2), dim3(256), 0, st, npar, nmax, kpar, q, m, ppar,
                     nodes, ser, serD);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Job-list forms for the root server: one launch covers the tables of many
// (request, species) jobs.  Job j owns blocks [start[j], start[j+1]) of the
// flattened grid, in the same (x, y) order as the one-job launch; every block
// runs the same device code as there, so a job's tables do not depend on the
// other jobs of its launch.
// ---------------------------------------------------------------------------
__device__ __forceinline__ int job_of(const int *__restrict__ start, int njobs, int b) {
  int lo = 0, hi = njobs - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

__global__ void k_bessel_max_jobs(const TabJob *__restrict__ jobs, const int *__restrict__ start,
                                  int njobs, const int *__restrict__ mtab) {
  __shared__ double red[BMAX_T];
  const int j = job_of(start, njobs, blockIdx.x);
  const TabJob &J = jobs[j];
  const int n = 1 + (blockIdx.x - start[j]);
  bessel_max_block(n, J.nperp, J.kperp, J.pperp, J.q, mtab, J.bmax + (n - 1), red);
}

__global__ void k_bessel_jobs(const TabJob *__restrict__ jobs, const int *__restrict__ start,
                              int njobs, const int *__restrict__ mtab) {
  const int j = job_of(start, njobs, blockIdx.x);
  const TabJob &J = jobs[j];
  const int b = blockIdx.x - start[j];
  const int tiles = (J.nperp + 1 + 63) / 64;
  const int bx = b % tiles, by = b / tiles;
  bessel_entry(bx * 64 + threadIdx.x, by - 1, J.nperp, J.nmax, J.kperp, J.q, J.pperp, mtab, J.bes);
}

__global__ void k_atab_jobs(const TabJob *__restrict__ jobs, const int *__restrict__ start,
                            int njobs) {
  const int j = job_of(start, njobs, blockIdx.x);
  const TabJob &J = jobs[j];
  const int b = blockIdx.x - start[j];
  const int tiles = (J.nperp + 1 + 63) / 64;
  const int bx = b % tiles, by = b / tiles;
  atab_entry(bx * 64 + threadIdx.x, by, J.nperp, J.nmax, J.pperp, J.bes, J.atab);
}

// moments_block for MB consecutive harmonics n0 .. n0+MB-1 (those above nmax
// are computed on n = nmax and not stored): every df0 load serves MB
// harmonics.  The weighted a-values of the block's harmonics and the pperp
// values are staged in LDS MCH rows at a time (one row per lane, its loads all
// in flight together), so the row loop reads them as LDS broadcasts instead of
// a chain of dependent scalar loads per row.  Sums run over iperp in order, as
// in moments_block.
constexpr int MB = 8;
constexpr int MCH = 64;
constexpr int MR = 4;
typedef const __attribute__((address_space(1))) double *gdptr_t;

__device__ __forceinline__ void moments_row(double (&g)[MB][3], double (&h)[MB][3],
                                            const double *la_row, double pe, double d1, double d2,
                                            double kom, double pa) {
  const double A = kom * (pe * d2 - pa * d1);
#pragma unroll
  for (int u = 0; u < MB; ++u) {
    const double a1 = la_row[3 * u + 0], a2 = la_row[3 * u + 1], a3 = la_row[3 * u + 2];
    g[u][0] += a1 * d1;
    h[u][0] += a1 * A;
    g[u][1] += a2 * d1;
    h[u][1] += a2 * A;
    g[u][2] += a3 * d1;
    h[u][2] += a3 * A;
  }
}

__device__ __forceinline__ void moments_block_mb(int bx, int n0, int nperp, int npar, int nmax,
                                                 double kpar, double m,
                                                 const double *__restrict__ pperp,
                                                 const double *__restrict__ ppar,
                                                 const double *__restrict__ df1,
                                                 const double *__restrict__ df2,
                                                 const double *__restrict__ atab,
                                                 double *__restrict__ nodes) {
  __shared__ double la[MCH][MB * 3];
  __shared__ double lpe[MCH];
  const int ipar = 1 + bx * 64 + threadIdx.x;
  const int ip = ipar <= npar - 1 ? ipar : npar - 1;
  const int W = nperp + 1;
  const double pa = ppar[ip];
  const double kom = kpar / m;
  // global (not flat) loads: their waits do not hold the LDS reads
  const gdptr_t f1 = (gdptr_t)df1, f2 = (gdptr_t)df2;
  double g[MB][3], h[MB][3];
#pragma unroll
  for (int u = 0; u < MB; ++u)
#pragma unroll
    for (int t = 0; t < 3; ++t) g[u][t] = h[u][t] = 0.0;
  for (int c0 = 1; c0 <= nperp - 1; c0 += MCH) {
    const int cn = min(MCH, nperp - c0);  // rows c0 .. c0 + cn - 1 <= nperp - 1
    __syncthreads();                       // the previous chunk is consumed
    if ((int)threadIdx.x < cn) {
      const int iperp = c0 + threadIdx.x;
      const double wperp = (iperp == nperp - 1) ? 1.0 : 2.0;
      double v[MB * 3];
#pragma unroll
      for (int u = 0; u < MB; ++u) {
        const double *a = atab + ((size_t)min(n0 + u, nmax) * W + iperp) * ATAB_STRIDE;
        v[3 * u + 0] = a[0];
        v[3 * u + 1] = a[1];
        v[3 * u + 2] = a[2];
      }
#pragma unroll
      for (int k = 0; k < MB * 3; ++k) la[threadIdx.x][k] = wperp * v[k];
      lpe[threadIdx.x] = pperp[iperp];
    }
    __syncthreads();
    // rows in groups of MR, the group's df loads issued before its sums, then
    // the remaining rows one by one
    int t0 = 0;
    for (; t0 + MR <= cn; t0 += MR) {
      double d1[MR], d2[MR];
#pragma unroll
      for (int r = 0; r < MR; ++r) {
        const size_t o = (size_t)(c0 + t0 + r) * (npar + 1) + ip;
        d1[r] = f1[o];
        d2[r] = f2[o];
      }
#pragma unroll
      for (int r = 0; r < MR; ++r) moments_row(g, h, la[t0 + r], lpe[t0 + r], d1[r], d2[r], kom, pa);
    }
    for (; t0 < cn; ++t0) {
      const size_t o = (size_t)(c0 + t0) * (npar + 1) + ip;
      moments_row(g, h, la[t0], lpe[t0], f1[o], f2[o], kom, pa);
    }
  }
  if (ipar > npar - 1) return;
#pragma unroll
  for (int u = 0; u < MB; ++u) {
    const int n = n0 + u;
    if (n > nmax) break;
    double *o = nodes + ((size_t)n * (npar - 1) + (ipar - 1)) * NODE_STRIDE;
    o[0] = g[u][0];
    o[1] = h[u][0];
    o[2] = g[u][1];
    o[3] = h[u][1];
    o[4] = g[u][2];
    o[5] = h[u][2];
  }
}

__global__ __launch_bounds__(64) void k_moments_jobs(const TabJob *__restrict__ jobs,
                                                     const int *__restrict__ start, int njobs) {
  const int j = job_of(start, njobs, blockIdx.x);
  const TabJob &J = jobs[j];
  const int b = blockIdx.x - start[j];
  const int tiles = (J.npar - 1 + 63) / 64;
  const int bx = b % tiles, by = b / tiles;
  moments_block_mb(bx, by * MB, J.nperp, J.npar, J.nmax, J.kpar, J.m, J.pperp, J.ppar, J.df1,
                   J.df2, J.atab, J.nodes);
}

// The same node tables on the matrix cores.  For one species the node sums
// are a real GEMM, nodes[(n, a)][(ipar, g/h)] = sum_iperp Aw[(n, a)][iperp]
// B[iperp][(ipar, g/h)], with Aw = w_perp a(n, iperp) (a = J^2, pperp^2 J'^2,
// pperp J J'), B(g) = df1 and B(h) = (kpar/m)(pperp df2 - ppar df1) per point,
// M = 3 (nmax+1), N = 2 (npar-1), K = nperp-1.  One wavefront computes a
// 16 x 16 MT tile with v_mfma_f64_16x16x4f64 (lane l holds A[l%16][l/16] and
// B[l/16][l%16], and D[l/16 + 4q][l%16] in element q), stepping K by 4;
// elements outside M, N or K enter as zeros.  The per-point B values are
// rounded as in moments_block; only the order of the sums over iperp differs.
constexpr int MT = 4;  // 16-column tiles per wavefront (one A operand feeds MT products)

__device__ __forceinline__ void moments_mfma_tile(int tile_c, int tile_r, int nperp, int npar,
                                                  int nmax, double kpar, double m,
                                                  const double *__restrict__ pperp,
                                                  const double *__restrict__ ppar,
                                                  const double *__restrict__ df1,
                                                  const double *__restrict__ df2,
                                                  const double *__restrict__ atab,
                                                  double *__restrict__ nodes) {
  typedef double double4v __attribute__((ext_vector_type(4)));
  const int l = threadIdx.x & 63;
  const int i = l & 15, kk = l >> 4;
  const int M = 3 * (nmax + 1), N = 2 * (npar - 1), W = nperp + 1;
  const double kom = kpar / m;
  // this lane's A row and B columns
  const int r = tile_r * 16 + i;
  const bool rv = r < M;
  const int an = rv ? r / 3 : 0, aa = rv ? r % 3 : 0;
  int ipc[MT], ghc[MT];
  bool cv[MT];
  double pac[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int c = (tile_c * MT + t) * 16 + i;
    cv[t] = c < N;
    ipc[t] = cv[t] ? 1 + (c >> 1) : 1;
    ghc[t] = c & 1;
    pac[t] = ppar[ipc[t]];
  }
  double4v acc[MT];
#pragma unroll
  for (int t = 0; t < MT; ++t) acc[t] = double4v{0.0, 0.0, 0.0, 0.0};
  for (int k0 = 1; k0 <= nperp - 1; k0 += 4) {
    const int iperp = k0 + kk;
    const bool kv = iperp <= nperp - 1;
    const int ik = kv ? iperp : 1;
    const double wperp = (ik == nperp - 1) ? 1.0 : 2.0;
    const double av = (rv && kv) ? wperp * atab[((size_t)an * W + ik) * ATAB_STRIDE + aa] : 0.0;
    const double pe = pperp[ik];
#pragma unroll
    for (int t = 0; t < MT; ++t) {
      const size_t o = (size_t)ik * (npar + 1) + ipc[t];
      const double d1 = df1[o], d2 = df2[o];
      const double bg = ghc[t] ? kom * (pe * d2 - pac[t] * d1) : d1;
      const double bv = (cv[t] && kv) ? bg : 0.0;
      acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < MT; ++t) {
    const int c = (tile_c * MT + t) * 16 + i;
    if (c >= N) continue;
    const int ipar = 1 + (c >> 1), gh = c & 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int rr = tile_r * 16 + kk + 4 * q;
      if (rr >= M) continue;
      const int n = rr / 3, a = rr % 3;
      nodes[((size_t)n * (npar - 1) + (ipar - 1)) * NODE_STRIDE + 2 * a + gh] = acc[t][q];
    }
  }
}

__global__ __launch_bounds__(64) void k_moments_mfma(int nperp, int npar, int nmax, double kpar,
                                                     double m, const double *__restrict__ pperp,
                                                     const double *__restrict__ ppar,
                                                     const double *__restrict__ df1,
                                                     const double *__restrict__ df2,
                                                     const double *__restrict__ atab,
                                                     double *__restrict__ nodes) {
  moments_mfma_tile(blockIdx.x, blockIdx.y, nperp, npar, nmax, kpar, m, pperp, ppar, df1, df2,
                    atab, nodes);
}

__global__ __launch_bounds__(64) void k_moments_mfma_jobs(const TabJob *__restrict__ jobs,
                                                          const int *__restrict__ start,
                                                          int njobs) {
  const int j = job_of(start, njobs, blockIdx.x);
  const TabJob &J = jobs[j];
  const int b = blockIdx.x - start[j];
  const int tc = (2 * (J.npar - 1) + 16 * MT - 1) / (16 * MT);
  moments_mfma_tile(b % tc, b / tc, J.nperp, J.npar, J.nmax, J.kpar, J.m, J.pperp, J.ppar, J.df1,
                    J.df2, J.atab, J.nodes);
}

// v into p[0..n-1] (a kernel rather than a memset node, so that the chain
// can also run inside a caller's graph capture)
__global__ void k_fill_i32(int *__restrict__ p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// ---------------------------------------------------------------------------
// Host launchers (called from alpsg_api.cpp).
// ---------------------------------------------------------------------------
hipError_t launch_fill_i32(hipStream_t st, int *p, int n, int v) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3((n + 255) / 256), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_derivative(hipStream_t st, int nspec, int nperp, int npar, const double *f0,
                             const double *pperp, const double *ppar, double *df1, double *df2) {
  dim3 grid((npar + 1 + 127) / 128, nperp + 1, nspec);
  hipLaunchKernelGGL(k_derivative, grid, dim3(128), 0, st, nspec, nperp, npar, f0, pperp, ppar,
                     df1, df2);
  return hipGetLastError();
}

hipError_t launch_int_ee(hipStream_t st, int nspec, int nperp, int npar, const double *pperp,
                         const double *ppar, const double *df1, const double *df2,
                         const double *qs, const double *ms, double *out) {
  hipLaunchKernelGGL(k_int_ee, dim3(nspec), dim3(256), 0, st, nperp, npar, pperp, ppar, df1, df2,
                     qs, ms, out);
  return hipGetLastError();
}

hipError_t launch_bessel_max(hipStream_t st, int nspec, int nperp, int ntry, double kperp,
                             const double *pperp, const double *qs, const int *mtab,
                             double *out) {
  hipLaunchKernelGGL(k_bessel_max, dim3(ntry, nspec), dim3(BMAX_T), 0, st, nperp, ntry, kperp, pperp,
                     qs, mtab, out);
  return hipGetLastError();
}

// J_n(z) on the (Gamma, pparbar) grid of a relativistic species, n = 0..nmax+1,
// z = (kperp m/(vA q)) sqrt(Gamma^2 - 1 - pparbar^2) (int_T_rel,
// ALPS_fns_rel.f90:1298-1322); NaN outside the subluminal cone, as there.
__global__ void k_rel_bessel(int npts, int nmax, double zfac, const double *__restrict__ grel,
                             const double *__restrict__ pbrel, const int *__restrict__ mtab,
                             double *__restrict__ out) {
#pragma clang fp contract(off)
  const int pt = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = blockIdx.y;
  if (pt >= npts || n > nmax + 1) return;
  const double g = grel[pt], pb = pbrel[pt];
  const double pperpbar = sqrt(g * g - 1.0 - pb * pb);
  out[(size_t)n * npts + pt] = dev_bessj(n, zfac * pperpbar, mtab);
}

hipError_t launch_rel_bessel(hipStream_t st, int npts, int nmax, double zfac, const double *grel,
                             const double *pbrel, const int *mtab, double *out) {
  hipLaunchKernelGGL(k_rel_bessel, dim3((npts + 255) / 256, nmax + 2), dim3(256), 0, st, npts,
                     nmax, zfac, grel, pbrel, mtab, out);
  return hipGetLastError();
}

hipError_t launch_species_k(hipStream_t st, int nperp, int npar, int nmax, double kperp,
                            double kpar, double q, double m, const double *pperp,
                            const double *ppar, const double *df1, const double *df2,
                            const int *mtab, double *bes, double *atab, double *nodes) {
  hipLaunchKernelGGL(k_bessel, dim3((nperp + 1 + 63) / 64, nmax + 3), dim3(64), 0, st, nperp,
                     nmax, kperp, q, pperp, mtab, bes);
  hipLaunchKernelGGL(k_atab, dim3((nperp + 1 + 63) / 64, nmax + 1), dim3(64), 0, st, nperp, nmax,
                     pperp, bes, atab);
  const int tc = (2 * (npar - 1) + 16 * MT - 1) / (16 * MT);
  const int tr = (3 * (nmax + 1) + 15) / 16;
  hipLaunchKernelGGL(k_moments_mfma, dim3(tc, tr), dim3(64), 0, st, nperp, npar, nmax, kpar, m,
                     pperp, ppar, df1, df2, atab, nodes);
  return hipGetLastError();
}

// Block counts per job of the four job-list launches (host side, shared with
// the caller that builds the start arrays).
int jobs_blocks_bessel_max(const TabJob &J) { return J.ntry; }
int jobs_blocks_bessel(const TabJob &J) { return ((J.nperp + 1 + 63) / 64) * (J.nmax + 3); }
int jobs_blocks_atab(const TabJob &J) { return ((J.nperp + 1 + 63) / 64) * (J.nmax + 1); }
int jobs_blocks_moments(const TabJob &J) {
  return ((2 * (J.npar - 1) + 16 * MT - 1) / (16 * MT)) * ((3 * (J.nmax + 1) + 15) / 16);
}

// start arrays: [njobs + 1] each, device copies; total = start[njobs]
hipError_t launch_bessel_max_jobs(hipStream_t st, const TabJob *jobs, const int *start, int njobs,
                                  int total, const int *mtab) {
  if (njobs <= 0 || total <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bessel_max_jobs, dim3(total), dim3(BMAX_T), 0, st, jobs, start, njobs, mtab);
  return hipGetLastError();
}

hipError_t launch_species_k_jobs(hipStream_t st, const TabJob *jobs, int njobs,
                                 const int *start_bes, int total_bes, const int *start_atab,
                                 int total_atab, const int *start_mom, int total_mom,
                                 int max_nperp, const int *mtab) {
  if (njobs <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_bessel_jobs, dim3(total_bes), dim3(64), 0, st, jobs, start_bes, njobs, mtab);
  hipLaunchKernelGGL(k_atab_jobs, dim3(total_atab), dim3(64), 0, st, jobs, start_atab, njobs);
  (void)max_nperp;
  hipLaunchKernelGGL(k_moments_mfma_jobs, dim3(total_mom), dim3(64), 0, st, jobs, start_mom, njobs);
  return hipGetLastError();
}

// Problem descriptors of the requests a server pass activates, packed on the
// host, to their slots of the device descriptor array (one block each).
__global__ void k_scatter_probs(const ProbDev *__restrict__ src, const int *__restrict__ slots,
                                int n, ProbDev *__restrict__ dst) {
  const int i = blockIdx.x;
  if (i >= n) return;
  const unsigned long long *a = reinterpret_cast<const unsigned long long *>(src + i);
  unsigned long long *b = reinterpret_cast<unsigned long long *>(dst + slots[i]);
  for (int w = threadIdx.x; w < (int)(sizeof(ProbDev) / 8); w += blockDim.x) b[w] = a[w];
}

hipError_t launch_scatter_probs(hipStream_t st, const ProbDev *src, const int *slots, int n,
                                ProbDev *dst) {
  if (n <= 0) return hipSuccess;
  static_assert(sizeof(ProbDev) % 8 == 0, "ProbDev is copied in 8-byte words");
  hipLaunchKernelGGL(k_scatter_probs, dim3(n), dim3(256), 0, st, src, slots, n, dst);
  return hipGetLastError();
}
