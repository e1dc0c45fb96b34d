// Closed-form susceptibility of bi-Maxwellian and cold species (use_bM = T),
// ALPS_NHDS.f90:59-866, for a batch of omegas.
//
// The reference evaluates these species with NHDS's calc_chi on the rank that
// holds n = 0 (ALPS_fns.f90:338-360) instead of integrating a table.  Here the
// omega-independent parts -- the cutoff nmaxrun and the scaled modified Bessel
// values I_n(z) e^-z and their z-derivatives -- are computed on the host once
// per wavevector (bessi_host, nhds_tables); k_nhds then sums Stix's Y tensor
// over n = -nmaxrun..nmaxrun for each omega, with the plasma dispersion
// function from the Faddeeva function w(z) (dev_wofz).
//
// k_nhds runs after k_sweep on the same stream and overwrites that kernel's
// partial slot 0 of the species (which k_sweep leaves zero for these species)
// with chi_NHDS / (n_s q_s), and the species' chi_low slots, so that k_assemble
// treats every species alike.
#include <cmath>
#include <vector>

#include "alpsg_device.h"

namespace {

// ---------------------------------------------------------------------------
// Host: BESSI (ALPS_NHDS.f90:751-866), I_n(x) e^-x
// ---------------------------------------------------------------------------
double bessi0_host(double x) {
  const double P1 = 1.0, P2 = 3.5156229, P3 = 3.0899424, P4 = 1.2067492, P5 = 0.2659732,
               P6 = 0.360768e-1, P7 = 0.45813e-2;
  const double Q1 = 0.39894228, Q2 = 0.1328592e-1, Q3 = 0.225319e-2, Q4 = -0.157565e-2,
               Q5 = 0.916281e-2, Q6 = -0.2057706e-1, Q7 = 0.2635537e-1, Q8 = -0.1647633e-1,
               Q9 = 0.392377e-2;
  double ax = std::fabs(x);
  if (ax < 3.75) {
    const double y = (x / 3.75) * (x / 3.75);
    return (P1 + y * (P2 + y * (P3 + y * (P4 + y * (P5 + y * (P6 + y * P7)))))) * std::exp(-ax);
  }
  const double y = 3.75 / ax, bx = 1.0 / std::sqrt(ax);
  ax = Q1 + y * (Q2 + y * (Q3 + y * (Q4 + y * (Q5 + y * (Q6 + y * (Q7 + y * (Q8 + y * Q9)))))));
  return ax * bx;
}

double bessi1_host(double x) {
  const double P1 = 0.5, P2 = 0.87890594, P3 = 0.51498869, P4 = 0.15084934, P5 = 0.2658733e-1,
               P6 = 0.301532e-2, P7 = 0.32411e-3;
  const double Q1 = 0.39894228, Q2 = -0.3988024e-1, Q3 = -0.362018e-2, Q4 = 0.163801e-2,
               Q5 = -0.1031555e-1, Q6 = 0.2282967e-1, Q7 = -0.2895312e-1, Q8 = 0.1787654e-1,
               Q9 = -0.420059e-2;
  double ax = std::fabs(x);
  if (ax < 3.75) {
    const double y = (x / 3.75) * (x / 3.75);
    return x * (P1 + y * (P2 + y * (P3 + y * (P4 + y * (P5 + y * (P6 + y * P7)))))) *
           std::exp(-ax);
  }
  const double y = 3.75 / ax, bx = 1.0 / std::sqrt(ax);
  ax = Q1 + y * (Q2 + y * (Q3 + y * (Q4 + y * (Q5 + y * (Q6 + y * (Q7 + y * (Q8 + y * Q9)))))));
  return ax * bx;
}

// Miller's backward recurrence from m = 2 (n + int(sqrt(float(40 n)))) with
// rescaling by 2^-512 once the exponent passes maxexponent/2
double bessi_host(int n, double x) {
  if (n < 0) n = -n;
  if (n == 0) return bessi0_host(x);
  if (n == 1) return bessi1_host(x);
  if (x == 0.0) return 0.0;
  const double tox = 2.0 / x;
  double bip = 0.0, bi = 1.0, res = 0.0;
  const int m = 2 * (n + (int)std::sqrt((float)(40 * n)));
  for (int j = m; j >= 1; --j) {
    const double bim = bip + ((double)j * tox) * bi;
    bip = bi;
    bi = bim;
    int e;
    std::frexp(bi, &e);
    if (e > 512) {
      bi = std::ldexp(bi, -512);
      bip = std::ldexp(bip, -512);
      res = std::ldexp(res, -512);
    }
    if (j == n) res = bip;
  }
  return bessi0_host(x) * (res / bi);
}

// ---------------------------------------------------------------------------
// Device: WOFZ, the Faddeeva function w(z) = exp(-z^2) erfc(-iz) by the
// algorithm of Poppe & Wijers (ACM TOMS 680), with the reference's
// default-real constants (ALPS_NHDS.f90:541-744).  Returns false on overflow.
// ---------------------------------------------------------------------------
__device__ double ipow_d(double b, int e) {
  double r = 1.0;
  while (e > 0) {
    if (e & 1) r *= b;
    b *= b;
    e >>= 1;
  }
  return r;
}

__device__ bool dev_wofz(double xi, double yi, double &u, double &v) {
#pragma clang fp contract(off)
  const double FACTOR = 1.12837916709551257388, RMAXREAL = 0.5e+154, RMAXEXP = 708.503061461606,
               RMAXGONI = 3.53711887601422e+15;
  double u2 = 0.0, v2 = 0.0;
  u = v = 0.0;
  const double xabs = fabs(xi), yabs = fabs(yi);
  const double x = xabs / (double)6.3f, y = yabs / (double)4.4f;
  if (xabs > RMAXREAL || yabs > RMAXREAL) return false;
  double qrho = x * x + y * y;
  double xquad = xabs * xabs - yabs * yabs;
  const double yquad = 2 * xabs * yabs;
  const bool a = qrho < 0.085264;
  if (a) {
    qrho = (1 - (double)0.85f * y) * sqrt(qrho);
    const int nn = (int)llround(6 + 72 * qrho);
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
    double sq, cq;
    sincos(yquad, &sq, &cq);
    u2 = daux * cq;
    v2 = -daux * sq;
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
      kapn = (int)llround(7 + 34 * qrho);
      nu = (int)llround(16 + 26 * qrho);
    }
    const bool b = h > 0.0;
    if (b) qlambda = ipow_d(h2, kapn);
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
      if (yquad > RMAXGONI || xquad > RMAXEXP) {
        u = v = 0.0;
        return false;
      }
      const double w1 = 2 * exp(xquad);
      double sq, cq;
      sincos(yquad, &sq, &cq);
      u2 = w1 * cq;
      v2 = -w1 * sq;
    }
    u = u2 - u;
    v = v2 - v;
    if (xi > 0.0) v = -v;
  } else {
    if (xi < 0.0) v = -v;
  }
  return true;
}

// plasma dispersion function Z(zeta) = i sqrt(pi) w(zeta), mirrored for
// kpar < 0 (dispfunct, ALPS_NHDS.f90:511-537)
__device__ cd dispfunct(cd zeta, bool kpos) {
  const double sqpi = 1.7724538509055160273;  // sqrt(4 atan(1))
  double u, v;
  if (kpos) {
    dev_wofz(zeta.re, zeta.im, u, v);
    return mk(-sqpi * v, sqpi * u);
  }
  dev_wofz(-zeta.re, -zeta.im, u, v);
  return mk(sqpi * v, -sqpi * u);
}

__device__ __forceinline__ cd cdivd(cd a, cd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    const double r = b.im / b.re, d = b.re + b.im * r;
    return cd{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  }
  const double r = b.re / b.im, d = b.re * r + b.im;
  return cd{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
}

__device__ __forceinline__ cd cdivr(cd a, double d) { return cd{a.re / d, a.im / d}; }

// shuffle sum over the NCH lanes of an omega group
template <int NCH>
__device__ __forceinline__ double gsum_n(double v) {
#pragma unroll
  for (int o = NCH / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Stix's Y tensor term n (calc_ypsilon, ALPS_NHDS.f90:250-375), the six
// components xx yy zz xy xz yz
__device__ void ypsilon(const SpecDev &S, bool kperp_norm, double kz, double kperp, cd x, int n,
                        double BInz, double dB, cd Y[6]) {
  const double alpha = S.bm_alpha, Omega = S.bm_Omega, vth = S.bm_vth, vd = S.bm_vd;
  const double nO = (1.0 * n) * Omega;
  const cd xd = mk(x.re - kz * vd, x.im);
  const cd resfac = mk(xd.re - nO, xd.im);
  const cd zeta = mk(resfac.re / (kz * vth), resfac.im / (kz * vth));
  const cd Z = dispfunct(zeta, !(kz < 0.0));
  cd An = mk(alpha - 1.0, 0.0);
  const cd ar = mk(alpha * resfac.re + nO, alpha * resfac.im);
  An = An + ((1.0 / (kz * vth)) * ar) * Z;
  const cd xn = mk(x.re - nO, x.im);
  cd Bn = mk((alpha * xn.re - (kz * vd - nO)) / kz, (alpha * xn.im) / kz);
  const cd q = xn * ar;
  const double dz = (kz * kz) * vth;
  Bn = Bn + mk(q.re / dz, q.im / dz) * Z;
  const double nn2 = 1.0 * (n * n);
  const double BmD = BInz - dB;
  const double d33 = ((kz * vth) * vth) * alpha;
  if (kperp_norm) {
    const double z = S.bm_z;
    Y[0] = cdivr((nn2 * BInz) * An, z);
    Y[1] = ((nn2 * BInz) / z + (2.0 * z) * BInz - (2.0 * z) * dB) * An;
    Y[2] = cdivr(((2.0 * xn) * BInz) * Bn, d33);
    Y[3] = -1.0 * ((mk(0.0, 1.0 * n) * BmD) * An);
    Y[4] = cdivr(((kperp * n) * BInz) * Bn, Omega * z);
    Y[5] = cdivr((mk(0.0, kperp) * BmD) * Bn, Omega);
  } else {
    const double zp = S.bm_zp, z = S.bm_z, k2 = kperp * kperp;
    Y[0] = cdivr((nn2 * BInz) * An, zp);
    Y[1] = ((nn2 * BInz) / zp + ((k2 * 2.0) * z) * BInz - ((k2 * 2.0) * z) * dB) * An;
    Y[2] = cdivr((((k2 * 2.0) * xn) * BInz) * Bn, d33);
    Y[3] = -1.0 * ((((mk(0.0, 1.0 * n) * BmD) * An) * kperp) * kperp);
    Y[4] = cdivr(((kperp * n) * BInz) * Bn, Omega * zp);
    Y[5] = cdivr((mk(0.0, k2 * kperp) * BmD) * Bn, Omega);
  }
}

// ---------------------------------------------------------------------------
// k_nhds: NCH lanes per omega split n; one workgroup (one wavefront) holds
// OMB = 64/NCH omegas of one species (blockIdx.y).
// ---------------------------------------------------------------------------
template <int NCH>
__global__ __launch_bounds__(64) void k_nhds(const ProbDev *__restrict__ probp, int nom,
                                             const double2 *__restrict__ om_in,
                                             const int *__restrict__ active,
                                             double2 *__restrict__ chiA,
                                             double2 *__restrict__ lowA,
                                             const int *__restrict__ pidx) {
  constexpr int OMB = 64 / NCH;
  __shared__ double2 lbuf[OMB][18];
  // with problem indices one omega per workgroup (NCH = 64)
  const ProbDev &Pd = (NCH == 64 && pidx != nullptr) ? probp[pidx[min((int)blockIdx.x, nom - 1)]] : *probp;
  const int s = blockIdx.y;
  if (s >= Pd.nspec) return;
  const SpecDev &S = Pd.sp[s];
  if (!S.usebm) return;
  const int lane = threadIdx.x;
  const int g = lane / NCH, c = lane % NCH;
  const int iom = blockIdx.x * OMB + g;
  const bool valid = (iom < nom) && (active == nullptr || active[iom] != 0);
  const double2 ov = valid ? om_in[iom] : make_double2(1.0, 0.5);
  const cd x = mk(ov.x, ov.y);
  const double kz = Pd.kpar, kperp = Pd.kperp;
  const bool kn = Pd.kperp_norm != 0;
  for (int t = c; t < 18; t += NCH) lbuf[g][t] = make_double2(0.0, 0.0);
  __syncthreads();
  cd Ys[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) Ys[k] = mk(0.0, 0.0);
  if (!S.bm_cold) {
    const int nr = S.bm_nrun;
    for (int n = -nr + c; n <= nr; n += NCH) {
      const double2 bi = S.bmi[n + nr];
      cd Y[6];
      ypsilon(S, kn, kz, kperp, x, n, bi.x, bi.y, Y);
#pragma unroll
      for (int k = 0; k < 6; ++k) Ys[k] = Ys[k] + Y[k];
      if (n >= -1 && n <= 1)
        for (int k = 0; k < 6; ++k) lbuf[g][6 * (n + 1) + k] = make_double2(Y[k].re, Y[k].im);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) Ys[k] = mk(gsum_n<NCH>(Ys[k].re), gsum_n<NCH>(Ys[k].im));
  }
  __syncthreads();
  if (!valid || c != 0) return;
  const double nq = S.dens * S.q;
  const double l2 = S.bm_ell * S.bm_ell;
  cd chi[6];
  if (S.bm_cold) {
    // calc_chi_cold (ALPS_NHDS.f90:379-464)
    const double il2 = 1.0 / l2, Om = S.bm_Omega, vd = S.bm_vd;
    const cd xd = mk(x.re - kz * vd, x.im);
    const cd den = xd * xd - mk(Om * Om, 0.0);
    const cd R = -1.0 * cdivd(il2 * xd, mk(xd.re + Om, xd.im));
    const cd L = -1.0 * cdivd(il2 * xd, mk(xd.re - Om, xd.im));
    const cd Pp = -1.0 * (il2 * (cdivd(x * x, xd * xd) +
                                 cdivd(mk((kperp * vd) * (kperp * vd), 0.0), den)));
    const cd J = -1.0 * cdivd(((il2 * kperp) * vd) * xd, den);
    const cd M = cdivd(mk(0.0, ((il2 * kperp) * vd) * Om), den);
    const cd RL = R + L, RmL = R - L;
    chi[0] = 0.5 * RL;
    chi[1] = 0.5 * RL;
    chi[2] = Pp;
    chi[3] = timesI(RmL) * (-0.5);
    chi[4] = J;
    chi[5] = M;
  } else {
    const double alpha = S.bm_alpha, vth = S.bm_vth, vd = S.bm_vd;
    const double dd = (((l2 * kz) * vth) * vth) * alpha;
    const cd dr = kn ? cdivr((2.0 * x) * vd, dd) : cdivr((((kperp * kperp) * 2.0) * x) * vd, dd);
    for (int k = 0; k < 6; ++k) chi[k] = cdivr(Ys[k], l2);
    chi[2] = dr + cdivr(Ys[2], l2);
    if (lowA != nullptr) {
      double2 *o = lowA + ((size_t)iom * Pd.nspec + s) * 18;
      for (int qq = 0; qq < 3; ++qq)
        for (int k = 0; k < 6; ++k) {
          const double2 a = lbuf[g][6 * qq + k];
          cd v = cdivr(mk(a.x, a.y), l2);
          if (qq == 1 && k == 2) v = v + dr;
          v = cdivr(v, nq);
          o[6 * qq + k] = make_double2(v.re, v.im);
        }
    }
  }
  if (S.bm_cold && lowA != nullptr) {
    double2 *o = lowA + ((size_t)iom * Pd.nspec + s) * 18;
    for (int t = 0; t < 18; ++t) o[t] = make_double2(0.0, 0.0);
  }
  // partial slot 0 of k_sweep: chi_NHDS / (n_s q_s) (ALPS_fns.f90:344-349)
  double2 *o = chiA + ((size_t)s * 6) * nom + iom;
  for (int k = 0; k < 6; ++k) {
    const cd v = cdivr(chi[k], nq);
    o[(size_t)k * nom] = make_double2(v.re, v.im);
  }
}

}  // namespace

// nmaxrun (calc_chi, ALPS_NHDS.f90:176-185) and the (I_n e^-z, dI_n/dz e^-z)
// table for n = -nmaxrun..nmaxrun of one species; returns nmaxrun.
int nhds_tables(double z, int bm_nmax, double bessel_zero, std::vector<double2> &tab) {
  int nr = 0;
  for (;;) {
    if (nr >= bm_nmax || bessi_host(nr, z) < bessel_zero) break;
    ++nr;
  }
  tab.resize(2 * nr + 1);
  for (int n = -nr; n <= nr; ++n) {
    const double B = 1.0 * bessi_host(n < 0 ? -n : n, z);
    const double dB = 5.e-1 * (bessi_host(n + 1, z) + bessi_host(n - 1, z));
    tab[n + nr] = make_double2(B, dB);
  }
  return nr;
}

hipError_t launch_nhds(hipStream_t st, const ProbDev *prob, int nspec, int nom, const double2 *om,
                       const int *active, double2 *chiA, double2 *lowA, const int *pidx) {
  if (nom <= 0) return hipSuccess;
#define NH(N)                                                                             \
  {                                                                                       \
    const int grid = (nom + (64 / N) - 1) / (64 / N);                                     \
    hipLaunchKernelGGL((k_nhds<N>), dim3(grid, nspec), dim3(64), 0, st, prob, nom, om, active, \
                       chiA, lowA, pidx);                                                 \
    return hipGetLastError();                                                             \
  }
  if (pidx != nullptr) NH(64)
  if (nom >= 1024) NH(8)
  if (nom >= 64) NH(32)
  NH(64)
#undef NH
}
