// Device-side data of the dispersion engine: one ProbDev per handle holds the
// per-species tables that the kernels read.  All tables live in HBM and are
// built on the device (alpsg_tables.hip); the structs only carry pointers and
// scalars, so a ProbDev can be copied to device memory as one block.
#pragma once

#include <hip/hip_runtime.h>

#define ALPSG_DMAXSPEC 8
#define ALPSG_DMAXFITS 4
#define NODE_STRIDE 6   // doubles per (n, ipar) node record: G1 H1 G2 H2 G3 H3 (ppar from the grid)
#define ATAB_STRIDE 4   // doubles per (n, iperp) entry: J^2, pperp^2 J'^2, pperp J J', J
#define FIT_STRIDE 8    // doubles per (iperp, fit) entry of the analytic-continuation table
#define CREC_STRIDE 13  // doubles per (n, ipar) record of the split-form sweep table
#define RSUM_STRIDE 8   // doubles per harmonic of the split-form row sums
#define GCUM_STRIDE 8   // doubles per (n, ipar) cumulative G/m sum (six used)
#define SER_Q 64        // coefficients of a far row's series in m w (degrees 0..SER_Q-1)
#define SER_STRIDE 8    // doubles per series coefficient (the six node products, 2 pad)
#define SER_RHO 0.55    // largest |m w| / min|kpar p +- n q| a far row's series is used at
#define GTAB_N 4200     // Gamma(m) table entries (m < GTAB_N), followed by Fact(0..20)
#define RFG_NU (GTAB_N - 21)  // orders nu of the series coefficients 1/(k! Gamma(nu+k+1)) that follow

struct SpecDev {
  int nperp, npar, nmax;
  int acmethod, nfits;
  int anum;           // 1-based species index: the case of distribution_analyt (ACmethod 0)
  int fit_type[ALPSG_DMAXFITS];
  int poly_kind, poly_order, poly_logfit;
  int fit_uniform;    // every fit Maxwellian with pperp-independent ppar parameters
  int fit_kappa;      // ACmethod 1 and every fit a kappa fit (fit_type 2)
  double q, m, dens;
  double dpperp;      // pp(2,2,1) - pp(1,2,1)
  double dppar;       // pp(2,2,2) - pp(2,1,2)  (integrate_res, determine_resonances)
  double dppar_abs;   // abs(...)               (integrate, funct_g, landau_integrate)
  double z;           // T-tensor Bessel scale: kperp/q (kperp_norm) or 1/q
  double int_ee;      // omega-independent ee term of this species
  double poly_log_max, poly_norm1, poly_norm2;
  const double *pperp;  // [nperp+1]
  const double *ppar;   // [npar+1]
  const double *atab;   // [(nmax+1)*(nperp+1)*ATAB_STRIDE]
  const double *nodes;  // [(nmax+1)*(npar-1)*NODE_STRIDE]
  // split form of the node sums for rows every lane integrates over the whole
  // grid (may be NULL): U/den = G/m + C/den with C = G (kpar p +- n q)/m + H;
  // crec[n][ipar] = {kpar p, the six coefficients of +n, of -n}, rsum[n] = the
  // omega-independent sums of the G/m terms over the interior nodes
  const double *crec;   // [(nmax+1)*(npar-1)*CREC_STRIDE]
  const double *rsum;   // [(nmax+1)*RSUM_STRIDE]
  // gcum[n][i] = the G/m terms of rsum summed over ipar = 1..i (gcum[n][0] =
  // 0): a partial-range row's G/m part over nodes a..b is gcum[b] - gcum[a-1]
  const double *gcum;   // [(nmax+1)*npar*GCUM_STRIDE]
  // Far rows (may be NULL).  For sign sg (0: +n, 1: -n) with d_i = kpar p_i
  // +- n q and D = min |d_i| over ipar = 1..npar-1 (serD[2n+sg], 0 if none),
  // the full-range node sum of a row whose every lane has |m w| <= SER_RHO D
  // is the polynomial sum_q ser[n][sg][q][k] z^q in z = m w / D (the geometric
  // series of 1/(m w - d_i) summed over the nodes), truncated where the
  // remainder falls below double rounding
  const double *ser;    // [(nmax+1)*2*SER_Q*SER_STRIDE]
  const double *serD;   // [(nmax+1)*2]
  const double *fitc;   // [(nperp+1)*nfits*FIT_STRIDE]
  // Landau sums of a species whose continuation is one Maxwellian with
  // pperp-independent (u2, u3) (NULL otherwise), per k: lsum[n][6] =
  // {sum h a_k pperp c0, k = 0..2; sum h a_k beta, k = 0..2} over iperp = 0..nperp,
  // with c0 = u1 exp(-pc pperp^2), beta the pperp difference of c0 that
  // landau_integrate takes and h its trapezoid weight; lz_c1 / lz_c2 the
  // smallest |c0| over iperp = 1..nperp-1 / 2..nperp (its zero test)
  const double *lsum;
  double lz_c1, lz_c2;
  const double *polyc;  // [(nperp+1)*(poly_order+1)]
  // closed-form (use_bM) species: NHDS parameters (ALPS_NHDS.f90:163-168)
  int usebm, bm_cold, bm_nrun;
  double bm_Omega, bm_ell, bm_vth, bm_vd, bm_alpha, bm_z, bm_zp;
  const double2 *bmi;   // [2*bm_nrun+1]: (I_n(z) e^-z, dI_n/dz e^-z), n = -nrun..nrun
  // relativistic species (ALPS_fns_rel.f90): the (Gamma, pparbar) grid as
  // derivative_f0_rel leaves it, point index ig*(npb+1)+ip
  int rel, ng, npb, nmax_rel;
  const double *grel, *pbrel, *f0rel;
  const double2 *dfrel;   // (df/dGamma, df/dpparbar)
  const int2 *rowlim;     // [ng+1]: first and last index of each row inside the subluminal cone
  const double *relbes;   // per k: J_n(z) at every point, n = 0..nmax_rel+1: [n][point]
  const double *gam_pp;   // Gamma at the (pperp, ppar) points: [iperp*(npar+1)+ipar]
  const double *df1, *df2;  // df0 on the (pperp, ppar) grid: [iperp*(npar+1)+ipar]
  double rel_pref;        // 2 pi (m/vA)^3 (q vA/(kpar m))
  double rel_zfac, rel_zbar;  // z = zfac * pperpbar; zbar of int_T_rel
  double rel_dg, rel_dpb;
  double rel_gmin, rel_gmax;  // the smallest and largest Gamma of the (pperp, ppar) points
  int rel_gmono;              // that Gamma non-decreasing in iperp along every ppar column
};

// One (request, species) job of the server's batched table launches: the
// Bessel cutoff search (bmax, ntry entries) or the per-k tables (bes, atab,
// nodes for n = 0..nmax).
struct TabJob {
  const double *pperp, *ppar, *df1, *df2;
  double *bes, *atab, *nodes, *bmax;
  double kperp, kpar, q, m;
  int nperp, npar, nmax, ntry;
};

struct ProbDev {
  int nspec, kperp_norm, P, NRI;
  double kperp, kpar, vA, Tlim;
  const double *gtab;  // Gamma(m) of ALPS_fns_rel.f90:1738 for m = 0..GTAB_N-1, then Fact(k), k = 0..20,
                       // then 1/(Fact(k) Gamma(nu+k+1)) at [nu*21 + k], nu = 0..RFG_NU-1
  int nmaxR;           // largest nmax over the relativistic species
  SpecDev sp[ALPSG_DMAXSPEC];
};

// ---------------------------------------------------------------------------
// Small complex helpers (double2 = (re, im)).
// ---------------------------------------------------------------------------
struct cd {
  double re, im;
};
__device__ __forceinline__ cd mk(double r, double i) { return cd{r, i}; }
__device__ __forceinline__ cd operator+(cd a, cd b) { return cd{a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ cd operator-(cd a, cd b) { return cd{a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ cd operator*(cd a, cd b) {
  return cd{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ cd operator*(double s, cd a) { return cd{s * a.re, s * a.im}; }
__device__ __forceinline__ cd operator*(cd a, double s) { return cd{s * a.re, s * a.im}; }
__device__ __forceinline__ cd conjc(cd a) { return cd{a.re, -a.im}; }
__device__ __forceinline__ cd timesI(cd a) { return cd{-a.im, a.re}; }
__device__ __forceinline__ double cabs2(cd a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ double cabsd(cd a) { return hypot(a.re, a.im); }

// Reciprocal of a double: hardware estimate refined by two Newton steps
// (full double precision for |x| in the normal range).
__device__ __forceinline__ double rcp_d(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  return r;
}
// Reciprocal with one Newton step (relative error below 1e-13 for |x| in the
// normal range): used in the node sweep, where it is summed over many nodes.
__device__ __forceinline__ double rcp1_d(double x) {
  double r = __builtin_amdgcn_rcp(x);
  const double e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}
// s / z for complex z
__device__ __forceinline__ cd cscale_inv(double s, cd z) {
  double inv = s * rcp_d(cabs2(z));
  return cd{z.re * inv, -z.im * inv};
}
// a / b, complex (Smith's algorithm)
__device__ __forceinline__ cd cdiv(cd a, cd b) {
  if (fabs(b.re) >= fabs(b.im)) {
    double r = b.im / b.re, d = b.re + b.im * r;
    return cd{(a.re + a.im * r) / d, (a.im - a.re * r) / d};
  } else {
    double r = b.re / b.im, d = b.re * r + b.im;
    return cd{(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
}
__device__ __forceinline__ cd cexpd(cd z) {
  double e = exp(z.re), s, c;
  sincos(z.im, &s, &c);
  return cd{e * c, e * s};
}
__device__ __forceinline__ cd clogd(cd z) { return cd{log(hypot(z.re, z.im)), atan2(z.im, z.re)}; }
// principal square root (branch cut on the negative real axis): with
// t = sqrt((|x| + |z|)/2), sqrt(z) = (t, y/2t) for x >= 0 and (|y|/2t, +-t)
// for x < 0, which avoids the cancellation of the half-angle forms
__device__ __forceinline__ cd csqrtd(cd z) {
  const double r = hypot(z.re, z.im);
  if (r == 0.0) return cd{0.0, 0.0};
  if (z.re >= 0.0) {
    const double t = sqrt(0.5 * (r + z.re));
    return cd{t, z.im / (2.0 * t)};
  }
  const double t = sqrt(0.5 * (r - z.re));
  return cd{fabs(z.im) / (2.0 * t), copysign(t, z.im)};
}
// principal z**p for real p
__device__ __forceinline__ cd cpowr(cd z, double p) {
  if (z.re == 0.0 && z.im == 0.0) return cd{0.0, 0.0};
  cd l = clogd(z);
  return cexpd(cd{p * l.re, p * l.im});
}
