// Batched evaluation of disp(om) (ALPS_fns.f90:252-636) in three launches.
//
//  k_sweep     the trapezoid sums of integrate() (ALPS_fns.f90:799-864) over
//              the ppar nodes of every harmonic row, on the pperp-reduced node
//              tables (alpsg_tables.hip).  The +n and -n rows of a harmonic
//              share one pass over the table; harmonic rows are dealt to the
//              W wavefronts of a workgroup.  Within a wavefront OMB = 64/NCH
//              omegas run side by side, NCH lanes per omega.  With NCH = 1 a
//              lane owns one omega and the node records are wave-uniform
//              scalar loads, so the sweep is pure FP64 VALU work.
//  k_resonant  the principal-value part of integrate_res (ALPS_fns.f90:870-1238)
//              and landau_integrate (ALPS_fns.f90:1327-1452) of the harmonics
//              whose resonance lies in or near the ppar grid; the NCH lanes of
//              an omega split the sub-step ladder and the pperp points.
//  k_assemble  species sums, ee term, n_s q_s normalisation, dielectric tensor,
//              wave tensor and det D (ALPS_fns.f90:479-624).
// Every branch decision (resonance detection, ipar_res, ipar_close, ladder
// positions) is evaluated without FMA contraction in the reference's
// operation order, so the device takes the same branches as the CPU
// restatement.
#include <algorithm>

#include "alpsg_device.h"
#include "alpsg_fit.h"
#include "alpsg_osc.h"

namespace {

constexpr double PI_D = 3.141592653589793238462643383279502884;

typedef const __attribute__((address_space(4))) double *cdptr;  // invariant (scalar-loadable)
typedef const __attribute__((address_space(1))) double *gdptr;  // global (per-lane loads)

struct RowClass {
  int a1, b1, a2, b2;  // ppar index ranges of the trapezoid (integrate calls)
  int pv;              // principal-value part present
  int landau;          // Landau multiplier 0, 1, 2 (full_integrate, ALPS_fns.f90:782-790)
  int found;
  int ipar_res, upper;
  double pr, pim;      // resonant momentum
};

// determine_resonances (ALPS_fns.f90:704-742) and the ipar_res / range logic of
// integrate_res (ALPS_fns.f90:950-1007) for one (n, omega).
template <typename PA>
__device__ __forceinline__ void classify(const SpecDev &S, int P, double kpar, double omr,
                                         double omi, int n, RowClass &rc, PA pa) {
#pragma clang fp contract(off)
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
  rc.a1 = 1;
  rc.b1 = npar - 1;
  bool found = (pa[1] <= pr && pr < pa[npar]);
  const double pos = (double)P;
  if (pr < pa[1] && pr >= (pa[1] - pos * d)) found = true;
  if (pr >= pa[npar - 1] && pr < (pa[npar - 1] + pos * d)) found = true;
  rc.found = found;
  if (!found) return;
  int ipr = 0;
  if (pa[1] <= pr && pr < pa[npar - 1]) {
    // the unique i in [1, npar-2] with pa[i] <= pr < pa[i+1] (strictly increasing grid)
    int k = 1 + (int)((pr - pa[1]) / d);
    k = max(1, min(k, npar - 2));
    while (k > 1 && pa[k] > pr) --k;
    while (k < npar - 2 && pa[k + 1] <= pr) ++k;
    ipr = k;
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

// The last ipar in [1, npar-1] with |ppar(ipar) - x| <= dppar/2, or -1 when
// there is none: the candidates around the nearest grid index are loaded
// together (one memory round trip, not a chain of dependent loads) and tested
// from the highest index down, as the source's loop does.
__device__ __forceinline__ int ipar_close_raw(const SpecDev &S, double x) {
#pragma clang fp contract(off)
  const gdptr pa = (gdptr)S.ppar;
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
  const int khi = min(k0 + 2, npar - 1), klo = max(k0 - 2, 1);
  double v[5];
#pragma unroll
  for (int u = 0; u < 5; ++u) v[u] = (khi - u >= klo) ? pa[khi - u] : 0.0;
  int best = -1;
#pragma unroll
  for (int u = 0; u < 5; ++u)
    if (best < 0 && khi - u >= klo && fabs(v[u] - x) <= half) best = khi - u;
  return best;
}

// funct_g's ipar_close (ALPS_fns.f90:1285-1294): the last ipar in [1,npar-1]
// with |ppar(ipar) - x| <= dppar/2 (0 when there is none), clamped to [2, npar-2].
__device__ __forceinline__ int ipar_close(const SpecDev &S, double x) {
  const int best = ipar_close_raw(S, x);
  return min(max(best, 2), S.npar - 2);
}

// node products [U1, p U1, p^2 U1, U2, U3, p U3], U_a = om*G_a + H_a, at the
// node with ppar p
__device__ __forceinline__ void node6(const double *r, double p, double omr, double omi,
                                      cd u[6]) {
  const cd U1 = mk(omr * r[0] + r[1], omi * r[0]);
  const cd U2 = mk(omr * r[2] + r[3], omi * r[2]);
  const cd U3 = mk(omr * r[4] + r[5], omi * r[4]);
  const double p2 = p * p;
  u[0] = U1;
  u[1] = p * U1;
  u[2] = p2 * U1;
  u[3] = U2;
  u[4] = U3;
  u[5] = p * U3;
}

// pperp-summed funct_g at real ppar x (without the -q/kpar factor), summed into
// acc with the complex weight wt: acc += wt * interp(x)
__device__ __forceinline__ void interp6_acc(const SpecDev &S, const double *__restrict__ rows,
                                            double x, double omr, double omi, cd wt, cd acc[6]) {
  const int ic = ipar_close(S, x);
  cd um[6], u0[6], up[6];
  node6(rows + (size_t)(ic - 2) * NODE_STRIDE, S.ppar[ic - 1], omr, omi, um);
  node6(rows + (size_t)(ic - 1) * NODE_STRIDE, S.ppar[ic], omr, omi, u0);
  node6(rows + (size_t)(ic)*NODE_STRIDE, S.ppar[ic + 1], omr, omi, up);
  const double f = 0.5 * (x - S.ppar[ic]) / S.dppar_abs;
#pragma unroll
  for (int k = 0; k < 6; ++k) acc[k] = acc[k] + wt * (u0[k] + f * (up[k] - um[k]));
}

__device__ __forceinline__ bool czero(cd a) { return a.re == 0.0 && a.im == 0.0; }

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
__device__ constexpr int MI[6] = {0, 1, 2, 0, 0, 1};
__device__ constexpr int MJ[6] = {0, 1, 2, 1, 2, 2};

// T-tensor constant C_k(n) (int_T, ALPS_fns.f90:1659-1703)
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

// one node of one row: acc += (w/den) * U with the b-weights (w is half the
// trapezoid weight: the factor 2 is applied to the row sum)
template <bool UNIT>
__device__ __forceinline__ void node_acc(double w, double dre, double dim, double dim2,
                                         const cd &U1, const cd &U2, const cd &U3, double p,
                                         double p2, cd acc[6]) {
  double inv = rcp1_d(__builtin_fma(dre, dre, dim2));
  if constexpr (!UNIT) inv = w * inv;
  const cd r = mk(dre * inv, -dim * inv);
  const cd u1 = U1 * r, u2 = U2 * r, u3 = U3 * r;
  acc[0] = acc[0] + u1;
  acc[1] = acc[1] + p * u1;
  acc[2] = acc[2] + p2 * u1;
  acc[3] = acc[3] + u2;
  acc[4] = acc[4] + u3;
  acc[5] = acc[5] + p * u3;
}

template <bool STD, bool PAIR, bool FOLD = false>
__device__ __forceinline__ void node_pair(const double *rec, double p, int i, double omr,
                                          double omi,
                                          double mom, double kpar, double nq, double dim,
                                          double dim2, int pl1, int pc1, int pl2, int pc2, int ml1,
                                          int mc1, int ml2, int mc2, cd ap[6], cd am[6],
                                          int npar = 0) {
  const double G1 = rec[0], H1 = rec[1], G2 = rec[2], H2 = rec[3];
  const double G3 = rec[4], H3 = rec[5], p2 = p * p;
  cd U1, U2, U3;
  {
#pragma clang fp contract(off)
    U1 = mk(omr * G1 + H1, omi * G1);
    U2 = mk(omr * G2 + H2, omi * G2);
    U3 = mk(omr * G3 + H3, omi * G3);
  }
  const double base = __builtin_fma(-kpar, p, mom);
  if constexpr (STD && FOLD) {
    // a full-range row with its two end nodes (weight 1/2) in the same loop
    const double w = (i == 1 || i == npar - 1) ? 0.5 : 1.0;
    node_acc<false>(w, base - nq, dim, dim2, U1, U2, U3, p, p2, ap);
    if (PAIR) node_acc<false>(w, base + nq, dim, dim2, U1, U2, U3, p, p2, am);
  } else if constexpr (STD) {
    node_acc<true>(1.0, base - nq, dim, dim2, U1, U2, U3, p, p2, ap);
    if (PAIR) node_acc<true>(1.0, base + nq, dim, dim2, U1, U2, U3, p, p2, am);
  } else {
    const bool inP = ((unsigned)(i - pl1) < (unsigned)pc1) | ((unsigned)(i - pl2) < (unsigned)pc2);
    const bool inM = ((unsigned)(i - ml1) < (unsigned)mc1) | ((unsigned)(i - ml2) < (unsigned)mc2);
    node_acc<false>(inP ? 1.0 : 0.0, inP ? base - nq : 1.0, dim, dim2, U1, U2, U3, p, p2, ap);
    if (PAIR)
      node_acc<false>(inM ? 1.0 : 0.0, inM ? base + nq : 1.0, dim, dim2, U1, U2, U3, p, p2, am);
  }
}

// interior nodes i = 2..npar-2 (weight 2) of the +n row (and the -n row when
// PAIR), lanes of an omega group strided by NCH.  STD: every interior node is in
// range; otherwise node i is in range when it lies strictly inside one of the
// two integrate() intervals of that row.
template <int NCH, bool STD, bool PAIR>
__device__ __forceinline__ void sweep_nodes(cdptr srows, cdptr pa, int npar, int c,
                                            double omr, double omi,
                                            double m, double kpar, double nq, double dim, int pl1,
                                            int pc1, int pl2, int pc2, int ml1, int mc1, int ml2,
                                            int mc2, cd ap[6], cd am[6]) {
  const double dim2 = dim * dim;
  const double mom = m * omr;
  if constexpr (STD && NCH > 1) {
    // several lanes per omega: the end nodes join the lanes' loop (weight 1/2)
    for (int i = 1 + c; i <= npar - 1; i += NCH) {
      double rec[NODE_STRIDE];
#pragma unroll
      for (int t = 0; t < NODE_STRIDE; ++t) rec[t] = srows[(i - 1) * NODE_STRIDE + t];
      node_pair<true, PAIR, true>(rec, pa[i], i, omr, omi, mom, kpar, nq, dim, dim2, 0, 0, 0, 0, 0, 0, 0,
                                  0, ap, am, npar);
    }
    return;
  }
  for (int i = 2 + c; i <= npar - 2; i += NCH) {
    const double G1 = srows[(i - 1) * NODE_STRIDE + 0], H1 = srows[(i - 1) * NODE_STRIDE + 1];
    const double G2 = srows[(i - 1) * NODE_STRIDE + 2], H2 = srows[(i - 1) * NODE_STRIDE + 3];
    const double G3 = srows[(i - 1) * NODE_STRIDE + 4], H3 = srows[(i - 1) * NODE_STRIDE + 5];
    const double p = pa[i], p2 = p * p;
    cd U1, U2, U3;
    {
      // mul + add keeps each instruction to one scalar (table) operand
#pragma clang fp contract(off)
      U1 = mk(omr * G1 + H1, omi * G1);
      U2 = mk(omr * G2 + H2, omi * G2);
      U3 = mk(omr * G3 + H3, omi * G3);
    }
    const double base = __builtin_fma(-kpar, p, mom);
    if constexpr (STD) {
      node_acc<true>(1.0, base - nq, dim, dim2, U1, U2, U3, p, p2, ap);
      if (PAIR) node_acc<true>(1.0, base + nq, dim, dim2, U1, U2, U3, p, p2, am);
    } else {
      const bool inP = ((unsigned)(i - pl1) < (unsigned)pc1) | ((unsigned)(i - pl2) < (unsigned)pc2);
      const bool inM = ((unsigned)(i - ml1) < (unsigned)mc1) | ((unsigned)(i - ml2) < (unsigned)mc2);
      node_acc<false>(inP ? 1.0 : 0.0, inP ? base - nq : 1.0, dim, dim2, U1, U2, U3, p, p2, ap);
      if (PAIR)
        node_acc<false>(inM ? 1.0 : 0.0, inM ? base + nq : 1.0, dim, dim2, U1, U2, U3, p, p2, am);
    }
  }
}

// sweep_nodes with the node records of K consecutive strides loaded before any
// of them is used, so each lane keeps K records in flight; the nodes are
// accumulated in the same order as sweep_nodes.
// sweep_nodes in chunks of K strides whose node records are all loaded before
// the first is used, so each lane keeps K records in flight; a remainder loop
// takes the last partial chunk.  The nodes are accumulated in the same order
// as sweep_nodes.
template <int NCH, bool STD, bool PAIR, int K>
__device__ __forceinline__ void sweep_nodes_pf(cdptr srows, cdptr pa, int npar, int c,
                                               double omr,
                                               double omi, double m, double kpar, double nq,
                                               double dim, int pl1, int pc1, int pl2, int pc2,
                                               int ml1, int mc1, int ml2, int mc2, cd ap[6],
                                               cd am[6]) {
  const double dim2 = dim * dim;
  const double mom = m * omr;
  // full-range rows take their end nodes (weight 1/2) in the same loop
  const int lo = STD ? 1 : 2, hi = STD ? npar - 1 : npar - 2;
  int i0 = lo + c;
  for (; i0 + (K - 1) * NCH <= hi; i0 += NCH * K) {
    double rec[K][NODE_STRIDE];
#pragma unroll
    for (int u = 0; u < K; ++u)
#pragma unroll
      for (int t = 0; t < NODE_STRIDE; ++t) rec[u][t] = srows[(i0 + u * NCH - 1) * NODE_STRIDE + t];
#pragma unroll
    for (int u = 0; u < K; ++u)
      node_pair<STD, PAIR, STD>(rec[u], pa[i0 + u * NCH], i0 + u * NCH, omr, omi, mom, kpar, nq,
                                dim, dim2, pl1, pc1,
                                pl2, pc2, ml1, mc1, ml2, mc2, ap, am, npar);
  }
  for (; i0 <= hi; i0 += NCH) {
    double rec[NODE_STRIDE];
#pragma unroll
    for (int t = 0; t < NODE_STRIDE; ++t) rec[t] = srows[(i0 - 1) * NODE_STRIDE + t];
    node_pair<STD, PAIR, STD>(rec, pa[i0], i0, omr, omi, mom, kpar, nq, dim, dim2, pl1, pc1, pl2,
                              pc2, ml1,
                              mc1, ml2, mc2, ap, am, npar);
  }
}

constexpr double SPLIT_MAX = 256.0;
constexpr int LPA = 1025;  // ppar points a sweep workgroup keeps in LDS (npar <= 1024)
constexpr int LPA_REQ = 2056;  // the request sweep's: every species' grid at once up to this

// The interior nodes of a full-range row pair in the split form (one lane per
// omega, records read as wave-uniform scalars): with r = 1/den,
// acc_k += C_k r for each sign; the G/m terms are the row's rsum, added last.
template <bool PAIR>
__device__ __forceinline__ void split_node(const double *r, double base, double nq, double dim2,
                                           double ndim, cd ap[6], cd am[6]) {
  {
    const double dre = base - nq;
    const double inv = rcp1_d(__builtin_fma(dre, dre, dim2));
    const double rr = dre * inv, ri = ndim * inv;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      ap[k].re = __builtin_fma(r[1 + k], rr, ap[k].re);
      ap[k].im = __builtin_fma(r[1 + k], ri, ap[k].im);
    }
  }
  if (PAIR) {
    const double dre = base + nq;
    const double inv = rcp1_d(__builtin_fma(dre, dre, dim2));
    const double rr = dre * inv, ri = ndim * inv;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      am[k].re = __builtin_fma(r[7 + k], rr, am[k].re);
      am[k].im = __builtin_fma(r[7 + k], ri, am[k].im);
    }
  }
}

// The interior nodes of a full-range row pair in the split form (one lane per
// omega, records read as wave-uniform scalars, the next record loaded while
// the current one is used): with r = 1/den, acc_k += C_k r for each sign; the
// G/m terms are the row's rsum, added last.
template <bool PAIR>
__device__ __forceinline__ void sweep_split(cdptr crow, cdptr rsum, int npar, double omr,
                                            double omi, double m, double nq, double dim,
                                            cd ap[6], cd am[6]) {
  const double dim2 = dim * dim, ndim = -dim;
  const double mom = m * omr;
  double cur[CREC_STRIDE];
#pragma unroll
  for (int t = 0; t < CREC_STRIDE; ++t) cur[t] = crow[1 * CREC_STRIDE + t];  // node 2
  for (int i = 2; i <= npar - 2; ++i) {
    const int in = (i < npar - 2) ? i + 1 : i;
    double nxt[CREC_STRIDE];
#pragma unroll
    for (int t = 0; t < CREC_STRIDE; ++t) nxt[t] = crow[(in - 1) * CREC_STRIDE + t];
    split_node<PAIR>(cur, mom - cur[0], nq, dim2, ndim, ap, am);
#pragma unroll
    for (int t = 0; t < CREC_STRIDE; ++t) cur[t] = nxt[t];
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    ap[k].re += rsum[k];
    if (PAIR) am[k].re += rsum[k];
  }
}

// A far row's node sum for one sign from its series table c (SER_STRIDE
// doubles per degree): acc_k = sum_{q<=J} c_q[k] z^q, J >= 1.  The
// coefficients are real, so the sum runs the second-order recurrence
// b_q = c_q + 2 Re(z) b_{q+1} - |z|^2 b_{q+2} (q = J..1) and ends with
// c_0 + z b_1 - |z|^2 b_2: two real FMAs per coefficient and product.
__device__ __forceinline__ void series_row(cdptr c, int J, double zr, double zi, cd acc[6]) {
  const double r2 = 2.0 * zr, sq = __builtin_fma(zr, zr, zi * zi);
  // b1 = b_{q+1}, b2 = b_{q+2}; two steps per pass, so the pair swaps roles
  // instead of moving
  double b1[6], b2[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    b1[k] = c[J * SER_STRIDE + k];
    b2[k] = 0.0;
  }
  int q = J - 1;
  for (; q >= 2; q -= 2) {
#pragma unroll
    for (int k = 0; k < 6; ++k)
      b2[k] = __builtin_fma(r2, b1[k], __builtin_fma(-sq, b2[k], c[q * SER_STRIDE + k]));
#pragma unroll
    for (int k = 0; k < 6; ++k)
      b1[k] = __builtin_fma(r2, b2[k], __builtin_fma(-sq, b1[k], c[(q - 1) * SER_STRIDE + k]));
  }
  if (q == 1) {
#pragma unroll
    for (int k = 0; k < 6; ++k) {
      const double t = __builtin_fma(r2, b1[k], __builtin_fma(-sq, b2[k], c[SER_STRIDE + k]));
      b2[k] = b1[k];
      b1[k] = t;
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) acc[k] = mk(__builtin_fma(zr, b1[k], __builtin_fma(-sq, b2[k], c[k])), zi * b1[k]);
}

// The smallest degree J >= 1 at which the series' remainder bound
// rho^(J+1) / (1 - rho) is below 2^-54 (rho <= SER_RHO < 1).
__device__ __forceinline__ int series_degree(double rho) {
  const double thr = 0x1p-54 * (1.0 - rho);
  double t = rho * rho;
  int J = 1;
  while (t > thr && J < SER_Q - 1) {
    t *= rho;
    ++J;
  }
  return J;
}

// The interior nodes of a partial-range row pair in the split form (one lane
// per omega, records read as wave-uniform scalars): every node i = 2..npar-2
// is visited and its reciprocal zeroed, by a select, outside the lane's
// interior ranges (an excluded node may sit on the resonance), so the C/den
// sums cover exactly the lane's ranges; the G/m sums over the same ranges
// come from the cumulative table, gcum[b] - gcum[a-1] per range.
template <bool PAIR>
__device__ __forceinline__ void sweep_split_masked(cdptr crow, const double *gcum, int npar,
                                                   double omr, double m, double nq, double dim,
                                                   int pl1, int pc1, int pl2, int pc2, int ml1,
                                                   int mc1, int ml2, int mc2, cd ap[6], cd am[6]) {
  const double dim2 = dim * dim, ndim = -dim;
  const double mom = m * omr;
  double cur[CREC_STRIDE];
#pragma unroll
  for (int t = 0; t < CREC_STRIDE; ++t) cur[t] = crow[1 * CREC_STRIDE + t];  // node 2
  for (int i = 2; i <= npar - 2; ++i) {
    const int in = (i < npar - 2) ? i + 1 : i;
    double nxt[CREC_STRIDE];
#pragma unroll
    for (int t = 0; t < CREC_STRIDE; ++t) nxt[t] = crow[(in - 1) * CREC_STRIDE + t];
    const double base = mom - cur[0];
    {
      const bool inP = ((unsigned)(i - pl1) < (unsigned)pc1) | ((unsigned)(i - pl2) < (unsigned)pc2);
      const double dre = base - nq;
      double inv = rcp1_d(__builtin_fma(dre, dre, dim2));
      inv = inP ? inv : 0.0;
      const double rr = dre * inv, ri = ndim * inv;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        ap[k].re = __builtin_fma(cur[1 + k], rr, ap[k].re);
        ap[k].im = __builtin_fma(cur[1 + k], ri, ap[k].im);
      }
    }
    if (PAIR) {
      const bool inM = ((unsigned)(i - ml1) < (unsigned)mc1) | ((unsigned)(i - ml2) < (unsigned)mc2);
      const double dre = base + nq;
      double inv = rcp1_d(__builtin_fma(dre, dre, dim2));
      inv = inM ? inv : 0.0;
      const double rr = dre * inv, ri = ndim * inv;
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        am[k].re = __builtin_fma(cur[7 + k], rr, am[k].re);
        am[k].im = __builtin_fma(cur[7 + k], ri, am[k].im);
      }
    }
#pragma unroll
    for (int t = 0; t < CREC_STRIDE; ++t) cur[t] = nxt[t];
  }
  auto range = [&](int pl, int pc, cd acc[6]) {
    if (pc > 0) {
      const double *hi = gcum + (size_t)(pl + pc - 1) * GCUM_STRIDE;
      const double *lo = gcum + (size_t)(pl - 1) * GCUM_STRIDE;
#pragma unroll
      for (int k = 0; k < 6; ++k) acc[k].re += hi[k] - lo[k];
    }
  };
  range(pl1, pc1, ap);
  range(pl2, pc2, ap);
  if (PAIR) {
    range(ml1, mc1, am);
    range(ml2, mc2, am);
  }
}

// endpoint contribution (weight 1) of node e for a row with denominator shift nq
__device__ __forceinline__ void endpoint_acc(const double *__restrict__ rows, cdptr pa,
                                             int e, int npar,
                                             double m, double kpar, double omr, double omi,
                                             double nq, cd acc[6]) {
  if (e < 1 || e > npar - 1) return;
  const double *r = rows + (size_t)(e - 1) * NODE_STRIDE;
  const cd U1 = mk(omr * r[0] + r[1], omi * r[0]);
  const cd U2 = mk(omr * r[2] + r[3], omi * r[2]);
  const cd U3 = mk(omr * r[4] + r[5], omi * r[4]);
  const double dim = m * omi;
  const double p = pa[e];
  node_acc<false>(0.5, (m * omr - kpar * p) - nq, dim, dim * dim, U1, U2, U3, p, p * p, acc);
}

}  // namespace

// ---------------------------------------------------------------------------
// k_sweep: trapezoid node sums of every harmonic row pair.
// The harmonic row pairs are dealt round-robin to the TW = W * gridDim.y
// wavefronts that share an omega block (W per workgroup, gridDim.y workgroups),
// so a small batch still spreads over many compute units.
// chiA[iom][TW][nspec][6] gets each wavefront's partial chi per species (before
// the n_s q_s factor); lowA[iom][nspec][3][6] (if non-null) the n = -1, 0, 1
// rows.
// ---------------------------------------------------------------------------
template <int NCH, int W>
// Four wavefronts per SIMD (a few spilled registers; three by its register
// count): with its ppar grid in LDS sized to the grid, the one-lane shape is
// register-, not LDS-bound, and the fourth wavefront hides more of the node
// loads' latency (the map batch's k_sweep 0.512 -> 0.484 ms)
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(4))) void k_sweep(const ProbDev *__restrict__ probp, int nom,
                                                  const double2 *__restrict__ om_in,
                                                  const int *__restrict__ active,
                                                  double2 *__restrict__ chiA,
                                                  double2 *__restrict__ lowA,
                                                  const int *__restrict__ pidx,
                                                  unsigned long long *__restrict__ work,
                                                  int lpa_n) {
  constexpr int OMB = 64 / NCH;
  __shared__ double lchi[W][12][64];
  // the ppar grid for the resonance classification: lpa_n doubles of dynamic
  // LDS (the grid's npar + 1, 0 if it does not fit LPA), so that a small grid
  // does not cost the occupancy of the largest
  extern __shared__ double lpa[];
  // one problem per launch, or (pidx, one omega per workgroup: NCH = 64) the
  // problem of each omega
  const ProbDev &Pd = (NCH == 64 && pidx != nullptr) ? probp[pidx[blockIdx.x]] : *probp;
  const int lane = threadIdx.x & 63;
  const int TW = W * gridDim.y;
  const int wv = blockIdx.y * W + (threadIdx.x >> 6);
  const int g = lane / NCH, c = lane % NCH;
  const int iom = blockIdx.x * OMB + g;
  const bool valid = (iom < nom) && (active == nullptr || active[iom] != 0);
  const double2 ov = valid ? om_in[iom] : make_double2(1.0, 0.5);
  const double omr = ov.x, omi = ov.y;
  const int nspec = Pd.nspec;
  const double kpar = Pd.kpar;
  int rowbase = 0;
  // work counts (counting launches only): omega-weighted node units (species,
  // harmonic, ppar node) of node sums, far-row signs and their series degrees
  unsigned w_nodes = 0, w_ser = 0, w_deg = 0;
  for (int s = 0; s < nspec; ++s) {
    const SpecDev &S = Pd.sp[s];
    const int npar = S.npar;
    const double m = S.m;
    // the largest |m w| of the wavefront's omegas (far rows compare it with D)
    double mw2 = (m * m) * __builtin_fma(omr, omr, omi * omi);
    if (NCH == 1 && S.ser != nullptr) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) mw2 = fmax(mw2, __shfl_xor(mw2, o, 64));
    }
    const double mwmax = sqrt(mw2);
    // the species' ppar grid in LDS for the resonance classification
    const double *pa = S.ppar;
    if (S.nmax >= 0 && npar + 1 <= lpa_n) {
      __syncthreads();
      for (int t = threadIdx.x; t <= npar; t += blockDim.x) lpa[t] = S.ppar[t];
      __syncthreads();
      pa = lpa;
    }
    // 2 pi dpperp dppar / 4 * q, times 2: node sums carry half the trapezoid weights
    const double scaleN = 2.0 * (2.0 * PI_D * S.dpperp * S.dppar_abs * 0.25 * S.q);
    // the species' running chi lives in LDS ([component][lane]) between rows
    double *lc = &lchi[threadIdx.x >> 6][0][lane];
#pragma unroll
    for (int t = 0; t < 12; ++t) lc[64 * t] = 0.0;
    // a row pair's node sums (before the row's T constants) into the running
    // chi in LDS, and the n = -1, 0, 1 rows into lowA
    auto finish_row = [&](int nn, const cd *xp, const cd *xm) {
      cd vp[6], vm[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        vp[k] = (scaleN * xp[ACC_OF_MODE[k]]) * tconst(Pd, S, nn, k);
        vm[k] = (scaleN * xm[ACC_OF_MODE[k]]) * tconst(Pd, S, -nn, k);
      }
      if (nn == 0) {
        lc[64 * 2] += vp[1].re;
        lc[64 * 3] += vp[1].im;
        lc[64 * 4] += vp[2].re;
        lc[64 * 5] += vp[2].im;
        lc[64 * 10] += vp[5].re;
        lc[64 * 11] += vp[5].im;
      } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const cd t = (mk(lc[64 * (2 * k)], lc[64 * (2 * k + 1)]) + vp[k]) + vm[k];
          lc[64 * (2 * k)] = t.re;
          lc[64 * (2 * k + 1)] = t.im;
        }
      }
      if (lowA != nullptr && valid && c == 0 && nn <= 1) {
        double2 *o = lowA + ((size_t)iom * nspec + s) * 18;
        if (nn == 0) {
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            const bool use = (k == 1 || k == 2 || k == 5);
            o[6 * 1 + k] = use ? make_double2(vp[k].re, vp[k].im) : make_double2(0.0, 0.0);
          }
        } else {
#pragma unroll
          for (int k = 0; k < 6; ++k) {
            o[6 * 2 + k] = make_double2(vp[k].re, vp[k].im);
            o[6 * 0 + k] = make_double2(vm[k].re, vm[k].im);
          }
        }
      }
    };
    for (int nn = 0; nn <= S.nmax; ++nn) {
      if (((rowbase + nn) % TW) != wv) continue;
      const double *rows = S.nodes + (size_t)nn * (npar - 1) * NODE_STRIDE;
      const cdptr srows = (cdptr)rows;
      const cdptr spa = (cdptr)S.ppar;  // the nodes' ppar values
      RowClass rp, rm;
      classify(S, Pd.P, kpar, omr, omi, nn, rp, pa);
      if (nn > 0) classify(S, Pd.P, kpar, omr, omi, -nn, rm, pa);
      const double nq = (1.0 * nn) * S.q;
      // interior (weight-2) ranges (a, b) of the two integrate() calls per sign
      const int pl1 = rp.a1 + 1, pc1 = max(rp.b1 - rp.a1 - 1, 0);
      const int pl2 = rp.a2 + 1, pc2 = max(rp.b2 - rp.a2 - 1, 0);
      const int ml1 = nn > 0 ? rm.a1 + 1 : 0, mc1 = nn > 0 ? max(rm.b1 - rm.a1 - 1, 0) : 0;
      const int ml2 = nn > 0 ? rm.a2 + 1 : 0, mc2 = nn > 0 ? max(rm.b2 - rm.a2 - 1, 0) : 0;
      cd ap[6], am[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) ap[k] = am[k] = mk(0.0, 0.0);
      const double dim = m * omi;
      // rows whose trapezoid covers the whole grid in every lane take the
      // uniform-weight loop (no per-node range tests)
      const bool stdp = (rp.a2 < 0) && rp.a1 == 1 && rp.b1 == npar - 1;
      const bool stdm = (nn == 0) || ((rm.a2 < 0) && rm.a1 == 1 && rm.b1 == npar - 1);
      // the split form sums G/m and C/den, which cancel to U/den when |den| is
      // large against m|w|: it is used where the cancellation factor
      // (kpar max|p| + |n q|) / (m |w|) stays below SPLIT_MAX for every lane
      const bool split_ok =
          NCH == 1 && S.crec != nullptr &&
          fabs(kpar) * fmax(fabs(pa[0]), fabs(pa[npar])) + fabs(nq) <=
              SPLIT_MAX * m * sqrt(omr * omr + omi * omi);
      const bool std_row = __all(stdp && stdm);
      // a far row (every lane's |m w| within SER_RHO of both signs' D) takes
      // its series, in arrays of its own, and is done
      if (NCH == 1 && S.ser != nullptr && std_row) {
        const double Dp = S.serD[2 * nn], Dm = (nn > 0) ? S.serD[2 * nn + 1] : Dp;
        if (Dp > 0.0 && Dm > 0.0 && mwmax <= SER_RHO * Dp && mwmax <= SER_RHO * Dm) {
          const cdptr cp = (cdptr)(S.ser + (size_t)(2 * nn) * SER_Q * SER_STRIDE);
          cd sp[6], sm[6];
          // (the degrees are wave-uniform: the loop and its table loads stay scalar)
          const int Jp = __builtin_amdgcn_readfirstlane(series_degree(mwmax / Dp));
          series_row(cp, Jp, (m * omr) / Dp, (m * omi) / Dp, sp);
          w_ser += 1;
          w_deg += Jp;
          if (nn > 0) {
            const int Jm = __builtin_amdgcn_readfirstlane(series_degree(mwmax / Dm));
            series_row(cp + SER_Q * SER_STRIDE, Jm, (m * omr) / Dm, (m * omi) / Dm, sm);
            w_ser += 1;
            w_deg += Jm;
          } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) sm[k] = mk(0.0, 0.0);
          }
          finish_row(nn, sp, sm);
          continue;
        }
      }
      if (std_row) {
        w_nodes += (nn > 0 ? 2 : 1) * (npar - 1);
        if (__all(split_ok)) {
          const cdptr crow = (cdptr)(S.crec + (size_t)nn * (npar - 1) * CREC_STRIDE);
          const cdptr rs = (cdptr)(S.rsum + (size_t)nn * RSUM_STRIDE);
          if (nn > 0)
            sweep_split<true>(crow, rs, npar, omr, omi, m, nq, dim, ap, am);
          else
            sweep_split<false>(crow, rs, npar, omr, omi, m, nq, dim, ap, am);
        } else if (nn > 0) {
          sweep_nodes<NCH, true, true>(srows, spa, npar, c, omr, omi, m, kpar, nq, dim, 0, 0, 0, 0, 0, 0,
                                       0, 0, ap, am);
        } else {
          sweep_nodes<NCH, true, false>(srows, spa, npar, c, omr, omi, m, kpar, nq, dim, 0, 0, 0, 0, 0,
                                        0, 0, 0, ap, am);
        }
      } else {
        w_nodes += (nn > 0 ? 2 : 1) * (npar - 1);
        if (NCH == 1 && S.gcum != nullptr && __all(split_ok)) {
          const cdptr crow = (cdptr)(S.crec + (size_t)nn * (npar - 1) * CREC_STRIDE);
          const double *gc = S.gcum + (size_t)nn * npar * GCUM_STRIDE;
          if (nn > 0)
            sweep_split_masked<true>(crow, gc, npar, omr, m, nq, dim, pl1, pc1, pl2, pc2, ml1, mc1,
                                     ml2, mc2, ap, am);
          else
            sweep_split_masked<false>(crow, gc, npar, omr, m, nq, dim, pl1, pc1, pl2, pc2, ml1, mc1,
                                      ml2, mc2, ap, am);
        } else {
          sweep_nodes<NCH, false, true>(srows, spa, npar, c, omr, omi, m, kpar, nq, dim, pl1, pc1, pl2,
                                        pc2, ml1, mc1, ml2, mc2, ap, am);
        }
      }
      // (with several lanes per omega a full-range row took its end nodes in the sweep)
      if (c == 0 && !(NCH > 1 && std_row)) {
        endpoint_acc(rows, spa, rp.a1, npar, m, kpar, omr, omi, nq, ap);
        endpoint_acc(rows, spa, rp.b1, npar, m, kpar, omr, omi, nq, ap);
        if (rp.a2 > 0) {
          endpoint_acc(rows, spa, rp.a2, npar, m, kpar, omr, omi, nq, ap);
          endpoint_acc(rows, spa, rp.b2, npar, m, kpar, omr, omi, nq, ap);
        }
        if (nn > 0) {
          endpoint_acc(rows, spa, rm.a1, npar, m, kpar, omr, omi, -nq, am);
          endpoint_acc(rows, spa, rm.b1, npar, m, kpar, omr, omi, -nq, am);
          if (rm.a2 > 0) {
            endpoint_acc(rows, spa, rm.a2, npar, m, kpar, omr, omi, -nq, am);
            endpoint_acc(rows, spa, rm.b2, npar, m, kpar, omr, omi, -nq, am);
          }
        }
      }
      if (NCH > 1) {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          ap[k] = gsumc<NCH>(ap[k]);
          if (nn > 0) am[k] = gsumc<NCH>(am[k]);
        }
      }
      finish_row(nn, ap, am);
    }
    rowbase += S.nmax + 1;
    if (valid && c == 0) {
      // omega-fastest layout: chiA[((wv * nspec + s) * 6 + k) * nom + iom]
      double2 *o = chiA + ((size_t)(wv * nspec + s) * 6) * nom + iom;
#pragma unroll
      for (int k = 0; k < 6; ++k) o[(size_t)k * nom] = make_double2(lc[64 * (2 * k)], lc[64 * (2 * k + 1)]);
    }
  }
  if (work != nullptr) {
    // each omega's counts (lane c = 0 of its group), summed over the wavefront
    const bool own = valid && c == 0;
    unsigned long long a = own ? w_nodes : 0u, b = own ? w_ser : 0u, d = own ? w_deg : 0u;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      a += __shfl_xor(a, o, 64);
      b += __shfl_xor(b, o, 64);
      d += __shfl_xor(d, o, 64);
    }
    if (lane == 0) {
      atomicAdd(work + 0, a);
      atomicAdd(work + 1, b);
      atomicAdd(work + 2, d);
    }
  }
}

// ---------------------------------------------------------------------------
// k_sweep_req: the node sums of k_sweep for batches in which every omega has
// its own problem (the root server: one workgroup per omega, descriptor
// probp[pidx[omega]]).  Such a batch reads each problem's node table once, so
// the sweep streams the tables from memory; here a wavefront works on
// RPW = 64 / NCH harmonic row pairs at once, NCH lanes per row, each lane
// keeping K node records in flight.  Row pair nn of species s goes to row slot
// (rowbase + nn) % (RPW * TW).  A lane multiplies its own node sums of a row by
// the row's T constants and keeps the running chi of its rows in registers
// (sum_rows T_row sum_nodes = sum_lanes sum_rows T_row sum_lane's nodes), so
// the lanes are summed once per species, in a fixed order, before the
// wavefront writes its partial chiA slot.  Every omega's result depends only on
// its own problem and the launch shape.
// ---------------------------------------------------------------------------
// One workgroup of k_sweep_req: omega bx, workgroup by of the omega's gy.
// Each wavefront wv writes partial slot wv of chiA (W gy slots per omega), or
// with RED the workgroup sums its wavefronts in order and writes slot by (gy
// slots per omega).
template <int NCH, int W, int K, bool RED>
__device__ __forceinline__ void sweep_req_body(const ProbDev *__restrict__ probp, int nom,
                                               const double2 *__restrict__ om_in,
                                               double2 *__restrict__ chiA,
                                               const int *__restrict__ pidx, int bx, int by, int gy) {
  constexpr int RPW = 64 / NCH;
  const int iom = bx;
  if (iom >= nom) return;
  const ProbDev &Pd = probp[pidx[iom]];
  const int lane = threadIdx.x & 63;
  const int TW = W * gy;
  const int wv = by * W + (threadIdx.x >> 6);
  const int g = lane / NCH, c = lane % NCH;
  const int slot = wv * RPW + g, TS = TW * RPW;
  const double2 ov = om_in[iom];
  const double omr = ov.x, omi = ov.y;
  const int nspec = Pd.nspec;
  const double kpar = Pd.kpar;
  __shared__ double lpa[LPA_REQ];
  __shared__ double red[ALPSG_DMAXSPEC][W][12];
  // the species' ppar grids (for the resonance classification) go to LDS all
  // at once when they fit, and the wavefronts' partial sums are combined once
  // for every species at the end: two barriers per launch instead of four per
  // species (the same sums in the same order)
  int tot = 0;
  for (int s = 0; s < nspec; ++s)
    if (Pd.sp[s].nmax >= 0) tot += Pd.sp[s].npar + 1;
  const bool all = tot <= LPA_REQ;
  if (all) {
    int o = 0;
    for (int s = 0; s < nspec; ++s) {
      const SpecDev &S = Pd.sp[s];
      if (S.nmax < 0) continue;
      for (int t = threadIdx.x; t <= S.npar; t += blockDim.x) lpa[o + t] = S.ppar[t];
      o += S.npar + 1;
    }
    __syncthreads();
  }
  int rowbase = 0, lo = 0;
  for (int s = 0; s < nspec; ++s) {
    const SpecDev &S = Pd.sp[s];
    const int npar = S.npar;
    const double m = S.m;
    const double *pa = S.ppar;
    if (S.nmax >= 0 && all) {
      pa = lpa + lo;
      lo += npar + 1;
    } else if (S.nmax >= 0 && npar + 1 <= LPA_REQ) {
      __syncthreads();
      for (int t = threadIdx.x; t <= npar; t += blockDim.x) lpa[t] = S.ppar[t];
      __syncthreads();
      pa = lpa;
    }
    const double scaleN = 2.0 * (2.0 * PI_D * S.dpperp * S.dppar_abs * 0.25 * S.q);
    double run[12];
#pragma unroll
    for (int t = 0; t < 12; ++t) run[t] = 0.0;
    // first row pair of this slot, then every TS-th
    int nn = ((slot - rowbase) % TS + TS) % TS;
    for (; nn <= S.nmax; nn += TS) {
      const double *rows = S.nodes + (size_t)nn * (npar - 1) * NODE_STRIDE;
      const cdptr srows = (cdptr)rows;
      const cdptr spa = (cdptr)S.ppar;  // the nodes' ppar values
      RowClass rp, rm;
      classify(S, Pd.P, kpar, omr, omi, nn, rp, pa);
      if (nn > 0) classify(S, Pd.P, kpar, omr, omi, -nn, rm, pa);
      const double nq = (1.0 * nn) * S.q;
      const int pl1 = rp.a1 + 1, pc1 = max(rp.b1 - rp.a1 - 1, 0);
      const int pl2 = rp.a2 + 1, pc2 = max(rp.b2 - rp.a2 - 1, 0);
      const int ml1 = nn > 0 ? rm.a1 + 1 : 0, mc1 = nn > 0 ? max(rm.b1 - rm.a1 - 1, 0) : 0;
      const int ml2 = nn > 0 ? rm.a2 + 1 : 0, mc2 = nn > 0 ? max(rm.b2 - rm.a2 - 1, 0) : 0;
      cd ap[6], am[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) ap[k] = am[k] = mk(0.0, 0.0);
      const double dim = m * omi;
      const bool stdp = (rp.a2 < 0) && rp.a1 == 1 && rp.b1 == npar - 1;
      const bool stdm = (nn == 0) || ((rm.a2 < 0) && rm.a1 == 1 && rm.b1 == npar - 1);
      if (stdp && stdm) {
        if (nn > 0)
          sweep_nodes_pf<NCH, true, true, K>(srows, spa, npar, c, omr, omi, m, kpar, nq, dim, 0, 0, 0, 0, 0, 0,
                                       0, 0, ap, am);
        else
          sweep_nodes_pf<NCH, true, false, K>(srows, spa, npar, c, omr, omi, m, kpar, nq, dim, 0, 0, 0, 0, 0,
                                        0, 0, 0, ap, am);
      } else {
        sweep_nodes_pf<NCH, false, true, K>(srows, spa, npar, c, omr, omi, m, kpar, nq, dim, pl1, pc1, pl2, pc2,
                                      ml1, mc1, ml2, mc2, ap, am);
      }
      // (a full-range row took its end nodes in the sweep)
      if (c == 0 && !(stdp && stdm)) {
        endpoint_acc(rows, spa, rp.a1, npar, m, kpar, omr, omi, nq, ap);
        endpoint_acc(rows, spa, rp.b1, npar, m, kpar, omr, omi, nq, ap);
        if (rp.a2 > 0) {
          endpoint_acc(rows, spa, rp.a2, npar, m, kpar, omr, omi, nq, ap);
          endpoint_acc(rows, spa, rp.b2, npar, m, kpar, omr, omi, nq, ap);
        }
        if (nn > 0) {
          endpoint_acc(rows, spa, rm.a1, npar, m, kpar, omr, omi, -nq, am);
          endpoint_acc(rows, spa, rm.b1, npar, m, kpar, omr, omi, -nq, am);
          if (rm.a2 > 0) {
            endpoint_acc(rows, spa, rm.a2, npar, m, kpar, omr, omi, -nq, am);
            endpoint_acc(rows, spa, rm.b2, npar, m, kpar, omr, omi, -nq, am);
          }
        }
      }
      // this lane's share of the row pair, times the row's T constants; the
      // lanes are summed once per species below
      cd vp[6], vm[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) {
        vp[k] = (scaleN * ap[ACC_OF_MODE[k]]) * tconst(Pd, S, nn, k);
        vm[k] = (scaleN * am[ACC_OF_MODE[k]]) * tconst(Pd, S, -nn, k);
      }
      if (nn == 0) {
        run[2] += vp[1].re;
        run[3] += vp[1].im;
        run[4] += vp[2].re;
        run[5] += vp[2].im;
        run[10] += vp[5].re;
        run[11] += vp[5].im;
      } else {
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const cd t = (mk(run[2 * k], run[2 * k + 1]) + vp[k]) + vm[k];
          run[2 * k] = t.re;
          run[2 * k + 1] = t.im;
        }
      }
    }
    rowbase += S.nmax + 1;
    // the lanes of the wavefront, combined pairwise in a fixed order, then
    // the wavefronts of the workgroup in order: one partial per workgroup
#pragma unroll
    for (int o = 1; o < 64; o <<= 1)
#pragma unroll
      for (int t = 0; t < 12; ++t) run[t] += __shfl_xor(run[t], o, 64);
    if constexpr (RED) {
      if (all) {
        if (lane == 0)
#pragma unroll
          for (int t = 0; t < 12; ++t) red[s][threadIdx.x >> 6][t] = run[t];
      } else {
        __syncthreads();
        if (lane == 0)
#pragma unroll
          for (int t = 0; t < 12; ++t) red[0][threadIdx.x >> 6][t] = run[t];
        __syncthreads();
        if (threadIdx.x < 12) {
          double v = red[0][0][threadIdx.x];
#pragma unroll
          for (int w = 1; w < W; ++w) v += red[0][w][threadIdx.x];
          double *o = reinterpret_cast<double *>(
              chiA + ((size_t)(by * nspec + s) * 6 + (threadIdx.x >> 1)) * nom + iom);
          o[threadIdx.x & 1] = v;
        }
      }
    } else if (lane == 0) {
      double2 *o = chiA + ((size_t)(wv * nspec + s) * 6) * nom + iom;
#pragma unroll
      for (int k = 0; k < 6; ++k) o[(size_t)k * nom] = make_double2(run[2 * k], run[2 * k + 1]);
    }
  }
  if constexpr (RED) {
    if (all) {
      // every species' partials of the wavefronts, summed in wavefront order
      __syncthreads();
      for (int t = threadIdx.x; t < 12 * nspec; t += blockDim.x) {
        const int s = t / 12, c = t % 12;
        double v = red[s][0][c];
#pragma unroll
        for (int w = 1; w < W; ++w) v += red[s][w][c];
        double *o = reinterpret_cast<double *>(chiA + ((size_t)(by * nspec + s) * 6 + (c >> 1)) * nom + iom);
        o[c & 1] = v;
      }
    }
  }
}

template <int NCH, int W, int K, bool RED>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(2))) void k_sweep_req(const ProbDev *__restrict__ probp, int nom,
                                                      const double2 *__restrict__ om_in,
                                                      double2 *__restrict__ chiA,
                                                      const int *__restrict__ pidx) {
  sweep_req_body<NCH, W, K, RED>(probp, nom, om_in, chiA, pidx, blockIdx.x, blockIdx.y, gridDim.y);
}

// ---------------------------------------------------------------------------
// k_resonant: principal-value and Landau parts of the resonant harmonics of
// species blockIdx.y; blockIdx.z selects a block of the candidate harmonics
// (harmonic n handled by block (n - nlo) % gridDim.z).
// chiB[iom][nspec][RB][6], lowB[iom][nspec][RB][3][6] (if non-null).
//
// The principal-value sums of integrate_res evaluate funct_g, a linear
// interpolation of node values (ALPS_fns.f90:1316-1317), at ~2 n_resonance_interval
// ladder points.  Being linear in the node values, the whole sum equals
// sum_i c_i u(i) over the few nodes i next to the resonance, where the complex
// coefficient c_i collects the ladder weights times the interpolation weights.
// The lanes of an omega walk contiguous blocks of the ladder, keep the three
// coefficients of the current interpolation cell in registers, and flush them
// into a per-omega LDS window when the cell changes; the node records are then
// read once per row.
// ---------------------------------------------------------------------------
namespace {

constexpr int PV_WIN = 40;  // nodes ipar_res-P-3 .. ipar_res+P+4 (P <= 16)

struct CellCache {
  int ic;
  cd cm, c0, cp;
};

// ipar_close along a monotone ladder: after the first point only the cells
// next to the previous one can be the last index within dppar/2 of x, so the
// candidates are checked in descending index order (the same answer as the full
// search); a miss falls back to the full search.
struct Cursor {
  int ic;  // last result, -1 before the first point
};

__device__ __forceinline__ int ipar_close_step(const SpecDev &S, Cursor &cu, double x, int dir) {
#pragma clang fp contract(off)
  if (cu.ic >= 0) {
    const double *pa = S.ppar;
    const double half = 0.5 * S.dppar_abs;
    const int npar = S.npar;
    int best = -1;
    // the candidates' grid values are loaded together (every index is inside
    // the grid), then the highest match is taken, as the descending search does
    if (dir > 0) {
      // x increases: candidates ic+3 .. ic (descending; two indices can tie)
      double v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) v[u] = (cu.ic + u <= npar - 1) ? pa[cu.ic + u] : 0.0;
#pragma unroll
      for (int u = 3; u >= 0; --u)
        if (best < 0 && cu.ic + u <= npar - 1 && fabs(v[u] - x) <= half) best = cu.ic + u;
    } else {
      // x decreases: candidates ic, ic-1, ic-2 (descending)
      double v[3];
#pragma unroll
      for (int u = 0; u < 3; ++u) v[u] = (cu.ic - u >= 1) ? pa[cu.ic - u] : 0.0;
#pragma unroll
      for (int u = 0; u < 3; ++u)
        if (best < 0 && cu.ic - u >= 1 && fabs(v[u] - x) <= half) best = cu.ic - u;
    }
    if (best >= 0) {
      int r = best;
      if (r >= npar - 1) r = npar - 2;
      if (r <= 1) r = 2;
      // the cursor keeps the unclamped index so that the next step searches near x
      cu.ic = best;
      return r;
    }
  }
  // full search (first point or a miss); the cursor keeps the unclamped index
  const int best = ipar_close_raw(S, x);
  cu.ic = best;
  return min(max(best, 2), S.npar - 2);
}

__device__ __forceinline__ void lds_cadd(double *w, int t, cd v) {
  atomicAdd(w + 2 * t, v.re);
  atomicAdd(w + 2 * t + 1, v.im);
}

__device__ __forceinline__ void cache_flush(CellCache &cc, double *win, int lo) {
  if (cc.ic >= 0) {
    lds_cadd(win, cc.ic - 1 - lo, cc.cm);
    lds_cadd(win, cc.ic - lo, cc.c0);
    lds_cadd(win, cc.ic + 1 - lo, cc.cp);
  }
  cc.ic = -1;
  cc.cm = cc.c0 = cc.cp = mk(0.0, 0.0);
}

// add wt * interp(x) to the coefficient window (through the cell cache)
__device__ __forceinline__ void coef_add(const SpecDev &S, CellCache &cc, Cursor &cu, double *win,
                                         int lo, double x, int dir, cd wt, double hinv) {
  const int ic = ipar_close_step(S, cu, x, dir);
  if (ic != cc.ic) {
    cache_flush(cc, win, lo);
    cc.ic = ic;
  }
  const double f = (x - S.ppar[ic]) * hinv;  // funct_g's 0.5 (x - p_c) / dppar
  cc.c0 = cc.c0 + wt;
  cc.cp = cc.cp + f * wt;
  cc.cm = cc.cm - f * wt;
}

// A monotone ladder's interpolation cells without per-point grid loads: the
// last grid index within dppar/2 of x (ipar_close's answer before clamping)
// can only move to the neighbour in the ladder's direction when a step is
// shorter than 0.9 dppar (fast; the grid is uniform), so the neighbour's grid
// value is kept in a register and each point makes the same |pa[k] - x| <=
// dppar/2 tests as the full search.  A point that matches neither index (a
// rounding gap between two cells), and every point of a coarser ladder, takes
// the full search.
//
// The ladder's weights are r_j = w_j / (t_j - i pim), t_j = x_j - pr, and a
// point adds r_j to cell c's c0 and f_j r_j to its cp, f_j = (x_j - p_c) /
// (2 dppar) (funct_g's interpolation fraction; the cell below takes -cp).
// Since t r = w + i pim r, the cell's sum of f r is (W + (pr - p_c + i pim) S)
// / (2 dppar) with S = sum r_j and W = sum w_j: a point adds r_j and w_j, and
// the cell's coefficients are formed when it is flushed.  The mirror ladder
// x = 2 pr - x_j (weights -conj(r_j)) flushes c0 = -conj(S), cp = (W - (pr -
// p_c + i pim) conj(S)) / (2 dppar).
struct Walk {
  int k;           // the last index within dppar/2 of the current x, -1: unknown
  double pk, pn;   // pa[k] and pa[k + dir]
  int r;           // the interpolation cell: k clamped to [2, npar-2], -1: none yet
  double pc;       // pa[r]
  cd s;            // sum of r_j over the cell's points
  double wsum;     // sum of w_j
};

template <bool MIRROR>
__device__ __forceinline__ void walk_flush(Walk &w, double *win, int lo, double pr, double pim,
                                           double hinv) {
  if (w.r >= 0) {
    const cd S = MIRROR ? mk(w.s.re, -w.s.im) : w.s;
    const cd z = mk(pr - w.pc, pim);
    const cd zS = z * S;
    const cd c0 = MIRROR ? mk(-S.re, -S.im) : S;
    const cd cp = MIRROR ? mk(hinv * (w.wsum - zS.re), -hinv * zS.im)
                         : mk(hinv * (zS.re + w.wsum), hinv * zS.im);
    lds_cadd(win, w.r - 1 - lo, mk(-cp.re, -cp.im));
    lds_cadd(win, w.r - lo, c0);
    lds_cadd(win, w.r + 1 - lo, cp);
  }
  w.r = -1;
  w.s = mk(0.0, 0.0);
  w.wsum = 0.0;
}

__device__ __forceinline__ void walk_reset(Walk &w) {
  w.k = -1;
  w.r = -1;
  w.pk = w.pn = w.pc = 0.0;
  w.s = mk(0.0, 0.0);
  w.wsum = 0.0;
}

template <int DIR>
__device__ __forceinline__ void walk_add(const SpecDev &S, Walk &w, double *win, int lo, double x,
                                         cd r, double wj, double half, double pr, double pim,
                                         double hinv, bool fast) {
#pragma clang fp contract(off)
  const int npar = S.npar;
  const gdptr pa = (gdptr)S.ppar;
  bool ok = false;
  if (fast && w.k >= 0) {
    if (DIR > 0) {
      if (w.k + 1 <= npar - 1 && fabs(w.pn - x) <= half) {
        w.k += 1;
        w.pk = w.pn;
        w.pn = (w.k + 1 <= npar - 1) ? pa[w.k + 1] : 0.0;
        ok = true;
      } else {
        ok = fabs(w.pk - x) <= half;
      }
    } else {
      if (fabs(w.pk - x) <= half) {
        ok = true;
      } else if (w.k - 1 >= 1 && fabs(w.pn - x) <= half) {
        w.k -= 1;
        w.pk = w.pn;
        w.pn = (w.k - 1 >= 1) ? pa[w.k - 1] : 0.0;
        ok = true;
      }
    }
  }
  int rc;
  if (ok) {
    rc = min(max(w.k, 2), npar - 2);
  } else {
    // full search: ipar_close before its clamp, kept for the next point
    const int best = ipar_close_raw(S, x);
    rc = min(max(best, 2), npar - 2);
    w.k = best;
    if (best >= 0) {
      w.pk = pa[best];
      const int kn = best + DIR;
      w.pn = (kn >= 1 && kn <= npar - 1) ? pa[kn] : 0.0;
    }
  }
  if (rc != w.r) {
    walk_flush<(DIR < 0)>(w, win, lo, pr, pim, hinv);
    w.r = rc;
    w.pc = pa[rc];
  }
  w.s = w.s + r;
  w.wsum += wj;
}

// Principal-value and tiny-rest coefficients of one resonant row (integrate_res
// after its two integrate() calls), this lane's share.
template <int NCH>
__device__ void pv_coefs(const SpecDev &S, const ProbDev &Pd, const RowClass &rc, int c,
                         double *win, int lo) {
  const int NRI = Pd.NRI, P = Pd.P;
  const double pr = rc.pr, pim = rc.pim;
  double capDelta, smdelta;
  {
#pragma clang fp contract(off)
    capDelta = pr - S.ppar[rc.ipar_res - P];
    smdelta = capDelta / (1.0 * NRI);
  }
  const double hinv = 0.5 / S.dppar_abs;
  CellCache cd1, cd2;
  Cursor u1, u2;
  u1.ic = u2.ic = -1;
  cd1.ic = cd2.ic = -1;
  cd1.cm = cd1.c0 = cd1.cp = cd2.cm = cd2.c0 = cd2.cp = mk(0.0, 0.0);
  if (fabs(pim) > Pd.Tlim) {
    const int blk = (NRI + 1 + NCH - 1) / NCH;
    const int j0 = c * blk, j1 = min(j0 + blk, NRI + 1);
    const double half = 0.5 * S.dppar_abs;
    const bool fast = fabs(smdelta) < 0.9 * S.dppar_abs;
    Walk wu, wd;
    walk_reset(wu);
    walk_reset(wd);
    for (int j = j0; j < j1; ++j) {
      double x, xm;
      {
#pragma clang fp contract(off)
        x = (j == 0) ? pr : ((j == NRI) ? pr + capDelta : pr + smdelta * j);
        xm = 2.0 * pr - x;
      }
      const double w = (j == 0 || j == NRI) ? 1.0 : 2.0;
      const cd r1 = cscale_inv(w, mk(x - pr, -pim));
      walk_add<1>(S, wu, win, lo, x, r1, w, half, pr, pim, hinv, fast);
      walk_add<-1>(S, wd, win, lo, xm, r1, w, half, pr, pim, hinv, fast);  // weight -conj(r1)
    }
    walk_flush<false>(wu, win, lo, pr, pim, hinv);
    walk_flush<true>(wd, win, lo, pr, pim, hinv);
  } else {
    const int blk = (NRI + NCH - 1) / NCH;
    const int j0 = 1 + c * blk, j1 = min(j0 + blk, NRI + 1);
    double sphi = 0.0;
    for (int j = j0; j < j1; ++j) {
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
      const double f = 2.0 * sphi / (2.0 * S.dppar);
      coef_add(S, cd1, u1, win, lo, xp, 1, mk(f, 0.0), hinv);
      coef_add(S, cd2, u2, win, lo, xq, -1, mk(-f, 0.0), hinv);
      if (pim != 0.0) {
        const double sg = (pim > 0.0) ? 1.0 : -1.0;
        cache_flush(cd1, win, lo);
        u1.ic = -1;
        coef_add(S, cd1, u1, win, lo, pr, 1, mk(0.0, sg * 2.0 * PI_D / smdelta), hinv);
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
  if (ntiny > 0) {
    const int blk = (ntiny + 1 + NCH - 1) / NCH;
    const int k0 = c * blk, k1 = min(k0 + blk, ntiny + 1);
    const double half = 0.5 * S.dppar_abs;
    const bool fast = fabs(cs) < 0.9 * S.dppar_abs;
    Walk wt;
    walk_reset(wt);
    for (int kk = k0; kk < k1; ++kk) {
      double x;
      {
#pragma clang fp contract(off)
        x = (kk == 0) ? x0 : x0 + cs * kk;
      }
      const double w = (kk == 0 || kk == ntiny) ? 1.0 : 2.0;
      walk_add<1>(S, wt, win, lo, x, cscale_inv(corr * w, mk(x - pr, -pim)), corr * w, half, pr, pim,
                  hinv, fast);
    }
    walk_flush<false>(wt, win, lo, pr, pim, hinv);
  }
  cache_flush(cd1, win, lo);
  cache_flush(cd2, win, lo);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// How the Landau pass evaluates the continuation of a species: by a call of
// eval_fit (LM_GEN), from one exponential per p and the per-k sums (LM_UNI: one
// Maxwellian fit with pperp-independent (u2, u3)), or by eval_fit_kappa inlined
// (LM_KAPPA: every fit a kappa fit).
enum { LM_GEN = 0, LM_UNI = 1, LM_KAPPA = 2 };

// Landau sums L_a = sum_iperp h a(n,iperp) K(iperp) of this lane, and the
// early-return flag of landau_integrate.  When every fit of the species is a
// Maxwellian whose ppar parameters do not depend on pperp, exp(-u2 (p-u3)^2) is
// the same for every iperp and is evaluated once (bit-identical values).
template <int NCH, int LM>
__device__ void landau_part(const SpecDev &S, const ProbDev &Pd, int an, const RowClass &rc,
                            double omr, double omi, int c, cd L[3], int &zero) {
  const double *__restrict__ arow = S.atab + (size_t)an * (S.nperp + 1) * ATAB_STRIDE;
  const int nperp = S.nperp;
  const double d = S.dppar_abs, dpp = S.dpperp;
  const cd p = mk(rc.pr, rc.pim);
  const cd pp_ = mk(rc.pr + d, rc.pim), pm_ = mk(rc.pr - d, rc.pim);
  const double qk = S.q / fabs(Pd.kpar), km = Pd.kpar / S.m;
  // single Maxwellian fit with pperp-independent (u2, u3): one exponential per p
  constexpr bool UNI = LM == LM_UNI;
  cd E0 = mk(0.0, 0.0), Ep = E0, Em = E0;
  if constexpr (UNI) {
    const double u2 = S.fitc[1], u3 = S.fitc[2];  // iperp = 0 entry
    cd dq, sq;
    dq = mk(p.re - u3, p.im);
    sq = dq * dq;
    E0 = cexpd(mk(-u2 * sq.re, -u2 * sq.im));
    dq = mk(pp_.re - u3, pp_.im);
    sq = dq * dq;
    Ep = cexpd(mk(-u2 * sq.re, -u2 * sq.im));
    dq = mk(pm_.re - u3, pm_.im);
    sq = dq * dq;
    Em = cexpd(mk(-u2 * sq.re, -u2 * sq.im));
  }
  if constexpr (UNI) {
    if (S.lsum != nullptr) {
      // eval_fit = c0(iperp) E(p): the sum over iperp is two omega-independent
      // per-k sums (k_lsum) times A = (E(p+d) - E(p-d))/2d and (om - km p) E(p);
      // the zero test needs only the smallest |c0| of its rows (a product
      // c0 x rounds to zero for some row iff it does for the smallest |c0|)
      if (c == 0) {
        const cd z1 = S.lz_c1 * Ep, z2 = S.lz_c1 * Em, z3 = S.lz_c2 * E0;
        if (czero(z1) || czero(z2) || czero(z3)) zero = 1;
        const double *ls = S.lsum + (size_t)an * 6;
        const cd A = (Ep - Em) * (1.0 / (2.0 * d));
        const cd B = (mk(omr, omi) - km * p) * E0;
#pragma unroll
        for (int k = 0; k < 3; ++k) L[k] = L[k] + qk * ((km * ls[k]) * A + ls[3 + k] * B);
      }
      return;
    }
  }
  auto F = [&](int ip, int which) -> cd {
    if constexpr (LM == LM_GEN) {
      return eval_fit(S, ip, which == 0 ? p : (which > 0 ? pp_ : pm_));
    } else if constexpr (LM == LM_KAPPA) {
      return eval_fit_kappa(S, ip, which == 0 ? p : (which > 0 ? pp_ : pm_));
    } else {
      const double c0 = S.fitc[(size_t)ip * FIT_STRIDE];
      return mk(0.0, 0.0) + c0 * (which == 0 ? E0 : (which > 0 ? Ep : Em));
    }
  };
  if constexpr (!UNI) {
    // every continuation value is an evaluation of the fit: each lane takes a
    // contiguous block of rows and carries f(iperp-1, p), f(iperp, p) from row
    // to row, so a row costs f(iperp, p +- dppar) and f(iperp+1, p), three
    // evaluations instead of four (the same values and the same row
    // formulas), made one at a time (unroll 1: their temporaries are not live
    // together)
    const int blk = (nperp + 1 + NCH - 1) / NCH;
    const int r0 = c * blk, r1 = min(r0 + blk, nperp + 1) - 1;
    if (r0 > r1) return;
    cd fm = (r0 >= 1) ? F(r0 - 1, 0) : mk(0.0, 0.0);  // f(ip - 1, p)
    cd f0 = F(r0, 0);                                // f(ip, p)
#pragma unroll 1
    for (int ip = r0; ip <= r1; ++ip) {
      double h;
      cd dfperp, dfpar, fpi, fpf, fn = mk(0.0, 0.0);
#pragma unroll 1
      for (int w = 0; w < 3; ++w) {
        if (w == 2 && ip >= nperp) break;  // f(nperp + 1, p) is not taken
        const cd v = F(w == 2 ? ip + 1 : ip, w == 0 ? 1 : (w == 1 ? -1 : 0));
        if (w == 0)
          fpi = v;  // f(ip, p + dppar)
        else if (w == 1)
          fpf = v;  // f(ip, p - dppar)
        else
          fn = v;   // f(ip + 1, p)
      }
      if (ip >= 1 && ip <= nperp - 1) {
        h = (ip == nperp - 1) ? 0.5 : 1.0;
        if (czero(fpi) || czero(fpf) || czero(fn)) zero = 1;
        dfperp = (fn - fm) * (1.0 / (2.0 * dpp));
        dfpar = (fpi - fpf) * (1.0 / (2.0 * d));
      } else if (ip == 0) {
        h = 0.5;
        dfperp = (fn - f0) * (1.0 / dpp);
        dfpar = (fpi - fpf) * (1.0 / (2.0 * d));
      } else {
        h = 0.5;
        dfperp = (f0 - fm) * (1.0 / dpp);
        dfpar = (fpi - fpf) * (1.0 / (2.0 * d));
      }
      const cd K = qk * ((S.pperp[ip] * dfpar - p * dfperp) * km + mk(omr, omi) * dfperp);
      const double *a = arow + (size_t)ip * ATAB_STRIDE;
      L[0] = L[0] + (h * a[0]) * K;
      L[1] = L[1] + (h * a[1]) * K;
      L[2] = L[2] + (h * a[2]) * K;
      fm = f0;
      f0 = fn;
    }
    return;
  }
  for (int ip = c; ip <= nperp; ip += NCH) {
    double h;
    cd dfperp, dfpar;
    if (ip >= 1 && ip <= nperp - 1) {
      h = (ip == nperp - 1) ? 0.5 : 1.0;
      const cd fpi = F(ip, 1), fpf = F(ip, -1);
      const cd fperpi = F(ip + 1, 0), fperpf = F(ip - 1, 0);
      if (czero(fpi) || czero(fpf) || czero(fperpi)) zero = 1;
      dfperp = (fperpi - fperpf) * (1.0 / (2.0 * dpp));
      dfpar = (fpi - fpf) * (1.0 / (2.0 * d));
    } else if (ip == 0) {
      h = 0.5;
      dfperp = (F(1, 0) - F(0, 0)) * (1.0 / dpp);
      dfpar = (F(0, 1) - F(0, -1)) * (1.0 / (2.0 * d));
    } else {
      h = 0.5;
      dfperp = (F(nperp, 0) - F(nperp - 1, 0)) * (1.0 / dpp);
      dfpar = (F(nperp, 1) - F(nperp, -1)) * (1.0 / (2.0 * d));
    }
    const cd K = qk * ((S.pperp[ip] * dfpar - p * dfperp) * km + mk(omr, omi) * dfperp);
    const double *a = arow + (size_t)ip * ATAB_STRIDE;
    L[0] = L[0] + (h * a[0]) * K;
    L[1] = L[1] + (h * a[1]) * K;
    L[2] = L[2] + (h * a[2]) * K;
  }
}

}  // namespace

// PART 0: principal-value terms; PART 1: Landau terms of species whose analytic
// continuation is one Maxwellian with pperp-independent (u2, u3); PART 3: Landau
// terms of species whose every fit is a kappa fit; PART 2: Landau terms of the
// other species.  PV writes partial slots rb, the Landau launches slots RB + rb
// of chiB / lowB (each species by exactly one of PART 1, 2, 3).
// One wavefront's share of k_resonant: omega block bx, species by, harmonic
// block rb of RB; win_lds holds (64 / NCH) windows of 2 pvw doubles.
template <int NCH, int PART>
__device__ __forceinline__ void resonant_body(const ProbDev *__restrict__ probp, int nom,
                                              const double2 *__restrict__ om_in,
                                              const int *__restrict__ active,
                                              double2 *__restrict__ chiB,
                                              double2 *__restrict__ lowB,
                                              const int *__restrict__ pidx, int RB, int rb, int bx,
                                              int by, int lane, double *win_lds,
                                              int pvw = PV_WIN) {
  constexpr int OMB = 64 / NCH;
  const ProbDev &Pd = (NCH == 64 && pidx != nullptr) ? probp[pidx[bx]] : *probp;
  const int g = lane / NCH, c = lane % NCH;
  const int iom = bx * OMB + g;
  const int s = by;
  const int slot = (PART == 0 ? 0 : RB) + rb;
  const bool valid = (iom < nom) && (active == nullptr || active[iom] != 0);
  const double2 ov = valid ? om_in[iom] : make_double2(1.0, 0.5);
  const double omr = ov.x, omi = ov.y;
  const int nspec = Pd.nspec;
  const SpecDev &S = Pd.sp[s];
  const int npar = S.npar, P = Pd.P;
  const double kpar = Pd.kpar;
  const bool uni_s = (S.fit_uniform != 0) && (S.nfits == 1) && (S.acmethod == 1);
  const bool kap_s = S.fit_kappa != 0;
  if ((PART == 1 && !uni_s) || (PART == 2 && (uni_s || kap_s)) || (PART == 3 && !kap_s)) return;
  double *win = win_lds + g * (2 * pvw);  // pvw >= 2P + 8, the window's extent
  cd res[6];
#pragma unroll
  for (int k = 0; k < 6; ++k) res[k] = mk(0.0, 0.0);
  double2 *lo = (lowB != nullptr && valid && c == 0)
                    ? lowB + (((size_t)iom * nspec + s) * (2 * RB) + slot) * 18
                    : nullptr;
  if (lo)
    for (int t = 0; t < 18; ++t) lo[t] = make_double2(0.0, 0.0);

  // harmonics whose p_res = (m w - n q)/kpar can satisfy determine_resonances:
  // Re p_res in [pa[1] - P dppar, max(pa[npar], pa[npar-1] + P dppar)), widened by one
  int nlo = 1, nhi = 0;
  if (valid && omr == omr) {
    const double L = S.ppar[1] - P * S.dppar;
    const double U = fmax(S.ppar[npar], S.ppar[npar - 1] + P * S.dppar);
    const double n1 = (S.m * omr - kpar * L) / S.q, n2 = (S.m * omr - kpar * U) / S.q;
    const double lo_d = fmin(n1, n2) - 1.0, hi_d = fmax(n1, n2) + 1.0;
    nlo = (int)fmax(floor(lo_d), -(double)S.nmax);
    nhi = (int)fmin(ceil(hi_d), (double)S.nmax);
  }
  const double scaleN = 2.0 * PI_D * S.dpperp * 0.25 * (-S.q / kpar);
  for (int n = nlo + rb; n <= nhi; n += RB) {
    RowClass rc;
    classify(S, P, kpar, omr, omi, n, rc, (gdptr)S.ppar);
    if (!rc.found || (PART == 0 && !rc.pv) || (PART != 0 && !rc.landau)) continue;
    const int an = n < 0 ? -n : n;
    const double *rows = S.nodes + (size_t)an * (npar - 1) * NODE_STRIDE;
    cd add[6];
#pragma unroll
    for (int k = 0; k < 6; ++k) add[k] = mk(0.0, 0.0);
    if constexpr (PART == 0) {
      const int wlo = max(1, rc.ipar_res - P - 3);
      const int whi = min(npar - 1, rc.ipar_res + P + 4);
      for (int t = c; t < 2 * pvw; t += NCH) win[t] = 0.0;
      wave_sync();
      pv_coefs<NCH>(S, Pd, rc, c, win, wlo);
      wave_sync();
      cd pv[6];
#pragma unroll
      for (int k = 0; k < 6; ++k) pv[k] = mk(0.0, 0.0);
      for (int i = wlo + c; i <= whi; i += NCH) {
        const cd cf = mk(win[2 * (i - wlo)], win[2 * (i - wlo) + 1]);
        cd u[6];
        node6(rows + (size_t)(i - 1) * NODE_STRIDE, S.ppar[i], omr, omi, u);
#pragma unroll
        for (int k = 0; k < 6; ++k) pv[k] = pv[k] + cf * u[k];
      }
      wave_sync();
      double capDelta, smdelta;
      {
#pragma clang fp contract(off)
        capDelta = rc.pr - S.ppar[rc.ipar_res - P];
        smdelta = capDelta / (1.0 * Pd.NRI);
      }
      const double sc = scaleN * smdelta;
#pragma unroll
      for (int k = 0; k < 6; ++k)
        add[k] = add[k] + (sc * gsumc<NCH>(pv[ACC_OF_MODE[k]])) * tconst(Pd, S, n, k);
    }
    if constexpr (PART != 0) {
      cd L[3] = {mk(0.0, 0.0), mk(0.0, 0.0), mk(0.0, 0.0)};
      int zero = 0;
      landau_part<NCH, PART == 1 ? LM_UNI : (PART == 3 ? LM_KAPPA : LM_GEN)>(S, Pd, an, rc, omr,
                                                                             omi, c, L, zero);
      zero = gor<NCH>(zero);
      if (!zero) {
        const cd p = mk(rc.pr, rc.pim);
        const cd Ls0 = gsumc<NCH>(L[0]), Ls1 = gsumc<NCH>(L[1]), Ls2 = gsumc<NCH>(L[2]);
        // T_res (int_T_res, ALPS_fns.f90:1768-1808): zz, xz, yz take p_res
        const cd Lm[6] = {Ls0, Ls1, Ls0 * (p * p), Ls2, Ls0 * p, Ls2 * p};
        const double f = -(double)rc.landau * S.dpperp * PI_D * 2.0 * PI_D;
#pragma unroll
        for (int k = 0; k < 6; ++k) add[k] = add[k] + timesI(f * (Lm[k] * tconst(Pd, S, n, k)));
      }
    }
    if (n == 0) {
      add[0] = add[3] = add[4] = mk(0.0, 0.0);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) res[k] = res[k] + add[k];
    if (lo && an <= 1) {
#pragma unroll
      for (int k = 0; k < 6; ++k) lo[6 * (n + 1) + k] = make_double2(add[k].re, add[k].im);
    }
  }
  if (valid && c == 0) {
    // omega-fastest layout: chiB[((slot * nspec + s) * 6 + k) * nom + iom]
    double2 *o = chiB + ((size_t)(slot * nspec + s) * 6) * nom + iom;
#pragma unroll
    for (int k = 0; k < 6; ++k) o[(size_t)k * nom] = make_double2(res[k].re, res[k].im);
  }
}

// pvw: principal-value window entries per omega (2P + 8, at most PV_WIN); the
// windows of part 0 are (64 / NCH) 2 pvw doubles of dynamic LDS, so that the
// usual P = 3 does not cost the occupancy of the largest P.  Held to four
// wavefronts per SIMD (128 registers, a few spilled): the map's PV and Landau
// passes run about 10 % faster than at three.  Part 2 (the Landau pass on
// eval_fit calls) is k_resonant_gen, with no such bound: its call needs the
// full register file, and a bound only spills it
template <int NCH, int PART>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_resonant(const ProbDev *__restrict__ probp, int nom,
                                                 const double2 *__restrict__ om_in,
                                                 const int *__restrict__ active,
                                                 double2 *__restrict__ chiB,
                                                 double2 *__restrict__ lowB,
                                                 const int *__restrict__ pidx, int pvw) {
  extern __shared__ double win_dyn[];
  resonant_body<NCH, PART>(probp, nom, om_in, active, chiB, lowB, pidx, gridDim.z, blockIdx.z,
                           blockIdx.x, blockIdx.y, threadIdx.x, win_dyn, pvw);
}

template <int NCH>
__global__ __launch_bounds__(64) void k_resonant_gen(const ProbDev *__restrict__ probp, int nom,
                                                     const double2 *__restrict__ om_in,
                                                     const int *__restrict__ active,
                                                     double2 *__restrict__ chiB,
                                                     double2 *__restrict__ lowB,
                                                     const int *__restrict__ pidx, int pvw) {
  extern __shared__ double win_dyn[];
  resonant_body<NCH, 2>(probp, nom, om_in, active, chiB, lowB, pidx, gridDim.z, blockIdx.z,
                        blockIdx.x, blockIdx.y, threadIdx.x, win_dyn, pvw);
}

// The Landau part of the kappa species (part 3): the lean power leaves the
// pass short of four wavefronts per SIMD; three, with a few spills outside
// the evaluation loop, run it 12 % faster than the two of its natural
// register count (15.7 against 17.7 ms on a 512 x 512 map, 256 x 512 grid).
template <int NCH>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) void k_resonant_kappa(
    const ProbDev *__restrict__ probp, int nom, const double2 *__restrict__ om_in,
    const int *__restrict__ active, double2 *__restrict__ chiB, double2 *__restrict__ lowB,
    const int *__restrict__ pidx) {
  __shared__ double win[(64 / NCH) * 2 * PV_WIN];
  resonant_body<NCH, 3>(probp, nom, om_in, active, chiB, lowB, pidx, gridDim.z, blockIdx.z,
                        blockIdx.x, blockIdx.y, threadIdx.x, win);
}

// The principal-value and Landau parts in one launch (small batches, where a
// launch costs more than the work): z < RB runs the principal-value part of
// harmonic block z, z >= RB the Landau part of block z - RB (the part-1, 2 or
// 3 form, whichever the species takes).  GEN: bit 0 set if a species takes
// part 2, bit 1 if one takes part 3.
template <int NCH, int GEN>
__global__ __launch_bounds__(64) void k_resonant_all(const ProbDev *__restrict__ probp, int nom,
                                                     const double2 *__restrict__ om_in,
                                                     double2 *__restrict__ chiB, int RB) {
  __shared__ double win[(64 / NCH) * 2 * PV_WIN];
  const int z = blockIdx.z, bx = blockIdx.x, by = blockIdx.y, lane = threadIdx.x;
  if (z < RB) {
    resonant_body<NCH, 0>(probp, nom, om_in, nullptr, chiB, nullptr, nullptr, RB, z, bx, by, lane, win);
    return;
  }
  resonant_body<NCH, 1>(probp, nom, om_in, nullptr, chiB, nullptr, nullptr, RB, z - RB, bx, by, lane,
                        win);
  if constexpr ((GEN & 1) != 0)
    resonant_body<NCH, 2>(probp, nom, om_in, nullptr, chiB, nullptr, nullptr, RB, z - RB, bx, by,
                          lane, win);
  if constexpr ((GEN & 2) != 0)
    resonant_body<NCH, 3>(probp, nom, om_in, nullptr, chiB, nullptr, nullptr, RB, z - RB, bx, by,
                          lane, win);
}

// ---------------------------------------------------------------------------
// k_assemble: a workgroup of 256 threads handles 256/(6 nspec) omegas; thread
// (omega, s, k) sums the partials of chi_sk, then one thread per omega forms
// the dielectric tensor, the wave tensor and det D.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void assemble_block(
    int blk, const ProbDev *__restrict__ probp, int nom, int W, int RB,
    const double2 *__restrict__ om_in, const int *__restrict__ active,
    const double2 *__restrict__ chiA, const double2 *__restrict__ chiB,
    const double2 *__restrict__ lowA, const double2 *__restrict__ lowB,
    const double2 *__restrict__ chiR, int RR, const double2 *__restrict__ lowR,
    double2 *__restrict__ det_out, double2 *__restrict__ chi_out, double2 *__restrict__ low_out,
    double2 *__restrict__ wave_out, int nspec, const int *__restrict__ pidx,
    double2 *det_lds = nullptr) {
  __shared__ double2 lchi[256];
  const int per = 6 * nspec;
  const int opb = 256 / per;  // omegas per block
  const int t = threadIdx.x;
  const int lo_om = t / per, sk = t % per;
  const int s = sk / 6, k = sk % 6;
  const int iom = blk * opb + lo_om;
  const bool in = (lo_om < opb) && (iom < nom) && (active == nullptr || active[iom] != 0);
  const ProbDev &Pd = pidx != nullptr ? probp[pidx[min(iom, nom - 1)]] : *probp;
  if (in) {
    // partial sums, loads issued in batches of 16
    cd v = mk(0.0, 0.0);
    const int tot = W + 2 * RB;
    for (int base = 0; base < tot; base += 16) {
      double2 buf[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int j = base + u;
        if (j < W)
          buf[u] = chiA[((size_t)(j * nspec + s) * 6 + k) * nom + iom];
        else if (j < tot)
          buf[u] = chiB[((size_t)((j - W) * nspec + s) * 6 + k) * nom + iom];
        else
          buf[u] = make_double2(0.0, 0.0);
      }
#pragma unroll
      for (int u = 0; u < 16; ++u) v = v + mk(buf[u].x, buf[u].y);
    }
    if (Pd.sp[s].rel)  // relativistic species: the partial slots of k_rel_nonres / k_rel_res
      for (int j = 0; j < RR; ++j) {
        const double2 a = chiR[((size_t)(j * nspec + s) * 6 + k) * nom + iom];
        v = v + mk(a.x, a.y);
      }
    lchi[t] = make_double2(v.re, v.im);
  }
  __syncthreads();
  if (!in || sk != 0) return;
  const double2 ov = om_in[iom];
  const cd om = mk(ov.x, ov.y);
  cd eps[6];
  for (int q = 0; q < 6; ++q) eps[q] = mk(0.0, 0.0);
  cd dn = (om * om) * (Pd.vA * Pd.vA);
  if (!Pd.kperp_norm) dn = dn * (Pd.kperp * Pd.kperp);
  for (int sp = 0; sp < nspec; ++sp) {
    const SpecDev &S = Pd.sp[sp];
    cd chi[6];
    for (int q = 0; q < 6; ++q) {
      const double2 a = lchi[lo_om * per + sp * 6 + q];
      chi[q] = mk(a.x, a.y);
    }
    const double ee = Pd.kperp_norm ? S.int_ee : Pd.kperp * Pd.kperp * S.int_ee;
    chi[2] = chi[2] + mk(ee, 0.0);
    const double norm = S.dens * S.q;
    for (int q = 0; q < 6; ++q) {
      chi[q] = chi[q] * norm;
      eps[q] = eps[q] + chi[q];
    }
    if (chi_out != nullptr) {
      double2 *o = chi_out + (size_t)iom * nspec * 9;
      cd v[3][3];
      for (int q = 0; q < 6; ++q) v[MI[q]][MJ[q]] = cdiv(chi[q], dn);
      v[1][0] = mk(-v[0][1].re, -v[0][1].im);
      v[2][0] = v[0][2];
      v[2][1] = mk(-v[1][2].re, -v[1][2].im);
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) o[sp + nspec * (i + 3 * j)] = make_double2(v[i][j].re, v[i][j].im);
    }
    if (low_out != nullptr) {
      double2 *o = low_out + (size_t)iom * nspec * 27;
      for (int qq = 0; qq < 3; ++qq) {
        cd v[3][3];
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) v[i][j] = mk(0.0, 0.0);
        for (int q = 0; q < 6; ++q) {
          if (qq == 1 && (q == 0 || q == 3 || q == 4)) continue;
          cd tt;
          if (S.rel) {
            // relativistic: one of k_rel_nonres / k_rel_res wrote the term; the
            // ee term goes into chi only (ALPS_fns.f90:481-486)
            const double2 a = lowR[((size_t)iom * nspec + sp) * 18 + 6 * qq + q];
            tt = mk(a.x, a.y);
          } else {
            const double2 a = lowA[((size_t)iom * nspec + sp) * 18 + 6 * qq + q];
            tt = mk(a.x, a.y);
            for (int b = 0; b < 2 * RB; ++b) {
              const double2 u = lowB[(((size_t)iom * nspec + sp) * (2 * RB) + b) * 18 + 6 * qq + q];
              tt = tt + mk(u.x, u.y);
            }
            if (qq == 1 && q == 2) tt = tt + mk(ee, 0.0);
          }
          v[MI[q]][MJ[q]] = cdiv(tt * norm, dn);
        }
        v[1][0] = mk(-v[0][1].re, -v[0][1].im);
        v[2][0] = v[0][2];
        v[2][1] = mk(-v[1][2].re, -v[1][2].im);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            o[sp + nspec * (i + 3 * (j + 3 * qq))] = make_double2(v[i][j].re, v[i][j].im);
      }
    }
  }
  // dielectric tensor, wave tensor, determinant (ALPS_fns.f90:566-624)
  const cd unit = Pd.kperp_norm ? (om * Pd.vA) * (om * Pd.vA)
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
  const cd w11 = (eps[0] + unit) - mk(enz2, 0.0);
  const cd w22 = ((eps[1] + unit) - mk(enz2, 0.0)) - mk(enx2, 0.0);
  const cd w33 = (eps[2] + unit) - mk(enx2, 0.0);
  const cd w13 = eps[4] + mk(enxnz, 0.0);
  const cd w12 = eps[3];
  const cd w23 = eps[5];
  const cd det = w11 * (w22 * w33 + w23 * w23) + 2.0 * w12 * w23 * w13 - (w13 * w13) * w22 +
                 (w12 * w12) * w33;
  det_out[iom] = make_double2(det.re, det.im);
  if (det_lds != nullptr) det_lds[iom] = make_double2(det.re, det.im);
  if (wave_out != nullptr) {
    double2 *o = wave_out + (size_t)iom * 9;
    const cd Wt[3][3] = {{w11, w12, w13},
                         {mk(-w12.re, -w12.im), w22, w23},
                         {w13, mk(-w23.re, -w23.im), w33}};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) o[i + 3 * j] = make_double2(Wt[i][j].re, Wt[i][j].im);
  }
}

__global__ __launch_bounds__(256) void k_assemble(
    const ProbDev *__restrict__ probp, int nom, int W, int RB, const double2 *__restrict__ om_in,
    const int *__restrict__ active, const double2 *__restrict__ chiA,
    const double2 *__restrict__ chiB, const double2 *__restrict__ lowA,
    const double2 *__restrict__ lowB, const double2 *__restrict__ chiR, int RR,
    const double2 *__restrict__ lowR, double2 *__restrict__ det_out, double2 *__restrict__ chi_out,
    double2 *__restrict__ low_out, double2 *__restrict__ wave_out, int nspec,
    const int *__restrict__ pidx) {
  assemble_block(blockIdx.x, probp, nom, W, RB, om_in, active, chiA, chiB, lowA, lowB, chiR, RR,
                 lowR, det_out, chi_out, low_out, wave_out, nspec, pidx);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
// npar1: the ppar points of the problem's grids (npar + 1), staged in LDS
// when at most LPA
hipError_t launch_sweep(hipStream_t st, int nch, int W, int RC, const ProbDev *prob, int nom,
                        const double2 *om, const int *active, double2 *chiA, double2 *lowA,
                        int npar1, const int *pidx, unsigned long long *work) {
  if (nom <= 0) return hipSuccess;
  if (pidx != nullptr && nch != 64) return hipErrorInvalidValue;
  const int lpa_n = (npar1 >= 1 && npar1 <= LPA) ? npar1 : 0;
#define SW(N, WV)                                                                                \
  if (nch == N && W == WV) {                                                                     \
    const int grid = (nom + (64 / N) - 1) / (64 / N);                                            \
    hipLaunchKernelGGL((k_sweep<N, WV>), dim3(grid, RC), dim3(64 * WV), lpa_n * sizeof(double),  \
                       st, prob, nom, om, active, chiA, lowA, pidx, work, lpa_n);                \
    return hipGetLastError();                                                                    \
  }
  SW(1, 1) SW(1, 2) SW(1, 4) SW(2, 4) SW(4, 4) SW(8, 4) SW(16, 4) SW(16, 8) SW(32, 8) SW(64, 4)
  SW(64, 8)
#undef SW
  return hipErrorInvalidValue;
}

// the per-omega-problem sweep (k_sweep_req): nch lanes per row pair, W
// wavefronts per workgroup, RC workgroups per omega; W RC partial slots per
// omega, or RC with reduce (the 32/4/4 shape only)
hipError_t launch_sweep_req(hipStream_t st, int nch, int W, int RC, int K, const ProbDev *prob,
                            int nom, const double2 *om, double2 *chiA, const int *pidx,
                            bool reduce) {
  if (nom <= 0) return hipSuccess;
  if (pidx == nullptr) return hipErrorInvalidValue;
  if (reduce) {
    if (!(nch == 32 && W == 4 && (K == 4 || K == 2))) return hipErrorInvalidValue;
    if (K == 4)
      hipLaunchKernelGGL((k_sweep_req<32, 4, 4, true>), dim3(nom, RC), dim3(256), 0, st, prob, nom,
                         om, chiA, pidx);
    else
      hipLaunchKernelGGL((k_sweep_req<32, 4, 2, true>), dim3(nom, RC), dim3(256), 0, st, prob, nom,
                         om, chiA, pidx);
    return hipGetLastError();
  }
#define SQ(N, WV, KK)                                                                             \
  if (nch == N && W == WV && K == KK) {                                                           \
    hipLaunchKernelGGL((k_sweep_req<N, WV, KK, false>), dim3(nom, RC), dim3(64 * WV), 0, st, prob, \
                       nom, om, chiA, pidx);                                                      \
    return hipGetLastError();                                                                     \
  }
  SQ(16, 8, 1) SQ(32, 8, 1) SQ(16, 8, 2) SQ(32, 8, 2) SQ(16, 8, 4) SQ(32, 8, 4) SQ(32, 4, 2)
  SQ(32, 4, 4) SQ(64, 4, 2) SQ(64, 4, 4)
#undef SQ
  return hipErrorInvalidValue;
}

// lgen: bit 0 set if a species takes the part-2 Landau form, bit 1 if one
// takes the part-3 (kappa) form; P: positions_principal of every problem of
// the launch (<= 0: unknown, the largest window)
hipError_t launch_resonant(hipStream_t st, int nch, int RB, const ProbDev *prob, int nspec, int nom,
                           const double2 *om, const int *active, double2 *chiB, double2 *lowB,
                           int lgen, const int *pidx, int P) {
  if (nom <= 0) return hipSuccess;
  if (pidx != nullptr && nch != 64) return hipErrorInvalidValue;
  const int pvw = (P >= 1 && 2 * P + 8 <= PV_WIN) ? 2 * P + 8 : PV_WIN;
#define RS(N)                                                                                  \
  if (nch == N) {                                                                              \
    const int grid = (nom + (64 / N) - 1) / (64 / N);                                          \
    hipLaunchKernelGGL((k_resonant<N, 0>), dim3(grid, nspec, RB), dim3(64),                    \
                       (64 / N) * 2 * pvw * sizeof(double), st, prob, nom, om, active, chiB,   \
                       lowB, pidx, pvw);                                                       \
    hipLaunchKernelGGL((k_resonant<N, 1>), dim3(grid, nspec, RB), dim3(64), 0, st, prob, nom,  \
                       om, active, chiB, lowB, pidx, pvw);                                     \
    if (lgen & 1)                                                                              \
      hipLaunchKernelGGL((k_resonant_gen<N>), dim3(grid, nspec, RB), dim3(64), 0, st, prob, nom, \
                         om, active, chiB, lowB, pidx, pvw);                                   \
    if (lgen & 2)                                                                              \
      hipLaunchKernelGGL((k_resonant_kappa<N>), dim3(grid, nspec, RB), dim3(64), 0, st, prob,   \
                         nom, om, active, chiB, lowB, pidx);                                   \
    return hipGetLastError();                                                                  \
  }
  RS(4) RS(8) RS(16) RS(32) RS(64)
#undef RS
  return hipErrorInvalidValue;
}

hipError_t launch_assemble(hipStream_t st, const ProbDev *prob, int nspec, int nom, int W, int RB,
                           const double2 *om, const int *active, const double2 *chiA,
                           const double2 *chiB, const double2 *lowA, const double2 *lowB,
                           const double2 *chiR, int RR, const double2 *lowR, double2 *det,
                           double2 *chi, double2 *low, double2 *wave, const int *pidx) {
  if (nom <= 0) return hipSuccess;
  const int per = 6 * nspec;
  const int opb = 256 / per;
  hipLaunchKernelGGL(k_assemble, dim3((nom + opb - 1) / opb), dim3(256), 0, st, prob, nom, W, RB,
                     om, active, chiA, chiB, lowA, lowB, chiR, RR, lowR, det, chi, low, wave, nspec,
                     pidx);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// map_search post-processing (ALPS_fns.f90:3723-3742, find_minima :3860-3966)
// on the device: log10|D| with the reference's NaN / Inf sentinels, and the
// strict four-neighbour minima of log10|D| (one-sided at the edges).  Point
// (ir, ii), 1-based, is element (ir-1) + nr (ii-1).
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ void map_value(double2 &c, double &v) {
  v = log10(hypot(c.x, c.y));
  const double tmp = c.x;
  const bool same = (tmp == c.x) && (0.0 == c.y);  // cmplx(tmp, 0) == cal
  bool sentinel;
  if (c.y != 0.0)
    sentinel = same;
  else
    sentinel = !same;  // a NaN real part
  if (sentinel) {
    c = make_double2(999999.0, 0.0);
    v = 999999.0;
  }
  if (fabs(tmp) > 1.e100) {
    c = make_double2(899999.0, 0.0);
    v = 899999.0;
  }
}

__device__ __forceinline__ double map_val_at(const double2 *__restrict__ cal, size_t k) {
  double2 c = cal[k];
  double v;
  map_value(c, v);
  return v;
}

__global__ __launch_bounds__(256) void k_map_post(int nr, int ni, const double2 *__restrict__ cal,
                                                  double2 *__restrict__ calo, double *__restrict__ val,
                                                  unsigned char *__restrict__ hit) {
  const size_t N = (size_t)nr * ni;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  double2 c = cal[k];
  double v;
  map_value(c, v);
  calo[k] = c;
  val[k] = v;
  const int ir = (int)(k % nr) + 1, ii = (int)(k / nr) + 1;
  bool h = false;
  if (nr > 1 && ni > 1) {
    bool rok;
    if (ir == 1)
      rok = v < map_val_at(cal, k + 1);
    else if (ir == nr)
      rok = v < map_val_at(cal, k - 1);
    else
      rok = v < map_val_at(cal, k - 1) && v < map_val_at(cal, k + 1);
    if (rok) {
      if (ii == 1)
        h = v < map_val_at(cal, k + nr);
      else if (ii == ni)
        h = v < map_val_at(cal, k - nr);
      else
        h = v < map_val_at(cal, k - nr) && v < map_val_at(cal, k + nr);
    }
  }
  hit[k] = h ? 1 : 0;
}

// The linear map grid (ALPS_fns.f90:3650-3700 with loggridw = loggridg = F):
// point (ir, ii) at k = (ir-1) + nr (ii-1) is (omi + dr (ir-1), gami + di (ii-1)),
// one product and one sum each, rounded as the host computes them.
__global__ __launch_bounds__(256) void k_map_grid(int nr, int ni, double omi, double dr,
                                                  double gami, double di,
                                                  double2 *__restrict__ om) {
#pragma clang fp contract(off)
  const size_t N = (size_t)nr * ni;
  const size_t k = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= N) return;
  const int ir = (int)(k % nr) + 1, ii = (int)(k / nr) + 1;
  om[k] = make_double2(omi + dr * (1.0 * (ir - 1)), gami + di * (1.0 * (ii - 1)));
}

}  // namespace

hipError_t launch_map_grid(hipStream_t st, int nr, int ni, double omi, double dr, double gami,
                           double di, double2 *om) {
  const size_t N = (size_t)nr * ni;
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_map_grid, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, nr, ni, omi,
                     dr, gami, di, om);
  return hipGetLastError();
}

hipError_t launch_map_post(hipStream_t st, int nr, int ni, const double2 *cal, double2 *calo,
                           double *val, unsigned char *hit) {
  const size_t N = (size_t)nr * ni;
  if (N == 0) return hipSuccess;
  hipLaunchKernelGGL(k_map_post, dim3((unsigned)((N + 255) / 256)), dim3(256), 0, st, nr, ni, cal,
                     calo, val, hit);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------
// One round of many secant_osc solves on the device: every unfinished solve
// takes det D at the frequency it asked for, steps its state machine
// (alpsg_osc.h, the same code the host runs) and writes the next frequency it
// needs.  One workgroup; round_ctr counts the rounds (read by all threads,
// advanced by thread 0 after a barrier), remaining gets the number of threads
// that still hold an unfinished solve.
// ---------------------------------------------------------------------------
namespace {

__device__ __forceinline__ void osc_step(int n, alpsg_osc::OscMachine *__restrict__ S,
                                         const double2 *__restrict__ det, double2 *__restrict__ om,
                                         int *__restrict__ round_ctr, int *__restrict__ remaining) {
  const int round = *round_ctr;
  int left = 0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    alpsg_osc::OscMachine m = S[r];
    if (!m.done) {
      m.feed(alpsg_osc::zmk(det[r].x, det[r].y));
      if (m.done) {
        m.done_round = round;
      } else {
        const alpsg_osc::zc w = m.want();
        om[r] = make_double2(w.re, w.im);
        left = 1;
      }
      S[r] = m;
    }
  }
  const int tot = __syncthreads_count(left);
  if (threadIdx.x == 0) {
    *remaining = tot;
    *round_ctr = round + 1;
  }
}

__global__ __launch_bounds__(1024) void k_osc_round(int n, alpsg_osc::OscMachine *__restrict__ S,
                                                    const double2 *__restrict__ det,
                                                    double2 *__restrict__ om, int *__restrict__ round_ctr,
                                                    int *__restrict__ remaining) {
  osc_step(n, S, det, om, round_ctr, remaining);
}

// k_assemble for a small batch in one workgroup (its omegas in chunks of
// 256 / (6 nspec)), followed by the secant step of every solve (nom <= 256).
// Each thread's solve state is loaded first, beside the count of unfinished
// solves, and det D reaches the step through LDS, so that the round waits on
// global memory for the partial sums only.
__global__ __launch_bounds__(256) void k_assemble_osc(
    const ProbDev *__restrict__ probp, int nom, int W, int RB, const double2 *__restrict__ om_in,
    const double2 *__restrict__ chiA, const double2 *__restrict__ chiB, double2 *__restrict__ det,
    int nspec, alpsg_osc::OscMachine *__restrict__ S, double2 *__restrict__ om_next,
    int *__restrict__ round_ctr, int *__restrict__ remaining) {
  __shared__ double2 det_l[256];
  const int t = threadIdx.x;
  alpsg_osc::OscMachine m;
  if (t < nom) m = S[t];
  // a round after every solve has finished does nothing (its values would be
  // ignored); the count is read before any thread can change it
  const int left = *remaining;
  const int round = *round_ctr;
  if (left == 0) return;
  const int opb = 256 / (6 * nspec);
  for (int blk = 0; blk * opb < nom; ++blk) {
    assemble_block(blk, probp, nom, W, RB, om_in, nullptr, chiA, chiB, nullptr, nullptr, nullptr, 0,
                   nullptr, det, nullptr, nullptr, nullptr, nspec, nullptr, det_l);
    __syncthreads();
  }
  int more = 0;
  if (t < nom && !m.done) {
    m.feed(alpsg_osc::zmk(det_l[t].x, det_l[t].y));
    if (m.done) {
      m.done_round = round;
    } else {
      const alpsg_osc::zc w = m.want();
      om_next[t] = make_double2(w.re, w.im);
      more = 1;
    }
    S[t] = m;
  }
  const int tot = __syncthreads_count(more);
  if (t == 0) {
    *remaining = tot;
    *round_ctr = round + 1;
  }
}

// The request sweep and both resonant parts of a small batch in one launch:
// workgroups [0, nom RC) run the sweep (omega b % nom, workgroup b / nom of
// RC), the others four resonant wavefronts each (every part and harmonic
// block).
template <int NCHS, int K, int NCHR, int GEN>
__global__ __launch_bounds__(256) void k_sweep_res(const ProbDev *__restrict__ probp, int nom,
                                                   const double2 *__restrict__ om,
                                                   double2 *__restrict__ chiA,
                                                   double2 *__restrict__ chiB,
                                                   const int *__restrict__ zidx, int RC, int RB,
                                                   int nspec, const int *__restrict__ remaining) {
  constexpr int OMB = 64 / NCHR;
  __shared__ double win[4][OMB * 2 * PV_WIN];
  if (*remaining == 0) return;  // every solve has finished: nothing to evaluate
  const int A = nom * RC;
  const int b = blockIdx.x;
  if (b < A) {
    sweep_req_body<NCHS, 4, K, true>(probp, nom, om, chiA, zidx, b % nom, b / nom, RC);
    return;
  }
  const int gx = (nom + OMB - 1) / OMB;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int vb = (b - A) * 4 + wave;
  if (vb >= gx * nspec * 2 * RB) return;
  const int x = vb % gx, y = (vb / gx) % nspec, z = vb / (gx * nspec);
  if (z < RB) {
    resonant_body<NCHR, 0>(probp, nom, om, nullptr, chiB, nullptr, nullptr, RB, z, x, y, lane,
                           win[wave]);
    return;
  }
  resonant_body<NCHR, 1>(probp, nom, om, nullptr, chiB, nullptr, nullptr, RB, z - RB, x, y, lane,
                         win[wave]);
  if constexpr ((GEN & 1) != 0)
    resonant_body<NCHR, 2>(probp, nom, om, nullptr, chiB, nullptr, nullptr, RB, z - RB, x, y, lane,
                           win[wave]);
  if constexpr ((GEN & 2) != 0)
    resonant_body<NCHR, 3>(probp, nom, om, nullptr, chiB, nullptr, nullptr, RB, z - RB, x, y, lane,
                           win[wave]);
}

}  // namespace

// a secant round of a small batch in two launches: the request sweep (32
// lanes per row pair, 4 records in flight, 4 wavefronts per workgroup, RC
// workgroups per omega) with both resonant parts (nchr lanes per omega, RB
// harmonic blocks), then the assembly and the secant step in one workgroup;
// ctl = {round, unfinished}
hipError_t launch_round(hipStream_t st, int nchr, int RB, int RC, const ProbDev *prob, int nspec,
                        int nom, double2 *om, double2 *chiA, double2 *chiB, double2 *det,
                        const int *zidx, int lgen, alpsg_osc::OscMachine *S, int *ctl) {
  if (nom <= 0) return hipSuccess;
  if (nom > 256 || 6 * nspec > 256) return hipErrorInvalidValue;
  bool ok = false;
#define RDG(N, G)                                                                                  \
  if (lgen == G)                                                                                   \
    hipLaunchKernelGGL((k_sweep_res<32, 4, N, G>), dim3(grid), dim3(256), 0, st, prob, nom, om,    \
                       chiA, chiB, zidx, RC, RB, nspec, ctl + 1);
#define RD(N)                                                                                      \
  if (nchr == N) {                                                                                 \
    const int gx = (nom + (64 / N) - 1) / (64 / N);                                                \
    const int grid = nom * RC + (gx * nspec * 2 * RB + 3) / 4;                                     \
    RDG(N, 0) RDG(N, 1) RDG(N, 2) RDG(N, 3)                                                        \
    ok = true;                                                                                     \
  }
  if (lgen < 0 || lgen > 3) return hipErrorInvalidValue;
  RD(4) RD(8) RD(16) RD(32) RD(64)
#undef RD
#undef RDG
  if (!ok) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_assemble_osc, dim3(1), dim3(256), 0, st, prob, nom, RC, RB, om, chiA, chiB,
                     det, nspec, S, om, ctl, ctl + 1);
  return hipGetLastError();
}

hipError_t launch_resonant_all(hipStream_t st, int nch, int RB, const ProbDev *prob, int nspec,
                               int nom, const double2 *om, double2 *chiB, int lgen) {
  if (nom <= 0) return hipSuccess;
  if (lgen < 0 || lgen > 3) return hipErrorInvalidValue;
#define RAG(N, G)                                                                              \
  if (lgen == G)                                                                               \
    hipLaunchKernelGGL((k_resonant_all<N, G>), dim3(grid, nspec, 2 * RB), dim3(64), 0, st,     \
                       prob, nom, om, chiB, RB);
#define RA(N)                                                                                  \
  if (nch == N) {                                                                              \
    const int grid = (nom + (64 / N) - 1) / (64 / N);                                          \
    RAG(N, 0) RAG(N, 1) RAG(N, 2) RAG(N, 3)                                                    \
    return hipGetLastError();                                                                  \
  }
  RA(4) RA(8) RA(16) RA(32) RA(64)
#undef RA
#undef RAG
  return hipErrorInvalidValue;
}

// the assembly of at most 256 omegas in one workgroup, then one secant step
// of each of their solves (om is read as this round's frequencies and
// overwritten with the next ones)
hipError_t launch_assemble_osc(hipStream_t st, const ProbDev *prob, int nspec, int nom, int W, int RB,
                               double2 *om, const double2 *chiA, const double2 *chiB, double2 *det,
                               alpsg_osc::OscMachine *S, int *round_ctr, int *remaining) {
  if (nom <= 0) return hipSuccess;
  if (nom > 256 || 6 * nspec > 256) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_assemble_osc, dim3(1), dim3(256), 0, st, prob, nom, W, RB, om, chiA, chiB, det,
                     nspec, S, om, round_ctr, remaining);
  return hipGetLastError();
}

hipError_t launch_osc_round(hipStream_t st, int n, alpsg_osc::OscMachine *S, const double2 *det,
                            double2 *om, int *round_ctr, int *remaining) {
  if (n <= 0) return hipSuccess;
  const int threads = std::min(1024, (n + 63) / 64 * 64);
  hipLaunchKernelGGL(k_osc_round, dim3(1), dim3(threads), 0, st, n, S, det, om, round_ctr, remaining);
  return hipGetLastError();
}
