// Host side of the C-ABI declared in include/alpsg.h.
//
// A handle owns one HIP stream, the device copies of the plasma (grids, f0,
// df0, fit tables), the per-k tables and the staging buffers of the batched
// entry points.  Nothing here evaluates physics on the CPU: every number that
// reaches det D is computed by the kernels of alpsg_tables.hip and
// alpsg_disp.hip; the host only moves data, sequences launches and runs the
// scalar control flow of the root searches (secant_osc state, map minima,
// k-path stepping), which the reference also runs on one rank.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <memory>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <string>
#include <vector>

#include "../../include/alpsg.h"
#include "alpsg_analyt.h"
#include "alpsg_device.h"
#include "alpsg_guard.h"
#include "alpsg_lean.h"
#include "alpsg_osc.h"
#include "alpsg_relhost.h"

// kernels (alpsg_tables.hip, alpsg_disp.hip)
hipError_t launch_derivative(hipStream_t st, int nspec, int nperp, int npar, const double *f0,
                             const double *pperp, const double *ppar, double *df1, double *df2);
hipError_t launch_int_ee(hipStream_t st, int nspec, int nperp, int npar, const double *pperp,
                         const double *ppar, const double *df1, const double *df2,
                         const double *qs, const double *ms, double *out);
hipError_t launch_bessel_max(hipStream_t st, int nspec, int nperp, int ntry, double kperp,
                             const double *pperp, const double *qs, const int *mtab, double *out);
hipError_t launch_species_k(hipStream_t st, int nperp, int npar, int nmax, double kperp,
                            double kpar, double q, double m, const double *pperp,
                            const double *ppar, const double *df1, const double *df2,
                            const int *mtab, double *bes, double *atab, double *nodes);
hipError_t launch_sweep(hipStream_t st, int nch, int W, int RC, const ProbDev *prob, int nom,
                        const double2 *om, const int *active, double2 *chiA, double2 *lowA,
                        int npar1, const int *pidx = nullptr, unsigned long long *work = nullptr);
hipError_t launch_sweep_req(hipStream_t st, int nch, int W, int RC, int K, const ProbDev *prob,
                            int nom, const double2 *om, double2 *chiA, const int *pidx,
                            bool reduce);
hipError_t launch_resonant(hipStream_t st, int nch, int RB, const ProbDev *prob, int nspec, int nom,
                           const double2 *om, const int *active, double2 *chiB, double2 *lowB,
                           int lgen, const int *pidx = nullptr, int P = 0);
hipError_t launch_nhds(hipStream_t st, const ProbDev *prob, int nspec, int nom, const double2 *om,
                       const int *active, double2 *chiA, double2 *lowA, const int *pidx = nullptr);
int nhds_tables(double z, int bm_nmax, double bessel_zero, std::vector<double2> &tab);
hipError_t launch_lsum(hipStream_t st, int nperp, int nmax, double dpp, const double *pperp,
                       const double *fitc, const double *atab, double *lsum);
hipError_t launch_fill_i32(hipStream_t st, int *p, int n, int v);
hipError_t launch_split_tables(hipStream_t st, int npar, int nmax, double kpar, double q, double m,
                               const double *ppar, const double *nodes, double *crec,
                               double *rsum, double *gcum);
hipError_t launch_series_tables(hipStream_t st, int npar, int nmax, double kpar, double q, double m,
                                const double *ppar, const double *nodes, double *ser, double *serD);
hipError_t launch_rel_bessel(hipStream_t st, int npts, int nmax, double zfac, const double *grel,
                             const double *pbrel, const int *mtab, double *out);
hipError_t launch_map_post(hipStream_t st, int nr, int ni, const double2 *cal, double2 *calo,
                           double *val, unsigned char *hit);
hipError_t launch_map_grid(hipStream_t st, int nr, int ni, double omi, double dr, double gami,
                           double di, double2 *om);
hipError_t launch_osc_round(hipStream_t st, int n, alpsg_osc::OscMachine *S, const double2 *det,
                            double2 *om, int *round_ctr, int *remaining);
hipError_t launch_resonant_all(hipStream_t st, int nch, int RB, const ProbDev *prob, int nspec,
                               int nom, const double2 *om, double2 *chiB, int lgen);
hipError_t launch_round(hipStream_t st, int nchr, int RB, int RC, const ProbDev *prob, int nspec,
                        int nom, double2 *om, double2 *chiA, double2 *chiB, double2 *det,
                        const int *zidx, int lgen, alpsg_osc::OscMachine *S, int *ctl);
hipError_t launch_assemble_osc(hipStream_t st, const ProbDev *prob, int nspec, int nom, int W, int RB,
                               double2 *om, const double2 *chiA, const double2 *chiB, double2 *det,
                               alpsg_osc::OscMachine *S, int *round_ctr, int *remaining);
hipError_t launch_rel(hipStream_t st, const ProbDev *prob, int nspec, int nmaxR, int NB, int nom,
                      const double2 *om, const int *active, double2 *chiR, double2 *lowR,
                      int *flags, int *err, const int *pidx = nullptr);
int rel_slots(int NB, int nmaxR);
hipError_t launch_assemble(hipStream_t st, const ProbDev *prob, int nspec, int nom, int W, int RB,
                           const double2 *om, const int *active, const double2 *chiA,
                           const double2 *chiB, const double2 *lowA, const double2 *lowB,
                           const double2 *chiR, int RR, const double2 *lowR, double2 *det,
                           double2 *chi, double2 *low, double2 *wave, const int *pidx = nullptr);

namespace {

thread_local std::string g_err;
const int MTAB_N = 4096;

int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess)                                                             \
      return fail(ALPSG_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t cnt) {
    if (cnt <= n && p) return hipSuccess;
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipMalloc((void **)&p, std::max<size_t>(cnt, 1) * sizeof(T));
    if (e == hipSuccess) n = std::max<size_t>(cnt, 1);
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
};

// pinned host staging for the small transfers of every disp call
struct PinBuf {
  double2 *p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t cnt) {
    if (cnt <= n && p) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipHostMalloc((void **)&p, std::max<size_t>(cnt, 1) * sizeof(double2), hipHostMallocDefault);
    if (e == hipSuccess) n = std::max<size_t>(cnt, 1);
    return e;
  }
  ~PinBuf() {
    if (p) (void)hipHostFree(p);
  }
};

// pinned host bytes, grown on demand (the map's grid and outputs)
struct PinBytes {
  unsigned char *p = nullptr;
  size_t n = 0;
  hipError_t ensure(size_t cnt) {
    if (cnt <= n && p) return hipSuccess;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    n = 0;
    hipError_t e = hipHostMalloc((void **)&p, std::max<size_t>(cnt, 1), hipHostMallocDefault);
    if (e == hipSuccess) n = std::max<size_t>(cnt, 1);
    return e;
  }
  ~PinBytes() {
    if (p) (void)hipHostFree(p);
  }
};

}  // namespace

// the error text of the calling thread, for the other translation units
int alpsg_internal_fail(int code, const std::string &msg) { return fail(code, msg); }

struct alpsg_ctx {
  int device = 0;  // the device current at alpsg_create, which owns the stream and tables
  hipStream_t st = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t ev_sweep = nullptr, ev_res = nullptr;  // recorded after k_sweep, after k_resonant
  hipEvent_t ev_done = nullptr;                     // end of a small disp call's work
  // alpsg_map_search_refine: the map's D and log10|D| go to the host on their
  // own stream (default priority) while the secant rounds run on st
  hipStream_t st_copy = nullptr;
  hipEvent_t ev_map = nullptr, ev_copy = nullptr;
  int nspec = 0, nperp = 0, npar = 0;
  int kperp_norm = 1, P = 5, NRI = 100;
  double vA = 0, Tlim = 0.01, Bessel_zero = 1e-45;
  bool have_k = false;
  double kperp = 0, kpar = 0;
  int nmax[ALPSG_MAXSPEC] = {0};
  long long evals = 0;
  double last_ms = 0.0;
  double last_split[3] = {0.0, 0.0, 0.0};  // k_sweep, k_resonant (all parts), k_assemble
  // host copies
  std::vector<double> qs, ms, ns, pperp_h, ppar_h;
  // static device data
  DevBuf<double> pperp, ppar, f0, df1, df2, qs_d, ms_d, int_ee, fitc, polyc;
  DevBuf<int> mtab;
  std::vector<double> int_ee_h;
  // per-k device data
  DevBuf<double> bes[ALPSG_MAXSPEC], atab[ALPSG_MAXSPEC], nodes[ALPSG_MAXSPEC];
  DevBuf<double> crec[ALPSG_MAXSPEC], rsum[ALPSG_MAXSPEC];  // split-form sweep tables
  DevBuf<double> gcum[ALPSG_MAXSPEC];  // cumulative G/m sums (partial-range split rows)
  DevBuf<double> ser[ALPSG_MAXSPEC], serD[ALPSG_MAXSPEC];  // far-row series tables
  bool series_on = true;    // far rows of the node sweep take their series
  bool ser_pending = false;  // the current k's series tables are not built yet
  bool ser_built = false;    // they are (the buffers hold this k's tables)
  DevBuf<double> lsum[ALPSG_MAXSPEC];  // Landau sums of single-Maxwellian species
  DevBuf<int> zidx;  // zeros: the problem index of every omega of a small batch
  DevBuf<unsigned long long> work_d;  // the sweep's work counts (alpsg_sweep_work)
  PinBuf om_pin, det_pin;  // staging of om and det for batches of at most PIN_MAX omegas
  // captured disp chains of small batches (index n = 1..REQ_MAX_N), with the
  // shape and buffers they were captured with; graphs_ok turns off after a
  // failed capture
  struct Graph {
    hipGraphExec_t exec = nullptr;
    int shape[5] = {0, 0, 0, 0, 0};
    const void *ptrs[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  } graphs[17];
  bool graphs_ok = true;
  DevBuf<double> bmax;
  // problem descriptor
  ProbDev prob{};
  DevBuf<ProbDev> prob_d;
  // batch staging
  DevBuf<double2> om_d, det_d, chi_d, low_d, wave_d, chiA_d, chiB_d, lowA_d, lowB_d;
  DevBuf<int> act_d;
  // map_search outputs on the device: D with sentinels, log10|D|, minimum flags
  DevBuf<double2> mcal_d;
  DevBuf<double> mval_d;
  DevBuf<unsigned char> mhit_d;
  PinBytes map_pin;  // the grid, then D, log10|D| and the minimum flags
  // secant_osc rounds on the device: the solves' state machines, the round
  // counter and unfinished count ([0], [1]) and its pinned copy, and graphs of
  // OSC_K rounds keyed by batch size, launch shape and buffers
  DevBuf<alpsg_osc::OscMachine> osc_d;
  DevBuf<int> oscctl_d;
  int *oscctl_pin = nullptr;
  struct OscGraph {
    hipGraphExec_t exec = nullptr;
    int n = 0;
    int shape[5] = {0, 0, 0, 0, 0};
    const void *ptrs[9] = {nullptr};
  } osc_graphs[8];
  int osc_graph_next = 0;
  bool osc_graphs_ok = true;
  std::vector<double> root_ms;  // completion time of each root of the last secant batch
  std::vector<double> scan_ms;  // end of each k step of the last om_scan (ms from its start)
  std::vector<int> scan_iters;  // [step][root] iterations of its solves, -1 if not followed
  bool plain_roots = false;     // root searches step on the host, one disp launch per omega
  // Landau forms the species take: bit 0 a call of eval_fit, bit 1 the inlined
  // kappa evaluation (species whose every fit is a kappa fit)
  int lgen = 0;
  // closed-form (use_bM) species and their &bM_spec parameters
  int usebm[ALPSG_MAXSPEC] = {0};
  bool any_bm = false;
  int bm_nmax[ALPSG_MAXSPEC] = {0};
  double bm_bz[ALPSG_MAXSPEC] = {0}, bm_beta[ALPSG_MAXSPEC] = {0}, bm_alpha[ALPSG_MAXSPEC] = {0},
         bm_pdrift[ALPSG_MAXSPEC] = {0};
  DevBuf<double2> bmi[ALPSG_MAXSPEC];
  // relativistic species: (Gamma, pparbar) tables, per-k J_n tables, Gamma(m)
  int relsp[ALPSG_MAXSPEC] = {0};
  bool any_rel = false;
  int ngamma = 0, npparbar = 0;
  DevBuf<double> grel[ALPSG_MAXSPEC], pbrel[ALPSG_MAXSPEC], f0rel[ALPSG_MAXSPEC],
      relbes[ALPSG_MAXSPEC], gampp[ALPSG_MAXSPEC];
  DevBuf<double2> dfrel[ALPSG_MAXSPEC];
  DevBuf<int2> rowlim[ALPSG_MAXSPEC];
  DevBuf<double> gtab;
  DevBuf<double2> chiR_d, lowR_d;
  DevBuf<int> flags_d, err_d;
  int rel_NB = 1;
  std::vector<double> current_int;  // parallel current density per species
  bool have_current = false;
  // launch shapes chosen by alpsg_autotune (or pinned by alpsg_set_shape),
  // keyed by the batch-size bucket floor(log2(n))
  int tuned_valid[32] = {0};
  int tuned_shape[32][5] = {{0}};

  ~alpsg_ctx() {
    if (st) (void)hipStreamSynchronize(st);
    DevBuf<double> *all[] = {&pperp, &ppar, &f0, &df1, &df2, &qs_d, &ms_d, &int_ee, &fitc, &polyc,
                             &bmax};
    for (auto *b : all) b->release();
    for (auto &g : graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    for (auto &g : osc_graphs)
      if (g.exec) (void)hipGraphExecDestroy(g.exec);
    osc_d.release();
    oscctl_d.release();
    if (oscctl_pin) (void)hipHostFree(oscctl_pin);
    work_d.release();
    zidx.release();
    mcal_d.release();
    mval_d.release();
    mhit_d.release();
    for (int s = 0; s < ALPSG_MAXSPEC; ++s) {
      bes[s].release();
      atab[s].release();
      nodes[s].release();
      crec[s].release();
      rsum[s].release();
      gcum[s].release();
      ser[s].release();
      serD[s].release();
      bmi[s].release();
      grel[s].release();
      pbrel[s].release();
      f0rel[s].release();
      relbes[s].release();
      gampp[s].release();
      dfrel[s].release();
      rowlim[s].release();
    }
    mtab.release();
    prob_d.release();
    om_d.release();
    det_d.release();
    chi_d.release();
    low_d.release();
    wave_d.release();
    chiA_d.release();
    chiB_d.release();
    lowA_d.release();
    lowB_d.release();
    act_d.release();
    gtab.release();
    chiR_d.release();
    lowR_d.release();
    flags_d.release();
    err_d.release();
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_sweep) (void)hipEventDestroy(ev_sweep);
    if (ev_res) (void)hipEventDestroy(ev_res);
    if (ev_done) (void)hipEventDestroy(ev_done);
    if (ev_map) (void)hipEventDestroy(ev_map);
    if (ev_copy) (void)hipEventDestroy(ev_copy);
    if (st_copy) (void)hipStreamDestroy(st_copy);
    if (st) (void)hipStreamDestroy(st);
  }
};

namespace {

// Fortran column-major index helpers
inline size_t f_pp(int nspec, int nperp, int npar, int s, int iperp, int ipar, int c) {
  return (size_t)s + (size_t)nspec * (iperp + (size_t)(nperp + 1) * (ipar + (size_t)(npar + 1) * c));
}

// Launch shape of the three disp kernels for a batch of n omegas: lanes per
// omega in the node sweep (nch_a), its wavefronts per workgroup (w_a) and
// workgroups per omega block (rc_a); lanes per omega in the resonant kernel
// (nch_b) and its harmonic blocks (rb).
struct LaunchShape {
  int nch_a, w_a, rc_a, nch_b, rb;
};

int size_bucket(int n) {
  int b = 0;
  while ((2 << b) <= n && b < 31) ++b;
  return b;
}

LaunchShape pick_shape(int n) {
  if (n >= 32768) return {1, 4, 1, 16, 1};
  if (n >= 8192) return {2, 4, 1, 16, 1};
  if (n >= 2048) return {8, 4, 1, 32, 1};
  if (n >= 256) return {16, 8, 1, 32, 2};
  if (n >= 32) return {32, 8, 2, 64, 4};
  return {64, 8, 4, 64, 8};
}

LaunchShape shape_for(const alpsg_ctx *h, int n) {
  const int b = size_bucket(n);
  if (h->tuned_valid[b]) {
    const int *t = h->tuned_shape[b];
    return LaunchShape{t[0], t[1], t[2], t[3], t[4]};
  }
  return pick_shape(n);
}

// the small-batch sweep (k_sweep_req<REQ_NCH, REQ_W, REQ_K> over REQ_TW / REQ_W
// workgroups per omega) for batches of at most REQ_MAX_N omegas
constexpr int REQ_MAX_N = 16, REQ_NCH = 32, REQ_W = 4, REQ_TW = 64, REQ_K = 4;
constexpr int PIN_MAX = 4096;  // batches up to this size copy om and det through pinned memory

// (lanes per omega, wavefronts per workgroup) pairs instantiated for k_sweep
const int kSweepVariants[][2] = {{1, 1}, {1, 2}, {1, 4}, {2, 4}, {4, 4}, {8, 4},
                                 {16, 4}, {16, 8}, {32, 8}, {64, 4}, {64, 8}};
const int kResonantLanes[] = {4, 8, 16, 32, 64};

bool shape_is_valid(const int *t) {
  bool a = false, b = false;
  for (auto &v : kSweepVariants) a = a || (v[0] == t[0] && v[1] == t[1]);
  for (int v : kResonantLanes) b = b || (v == t[3]);
  return a && b && t[2] >= 1 && t[2] <= 64 && t[4] >= 1 && t[4] <= 16 && t[1] * t[2] <= 128;
}

// The device buffers of one disp chain (det, the partial-chi slots, and for
// small batches the zero problem indices of the request sweep).
struct DispBufs {
  double2 *det, *chiA, *chiB, *chi, *low, *lowA, *lowB, *wave;
  const int *zidx;
};

// The disp kernels on n omegas resident in om_dev, into the buffers of B
// (sized by the caller; nothing is allocated here).  Relativistic species take
// the handle's own buffers, so callers with B outside the handle exclude them.
int disp_chain(alpsg_ctx *h, int n, const double2 *om_dev, const int *active_dev,
               const DispBufs &B, bool want_chi, bool want_low, bool want_wave, LaunchShape sh) {
  const int ns = h->nspec;
  const int TW = sh.w_a * sh.rc_a;
  // Small batches (a few roots' secant steps, the one-omega plain path): the
  // sweep that keeps several node records in flight per lane and spreads each
  // omega's row pairs over REQ_TW wavefronts (problem index 0 for every omega).
  // The n_s-row partial layout is the same, so the rest of the chain is unchanged.
  const bool req = !want_low && active_dev == nullptr && n <= REQ_MAX_N;
  if (req) {
    HIPCHK(launch_sweep_req(h->st, REQ_NCH, REQ_W, REQ_TW / REQ_W, REQ_K, h->prob_d.p, n, om_dev,
                            B.chiA, B.zidx, true));
  } else {
    HIPCHK(launch_sweep(h->st, sh.nch_a, sh.w_a, sh.rc_a, h->prob_d.p, n, om_dev, active_dev,
                        B.chiA, want_low ? B.lowA : nullptr, h->npar + 1));
  }
  const int TWa = req ? REQ_TW / REQ_W : TW;  // the request sweep writes one partial per workgroup
  HIPCHK(hipEventRecord(h->ev_sweep, h->st));
  if (req)
    HIPCHK(launch_resonant_all(h->st, sh.nch_b, sh.rb, h->prob_d.p, ns, n, om_dev, B.chiB,
                               h->lgen));
  else
    HIPCHK(launch_resonant(h->st, sh.nch_b, sh.rb, h->prob_d.p, ns, n, om_dev, active_dev, B.chiB,
                           want_low ? B.lowB : nullptr, h->lgen, nullptr, h->P));
  if (h->any_bm)
    HIPCHK(launch_nhds(h->st, h->prob_d.p, ns, n, om_dev, active_dev, B.chiA,
                       want_low ? B.lowA : nullptr));
  int RR = 0;
  if (h->any_rel) {
    const int nmaxR = h->prob.nmaxR;
    const int NB = std::max(1, std::min(2 * nmaxR + 1, 1024 / std::max(n, 1)));
    RR = rel_slots(NB, nmaxR);
    HIPCHK(h->chiR_d.ensure((size_t)RR * ns * 6 * n));
    HIPCHK(h->flags_d.ensure((size_t)n * ns * (2 * nmaxR + 1)));
    if (want_low) HIPCHK(h->lowR_d.ensure((size_t)n * ns * 18));
    HIPCHK(launch_rel(h->st, h->prob_d.p, ns, nmaxR, NB, n, om_dev, active_dev, h->chiR_d.p,
                      want_low ? h->lowR_d.p : nullptr, h->flags_d.p, h->err_d.p));
  }
  HIPCHK(hipEventRecord(h->ev_res, h->st));
  HIPCHK(launch_assemble(h->st, h->prob_d.p, ns, n, TWa, sh.rb, om_dev, active_dev, B.chiA, B.chiB,
                         want_low ? B.lowA : nullptr, want_low ? B.lowB : nullptr,
                         RR ? h->chiR_d.p : nullptr, RR, (RR && want_low) ? h->lowR_d.p : nullptr,
                         B.det, want_chi ? B.chi : nullptr, want_low ? B.low : nullptr,
                         want_wave ? B.wave : nullptr));
  return ALPSG_OK;
}

// Runs the disp kernels on n omegas already resident in om_dev, in the
// handle's buffers (grown on demand).
// Batches of at least SER_BUILD_MIN omegas build the current k's far-row
// series tables (k_series, a few microseconds) if they are not built yet; the
// descriptor goes to the device again with their pointers.  Smaller batches
// (secant steps, scans: a few evaluations per k) use them once they exist.
constexpr int SER_BUILD_MIN = 64;

int ensure_series(alpsg_ctx *h, int n) {
  if (!h->ser_pending || !h->series_on || n < SER_BUILD_MIN) return ALPSG_OK;
  const double kpar = h->prob.kpar;
  for (int s = 0; s < h->nspec; ++s) {
    SpecDev &S = h->prob.sp[s];
    if (S.crec == nullptr || h->ser[s].p == nullptr) continue;
    HIPCHK(launch_series_tables(h->st, S.npar, S.nmax, kpar, h->qs[s], h->ms[s], S.ppar, S.nodes,
                                h->ser[s].p, h->serD[s].p));
    S.ser = h->ser[s].p;
    S.serD = h->serD[s].p;
  }
  HIPCHK(hipMemcpyAsync(h->prob_d.p, &h->prob, sizeof(ProbDev), hipMemcpyHostToDevice, h->st));
  h->ser_pending = false;
  h->ser_built = true;
  return ALPSG_OK;
}

int disp_device(alpsg_ctx *h, int n, const double2 *om_dev, const int *active_dev, bool want_chi,
                bool want_low, bool want_wave, LaunchShape sh) {
  if (const int rc = ensure_series(h, n)) return rc;
  const int ns = h->nspec;
  const int TW = sh.w_a * sh.rc_a;
  HIPCHK(h->det_d.ensure(n));
  HIPCHK(h->chiA_d.ensure((size_t)n * TW * ns * 6));
  HIPCHK(h->chiB_d.ensure((size_t)n * ns * 2 * sh.rb * 6));
  if (want_chi) HIPCHK(h->chi_d.ensure((size_t)n * ns * 9));
  if (want_low) {
    HIPCHK(h->low_d.ensure((size_t)n * ns * 27));
    HIPCHK(h->lowA_d.ensure((size_t)n * ns * 18));
    HIPCHK(h->lowB_d.ensure((size_t)n * ns * 2 * sh.rb * 18));
  }
  if (want_wave) HIPCHK(h->wave_d.ensure((size_t)n * 9));
  const bool req = !want_low && active_dev == nullptr && n <= REQ_MAX_N;
  if (req) {
    HIPCHK(h->chiA_d.ensure((size_t)n * REQ_TW * ns * 6));
    if (h->zidx.n < (size_t)REQ_MAX_N) {
      HIPCHK(h->zidx.ensure(REQ_MAX_N));
      HIPCHK(hipMemsetAsync(h->zidx.p, 0, sizeof(int) * REQ_MAX_N, h->st));
    }
  }
  DispBufs B{h->det_d.p, h->chiA_d.p, h->chiB_d.p, h->chi_d.p, h->low_d.p, h->lowA_d.p,
             h->lowB_d.p, h->wave_d.p, h->zidx.p};
  return disp_chain(h, n, om_dev, active_dev, B, want_chi, want_low, want_wave, sh);
}

}  // namespace

// det D for n omegas with every device buffer carved from the caller's
// scratch (scratch_bytes at scratch_dev, on the handle's device): the chain
// and launch shape of alpsg_disp_batch, nothing allocated.  Layout, each part
// 256-byte aligned: om [n], det [n], chiA [n TW nspec 6], chiB [n nspec 2 RB 6]
// (double2), and for n <= REQ_MAX_N the REQ_MAX_N zero problem indices.
int alpsg_disp_on_scratch(alpsg_ctx *h, int n, const double *om, double *det,
                          unsigned char *scratch_dev, size_t scratch_bytes) {
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  if (n < 0) return fail(ALPSG_E_ARG, "negative batch size");
  if (n == 0) return ALPSG_OK;
  if (h->any_rel || h->any_bm)
    return fail(ALPSG_E_UNSUPPORTED,
                "the no-allocation disp takes tabulated non-relativistic species only");
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev != h->device) return fail(ALPSG_E_STATE, "call on the device the handle was created on");
  const LaunchShape sh = shape_for(h, n);
  const int ns = h->nspec;
  const bool req = n <= REQ_MAX_N;
  const size_t TW = req ? (size_t)REQ_TW : (size_t)sh.w_a * sh.rc_a;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  const size_t b_om = up(sizeof(double2) * n), b_det = b_om;
  const size_t b_a = up(sizeof(double2) * (size_t)n * TW * ns * 6);
  const size_t b_b = up(sizeof(double2) * (size_t)n * ns * 2 * sh.rb * 6);
  const size_t b_z = req ? up(sizeof(int) * REQ_MAX_N) : 0;
  const size_t need = b_om + b_det + b_a + b_b + b_z;
  if (need > scratch_bytes)
    return fail(ALPSG_E_ARG, "batch of " + std::to_string(n) + " omegas needs " +
                                 std::to_string(need) + " bytes of scratch; the no-allocation path has " +
                                 std::to_string(scratch_bytes));
  unsigned char *q = scratch_dev;
  double2 *om_d = reinterpret_cast<double2 *>(q);
  q += b_om;
  double2 *det_d = reinterpret_cast<double2 *>(q);
  q += b_det;
  double2 *chiA = reinterpret_cast<double2 *>(q);
  q += b_a;
  double2 *chiB = reinterpret_cast<double2 *>(q);
  q += b_b;
  int *zidx = req ? reinterpret_cast<int *>(q) : nullptr;
  HIPCHK(hipMemcpyAsync(om_d, om, sizeof(double2) * n, hipMemcpyHostToDevice, h->st));
  if (req) HIPCHK(launch_fill_i32(h->st, zidx, REQ_MAX_N, 0));
  const DispBufs B{det_d, chiA, chiB, nullptr, nullptr, nullptr, nullptr, nullptr, zidx};
  const int rc = disp_chain(h, n, om_d, nullptr, B, false, false, false, sh);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(det, det_d, sizeof(double2) * n, hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->evals += n;
  return ALPSG_OK;
}

namespace {

// Autotuner: times the node-sweep variants (lanes per omega x wavefronts per
// workgroup x workgroups per omega block) and the resonant-kernel variants
// (lanes per omega x harmonic blocks) separately on the batch, keeps the
// fastest of each for the batch-size bucket of n.
int autotune(alpsg_ctx *h, int n, const double2 *om_host, int *shape_out) {
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  if (n < 1) return fail(ALPSG_E_ARG, "n must be >= 1");
  if (const int rc = ensure_series(h, n)) return rc;  // the variants are timed as they will run
  std::vector<double2> om(n);
  if (om_host) {
    for (int i = 0; i < n; ++i) om[i] = om_host[i];
  } else {
    // a spread of frequencies: resonant and non-resonant rows, both signs of Im
    for (int i = 0; i < n; ++i) {
      const double t = (i + 0.5) / n;
      om[i] = make_double2(0.05 + 1.95 * t, -0.3 + 0.32 * std::fmod(7.0 * t, 1.0));
    }
  }
  const int ns = h->nspec;
  HIPCHK(h->om_d.ensure(n));
  HIPCHK(hipMemcpyAsync(h->om_d.p, om.data(), sizeof(double2) * n, hipMemcpyHostToDevice, h->st));
  const size_t budget = (size_t)512 << 20;
  auto time_launch = [&](auto &&launch, float &best) -> int {
    HIPCHK(launch());  // warm-up
    float tot = 0.f;
    for (int r = 0; r < 2; ++r) {
      HIPCHK(hipEventRecord(h->ev0, h->st));
      HIPCHK(launch());
      HIPCHK(hipEventRecord(h->ev1, h->st));
      HIPCHK(hipEventSynchronize(h->ev1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
      tot += ms;
    }
    best = tot / 2;
    return ALPSG_OK;
  };
  std::vector<std::pair<float, LaunchShape>> candA, candB;
  const int rcs[] = {1, 2, 3, 4, 6, 8, 12, 16};
  for (auto &v : kSweepVariants) {
    const int omb = 64 / v[0];
    const int blocks = (n + omb - 1) / omb;
    for (int rc : rcs) {
      const int TW = v[1] * rc;
      if (TW > 128) continue;
      if ((size_t)n * TW * ns * 6 * sizeof(double2) > budget) continue;
      if (rc > 1 && (size_t)blocks * rc > 65536) continue;
      HIPCHK(h->chiA_d.ensure((size_t)n * TW * ns * 6));
      float t;
      int rc_ = time_launch(
          [&]() {
            return launch_sweep(h->st, v[0], v[1], rc, h->prob_d.p, n, h->om_d.p, nullptr,
                                h->chiA_d.p, nullptr, h->npar + 1);
          },
          t);
      if (rc_) return rc_;
      candA.push_back({t, LaunchShape{v[0], v[1], rc, 0, 0}});
    }
  }
  const int rbs[] = {1, 2, 4, 8};
  for (int nb : kResonantLanes)
    for (int rb : rbs) {
      if ((size_t)n * ns * 2 * rb * 6 * sizeof(double2) > budget) continue;
      HIPCHK(h->chiB_d.ensure((size_t)n * ns * 2 * rb * 6));
      float t;
      int rc_ = time_launch(
          [&]() {
            return launch_resonant(h->st, nb, rb, h->prob_d.p, ns, n, h->om_d.p, nullptr,
                                   h->chiB_d.p, nullptr, h->lgen, nullptr, h->P);
          },
          t);
      if (rc_) return rc_;
      candB.push_back({t, LaunchShape{0, 0, 0, nb, rb}});
    }
  if (candA.empty() || candB.empty()) return fail(ALPSG_E_ARG, "batch too large to tune");
  auto by_time = [](const std::pair<float, LaunchShape> &a, const std::pair<float, LaunchShape> &b) {
    return a.first < b.first;
  };
  std::sort(candA.begin(), candA.end(), by_time);
  std::sort(candB.begin(), candB.end(), by_time);
  // the assemble pass reads W*RC + 2*RB partials per omega: settle the
  // choice among the leaders with the three kernels timed together
  LaunchShape best{};
  float tbest = 1e30f;
  for (size_t a = 0; a < std::min<size_t>(3, candA.size()); ++a)
    for (size_t b = 0; b < std::min<size_t>(2, candB.size()); ++b) {
      LaunchShape sh = candA[a].second;
      sh.nch_b = candB[b].second.nch_b;
      sh.rb = candB[b].second.rb;
      float t;
      int rc_ = time_launch(
          [&]() -> hipError_t {
            return disp_device(h, n, h->om_d.p, nullptr, false, false, false, sh) == ALPSG_OK
                       ? hipSuccess
                       : hipErrorUnknown;
          },
          t);
      if (rc_) return rc_;
      if (t < tbest) {
        tbest = t;
        best = sh;
      }
    }
  const int b = size_bucket(n);
  const int sh[5] = {best.nch_a, best.w_a, best.rc_a, best.nch_b, best.rb};
  for (int k = 0; k < 5; ++k) h->tuned_shape[b][k] = sh[k];
  h->tuned_valid[b] = 1;
  if (shape_out)
    for (int k = 0; k < 5; ++k) shape_out[k] = sh[k];
  return ALPSG_OK;
}

// det D of a small batch (n <= REQ_MAX_N, det only, no closed-form or
// relativistic species) by replaying the captured chain pinned copy in ->
// k_sweep_req -> k_resonant parts -> k_assemble -> pinned copy out.  Returns
// 1 when the chain cannot be captured (the caller then runs it eagerly).
int graph_disp(alpsg_ctx *h, int n, const double *om, double *det) {
  const int ns = h->nspec;
  LaunchShape sh = shape_for(h, n);
  const int shp[5] = {sh.nch_a, sh.w_a, sh.rc_a, sh.nch_b, sh.rb};
  // every buffer at its small-batch size before a capture (no allocation inside)
  HIPCHK(h->om_d.ensure(REQ_MAX_N));
  HIPCHK(h->det_d.ensure(REQ_MAX_N));
  HIPCHK(h->chiA_d.ensure((size_t)REQ_MAX_N * REQ_TW * ns * 6));
  HIPCHK(h->chiB_d.ensure((size_t)REQ_MAX_N * ns * 2 * 16 * 6));
  HIPCHK(h->om_pin.ensure(PIN_MAX));
  HIPCHK(h->det_pin.ensure(PIN_MAX));
  if (h->zidx.n < (size_t)REQ_MAX_N) {
    HIPCHK(h->zidx.ensure(REQ_MAX_N));
    HIPCHK(hipMemsetAsync(h->zidx.p, 0, sizeof(int) * REQ_MAX_N, h->st));
  }
  const void *ptrs[6] = {h->om_d.p, h->det_d.p, h->chiA_d.p, h->chiB_d.p, h->prob_d.p, h->zidx.p};
  alpsg_ctx::Graph &g = h->graphs[n];
  bool same = g.exec != nullptr;
  for (int k = 0; k < 5 && same; ++k) same = g.shape[k] == shp[k];
  for (int k = 0; k < 6 && same; ++k) same = g.ptrs[k] == ptrs[k];
  if (!same) {
    if (g.exec) {
      HIPCHK(hipStreamSynchronize(h->st));
      (void)hipGraphExecDestroy(g.exec);
      g.exec = nullptr;
    }
    HIPCHK(hipStreamSynchronize(h->st));
    if (hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal) != hipSuccess) return 1;
    hipError_t e = hipMemcpyAsync(h->om_d.p, h->om_pin.p, sizeof(double2) * n,
                                  hipMemcpyHostToDevice, h->st);
    if (e == hipSuccess)
      e = launch_sweep_req(h->st, REQ_NCH, REQ_W, REQ_TW / REQ_W, REQ_K, h->prob_d.p, n, h->om_d.p,
                           h->chiA_d.p, h->zidx.p, true);
    if (e == hipSuccess)
      e = launch_resonant_all(h->st, sh.nch_b, sh.rb, h->prob_d.p, ns, n, h->om_d.p, h->chiB_d.p,
                              h->lgen);
    if (e == hipSuccess)
      e = launch_assemble(h->st, h->prob_d.p, ns, n, REQ_TW / REQ_W, sh.rb, h->om_d.p, nullptr, h->chiA_d.p,
                          h->chiB_d.p, nullptr, nullptr, nullptr, 0, nullptr, h->det_d.p, nullptr,
                          nullptr, nullptr);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h->det_pin.p, h->det_d.p, sizeof(double2) * n, hipMemcpyDeviceToHost,
                         h->st);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(h->st, &graph);
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess && e2 == hipSuccess && graph != nullptr &&
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
      g.exec = exec;
      for (int k = 0; k < 5; ++k) g.shape[k] = shp[k];
      for (int k = 0; k < 6; ++k) g.ptrs[k] = ptrs[k];
    }
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (!g.exec) return 1;
  }
  std::memcpy(h->om_pin.p, om, sizeof(double2) * n);
  HIPCHK(hipGraphLaunch(g.exec, h->st));
  HIPCHK(hipEventRecord(h->ev_done, h->st));
  hipError_t q;
  while ((q = hipEventQuery(h->ev_done)) == hipErrorNotReady) {
  }
  HIPCHK(q);
  std::memcpy(det, h->det_pin.p, sizeof(double2) * n);
  h->evals += n;
  return ALPSG_OK;
}

int run_disp(alpsg_ctx *h, int n, const double *om, double *det, double *chi, double *low,
             double *wave) {
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  if (n < 0) return fail(ALPSG_E_ARG, "negative batch size");
  if (n == 0) return ALPSG_OK;
  if (h->graphs_ok && n <= REQ_MAX_N && !chi && !low && !wave && !h->any_bm && !h->any_rel) {
    const int rc = graph_disp(h, n, om, det);
    if (rc != 1) return rc;
    h->graphs_ok = false;  // capture is not available: run eagerly from now on
  }
  HIPCHK(h->om_d.ensure(n));
  const bool pin = n <= PIN_MAX;
  if (pin) {
    HIPCHK(h->om_pin.ensure(PIN_MAX));
    HIPCHK(h->det_pin.ensure(PIN_MAX));
    std::memcpy(h->om_pin.p, om, sizeof(double2) * n);
    HIPCHK(hipMemcpyAsync(h->om_d.p, h->om_pin.p, sizeof(double2) * n, hipMemcpyHostToDevice, h->st));
  } else {
    HIPCHK(hipMemcpyAsync(h->om_d.p, om, sizeof(double2) * n, hipMemcpyHostToDevice, h->st));
  }
  LaunchShape sh = shape_for(h, n);
  HIPCHK(hipEventRecord(h->ev0, h->st));
  int rc = disp_device(h, n, h->om_d.p, nullptr, chi != nullptr, low != nullptr, wave != nullptr, sh);
  if (rc) return rc;
  HIPCHK(hipEventRecord(h->ev1, h->st));
  HIPCHK(hipMemcpyAsync(pin ? (void *)h->det_pin.p : (void *)det, h->det_d.p, sizeof(double2) * n,
                        hipMemcpyDeviceToHost, h->st));
  if (chi)
    HIPCHK(hipMemcpyAsync(chi, h->chi_d.p, sizeof(double2) * n * h->nspec * 9,
                          hipMemcpyDeviceToHost, h->st));
  if (low)
    HIPCHK(hipMemcpyAsync(low, h->low_d.p, sizeof(double2) * n * h->nspec * 27,
                          hipMemcpyDeviceToHost, h->st));
  if (wave)
    HIPCHK(hipMemcpyAsync(wave, h->wave_d.p, sizeof(double2) * n * 9, hipMemcpyDeviceToHost,
                          h->st));
  int relerr = 0;
  if (h->any_rel)
    HIPCHK(hipMemcpyAsync(&relerr, h->err_d.p, sizeof(int), hipMemcpyDeviceToHost, h->st));
  if (pin && !h->any_rel && !chi && !low && !wave) {
    // a root-search step waits on this result: poll for it rather than sleep
    HIPCHK(hipEventRecord(h->ev_done, h->st));
    hipError_t q;
    while ((q = hipEventQuery(h->ev_done)) == hipErrorNotReady) {
    }
    HIPCHK(q);
  } else {
    HIPCHK(hipStreamSynchronize(h->st));
  }
  if (pin) std::memcpy(det, h->det_pin.p, sizeof(double2) * n);
  if (relerr) {
    const int zero = 0;
    HIPCHK(hipMemcpy(h->err_d.p, &zero, sizeof(int), hipMemcpyHostToDevice));
    return fail(ALPSG_E_ARG,
                "a resonance's principal-value window exceeds the subluminal part of a Gamma row "
                "(the reference stops with alps_error(8)); refine npparbar or lower "
                "positions_principal");
  }
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  h->last_ms = ms;
  float a = 0.f, b = 0.f, c = 0.f;
  HIPCHK(hipEventElapsedTime(&a, h->ev0, h->ev_sweep));
  HIPCHK(hipEventElapsedTime(&b, h->ev_sweep, h->ev_res));
  HIPCHK(hipEventElapsedTime(&c, h->ev_res, h->ev1));
  h->last_split[0] = a;
  h->last_split[1] = b;
  h->last_split[2] = c;
  h->evals += n;
  return ALPSG_OK;
}

typedef std::complex<double> cplx;

const int OSC_K = 8;         // secant rounds per graph replay
const int OSC_MAX_N = 1024;  // solves of one device-round batch (one workgroup steps them)

// secant_osc for many roots with the state machines on the device: a graph
// holds OSC_K rounds of [disp of every solve's requested frequency, one step of
// every unfinished solve], and the host only replays it until no solve is left.
// Every solve's frequency is evaluated in every round, also after it finished
// (its value is then ignored), so the batch and the launch shape stay fixed.
// Returns -1 when the batch cannot run this way (the caller then steps the
// solves on the host).
// work (may be NULL): host work of the caller, run once while the first rounds
// run on the device (then cleared)
int osc_graph_batch(alpsg_ctx *h, const alpsg_root_params *rp, int n, cplx *om, cplx *Dfin,
                    int *iters, std::function<int()> *work) {
  if (!h->graphs_ok || !h->osc_graphs_ok || h->any_bm || h->any_rel || n > OSC_MAX_N) return -1;
  const int ns = h->nspec;
  LaunchShape sh = shape_for(h, n);
  const bool req = n <= REQ_MAX_N;
  const int TW = req ? REQ_TW / REQ_W : sh.w_a * sh.rc_a;  // partial slots of the sweep
  const int shp[5] = {sh.nch_a, sh.w_a, sh.rc_a, sh.nch_b, sh.rb};
  HIPCHK(h->om_d.ensure(std::max(n, REQ_MAX_N)));
  HIPCHK(h->det_d.ensure(std::max(n, REQ_MAX_N)));
  HIPCHK(h->chiA_d.ensure((size_t)n * TW * ns * 6));
  HIPCHK(h->chiB_d.ensure((size_t)n * ns * 2 * sh.rb * 6));
  if (req && h->zidx.n < (size_t)REQ_MAX_N) {
    HIPCHK(h->zidx.ensure(REQ_MAX_N));
    HIPCHK(hipMemsetAsync(h->zidx.p, 0, sizeof(int) * REQ_MAX_N, h->st));
  }
  HIPCHK(h->osc_d.ensure(n));
  HIPCHK(h->oscctl_d.ensure(2));
  if (!h->oscctl_pin)
    HIPCHK(hipHostMalloc((void **)&h->oscctl_pin, 2 * sizeof(int), hipHostMallocDefault));
  std::vector<alpsg_osc::OscMachine> S(n);
  std::vector<double2> w(n);
  for (int r = 0; r < n; ++r) {
    S[r].start(alpsg_osc::zmk(om[r].real(), om[r].imag()), rp->numiter, rp->D_threshold, rp->D_prec);
    const alpsg_osc::zc x = S[r].want();
    w[r] = make_double2(x.re, x.im);
  }
  const int ctl0[2] = {0, n};  // round, unfinished solves
  HIPCHK(hipMemcpyAsync(h->osc_d.p, S.data(), sizeof(alpsg_osc::OscMachine) * n,
                        hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->om_d.p, w.data(), sizeof(double2) * n, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->oscctl_d.p, ctl0, sizeof(ctl0), hipMemcpyHostToDevice, h->st));
  const void *ptrs[9] = {h->om_d.p,  h->det_d.p,  h->chiA_d.p,    h->chiB_d.p,   h->prob_d.p,
                         req ? h->zidx.p : nullptr, h->osc_d.p, h->oscctl_d.p, h->oscctl_pin};
  alpsg_ctx::OscGraph *g = nullptr;
  for (auto &c : h->osc_graphs) {
    if (!c.exec || c.n != n) continue;
    bool same = true;
    for (int k = 0; k < 5 && same; ++k) same = c.shape[k] == shp[k];
    for (int k = 0; k < 9 && same; ++k) same = c.ptrs[k] == ptrs[k];
    if (same) {
      g = &c;
      break;
    }
  }
  if (!g) {
    g = &h->osc_graphs[h->osc_graph_next];
    h->osc_graph_next = (h->osc_graph_next + 1) % 8;
    HIPCHK(hipStreamSynchronize(h->st));
    if (g->exec) {
      (void)hipGraphExecDestroy(g->exec);
      g->exec = nullptr;
    }
    if (hipStreamBeginCapture(h->st, hipStreamCaptureModeThreadLocal) != hipSuccess) {
      (void)hipGetLastError();
      h->osc_graphs_ok = false;
      return -1;
    }
    hipError_t e = hipSuccess;
    for (int k = 0; k < OSC_K && e == hipSuccess; ++k) {
      if (req && REQ_W == 4 && REQ_NCH == 32 && REQ_K == 4) {
        // the whole round in two launches
        e = launch_round(h->st, sh.nch_b, sh.rb, REQ_TW / REQ_W, h->prob_d.p, ns, n, h->om_d.p,
                         h->chiA_d.p, h->chiB_d.p, h->det_d.p, h->zidx.p, h->lgen,
                         h->osc_d.p, h->oscctl_d.p);
        continue;
      }
      if (req)
        e = launch_sweep_req(h->st, REQ_NCH, REQ_W, REQ_TW / REQ_W, REQ_K, h->prob_d.p, n, h->om_d.p,
                             h->chiA_d.p, h->zidx.p, true);
      else
        e = launch_sweep(h->st, sh.nch_a, sh.w_a, sh.rc_a, h->prob_d.p, n, h->om_d.p, nullptr,
                         h->chiA_d.p, nullptr, h->npar + 1);
      if (n <= 256) {
        // two launches per round: both resonant parts, then the assembly and
        // the secant step in one workgroup
        if (e == hipSuccess)
          e = launch_resonant_all(h->st, sh.nch_b, sh.rb, h->prob_d.p, ns, n, h->om_d.p,
                                  h->chiB_d.p, h->lgen);
        if (e == hipSuccess)
          e = launch_assemble_osc(h->st, h->prob_d.p, ns, n, TW, sh.rb, h->om_d.p, h->chiA_d.p,
                                  h->chiB_d.p, h->det_d.p, h->osc_d.p, h->oscctl_d.p,
                                  h->oscctl_d.p + 1);
        continue;
      }
      if (e == hipSuccess)
        e = launch_resonant(h->st, sh.nch_b, sh.rb, h->prob_d.p, ns, n, h->om_d.p, nullptr,
                            h->chiB_d.p, nullptr, h->lgen, nullptr, h->P);
      if (e == hipSuccess)
        e = launch_assemble(h->st, h->prob_d.p, ns, n, TW, sh.rb, h->om_d.p, nullptr, h->chiA_d.p,
                            h->chiB_d.p, nullptr, nullptr, nullptr, 0, nullptr, h->det_d.p, nullptr,
                            nullptr, nullptr);
      if (e == hipSuccess)
        e = launch_osc_round(h->st, n, h->osc_d.p, h->det_d.p, h->om_d.p, h->oscctl_d.p,
                             h->oscctl_d.p + 1);
    }
    if (e == hipSuccess)
      e = hipMemcpyAsync(h->oscctl_pin, h->oscctl_d.p, 2 * sizeof(int), hipMemcpyDeviceToHost, h->st);
    hipGraph_t graph = nullptr;
    const hipError_t e2 = hipStreamEndCapture(h->st, &graph);
    hipGraphExec_t exec = nullptr;
    if (e == hipSuccess && e2 == hipSuccess && graph != nullptr &&
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess) {
      g->exec = exec;
      g->n = n;
      for (int k = 0; k < 5; ++k) g->shape[k] = shp[k];
      for (int k = 0; k < 9; ++k) g->ptrs[k] = ptrs[k];
    }
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (!g->exec) {
      h->osc_graphs_ok = false;
      return -1;
    }
  }
  // the bound on rounds: two initial evaluations and at most four per iteration
  const int maxrep = (2 + 4 * std::max(rp->numiter, 1)) / OSC_K + 2;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<double> trep;
  bool finished = false;
  for (int rep = 0; rep < maxrep; ++rep) {
    HIPCHK(hipGraphLaunch(g->exec, h->st));
    HIPCHK(hipEventRecord(h->ev_done, h->st));
    if (work && *work) {  // runs while the rounds just launched run on the device
      std::function<int()> w;
      w.swap(*work);
      const int rc = w();
      if (rc) {
        (void)hipStreamSynchronize(h->st);
        return rc;
      }
    }
    hipError_t q;
    while ((q = hipEventQuery(h->ev_done)) == hipErrorNotReady) {
    }
    HIPCHK(q);
    trep.push_back(std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count());
    if (h->oscctl_pin[1] == 0) {
      finished = true;
      break;
    }
  }
  // evaluations made: with the fused round (n <= REQ_MAX_N) the rounds after the
  // last solve finished return at once, and the round counter stops there
  const bool fused = req && REQ_W == 4 && REQ_NCH == 32 && REQ_K == 4;
  const long long rounds_run = fused ? (long long)h->oscctl_pin[0] : (long long)OSC_K * trep.size();
  h->evals += (long long)n * rounds_run;
  if (!finished) return fail(ALPSG_E_DEVICE, "secant rounds on the device did not finish");
  HIPCHK(hipMemcpy(S.data(), h->osc_d.p, sizeof(alpsg_osc::OscMachine) * n, hipMemcpyDeviceToHost));
  h->root_ms.assign(n, 0.0);
  for (int r = 0; r < n; ++r) {
    const alpsg_osc::zc x = S[r].result();
    om[r] = cplx(x.re, x.im);
    if (iters) iters[r] = S[r].iter;
    const int rep = std::min((int)trep.size() - 1, std::max(0, S[r].done_round) / OSC_K);
    h->root_ms[r] = trep[rep];
  }
  if (Dfin) {
    int rc = run_disp(h, n, reinterpret_cast<const double *>(om), reinterpret_cast<double *>(Dfin),
                      nullptr, nullptr, nullptr);
    if (rc) return rc;
  }
  return ALPSG_OK;
}

// secant_osc (ALPS_fns.f90:1919-2101) for many independent roots: each root is
// a state machine (alpsg_osc.h) whose disp() requests are served in batches.
int secant_osc_batch(alpsg_ctx *h, const alpsg_root_params *rp, int nroots, cplx *om, cplx *Dfin,
                     int *iters, std::function<int()> *work = nullptr) {
  if (!h->plain_roots) {
    const int rc = osc_graph_batch(h, rp, nroots, om, Dfin, iters, work);
    if (rc != -1) return rc;
  }
  std::vector<alpsg_osc::OscMachine> S(nroots);
  std::vector<cplx> req(nroots), res(nroots);
  std::vector<int> who;
  const auto t0 = std::chrono::steady_clock::now();
  h->root_ms.assign(nroots, 0.0);
  for (int r = 0; r < nroots; ++r)
    S[r].start(alpsg_osc::zmk(om[r].real(), om[r].imag()), rp->numiter, rp->D_threshold, rp->D_prec);
  for (;;) {
    who.clear();
    for (int r = 0; r < nroots; ++r) {
      if (S[r].done) continue;
      const alpsg_osc::zc w = S[r].want();
      req[who.size()] = cplx(w.re, w.im);
      who.push_back(r);
    }
    if (who.empty()) break;
    int rc = run_disp(h, (int)who.size(), reinterpret_cast<const double *>(req.data()),
                      reinterpret_cast<double *>(res.data()), nullptr, nullptr, nullptr);
    if (rc) return rc;
    const double now_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (size_t t = 0; t < who.size(); ++t) {
      h->root_ms[who[t]] = now_ms;
      S[who[t]].feed(alpsg_osc::zmk(res[t].real(), res[t].imag()));
    }
  }
  for (int r = 0; r < nroots; ++r) {
    const alpsg_osc::zc w = S[r].result();
    om[r] = cplx(w.re, w.im);
    if (iters) iters[r] = S[r].iter;
  }
  if (Dfin) {
    int rc = run_disp(h, nroots, reinterpret_cast<const double *>(om),
                      reinterpret_cast<double *>(Dfin), nullptr, nullptr, nullptr);
    if (rc) return rc;
  }
  return ALPSG_OK;
}

// secant (ALPS_fns.f90:1815-1917) for many independent roots.  Phases: 0 =
// need D(prevom) and 1 = need D(om) before the loop, 2 = need D(om) in an
// iteration, 3 = need D(prevom + 1e-8) after a stalled difference.
struct SecState {
  cplx om, prevom, Dprev, D, minom, minD;
  int iter, phase, done;
};

int secant_batch(alpsg_ctx *h, const alpsg_root_params *rp, int nroots, cplx *om, int *iters) {
  std::vector<SecState> S(nroots);
  std::vector<cplx> req(nroots), res(nroots);
  std::vector<int> who;
  const auto t0 = std::chrono::steady_clock::now();
  h->root_ms.assign(nroots, 0.0);
  for (int r = 0; r < nroots; ++r) {
    SecState &s = S[r];
    s.om = om[r];
    s.prevom = s.om * (1.0 - rp->D_prec);
    s.phase = 0;
    s.done = 0;
    s.iter = 0;
  }
  auto start_iter = [&](SecState &s) {
    if (s.iter <= rp->numiter - 1) {
      s.iter = s.iter + 1;
      s.phase = 2;
    } else {
      s.done = 1;
    }
  };
  auto step = [&](SecState &s) {
    cplx jump;
    bool go = true;
    if (std::abs(s.D) < rp->D_threshold) {
      jump = 0.0;
      go = false;
    } else {
      jump = s.D * (s.om - s.prevom) / (s.D - s.Dprev);
    }
    s.prevom = s.om;
    s.om = s.om - jump;
    s.Dprev = s.D;
    if (std::abs(s.D) < std::abs(s.minD)) {  // the reference records the updated om
      s.minom = s.om;
      s.minD = s.D;
    }
    if (go)
      start_iter(s);
    else
      s.done = 1;
  };
  for (;;) {
    who.clear();
    for (int r = 0; r < nroots; ++r) {
      SecState &s = S[r];
      if (s.done) continue;
      req[who.size()] = (s.phase == 0 || s.phase == 3) ? s.prevom : s.om;
      who.push_back(r);
    }
    if (who.empty()) break;
    int rc = run_disp(h, (int)who.size(), reinterpret_cast<const double *>(req.data()),
                      reinterpret_cast<double *>(res.data()), nullptr, nullptr, nullptr);
    if (rc) return rc;
    const double now_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (size_t t = 0; t < who.size(); ++t) {
      SecState &s = S[who[t]];
      const cplx D = res[t];
      h->root_ms[who[t]] = now_ms;
      switch (s.phase) {
        case 0:
          s.Dprev = D;
          s.phase = 1;
          break;
        case 1:
          s.minD = D;
          s.minom = s.om;
          if (std::abs(s.Dprev) < std::abs(s.minD)) {
            s.minom = s.prevom;
            s.minD = s.Dprev;
          }
          start_iter(s);
          break;
        case 2:
          s.D = D;
          if (std::abs(s.D - s.Dprev) < 1.e-80) {
            s.prevom = s.prevom + 1.e-8;
            s.phase = 3;
            break;
          }
          step(s);
          break;
        default:
          s.Dprev = D;
          step(s);
          break;
      }
    }
  }
  for (int r = 0; r < nroots; ++r) {
    if (S[r].iter >= rp->numiter) S[r].om = S[r].minom;
    om[r] = S[r].om;
    if (iters) iters[r] = S[r].iter;
  }
  return ALPSG_OK;
}

// rtsec (ALPS_fns.f90:2105-2195) for many independent roots.  Phases: 0 =
// need D(x1), 1 = need D(x2), 2 = need D(rts) in an iteration.
struct RtsecState {
  cplx x1, x2, xl, fl, f, rts, dx;
  int j, phase, done;
};

int rtsec_batch(alpsg_ctx *h, const alpsg_root_params *rp, int nroots, cplx *om, int *iters) {
  std::vector<RtsecState> S(nroots);
  std::vector<cplx> req(nroots), res(nroots);
  std::vector<int> who;
  const auto t0 = std::chrono::steady_clock::now();
  h->root_ms.assign(nroots, 0.0);
  for (int r = 0; r < nroots; ++r) {
    RtsecState &s = S[r];
    s.x1 = om[r] * 1.0;
    s.x2 = om[r] * (1.0 + rp->D_prec);
    s.j = 0;
    s.phase = 0;
    s.done = 0;
  }
  auto next = [&](RtsecState &s) {
    // one pass of the do-loop body up to the disp call
    if (s.j + 1 > rp->numiter - 1) {
      s.done = 1;
      return;
    }
    s.j = s.j + 1;
    if (std::abs(s.f - s.fl) > 1.e-40)
      s.dx = (s.xl - s.rts) * s.f / (s.f - s.fl);
    else
      s.dx = (s.x2 - s.x1) / 25.0;
    s.xl = s.rts;
    s.fl = s.f;
    s.rts = s.rts + s.dx / 2.0;
    s.phase = 2;
  };
  for (;;) {
    who.clear();
    for (int r = 0; r < nroots; ++r) {
      RtsecState &s = S[r];
      if (s.done) continue;
      req[who.size()] = s.phase == 0 ? s.x1 : (s.phase == 1 ? s.x2 : s.rts);
      who.push_back(r);
    }
    if (who.empty()) break;
    int rc = run_disp(h, (int)who.size(), reinterpret_cast<const double *>(req.data()),
                      reinterpret_cast<double *>(res.data()), nullptr, nullptr, nullptr);
    if (rc) return rc;
    const double now_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    for (size_t t = 0; t < who.size(); ++t) {
      RtsecState &s = S[who[t]];
      const cplx D = res[t];
      h->root_ms[who[t]] = now_ms;
      switch (s.phase) {
        case 0:
          s.fl = D;
          s.phase = 1;
          break;
        case 1:
          s.f = D;
          if (std::abs(s.fl) < std::abs(s.f)) {
            s.rts = s.x1;
            s.xl = s.x2;
            std::swap(s.fl, s.f);
          } else {
            s.xl = s.x1;
            s.rts = s.x2;
          }
          next(s);
          break;
        default:
          s.f = D;
          if (std::abs(s.dx) < rp->D_tol || std::abs(s.f) == 0.0)
            s.done = 1;
          else
            next(s);
          break;
      }
    }
  }
  for (int r = 0; r < nroots; ++r) {
    om[r] = S[r].rts;
    if (iters) iters[r] = S[r].j;
  }
  return ALPSG_OK;
}

// The root solver the &system namelist selects (secant_method, as in
// om_scan, ALPS_fns.f90:2507-2514, and refine_guess, ALPS_fns.f90:3830-3843),
// then the "final instance of disp" at each root when Dfin is given.
// work (may be NULL): host work the device-stepped secant_osc runs while its
// first rounds run; left set for the caller otherwise
int root_batch(alpsg_ctx *h, const alpsg_root_params *rp, int nroots, cplx *om, cplx *Dfin,
               int *iters, std::function<int()> *work = nullptr) {
  int rc;
  switch (rp->secant_method) {
    case 0:
      rc = secant_batch(h, rp, nroots, om, iters);
      break;
    case 1:
      rc = rtsec_batch(h, rp, nroots, om, iters);
      break;
    case 2:
      return secant_osc_batch(h, rp, nroots, om, Dfin, iters, work);
    default:
      return fail(ALPSG_E_ARG, "secant_method must be 0, 1 or 2");
  }
  if (rc) return rc;
  if (Dfin)
    return run_disp(h, nroots, reinterpret_cast<const double *>(om),
                    reinterpret_cast<double *>(Dfin), nullptr, nullptr, nullptr);
  return ALPSG_OK;
}

// calc_eigen (ALPS_fns.f90:2605-2899) for one root from the chi0, chi0_low and
// wave that disp returned at om (c0, w0), real(om) (c1) and real(om*1.000001)
// (c2, l2); chi arrays in the Fortran layouts of alpsg_disp_batch.
void calc_eigen_one(const alpsg_ctx *h, cplx omega, const cplx *c0, const cplx *w0, const cplx *c1,
                    const cplx *c2, const cplx *l2, bool eigen, bool heat, cplx *electric,
                    cplx *magnetic, cplx *vmean, cplx *ds, double *Ps, double *Ps_split,
                    double *ewave_out) {
  const int nspec = h->nspec;
  const double kperp = h->kperp, kpar = h->kpar, vA = h->vA;
  const cplx ii(0.0, 1.0);
  auto W = [&](int i, int j) { return w0[i + 3 * j]; };  // wave(i+1, j+1)
  auto CHI = [&](const cplx *c, int s, int i, int j) { return c[s + nspec * (i + 3 * j)]; };
  auto LOW = [&](int s, int i, int j, int q) { return l2[s + nspec * (i + 3 * (j + 3 * (q + 1)))]; };
  electric[0] = 1.0;
  electric[2] = -electric[0] * (W(1, 0) * W(2, 1) - W(2, 0) * W(1, 1));
  electric[2] = electric[2] / (W(1, 2) * W(2, 1) - W(2, 2) * W(1, 1));
  if (std::abs(W(2, 1)) != 0.0) {
    electric[1] = -electric[2] * W(2, 2) - electric[0] * W(2, 0);
    electric[1] = electric[1] / W(2, 1);
  } else {
    electric[1] = W(1, 0) * W(0, 2) - W(0, 0) * W(1, 2);
    electric[1] = electric[1] / (W(1, 2) * W(0, 1) - W(1, 1) * W(0, 2));
  }
  magnetic[0] = -1.0 * kpar * electric[1] / (omega * vA);
  magnetic[1] = -1.0 * (kperp * electric[2] - kpar * electric[0]) / (omega * vA);
  magnetic[2] = kperp * electric[1] / (omega * vA);
  if (eigen) {
    for (int s = 0; s < nspec; ++s) {
      const double pf = h->current_int[s] / (h->ns[s] * h->qs[s]);
      const double pre = (vA * vA) / (h->qs[s] * h->ns[s]);
      auto sumec = [&](int j) {
        return electric[0] * CHI(c0, s, j, 0) + electric[1] * CHI(c0, s, j, 1) +
               electric[2] * CHI(c0, s, j, 2);
      };
      cplx *v = vmean + 3 * s;
      if (pf == 0.0) {
        for (int j = 0; j < 3; ++j) v[j] = -(pre * ii * omega * sumec(j));
      } else {
        for (int j = 0; j < 2; ++j) v[j] = -(pre * ii * omega * sumec(j));
        v[2] = -(pre * ii * omega * sumec(2)) - pf * kperp * v[0] / (omega - kpar * pf);
        v[2] = v[2] / (1.0 + (kpar * pf) / (omega - kpar * pf));
      }
      ds[s] = (1.0 / vA) * (v[0] * kperp + v[2] * kpar) / (omega - kpar * pf);
    }
  }
  if (!heat) return;
  cplx chihold[3][3], chih[3][3], dchih[3][3], term1[3];
  for (int s = 0; s < nspec; ++s) Ps[s] = 0.0;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      cplx a = 0.0, b = 0.0, a2 = 0.0, b2 = 0.0;
      for (int s = 0; s < nspec; ++s) {
        a = a + CHI(c1, s, i, j);
        b = b + std::conj(CHI(c1, s, j, i));
        a2 = a2 + CHI(c2, s, i, j);
        b2 = b2 + std::conj(CHI(c2, s, j, i));
      }
      chihold[i][j] = 0.5 * (a + b);
      chih[i][j] = 0.5 * (a2 + b2);
      dchih[i][j] = (1.000001 * chih[i][j] - chihold[i][j]) / 0.000001;
    }
  for (int s = 0; s < nspec; ++s) {
    cplx t[3];
    for (int i = 0; i < 3; ++i) {
      t[i] = 0.0;
      for (int k = 0; k < 3; ++k) {
        const cplx chia = -0.5 * ii * (CHI(c1, s, k, i) - std::conj(CHI(c1, s, i, k)));
        t[i] = t[i] + std::conj(electric[k]) * chia;
      }
    }
    Ps[s] = (t[0] * electric[0] + t[1] * electric[1] + t[2] * electric[2]).real();
  }
  for (int i = 0; i < 3; ++i)
    term1[i] = std::conj(electric[0]) * dchih[0][i] + std::conj(electric[1]) * dchih[1][i] +
               std::conj(electric[2]) * dchih[2][i];
  const double ewave =
      ((term1[0] * electric[0] + term1[1] * electric[1] + term1[2] * electric[2]) +
       (magnetic[0] * std::conj(magnetic[0]) + magnetic[1] * std::conj(magnetic[1]) +
        magnetic[2] * std::conj(magnetic[2])))
          .real();
  for (int s = 0; s < nspec; ++s) Ps[s] = Ps[s] / ewave;
  const cplx hi = -0.5 * ii;
  const cplx exy[3] = {electric[0], electric[1], 0.0};
  for (int s = 0; s < nspec; ++s) {
    double *o = Ps_split + 4 * s;
    double v = (hi * std::conj(electric[1]) * electric[1] * (LOW(s, 1, 1, 0) - std::conj(LOW(s, 1, 1, 0)))).real();
    v = (v + hi * (electric[2] * std::conj(electric[1]) * LOW(s, 1, 2, 0) -
                   std::conj(electric[2]) * electric[1] * std::conj(LOW(s, 1, 2, 0)))).real();
    o[0] = v;
    v = (hi * (electric[1] * std::conj(electric[2]) * LOW(s, 2, 1, 0) -
               std::conj(electric[1]) * electric[2] * std::conj(LOW(s, 2, 1, 0)))).real();
    v = (v + hi * std::conj(electric[2]) * electric[2] * (LOW(s, 2, 2, 0) - std::conj(LOW(s, 2, 2, 0)))).real();
    o[1] = v;
    for (int q = 1; q >= -1; q -= 2) {
      cplx t[3];
      for (int i = 0; i < 3; ++i) {
        t[i] = 0.0;
        for (int k = 0; k < 3; ++k)
          t[i] = t[i] + std::conj(exy[k]) * (hi * (LOW(s, k, i, q) - std::conj(LOW(s, i, k, q))));
      }
      o[q == 1 ? 2 : 3] = (t[0] * exy[0] + t[1] * exy[1] + t[2] * exy[2]).real();
    }
    for (int k = 0; k < 4; ++k) o[k] = o[k] / ewave;
  }
  *ewave_out = ewave;
}

int calc_eigen_batch(alpsg_ctx *h, int n, const cplx *om, bool eigen, bool heat, cplx *ef, cplx *bf,
                     cplx *Us, cplx *ds, double *Ps, double *Ps_split, double *W_EM) {
  if (eigen && !h->have_current)
    return fail(ALPSG_E_STATE, "current_int unknown: pass f0 or current_int in the plasma description");
  const int ns = h->nspec;
  const int m = heat ? 3 * n : n;
  std::vector<cplx> w(m), det(m), chi((size_t)m * ns * 9), low((size_t)m * ns * 27), wave((size_t)m * 9);
  for (int r = 0; r < n; ++r) {
    w[r] = om[r];
    if (heat) {
      w[n + r] = cplx(om[r].real(), 0.0);
      w[2 * n + r] = cplx((om[r] * 1.000001).real(), 0.0);
    }
  }
  int rc = run_disp(h, m, reinterpret_cast<const double *>(w.data()), reinterpret_cast<double *>(det.data()),
                    reinterpret_cast<double *>(chi.data()), reinterpret_cast<double *>(low.data()),
                    reinterpret_cast<double *>(wave.data()));
  if (rc) return rc;
  std::vector<cplx> dummy(3 * ns + 8);
  std::vector<double> dd(4 * ns + 4);
  for (int r = 0; r < n; ++r) {
    const cplx *c0 = &chi[(size_t)r * ns * 9], *w0 = &wave[(size_t)r * 9];
    const cplx *c1 = heat ? &chi[(size_t)(n + r) * ns * 9] : nullptr;
    const cplx *c2 = heat ? &chi[(size_t)(2 * n + r) * ns * 9] : nullptr;
    const cplx *l2 = heat ? &low[(size_t)(2 * n + r) * ns * 27] : nullptr;
    cplx e3[3], b3[3];
    double ew = 0.0;
    calc_eigen_one(h, om[r], c0, w0, c1, c2, l2, eigen, heat, e3, b3,
                   eigen && Us ? Us + (size_t)r * 3 * ns : dummy.data(),
                   eigen && ds ? ds + (size_t)r * ns : dummy.data() + 3 * ns,
                   heat && Ps ? Ps + (size_t)r * ns : dd.data(),
                   heat && Ps_split ? Ps_split + (size_t)r * 4 * ns : dd.data(), &ew);
    for (int k = 0; k < 3; ++k) {
      if (ef) ef[(size_t)r * 3 + k] = e3[k];
      if (bf) bf[(size_t)r * 3 + k] = b3[k];
    }
    if (heat && W_EM) W_EM[r] = ew;
  }
  return ALPSG_OK;
}

}  // namespace

int rel_host_tables(const alpsg_plasma_desc *d, int s, int r, int nspec_rel, const double *pe,
                    const double *pa, RelHost &o) {
  const int ngam = d->ngamma, npb = d->npparbar, nperp = d->nperp, npar = d->npar;
  // Fortran layout (1:nspec_rel, 0:ngamma, 0:npparbar[, 1:2])
  auto frel = [&](const double *a, int ig, int ip, int c) {
    return a[(size_t)r + (size_t)nspec_rel * (ig + (size_t)(ngam + 1) * (ip + (size_t)(npb + 1) * c))];
  };
  const size_t np = (size_t)(ngam + 1) * (npb + 1);
  o.G.resize(np);
  o.PB.resize(np);
  o.F.resize(np);
  o.DF.resize(np);
  for (int ig = 0; ig <= ngam; ++ig)
    for (int ip = 0; ip <= npb; ++ip) {
      const size_t k = (size_t)ig * (npb + 1) + ip;
      o.G[k] = frel(d->gamma_rel, ig, ip, 0);
      o.PB[k] = frel(d->pparbar_rel, ig, ip, 0);
      o.F[k] = frel(d->f0_rel, ig, ip, 0);
      o.DF[k] = make_double2(frel(d->df0_rel, ig, ip, 0), frel(d->df0_rel, ig, ip, 1));
    }
  for (int ip = 0; ip < npb; ++ip)
    if (!(o.PB[(size_t)2 * (npb + 1) + ip + 1] > o.PB[(size_t)2 * (npb + 1) + ip]))
      return fail(ALPSG_E_ARG, "pparbar_rel rows must increase");
  o.lim.resize(ngam + 1);
  for (int ig = 0; ig <= ngam; ++ig) {
    int lo = 1, up = npb - 1;
    bool fl = false, fu = false;
    const double *f = &o.F[(size_t)ig * (npb + 1)];
    for (int ip = 1; ip <= npb - 1; ++ip) {
      if (!fl && f[ip - 1] <= -1.0 && f[ip] > -1.0) {
        lo = ip;
        fl = true;
      }
      if (!fu && f[ip] > -1.0 && f[ip + 1] <= -1.0) {
        up = ip;
        fu = true;
      }
    }
    o.lim[ig] = make_int2(lo, up);
  }
  const double vA = d->vA, m = d->ms[s], q = d->qs[s];
  o.dg = o.G[(size_t)2 * (npb + 1) + 2] - o.G[(size_t)1 * (npb + 1) + 2];
  o.dpb = o.PB[(size_t)2 * (npb + 1) + 2] - o.PB[(size_t)2 * (npb + 1) + 1];
  double ee = 0.0;
  {
    auto term = [&](int ig, int ip) {
      const size_t k = (size_t)ig * (npb + 1) + ip;
      return o.PB[k] * o.DF[k].y;
    };
    int ig = ngam - 1;
    ee = ee + term(ig, o.lim[ig].x);
    ee = ee + term(ig, o.lim[ig].y);
    for (int ip = o.lim[ig].x + 1; ip <= o.lim[ig].y - 1; ++ip) ee = ee + 2.0 * term(ig, ip);
    for (ig = 1; ig <= ngam - 2; ++ig) {
      for (int ip = o.lim[ig].x + 1; ip <= o.lim[ig].y - 1; ++ip) ee = ee + 4.0 * term(ig, ip);
      ee = ee + 2.0 * term(ig, o.lim[ig].x);
      ee = ee + 2.0 * term(ig, o.lim[ig].y);
    }
    const double PI = 3.141592653589793238462643383279502884;
    ee = ee * 2.0 * PI * q / m;
    ee = ee * o.dg * o.dpb * 0.25 * ((m / vA) * (m / vA) * (m / vA));
  }
  o.ee = ee;
  o.gpp.resize((size_t)(nperp + 1) * (npar + 1));
  for (int i = 0; i <= nperp; ++i)
    for (int j = 0; j <= npar; ++j) {
      const double a = pe[i], b = pa[j];
      o.gpp[(size_t)i * (npar + 1) + j] = std::sqrt((a * a + b * b) * (vA * vA) / (m * m) + 1.0);
    }
  o.gmin = *std::min_element(o.gpp.begin(), o.gpp.end());
  o.gmax = *std::max_element(o.gpp.begin(), o.gpp.end());
  o.gmono = true;
  for (int i = 1; i <= nperp && o.gmono; ++i)
    for (int j = 0; j <= npar; ++j)
      if (o.gpp[(size_t)i * (npar + 1) + j] < o.gpp[(size_t)(i - 1) * (npar + 1) + j]) o.gmono = false;
  return ALPSG_OK;
}

void rel_fit_rows(const alpsg_plasma_desc *d, int s, int r, int nspec_rel, double *fitc) {
  // Juettner fits in Gamma (ALPS_analyt.f90:238-252) on the igamma rows, in
  // the Maxwell form c0*exp(-c1*(p-c2)^2): type 4 has c1 = c2 = 0
  const int nspec = d->nspec, ngam = d->ngamma, npb = d->npparbar, npf = d->param_fit_nperp;
  for (int ig = 0; ig <= ngam; ++ig) {
    const double gv =
        d->gamma_rel[(size_t)r + (size_t)nspec_rel * (ig + (size_t)(ngam + 1) * 1)];
    (void)npb;
    for (int f = 0; f < d->n_fits[s]; ++f) {
      auto prm = [&](int k) {
        return d->param_fit[(size_t)s + (size_t)nspec * (ig + (size_t)(npf + 1) * (k + 5 * (size_t)f))];
      };
      const double pc = d->perp_correction[s + (size_t)nspec * f];
      const int ft = d->fit_type[s + (size_t)nspec * f];
      double *c = &fitc[((size_t)ig * d->n_fits[s] + f) * FIT_STRIDE];
      c[0] = prm(0) * std::exp(-pc * gv);
      c[1] = ft == 5 ? prm(1) : 0.0;
      c[2] = ft == 5 ? prm(2) : 0.0;
    }
  }
}

void rel_gamma_table(std::vector<double> &gt) {
  // Gamma(m) by the Lanczos series of ALPS_fns_rel.f90:1738-1791 and Fact(k)
  // (:1557-1574), as CBESSJ divides by them
  gt.assign(GTAB_N + 21 + (size_t)RFG_NU * 21, 0.0);
  const double cof[6] = {76.18009173, -86.50532033, 24.01409822, -1.231739516, 0.120858003e-2,
                         -0.536382e-5};
  for (int mm = 1; mm < GTAB_N; ++mm) {
    double x = (1.0 * mm) - 1.0, tmp = x + 5.5, ser = 1.0;
    tmp = (x + 0.5) * std::log(tmp) - tmp;
    for (int j = 0; j < 6; ++j) {
      x = x + 1.0;
      ser = ser + cof[j] / x;
    }
    gt[mm] = std::exp(tmp + std::log(2.50662827465 * ser));
  }
  for (int k = 0; k <= 20; ++k) {
    double f = 1.0;
    for (int i = 2; i <= k; ++i) f = f * (1.0 * i);
    gt[GTAB_N + k] = f;
  }
  // the coefficients of the CBESSJ series, (-z^2/4)^k / (k! Gamma(nu+k+1)),
  // as one reciprocal per (nu, k): an overflowing Gamma gives 0, as the
  // source's division by it does
  for (int nu = 0; nu < RFG_NU; ++nu)
    for (int k = 0; k <= 20; ++k)
      gt[GTAB_N + 21 + (size_t)nu * 21 + k] = 1.0 / (gt[GTAB_N + k] * gt[nu + k + 1]);
}

extern "C" {

const char *alpsg_last_error(void) { return g_err.c_str(); }

int alpsg_create(const alpsg_plasma_desc *d, alpsg_handle *out) {
  ALPSG_TRY
  if (!d || !out) return fail(ALPSG_E_ARG, "null argument");
  *out = nullptr;
  const int nspec = d->nspec, nperp = d->nperp, npar = d->npar;
  if (nspec < 1 || nspec > ALPSG_MAXSPEC) return fail(ALPSG_E_ARG, "nspec out of range 1..8");
  if (!d->pp || (!d->f0 && !d->df0) || !d->ns || !d->qs || !d->ms || !d->ACmethod || !d->n_fits)
    return fail(ALPSG_E_ARG, "missing plasma array");
  const int P = d->positions_principal;
  if (P < 1 || P > 16) return fail(ALPSG_E_UNSUPPORTED, "positions_principal must be in 1..16");
  if (nperp < 4 || npar < 2 * P + 8) return fail(ALPSG_E_ARG, "grid too small for positions_principal");
  if (d->n_resonance_interval < 2) return fail(ALPSG_E_ARG, "n_resonance_interval < 2");
  int nrel = 0;
  for (int s = 0; s < nspec; ++s) {
    if (d->qs[s] == 0.0 || d->ms[s] <= 0.0) return fail(ALPSG_E_ARG, "species charge/mass");
    if (d->usebM && d->usebM[s]) {
      // closed-form species: no table, fits or continuation are read
      if (!d->bM_nmax || !d->bM_Bessel_zero || !d->bM_beta || !d->bM_alpha || !d->bM_pdrift)
        return fail(ALPSG_E_ARG, "missing bM_spec arrays for a use_bM species");
      if (d->bM_beta[s] < 0.0 || (d->bM_beta[s] > 0.0 && !(d->bM_alpha[s] > 0.0)) || d->ns[s] <= 0.0)
        return fail(ALPSG_E_ARG, "bM_spec parameters");
      if (d->bM_beta[s] == 0.0 && !d->kperp_norm)
        return fail(ALPSG_E_UNSUPPORTED, "cold-plasma species need kperp_norm (the reference stops)");
      continue;
    }
    if (d->relativistic && d->relativistic[s]) {
      // Gamma-grid species: Juettner fits in Gamma (types 4, 5) indexed by igamma
      ++nrel;
      if (!d->gamma_rel || !d->pparbar_rel || !d->f0_rel || !d->df0_rel)
        return fail(ALPSG_E_ARG, "missing relativistic tables");
      if (d->ngamma < 4 || d->npparbar < 2 * P + 8)
        return fail(ALPSG_E_ARG, "relativistic grid too small (ngamma >= 4, npparbar >= 2P+8)");
      if (d->ACmethod[s] != 1)
        return fail(ALPSG_E_UNSUPPORTED, "relativistic species need fit continuation (ACmethod 1)");
      if (d->n_fits[s] < 1 || d->n_fits[s] > ALPSG_MAXFITS || d->n_fits[s] > d->max_fits)
        return fail(ALPSG_E_ARG, "n_fits out of range");
      if (!d->fit_type || !d->perp_correction || !d->param_fit) return fail(ALPSG_E_ARG, "missing fit arrays");
      for (int f = 0; f < d->n_fits[s]; ++f) {
        const int ft = d->fit_type[s + (size_t)nspec * f];
        if (ft != 4 && ft != 5)
          return fail(ALPSG_E_UNSUPPORTED, "relativistic species need fit_type 4 or 5");
      }
      if (d->param_fit_nperp < std::max(nperp, d->ngamma)) return fail(ALPSG_E_ARG, "param_fit_nperp < max(nperp, ngamma)");
      continue;
    }
    if (d->ACmethod[s] < 0 || d->ACmethod[s] > 2) return fail(ALPSG_E_ARG, "ACmethod must be 0, 1 or 2");
    if (d->ACmethod[s] == 0) {
      // the pre-coded f0 of distribution_analyt (ALPS_analyt.f90:77-79)
      if (!distribution_analyt_defined(s + 1))
        return fail(ALPSG_E_UNSUPPORTED, "ACmethod 0: distribution_analyt defines no f0 for this species");
    } else if (d->ACmethod[s] == 1) {
      if (d->n_fits[s] < 1 || d->n_fits[s] > ALPSG_MAXFITS || d->n_fits[s] > d->max_fits)
        return fail(ALPSG_E_ARG, "n_fits out of range");
      if (!d->fit_type || !d->perp_correction || !d->param_fit) return fail(ALPSG_E_ARG, "missing fit arrays");
      for (int f = 0; f < d->n_fits[s]; ++f) {
        int ft = d->fit_type[s + (size_t)nspec * f];
        if (ft != 1 && ft != 2 && ft != 3 && ft != 6)
          return fail(ALPSG_E_UNSUPPORTED, "fit_type must be 1, 2, 3 or 6 (4 and 5 are relativistic)");
      }
      if (d->param_fit_nperp < nperp) return fail(ALPSG_E_ARG, "param_fit_nperp < nperp");
    } else {
      if (!d->poly_kind || !d->poly_order || !d->logfit || !d->poly_log_max || !d->poly_fit_coeffs)
        return fail(ALPSG_E_ARG, "missing Chebyshev arrays");
      if (d->poly_kind[s] != 1) return fail(ALPSG_E_UNSUPPORTED, "poly_kind must be 1");
      if (d->poly_order[s] < 1 || d->poly_order[s] > d->poly_max_order) return fail(ALPSG_E_ARG, "poly_order");
    }
  }
  std::unique_ptr<alpsg_ctx> hold(new alpsg_ctx());
  alpsg_ctx *h = hold.get();
  auto bail = [&](int code) { return code; };
  h->nspec = nspec;
  h->nperp = nperp;
  h->npar = npar;
  h->kperp_norm = d->kperp_norm ? 1 : 0;
  h->P = P;
  h->NRI = d->n_resonance_interval;
  h->vA = d->vA;
  h->Tlim = d->Tlim;
  h->Bessel_zero = d->Bessel_zero;
  h->qs.assign(d->qs, d->qs + nspec);
  h->ms.assign(d->ms, d->ms + nspec);
  h->ns.assign(d->ns, d->ns + nspec);
  // separable, uniform, increasing grid: pperp(iperp) from pp(s,iperp,1,1), ppar(ipar) from pp(s,2,ipar,2)
  h->pperp_h.resize((size_t)nspec * (nperp + 1));
  h->ppar_h.resize((size_t)nspec * (npar + 1));
  for (int s = 0; s < nspec; ++s) {
    for (int i = 0; i <= nperp; ++i) h->pperp_h[(size_t)s * (nperp + 1) + i] = d->pp[f_pp(nspec, nperp, npar, s, i, 1, 0)];
    for (int j = 0; j <= npar; ++j) h->ppar_h[(size_t)s * (npar + 1) + j] = d->pp[f_pp(nspec, nperp, npar, s, 2, j, 1)];
    if (d->usebM && d->usebM[s]) continue;  // the closed form reads no grid
    for (int i = 0; i <= nperp; ++i)
      for (int j = 0; j <= npar; ++j) {
        if (d->pp[f_pp(nspec, nperp, npar, s, i, j, 0)] != h->pperp_h[(size_t)s * (nperp + 1) + i] ||
            d->pp[f_pp(nspec, nperp, npar, s, i, j, 1)] != h->ppar_h[(size_t)s * (npar + 1) + j])
          return bail(fail(ALPSG_E_UNSUPPORTED, "momentum grid is not a tensor product of pperp and ppar"));
      }
    const double *pa = &h->ppar_h[(size_t)s * (npar + 1)];
    const double dpa = pa[2] - pa[1];
    if (!(dpa > 0)) return bail(fail(ALPSG_E_ARG, "ppar grid must increase"));
    for (int j = 0; j < npar; ++j)
      if (!(pa[j + 1] > pa[j]) || std::fabs((pa[j + 1] - pa[j]) - dpa) > 1e-6 * dpa)
        return bail(fail(ALPSG_E_UNSUPPORTED, "ppar grid is not uniform"));
  }
  HIPCHK(hipGetDevice(&h->device));
  HIPCHK(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking));
  HIPCHK(hipEventCreate(&h->ev0));
  HIPCHK(hipEventCreate(&h->ev1));
  HIPCHK(hipEventCreate(&h->ev_sweep));
  HIPCHK(hipEventCreate(&h->ev_res));
  HIPCHK(hipEventCreateWithFlags(&h->ev_done, hipEventDisableTiming));
  HIPCHK(hipStreamCreateWithFlags(&h->st_copy, hipStreamNonBlocking));
  HIPCHK(hipEventCreateWithFlags(&h->ev_map, hipEventDisableTiming));
  HIPCHK(hipEventCreateWithFlags(&h->ev_copy, hipEventDisableTiming));
  const size_t ngrid = (size_t)nspec * (nperp + 1) * (npar + 1);
  HIPCHK(h->pperp.ensure(h->pperp_h.size()));
  HIPCHK(h->ppar.ensure(h->ppar_h.size()));
  HIPCHK(h->df1.ensure(ngrid));
  HIPCHK(h->df2.ensure(ngrid));
  HIPCHK(h->qs_d.ensure(nspec));
  HIPCHK(h->ms_d.ensure(nspec));
  HIPCHK(h->int_ee.ensure(nspec));
  HIPCHK(hipMemcpyAsync(h->pperp.p, h->pperp_h.data(), sizeof(double) * h->pperp_h.size(), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->ppar.p, h->ppar_h.data(), sizeof(double) * h->ppar_h.size(), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->qs_d.p, d->qs, sizeof(double) * nspec, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->ms_d.p, d->ms, sizeof(double) * nspec, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemsetAsync(h->df1.p, 0, sizeof(double) * ngrid, h->st));
  HIPCHK(hipMemsetAsync(h->df2.p, 0, sizeof(double) * ngrid, h->st));
  if (d->f0) {
    // f0 to the [s][iperp][ipar] layout, derivatives on the device
    std::vector<double> f0c(ngrid);
    for (int s = 0; s < nspec; ++s)
      for (int i = 0; i <= nperp; ++i)
        for (int j = 0; j <= npar; ++j)
          f0c[((size_t)s * (nperp + 1) + i) * (npar + 1) + j] =
              d->f0[(size_t)s + (size_t)nspec * (i + (size_t)(nperp + 1) * j)];
    HIPCHK(h->f0.ensure(ngrid));
    HIPCHK(hipMemcpyAsync(h->f0.p, f0c.data(), sizeof(double) * ngrid, hipMemcpyHostToDevice, h->st));
    HIPCHK(launch_derivative(h->st, nspec, nperp, npar, h->f0.p, h->pperp.p, h->ppar.p, h->df1.p, h->df2.p));
    HIPCHK(hipStreamSynchronize(h->st));
  } else {
    std::vector<double> a(ngrid, 0.0), b(ngrid, 0.0);
    for (int s = 0; s < nspec; ++s)
      for (int i = 1; i <= nperp - 1; ++i)
        for (int j = 1; j <= npar - 1; ++j) {
          size_t src = (size_t)s + (size_t)nspec * ((i - 1) + (size_t)(nperp - 1) * ((j - 1)));
          size_t str = (size_t)nspec * (nperp - 1) * (npar - 1);
          a[((size_t)s * (nperp + 1) + i) * (npar + 1) + j] = d->df0[src];
          b[((size_t)s * (nperp + 1) + i) * (npar + 1) + j] = d->df0[src + str];
        }
    HIPCHK(hipMemcpyAsync(h->df1.p, a.data(), sizeof(double) * ngrid, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipMemcpyAsync(h->df2.p, b.data(), sizeof(double) * ngrid, hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
  }
  HIPCHK(launch_int_ee(h->st, nspec, nperp, npar, h->pperp.p, h->ppar.p, h->df1.p, h->df2.p, h->qs_d.p, h->ms_d.p, h->int_ee.p));
  h->int_ee_h.resize(nspec);
  HIPCHK(hipMemcpyAsync(h->int_ee_h.data(), h->int_ee.p, sizeof(double) * nspec, hipMemcpyDeviceToHost, h->st));

  // analytic-continuation tables
  std::vector<size_t> fit_off(nspec + 1, 0);
  auto bm = [&](int s) { return d->usebM != nullptr && d->usebM[s] != 0; };
  auto relf = [&](int s) { return !bm(s) && d->relativistic != nullptr && d->relativistic[s] != 0; };
  // index of each species among the relativistic ones (determine_sproc_rel, ALPS_fns_rel.f90:437-455)
  int rel_index[ALPSG_MAXSPEC];
  int nspec_rel = 0;
  for (int s = 0; s < nspec; ++s) rel_index[s] = (d->relativistic && d->relativistic[s]) ? nspec_rel++ : -1;
  const int ngam = d->ngamma, npb = d->npparbar;
  // Fortran layout (1:nspec_rel, 0:ngamma, 0:npparbar[, 1:2])
  auto frel = [&](const double *a, int r, int ig, int ip, int c) {
    return a[(size_t)r + (size_t)nspec_rel * (ig + (size_t)(ngam + 1) * (ip + (size_t)(npb + 1) * c))];
  };
  for (int s = 0; s < nspec; ++s) {
    const int rows = relf(s) ? ngam + 1 : nperp + 1;
    fit_off[s + 1] = fit_off[s] + ((d->ACmethod[s] == 1 && !bm(s)) ? (size_t)rows * d->n_fits[s] * FIT_STRIDE : 0);
  }
  std::vector<double> fitc(std::max<size_t>(fit_off[nspec], 1), 0.0);
  int maxorder = 0;
  for (int s = 0; s < nspec; ++s)
    if (d->ACmethod[s] == 2 && !bm(s)) maxorder = std::max(maxorder, d->poly_order[s]);
  std::vector<double> polyc((size_t)nspec * (nperp + 1) * (maxorder + 1), 0.0);
  const int npf = d->param_fit_nperp;
  for (int s = 0; s < nspec; ++s) {
    if (bm(s)) continue;
    if (relf(s)) {
      rel_fit_rows(d, s, rel_index[s], nspec_rel, &fitc[fit_off[s]]);
      continue;
    }
    if (d->ACmethod[s] == 1) {
      for (int i = 0; i <= nperp; ++i) {
        const double pv = h->pperp_h[(size_t)s * (nperp + 1) + i];
        for (int f = 0; f < d->n_fits[s]; ++f) {
          auto prm = [&](int k) {
            return d->param_fit[(size_t)s + (size_t)nspec * (i + (size_t)(npf + 1) * (k + 5 * (size_t)f))];
          };
          const double pc = d->perp_correction[s + (size_t)nspec * f];
          const int ft = d->fit_type[s + (size_t)nspec * f];
          double *c = &fitc[fit_off[s] + ((size_t)i * d->n_fits[s] + f) * FIT_STRIDE];
          c[1] = prm(1);
          c[2] = prm(2);
          if (ft == 1) {
            c[0] = prm(0) * std::exp(-pc * (pv * pv));
          } else if (ft == 2) {
            c[0] = prm(0);
            c[3] = prm(3);
            c[4] = pc * prm(4) * (pv * pv);
          } else if (ft == 3) {
            // Juettner in (pperp, ppar), ALPS_analyt.f90:238-242
            c[0] = prm(0);
            c[3] = pv * pv;
            c[4] = d->vA * d->vA / (d->ms[s] * d->ms[s]);
          } else {
            c[0] = prm(0);
            c[3] = prm(3) * pc * (pv * pv);
          }
        }
      }
    } else {
      for (int i = 0; i <= nperp; ++i)
        for (int n = 0; n <= d->poly_order[s]; ++n)
          polyc[((size_t)s * (nperp + 1) + i) * (maxorder + 1) + n] =
              d->poly_fit_coeffs[(size_t)s + (size_t)nspec * (i + (size_t)(nperp + 1) * n)];
    }
  }
  HIPCHK(h->fitc.ensure(fitc.size()));
  HIPCHK(h->polyc.ensure(polyc.size()));
  HIPCHK(hipMemcpyAsync(h->fitc.p, fitc.data(), sizeof(double) * fitc.size(), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemcpyAsync(h->polyc.p, polyc.data(), sizeof(double) * polyc.size(), hipMemcpyHostToDevice, h->st));
  // Miller start indices m(n) = 2*((n+int(sqrt(float(40n))))/2) (ALPS_fns_rel.f90:1622)
  std::vector<int> mt(MTAB_N, 0);
  for (int n = 1; n < MTAB_N; ++n) mt[n] = 2 * ((n + (int)std::sqrt((float)(40 * n))) / 2);
  HIPCHK(h->mtab.ensure(MTAB_N));
  HIPCHK(hipMemcpyAsync(h->mtab.p, mt.data(), sizeof(int) * MTAB_N, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  if (nrel > 0) {
    std::vector<double> gt;
    rel_gamma_table(gt);
    HIPCHK(h->gtab.ensure(gt.size()));
    HIPCHK(hipMemcpy(h->gtab.p, gt.data(), sizeof(double) * gt.size(), hipMemcpyHostToDevice));
    HIPCHK(h->err_d.ensure(1));
    HIPCHK(hipMemset(h->err_d.p, 0, sizeof(int)));
    h->ngamma = d->ngamma;
    h->npparbar = d->npparbar;
  }

  // current_int (derivative_f0, ALPS_fns.f90:157-190): rectangle sums over the
  // table in the reference's order; closed-form species n q pdrift / m
  h->current_int.assign(nspec, 0.0);
  if (d->current_int) {
    h->current_int.assign(d->current_int, d->current_int + nspec);
    h->have_current = true;
  } else if (d->f0) {
    const double PI = 3.141592653589793238462643383279502884;
    for (int s = 0; s < nspec; ++s) {
      if (d->usebM && d->usebM[s]) {
        h->current_int[s] = d->ns[s] * d->qs[s] * d->bM_pdrift[s] / d->ms[s];
        continue;
      }
      const double dpp = d->pp[f_pp(nspec, nperp, npar, s, 2, 2, 0)] - d->pp[f_pp(nspec, nperp, npar, s, 1, 2, 0)];
      const double dpa = std::fabs(d->pp[f_pp(nspec, nperp, npar, s, 2, 2, 1)] -
                                   d->pp[f_pp(nspec, nperp, npar, s, 2, 1, 1)]);
      double c = 0.0;
      for (int i = 0; i <= nperp; ++i)
        for (int j = 0; j <= npar; ++j)
          c = c + (d->ns[s] * d->qs[s] / d->ms[s]) * d->pp[f_pp(nspec, nperp, npar, s, i, j, 0)] *
                      d->pp[f_pp(nspec, nperp, npar, s, i, j, 1)] *
                      d->f0[(size_t)s + (size_t)nspec * (i + (size_t)(nperp + 1) * j)] * 2.0 * PI *
                      dpp * dpa;
      h->current_int[s] = c;
    }
    h->have_current = true;
  }

  // problem descriptor (per-k pointers filled by set_k)
  ProbDev &pd = h->prob;
  std::memset(&pd, 0, sizeof(pd));
  h->lgen = 0;
  pd.nspec = nspec;
  pd.kperp_norm = h->kperp_norm;
  pd.P = P;
  pd.NRI = h->NRI;
  pd.vA = d->vA;
  pd.Tlim = d->Tlim;
  pd.gtab = h->gtab.p;
  for (int s = 0; s < nspec; ++s) {
    SpecDev &S = pd.sp[s];
    const double *pe = &h->pperp_h[(size_t)s * (nperp + 1)];
    const double *pa = &h->ppar_h[(size_t)s * (npar + 1)];
    S.nperp = nperp;
    S.npar = npar;
    S.acmethod = d->ACmethod[s];
    S.anum = s + 1;
    S.nfits = d->ACmethod[s] == 1 ? d->n_fits[s] : 0;
    for (int f = 0; f < S.nfits; ++f) S.fit_type[f] = d->fit_type[s + (size_t)nspec * f];
    S.q = d->qs[s];
    S.m = d->ms[s];
    S.dens = d->ns[s];
    S.dpperp = pe[2] - pe[1];
    S.dppar = pa[2] - pa[1];
    S.dppar_abs = std::fabs(pa[2] - pa[1]);
    S.int_ee = h->int_ee_h[s];
    S.pperp = h->pperp.p + (size_t)s * (nperp + 1);
    S.ppar = h->ppar.p + (size_t)s * (npar + 1);
    S.fitc = h->fitc.p + fit_off[s];
    if (bm(s)) {
      // closed form (ALPS_NHDS.f90:163-168): no table terms, no ee term
      h->usebm[s] = 1;
      h->any_bm = true;
      h->bm_nmax[s] = d->bM_nmax[s];
      h->bm_bz[s] = d->bM_Bessel_zero[s];
      h->bm_beta[s] = d->bM_beta[s];
      h->bm_alpha[s] = d->bM_alpha[s];
      h->bm_pdrift[s] = d->bM_pdrift[s];
      S.usebm = 1;
      S.bm_cold = d->bM_beta[s] == 0.0 ? 1 : 0;
      S.bm_Omega = d->qs[s] / d->ms[s];
      S.bm_ell = std::sqrt(d->ms[s] / (d->ns[s] * d->qs[s] * d->qs[s]));
      S.bm_vth = std::sqrt(d->bM_beta[s] / (d->ns[s] * d->ms[s]));
      S.bm_vd = d->bM_pdrift[s] / d->ms[s];
      S.bm_alpha = d->bM_alpha[s];
      S.int_ee = 0.0;
      S.acmethod = 1;
      S.nfits = 1;
      S.fit_type[0] = 1;
      S.fit_uniform = 1;
      S.nmax = -1;
      continue;
    }
    if (relf(s)) {
      // (Gamma, pparbar) tables, subluminal row limits, Gamma of the (pperp, ppar)
      // points and the omega-independent int_ee_rel (ALPS_fns_rel.f90:1097-1215)
      RelHost R;
      {
        const int rc = rel_host_tables(d, s, rel_index[s], nspec_rel, pe, pa, R);
        if (rc) return bail(rc);
      }
      const size_t np = R.G.size();
      const std::vector<double> &G = R.G, &PBv = R.PB, &F = R.F, &gpp = R.gpp;
      const std::vector<double2> &DF = R.DF;
      const std::vector<int2> &lim = R.lim;
      const double dg = R.dg, dpb = R.dpb, ee = R.ee;
      HIPCHK(h->grel[s].ensure(np));
      HIPCHK(h->pbrel[s].ensure(np));
      HIPCHK(h->f0rel[s].ensure(np));
      HIPCHK(h->dfrel[s].ensure(np));
      HIPCHK(h->rowlim[s].ensure(ngam + 1));
      HIPCHK(h->gampp[s].ensure(gpp.size()));
      HIPCHK(hipMemcpy(h->grel[s].p, G.data(), sizeof(double) * np, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(h->pbrel[s].p, PBv.data(), sizeof(double) * np, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(h->f0rel[s].p, F.data(), sizeof(double) * np, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(h->dfrel[s].p, DF.data(), sizeof(double2) * np, hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(h->rowlim[s].p, lim.data(), sizeof(int2) * (ngam + 1), hipMemcpyHostToDevice));
      HIPCHK(hipMemcpy(h->gampp[s].p, gpp.data(), sizeof(double) * gpp.size(), hipMemcpyHostToDevice));
      h->relsp[s] = 1;
      h->any_rel = true;
      S.rel = 1;
      S.ng = ngam;
      S.npb = npb;
      S.grel = h->grel[s].p;
      S.pbrel = h->pbrel[s].p;
      S.f0rel = h->f0rel[s].p;
      S.dfrel = h->dfrel[s].p;
      S.rowlim = h->rowlim[s].p;
      S.gam_pp = h->gampp[s].p;
      S.df1 = h->df1.p + (size_t)s * (nperp + 1) * (npar + 1);
      S.df2 = h->df2.p + (size_t)s * (nperp + 1) * (npar + 1);
      S.rel_dg = dg;
      S.rel_dpb = dpb;
      S.rel_gmin = R.gmin;
      S.rel_gmax = R.gmax;
      S.rel_gmono = R.gmono ? 1 : 0;
      S.int_ee = ee;
      for (int f = 0; f < S.nfits; ++f) S.fit_type[f] = 1;  // the Maxwell form of the table
      S.fit_uniform = 1;
      S.nmax = -1;
      continue;
    }
    if (d->ACmethod[s] == 1) {
      bool uni = true;
      for (int f = 0; f < d->n_fits[s] && uni; ++f) {
        if (d->fit_type[s + (size_t)nspec * f] != 1) uni = false;
        for (int k = 1; k <= 2 && uni; ++k) {
          const double v0 = d->param_fit[(size_t)s + (size_t)nspec * (0 + (size_t)(npf + 1) * (k + 5 * (size_t)f))];
          for (int i = 1; i <= nperp; ++i)
            if (d->param_fit[(size_t)s + (size_t)nspec * (i + (size_t)(npf + 1) * (k + 5 * (size_t)f))] != v0) {
              uni = false;
              break;
            }
        }
      }
      S.fit_uniform = uni ? 1 : 0;
      if (uni && d->n_fits[s] == 1) {
        // the zero test of landau_integrate (ALPS_fns.f90:1404-1405) over its
        // rows: c0 = u1 exp(-pc pperp^2) on iperp = 1..nperp-1 and 2..nperp
        double c1 = INFINITY, c2 = INFINITY;
        for (int i = 1; i <= nperp; ++i) {
          const double c0 = std::fabs(fitc[fit_off[s] + (size_t)i * FIT_STRIDE]);
          if (i <= nperp - 1) c1 = std::min(c1, c0);
          if (i >= 2) c2 = std::min(c2, c0);
        }
        S.lz_c1 = c1;
        S.lz_c2 = c2;
      }
    }
    if (S.acmethod == 1 && !S.fit_uniform && S.nfits >= 1) {
      // every fit a kappa fit whose exponent u4 stays in the lean power's domain
      S.fit_kappa = 1;
      for (int f = 0; f < S.nfits; ++f) {
        if (S.fit_type[f] != 2) S.fit_kappa = 0;
        for (int i = 0; i <= nperp && S.fit_kappa; ++i) {
          const double u4 = d->param_fit[(size_t)s + (size_t)nspec * (i + (size_t)(npf + 1) * (3 + 5 * (size_t)f))];
          if (!(std::fabs(u4) <= ALPSG_LEAN_MAXC)) S.fit_kappa = 0;
        }
      }
    }
    if (S.fit_kappa)
      h->lgen |= 2;
    else if (!(S.acmethod == 1 && S.fit_uniform && S.nfits == 1))
      h->lgen |= 1;
    if (d->ACmethod[s] == 2) {
      S.poly_kind = d->poly_kind[s];
      S.poly_order = d->poly_order[s];
      S.poly_logfit = d->logfit[s] ? 1 : 0;
      S.poly_log_max = d->poly_log_max[s];
      S.poly_norm1 = 5.e-1 * (pa[npar] + pa[0]);
      S.poly_norm2 = 5.e-1 * (pa[npar] - pa[0]);
      S.polyc = h->polyc.p + (size_t)s * (nperp + 1) * (maxorder + 1);
    }
  }
  HIPCHK(h->prob_d.ensure(1));
  *out = hold.release();
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_destroy(alpsg_handle h) {
  ALPSG_TRY
  delete h;
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_set_k(alpsg_handle h, double kperp, double kpar, const int *nmax_in, int *nmax_out) {
  ALPSG_TRY
  if (!h) return fail(ALPSG_E_ARG, "null handle");
  if (!(kpar != 0.0) || !std::isfinite(kpar) || !std::isfinite(kperp))
    return fail(ALPSG_E_ARG, "kpar must be finite and nonzero");
  const int nspec = h->nspec, nperp = h->nperp, npar = h->npar;
  int nmax[ALPSG_MAXSPEC];
  bool all_bm = true;
  for (int s = 0; s < nspec; ++s) all_bm = all_bm && h->usebm[s];
  if (nmax_in) {
    for (int s = 0; s < nspec; ++s) {
      if (h->usebm[s]) {
        nmax[s] = 1;
        continue;
      }
      if (nmax_in[s] < 1 || nmax_in[s] >= MTAB_N - 2) return fail(ALPSG_E_ARG, "nmax_in out of range");
      nmax[s] = nmax_in[s];
    }
  } else if (all_bm) {
    for (int s = 0; s < nspec; ++s) nmax[s] = 1;
  } else {
    // determine_nmax: first n >= 1 with max_iperp |J_n| <= Bessel_zero.  The
    // device evaluates n = 1..ntry, where ntry bounds the answer through
    // |J_n(x)| <= (x/2)^n / n!.
    double xmax = 0.0;
    for (int s = 0; s < nspec; ++s)
      for (int i = 0; i <= nperp && !h->usebm[s]; ++i)
        xmax = std::max(xmax, std::fabs(kperp * h->pperp_h[(size_t)s * (nperp + 1) + i] / h->qs[s]));
    int ntry = 1;
    double lb = 0.0, lz = std::log(std::max(h->Bessel_zero, 1e-300)) - std::log(1e6);
    while (ntry < MTAB_N - 16) {
      lb = ntry * std::log(std::max(xmax, 1e-300) / 2.0) - std::lgamma(ntry + 1.0);
      if (lb < lz && ntry > xmax) break;
      ++ntry;
    }
    ntry = std::min(ntry + 8, MTAB_N - 4);
    HIPCHK(h->bmax.ensure((size_t)nspec * ntry));
    HIPCHK(launch_bessel_max(h->st, nspec, nperp, ntry, kperp, h->pperp.p, h->qs_d.p, h->mtab.p, h->bmax.p));
    std::vector<double> bm((size_t)nspec * ntry);
    HIPCHK(hipMemcpyAsync(bm.data(), h->bmax.p, sizeof(double) * bm.size(), hipMemcpyDeviceToHost, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
    for (int s = 0; s < nspec; ++s) {
      if (h->usebm[s]) {  // determine_nmax, ALPS_fns.f90:4012-4013
        nmax[s] = 1;
        continue;
      }
      int n = 0;
      for (int t = 1; t <= ntry; ++t)
        if (!(bm[(size_t)s * ntry + (t - 1)] > h->Bessel_zero)) {
          n = t;
          break;
        }
      if (n == 0) return fail(ALPSG_E_ARG, "Bessel cutoff not reached (kperp too large?)");
      nmax[s] = n;
    }
  }
  int nmaxR = 0;
  for (int s = 0; s < nspec; ++s) {
    SpecDev &S = h->prob.sp[s];
    const int nm = nmax[s];
    h->nmax[s] = nm;
    if (h->usebm[s]) {
      // omega-independent part of calc_chi (ALPS_NHDS.f90:163-185)
      if (!S.bm_cold) {
        const double a = kperp * S.bm_vth / S.bm_Omega;
        S.bm_z = ((0.5 * a) * a) * S.bm_alpha;
        const double b = S.bm_vth / S.bm_Omega;
        S.bm_zp = ((0.5 * b) * b) * S.bm_alpha;
        std::vector<double2> tab;
        S.bm_nrun = nhds_tables(S.bm_z, h->bm_nmax[s], h->bm_bz[s], tab);
        HIPCHK(h->bmi[s].ensure(tab.size()));
        HIPCHK(hipMemcpyAsync(h->bmi[s].p, tab.data(), sizeof(double2) * tab.size(),
                              hipMemcpyHostToDevice, h->st));
        S.bmi = h->bmi[s].p;
      }
      S.nmax = -1;
      continue;
    }
    HIPCHK(h->bes[s].ensure((size_t)(nm + 3) * (nperp + 1)));
    HIPCHK(h->atab[s].ensure((size_t)(nm + 1) * (nperp + 1) * ATAB_STRIDE));
    HIPCHK(h->nodes[s].ensure((size_t)(nm + 1) * (npar - 1) * NODE_STRIDE));
    HIPCHK(launch_species_k(h->st, nperp, npar, nm, kperp, kpar, h->qs[s], h->ms[s],
                            h->pperp.p + (size_t)s * (nperp + 1), h->ppar.p + (size_t)s * (npar + 1),
                            h->df1.p + (size_t)s * (nperp + 1) * (npar + 1),
                            h->df2.p + (size_t)s * (nperp + 1) * (npar + 1), h->mtab.p, h->bes[s].p,
                            h->atab[s].p, h->nodes[s].p));
    S.nmax = nm;
    S.z = h->kperp_norm ? kperp / h->qs[s] : 1.0 / h->qs[s];
    S.atab = h->atab[s].p;
    S.nodes = h->nodes[s].p;
    S.crec = S.rsum = S.gcum = nullptr;
    S.ser = S.serD = nullptr;
    S.lsum = nullptr;
    if (!h->relsp[s] && S.acmethod == 1 && S.fit_uniform && S.nfits == 1) {
      HIPCHK(h->lsum[s].ensure((size_t)(nm + 1) * 6));
      HIPCHK(launch_lsum(h->st, nperp, nm, S.dpperp, h->pperp.p + (size_t)s * (nperp + 1), S.fitc,
                         h->atab[s].p, h->lsum[s].p));
      S.lsum = h->lsum[s].p;
    }
    if (!h->relsp[s]) {
      HIPCHK(h->crec[s].ensure((size_t)(nm + 1) * (npar - 1) * CREC_STRIDE));
      HIPCHK(h->rsum[s].ensure((size_t)(nm + 1) * RSUM_STRIDE));
      HIPCHK(h->gcum[s].ensure((size_t)(nm + 1) * npar * GCUM_STRIDE));
      HIPCHK(h->ser[s].ensure((size_t)(nm + 1) * 2 * SER_Q * SER_STRIDE));
      HIPCHK(h->serD[s].ensure((size_t)(nm + 1) * 2));
      HIPCHK(launch_split_tables(h->st, npar, nm, kpar, h->qs[s], h->ms[s],
                                 h->ppar.p + (size_t)s * (npar + 1), h->nodes[s].p, h->crec[s].p,
                                 h->rsum[s].p, h->gcum[s].p));
      S.crec = h->crec[s].p;
      S.rsum = h->rsum[s].p;
      S.gcum = h->gcum[s].p;
    }
    if (h->relsp[s]) {
      // the harmonics run in k_rel_nonres / k_rel_res, not in the sweep
      const double m = h->ms[s], q = h->qs[s], vA = h->vA;
      S.nmax_rel = nm;
      S.nmax = -1;
      S.rel_pref = ((2.0 * 3.141592653589793238462643383279502884) * ((m / vA) * (m / vA) * (m / vA))) *
                   (q * vA / (kpar * m));
      S.rel_zfac = kperp * m / (vA * q);
      S.rel_zbar = h->kperp_norm ? kperp * m / (vA * q) : m / (vA * q);
      const int npts = (h->ngamma + 1) * (h->npparbar + 1);
      HIPCHK(h->relbes[s].ensure((size_t)(nm + 2) * npts));
      HIPCHK(launch_rel_bessel(h->st, npts, nm, S.rel_zfac, S.grel, S.pbrel, h->mtab.p,
                               h->relbes[s].p));
      S.relbes = h->relbes[s].p;
      nmaxR = std::max(nmaxR, nm);
    }
  }
  h->prob.kperp = kperp;
  h->prob.kpar = kpar;
  h->prob.nmaxR = nmaxR;
  // the far-row series tables of this k are built by the first batch that is
  // large enough to gain from them (ensure_series)
  h->ser_pending = true;
  h->ser_built = false;
  HIPCHK(hipMemcpyAsync(h->prob_d.p, &h->prob, sizeof(ProbDev), hipMemcpyHostToDevice, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  h->kperp = kperp;
  h->kpar = kpar;
  h->have_k = true;
  if (nmax_out)
    for (int s = 0; s < nspec; ++s) nmax_out[s] = nmax[s];
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_calc_eigen(alpsg_handle h, int n, const double *om, int eigen, int heat, double *ef,
                     double *bf, double *Us, double *ds, double *Ps, double *Ps_split,
                     double *W_EM) {
  ALPSG_TRY
  if (!h || (n > 0 && !om)) return fail(ALPSG_E_ARG, "null argument");
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  if (n <= 0) return ALPSG_OK;
  return calc_eigen_batch(h, n, reinterpret_cast<const cplx *>(om), eigen != 0, heat != 0,
                          reinterpret_cast<cplx *>(ef), reinterpret_cast<cplx *>(bf),
                          reinterpret_cast<cplx *>(Us), reinterpret_cast<cplx *>(ds), Ps, Ps_split,
                          W_EM);
  ALPSG_CATCH
}

int alpsg_autotune(alpsg_handle h, int n, const double *om, int *shape_out) {
  ALPSG_TRY
  if (!h) return fail(ALPSG_E_ARG, "null handle");
  return autotune(h, n, reinterpret_cast<const double2 *>(om), shape_out);
  ALPSG_CATCH
}

int alpsg_set_shape(alpsg_handle h, int n, const int *shape) {
  ALPSG_TRY
  if (!h) return fail(ALPSG_E_ARG, "null handle");
  if (n < 1) return fail(ALPSG_E_ARG, "n must be >= 1");
  const int b = size_bucket(n);
  if (!shape) {
    h->tuned_valid[b] = 0;
    return ALPSG_OK;
  }
  if (!shape_is_valid(shape)) return fail(ALPSG_E_ARG, "launch shape not instantiated");
  for (int k = 0; k < 5; ++k) h->tuned_shape[b][k] = shape[k];
  h->tuned_valid[b] = 1;
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_get_shape(alpsg_handle h, int n, int *shape_out) {
  ALPSG_TRY
  if (!h || !shape_out) return fail(ALPSG_E_ARG, "null argument");
  if (n < 1) return fail(ALPSG_E_ARG, "n must be >= 1");
  LaunchShape sh = shape_for(h, n);
  const int t[5] = {sh.nch_a, sh.w_a, sh.rc_a, sh.nch_b, sh.rb};
  for (int k = 0; k < 5; ++k) shape_out[k] = t[k];
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_get_nmax(alpsg_handle h, int *nmax_out) {
  ALPSG_TRY
  if (!h || !nmax_out) return fail(ALPSG_E_ARG, "null argument");
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  for (int s = 0; s < h->nspec; ++s) nmax_out[s] = h->nmax[s];
  return ALPSG_OK;
  ALPSG_CATCH
}

long long alpsg_eval_count(alpsg_handle h) { return h ? h->evals : -1; }

int alpsg_last_root_times(alpsg_handle h, int nroots, double *ms_out) {
  ALPSG_TRY
  if (!h || !ms_out) return fail(ALPSG_E_ARG, "null argument");
  if (nroots != (int)h->root_ms.size()) return fail(ALPSG_E_ARG, "nroots differs from the last secant batch");
  for (int r = 0; r < nroots; ++r) ms_out[r] = h->root_ms[r];
  return ALPSG_OK;
  ALPSG_CATCH
}

double alpsg_last_kernel_ms(alpsg_handle h) { return h ? h->last_ms : -1.0; }

int alpsg_last_scan_steps(alpsg_handle h, int nsteps, int nroots, double *ms_out, int *iters_out) {
  ALPSG_TRY
  if (!h) return fail(ALPSG_E_ARG, "null handle");
  if (nsteps != (int)h->scan_ms.size() || (size_t)nsteps * nroots != h->scan_iters.size())
    return fail(ALPSG_E_ARG, "nsteps or nroots differs from the last om_scan");
  for (int t = 0; ms_out && t < nsteps; ++t) ms_out[t] = h->scan_ms[t];
  for (size_t t = 0; iters_out && t < h->scan_iters.size(); ++t) iters_out[t] = h->scan_iters[t];
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_set_plain_roots(alpsg_handle h, int plain) {
  ALPSG_TRY
  if (!h) return fail(ALPSG_E_ARG, "null handle");
  h->plain_roots = plain != 0;
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_set_sweep_series(alpsg_handle h, int on) {
  ALPSG_TRY
  if (!h) return fail(ALPSG_E_ARG, "null handle");
  h->series_on = on != 0;
  if (h->have_k && h->ser_built) {
    // this k's tables exist: point the descriptor at them or away
    for (int s = 0; s < h->nspec; ++s) {
      SpecDev &S = h->prob.sp[s];
      const bool has = S.crec != nullptr && h->ser[s].p != nullptr;
      S.ser = (h->series_on && has) ? h->ser[s].p : nullptr;
      S.serD = (h->series_on && has) ? h->serD[s].p : nullptr;
    }
    HIPCHK(hipMemcpyAsync(h->prob_d.p, &h->prob, sizeof(ProbDev), hipMemcpyHostToDevice, h->st));
    HIPCHK(hipStreamSynchronize(h->st));
  }
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_sweep_work(alpsg_handle h, int n, const double *om, double *work3) {
  ALPSG_TRY
  if (!h || !work3 || (n > 0 && !om)) return fail(ALPSG_E_ARG, "null argument");
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  if (n < 0) return fail(ALPSG_E_ARG, "negative batch size");
  for (int k = 0; k < 3; ++k) work3[k] = 0.0;
  if (n == 0) return ALPSG_OK;
  const LaunchShape sh = shape_for(h, n);
  if (const int rc = ensure_series(h, n)) return rc;
  HIPCHK(h->om_d.ensure(n));
  HIPCHK(h->chiA_d.ensure((size_t)n * sh.w_a * sh.rc_a * h->nspec * 6));
  HIPCHK(h->work_d.ensure(3));
  HIPCHK(hipMemcpyAsync(h->om_d.p, om, sizeof(double2) * n, hipMemcpyHostToDevice, h->st));
  HIPCHK(hipMemsetAsync(h->work_d.p, 0, sizeof(unsigned long long) * 3, h->st));
  HIPCHK(launch_sweep(h->st, sh.nch_a, sh.w_a, sh.rc_a, h->prob_d.p, n, h->om_d.p, nullptr,
                      h->chiA_d.p, nullptr, h->npar + 1, nullptr, h->work_d.p));
  unsigned long long w[3] = {0, 0, 0};
  HIPCHK(hipMemcpyAsync(w, h->work_d.p, sizeof(w), hipMemcpyDeviceToHost, h->st));
  HIPCHK(hipStreamSynchronize(h->st));
  for (int k = 0; k < 3; ++k) work3[k] = (double)w[k];
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_last_kernel_split(alpsg_handle h, double *ms3) {
  ALPSG_TRY
  if (!h || !ms3) return fail(ALPSG_E_ARG, "null argument");
  for (int k = 0; k < 3; ++k) ms3[k] = h->last_split[k];
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_disp_batch(alpsg_handle h, int n, const double *om, double *det, double *chi0,
                     double *chi0_low, double *wave) {
  ALPSG_TRY
  if (!h || (n > 0 && (!om || !det))) return fail(ALPSG_E_ARG, "null argument");
  return run_disp(h, n, om, det, chi0, chi0_low, wave);
  ALPSG_CATCH
}

int alpsg_disp_plain(alpsg_handle h, int n, const double *om, double *det) {
  ALPSG_TRY
  if (!h || (n > 0 && (!om || !det))) return fail(ALPSG_E_ARG, "null argument");
  for (int i = 0; i < n; ++i) {
    int rc = run_disp(h, 1, om + 2 * i, det + 2 * i, nullptr, nullptr, nullptr);
    if (rc) return rc;
  }
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_secant_batch(alpsg_handle h, const alpsg_root_params *rp, int nroots, double *om,
                       double *D_final, int *iters) {
  ALPSG_TRY
  if (!h || !rp || (nroots > 0 && !om)) return fail(ALPSG_E_ARG, "null argument");
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  if (nroots <= 0) return ALPSG_OK;
  return root_batch(h, rp, nroots, reinterpret_cast<cplx *>(om), reinterpret_cast<cplx *>(D_final),
                    iters);
  ALPSG_CATCH
}

}  // extern "C"

namespace {

// omega of map point (ir, ii), 1-based (ALPS_fns.f90:3650-3700): the products
// and sums of the grid that alpsg_map_search fills
cplx map_point(double omi, double omf, double gami, double gamf, int nr, int ni, int loggridw,
               int loggridg, int ir, int ii) {
#pragma clang fp contract(off)
  double dr = omf - omi, di = gamf - gami;
  if (nr > 1) dr = (omf - omi) / (1.0 * (nr - 1));
  if (ni > 1) di = (gamf - gami) / (1.0 * (ni - 1));
  double wr, wi;
  if (loggridw) {
    wr = omi;
    if (nr > 1) wr = omi * std::pow(omf / omi, (1.0 * (ir - 1)) / (1.0 * (nr - 1)));
  } else {
    wr = omi + dr * (1.0 * (ir - 1));
  }
  if (loggridg) {
    wi = gami;
    if (ni > 1) wi = gami * std::pow(gamf / gami, (1.0 * (ii - 1)) / (1.0 * (ni - 1)));
  } else {
    wi = gami + di * (1.0 * (ii - 1));
  }
  return cplx(wr, wi);
}

// map_search's grid, det D, sentinels and minima.  With late != NULL the
// host waits only for the minimum flags: D and log10|D| go to pinned staging
// on st_copy, and *late is set to the host work that copies them into cal_out
// / val_out (the caller runs it before it returns).
int map_search_core(alpsg_ctx *h, double omi, double omf, double gami, double gamf, int nr, int ni,
                    int loggridw, int loggridg, double *om_out, double *cal_out, double *val_out,
                    int numroots, int *iroots, int *nfound, std::function<int()> *late) {
  const size_t N = (size_t)nr * ni;
  // pinned staging: grid [N] complex, D [N] complex, log10|D| [N], flags [N]
  HIPCHK(h->map_pin.ensure(N * (sizeof(cplx) * 2 + sizeof(double) + 1)));
  cplx *om = reinterpret_cast<cplx *>(h->map_pin.p);
  cplx *cal_h = om + N;
  double *val_h = reinterpret_cast<double *>(cal_h + N);
  unsigned char *hit = reinterpret_cast<unsigned char *>(val_h + N);
  double dr = omf - omi, di = gamf - gami;
  if (nr > 1) dr = (omf - omi) / (1.0 * (nr - 1));
  if (ni > 1) di = (gamf - gami) / (1.0 * (ni - 1));
  // the grid (ALPS_fns.f90:3650-3700) into g[(ir-1) + nr (ii-1)]
  auto fill_grid = [&](cplx *g) {
#pragma clang fp contract(off)
    for (int ir = 1; ir <= nr; ++ir) {
      double wr;
      if (loggridw) {
        wr = omi;
        if (nr > 1) wr = omi * std::pow(omf / omi, (1.0 * (ir - 1)) / (1.0 * (nr - 1)));
      } else {
        wr = omi + dr * (1.0 * (ir - 1));
      }
      for (int ii = 1; ii <= ni; ++ii) {
        double wi;
        if (loggridg) {
          wi = gami;
          if (ni > 1) wi = gami * std::pow(gamf / gami, (1.0 * (ii - 1)) / (1.0 * (ni - 1)));
        } else {
          wi = gami + di * (1.0 * (ii - 1));
        }
        g[(size_t)(ir - 1) + (size_t)nr * (ii - 1)] = cplx(wr, wi);
      }
    }
  };
  // det D of every grid point, then on the device log10|D| with the NaN / Inf
  // sentinels of ALPS_fns.f90:3723-3742 and the minimum flags of find_minima.
  // A linear grid is built on the device (the same products and sums), and the
  // caller's copy on the host while the device works; a log grid (std::pow) on
  // the host.
  const bool dev_grid = !loggridw && !loggridg;
  HIPCHK(h->om_d.ensure(N));
  if (dev_grid) {
    HIPCHK(launch_map_grid(h->st, nr, ni, omi, dr, gami, di, h->om_d.p));
  } else {
    fill_grid(om);
    HIPCHK(hipMemcpyAsync(h->om_d.p, om, sizeof(cplx) * N, hipMemcpyHostToDevice, h->st));
  }
  LaunchShape sh = shape_for(h, (int)N);
  HIPCHK(hipEventRecord(h->ev0, h->st));
  int rc = disp_device(h, (int)N, h->om_d.p, nullptr, false, false, false, sh);
  if (rc) return rc;
  HIPCHK(hipEventRecord(h->ev1, h->st));
  HIPCHK(h->mcal_d.ensure(N));
  HIPCHK(h->mval_d.ensure(N));
  HIPCHK(h->mhit_d.ensure(N));
  HIPCHK(launch_map_post(h->st, nr, ni, h->det_d.p, h->mcal_d.p, h->mval_d.p, h->mhit_d.p));
  HIPCHK(hipMemcpyAsync(hit, h->mhit_d.p, N, hipMemcpyDeviceToHost, h->st));
  const bool defer = late != nullptr && (cal_out || val_out);
  hipStream_t cs = h->st;
  if (defer) {
    HIPCHK(hipEventRecord(h->ev_map, h->st));
    HIPCHK(hipStreamWaitEvent(h->st_copy, h->ev_map, 0));
    cs = h->st_copy;
  }
  if (cal_out)
    HIPCHK(hipMemcpyAsync(cal_h, h->mcal_d.p, sizeof(cplx) * N, hipMemcpyDeviceToHost, cs));
  if (val_out)
    HIPCHK(hipMemcpyAsync(val_h, h->mval_d.p, sizeof(double) * N, hipMemcpyDeviceToHost, cs));
  if (defer) HIPCHK(hipEventRecord(h->ev_copy, h->st_copy));
  int relerr = 0;
  if (h->any_rel)
    HIPCHK(hipMemcpyAsync(&relerr, h->err_d.p, sizeof(int), hipMemcpyDeviceToHost, h->st));
  if (om_out) {
    if (dev_grid)
      fill_grid(reinterpret_cast<cplx *>(om_out));
    else
      std::memcpy(om_out, om, sizeof(cplx) * N);
  }
  HIPCHK(hipStreamSynchronize(h->st));
  if (defer) {
    *late = [h, cal_out, val_out, cal_h, val_h, N]() -> int {
      HIPCHK(hipEventSynchronize(h->ev_copy));
      if (cal_out) std::memcpy(cal_out, cal_h, sizeof(cplx) * N);
      if (val_out) std::memcpy(val_out, val_h, sizeof(double) * N);
      return ALPSG_OK;
    };
  } else {
    if (cal_out) std::memcpy(cal_out, cal_h, sizeof(cplx) * N);
    if (val_out) std::memcpy(val_out, val_h, sizeof(double) * N);
  }
  if (relerr) {
    if (defer) {
      std::function<int()> w;
      w.swap(*late);
      (void)w();
    }
    const int zero = 0;
    HIPCHK(hipMemcpy(h->err_d.p, &zero, sizeof(int), hipMemcpyHostToDevice));
    return fail(ALPSG_E_ARG,
                "a resonance's principal-value window exceeds the subluminal part of a Gamma row "
                "(the reference stops with alps_error(8)); refine npparbar or lower "
                "positions_principal");
  }
  float ms = 0.f, a = 0.f, b = 0.f, c = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, h->ev0, h->ev1));
  HIPCHK(hipEventElapsedTime(&a, h->ev0, h->ev_sweep));
  HIPCHK(hipEventElapsedTime(&b, h->ev_sweep, h->ev_res));
  HIPCHK(hipEventElapsedTime(&c, h->ev_res, h->ev1));
  h->last_ms = ms;
  h->last_split[0] = a;
  h->last_split[1] = b;
  h->last_split[2] = c;
  h->evals += (long long)N;
  // the minima in find_minima's order: ii from ni down to 1, ir from 1 to nr
  int n = 0;
  for (int ii = ni; ii >= 1; --ii) {
    const unsigned char *row = hit + (size_t)nr * (ii - 1);
    for (int ir = 1; ir <= nr; ++ir) {
      if (!row[ir - 1]) continue;
      if (iroots && n < numroots) {
        iroots[2 * n] = ir;
        iroots[2 * n + 1] = ii;
      }
      ++n;
    }
  }
  if (nfound) *nfound = n;
  return ALPSG_OK;
}

// runs host work left by a deferred map search, if any
int finish_host_work(std::function<int()> &work) {
  if (!work) return ALPSG_OK;
  std::function<int()> w;
  w.swap(work);
  return w();
}

}  // namespace

extern "C" {

int alpsg_map_search(alpsg_handle h, double omi, double omf, double gami, double gamf, int nr,
                     int ni, int loggridw, int loggridg, double *om_out, double *cal_out,
                     double *val_out, int numroots, int *iroots, int *nfound) {
  ALPSG_TRY
  if (!h) return fail(ALPSG_E_ARG, "null handle");
  if (nr < 1 || ni < 1) return fail(ALPSG_E_ARG, "nr, ni must be >= 1");
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  if ((size_t)nr * ni > (size_t)INT32_MAX) return fail(ALPSG_E_ARG, "map too large");
  return map_search_core(h, omi, omf, gami, gamf, nr, ni, loggridw, loggridg, om_out, cal_out,
                         val_out, numroots, iroots, nfound, nullptr);
  ALPSG_CATCH
}

int alpsg_map_search_refine(alpsg_handle h, const alpsg_root_params *rp, double omi, double omf,
                            double gami, double gamf, int nr, int ni, int loggridw, int loggridg,
                            double *om_out, double *cal_out, double *val_out, int numroots,
                            int *iroots, int *nfound, double *roots, double *D_final, int *iters) {
  ALPSG_TRY
  if (!h || !rp) return fail(ALPSG_E_ARG, "null argument");
  if (nr < 1 || ni < 1) return fail(ALPSG_E_ARG, "nr, ni must be >= 1");
  if (numroots < 0) return fail(ALPSG_E_ARG, "numroots must be >= 0");
  if (numroots > 0 && !roots) return fail(ALPSG_E_ARG, "null roots");
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  if ((size_t)nr * ni > (size_t)INT32_MAX) return fail(ALPSG_E_ARG, "map too large");
  if (rp->secant_method < 0 || rp->secant_method > 2)
    return fail(ALPSG_E_ARG, "secant_method must be 0, 1 or 2");
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int> ir(2 * (size_t)numroots + 2);
  int nf = 0;
  std::function<int()> late;
  int rc = map_search_core(h, omi, omf, gami, gamf, nr, ni, loggridw, loggridg, om_out, cal_out,
                           val_out, numroots, ir.data(), &nf, &late);
  if (rc) return rc;
  // the kernel times this call reports are the map batch's
  const double map_ms = h->last_ms, map_split[3] = {h->last_split[0], h->last_split[1],
                                                    h->last_split[2]};
  if (iroots)
    for (int i = 0; i < 2 * std::min(nf, numroots); ++i) iroots[i] = ir[i];
  if (nfound) *nfound = nf;
  // refine_guess (ALPS_fns.f90:3793-3856) on the recorded minima
  const int n = std::min(nf, numroots);
  cplx *w = reinterpret_cast<cplx *>(roots);
  for (int r = 0; r < n; ++r)
    w[r] = map_point(omi, omf, gami, gamf, nr, ni, loggridw, loggridg, ir[2 * r], ir[2 * r + 1]);
  // the solvers that step on the host take the grids first; the device-stepped
  // secant_osc copies them while its first rounds run
  if (rp->secant_method != 2 || h->plain_roots || n == 0) {
    rc = finish_host_work(late);
    if (rc) return rc;
  }
  const double t_ref = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (n > 0) {
    rc = root_batch(h, rp, n, w, reinterpret_cast<cplx *>(D_final), iters, &late);
    const int rc2 = finish_host_work(late);
    if (rc) return rc;
    if (rc2) return rc2;
    for (auto &t : h->root_ms) t += t_ref;
  }
  h->last_ms = map_ms;
  for (int k = 0; k < 3; ++k) h->last_split[k] = map_split[k];
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_om_scan(alpsg_handle h, const alpsg_root_params *rp, int scan_type, int log_scan,
                  double diff, double diff2, int n_out, int n_res, int nroots, double *roots,
                  double *kperp_out, double *kpar_out, double *roots_out) {
  ALPSG_TRY
  if (!h || !rp || !roots || !kperp_out || !kpar_out || !roots_out) return fail(ALPSG_E_ARG, "null argument");
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  if (scan_type < 0 || scan_type > 4 || n_out < 1 || n_res < 1 || nroots < 1) return fail(ALPSG_E_ARG, "scan parameters");
  cplx *R = reinterpret_cast<cplx *>(roots);
  cplx *RO = reinterpret_cast<cplx *>(roots_out);
  std::vector<int> jump(nroots, 1);
  int row = 0;
  kperp_out[0] = h->kperp;
  kpar_out[0] = h->kpar;
  for (int r = 0; r < nroots; ++r) RO[r] = R[r];
  const double kperp_last = h->kperp, kpar_last = h->kpar;
  const double theta_0 = std::atan(kperp_last / kpar_last);
  const double k_0 = std::sqrt(kperp_last * kperp_last + kpar_last * kpar_last);
  const int nt = n_out * n_res;
  std::vector<cplx> om;
  std::vector<int> idx;
  const auto t0 = std::chrono::steady_clock::now();
  h->scan_ms.assign(nt, 0.0);
  h->scan_iters.assign((size_t)nt * nroots, -1);
  std::vector<int> its;
  for (int it = 1; it <= nt; ++it) {
    double kperp = h->kperp, kpar = h->kpar;
    switch (scan_type) {
      case 0:
        if (log_scan) {
          kperp = std::pow(10.0, std::log10(kperp_last) + diff * it);
          kpar = std::pow(10.0, std::log10(kpar_last) + diff2 * it);
        } else {
          kperp = kperp_last + diff * it;
          kpar = kpar_last + diff2 * it;
        }
        break;
      case 1: {
        double th = log_scan ? std::pow(10.0, std::log10(theta_0) + diff * it) : theta_0 + diff * it;
        kperp = k_0 * std::sin(th);
        kpar = k_0 * std::cos(th);
      } break;
      case 2: {
        double kt = log_scan ? std::pow(10.0, std::log10(k_0) + diff * it) : (k_0) + diff * it;
        kperp = kt * std::sin(theta_0);
        kpar = kt * std::cos(theta_0);
      } break;
      case 3:
        kperp = log_scan ? std::pow(10.0, std::log10(kperp_last) + diff * it) : kperp_last + diff * it;
        break;
      case 4:
        kpar = log_scan ? std::pow(10.0, std::log10(kpar_last) + diff * it) : kpar_last + diff * it;
        break;
    }
    // scan type 4 keeps kperp, so nmax and the Bessel table are kept as well
    int rc;
    if (scan_type != 4) {
      rc = alpsg_set_k(h, kperp, kpar, nullptr, nullptr);
    } else {
      int nm[ALPSG_MAXSPEC];
      for (int s = 0; s < h->nspec; ++s) nm[s] = h->nmax[s];
      rc = alpsg_set_k(h, kperp, kpar, nm, nullptr);
    }
    if (rc) return rc;
    bool any = false;
    for (int r = 0; r < nroots; ++r) any = any || jump[r];
    if (!any) return fail(ALPSG_E_NOTFOUND, "all roots dropped during the scan (alps_error(9))");
    om.clear();
    idx.clear();
    for (int r = 0; r < nroots; ++r)
      if (jump[r]) {
        om.push_back(R[r]);
        idx.push_back(r);
      }
    std::vector<cplx> Df(om.size());
    its.assign(om.size(), 0);
    rc = root_batch(h, rp, (int)om.size(), om.data(), Df.data(), its.data());
    if (rc) return rc;
    for (size_t t = 0; t < idx.size(); ++t) R[idx[t]] = om[t];
    for (size_t t = 0; t < idx.size(); ++t) h->scan_iters[(size_t)(it - 1) * nroots + idx[t]] = its[t];
    for (size_t t = 0; t < idx.size(); ++t) {
      int r = idx[t];
      double re = R[r].real();
      if (std::isnan(re) || std::fabs(re) >= 1.79769313486231570815e+308) jump[r] = 0;
      for (int imm = 0; imm < r; ++imm)
        if (std::abs(R[r] - R[imm]) < rp->D_gap) {
          R[r] = 0.0;
          jump[r] = 0;
        }
    }
    if (it % n_res == 0) {
      ++row;
      kperp_out[row] = h->kperp;
      kpar_out[row] = h->kpar;
      for (int r = 0; r < nroots; ++r) RO[(size_t)row * nroots + r] = R[r];
    }
    h->scan_ms[it - 1] =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return ALPSG_OK;
  ALPSG_CATCH
}

}  // extern "C"

// om_double_scan (ALPS_fns.f90:2900-3591): scan 1 outer; from every output
// point of scan 1, scan 2 inner.  Each k step solves the roots still followed
// in one batch (the reference solves them one after another; a root's solve
// reads only its own previous root, and the root-too-close checks run in root
// order after the batch, so the results are the same).
namespace {

void scan_k(const alpsg_scan_params *sc, int it, double kperp_last, double kpar_last, double &kperp,
            double &kpar) {
  const double theta_0 = std::atan(kperp_last / kpar_last);
  const double k_0 = std::sqrt(kperp_last * kperp_last + kpar_last * kpar_last);
  switch (sc->scan_type) {
    case 0:
      if (sc->log_scan) {
        kperp = std::pow(10.0, std::log10(kperp_last) + sc->diff * it);
        kpar = std::pow(10.0, std::log10(kpar_last) + sc->diff2 * it);
      } else {
        kperp = kperp_last + sc->diff * it;
        kpar = kpar_last + sc->diff2 * it;
      }
      break;
    case 1: {
      const double th = sc->log_scan ? std::pow(10.0, std::log10(theta_0) + sc->diff * it)
                                     : theta_0 + sc->diff * it;
      kperp = k_0 * std::sin(th);
      kpar = k_0 * std::cos(th);
    } break;
    case 2: {
      const double kt = sc->log_scan ? std::pow(10.0, std::log10(k_0) + sc->diff * it) : (k_0) + sc->diff * it;
      kperp = kt * std::sin(theta_0);
      kpar = kt * std::cos(theta_0);
    } break;
    case 3:
      kperp = sc->log_scan ? std::pow(10.0, std::log10(kperp_last) + sc->diff * it) : kperp_last + sc->diff * it;
      break;
    case 4:
      kpar = sc->log_scan ? std::pow(10.0, std::log10(kpar_last) + sc->diff * it) : kpar_last + sc->diff * it;
      break;
  }
}

// a kpar-only scan keeps nmax and the Bessel table (ALPS_fns.f90:3224-3231)
int move_k(alpsg_ctx *h, int type, double kperp, double kpar) {
  if (type != 4) return alpsg_set_k(h, kperp, kpar, nullptr, nullptr);
  int nm[ALPSG_MAXSPEC];
  for (int s = 0; s < h->nspec; ++s) nm[s] = h->usebm[s] ? 1 : h->nmax[s];
  return alpsg_set_k(h, kperp, kpar, nm, nullptr);
}

}  // namespace

extern "C" int alpsg_om_double_scan(alpsg_handle h, const alpsg_root_params *rp,
                                    const alpsg_scan_params *s1, const alpsg_scan_params *s2,
                                    int nroots, double *roots, int eigen, int heat,
                                    double *kperp_out, double *kpar_out, double *roots_out,
                                    int *written, double *ef, double *bf, double *Us, double *ds,
                                    double *Ps, double *Ps_split, double *W_EM) {
  ALPSG_TRY
  if (!h || !rp || !s1 || !s2 || !roots || !kperp_out || !kpar_out || !roots_out || !written)
    return fail(ALPSG_E_ARG, "null argument");
  if (!h->have_k) return fail(ALPSG_E_STATE, "alpsg_set_k has not been called");
  if (nroots < 1) return fail(ALPSG_E_ARG, "nroots < 1");
  for (const alpsg_scan_params *sc : {s1, s2})
    if (sc->scan_type < 0 || sc->scan_type > 4 || sc->n_out < 1 || sc->n_res < 1)
      return fail(ALPSG_E_ARG, "scan parameters");
  if (s1->scan_type == s2->scan_type) return fail(ALPSG_E_ARG, "both scans of the same type (alps_error(5))");
  if (s1->scan_type == 0 || s2->scan_type == 0)
    return fail(ALPSG_E_ARG, "scan type 0 in a double scan (alps_error(6))");
  if ((eigen || heat) && (!ef || !bf))
    return fail(ALPSG_E_ARG, "ef and bf are needed with eigen or heat");
  if (eigen && (!Us || !ds)) return fail(ALPSG_E_ARG, "Us and ds are needed with eigen");
  if (heat && (!Ps || !Ps_split || !W_EM)) return fail(ALPSG_E_ARG, "Ps, Ps_split and W_EM are needed with heat");
  const int ns = h->nspec;
  const int nt = s1->n_out * s1->n_res, nt2 = s2->n_out * s2->n_res;
  const int nrow2 = s2->n_out + 1;
  const size_t cells = (size_t)(s1->n_out + 1) * nrow2 * nroots;
  for (size_t c = 0; c < cells; ++c) written[c] = 0;
  cplx *R = reinterpret_cast<cplx *>(roots);
  cplx *RO = reinterpret_cast<cplx *>(roots_out);
  std::vector<int> jump(nroots, 1);
  std::vector<cplx> omlast(R, R + nroots), wroots(R, R + nroots);
  const double kperp_last = h->kperp, kpar_last = h->kpar;
  std::vector<cplx> om, Df;
  std::vector<int> idx;
  // solve the followed roots at the handle's k from `from`; results into wroots
  auto solve = [&](const std::vector<cplx> &from) -> int {
    om.clear();
    idx.clear();
    for (int r = 0; r < nroots; ++r)
      if (jump[r]) {
        om.push_back(from[r]);
        idx.push_back(r);
      }
    Df.assign(om.size(), cplx(0.0, 0.0));
    const int rc = root_batch(h, rp, (int)om.size(), om.data(), Df.data(), nullptr);
    if (rc) return rc;
    for (size_t t = 0; t < idx.size(); ++t) wroots[idx[t]] = om[t];
    return ALPSG_OK;
  };
  auto bad = [](cplx w) { return std::isnan(w.real()) || std::fabs(w.real()) >= 1.79769313486231570815e+308; };
  int rc = ALPSG_OK;
  for (int it = 0; it <= nt; ++it) {
    double kperp = h->kperp, kpar = h->kpar;
    scan_k(s1, it, kperp_last, kpar_last, kperp, kpar);
    if ((rc = move_k(h, s1->scan_type, kperp, kpar))) return rc;
    bool any = false;
    for (int r = 0; r < nroots; ++r) any = any || jump[r];
    if (!any) return fail(ALPSG_E_NOTFOUND, "all roots dropped during the scan (alps_error(9))");
    if ((rc = solve(omlast))) return rc;
    for (size_t t = 0; t < idx.size(); ++t) {
      const int r = idx[t];
      omlast[r] = wroots[r];
      if (bad(wroots[r])) jump[r] = 0;
      for (int imm = 0; imm < r; ++imm)
        if (std::abs(wroots[r] - wroots[imm]) < rp->D_gap) {
          wroots[r] = 0.0;
          jump[r] = 0;
        }
    }
    if (it % s1->n_res == 0) {
      const double kperpi = h->kperp, kpari = h->kpar;
      wroots = omlast;
      for (int it2 = 0; it2 <= nt2; ++it2) {
        kperp = h->kperp;
        kpar = h->kpar;
        scan_k(s2, it2, kperpi, kpari, kperp, kpar);
        if ((rc = move_k(h, s2->scan_type, kperp, kpar))) return rc;
        const bool out = (it2 % s2->n_res == 0);
        const int row = (it / s1->n_res) * nrow2 + it2 / s2->n_res;
        std::vector<cplx> from = wroots;
        if ((rc = solve(from))) return rc;
        if (out && (eigen || heat) && !idx.empty()) {
          const size_t m = idx.size();
          std::vector<cplx> e3(3 * m), b3(3 * m), us(3 * ns * m), dd(ns * m);
          std::vector<double> ps(ns * m), psp(4 * ns * m), ew(m);
          std::vector<cplx> w(m);
          for (size_t t = 0; t < m; ++t) w[t] = wroots[idx[t]];
          if ((rc = calc_eigen_batch(h, (int)m, w.data(), eigen != 0, heat != 0, e3.data(), b3.data(),
                                     eigen ? us.data() : nullptr, eigen ? dd.data() : nullptr,
                                     heat ? ps.data() : nullptr, heat ? psp.data() : nullptr,
                                     heat ? ew.data() : nullptr)))
            return rc;
          for (size_t t = 0; t < m; ++t) {
            const size_t o = (size_t)row * nroots + idx[t];
            for (int k = 0; k < 3; ++k) {
              reinterpret_cast<cplx *>(ef)[3 * o + k] = e3[3 * t + k];
              reinterpret_cast<cplx *>(bf)[3 * o + k] = b3[3 * t + k];
            }
            if (eigen) {
              for (int k = 0; k < 3 * ns; ++k) reinterpret_cast<cplx *>(Us)[3 * ns * o + k] = us[3 * ns * t + k];
              for (int k = 0; k < ns; ++k) reinterpret_cast<cplx *>(ds)[ns * o + k] = dd[ns * t + k];
            }
            if (heat) {
              for (int k = 0; k < ns; ++k) Ps[ns * o + k] = ps[ns * t + k];
              for (int k = 0; k < 4 * ns; ++k) Ps_split[4 * ns * o + k] = psp[4 * ns * t + k];
              W_EM[o] = ew[t];
            }
          }
        }
        for (size_t t = 0; t < idx.size(); ++t) {
          const int r = idx[t];
          if (bad(wroots[r])) jump[r] = 0;
          if (std::abs(Df[t]) > 1.e100) jump[r] = 0;  // infty check
          for (int imm = 0; imm < r; ++imm)
            if (std::abs(wroots[r] - wroots[imm]) < rp->D_gap) {
              wroots[r] = 0.0;
              jump[r] = 0;
            }
          if (out) {
            kperp_out[row] = h->kperp;
            kpar_out[row] = h->kpar;
            RO[(size_t)row * nroots + r] = wroots[r];
            written[(size_t)row * nroots + r] = 1;
          }
        }
      }
    }
    // the second scan's coordinates back to their start (itt = 0,
    // ALPS_fns.f90:3537-3584; a kcstq second scan resets with the k1_k2 form)
    kperp = h->kperp;
    kpar = h->kpar;
    if (s2->scan_type == 2) {
      alpsg_scan_params k12 = *s2;
      k12.scan_type = 0;
      scan_k(&k12, 0, kperp_last, kpar_last, kperp, kpar);
    } else {
      scan_k(s2, 0, kperp_last, kpar_last, kperp, kpar);
    }
    if ((rc = move_k(h, s2->scan_type, kperp, kpar))) return rc;
  }
  for (int r = 0; r < nroots; ++r) R[r] = omlast[r];
  return ALPSG_OK;
  ALPSG_CATCH
}
