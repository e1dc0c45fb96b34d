// Continuous-batching root server.
//
// Independent root solves -- each with its own plasma, grid size, wavevector,
// initial frequency and root controls -- arrive with alpsg_server_submit.  A
// service thread admits them into a fixed HBM arena as its space and the
// problem slots allow (first come first served, by each request's footprint),
// builds their tables on the device, and advances every admitted solve by one
// secant_osc step per device pass: the pending frequencies of all active solves
// go into one launch of k_sweep / k_resonant / k_assemble in their
// one-omega-per-workgroup form, where each workgroup reads its own solve's
// problem descriptor (pidx).  Solves of uneven nmax and nperp x npar therefore
// share launches.  A finished solve frees its arena blocks and slot at once,
// and waiting requests take them in the following pass.
//
// A plasma (the omega-independent block: f0, df0, fit tables) is uploaded once
// per residency and shared by every request that names it; registered plasmas
// stay resident between requests and are evicted, least recently used first,
// when the arena needs their space.
//
// Admission is pipelined with the passes and batched: in pass i every newly
// admitted request adds its max|J_n| jobs (determine_nmax,
// ALPS_fns.f90:3971-4075) to one job-list launch; after pass i's
// synchronisation the host picks nmax, and one job-list launch per table kind
// builds the per-k tables of all of them for pass i+1.  The launch shape of a
// pass is fixed and job-list launches run the single-problem device code per
// job, so a solve's arithmetic does not depend on the other solves: its result
// is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/alpsg.h"
#include "alpsg_analyt.h"
#include "alpsg_device.h"
#include "alpsg_guard.h"
#include "alpsg_lean.h"
#include "alpsg_osc.h"
#include "alpsg_relhost.h"

// shared with alpsg_api.cpp / alpsg_tables.hip / alpsg_disp.hip
int alpsg_internal_fail(int code, const std::string &msg);
hipError_t launch_derivative(hipStream_t st, int nspec, int nperp, int npar, const double *f0,
                             const double *pperp, const double *ppar, double *df1, double *df2);
hipError_t launch_int_ee(hipStream_t st, int nspec, int nperp, int npar, const double *pperp,
                         const double *ppar, const double *df1, const double *df2,
                         const double *qs, const double *ms, double *out);
int jobs_blocks_bessel_max(const TabJob &J);
int jobs_blocks_bessel(const TabJob &J);
int jobs_blocks_atab(const TabJob &J);
int jobs_blocks_moments(const TabJob &J);
hipError_t launch_bessel_max_jobs(hipStream_t st, const TabJob *jobs, const int *start, int njobs,
                                  int total, const int *mtab);
hipError_t launch_species_k_jobs(hipStream_t st, const TabJob *jobs, int njobs,
                                 const int *start_bes, int total_bes, const int *start_atab,
                                 int total_atab, const int *start_mom, int total_mom,
                                 int max_nperp, const int *mtab);
hipError_t launch_scatter_probs(hipStream_t st, const ProbDev *src, const int *slots, int n,
                                ProbDev *dst);
hipError_t launch_sweep_req(hipStream_t st, int nch, int W, int RC, int K, const ProbDev *prob,
                            int nom, const double2 *om, double2 *chiA, const int *pidx,
                            bool reduce);
hipError_t launch_resonant(hipStream_t st, int nch, int RB, const ProbDev *prob, int nspec, int nom,
                           const double2 *om, const int *active, double2 *chiB, double2 *lowB,
                           int lgen, const int *pidx, int P);
hipError_t launch_assemble(hipStream_t st, const ProbDev *prob, int nspec, int nom, int W, int RB,
                           const double2 *om, const int *active, const double2 *chiA,
                           const double2 *chiB, const double2 *lowA, const double2 *lowB,
                           const double2 *chiR, int RR, const double2 *lowR, double2 *det,
                           double2 *chi, double2 *low, double2 *wave, const int *pidx);
hipError_t launch_nhds(hipStream_t st, const ProbDev *prob, int nspec, int nom, const double2 *om,
                       const int *active, double2 *chiA, double2 *lowA, const int *pidx);
int nhds_tables(double z, int bm_nmax, double bessel_zero, std::vector<double2> &tab);
hipError_t launch_rel_bessel(hipStream_t st, int npts, int nmax, double zfac, const double *grel,
                             const double *pbrel, const int *mtab, double *out);
hipError_t launch_rel(hipStream_t st, const ProbDev *prob, int nspec, int nmaxR, int NB, int nom,
                      const double2 *om, const int *active, double2 *chiR, double2 *lowR,
                      int *flags, int *err, const int *pidx);
int rel_slots(int NB, int nmaxR);

namespace {

typedef std::complex<double> cplx;
using alpsg_osc::OscMachine;

#define SCHK(expr)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return alpsg_internal_fail(ALPSG_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int MTAB_S = 4096;
// the fixed launch shape of a pass: k_sweep_req with SW_NCH lanes per row pair,
// SW_W wavefronts per workgroup, SW_RC workgroups per omega and SW_K node records
// in flight per lane; k_resonant<64, *> over RS_RB harmonic blocks
#ifndef ALPSG_SERVE_SWEEP
#define ALPSG_SERVE_SWEEP 32, 4, 4, 2
#endif
constexpr int SW_SHAPE[4] = {ALPSG_SERVE_SWEEP};
constexpr int SW_NCH = SW_SHAPE[0], SW_W = SW_SHAPE[1], SW_RC = SW_SHAPE[2], SW_K = SW_SHAPE[3];
#ifndef ALPSG_SERVE_RB
#define ALPSG_SERVE_RB 8
#endif
constexpr int RS_RB = ALPSG_SERVE_RB;
constexpr int REL_NB = 4;  // harmonic blocks of k_rel_nonres in a pass (fixed, as every shape here)
constexpr size_t ALIGN = 256;
constexpr int MAX_JOBS = 8192;               // (request, species) jobs per table launch
constexpr size_t STAGE_CAP = (size_t)1 << 20;  // doubles read back per pass (int_ee, max|J_n|)

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

size_t align_up(size_t b) { return (b + ALIGN - 1) / ALIGN * ALIGN; }

// first-fit allocator over one device allocation
struct Arena {
  char *base = nullptr;
  size_t cap = 0, used = 0;
  std::map<size_t, size_t> free_;  // offset -> size
  void init(char *b, size_t c) {
    base = b;
    cap = c;
    used = 0;
    free_.clear();
    free_[0] = c;
  }
  size_t alloc(size_t sz) {  // offset, or SIZE_MAX when no free block is large enough
    sz = align_up(sz);
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= sz) {
        const size_t off = it->first, rest = it->second - sz;
        free_.erase(it);
        if (rest) free_[off + sz] = rest;
        used += sz;
        return off;
      }
    return SIZE_MAX;
  }
  void release(size_t off, size_t sz) {
    sz = align_up(sz);
    used -= sz;
    auto it = free_.emplace(off, sz).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

// bump layout of several arrays inside one arena block
struct Layout {
  size_t size = 0;
  size_t add(size_t bytes) {
    const size_t o = size;
    size += align_up(bytes);
    return o;
  }
};

// A plasma: host copy (immutable once built) and its device block.
struct Plasma {
  int id = 0;
  int nspec = 0, nperp = 0, npar = 0, P = 3, NRI = 100, kperp_norm = 1;
  double vA = 0, Tlim = 0.01, Bessel_zero = 1e-45;
  std::vector<double> qs, ms, ns, pperp, ppar;
  std::vector<int> acm, nfits, ftype, funi, fkap, porder, plog;
  std::vector<double> plogmax;
  std::vector<size_t> fit_off;
  int maxorder = 0;
  // Landau forms the species take: bit 0 a call of eval_fit, bit 1 the inlined
  // kappa evaluation (species whose every fit is a kappa fit)
  int lgen = 0;
  // closed-form (use_bM) species: the &bM_spec parameters (ALPS_NHDS.f90:163-168)
  std::vector<int> isbm, bm_cold, bm_nmax;
  std::vector<double> bm_Omega, bm_ell, bm_vth, bm_vd, bm_alpha, bm_bz;
  // relativistic species: (Gamma, pparbar) tables in the device block
  std::vector<int> isrel;
  int ngamma = 0, npparbar = 0;
  std::vector<double> rel_dg, rel_dpb, rel_ee, rel_gmin, rel_gmax;
  std::vector<int> rel_gmono;
  std::vector<size_t> o_grel, o_pbrel, o_f0rel, o_dfrel, o_rowlim, o_gampp;
  bool any_bm = false, any_rel = false;
  // device block: the uploaded prefix (host image img), then df1, df2
  std::vector<char> img;
  size_t o_pperp = 0, o_ppar = 0, o_f0 = 0, o_qs = 0, o_ms = 0, o_fit = 0, o_poly = 0, n_up = 0;
  size_t o_df1 = 0, o_df2 = 0, sz = 0;
  std::atomic<bool> registered{false};
  // service-thread state
  size_t off = SIZE_MAX;  // SIZE_MAX: not resident
  int ee_state = 0;       // 0 absent, 1 int_ee in this pass's readback, 2 int_ee on the host
  size_t ee_stage = 0;
  std::vector<double> int_ee;
  int users = 0;  // admitted requests reading the block
  long long last_use = 0;
  bool listed = false;  // in the server's resident list
};

struct Request {
  long long tag = 0;
  double kperp = 0, kpar = 0;
  cplx guess;
  alpsg_root_params rp{};
  double t_submit = 0;
  std::shared_ptr<Plasma> pl;
  int ntry = 0;
  int slot = -1, state = 0;  // 0 waiting, 1 sizing, 2 active
  size_t stage = 0;          // offset of its max|J_n| rows in the pass readback
  std::vector<double> bmax;  // [nspec][ntry]
  size_t off_k = SIZE_MAX, sz_k = 0;
  int nmax[ALPSG_MAXSPEC] = {0};
  std::vector<double2> bmi_h;  // closed-form species' Bessel-I tables (uploaded at admission)
  int nmaxR = -1;              // largest nmax of its relativistic species, -1 without one
  bool failed = false;         // its relativistic evaluation hit alps_error(8)
  OscMachine m;
};

struct Result {
  long long tag;
  cplx root;
  int iters, status;
  double latency_ms;
};

// one job-list launch's host and device buffers
struct JobBuf {
  TabJob *jobs_h = nullptr, *jobs_d = nullptr;
  int *start_h = nullptr, *start_d = nullptr;  // 3 arrays of MAX_JOBS + 1
  hipError_t init() {
    hipError_t e = hipHostMalloc((void **)&jobs_h, sizeof(TabJob) * MAX_JOBS, hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void **)&jobs_d, sizeof(TabJob) * MAX_JOBS);
    if (e == hipSuccess)
      e = hipHostMalloc((void **)&start_h, sizeof(int) * 3 * (MAX_JOBS + 1), hipHostMallocDefault);
    if (e == hipSuccess) e = hipMalloc((void **)&start_d, sizeof(int) * 3 * (MAX_JOBS + 1));
    return e;
  }
  void free_all() {
    if (jobs_h) (void)hipHostFree(jobs_h);
    if (start_h) (void)hipHostFree(start_h);
    if (jobs_d) (void)hipFree(jobs_d);
    if (start_d) (void)hipFree(start_d);
  }
};

}  // namespace

struct alpsg_server {
  hipStream_t st = nullptr;
  char *arena_mem = nullptr;
  Arena arena;
  int max_slots = 0;
  ProbDev *probs_d = nullptr;
  std::vector<int> free_slots;
  int *mtab_d = nullptr;
  // pass buffers
  double2 *om_d = nullptr, *det_d = nullptr, *chiA = nullptr, *chiB = nullptr;
  int *pidx_d = nullptr;
  double2 *om_h = nullptr, *det_h = nullptr;  // pinned
  int *pidx_h = nullptr;                      // pinned
  double *stage_h = nullptr, *stage_d = nullptr;  // per-pass readback of int_ee and max|J_n|
  JobBuf jb1, jb2;                                // stage-1 and stage-2 job lists
  ProbDev *newp_h = nullptr, *newp_d = nullptr;   // descriptors activated by a pass
  double *gtab_d = nullptr;                        // Gamma(m) and Fact(k) of the relativistic Landau term
  double2 *chiR = nullptr;                         // relativistic partial slots of a pass (grown on demand)
  size_t chiR_n = 0;
  int *flags_d = nullptr;                          // relativistic resonance flags of a pass
  size_t flags_n = 0;
  int *err_d = nullptr, *err_h = nullptr;          // alps_error(8) flags, one per omega of a pass
  int *newslot_h = nullptr, *newslot_d = nullptr;
  // plasmas and requests
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::map<int, std::shared_ptr<Plasma>> registry;  // registered plasmas by id
  int next_id = 1;
  std::vector<std::shared_ptr<Plasma>> resident;     // plasmas with a device block
  std::deque<std::unique_ptr<Request>> queue;        // submitted, not admitted
  std::vector<std::unique_ptr<Request>> live;        // admitted (sizing or active)
  std::deque<Result> results;
  long long submitted = 0, finished = 0;
  bool stop = false, failed = false, gc = false;
  std::string fail_msg;
  std::thread th;
  // statistics, written by the service thread only
  long long passes = 0, evals = 0, uploads = 0, evictions = 0;
  // algorithmic units of the evaluated omegas: node records of the sweep
  // (species, harmonic, ppar node), relativistic 2-D points x harmonics, and
  // relativistic Gamma rows x harmonics
  double work[3] = {0.0, 0.0, 0.0};
  // device time (ms) by kernel class, from events around each group's launches:
  // the node sweep, the resonant parts (and closed-form species), the
  // relativistic kernels, the assembly
  double kt[4] = {0.0, 0.0, 0.0, 0.0};
  int pass_groups = 0;
  static constexpr int NEV = 8;  // groups per pass with events (species counts 1..8)
  hipEvent_t ev[NEV][5] = {};
  // host time (ms): whole passes, admission work, disp launch enqueue, waits on the device
  double t_pass = 0, t_admit = 0, t_launch = 0, t_wait = 0;
  // the copy alpsg_server_stats reads: the service thread publishes it under
  // mu at the end of every pass, so a stats call never reads state that a
  // running pass is changing
  struct Snap {
    long long passes = 0, evals = 0, uploads = 0, evictions = 0, live = 0, arena_used = 0;
    double t_pass = 0, t_admit = 0, t_launch = 0, t_wait = 0;
    double work[3] = {0.0, 0.0, 0.0};
    double kt[4] = {0.0, 0.0, 0.0, 0.0};
  } snap;
};

namespace {

// Host copy of a plasma (tabulated, non-relativistic species with fits 1/2/6 or
// Chebyshev continuation), the checks of alpsg_create, and the host image of
// its device block.
int prepare_plasma(const alpsg_plasma_desc *d, Plasma &r) {
  if (!d) return alpsg_internal_fail(ALPSG_E_ARG, "null plasma description");
  const int nspec = d->nspec, nperp = d->nperp, npar = d->npar;
  if (nspec < 1 || nspec > ALPSG_MAXSPEC) return alpsg_internal_fail(ALPSG_E_ARG, "nspec out of range 1..8");
  if (!d->pp || !d->f0 || !d->ns || !d->qs || !d->ms || !d->ACmethod || !d->n_fits)
    return alpsg_internal_fail(ALPSG_E_ARG, "served plasmas need pp, f0, ns, qs, ms, ACmethod, n_fits");
  const int P = d->positions_principal;
  if (P < 1 || P > 16) return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "positions_principal must be in 1..16");
  if (nperp < 4 || npar < 2 * P + 8) return alpsg_internal_fail(ALPSG_E_ARG, "grid too small for positions_principal");
  if (nperp > 2048) return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "served plasmas need nperp <= 2048");
  if (d->n_resonance_interval < 2) return alpsg_internal_fail(ALPSG_E_ARG, "n_resonance_interval < 2");
  r.nspec = nspec;
  r.nperp = nperp;
  r.npar = npar;
  r.P = P;
  r.NRI = d->n_resonance_interval;
  r.kperp_norm = d->kperp_norm ? 1 : 0;
  r.vA = d->vA;
  r.Tlim = d->Tlim;
  r.Bessel_zero = d->Bessel_zero;
  r.qs.assign(d->qs, d->qs + nspec);
  r.ms.assign(d->ms, d->ms + nspec);
  r.ns.assign(d->ns, d->ns + nspec);
  auto PP = [&](int s, int i, int j, int c) {
    return d->pp[(size_t)s + (size_t)nspec * (i + (size_t)(nperp + 1) * (j + (size_t)(npar + 1) * c))];
  };
  r.pperp.resize((size_t)nspec * (nperp + 1));
  r.ppar.resize((size_t)nspec * (npar + 1));
  r.isbm.assign(nspec, 0);
  r.bm_cold.assign(nspec, 0);
  r.bm_nmax.assign(nspec, 0);
  r.bm_Omega.assign(nspec, 0.0);
  r.bm_ell.assign(nspec, 0.0);
  r.bm_vth.assign(nspec, 0.0);
  r.bm_vd.assign(nspec, 0.0);
  r.bm_alpha.assign(nspec, 0.0);
  r.bm_bz.assign(nspec, 0.0);
  r.isrel.assign(nspec, 0);
  r.rel_dg.assign(nspec, 0.0);
  r.rel_gmin.assign(nspec, 0.0);
  r.rel_gmono.assign(nspec, 0);
  r.rel_gmax.assign(nspec, 0.0);
  r.rel_dpb.assign(nspec, 0.0);
  r.rel_ee.assign(nspec, 0.0);
  r.o_grel.assign(nspec, 0);
  r.o_pbrel.assign(nspec, 0);
  r.o_f0rel.assign(nspec, 0);
  r.o_dfrel.assign(nspec, 0);
  r.o_rowlim.assign(nspec, 0);
  r.o_gampp.assign(nspec, 0);
  r.any_bm = r.any_rel = false;
  for (int s = 0; s < nspec; ++s) {
    if (d->qs[s] == 0.0 || d->ms[s] <= 0.0) return alpsg_internal_fail(ALPSG_E_ARG, "species charge/mass");
    for (int i = 0; i <= nperp; ++i) r.pperp[(size_t)s * (nperp + 1) + i] = PP(s, i, 1, 0);
    for (int j = 0; j <= npar; ++j) r.ppar[(size_t)s * (npar + 1) + j] = PP(s, 2, j, 1);
    if (d->usebM && d->usebM[s]) {
      // closed-form species: no table, fits or continuation are read (as alpsg_create)
      if (!d->bM_nmax || !d->bM_Bessel_zero || !d->bM_beta || !d->bM_alpha || !d->bM_pdrift)
        return alpsg_internal_fail(ALPSG_E_ARG, "missing bM_spec arrays for a use_bM species");
      if (d->bM_beta[s] < 0.0 || (d->bM_beta[s] > 0.0 && !(d->bM_alpha[s] > 0.0)) || d->ns[s] <= 0.0)
        return alpsg_internal_fail(ALPSG_E_ARG, "bM_spec parameters");
      if (d->bM_beta[s] == 0.0 && !d->kperp_norm)
        return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "cold-plasma species need kperp_norm (the reference stops)");
      r.isbm[s] = 1;
      r.any_bm = true;
      r.bm_cold[s] = d->bM_beta[s] == 0.0 ? 1 : 0;
      r.bm_nmax[s] = d->bM_nmax[s];
      r.bm_bz[s] = d->bM_Bessel_zero[s];
      r.bm_Omega[s] = d->qs[s] / d->ms[s];
      r.bm_ell[s] = std::sqrt(d->ms[s] / (d->ns[s] * d->qs[s] * d->qs[s]));
      r.bm_vth[s] = std::sqrt(d->bM_beta[s] / (d->ns[s] * d->ms[s]));
      r.bm_vd[s] = d->bM_pdrift[s] / d->ms[s];
      r.bm_alpha[s] = d->bM_alpha[s];
      continue;
    }
    if (d->relativistic && d->relativistic[s]) {
      // Gamma-grid species: Juettner fits in Gamma (types 4, 5) indexed by igamma
      if (!d->gamma_rel || !d->pparbar_rel || !d->f0_rel || !d->df0_rel)
        return alpsg_internal_fail(ALPSG_E_ARG, "missing relativistic tables");
      if (d->ngamma < 4 || d->npparbar < 2 * P + 8)
        return alpsg_internal_fail(ALPSG_E_ARG, "relativistic grid too small (ngamma >= 4, npparbar >= 2P+8)");
      if (d->ACmethod[s] != 1)
        return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "relativistic species need fit continuation (ACmethod 1)");
      if (d->n_fits[s] < 1 || d->n_fits[s] > ALPSG_MAXFITS || d->n_fits[s] > d->max_fits)
        return alpsg_internal_fail(ALPSG_E_ARG, "n_fits out of range");
      if (!d->fit_type || !d->perp_correction || !d->param_fit)
        return alpsg_internal_fail(ALPSG_E_ARG, "missing fit arrays");
      for (int f = 0; f < d->n_fits[s]; ++f) {
        const int ft = d->fit_type[s + (size_t)nspec * f];
        if (ft != 4 && ft != 5)
          return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "relativistic species need fit_type 4 or 5");
      }
      if (d->param_fit_nperp < std::max(nperp, d->ngamma))
        return alpsg_internal_fail(ALPSG_E_ARG, "param_fit_nperp < max(nperp, ngamma)");
      r.isrel[s] = 1;
      r.any_rel = true;
      r.ngamma = d->ngamma;
      r.npparbar = d->npparbar;
    }
    for (int j = 0; j <= npar; ++j)
      for (int i = 0; i <= nperp; ++i)
        if (PP(s, i, j, 0) != r.pperp[(size_t)s * (nperp + 1) + i] ||
            PP(s, i, j, 1) != r.ppar[(size_t)s * (npar + 1) + j])
          return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "momentum grid is not a tensor product of pperp and ppar");
    const double *pa = &r.ppar[(size_t)s * (npar + 1)];
    const double dpa = pa[2] - pa[1];
    if (!(dpa > 0)) return alpsg_internal_fail(ALPSG_E_ARG, "ppar grid must increase");
    for (int j = 0; j < npar; ++j)
      if (!(pa[j + 1] > pa[j]) || std::fabs((pa[j + 1] - pa[j]) - dpa) > 1e-6 * dpa)
        return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "ppar grid is not uniform");
  }
  const int npf = d->param_fit_nperp;
  auto prm = [&](int s, int i, int k, int f) {
    return d->param_fit[(size_t)s + (size_t)nspec * (i + (size_t)(npf + 1) * (k + 5 * (size_t)f))];
  };
  r.acm.assign(d->ACmethod, d->ACmethod + nspec);
  r.nfits.assign(nspec, 0);
  r.ftype.assign((size_t)nspec * ALPSG_MAXFITS, 0);
  r.funi.assign(nspec, 0);
  r.fkap.assign(nspec, 0);
  r.porder.assign(nspec, 0);
  r.plog.assign(nspec, 0);
  r.plogmax.assign(nspec, 0.0);
  r.fit_off.assign(nspec + 1, 0);
  r.maxorder = 0;
  r.lgen = 0;
  for (int s = 0; s < nspec; ++s) {
    if (r.isbm[s]) {  // no fit table (the closed form has its own continuation)
      r.fit_off[s + 1] = r.fit_off[s];
      continue;
    }
    if (r.isrel[s]) {  // fit rows over Gamma, in the Maxwell form of the table
      r.nfits[s] = d->n_fits[s];
      for (int f = 0; f < r.nfits[s]; ++f) r.ftype[(size_t)s * ALPSG_MAXFITS + f] = 1;
      r.funi[s] = 1;
      r.fit_off[s + 1] = r.fit_off[s] + (size_t)(r.ngamma + 1) * r.nfits[s] * FIT_STRIDE;
      continue;
    }
    if (r.acm[s] == 1) {
      const int nf = d->n_fits[s];
      if (nf < 1 || nf > ALPSG_MAXFITS || nf > d->max_fits) return alpsg_internal_fail(ALPSG_E_ARG, "n_fits out of range");
      if (!d->fit_type || !d->perp_correction || !d->param_fit) return alpsg_internal_fail(ALPSG_E_ARG, "missing fit arrays");
      if (npf < nperp) return alpsg_internal_fail(ALPSG_E_ARG, "param_fit_nperp < nperp");
      r.nfits[s] = nf;
      bool uni = true;
      for (int f = 0; f < nf; ++f) {
        const int ft = d->fit_type[s + (size_t)nspec * f];
        if (ft != 1 && ft != 2 && ft != 3 && ft != 6)
          return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "fit_type must be 1, 2, 3 or 6 (4 and 5 are relativistic)");
        r.ftype[(size_t)s * ALPSG_MAXFITS + f] = ft;
        // one Maxwellian whose (u2, u3) do not depend on pperp: the uniform Landau part
        if (ft != 1) uni = false;
        for (int k = 1; k <= 2 && uni; ++k)
          for (int i = 1; i <= nperp; ++i)
            if (prm(s, i, k, f) != prm(s, 0, k, f)) {
              uni = false;
              break;
            }
      }
      r.funi[s] = uni ? 1 : 0;
      // every fit a kappa fit whose exponent u4 stays in the lean power's domain
      bool kap = !uni && nf >= 1;
      for (int f = 0; f < nf; ++f) {
        if (r.ftype[(size_t)s * ALPSG_MAXFITS + f] != 2) kap = false;
        for (int i = 0; i <= nperp && kap; ++i)
          if (!(std::fabs(prm(s, i, 3, f)) <= ALPSG_LEAN_MAXC)) kap = false;
      }
      r.fkap[s] = kap ? 1 : 0;
      if (kap)
        r.lgen |= 2;
      else if (!(uni && nf == 1))
        r.lgen |= 1;
    } else if (r.acm[s] == 2) {
      if (!d->poly_kind || !d->poly_order || !d->logfit || !d->poly_log_max || !d->poly_fit_coeffs)
        return alpsg_internal_fail(ALPSG_E_ARG, "missing Chebyshev arrays");
      if (d->poly_kind[s] != 1) return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "poly_kind must be 1");
      if (d->poly_order[s] < 1 || d->poly_order[s] > d->poly_max_order) return alpsg_internal_fail(ALPSG_E_ARG, "poly_order");
      r.porder[s] = d->poly_order[s];
      r.plog[s] = d->logfit[s] ? 1 : 0;
      r.plogmax[s] = d->poly_log_max[s];
      r.maxorder = std::max(r.maxorder, r.porder[s]);
      r.lgen |= 1;
    } else if (r.acm[s] == 0) {
      // the pre-coded f0 of distribution_analyt (ALPS_analyt.f90:77-79)
      if (!distribution_analyt_defined(s + 1))
        return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "ACmethod 0: distribution_analyt defines no f0 for this species");
      r.lgen |= 1;
    } else {
      return alpsg_internal_fail(ALPSG_E_ARG, "ACmethod must be 0, 1 or 2");
    }
    r.fit_off[s + 1] = r.fit_off[s] + (r.acm[s] == 1 ? (size_t)(nperp + 1) * r.nfits[s] * FIT_STRIDE : 0);
  }
  // block layout: the uploaded prefix, then df0
  const size_t ng = (size_t)nspec * (nperp + 1) * (npar + 1);
  const size_t nfit = std::max<size_t>(r.fit_off[nspec], 1);
  const size_t npoly = (size_t)nspec * (nperp + 1) * (r.maxorder + 1);
  Layout L;
  r.o_pperp = L.add(sizeof(double) * r.pperp.size());
  r.o_ppar = L.add(sizeof(double) * r.ppar.size());
  r.o_f0 = L.add(sizeof(double) * ng);
  r.o_qs = L.add(sizeof(double) * nspec);
  r.o_ms = L.add(sizeof(double) * nspec);
  r.o_fit = L.add(sizeof(double) * nfit);
  r.o_poly = L.add(sizeof(double) * npoly);
  // relativistic species: (Gamma, pparbar) grid, f0, df0, row limits, Gamma at
  // the (pperp, ppar) points
  int rel_index[ALPSG_MAXSPEC];
  int nspec_rel = 0;
  for (int s = 0; s < nspec; ++s) rel_index[s] = (d->relativistic && d->relativistic[s]) ? nspec_rel++ : -1;
  std::vector<RelHost> rh(nspec);
  for (int s = 0; s < nspec; ++s) {
    if (!r.isrel[s]) continue;
    const int rc = rel_host_tables(d, s, rel_index[s], nspec_rel, &r.pperp[(size_t)s * (nperp + 1)],
                                   &r.ppar[(size_t)s * (npar + 1)], rh[s]);
    if (rc) return rc;
    const size_t np = rh[s].G.size();
    r.o_grel[s] = L.add(sizeof(double) * np);
    r.o_pbrel[s] = L.add(sizeof(double) * np);
    r.o_f0rel[s] = L.add(sizeof(double) * np);
    r.o_dfrel[s] = L.add(sizeof(double2) * np);
    r.o_rowlim[s] = L.add(sizeof(int2) * rh[s].lim.size());
    r.o_gampp[s] = L.add(sizeof(double) * rh[s].gpp.size());
    r.rel_dg[s] = rh[s].dg;
    r.rel_gmin[s] = rh[s].gmin;
    r.rel_gmono[s] = rh[s].gmono ? 1 : 0;
    r.rel_gmax[s] = rh[s].gmax;
    r.rel_dpb[s] = rh[s].dpb;
    r.rel_ee[s] = rh[s].ee;
  }
  r.n_up = L.size;
  r.o_df1 = L.add(sizeof(double) * ng);
  r.o_df2 = L.add(sizeof(double) * ng);
  r.sz = L.size;
  r.img.assign(r.n_up, 0);
  auto H = [&](size_t o) { return reinterpret_cast<double *>(r.img.data() + o); };
  std::memcpy(H(r.o_pperp), r.pperp.data(), sizeof(double) * r.pperp.size());
  std::memcpy(H(r.o_ppar), r.ppar.data(), sizeof(double) * r.ppar.size());
  std::memcpy(H(r.o_qs), r.qs.data(), sizeof(double) * nspec);
  std::memcpy(H(r.o_ms), r.ms.data(), sizeof(double) * nspec);
  double *f0 = H(r.o_f0);
  for (int s = 0; s < nspec; ++s)
    for (int j = 0; j <= npar; ++j)
      for (int i = 0; i <= nperp; ++i)
        f0[((size_t)s * (nperp + 1) + i) * (npar + 1) + j] =
            d->f0[(size_t)s + (size_t)nspec * (i + (size_t)(nperp + 1) * j)];
  // fit_function coefficients (ALPS_analyt.f90:185-266) with the pperp factor
  // folded in, as alpsg_create tabulates them
  double *fitc = H(r.o_fit);
  double *polyc = H(r.o_poly);
  for (int s = 0; s < nspec; ++s) {
    if (!r.isrel[s]) continue;
    const RelHost &R = rh[s];
    std::memcpy(r.img.data() + r.o_grel[s], R.G.data(), sizeof(double) * R.G.size());
    std::memcpy(r.img.data() + r.o_pbrel[s], R.PB.data(), sizeof(double) * R.PB.size());
    std::memcpy(r.img.data() + r.o_f0rel[s], R.F.data(), sizeof(double) * R.F.size());
    std::memcpy(r.img.data() + r.o_dfrel[s], R.DF.data(), sizeof(double2) * R.DF.size());
    std::memcpy(r.img.data() + r.o_rowlim[s], R.lim.data(), sizeof(int2) * R.lim.size());
    std::memcpy(r.img.data() + r.o_gampp[s], R.gpp.data(), sizeof(double) * R.gpp.size());
    rel_fit_rows(d, s, rel_index[s], nspec_rel, &fitc[r.fit_off[s]]);
  }
  for (int s = 0; s < nspec; ++s) {
    if (r.isbm[s] || r.isrel[s]) continue;
    if (r.acm[s] == 1) {
      for (int i = 0; i <= nperp; ++i) {
        const double pv = r.pperp[(size_t)s * (nperp + 1) + i];
        for (int f = 0; f < r.nfits[s]; ++f) {
          const double pc = d->perp_correction[s + (size_t)nspec * f];
          const int ft = r.ftype[(size_t)s * ALPSG_MAXFITS + f];
          double *c = &fitc[r.fit_off[s] + ((size_t)i * r.nfits[s] + f) * FIT_STRIDE];
          c[1] = prm(s, i, 1, f);
          c[2] = prm(s, i, 2, f);
          if (ft == 1) {
            c[0] = prm(s, i, 0, f) * std::exp(-pc * (pv * pv));
          } else if (ft == 2) {
            c[0] = prm(s, i, 0, f);
            c[3] = prm(s, i, 3, f);
            c[4] = pc * prm(s, i, 4, f) * (pv * pv);
          } else if (ft == 3) {
            // Juettner in (pperp, ppar), ALPS_analyt.f90:238-242
            c[0] = prm(s, i, 0, f);
            c[3] = pv * pv;
            c[4] = d->vA * d->vA / (d->ms[s] * d->ms[s]);
          } else {
            c[0] = prm(s, i, 0, f);
            c[3] = prm(s, i, 3, f) * pc * (pv * pv);
          }
        }
      }
    } else if (r.acm[s] == 2) {
      for (int i = 0; i <= nperp; ++i)
        for (int n = 0; n <= r.porder[s]; ++n)
          polyc[((size_t)s * (nperp + 1) + i) * (r.maxorder + 1) + n] =
              d->poly_fit_coeffs[(size_t)s + (size_t)nspec * (i + (size_t)(nperp + 1) * n)];
    }
  }
  return ALPSG_OK;
}

// determine_nmax's search bound (ALPS_fns.f90:3971-4075), as alpsg_set_k
// bounds it: |J_n(x)| <= (x/2)^n / n!
int ntry_bound(const Plasma &P, double kperp) {
  double xmax = 0.0;
  for (int s = 0; s < P.nspec; ++s)
    for (int i = 0; i <= P.nperp; ++i)
      xmax = std::max(xmax, std::fabs(kperp * P.pperp[(size_t)s * (P.nperp + 1) + i] / P.qs[s]));
  int ntry = 1;
  const double lz = std::log(std::max(P.Bessel_zero, 1e-300)) - std::log(1e6);
  while (ntry < MTAB_S - 16) {
    const double lb = ntry * std::log(std::max(xmax, 1e-300) / 2.0) - std::lgamma(ntry + 1.0);
    if (lb < lz && ntry > xmax) break;
    ++ntry;
  }
  return std::min(ntry + 8, MTAB_S - 4);
}

void evict(alpsg_server *S, Plasma &P) {
  S->arena.release(P.off, P.sz);
  P.off = SIZE_MAX;
  P.ee_state = 0;
  S->evictions += 1;
}

// arena block, evicting resident plasmas no admitted request reads (least
// recently used first) until it fits; SIZE_MAX if it cannot
size_t alloc_evict(alpsg_server *S, size_t sz) {
  size_t off = S->arena.alloc(sz);
  while (off == SIZE_MAX) {
    Plasma *lru = nullptr;
    for (auto &p : S->resident)
      if (p->off != SIZE_MAX && p->users == 0 && p->ee_state != 1 &&
          (lru == nullptr || p->last_use < lru->last_use))
        lru = p.get();
    if (!lru) return SIZE_MAX;
    evict(S, *lru);
    off = S->arena.alloc(sz);
  }
  return off;
}

void drop_nonresident(alpsg_server *S) {
  S->resident.erase(std::remove_if(S->resident.begin(), S->resident.end(),
                                   [](const std::shared_ptr<Plasma> &p) {
                                     if (p->off != SIZE_MAX) return false;
                                     p->listed = false;
                                     return true;
                                   }),
                    S->resident.end());
}

// a plasma's device block: upload, df0 and int_ee (read back with the pass);
// returns 1 when the arena has no room
int upload_plasma(alpsg_server *S, const std::shared_ptr<Plasma> &pp, size_t &stage_used) {
  Plasma &P = *pp;
  if (stage_used + P.nspec > STAGE_CAP) return 1;
  const size_t off = alloc_evict(S, P.sz);
  if (off == SIZE_MAX) return 1;
  P.off = off;
  char *b = S->arena.base + off;
  auto D = [&](size_t o) { return reinterpret_cast<double *>(b + o); };
  SCHK(hipMemcpyAsync(b, P.img.data(), P.n_up, hipMemcpyHostToDevice, S->st));
  SCHK(hipMemsetAsync(D(P.o_df1), 0, P.sz - P.o_df1, S->st));
  SCHK(launch_derivative(S->st, P.nspec, P.nperp, P.npar, D(P.o_f0), D(P.o_pperp), D(P.o_ppar),
                         D(P.o_df1), D(P.o_df2)));
  SCHK(launch_int_ee(S->st, P.nspec, P.nperp, P.npar, D(P.o_pperp), D(P.o_ppar), D(P.o_df1),
                     D(P.o_df2), D(P.o_qs), D(P.o_ms), S->stage_d + stage_used));
  P.ee_stage = stage_used;
  stage_used += P.nspec;
  P.ee_state = 1;
  S->uploads += 1;
  if (!P.listed) {
    P.listed = true;
    S->resident.push_back(pp);
  }
  return 0;
}

// fill a prefix-sum start array of one job kind; returns the block total
int fill_start(const TabJob *jobs, int n, int *start, int (*blocks)(const TabJob &)) {
  long long t = 0;
  for (int j = 0; j < n; ++j) {
    start[j] = (int)t;
    t += blocks(jobs[j]);
  }
  start[n] = (int)t;
  return (int)t;
}

// per-k block of an admitted request: nmax from max|J_n| (determine_nmax),
// bes/atab/nodes jobs and its problem descriptor; returns 1 when the arena
// has no room, ALPSG_E_ARG when the Bessel cutoff is not reached
int admit_k(alpsg_server *S, Request &r, int &njobs, int &max_nperp, int &nnew) {
  const Plasma &P = *r.pl;
  const int ns = P.nspec, nperp = P.nperp, npar = P.npar;
  if (njobs + ns > MAX_JOBS) return 1;
  for (int s = 0; s < ns; ++s) {
    if (P.isbm[s]) {  // determine_nmax, ALPS_fns.f90:4012-4013
      r.nmax[s] = 1;
      continue;
    }
    int n = 0;
    for (int t = 1; t <= r.ntry; ++t)
      if (!(r.bmax[(size_t)s * r.ntry + (t - 1)] > P.Bessel_zero)) {
        n = t;
        break;
      }
    if (n == 0) return alpsg_internal_fail(ALPSG_E_ARG, "Bessel cutoff not reached (kperp too large?)");
    r.nmax[s] = n;
  }
  // closed-form species: the omega-independent part of calc_chi
  // (ALPS_NHDS.f90:163-185) and its Bessel-I table at this kperp
  double bm_z[ALPSG_MAXSPEC] = {0}, bm_zp[ALPSG_MAXSPEC] = {0};
  int bm_nrun[ALPSG_MAXSPEC] = {0};
  size_t bm_off[ALPSG_MAXSPEC] = {0};
  r.bmi_h.clear();
  for (int s = 0; s < ns; ++s) {
    if (!P.isbm[s] || P.bm_cold[s]) continue;
    const double a = r.kperp * P.bm_vth[s] / P.bm_Omega[s];
    bm_z[s] = ((0.5 * a) * a) * P.bm_alpha[s];
    const double b = P.bm_vth[s] / P.bm_Omega[s];
    bm_zp[s] = ((0.5 * b) * b) * P.bm_alpha[s];
    std::vector<double2> tab;
    bm_nrun[s] = nhds_tables(bm_z[s], P.bm_nmax[s], P.bm_bz[s], tab);
    bm_off[s] = r.bmi_h.size();
    r.bmi_h.insert(r.bmi_h.end(), tab.begin(), tab.end());
  }
  const int npts_rel = (P.ngamma + 1) * (P.npparbar + 1);
  Layout L;
  size_t ob[ALPSG_MAXSPEC], oa[ALPSG_MAXSPEC], on[ALPSG_MAXSPEC], orb[ALPSG_MAXSPEC], obm[ALPSG_MAXSPEC];
  for (int s = 0; s < ns; ++s) {
    const int nm = r.nmax[s];
    ob[s] = oa[s] = on[s] = orb[s] = obm[s] = 0;
    if (P.isbm[s]) {
      if (!P.bm_cold[s]) obm[s] = L.add(sizeof(double2) * (size_t)(2 * bm_nrun[s] + 1));
      continue;
    }
    ob[s] = L.add(sizeof(double) * (size_t)(nm + 3) * (nperp + 1));
    oa[s] = L.add(sizeof(double) * (size_t)(nm + 1) * (nperp + 1) * ATAB_STRIDE);
    on[s] = L.add(sizeof(double) * (size_t)(nm + 1) * (npar - 1) * NODE_STRIDE);
    if (P.isrel[s]) orb[s] = L.add(sizeof(double) * (size_t)(nm + 2) * npts_rel);
  }
  const size_t off = alloc_evict(S, L.size);
  if (off == SIZE_MAX) return 1;
  r.off_k = off;
  r.sz_k = L.size;
  char *bs = S->arena.base + P.off, *bk = S->arena.base + off;
  auto PS = [&](size_t o) { return reinterpret_cast<double *>(bs + o); };
  auto PK = [&](size_t o) { return reinterpret_cast<double *>(bk + o); };
  ProbDev &pd = S->newp_h[nnew];
  std::memset(&pd, 0, sizeof(pd));
  pd.nspec = ns;
  pd.kperp_norm = P.kperp_norm;
  pd.P = P.P;
  pd.NRI = P.NRI;
  pd.vA = P.vA;
  pd.Tlim = P.Tlim;
  pd.kperp = r.kperp;
  pd.kpar = r.kpar;
  pd.gtab = S->gtab_d;
  pd.nmaxR = 0;
  for (int s = 0; s < ns; ++s) {
    SpecDev &sd = pd.sp[s];
    const double *pe = &P.pperp[(size_t)s * (nperp + 1)];
    const double *pa = &P.ppar[(size_t)s * (npar + 1)];
    const int nm = r.nmax[s];
    double *pperp_d = PS(P.o_pperp) + (size_t)s * (nperp + 1);
    double *ppar_d = PS(P.o_ppar) + (size_t)s * (npar + 1);
    if (P.isbm[s]) {
      // closed form (ALPS_NHDS.f90:163-168): no table terms, no ee term
      sd.nperp = nperp;
      sd.npar = npar;
      sd.q = P.qs[s];
      sd.m = P.ms[s];
      sd.dens = P.ns[s];
      sd.pperp = pperp_d;
      sd.ppar = ppar_d;
      sd.usebm = 1;
      sd.bm_cold = P.bm_cold[s];
      sd.bm_Omega = P.bm_Omega[s];
      sd.bm_ell = P.bm_ell[s];
      sd.bm_vth = P.bm_vth[s];
      sd.bm_vd = P.bm_vd[s];
      sd.bm_alpha = P.bm_alpha[s];
      sd.int_ee = 0.0;
      sd.acmethod = 1;
      sd.nfits = 1;
      sd.fit_type[0] = 1;
      sd.fit_uniform = 1;
      sd.nmax = -1;
      if (!P.bm_cold[s]) {
        sd.bm_z = bm_z[s];
        sd.bm_zp = bm_zp[s];
        sd.bm_nrun = bm_nrun[s];
        double2 *t = reinterpret_cast<double2 *>(bk + obm[s]);
        SCHK(hipMemcpyAsync(t, r.bmi_h.data() + bm_off[s], sizeof(double2) * (size_t)(2 * bm_nrun[s] + 1),
                            hipMemcpyHostToDevice, S->st));
        sd.bmi = t;
      }
      continue;
    }
    TabJob &J = S->jb2.jobs_h[njobs++];
    J.pperp = pperp_d;
    J.ppar = ppar_d;
    J.df1 = PS(P.o_df1) + (size_t)s * (nperp + 1) * (npar + 1);
    J.df2 = PS(P.o_df2) + (size_t)s * (nperp + 1) * (npar + 1);
    J.bes = PK(ob[s]);
    J.atab = PK(oa[s]);
    J.nodes = PK(on[s]);
    J.bmax = nullptr;
    J.kperp = r.kperp;
    J.kpar = r.kpar;
    J.q = P.qs[s];
    J.m = P.ms[s];
    J.nperp = nperp;
    J.npar = npar;
    J.nmax = nm;
    J.ntry = 0;
    max_nperp = std::max(max_nperp, nperp);
    sd.nperp = nperp;
    sd.npar = npar;
    sd.nmax = nm;
    sd.acmethod = P.acm[s];
    sd.anum = s + 1;
    sd.nfits = P.acm[s] == 1 ? P.nfits[s] : 0;
    for (int f = 0; f < sd.nfits; ++f) sd.fit_type[f] = P.ftype[(size_t)s * ALPSG_MAXFITS + f];
    sd.fit_uniform = P.funi[s];
    sd.fit_kappa = P.fkap[s];
    sd.q = P.qs[s];
    sd.m = P.ms[s];
    sd.dens = P.ns[s];
    sd.dpperp = pe[2] - pe[1];
    sd.dppar = pa[2] - pa[1];
    sd.dppar_abs = std::fabs(pa[2] - pa[1]);
    sd.z = P.kperp_norm ? r.kperp / P.qs[s] : 1.0 / P.qs[s];
    sd.int_ee = P.int_ee[s];
    sd.pperp = pperp_d;
    sd.ppar = ppar_d;
    sd.atab = PK(oa[s]);
    sd.nodes = PK(on[s]);
    sd.fitc = PS(P.o_fit) + P.fit_off[s];
    if (P.acm[s] == 2 && !P.isrel[s]) {
      sd.poly_kind = 1;
      sd.poly_order = P.porder[s];
      sd.poly_logfit = P.plog[s];
      sd.poly_log_max = P.plogmax[s];
      sd.poly_norm1 = 5.e-1 * (pa[npar] + pa[0]);
      sd.poly_norm2 = 5.e-1 * (pa[npar] - pa[0]);
      sd.polyc = PS(P.o_poly) + (size_t)s * (nperp + 1) * (P.maxorder + 1);
    }
    if (P.isrel[s]) {
      // the harmonics run in k_rel_nonres / k_rel_res, not in the sweep
      const double m = P.ms[s], q = P.qs[s], vA = P.vA;
      sd.rel = 1;
      sd.ng = P.ngamma;
      sd.npb = P.npparbar;
      sd.grel = PS(P.o_grel[s]);
      sd.pbrel = PS(P.o_pbrel[s]);
      sd.f0rel = PS(P.o_f0rel[s]);
      sd.dfrel = reinterpret_cast<const double2 *>(bs + P.o_dfrel[s]);
      sd.rowlim = reinterpret_cast<const int2 *>(bs + P.o_rowlim[s]);
      sd.gam_pp = PS(P.o_gampp[s]);
      sd.df1 = PS(P.o_df1) + (size_t)s * (nperp + 1) * (npar + 1);
      sd.df2 = PS(P.o_df2) + (size_t)s * (nperp + 1) * (npar + 1);
      sd.rel_dg = P.rel_dg[s];
      sd.rel_gmin = P.rel_gmin[s];
      sd.rel_gmono = P.rel_gmono[s];
      sd.rel_gmax = P.rel_gmax[s];
      sd.rel_dpb = P.rel_dpb[s];
      sd.fit_uniform = 1;
      sd.nmax_rel = nm;
      sd.nmax = -1;
      sd.rel_pref = ((2.0 * 3.141592653589793238462643383279502884) * ((m / vA) * (m / vA) * (m / vA))) *
                    (q * vA / (r.kpar * m));
      sd.rel_zfac = r.kperp * m / (vA * q);
      sd.rel_zbar = P.kperp_norm ? r.kperp * m / (vA * q) : m / (vA * q);
      double *rb = PK(orb[s]);
      SCHK(launch_rel_bessel(S->st, npts_rel, nm, sd.rel_zfac, sd.grel, sd.pbrel, S->mtab_d, rb));
      sd.relbes = rb;
      pd.nmaxR = std::max(pd.nmaxR, nm);
    }
  }
  r.nmaxR = P.any_rel ? pd.nmaxR : -1;
  S->newslot_h[nnew++] = r.slot;
  r.m.start(alpsg_osc::zmk(r.guess.real(), r.guess.imag()), r.rp.numiter, r.rp.D_threshold, r.rp.D_prec);
  r.state = 2;
  return 0;
}

void release(alpsg_server *S, Request &r) {
  if (r.off_k != SIZE_MAX) S->arena.release(r.off_k, r.sz_k);
  r.off_k = SIZE_MAX;
  if (r.state >= 1 && r.pl) r.pl->users -= 1;
  if (r.slot >= 0) S->free_slots.push_back(r.slot);
  r.slot = -1;
}

// the end of a pass: its results and a snapshot of the statistics, under mu
void post_results(alpsg_server *S, std::vector<Result> &done) {
  std::lock_guard<std::mutex> lk(S->mu);
  alpsg_server::Snap &sn = S->snap;
  sn.passes = S->passes;
  sn.evals = S->evals;
  sn.uploads = S->uploads;
  sn.evictions = S->evictions;
  sn.live = (long long)S->live.size();
  sn.arena_used = (long long)S->arena.used;
  sn.t_pass = S->t_pass;
  sn.t_admit = S->t_admit;
  sn.t_launch = S->t_launch;
  sn.t_wait = S->t_wait;
  for (int k = 0; k < 3; ++k) sn.work[k] = S->work[k];
  for (int k = 0; k < 4; ++k) sn.kt[k] = S->kt[k];
  if (done.empty()) return;
  for (auto &d : done) S->results.push_back(d);
  S->finished += (long long)done.size();
  S->cv_done.notify_all();
  done.clear();
}

// one pass: admissions, the batched disp of every active solve, results
int server_pass(alpsg_server *S) {
  const double tp0 = now_ms();
  std::vector<Result> done;
  // stage 1 for waiting requests, first come first served, while slots, job
  // room, readback room and arena space last
  std::vector<Request *> sizing_now;
  std::vector<Plasma *> uploaded_now;
  size_t stage_used = 0;
  int nj1 = 0;
  for (;;) {
    Request *r = nullptr;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (S->queue.empty()) break;
      r = S->queue.front().get();
    }
    Plasma &P = *r->pl;
    const size_t need = (size_t)P.nspec * r->ntry;
    if (S->free_slots.empty() || nj1 + P.nspec > MAX_JOBS || stage_used + P.nspec + need > STAGE_CAP)
      break;
    int rc = 0;
    if (P.off == SIZE_MAX) {
      rc = upload_plasma(S, r->pl, stage_used);
      if (rc > 1) return rc;
      if (rc == 0) uploaded_now.push_back(&P);
    }
    if (rc == 1) {
      if (!S->live.empty() || !sizing_now.empty()) break;  // space frees as solves finish
      // nothing admitted can free space: the request does not fit the arena
      std::unique_ptr<Request> own;
      {
        std::lock_guard<std::mutex> lk(S->mu);
        own = std::move(S->queue.front());
        S->queue.pop_front();
      }
      done.push_back(Result{own->tag, cplx(0.0, 0.0), 0, ALPSG_E_ARG, now_ms() - own->t_submit});
      continue;
    }
    P.users += 1;
    P.last_use = S->passes;
    r->slot = S->free_slots.back();
    S->free_slots.pop_back();
    r->stage = stage_used;
    stage_used += need;
    const double *pe = reinterpret_cast<const double *>(S->arena.base + P.off + P.o_pperp);
    for (int s = 0; s < P.nspec; ++s) {
      TabJob &J = S->jb1.jobs_h[nj1++];
      std::memset(&J, 0, sizeof(J));
      J.pperp = pe + (size_t)s * (P.nperp + 1);
      J.bmax = S->stage_d + r->stage + (size_t)s * r->ntry;
      J.kperp = r->kperp;
      J.q = P.qs[s];
      J.nperp = P.nperp;
      J.ntry = r->ntry;
    }
    r->state = 1;
    sizing_now.push_back(r);
    std::lock_guard<std::mutex> lk(S->mu);
    S->live.push_back(std::move(S->queue.front()));
    S->queue.pop_front();
  }
  if (nj1 > 0) {
    const int total = fill_start(S->jb1.jobs_h, nj1, S->jb1.start_h, jobs_blocks_bessel_max);
    SCHK(hipMemcpyAsync(S->jb1.jobs_d, S->jb1.jobs_h, sizeof(TabJob) * nj1, hipMemcpyHostToDevice, S->st));
    SCHK(hipMemcpyAsync(S->jb1.start_d, S->jb1.start_h, sizeof(int) * (nj1 + 1), hipMemcpyHostToDevice, S->st));
    SCHK(launch_bessel_max_jobs(S->st, S->jb1.jobs_d, S->jb1.start_d, nj1, total, S->mtab_d));
  }
  if (stage_used > 0)
    SCHK(hipMemcpyAsync(S->stage_h, S->stage_d, sizeof(double) * stage_used, hipMemcpyDeviceToHost, S->st));
  // the batch: one pending omega per active solve, grouped by species count
  std::vector<Request *> act;
  for (auto &p : S->live)
    if (p->state == 2 && !p->m.done && !p->failed) act.push_back(p.get());
  std::stable_sort(act.begin(), act.end(), [](Request *a, Request *b) { return a->pl->nspec < b->pl->nspec; });
  const int nb = (int)act.size();
  const double t0 = now_ms();
  S->t_admit += t0 - tp0;
  for (int i = 0; i < nb; ++i) {
    const alpsg_osc::zc wz = act[i]->m.want();
    const cplx w(wz.re, wz.im);
    S->om_h[i] = make_double2(w.real(), w.imag());
    S->pidx_h[i] = act[i]->slot;
    const Plasma &P = *act[i]->pl;
    for (int s = 0; s < P.nspec; ++s) {
      if (P.isbm[s]) continue;
      const double rows = 2.0 * act[i]->nmax[s] + 1.0;
      if (P.isrel[s]) {
        S->work[1] += rows * (P.nperp - 1) * (P.npar - 1);
        S->work[2] += rows * (P.ngamma - 1);
      } else {
        S->work[0] += rows * (P.npar - 1);
      }
    }
  }
  bool pass_rel = false;
  if (nb > 0) {
    // relativistic partial slots and flags: sized for the largest group first
    size_t need_r = 0, need_f = 0;
    for (int i0 = 0, i1 = 0; i0 < nb; i0 = i1) {
      const int ns = act[i0]->pl->nspec;
      int relmax = -1;
      for (i1 = i0; i1 < nb && act[i1]->pl->nspec == ns; ++i1) relmax = std::max(relmax, act[i1]->nmaxR);
      if (relmax < 0) continue;
      const size_t n = (size_t)(i1 - i0);
      need_r = std::max(need_r, n * rel_slots(REL_NB, relmax) * ns * 6);
      need_f = std::max(need_f, n * ns * (size_t)(2 * relmax + 1));
      pass_rel = true;
    }
    if (need_r > S->chiR_n || need_f > S->flags_n) {
      SCHK(hipStreamSynchronize(S->st));
      if (need_r > S->chiR_n) {
        if (S->chiR) (void)hipFree(S->chiR);
        S->chiR = nullptr;
        S->chiR_n = 0;
        SCHK(hipMalloc((void **)&S->chiR, sizeof(double2) * need_r));
        S->chiR_n = need_r;
      }
      if (need_f > S->flags_n) {
        if (S->flags_d) (void)hipFree(S->flags_d);
        S->flags_d = nullptr;
        S->flags_n = 0;
        SCHK(hipMalloc((void **)&S->flags_d, sizeof(int) * need_f));
        S->flags_n = need_f;
      }
    }
    SCHK(hipMemcpyAsync(S->om_d, S->om_h, sizeof(double2) * nb, hipMemcpyHostToDevice, S->st));
    SCHK(hipMemcpyAsync(S->pidx_d, S->pidx_h, sizeof(int) * nb, hipMemcpyHostToDevice, S->st));
    if (pass_rel) SCHK(hipMemsetAsync(S->err_d, 0, sizeof(int) * nb, S->st));
    int i0 = 0, ng = 0;
    while (i0 < nb) {
      int i1 = i0;
      const int ns = act[i0]->pl->nspec;
      hipEvent_t *ev = ng < alpsg_server::NEV ? S->ev[ng] : nullptr;
      ++ng;
      if (ev) SCHK(hipEventRecord(ev[0], S->st));
      int gen = 0;
      bool bm = false;
      int relmax = -1;
      for (; i1 < nb && act[i1]->pl->nspec == ns; ++i1) {
        gen |= act[i1]->pl->lgen;
        if (act[i1]->pl->any_bm) bm = true;
        relmax = std::max(relmax, act[i1]->nmaxR);
      }
      const int n = i1 - i0;
      SCHK(launch_sweep_req(S->st, SW_NCH, SW_W, SW_RC, SW_K, S->probs_d, n, S->om_d + i0,
                            S->chiA, S->pidx_d + i0, false));
      if (ev) SCHK(hipEventRecord(ev[1], S->st));
      SCHK(launch_resonant(S->st, 64, RS_RB, S->probs_d, ns, n, S->om_d + i0, nullptr, S->chiB,
                           nullptr, gen, S->pidx_d + i0, 0));
      // closed-form species: partial slot 0 of their chi (the sweep left it zero)
      if (bm)
        SCHK(launch_nhds(S->st, S->probs_d, ns, n, S->om_d + i0, nullptr, S->chiA, nullptr,
                         S->pidx_d + i0));
      // relativistic species: REL_NB + 2 relmax + 1 partial slots; an omega whose
      // problem has fewer harmonics gets zeros in the slots past its own
      int RR = 0;
      if (ev) SCHK(hipEventRecord(ev[2], S->st));
      if (relmax >= 0) {
        RR = rel_slots(REL_NB, relmax);
        SCHK(launch_rel(S->st, S->probs_d, ns, relmax, REL_NB, n, S->om_d + i0, nullptr, S->chiR,
                        nullptr, S->flags_d, S->err_d + i0, S->pidx_d + i0));
      }
      if (ev) SCHK(hipEventRecord(ev[3], S->st));
      SCHK(launch_assemble(S->st, S->probs_d, ns, n, SW_W * SW_RC, RS_RB, S->om_d + i0, nullptr,
                           S->chiA, S->chiB, nullptr, nullptr, RR ? S->chiR : nullptr, RR, nullptr,
                           S->det_d + i0, nullptr, nullptr, nullptr, S->pidx_d + i0));
      if (ev) SCHK(hipEventRecord(ev[4], S->st));
      i0 = i1;
    }
    S->pass_groups = std::min(ng, (int)alpsg_server::NEV);
    SCHK(hipMemcpyAsync(S->det_h, S->det_d, sizeof(double2) * nb, hipMemcpyDeviceToHost, S->st));
    if (pass_rel) SCHK(hipMemcpyAsync(S->err_h, S->err_d, sizeof(int) * nb, hipMemcpyDeviceToHost, S->st));
  }
  const double t1 = now_ms();
  SCHK(hipStreamSynchronize(S->st));
  const double t2 = now_ms();
  S->t_launch += t1 - t0;
  S->t_wait += t2 - t1;
  for (int g = 0; g < (nb > 0 ? S->pass_groups : 0); ++g)
    for (int k = 0; k < 4; ++k) {
      float ms = 0.0f;
      SCHK(hipEventElapsedTime(&ms, S->ev[g][k], S->ev[g][k + 1]));
      S->kt[k] += ms;
    }
  S->passes += 1;
  S->evals += nb;
  // the readback of this pass: int_ee of the uploaded plasmas, max|J_n| of the admissions
  for (Plasma *P : uploaded_now) {
    P->int_ee.assign(S->stage_h + P->ee_stage, S->stage_h + P->ee_stage + P->nspec);
    for (int s = 0; s < P->nspec; ++s) {
      if (P->isbm[s]) P->int_ee[s] = 0.0;            // the closed form has no ee term
      if (P->isrel[s]) P->int_ee[s] = P->rel_ee[s];  // int_ee_rel on the (Gamma, pparbar) grid
    }
    P->ee_state = 2;
  }
  for (Request *r : sizing_now)
    r->bmax.assign(S->stage_h + r->stage, S->stage_h + r->stage + (size_t)r->pl->nspec * r->ntry);
  // stage 2 for every sized request, oldest first, while job room and space last
  const double t3 = now_ms();
  int nj2 = 0, max_nperp = 4, nnew = 0;
  for (auto it = S->live.begin(); it != S->live.end();) {
    Request &r = **it;
    if (r.state != 1) {
      ++it;
      continue;
    }
    const int rc = admit_k(S, r, nj2, max_nperp, nnew);
    bool drop = false;
    if (rc == 1) {
      bool any_active = nnew > 0;
      for (auto &p : S->live) any_active = any_active || p->state == 2;
      if (any_active || nj2 + r.pl->nspec > MAX_JOBS) break;  // room frees up later
      drop = true;  // nothing can free space: the request does not fit the arena
    } else if (rc == ALPSG_E_ARG) {
      drop = true;  // this request's own data (Bessel cutoff not reached)
    } else if (rc) {
      return rc;
    }
    if (drop) {
      done.push_back(Result{r.tag, cplx(0.0, 0.0), 0, ALPSG_E_ARG, now_ms() - r.t_submit});
      release(S, r);
      it = S->live.erase(it);
      continue;
    }
    ++it;
  }
  if (nj2 > 0) {
    int *sb = S->jb2.start_h, *sa = sb + (MAX_JOBS + 1), *sm = sa + (MAX_JOBS + 1);
    const int tb = fill_start(S->jb2.jobs_h, nj2, sb, jobs_blocks_bessel);
    const int ta = fill_start(S->jb2.jobs_h, nj2, sa, jobs_blocks_atab);
    const int tm = fill_start(S->jb2.jobs_h, nj2, sm, jobs_blocks_moments);
    SCHK(hipMemcpyAsync(S->jb2.jobs_d, S->jb2.jobs_h, sizeof(TabJob) * nj2, hipMemcpyHostToDevice, S->st));
    SCHK(hipMemcpyAsync(S->jb2.start_d, S->jb2.start_h, sizeof(int) * 3 * (MAX_JOBS + 1),
                        hipMemcpyHostToDevice, S->st));
    int *db = S->jb2.start_d, *da = db + (MAX_JOBS + 1), *dm = da + (MAX_JOBS + 1);
    SCHK(launch_species_k_jobs(S->st, S->jb2.jobs_d, nj2, db, tb, da, ta, dm, tm, max_nperp, S->mtab_d));
  }
  if (nnew > 0) {
    SCHK(hipMemcpyAsync(S->newp_d, S->newp_h, sizeof(ProbDev) * nnew, hipMemcpyHostToDevice, S->st));
    SCHK(hipMemcpyAsync(S->newslot_d, S->newslot_h, sizeof(int) * nnew, hipMemcpyHostToDevice, S->st));
    SCHK(launch_scatter_probs(S->st, S->newp_d, S->newslot_d, nnew, S->probs_d));
  }
  S->t_admit += now_ms() - t3;
  // feed the results; retire finished solves
  for (int i = 0; i < nb; ++i) {
    if (pass_rel && S->err_h[i] != 0) {  // alps_error(8): the solve stops with an error
      act[i]->failed = true;
      continue;
    }
    act[i]->m.feed(alpsg_osc::zmk(S->det_h[i].x, S->det_h[i].y));
  }
  for (auto it = S->live.begin(); it != S->live.end();) {
    Request &r = **it;
    if (r.state == 2 && r.failed) {
      done.push_back(Result{r.tag, cplx(0.0, 0.0), r.m.iter, ALPSG_E_ARG, now_ms() - r.t_submit});
      release(S, r);
      it = S->live.erase(it);
      continue;
    }
    if (r.state == 2 && r.m.done) {
      // a solve that used up numiter iterations returns its best point, as
      // secant_osc does after "Maximum iteration reached" (ALPS_fns.f90:2093-2099)
      const int st = r.m.iter >= r.m.numiter ? ALPSG_MAXITER : ALPSG_OK;
      done.push_back(Result{r.tag, cplx(r.m.result().re, r.m.result().im), r.m.iter, st,
                            now_ms() - r.t_submit});
      release(S, r);
      it = S->live.erase(it);
    } else {
      ++it;
    }
  }
  // device blocks of plasmas that are no longer registered and no request reads
  for (auto &p : S->resident)
    if (p->off != SIZE_MAX && p->users == 0 && p->ee_state != 1 && !p->registered.load()) evict(S, *p);
  drop_nonresident(S);
  S->t_pass += now_ms() - tp0;
  post_results(S, done);
  return ALPSG_OK;
}

void service_loop(alpsg_server *S) {
  (void)hipSetDevice(0);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv.wait(lk, [&] { return S->stop || S->gc || !S->queue.empty() || !S->live.empty(); });
      if (S->stop && S->queue.empty() && S->live.empty()) return;
      S->gc = false;
    }
    int rc;
    try {
      rc = server_pass(S);
    } catch (const std::exception &e) {
      rc = alpsg_internal_fail(ALPSG_E_DEVICE, std::string("internal error in a pass: ") + e.what());
    } catch (...) {
      rc = alpsg_internal_fail(ALPSG_E_DEVICE, "internal error in a pass");
    }
    if (rc) {
      std::lock_guard<std::mutex> lk(S->mu);
      S->failed = true;
      S->fail_msg = alpsg_last_error();
      S->cv_done.notify_all();
      return;
    }
  }
}

int make_request(std::shared_ptr<Plasma> pl, double kperp, double kpar, const double *om_guess,
                 const alpsg_root_params *rp, long long tag, std::unique_ptr<Request> &out) {
  if (!om_guess || !rp) return alpsg_internal_fail(ALPSG_E_ARG, "null argument");
  if (!(kpar != 0.0) || !std::isfinite(kpar) || !std::isfinite(kperp))
    return alpsg_internal_fail(ALPSG_E_ARG, "kpar must be finite and nonzero");
  if (!std::isfinite(om_guess[0]) || !std::isfinite(om_guess[1]))
    return alpsg_internal_fail(ALPSG_E_ARG, "initial frequency must be finite");
  if (rp->secant_method != 2)
    return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "the server runs secant_osc (secant_method 2)");
  if (rp->numiter < 1 || rp->numiter > 100000) return alpsg_internal_fail(ALPSG_E_ARG, "numiter out of range");
  std::unique_ptr<Request> r(new Request());
  r->tag = tag;
  r->kperp = kperp;
  r->kpar = kpar;
  r->guess = cplx(om_guess[0], om_guess[1]);
  r->rp = *rp;
  r->ntry = ntry_bound(*pl, kperp);
  r->pl = std::move(pl);
  out = std::move(r);
  return ALPSG_OK;
}

// queue prepared requests (all or none)
int enqueue(alpsg_server *S, std::vector<std::unique_ptr<Request>> &rs) {
  const double t = now_ms();
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->failed) return alpsg_internal_fail(ALPSG_E_DEVICE, "server stopped: " + S->fail_msg);
  if (S->stop) return alpsg_internal_fail(ALPSG_E_STATE, "server is shutting down");
  for (auto &r : rs) {
    r->t_submit = t;
    S->queue.push_back(std::move(r));
  }
  S->submitted += (long long)rs.size();
  S->cv.notify_all();
  return ALPSG_OK;
}

}  // namespace

extern "C" {

int alpsg_server_create(double hbm_bytes, int max_slots, alpsg_server **out) {
  ALPSG_TRY
  if (!out) return alpsg_internal_fail(ALPSG_E_ARG, "null argument");
  *out = nullptr;
  if (max_slots < 1 || max_slots > 65536) return alpsg_internal_fail(ALPSG_E_ARG, "max_slots out of range 1..65536");
  if (!(hbm_bytes >= (double)(1 << 20)) || !(hbm_bytes <= 256e9))
    return alpsg_internal_fail(ALPSG_E_ARG, "HBM budget out of range 1 MiB..256 GB");
  std::unique_ptr<alpsg_server> S(new alpsg_server());
  S->max_slots = max_slots;
  // on failure below, alpsg_server_destroy frees what exists
  struct Guard {
    std::unique_ptr<alpsg_server> &p;
    ~Guard() {
      if (p) alpsg_server_destroy(p.release());
    }
  } guard{S};
  SCHK(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking));
  for (auto &row : S->ev)
    for (auto &e : row) SCHK(hipEventCreate(&e));
  const size_t cap = (size_t)hbm_bytes / ALIGN * ALIGN;
  SCHK(hipMalloc((void **)&S->arena_mem, cap));
  S->arena.init(S->arena_mem, cap);
  SCHK(hipMalloc((void **)&S->probs_d, sizeof(ProbDev) * max_slots));
  for (int i = max_slots - 1; i >= 0; --i) S->free_slots.push_back(i);
  // Miller start indices m(n) = 2*((n+int(sqrt(float(40n))))/2) (ALPS_fns_rel.f90:1622)
  std::vector<int> mt(MTAB_S, 0);
  for (int n = 1; n < MTAB_S; ++n) mt[n] = 2 * ((n + (int)std::sqrt((float)(40 * n))) / 2);
  SCHK(hipMalloc((void **)&S->mtab_d, sizeof(int) * MTAB_S));
  SCHK(hipMemcpy(S->mtab_d, mt.data(), sizeof(int) * MTAB_S, hipMemcpyHostToDevice));
  const size_t mb = (size_t)max_slots;
  SCHK(hipMalloc((void **)&S->om_d, sizeof(double2) * mb));
  SCHK(hipMalloc((void **)&S->det_d, sizeof(double2) * mb));
  SCHK(hipMalloc((void **)&S->pidx_d, sizeof(int) * mb));
  SCHK(hipMalloc((void **)&S->chiA, sizeof(double2) * mb * (SW_W * SW_RC) * ALPSG_MAXSPEC * 6));
  SCHK(hipMalloc((void **)&S->chiB, sizeof(double2) * mb * (2 * RS_RB) * ALPSG_MAXSPEC * 6));
  SCHK(hipHostMalloc((void **)&S->om_h, sizeof(double2) * mb, hipHostMallocDefault));
  SCHK(hipHostMalloc((void **)&S->det_h, sizeof(double2) * mb, hipHostMallocDefault));
  SCHK(hipHostMalloc((void **)&S->pidx_h, sizeof(int) * mb, hipHostMallocDefault));
  SCHK(hipHostMalloc((void **)&S->stage_h, sizeof(double) * STAGE_CAP, hipHostMallocDefault));
  SCHK(hipMalloc((void **)&S->stage_d, sizeof(double) * STAGE_CAP));
  SCHK(S->jb1.init());
  SCHK(S->jb2.init());
  SCHK(hipHostMalloc((void **)&S->newp_h, sizeof(ProbDev) * mb, hipHostMallocDefault));
  SCHK(hipMalloc((void **)&S->newp_d, sizeof(ProbDev) * mb));
  SCHK(hipHostMalloc((void **)&S->newslot_h, sizeof(int) * mb, hipHostMallocDefault));
  SCHK(hipMalloc((void **)&S->newslot_d, sizeof(int) * mb));
  SCHK(hipMalloc((void **)&S->err_d, sizeof(int) * mb));
  SCHK(hipHostMalloc((void **)&S->err_h, sizeof(int) * mb, hipHostMallocDefault));
  {
    std::vector<double> gt;
    rel_gamma_table(gt);
    SCHK(hipMalloc((void **)&S->gtab_d, sizeof(double) * gt.size()));
    SCHK(hipMemcpy(S->gtab_d, gt.data(), sizeof(double) * gt.size(), hipMemcpyHostToDevice));
  }
  S->th = std::thread(service_loop, S.get());
  *out = S.release();
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_server_add_plasma(alpsg_server *S, const alpsg_plasma_desc *d, int *plasma_id) {
  ALPSG_TRY
  if (!S || !d || !plasma_id) return alpsg_internal_fail(ALPSG_E_ARG, "null argument");
  auto pl = std::make_shared<Plasma>();
  const int rc = prepare_plasma(d, *pl);
  if (rc) return rc;
  pl->registered = true;
  std::lock_guard<std::mutex> lk(S->mu);
  pl->id = S->next_id++;
  S->registry[pl->id] = pl;
  *plasma_id = pl->id;
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_server_remove_plasma(alpsg_server *S, int plasma_id) {
  ALPSG_TRY
  if (!S) return alpsg_internal_fail(ALPSG_E_ARG, "null handle");
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->registry.find(plasma_id);
  if (it == S->registry.end()) return alpsg_internal_fail(ALPSG_E_ARG, "unknown plasma id");
  it->second->registered = false;
  S->registry.erase(it);
  S->gc = true;  // its device block goes once no request reads it
  S->cv.notify_all();
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_server_submit_batch(alpsg_server *S, int n, const int *plasma_ids, const double *kperp,
                              const double *kpar, const double *om_guess,
                              const alpsg_root_params *rp, const long long *tags) {
  ALPSG_TRY
  if (!S) return alpsg_internal_fail(ALPSG_E_ARG, "null handle");
  if (n < 0) return alpsg_internal_fail(ALPSG_E_ARG, "negative request count");
  if (n == 0) return ALPSG_OK;
  if (!plasma_ids || !kperp || !kpar || !om_guess || !rp || !tags)
    return alpsg_internal_fail(ALPSG_E_ARG, "null argument");
  std::vector<std::shared_ptr<Plasma>> pls((size_t)n);
  {
    std::lock_guard<std::mutex> lk(S->mu);
    for (int i = 0; i < n; ++i) {
      auto it = S->registry.find(plasma_ids[i]);
      if (it == S->registry.end()) return alpsg_internal_fail(ALPSG_E_ARG, "unknown plasma id");
      pls[i] = it->second;
    }
  }
  std::vector<std::unique_ptr<Request>> rs((size_t)n);
  for (int i = 0; i < n; ++i) {
    const int rc = make_request(std::move(pls[i]), kperp[i], kpar[i], om_guess + 2 * (size_t)i, rp,
                                tags[i], rs[i]);
    if (rc) return rc;
  }
  return enqueue(S, rs);
  ALPSG_CATCH
}

int alpsg_server_submit_k(alpsg_server *S, int plasma_id, double kperp, double kpar,
                          const double *om_guess, const alpsg_root_params *rp, long long tag) {
  ALPSG_TRY
  return alpsg_server_submit_batch(S, 1, &plasma_id, &kperp, &kpar, om_guess, rp, &tag);
  ALPSG_CATCH
}

int alpsg_server_submit(alpsg_server *S, const alpsg_plasma_desc *d, double kperp, double kpar,
                        const double *om_guess, const alpsg_root_params *rp, long long tag) {
  ALPSG_TRY
  if (!S || !d) return alpsg_internal_fail(ALPSG_E_ARG, "null argument");
  auto pl = std::make_shared<Plasma>();  // unregistered: freed with its request
  const int rc = prepare_plasma(d, *pl);
  if (rc) return rc;
  std::vector<std::unique_ptr<Request>> rs(1);
  const int rc2 = make_request(std::move(pl), kperp, kpar, om_guess, rp, tag, rs[0]);
  if (rc2) return rc2;
  return enqueue(S, rs);
  ALPSG_CATCH
}

int alpsg_server_poll(alpsg_server *S, int max_results, int wait, long long *tags, double *roots,
                      int *iters, int *status, double *latency_ms, int *n_out) {
  ALPSG_TRY
  if (!S || !n_out) return alpsg_internal_fail(ALPSG_E_ARG, "null argument");
  std::unique_lock<std::mutex> lk(S->mu);
  if (wait)
    S->cv_done.wait(lk, [&] { return !S->results.empty() || S->failed || S->finished >= S->submitted; });
  int n = 0;
  while (n < max_results && !S->results.empty()) {
    const Result &r = S->results.front();
    if (tags) tags[n] = r.tag;
    if (roots) {
      roots[2 * n] = r.root.real();
      roots[2 * n + 1] = r.root.imag();
    }
    if (iters) iters[n] = r.iters;
    if (status) status[n] = r.status;
    if (latency_ms) latency_ms[n] = r.latency_ms;
    S->results.pop_front();
    ++n;
  }
  *n_out = n;
  if (S->failed && n == 0) return alpsg_internal_fail(ALPSG_E_DEVICE, "server stopped: " + S->fail_msg);
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_server_stats(alpsg_server *S, long long *counts, double *times_ms) {
  ALPSG_TRY
  if (!S) return alpsg_internal_fail(ALPSG_E_ARG, "null handle");
  std::lock_guard<std::mutex> lk(S->mu);
  if (counts) {
    counts[0] = S->submitted;
    counts[1] = S->finished;
    counts[2] = S->snap.passes;
    counts[3] = S->snap.evals;
    counts[4] = S->snap.live;
    counts[5] = (long long)S->queue.size();
    counts[6] = (long long)S->arena.cap;
    counts[7] = S->snap.arena_used;
    counts[8] = S->snap.uploads;
    counts[9] = S->snap.evictions;
  }
  if (times_ms) {
    times_ms[0] = S->snap.t_pass;
    times_ms[1] = S->snap.t_admit;
    times_ms[2] = S->snap.t_launch;
    times_ms[3] = S->snap.t_wait;
  }
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_server_work(alpsg_server *S, double *work7) {
  ALPSG_TRY
  if (!S || !work7) return alpsg_internal_fail(ALPSG_E_ARG, "null argument");
  std::lock_guard<std::mutex> lk(S->mu);
  for (int k = 0; k < 3; ++k) work7[k] = S->snap.work[k];
  for (int k = 0; k < 4; ++k) work7[3 + k] = S->snap.kt[k];
  return ALPSG_OK;
  ALPSG_CATCH
}

int alpsg_server_destroy(alpsg_server *S) {
  ALPSG_TRY
  if (!S) return ALPSG_OK;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    S->stop = true;
    S->cv.notify_all();
  }
  if (S->th.joinable()) S->th.join();
  if (S->st) (void)hipStreamSynchronize(S->st);
  if (S->arena_mem) (void)hipFree(S->arena_mem);
  if (S->probs_d) (void)hipFree(S->probs_d);
  if (S->mtab_d) (void)hipFree(S->mtab_d);
  if (S->om_d) (void)hipFree(S->om_d);
  if (S->det_d) (void)hipFree(S->det_d);
  if (S->pidx_d) (void)hipFree(S->pidx_d);
  if (S->chiA) (void)hipFree(S->chiA);
  if (S->chiB) (void)hipFree(S->chiB);
  if (S->om_h) (void)hipHostFree(S->om_h);
  if (S->det_h) (void)hipHostFree(S->det_h);
  if (S->pidx_h) (void)hipHostFree(S->pidx_h);
  if (S->stage_h) (void)hipHostFree(S->stage_h);
  if (S->stage_d) (void)hipFree(S->stage_d);
  S->jb1.free_all();
  S->jb2.free_all();
  if (S->newp_h) (void)hipHostFree(S->newp_h);
  if (S->newp_d) (void)hipFree(S->newp_d);
  if (S->newslot_h) (void)hipHostFree(S->newslot_h);
  if (S->newslot_d) (void)hipFree(S->newslot_d);
  if (S->err_d) (void)hipFree(S->err_d);
  if (S->err_h) (void)hipHostFree(S->err_h);
  if (S->gtab_d) (void)hipFree(S->gtab_d);
  if (S->chiR) (void)hipFree(S->chiR);
  if (S->flags_d) (void)hipFree(S->flags_d);
  for (auto &row : S->ev)
    for (auto &e : row)
      if (e) (void)hipEventDestroy(e);
  if (S->st) (void)hipStreamDestroy(S->st);
  delete S;
  return ALPSG_OK;
  ALPSG_CATCH
}

}  // extern "C"
