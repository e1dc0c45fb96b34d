// Continuous-batching root server.
//
// Independent root solves -- each with its own plasma tables, grid size,
// wavevector, initial frequency and root controls -- arrive with
// alpsg_server_submit.  A service thread admits them into a fixed HBM arena as
// its space and the problem slots allow (first fit, by each request's
// footprint), builds their tables on the device, and advances every admitted
// solve by one secant_osc step per device pass: the pending frequencies of all
// active solves go into one launch of k_sweep / k_resonant / k_assemble in
// their one-omega-per-workgroup form, where each workgroup reads its own
// solve's problem descriptor (pidx).  Solves of uneven nmax and nperp x npar
// therefore share launches.  A finished solve frees its arena blocks and slot
// at once, and the next waiting request takes them in the following pass.
//
// Admission is pipelined with the passes: pass i uploads a new request's
// tables and evaluates max |J_n| over its pperp grid (determine_nmax,
// ALPS_fns.f90:3971-4075); after pass i's synchronisation the host picks nmax,
// enqueues the per-k tables, and the solve joins pass i+1.  The launch shape of
// a pass is fixed, so a solve's arithmetic does not depend on the other solves
// in the pass: its result is the same alone and in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
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
#include "alpsg_device.h"
#include "alpsg_osc.h"

// shared with alpsg_api.cpp
int alpsg_internal_fail(int code, const std::string &msg);
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
                        const int *pidx);
hipError_t launch_resonant(hipStream_t st, int nch, int RB, const ProbDev *prob, int nspec, int nom,
                           const double2 *om, const int *active, double2 *chiB, double2 *lowB,
                           bool general, const int *pidx);
hipError_t launch_assemble(hipStream_t st, const ProbDev *prob, int nspec, int nom, int W, int RB,
                           const double2 *om, const int *active, const double2 *chiA,
                           const double2 *chiB, const double2 *lowA, const double2 *lowB,
                           const double2 *chiR, int RR, const double2 *lowR, double2 *det,
                           double2 *chi, double2 *low, double2 *wave, const int *pidx);

namespace {

using alpsg_host::cplx;
using alpsg_host::OscMachine;

#define SCHK(expr)                                                                           \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess)                                                                    \
      return alpsg_internal_fail(ALPSG_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

constexpr int MTAB_S = 4096;
// the fixed launch shape of a pass: k_sweep<64, 8> over 4 workgroups per omega,
// k_resonant<64, *> over 8 harmonic blocks
constexpr int SW_W = 8, SW_RC = 4, RS_RB = 8;
constexpr size_t ALIGN = 256;

double now_ms() {
  using namespace std::chrono;
  return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// first-fit allocator over one device allocation
struct Arena {
  char *base = nullptr;
  size_t cap = 0;
  std::map<size_t, size_t> free_;  // offset -> size
  void init(char *b, size_t c) {
    base = b;
    cap = c;
    free_.clear();
    free_[0] = c;
  }
  // returns offset or SIZE_MAX
  size_t alloc(size_t sz) {
    sz = (sz + ALIGN - 1) / ALIGN * ALIGN;
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= sz) {
        const size_t off = it->first, rest = it->second - sz;
        free_.erase(it);
        if (rest) free_[off + sz] = rest;
        return off;
      }
    return SIZE_MAX;
  }
  void release(size_t off, size_t sz) {
    sz = (sz + ALIGN - 1) / ALIGN * ALIGN;
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
  size_t free_bytes() const {
    size_t t = 0;
    for (auto &f : free_) t += f.second;
    return t;
  }
};

// bump layout of several arrays inside one arena block
struct Layout {
  size_t size = 0;
  size_t add(size_t bytes) {
    const size_t o = size;
    size += (bytes + ALIGN - 1) / ALIGN * ALIGN;
    return o;
  }
};

struct Request {
  long long tag = 0;
  double kperp = 0, kpar = 0;
  cplx guess;
  alpsg_root_params rp{};
  double t_submit = 0;
  // host copy of the plasma (species-major, C layout)
  int nspec = 0, nperp = 0, npar = 0, P = 3, NRI = 100, kperp_norm = 1;
  double vA = 0, Tlim = 0.01, Bessel_zero = 1e-45;
  std::vector<double> qs, ms, ns, pperp, ppar;
  std::vector<int> acm, nfits, ftype, funi, porder, plog;
  std::vector<double> plogmax;
  std::vector<size_t> fit_off;
  int maxorder = 0;
  bool general = false;  // a species needs the general Landau part
  // omega-independent block: host image of the uploaded prefix, then device-only arrays
  std::vector<char> img;
  size_t o_pperp = 0, o_ppar = 0, o_f0 = 0, o_qs = 0, o_ms = 0, o_fit = 0, o_poly = 0, n_up = 0;
  size_t o_df1 = 0, o_df2 = 0, o_ee = 0, o_bmax = 0, sz_s = 0;
  int ntry = 0;
  // device state
  int slot = -1, state = 0;  // 0 waiting, 1 sizing, 2 active
  size_t off_s = SIZE_MAX, off_k = SIZE_MAX, sz_k = 0;
  int stage = -1;
  std::vector<double> sized;  // int_ee[nspec] then max|J_n| [nspec][ntry], read back
  int nmax[ALPSG_MAXSPEC] = {0};
  OscMachine m;
};

struct Result {
  long long tag;
  cplx root;
  int iters, status;
  double latency_ms;
};

}  // namespace

struct alpsg_server {
  hipStream_t st = nullptr;
  char *arena_mem = nullptr;
  Arena arena;
  int max_slots = 0, max_batch = 0;
  static constexpr int max_admit = 64;  // admissions started per pass
  ProbDev *probs_d = nullptr;
  std::vector<ProbDev> probs_h;
  std::vector<int> free_slots;
  int *mtab_d = nullptr;
  // pass buffers
  double2 *om_d = nullptr, *det_d = nullptr, *chiA = nullptr, *chiB = nullptr;
  int *pidx_d = nullptr;
  double2 *om_h = nullptr, *det_h = nullptr;  // pinned
  int *pidx_h = nullptr;                      // pinned
  double *stage_h = nullptr;                  // pinned: one row per admission of a pass
  size_t stage_row = 0;                       // doubles per row
  // requests
  std::mutex mu;
  std::condition_variable cv, cv_done;
  std::deque<std::unique_ptr<Request>> queue;  // submitted, not admitted
  std::vector<std::unique_ptr<Request>> live;  // admitted (sizing or active)
  std::deque<Result> results;
  long long submitted = 0, finished = 0;
  bool stop = false, failed = false;
  std::string fail_msg;
  std::thread th;
  // statistics
  long long passes = 0, evals = 0;
  double busy_ms = 0;
};

namespace {

constexpr int MAXSPEC_S = ALPSG_MAXSPEC;
constexpr size_t STAGE_ROW = (size_t)MAXSPEC_S * (MTAB_S + 1);

// Host copy of a served request's plasma (tabulated, non-relativistic species
// with fits 1/2/6 or Chebyshev continuation), the checks of alpsg_create, and
// the host image of its omega-independent arena block.
int prepare_request(const alpsg_plasma_desc *d, Request &r) {
  const int nspec = d->nspec, nperp = d->nperp, npar = d->npar;
  if (nspec < 1 || nspec > ALPSG_MAXSPEC) return alpsg_internal_fail(ALPSG_E_ARG, "nspec out of range 1..8");
  if (!d->pp || !d->f0 || !d->ns || !d->qs || !d->ms || !d->ACmethod || !d->n_fits)
    return alpsg_internal_fail(ALPSG_E_ARG, "served requests need pp, f0, ns, qs, ms, ACmethod, n_fits");
  const int P = d->positions_principal;
  if (P < 1 || P > 16) return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "positions_principal must be in 1..16");
  if (nperp < 4 || npar < 2 * P + 8) return alpsg_internal_fail(ALPSG_E_ARG, "grid too small for positions_principal");
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
  for (int s = 0; s < nspec; ++s) {
    if ((d->usebM && d->usebM[s]) || (d->relativistic && d->relativistic[s]))
      return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "served requests take tabulated non-relativistic species");
    if (d->qs[s] == 0.0 || d->ms[s] <= 0.0) return alpsg_internal_fail(ALPSG_E_ARG, "species charge/mass");
    for (int i = 0; i <= nperp; ++i) r.pperp[(size_t)s * (nperp + 1) + i] = PP(s, i, 1, 0);
    for (int j = 0; j <= npar; ++j) r.ppar[(size_t)s * (npar + 1) + j] = PP(s, 2, j, 1);
    for (int i = 0; i <= nperp; ++i)
      for (int j = 0; j <= npar; ++j)
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
  r.porder.assign(nspec, 0);
  r.plog.assign(nspec, 0);
  r.plogmax.assign(nspec, 0.0);
  r.fit_off.assign(nspec + 1, 0);
  r.maxorder = 0;
  r.general = false;
  for (int s = 0; s < nspec; ++s) {
    if (r.acm[s] == 1) {
      const int nf = d->n_fits[s];
      if (nf < 1 || nf > ALPSG_MAXFITS || nf > d->max_fits) return alpsg_internal_fail(ALPSG_E_ARG, "n_fits out of range");
      if (!d->fit_type || !d->perp_correction || !d->param_fit) return alpsg_internal_fail(ALPSG_E_ARG, "missing fit arrays");
      if (npf < nperp) return alpsg_internal_fail(ALPSG_E_ARG, "param_fit_nperp < nperp");
      r.nfits[s] = nf;
      bool uni = true;
      for (int f = 0; f < nf; ++f) {
        const int ft = d->fit_type[s + (size_t)nspec * f];
        if (ft != 1 && ft != 2 && ft != 6) return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "fit_type must be 1, 2 or 6");
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
      if (!(uni && nf == 1)) r.general = true;
    } else if (r.acm[s] == 2) {
      if (!d->poly_kind || !d->poly_order || !d->logfit || !d->poly_log_max || !d->poly_fit_coeffs)
        return alpsg_internal_fail(ALPSG_E_ARG, "missing Chebyshev arrays");
      if (d->poly_kind[s] != 1) return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "poly_kind must be 1");
      if (d->poly_order[s] < 1 || d->poly_order[s] > d->poly_max_order) return alpsg_internal_fail(ALPSG_E_ARG, "poly_order");
      r.porder[s] = d->poly_order[s];
      r.plog[s] = d->logfit[s] ? 1 : 0;
      r.plogmax[s] = d->poly_log_max[s];
      r.maxorder = std::max(r.maxorder, r.porder[s]);
      r.general = true;
    } else {
      return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "ACmethod must be 1 or 2");
    }
    r.fit_off[s + 1] = r.fit_off[s] + (r.acm[s] == 1 ? (size_t)(nperp + 1) * r.nfits[s] * FIT_STRIDE : 0);
  }
  // block layout: the uploaded prefix, then df0, int_ee and max|J_n| (read back together)
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
  r.n_up = L.size;
  r.o_df1 = L.add(sizeof(double) * ng);
  r.o_df2 = L.add(sizeof(double) * ng);
  r.o_ee = L.size;  // int_ee and max|J_n| contiguous
  L.add(sizeof(double) * nspec);
  r.o_bmax = r.o_ee + sizeof(double) * nspec;
  r.img.assign(r.n_up, 0);
  auto H = [&](size_t o) { return reinterpret_cast<double *>(r.img.data() + o); };
  std::memcpy(H(r.o_pperp), r.pperp.data(), sizeof(double) * r.pperp.size());
  std::memcpy(H(r.o_ppar), r.ppar.data(), sizeof(double) * r.ppar.size());
  std::memcpy(H(r.o_qs), r.qs.data(), sizeof(double) * nspec);
  std::memcpy(H(r.o_ms), r.ms.data(), sizeof(double) * nspec);
  double *f0 = H(r.o_f0);
  for (int s = 0; s < nspec; ++s)
    for (int i = 0; i <= nperp; ++i)
      for (int j = 0; j <= npar; ++j)
        f0[((size_t)s * (nperp + 1) + i) * (npar + 1) + j] =
            d->f0[(size_t)s + (size_t)nspec * (i + (size_t)(nperp + 1) * j)];
  // fit_function coefficients (ALPS_analyt.f90:185-266) with the pperp factor
  // folded in, as alpsg_create tabulates them
  double *fitc = H(r.o_fit);
  double *polyc = H(r.o_poly);
  for (int s = 0; s < nspec; ++s) {
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
          } else {
            c[0] = prm(s, i, 0, f);
            c[3] = prm(s, i, 3, f) * pc * (pv * pv);
          }
        }
      }
    } else {
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
void size_request(Request &r) {
  double xmax = 0.0;
  for (int s = 0; s < r.nspec; ++s)
    for (int i = 0; i <= r.nperp; ++i)
      xmax = std::max(xmax, std::fabs(r.kperp * r.pperp[(size_t)s * (r.nperp + 1) + i] / r.qs[s]));
  int ntry = 1;
  const double lz = std::log(std::max(r.Bessel_zero, 1e-300)) - std::log(1e6);
  while (ntry < MTAB_S - 16) {
    const double lb = ntry * std::log(std::max(xmax, 1e-300) / 2.0) - std::lgamma(ntry + 1.0);
    if (lb < lz && ntry > xmax) break;
    ++ntry;
  }
  r.ntry = std::min(ntry + 8, MTAB_S - 4);
  r.sz_s = r.o_bmax + (sizeof(double) * (size_t)r.nspec * r.ntry + ALIGN - 1) / ALIGN * ALIGN;
}

// stage 1 of admission: the omega-independent block, df0, int_ee and max|J_n|
// (read back into pinned row 'stage'); returns 1 when the arena has no room
int admit_static(alpsg_server *S, Request &r, int stage) {
  const int ns = r.nspec, nperp = r.nperp, npar = r.npar;
  const size_t off = S->arena.alloc(r.sz_s);
  if (off == SIZE_MAX) return 1;
  r.off_s = off;
  char *b = S->arena.base + off;
  auto P = [&](size_t o) { return reinterpret_cast<double *>(b + o); };
  SCHK(hipMemcpyAsync(b, r.img.data(), r.n_up, hipMemcpyHostToDevice, S->st));
  SCHK(hipMemsetAsync(P(r.o_df1), 0, r.o_ee - r.o_df1, S->st));
  SCHK(launch_derivative(S->st, ns, nperp, npar, P(r.o_f0), P(r.o_pperp), P(r.o_ppar), P(r.o_df1), P(r.o_df2)));
  SCHK(launch_int_ee(S->st, ns, nperp, npar, P(r.o_pperp), P(r.o_ppar), P(r.o_df1), P(r.o_df2), P(r.o_qs),
                     P(r.o_ms), P(r.o_ee)));
  SCHK(launch_bessel_max(S->st, ns, nperp, r.ntry, r.kperp, P(r.o_pperp), P(r.o_qs), S->mtab_d, P(r.o_bmax)));
  r.stage = stage;
  SCHK(hipMemcpyAsync(S->stage_h + (size_t)stage * S->stage_row, P(r.o_ee),
                      sizeof(double) * (size_t)ns * (1 + r.ntry), hipMemcpyDeviceToHost, S->st));
  r.state = 1;
  return 0;
}

// stage 2: nmax from max|J_n| (determine_nmax), the per-k tables, the problem
// descriptor; returns 1 when the arena has no room
int admit_k(alpsg_server *S, Request &r) {
  const int ns = r.nspec, nperp = r.nperp, npar = r.npar;
  const double *ee = r.sized.data();
  const double *bm = ee + ns;
  for (int s = 0; s < ns; ++s) {
    int n = 0;
    for (int t = 1; t <= r.ntry; ++t)
      if (!(bm[(size_t)s * r.ntry + (t - 1)] > r.Bessel_zero)) {
        n = t;
        break;
      }
    if (n == 0) return alpsg_internal_fail(ALPSG_E_ARG, "Bessel cutoff not reached (kperp too large?)");
    r.nmax[s] = n;
  }
  Layout L;
  size_t ob[ALPSG_MAXSPEC], oa[ALPSG_MAXSPEC], on[ALPSG_MAXSPEC];
  for (int s = 0; s < ns; ++s) {
    const int nm = r.nmax[s];
    ob[s] = L.add(sizeof(double) * (size_t)(nm + 3) * (nperp + 1));
    oa[s] = L.add(sizeof(double) * (size_t)(nm + 1) * (nperp + 1) * ATAB_STRIDE);
    on[s] = L.add(sizeof(double) * (size_t)(nm + 1) * (npar - 1) * NODE_STRIDE);
  }
  const size_t off = S->arena.alloc(L.size);
  if (off == SIZE_MAX) return 1;
  r.off_k = off;
  r.sz_k = L.size;
  char *bs = S->arena.base + r.off_s, *bk = S->arena.base + off;
  auto PS = [&](size_t o) { return reinterpret_cast<double *>(bs + o); };
  auto PK = [&](size_t o) { return reinterpret_cast<double *>(bk + o); };
  ProbDev &pd = S->probs_h[r.slot];
  std::memset(&pd, 0, sizeof(pd));
  pd.nspec = ns;
  pd.kperp_norm = r.kperp_norm;
  pd.P = r.P;
  pd.NRI = r.NRI;
  pd.vA = r.vA;
  pd.Tlim = r.Tlim;
  pd.kperp = r.kperp;
  pd.kpar = r.kpar;
  for (int s = 0; s < ns; ++s) {
    SpecDev &sd = pd.sp[s];
    const double *pe = &r.pperp[(size_t)s * (nperp + 1)];
    const double *pa = &r.ppar[(size_t)s * (npar + 1)];
    const int nm = r.nmax[s];
    double *pperp_d = PS(r.o_pperp) + (size_t)s * (nperp + 1);
    double *ppar_d = PS(r.o_ppar) + (size_t)s * (npar + 1);
    SCHK(launch_species_k(S->st, nperp, npar, nm, r.kperp, r.kpar, r.qs[s], r.ms[s], pperp_d, ppar_d,
                          PS(r.o_df1) + (size_t)s * (nperp + 1) * (npar + 1),
                          PS(r.o_df2) + (size_t)s * (nperp + 1) * (npar + 1), S->mtab_d, PK(ob[s]),
                          PK(oa[s]), PK(on[s])));
    sd.nperp = nperp;
    sd.npar = npar;
    sd.nmax = nm;
    sd.acmethod = r.acm[s];
    sd.nfits = r.acm[s] == 1 ? r.nfits[s] : 0;
    for (int f = 0; f < sd.nfits; ++f) sd.fit_type[f] = r.ftype[(size_t)s * ALPSG_MAXFITS + f];
    sd.fit_uniform = r.funi[s];
    sd.q = r.qs[s];
    sd.m = r.ms[s];
    sd.dens = r.ns[s];
    sd.dpperp = pe[2] - pe[1];
    sd.dppar = pa[2] - pa[1];
    sd.dppar_abs = std::fabs(pa[2] - pa[1]);
    sd.z = r.kperp_norm ? r.kperp / r.qs[s] : 1.0 / r.qs[s];
    sd.int_ee = ee[s];
    sd.pperp = pperp_d;
    sd.ppar = ppar_d;
    sd.atab = PK(oa[s]);
    sd.nodes = PK(on[s]);
    sd.fitc = PS(r.o_fit) + r.fit_off[s];
    if (r.acm[s] == 2) {
      sd.poly_kind = 1;
      sd.poly_order = r.porder[s];
      sd.poly_logfit = r.plog[s];
      sd.poly_log_max = r.plogmax[s];
      sd.poly_norm1 = 5.e-1 * (pa[npar] + pa[0]);
      sd.poly_norm2 = 5.e-1 * (pa[npar] - pa[0]);
      sd.polyc = PS(r.o_poly) + (size_t)s * (nperp + 1) * (r.maxorder + 1);
    }
  }
  SCHK(hipMemcpyAsync(S->probs_d + r.slot, &pd, sizeof(ProbDev), hipMemcpyHostToDevice, S->st));
  r.m.start(r.guess, r.rp);
  r.state = 2;
  return 0;
}

void release(alpsg_server *S, Request &r) {
  if (r.off_s != SIZE_MAX) S->arena.release(r.off_s, r.sz_s);
  if (r.off_k != SIZE_MAX) S->arena.release(r.off_k, r.sz_k);
  r.off_s = r.off_k = SIZE_MAX;
  if (r.slot >= 0) S->free_slots.push_back(r.slot);
  r.slot = -1;
}

void post_results(alpsg_server *S, std::vector<Result> &done) {
  if (done.empty()) return;
  std::lock_guard<std::mutex> lk(S->mu);
  for (auto &d : done) S->results.push_back(d);
  S->finished += (long long)done.size();
  S->cv_done.notify_all();
  done.clear();
}

// one pass: admissions, the batched disp of every active solve, results
int server_pass(alpsg_server *S) {
  std::vector<Result> done;
  // stage 1 for waiting requests, first come first served, while slots and space last
  std::vector<Request *> sizing_now;
  for (;;) {
    Request *r = nullptr;
    {
      std::lock_guard<std::mutex> lk(S->mu);
      if (S->queue.empty()) break;
      r = S->queue.front().get();
    }
    if (S->free_slots.empty() || (int)sizing_now.size() >= S->max_admit) break;
    r->slot = S->free_slots.back();
    const int rc = admit_static(S, *r, (int)sizing_now.size());
    if (rc > 1) return rc;
    if (rc == 1) {
      r->slot = -1;
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
    S->free_slots.pop_back();
    sizing_now.push_back(r);
    std::lock_guard<std::mutex> lk(S->mu);
    S->live.push_back(std::move(S->queue.front()));
    S->queue.pop_front();
  }
  // the batch: one pending omega per active solve, grouped by species count
  std::vector<Request *> act;
  for (auto &p : S->live)
    if (p->state == 2 && !p->m.done) act.push_back(p.get());
  std::stable_sort(act.begin(), act.end(), [](Request *a, Request *b) { return a->nspec < b->nspec; });
  if ((int)act.size() > S->max_batch) act.resize(S->max_batch);
  const int nb = (int)act.size();
  const double t0 = now_ms();
  for (int i = 0; i < nb; ++i) {
    const cplx w = act[i]->m.want();
    S->om_h[i] = make_double2(w.real(), w.imag());
    S->pidx_h[i] = act[i]->slot;
  }
  if (nb > 0) {
    SCHK(hipMemcpyAsync(S->om_d, S->om_h, sizeof(double2) * nb, hipMemcpyHostToDevice, S->st));
    SCHK(hipMemcpyAsync(S->pidx_d, S->pidx_h, sizeof(int) * nb, hipMemcpyHostToDevice, S->st));
    int i0 = 0;
    while (i0 < nb) {
      int i1 = i0;
      const int ns = act[i0]->nspec;
      bool gen = false;
      for (; i1 < nb && act[i1]->nspec == ns; ++i1)
        if (act[i1]->general) gen = true;
      const int n = i1 - i0;
      SCHK(launch_sweep(S->st, 64, SW_W, SW_RC, S->probs_d, n, S->om_d + i0, nullptr, S->chiA,
                        nullptr, S->pidx_d + i0));
      SCHK(launch_resonant(S->st, 64, RS_RB, S->probs_d, ns, n, S->om_d + i0, nullptr, S->chiB,
                           nullptr, gen, S->pidx_d + i0));
      SCHK(launch_assemble(S->st, S->probs_d, ns, n, SW_W * SW_RC, RS_RB, S->om_d + i0, nullptr,
                           S->chiA, S->chiB, nullptr, nullptr, nullptr, 0, nullptr, S->det_d + i0,
                           nullptr, nullptr, nullptr, S->pidx_d + i0));
      i0 = i1;
    }
    SCHK(hipMemcpyAsync(S->det_h, S->det_d, sizeof(double2) * nb, hipMemcpyDeviceToHost, S->st));
  }
  SCHK(hipStreamSynchronize(S->st));
  S->passes += 1;
  S->evals += nb;
  S->busy_ms += now_ms() - t0;
  // int_ee and max|J_n| of this pass's admissions are on the host now
  for (Request *r : sizing_now) {
    const double *row = S->stage_h + (size_t)r->stage * S->stage_row;
    r->sized.assign(row, row + (size_t)r->nspec * (1 + r->ntry));
  }
  // stage 2 for every sized request, oldest first, while space lasts
  for (auto it = S->live.begin(); it != S->live.end();) {
    Request &r = **it;
    if (r.state != 1) {
      ++it;
      continue;
    }
    const int rc = admit_k(S, r);
    bool drop = false;
    if (rc == 1) {
      bool any_active = false;
      for (auto &p : S->live) any_active = any_active || p->state == 2;
      if (any_active) break;  // space frees as solves finish
      drop = true;            // nothing can free space: the request does not fit the arena
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
  // feed the results; retire finished solves
  for (int i = 0; i < nb; ++i) act[i]->m.feed(cplx(S->det_h[i].x, S->det_h[i].y));
  for (auto it = S->live.begin(); it != S->live.end();) {
    Request &r = **it;
    if (r.state == 2 && r.m.done) {
      done.push_back(Result{r.tag, r.m.result(), r.m.iter, ALPSG_OK, now_ms() - r.t_submit});
      release(S, r);
      it = S->live.erase(it);
    } else {
      ++it;
    }
  }
  post_results(S, done);
  return ALPSG_OK;
}

void service_loop(alpsg_server *S) {
  (void)hipSetDevice(0);
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(S->mu);
      S->cv.wait(lk, [&] { return S->stop || !S->queue.empty() || !S->live.empty(); });
      if (S->stop && S->queue.empty() && S->live.empty()) return;
    }
    const int rc = server_pass(S);
    if (rc) {
      std::lock_guard<std::mutex> lk(S->mu);
      S->failed = true;
      S->fail_msg = alpsg_last_error();
      S->cv_done.notify_all();
      return;
    }
  }
}

}  // namespace

extern "C" {

int alpsg_server_create(double hbm_bytes, int max_slots, alpsg_server **out) {
  if (!out) return alpsg_internal_fail(ALPSG_E_ARG, "null argument");
  *out = nullptr;
  if (max_slots < 1 || max_slots > 65536) return alpsg_internal_fail(ALPSG_E_ARG, "max_slots out of range 1..65536");
  if (!(hbm_bytes >= (double)(1 << 20)) || !(hbm_bytes <= 256e9))
    return alpsg_internal_fail(ALPSG_E_ARG, "HBM budget out of range 1 MiB..256 GB");
  std::unique_ptr<alpsg_server> S(new alpsg_server());
  S->max_slots = max_slots;
  S->max_batch = max_slots;
  // on failure below, the destructor path of alpsg_server_destroy frees what exists
  struct Guard {
    std::unique_ptr<alpsg_server> &p;
    ~Guard() {
      if (p) alpsg_server_destroy(p.release());
    }
  } guard{S};
  SCHK(hipStreamCreateWithFlags(&S->st, hipStreamNonBlocking));
  const size_t cap = (size_t)hbm_bytes / ALIGN * ALIGN;
  SCHK(hipMalloc((void **)&S->arena_mem, cap));
  S->arena.init(S->arena_mem, cap);
  SCHK(hipMalloc((void **)&S->probs_d, sizeof(ProbDev) * max_slots));
  S->probs_h.assign(max_slots, ProbDev{});
  for (int i = max_slots - 1; i >= 0; --i) S->free_slots.push_back(i);
  // Miller start indices m(n) = 2*((n+int(sqrt(float(40n))))/2) (ALPS_fns_rel.f90:1622)
  std::vector<int> mt(MTAB_S, 0);
  for (int n = 1; n < MTAB_S; ++n) mt[n] = 2 * ((n + (int)std::sqrt((float)(40 * n))) / 2);
  SCHK(hipMalloc((void **)&S->mtab_d, sizeof(int) * MTAB_S));
  SCHK(hipMemcpy(S->mtab_d, mt.data(), sizeof(int) * MTAB_S, hipMemcpyHostToDevice));
  const size_t mb = (size_t)S->max_batch;
  SCHK(hipMalloc((void **)&S->om_d, sizeof(double2) * mb));
  SCHK(hipMalloc((void **)&S->det_d, sizeof(double2) * mb));
  SCHK(hipMalloc((void **)&S->pidx_d, sizeof(int) * mb));
  SCHK(hipMalloc((void **)&S->chiA, sizeof(double2) * mb * (SW_W * SW_RC) * ALPSG_MAXSPEC * 6));
  SCHK(hipMalloc((void **)&S->chiB, sizeof(double2) * mb * (2 * RS_RB) * ALPSG_MAXSPEC * 6));
  SCHK(hipHostMalloc((void **)&S->om_h, sizeof(double2) * mb, hipHostMallocDefault));
  SCHK(hipHostMalloc((void **)&S->det_h, sizeof(double2) * mb, hipHostMallocDefault));
  SCHK(hipHostMalloc((void **)&S->pidx_h, sizeof(int) * mb, hipHostMallocDefault));
  S->stage_row = STAGE_ROW;
  SCHK(hipHostMalloc((void **)&S->stage_h, sizeof(double) * S->stage_row * alpsg_server::max_admit,
                     hipHostMallocDefault));
  S->th = std::thread(service_loop, S.get());
  *out = S.release();
  return ALPSG_OK;
}

int alpsg_server_submit(alpsg_server *S, const alpsg_plasma_desc *d, double kperp, double kpar,
                        const double *om_guess, const alpsg_root_params *rp, long long tag) {
  if (!S || !d || !om_guess || !rp) return alpsg_internal_fail(ALPSG_E_ARG, "null argument");
  if (!(kpar != 0.0) || !std::isfinite(kpar) || !std::isfinite(kperp))
    return alpsg_internal_fail(ALPSG_E_ARG, "kpar must be finite and nonzero");
  if (rp->secant_method != 2)
    return alpsg_internal_fail(ALPSG_E_UNSUPPORTED, "the server runs secant_osc (secant_method 2)");
  if (rp->numiter < 1 || rp->numiter > 100000) return alpsg_internal_fail(ALPSG_E_ARG, "numiter out of range");
  std::unique_ptr<Request> r(new Request());
  const int rc = prepare_request(d, *r);
  if (rc) return rc;
  r->tag = tag;
  r->kperp = kperp;
  r->kpar = kpar;
  r->guess = cplx(om_guess[0], om_guess[1]);
  r->rp = *rp;
  size_request(*r);
  r->t_submit = now_ms();
  std::lock_guard<std::mutex> lk(S->mu);
  if (S->failed) return alpsg_internal_fail(ALPSG_E_DEVICE, "server stopped: " + S->fail_msg);
  if (S->stop) return alpsg_internal_fail(ALPSG_E_STATE, "server is shutting down");
  S->queue.push_back(std::move(r));
  S->submitted += 1;
  S->cv.notify_all();
  return ALPSG_OK;
}

int alpsg_server_poll(alpsg_server *S, int max_results, int wait, long long *tags, double *roots,
                      int *iters, int *status, double *latency_ms, int *n_out) {
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
}

int alpsg_server_stats(alpsg_server *S, long long *counts, double *busy_ms) {
  if (!S) return alpsg_internal_fail(ALPSG_E_ARG, "null handle");
  std::lock_guard<std::mutex> lk(S->mu);
  if (counts) {
    counts[0] = S->submitted;
    counts[1] = S->finished;
    counts[2] = S->passes;
    counts[3] = S->evals;
    counts[4] = (long long)S->live.size();
    counts[5] = (long long)S->queue.size();
    counts[6] = (long long)S->arena.cap;
    counts[7] = (long long)S->arena.free_bytes();
  }
  if (busy_ms) *busy_ms = S->busy_ms;
  return ALPSG_OK;
}

int alpsg_server_destroy(alpsg_server *S) {
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
  if (S->st) (void)hipStreamDestroy(S->st);
  delete S;
  return ALPSG_OK;
}

}  // extern "C"
