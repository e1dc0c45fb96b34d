/*
 * alpsg.h -- C-ABI of the MI355X (gfx950) dispersion-tensor engine.
 *
 * The engine replaces the harmonic loop of ALPS's disp() and the root
 * searches that call it.  Every entry point is plain C: pointers, sizes and
 * an int status (0 = ok, never aborts).  Complex numbers are interleaved
 * (re, im) doubles, i.e. Fortran complex(c_double_complex) / C double _Complex.
 * Multi-dimensional arrays use the Fortran (column-major) layout of the
 * reference's own arrays, so a Fortran caller passes its alps_var arrays
 * unchanged.  One handle owns its device copies and one HIP stream; calls on
 * one handle are serialised by the caller.
 */
#ifndef ALPSG_H
#define ALPSG_H

#ifdef __cplusplus
extern "C" {
#endif

#define ALPSG_OK 0
#define ALPSG_E_ARG 1         /* invalid argument / shape */
#define ALPSG_E_UNSUPPORTED 2 /* feature outside the engine (Juettner fit 3, ...) */
#define ALPSG_E_DEVICE 3      /* HIP runtime error */
#define ALPSG_E_STATE 4       /* call order (e.g. disp before set_k) */
#define ALPSG_E_NOTFOUND 5    /* no root survived a scan (reference alps_error(9)) */
#define ALPSG_MAXITER 6       /* a served solve used up numiter iterations without
                                 |D| < D_threshold; its root is the best point
                                 found ("Maximum iteration reached",
                                 ALPS_fns.f90:2093-2099) */

#define ALPSG_MAXSPEC 8
#define ALPSG_MAXFITS 4

typedef struct alpsg_ctx *alpsg_handle;

/* Plasma description: the alps_var globals disp() reads (ALPS_var.f90:19-427).
   Array layouts are the Fortran ones:
     pp(1:nspec,0:nperp,0:npar,1:2)        ALPS_com.f90:202
     f0(1:nspec,0:nperp,0:npar)            ALPS.f90:72
     df0(1:nspec,1:nperp-1,1:npar-1,1:2)   ALPS_com.f90:200   (used when f0 == NULL)
     fit_type(1:nspec,1:max_fits), perp_correction(1:nspec,1:max_fits)
     param_fit(1:nspec,0:param_fit_nperp,1:5,1:max_fits)  ALPS_io.f90:188
     poly_fit_coeffs(1:nspec,0:nperp,0:poly_max_order)    ALPS_io.f90:270  */
typedef struct alpsg_plasma_desc {
  int nspec, nperp, npar;
  const double *pp;
  const double *f0;  /* may be NULL if df0 is given */
  const double *df0; /* may be NULL if f0 is given */
  const double *ns, *qs, *ms; /* [nspec] */
  double vA;
  int kperp_norm;
  int positions_principal;
  int n_resonance_interval;
  double Tlim;
  double Bessel_zero;
  const int *usebM;        /* [nspec] or NULL; nonzero: closed-form (NHDS) species */
  const int *relativistic; /* [nspec] or NULL; nonzero: relativistic species (see below) */
  const int *ACmethod;     /* [nspec]: 1 fits, 2 Chebyshev */
  const int *n_fits;       /* [nspec] */
  int max_fits;
  const int *fit_type;
  const double *perp_correction;
  const double *param_fit;
  int param_fit_nperp; /* upper bound of param_fit's 2nd dimension, max(nperp,ngamma) */
  const int *poly_kind, *poly_order, *logfit; /* [nspec] (may be NULL if no ACmethod 2) */
  const double *poly_log_max;                 /* [nspec] */
  const double *poly_fit_coeffs;
  int poly_max_order;
  /* closed-form species (usebM[s] != 0): the &bM_spec parameters
     (ALPS_var.f90:332-350), [nspec] each; may be NULL when no species uses it.
     bM_beta 0 selects the cold-plasma susceptibility. */
  const int *bM_nmax;
  const double *bM_Bessel_zero, *bM_beta, *bM_alpha, *bM_pdrift;
  /* relativistic species (relativistic[s] != 0): the (Gamma, pparbar) grid as
     derivative_f0_rel (ALPS_fns_rel.f90:19-408) leaves it, over the
     relativistic species only, in species order (ALPS_com.f90:247-250):
       gamma_rel, pparbar_rel, f0_rel (1:nspec_rel,0:ngamma,0:npparbar)
       df0_rel                         (1:nspec_rel,0:ngamma,0:npparbar,1:2)
     Their fits must be fit types 4 or 5 (Juettner in Gamma), with
     param_fit_nperp >= ngamma. */
  int ngamma, npparbar;
  const double *gamma_rel, *pparbar_rel, *f0_rel, *df0_rel;
  /* parallel current density per species (current_int, ALPS_fns.f90:157-190),
     [nspec] or NULL to compute it from f0; needed by alpsg_calc_eigen only. */
  const double *current_int;
} alpsg_plasma_desc;

/* Root-search controls (&system namelist, ALPS_var.f90:147-165). */
typedef struct alpsg_root_params {
  int secant_method; /* 0 secant, 1 rtsec, 2 secant_osc */
  int numiter;
  double D_threshold, D_prec, D_gap, D_tol;
} alpsg_root_params;

/* Error text of the last failing call on this thread. */
const char *alpsg_last_error(void);

/* Fit controls (&system namelist, ALPS_var.f90:288-297). */
typedef struct alpsg_fit_params {
  int maxsteps_fit;          /* 500 */
  double lambda_initial_fit; /* 1.0 */
  double lambdafac_fit;      /* 10.0 */
  double epsilon_fit;        /* 1e-8 */
} alpsg_fit_params;

/* determine_param_fit (ALPS_analyt.f90:365-728) on the device, called before
   alpsg_create as ALPS.f90:80 calls it before the disp() path: the
   Levenberg-Marquardt fit of every ACmethod-1 species row by row (pperp rows,
   or Gamma rows of a relativistic species), starting from the guesses in
   desc->param_fit(s,0,:,:); the Chebyshev coefficients of ACmethod-2 species;
   and the reset of use_bM, ACmethod-0 and ACmethod-2 species (param_fit 0,
   fit_type 1, perp_correction 1).  Needs desc->f0 (and the *_rel grids for
   relativistic species).  Outputs in the layouts of the description:
   param_fit (1:nspec,0:param_fit_nperp,1:5,1:max_fits), fit_type and
   perp_correction (1:nspec,1:max_fits), poly_fit_coeffs
   (1:nspec,0:nperp,0:poly_max_order; may be NULL without ACmethod 2); they may
   alias the description's arrays.  Rows of param_fit past a species' last
   fitted row are left as they are.  quality (may be NULL) gets the summed
   least-squares quality.  fp NULL selects the defaults above.  ALPSG_E_ARG
   for alps_error(2) (fewer points than parameters), a singular fit matrix
   (where the reference stops) or alps_error(10) (poly_kind other than 1). */
int alpsg_determine_param_fit(const alpsg_plasma_desc *desc, const alpsg_fit_params *fp,
                              double *param_fit, int *fit_type, double *perp_correction,
                              double *poly_fit_coeffs, double *quality);

/* Create a handle: uploads pp and f0 (or df0), builds df0 on the device with
   the central differences of derivative_f0 (ALPS_fns.f90:103-115), the
   omega-independent int_ee term (ALPS_fns.f90:1457-1555) and the fit tables.
   Replaces the state set up by ALPS.f90:70-89 for the disp() path. */
int alpsg_create(const alpsg_plasma_desc *desc, alpsg_handle *out);

int alpsg_destroy(alpsg_handle h);

/* Set the wavevector: determine_nmax (ALPS_fns.f90:3971-4075) unless nmax_in
   is non-NULL, then the Bessel table (determine_bessel_array,
   ALPS_fns.f90:4214-4255) and the pperp moment tables, all kept in HBM.
   nmax_out [nspec] may be NULL. */
int alpsg_set_k(alpsg_handle h, double kperp, double kpar, const int *nmax_in, int *nmax_out);

/* Batched disp(om) (ALPS_fns.f90:252-636): det D for n omegas.
   om, det: [2n].  Optional outputs (NULL to skip), per omega in the Fortran
   layout of the reference globals:
     chi0     (1:nspec,1:3,1:3)        -> [n][nspec*9*2]
     chi0_low (1:nspec,1:3,1:3,-1:1)   -> [n][nspec*27*2]
     wave     (1:3,1:3)                -> [n][9*2]                      */
int alpsg_disp_batch(alpsg_handle h, int n, const double *om, double *det, double *chi0,
                     double *chi0_low, double *wave);

/* The same evaluation with one launch per omega (the plain path). */
int alpsg_disp_plain(alpsg_handle h, int n, const double *om, double *det);

/* alpsg_disp_batch (det only) without any device allocation: no counterpart
   in the reference; for callers that cannot allocate (code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc).  The
   kernels and launch shape are those of alpsg_disp_batch; the buffers beyond
   om and det come from a fixed 32 MiB device array per GPU, held until the
   process ends.  A batch that needs more (n beyond a few thousand omegas,
   depending on nspec and the launch shape) is refused with ALPSG_E_ARG;
   handles with relativistic or closed-form species with
   ALPSG_E_UNSUPPORTED.  Calls on the handle's device only; calls that
   overlap on one GPU share the array and race, so the caller serialises
   them.  The handle itself (alpsg_create, alpsg_set_k) allocates as usual. */
int alpsg_disp_batch_noalloc(alpsg_handle h, int n, const double *om, double *det);

/* Batched root refinement for nroots independent roots at the handle's k with
   the solver rp->secant_method selects: secant (ALPS_fns.f90:1815-1917),
   rtsec (:2105-2195) or secant_osc (:1919-2101).  om [2*nroots] in/out.
   D_final [2*nroots] (optional) gets disp at the returned root (the "final instance of disp", ALPS_fns.f90:2523,
   3847); iters [nroots] (optional) the iteration count. */
int alpsg_secant_batch(alpsg_handle h, const alpsg_root_params *rp, int nroots, double *om,
                       double *D_final, int *iters);

/* map_search (ALPS_fns.f90:3595-3788): det D on the nr x ni omega grid in one
   launch, log10|D| with the reference's NaN/Inf sentinels, local minima
   (find_minima, ALPS_fns.f90:3860-3966, at most numroots recorded).
   om_out, cal_out: [2*nr*ni] and val_out [nr*ni] in Fortran order (ir fastest)
   (each may be NULL); iroots [2*numroots] gets (ir, ii) 1-based; *nfound the
   number of minima (may exceed numroots, as counted by the reference). */
int alpsg_map_search(alpsg_handle h, double omi, double omf, double gami, double gamf, int nr,
                     int ni, int loggridw, int loggridg, double *om_out, double *cal_out,
                     double *val_out, int numroots, int *iroots, int *nfound);

/* map_search with its refinement, the whole of the reference's map_search
   (ALPS_fns.f90:3595-3788: the map, find_minima, then refine_guess at :3781,
   :3793-3856) in one call: refine_guess's solver (rp->secant_method) runs on
   the first min(nfound, numroots) minima in find_minima's order, starting on
   the device as soon as the minimum flags reach the host, while D and
   log10|D| travel to the host beside the secant rounds.  Arguments as
   alpsg_map_search (iroots, nfound may be NULL); roots [2*numroots] gets the
   refined roots, D_final [2*numroots] and iters [numroots] (both optional)
   as alpsg_secant_batch.  The results equal alpsg_map_search followed by
   alpsg_secant_batch on the minima's grid frequencies; alpsg_last_root_times
   then counts from the start of this call, and alpsg_last_kernel_ms /
   alpsg_last_kernel_split report the map's batch. */
int alpsg_map_search_refine(alpsg_handle h, const alpsg_root_params *rp, double omi, double omf,
                            double gami, double gamf, int nr, int ni, int loggridw, int loggridg,
                            double *om_out, double *cal_out, double *val_out, int numroots,
                            int *iroots, int *nfound, double *roots, double *D_final, int *iters);

/* om_scan (ALPS_fns.f90:2198-2600) without eigenfunction/heating output:
   follows nroots roots along the wavevector path; diff/diff2 are the step
   sizes computed by scan_read (ALPS_io.f90:482-548).  roots [2*nroots] in/out.
   Outputs have n_out+1 rows (row 0 = the starting point):
   kperp_out, kpar_out [n_out+1]; roots_out [(n_out+1)*nroots*2]. */
int alpsg_om_scan(alpsg_handle h, const alpsg_root_params *rp, int scan_type, int log_scan,
                  double diff, double diff2, int n_out, int n_res, int nroots, double *roots,
                  double *kperp_out, double *kpar_out, double *roots_out);

/* One wavevector scan of a &scan_input namelist: scan_type 0..4 (k1_k2,
   theta, kcstq, kperp, kpara), log or linear steps, diff/diff2 as scan_read
   computes them (ALPS_io.f90:482-548), n_out output steps of n_res sub-steps. */
typedef struct alpsg_scan_params {
  int scan_type, log_scan;
  double diff, diff2;
  int n_out, n_res;
} alpsg_scan_params;

/* om_double_scan (ALPS_fns.f90:2900-3591): scan s1 outer, and from each of
   its output points the scan s2 inner; roots are followed with the
   rp->secant_method solver, all roots at a k in one batch.  roots [2*nroots]
   in/out (the roots along scan 1 at its last point).  Output point (i1, i2),
   i1 = 0..s1->n_out, i2 = 0..s2->n_out, is row i1*(s2->n_out+1)+i2:
   kperp_out, kpar_out [rows]; roots_out [rows*nroots*2]; written
   [rows*nroots] = 1 where the reference writes the .scan row of that root
   (the root was still followed).  With eigen/heat, calc_eigen's outputs for
   each (row, root): ef, bf [..][3] complex, Us [..][nspec][3] complex, ds
   [..][nspec] complex, Ps [..][nspec], Ps_split [..][nspec][4], W_EM [..];
   the arrays of the part not asked for may be NULL.  ALPSG_E_ARG for the
   reference's alps_error(5)/(6) (equal scan types, a type-0 scan),
   ALPSG_E_NOTFOUND for alps_error(9). */
int alpsg_om_double_scan(alpsg_handle h, const alpsg_root_params *rp, const alpsg_scan_params *s1,
                         const alpsg_scan_params *s2, int nroots, double *roots, int eigen,
                         int heat, double *kperp_out, double *kpar_out, double *roots_out,
                         int *written, double *ef, double *bf, double *Us, double *ds, double *Ps,
                         double *Ps_split, double *W_EM);

/* calc_eigen (ALPS_fns.f90:2605-2899) for n roots at the handle's k, batched:
   disp at every root (the "final instance of disp" the reference runs before
   it), and with heat != 0 at real(om) and real(om*1.000001).  Outputs per
   root r, in the Fortran layouts of om_scan's .eigen/.heat/.heat_mech rows:
     ef, bf [r][3] complex; Us [r][nspec][3] complex (Us(1:3,1:nspec));
     ds [r][nspec] complex; Ps [r][nspec]; Ps_split [r][nspec][4]
     (Ps_split(1:4,1:nspec)); W_EM [r].
   Arrays for the part not asked for (eigen == 0 or heat == 0) may be NULL. */
int alpsg_calc_eigen(alpsg_handle h, int n, const double *om, int eigen, int heat, double *ef,
                     double *bf, double *Us, double *ds, double *Ps, double *Ps_split,
                     double *W_EM);

/* Launch-shape autotuner.  A shape is 5 ints: lanes per omega in the node
   sweep, its wavefronts per workgroup, its workgroups per omega block
   (harmonic-row chunks), lanes per omega in the resonant kernels, and their
   harmonic blocks.  alpsg_autotune times the instantiated variants on a batch
   of n omegas at the current k (om [2n] or NULL for a built-in spread) and
   keeps the fastest for batches of n's size class (floor(log2 n));
   alpsg_set_shape pins a shape for that class (NULL restores the default);
   alpsg_get_shape reports the shape a batch of n would use. */
int alpsg_autotune(alpsg_handle h, int n, const double *om, int *shape_out);
int alpsg_set_shape(alpsg_handle h, int n, const int *shape);
int alpsg_get_shape(alpsg_handle h, int n, int *shape_out);

/* Diagnostics: nmax per species at the current k and the number of disp
   evaluations done by this handle. */
int alpsg_get_nmax(alpsg_handle h, int *nmax_out);
long long alpsg_eval_count(alpsg_handle h);

/* Wall time (ms, from the start of the last alpsg_secant_batch or
   alpsg_map_search_refine on this handle) at which each of its nroots roots
   finished its secant iterations. */
int alpsg_last_root_times(alpsg_handle h, int nroots, double *ms_out);

/* The last alpsg_om_scan on this handle, per wavevector step (nsteps =
   n_out*n_res): ms_out [nsteps] the wall time (ms from the scan's start) at
   which the step had solved all its roots; iters_out [nsteps*nroots] the
   iteration count of each root's solve at that step (numiter or more: the
   solve ended without |D| < D_threshold), -1 for a root no longer followed.
   Either may be NULL. */
int alpsg_last_scan_steps(alpsg_handle h, int nsteps, int nroots, double *ms_out, int *iters_out);

/* With plain != 0, the root searches of this handle (alpsg_secant_batch,
   alpsg_om_scan, alpsg_om_double_scan) step every
   solve on the host with one disp launch per frequency -- the plain form,
   timed for comparison -- instead of stepping the secant_osc machines on the
   device between the launches of a replayed graph.  The roots are the same. */
int alpsg_set_plain_roots(alpsg_handle h, int plain);

/* on != 0 (the default): in the one-lane-per-omega node sweep of large
   batches, rows whose every node is far from resonance for every omega of a
   wavefront (|m w| <= 0.55 min |kpar p_i -+ n q|) take their node sum from a
   per-k series table in m w, truncated below double rounding, instead of a
   pass over the nodes; on == 0: every row passes over its nodes.  Both give
   the reference's integrate() trapezoid sum (ALPS_fns.f90:799-864) to double
   rounding; the switch exists to compare them.  A k's tables are built by its
   first batch of 64 or more omegas (alpsg_disp_batch, alpsg_map_search,
   alpsg_autotune, alpsg_sweep_work). */
int alpsg_set_sweep_series(alpsg_handle h, int on);

/* The work of the node sweep (k_sweep, in the launch shape of an n-omega
   batch) over these n omegas, counted on the device: work3[0] the (species,
   harmonic, ppar node) units summed node by node, work3[1] the far-row signs
   summed by their series, work3[2] the sum of those series' degrees, each
   summed over the omegas.  The sweep's partial results are discarded. */
int alpsg_sweep_work(alpsg_handle h, int n, const double *om, double *work3);

/* Device time (ms) of the last alpsg_disp_batch / alpsg_map_search kernel
   sweep, measured with HIP events on the handle's stream. */
double alpsg_last_kernel_ms(alpsg_handle h);

/* The same interval split by kernel: ms3[0] the node sweep (k_sweep),
   ms3[1] the resonant-harmonic kernels (k_resonant, all parts), ms3[2] the
   assembly (k_assemble). */
int alpsg_last_kernel_split(alpsg_handle h, double *ms3);

/* Root server: continuous batching of independent secant_osc solves
   (ALPS_fns.f90:1919-2101), each with its own plasma, wavevector and initial
   frequency -- the solve the reference runs once per root and per scan step
   (refine_guess, ALPS_fns.f90:3791-3857; om_scan, :2198-2600) -- served from
   one device.  A service thread admits waiting requests into a fixed HBM
   arena of hbm_bytes (first come first served, by footprint: the plasma's
   momentum tables, then the request's per-k tables once nmax is known), at
   most max_slots solves at a time, and advances every admitted solve by one
   disp() per pass, with solves of uneven grids and nmax in shared launches.
   A request's result does not depend on the other requests in flight.

   Served species: tabulated f0 (non-relativistic or relativistic, as for
   alpsg_create) and closed-form use_bM species.  rp must ask for secant_osc
   (secant_method 2).

   alpsg_server_add_plasma copies a plasma description and returns its id;
   its device tables are built on first use and kept while the arena has room
   (least recently used unused plasmas are evicted when it has not).
   alpsg_server_remove_plasma forgets the id; requests already submitted with
   it still run.  alpsg_server_submit_k queues one solve on a registered
   plasma; alpsg_server_submit_batch queues n of them in one call (all or none;
   plasma_ids, kperp, kpar, tags [n], om_guess [2n], one rp for all);
   alpsg_server_submit does the same for a plasma given inline (copied before
   it returns, freed with the request).

   alpsg_server_poll returns finished solves (waits for one when wait != 0 and
   any are outstanding): tags [max_results], roots [2*max_results], iters,
   status (ALPSG_OK for a converged solve, ALPSG_MAXITER for one that ended at
   numiter with the best point as its root, or ALPSG_E_ARG for a request that
   cannot fit the arena,
   whose Bessel cutoff is not reached, or whose relativistic principal-value
   window leaves the subluminal cone, the reference's alps_error(8)) and
   latency_ms (submit to result);
   any may be NULL.  alpsg_server_stats gives counts[10] = {submitted,
   finished, passes, disp evaluations, admitted, waiting, arena bytes, arena
   bytes in use, plasma uploads, plasma evictions} and times_ms[4] = the
   service thread's time in passes, in admission work, in enqueueing the disp
   launches and in waiting for the device.
   No counterpart in the reference, which solves one root at a time. */
typedef struct alpsg_server alpsg_server;
int alpsg_server_create(double hbm_bytes, int max_slots, alpsg_server **out);
int alpsg_server_add_plasma(alpsg_server *s, const alpsg_plasma_desc *desc, int *plasma_id);
int alpsg_server_remove_plasma(alpsg_server *s, int plasma_id);
int alpsg_server_submit_k(alpsg_server *s, int plasma_id, double kperp, double kpar,
                          const double *om_guess, const alpsg_root_params *rp, long long tag);
int alpsg_server_submit_batch(alpsg_server *s, int n, const int *plasma_ids, const double *kperp,
                              const double *kpar, const double *om_guess,
                              const alpsg_root_params *rp, const long long *tags);
int alpsg_server_submit(alpsg_server *s, const alpsg_plasma_desc *desc, double kperp, double kpar,
                        const double *om_guess, const alpsg_root_params *rp, long long tag);
int alpsg_server_poll(alpsg_server *s, int max_results, int wait, long long *tags, double *roots,
                      int *iters, int *status, double *latency_ms, int *n_out);
int alpsg_server_stats(alpsg_server *s, long long *counts, double *times_ms);
/* The algorithmic units of every disp evaluation the server has run and the
   device time they took (a diagnostic for rooflines, no counterpart in the
   reference): work7[0] the node records of the sweep, sum over evaluations and
   tabulated non-relativistic species of (2 nmax + 1)(npar - 1); work7[1] the
   relativistic points x harmonics, (2 nmax + 1)(nperp - 1)(npar - 1); work7[2]
   the relativistic Gamma rows x harmonics, (2 nmax + 1)(ngamma - 1);
   work7[3..6] the device ms, from events around each pass's launches, of the
   node sweep, the resonant parts (with closed-form species), the relativistic
   kernels and the assembly. */
int alpsg_server_work(alpsg_server *s, double *work7);
int alpsg_server_destroy(alpsg_server *s);

#ifdef __cplusplus
}
#endif
#endif
