/*
 * alps_oracle.h -- CPU restatement of the ALPS dispersion-tensor path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load it.  The product (alps_amd/) never links or calls it.
 *
 * Layout of every array is plain C row-major with the index order of the
 * reference's Fortran arrays (species outermost here; see the PP/F0/DF0
 * macros in alps_oracle.c).  Complex numbers are C99 double complex.
 */
#ifndef ALPS_ORACLE_H
#define ALPS_ORACLE_H

#include <complex.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAXSPEC 8
#define ORC_MAXFITS 4

typedef struct {
  /* sizes */
  int nspec, nperp, npar;
  /* grid and distribution: pp[nspec][nperp+1][npar+1][2], f0[nspec][nperp+1][npar+1],
     df0[nspec][nperp+1][npar+1][2] (only interior points are meaningful) */
  double *pp;
  double *f0;
  double *df0;
  /* species: density, charge, mass */
  double ns[ORC_MAXSPEC], qs[ORC_MAXSPEC], ms[ORC_MAXSPEC];
  double vA;
  /* numerics (ALPS_var.f90 names) */
  int kperp_norm;             /* logical kperp_norm */
  int positions_principal;
  int n_resonance_interval;
  double Tlim;
  double Bessel_zero;
  /* analytic continuation: ACmethod 1 (fits) or 2 (Chebyshev) */
  int ACmethod[ORC_MAXSPEC];
  int n_fits[ORC_MAXSPEC];
  int fit_type[ORC_MAXSPEC][ORC_MAXFITS];
  double perp_correction[ORC_MAXSPEC][ORC_MAXFITS];
  /* param_fit[nspec][nperp+1][5][ORC_MAXFITS] */
  double *param_fit;
  int poly_kind[ORC_MAXSPEC];
  int poly_order[ORC_MAXSPEC];
  int logfit[ORC_MAXSPEC];
  double poly_log_max[ORC_MAXSPEC];
  /* poly_fit_coeffs[nspec][nperp+1][poly_stride] */
  double *poly_fit_coeffs;
  int poly_stride;
  /* wavevector state */
  double kperp, kpar;
  int nmax[ORC_MAXSPEC];
  /* bessel[s] : (nmax[s]+3) x (nperp+1), row n+1 holds order n (n = -1 .. nmax+1) */
  double *bessel[ORC_MAXSPEC];
  /* root-search controls */
  int numiter;
  double D_threshold, D_prec, D_gap, D_tol;
  int secant_method;
  int nproc;                  /* rank count whose nmax bump is emulated (1 = none) */
  /* statistics */
  long long n_disp_calls;
  int n_threads;              /* OpenMP threads used inside disp (0 = default) */
  /* closed-form species (use_bM = T) and their &bM_spec parameters
     (ALPS_var.f90:332-350; defaults ALPS_io.f90:145-150) */
  int usebM[ORC_MAXSPEC];
  int bM_nmax[ORC_MAXSPEC];
  double bM_Bessel_zero[ORC_MAXSPEC], bM_beta[ORC_MAXSPEC], bM_alpha[ORC_MAXSPEC];
  double bM_pdrift[ORC_MAXSPEC];
  /* relativistic species (ALPS_var.f90: relativistic, ngamma, npparbar, the
     *_rel grids of derivative_f0_rel) and the upper bound of param_fit's
     second dimension, max(nperp, ngamma) (0: nperp) */
  int relativistic[ORC_MAXSPEC];
  int ngamma, npparbar;
  int param_np;
  /* per species (NULL unless relativistic): [ngamma+1][npparbar+1], df0_rel [..][..][2] */
  double *gamma_rel[ORC_MAXSPEC], *pparbar_rel[ORC_MAXSPEC], *f0_rel[ORC_MAXSPEC],
      *df0_rel[ORC_MAXSPEC];
  /* parallel current density per species (derivative_f0, ALPS_fns.f90:157-190) */
  double current_int[ORC_MAXSPEC];
} orc_plasma;

/* Branch counters incremented by the hot-path restatement (test instrumentation). */
enum {
  ORC_BR_NONRES = 0,     /* full_integrate without resonance */
  ORC_BR_RES_POS,        /* resonant, Im om > 0 */
  ORC_BR_RES_NEG,        /* resonant, Im om < 0 (2 x Landau) */
  ORC_BR_RES_REAL,       /* resonant, Im om = 0 (1 x Landau) */
  ORC_BR_EDGE_LOW,       /* integrate_res edge fall-back, low side */
  ORC_BR_EDGE_HIGH,      /* integrate_res edge fall-back, high side */
  ORC_BR_PV_TLIM,        /* |Im p_res| > Tlim: Eq. (3.5) ladder */
  ORC_BR_PV_ANALYTIC,    /* |Im p_res| <= Tlim: Eq. (3.6) analytic form */
  ORC_BR_TINY_REST,      /* tiny-rest integration */
  ORC_BR_LANDAU_ZERO,    /* landau_integrate early return */
  ORC_BR_REL_RES,        /* relativistic: integrate_res_rel */
  ORC_BR_REL_PV_TLIM,    /* relativistic: principal value, Tlim ladder */
  ORC_BR_REL_PV_ANALYTIC,/* relativistic: principal value, Eq. (3.7) form */
  ORC_BR_REL_TINY,       /* relativistic: tiny rest */
  ORC_BR_REL_EDGE,       /* relativistic: window at the edge of the subluminal row (no PV) */
  ORC_BR_REL_LANDAU,     /* relativistic: Landau row inside the cone */
  ORC_NBRANCH
};
extern long long orc_branch_counts[ORC_NBRANCH];
void orc_reset_branch_counts(void);

/* Bessel functions (ALPS_fns_rel.f90:1580-1735). */
double orc_bessj(int n, double x);
double orc_bessj0(double x);
double orc_bessj1(double x);

/* Setup (ALPS_fns.f90:34-118, 3971-4075, 4214-4255). */
void orc_derivative_f0(orc_plasma *P);
int orc_determine_nmax(orc_plasma *P);
int orc_set_k(orc_plasma *P, double kperp, double kpar);
void orc_free_k(orc_plasma *P);

/* Hot path pieces (ALPS_fns.f90:252-1812, ALPS_analyt.f90:32-363). */
void orc_determine_resonances(const orc_plasma *P, int sproc, double complex om, int nn,
                              int *found_plus, int *found_minus);
double complex orc_full_integrate(const orc_plasma *P, int sproc, double complex om, int nn,
                                  int mode, int found_res);
double complex orc_integrate(const orc_plasma *P, int sproc, double complex om, int nn,
                             int mode, int iparmin, int iparmax);
double complex orc_integrate_res(const orc_plasma *P, int sproc, double complex om, int nn,
                                 int mode);
double complex orc_landau_integrate(const orc_plasma *P, int sproc, double complex om, int nn,
                                    int mode);
double complex orc_funct_g(const orc_plasma *P, int sproc, double ppar_real, int iperp,
                           double complex om, int nn, int mode);
double complex orc_int_ee(const orc_plasma *P, int sproc);
double complex orc_eval_fit(const orc_plasma *P, int is, int iperp, double complex ppar_valC);
/* distribution_analyt (distribution/distribution_analyt.f90:30-96), as shipped */
double complex orc_distribution_analyt(int is, double pperp, double complex ppar);

/* disp(om): det D; optional chi0 [nspec*9], chi0_low [nspec*27] (n = -1,0,1), wave [9],
   all row-major [s][i][j] / [s][i][j][n+1]. */
double complex orc_disp(orc_plasma *P, double complex om, double complex *chi0,
                        double complex *chi0_low, double complex *wave);

/* Closed-form bi-Maxwellian / cold susceptibility (ALPS_NHDS.f90, alps_oracle_nhds.c).
   chi [3][3], chi_low [3][3][3] (n = -1, 0, 1 last); j is the 1-based species. */
double orc_bessi(int n, double x);
double orc_besselI(int n, double x);
void orc_wofz(double xi, double yi, double *u, double *v, int *flag);
int orc_nhds_nmaxrun(const orc_plasma *P, int j, double kperp);
int orc_calc_chi(const orc_plasma *P, int j, double kz, double kperp, double complex x,
                 double complex chi[3][3], double complex chi_low[3][3][3]);

/* Relativistic path (ALPS_fns_rel.f90, alps_oracle_rel.c).  orc_rel_error is
   set to 8 where the reference calls alps_error(8). */
extern int orc_rel_error;
double orc_gamma_fn(double x);
double complex orc_cbessj(double complex z, int nu);
double complex orc_integrate_res_rel(const orc_plasma *P, int s, double complex om, int nn, int mode);
double complex orc_landau_integrate_rel(const orc_plasma *P, int s, double complex om, int nn, int mode);
double complex orc_int_ee_rel(const orc_plasma *P, int s);
void orc_determine_resonances_rel(const orc_plasma *P, int s, double complex om, int nn, int *fp,
                                  int *fm);

/* current_int of derivative_f0 (ALPS_fns.f90:157-190), from f0 */
void orc_current_int(orc_plasma *P);

/* calc_eigen (ALPS_fns.f90:2605-2899) at omega, which must be the last
   argument of orc_disp (its chi0 and wave are passed in, row-major as orc_disp
   writes them).  Outputs in the Fortran layouts: ef[3], bf[3], Us[3*nspec]
   (Us(1:3,1:nspec)), ds[nspec], Ps[nspec], Ps_split[4*nspec]
   (Ps_split(1:4,1:nspec)), *ewave.  Calls orc_disp twice when heat != 0. */
void orc_calc_eigen(orc_plasma *P, double complex omega, const double complex *chi0,
                    const double complex *wave, int eigen, int heat, double complex *ef,
                    double complex *bf, double complex *Us, double complex *ds, double *Ps,
                    double *Ps_split, double *ewave);

/* Root finders (ALPS_fns.f90:1815-2195). Return the number of iterations.
   The *_fn forms take the function whose root they seek, as the reference's
   rtsec(func, ...) does (ALPS_fns.f90:2105); the plain forms pass disp. */
typedef double complex (*orc_fn)(void *ctx, double complex om);
typedef struct {
  int numiter;
  double D_threshold, D_prec, D_tol;
} orc_root_ctl;
int orc_secant_fn(orc_fn fn, void *ctx, const orc_root_ctl *c, double complex *om);
int orc_secant_osc_fn(orc_fn fn, void *ctx, const orc_root_ctl *c, double complex *om);
double complex orc_rtsec_fn(orc_fn fn, void *ctx, const orc_root_ctl *c, double complex xin,
                            int *iflag);
int orc_secant(orc_plasma *P, double complex *om);
int orc_secant_osc(orc_plasma *P, double complex *om);
double complex orc_rtsec(orc_plasma *P, double complex xin, int *iflag);

/* Closed-form test functions of the solvers' pins (test infrastructure):
   kind 0 a (om - r1), kind 1 a (om^2 - r2), kind 2 a (om - r1)(om - r2) /
   (om - pole), kind 3 a exp(om); complex parameters as (re, im).  Each call
   records its argument in traj [2*cap] and counts itself in n. */
typedef struct {
  int kind;
  double a[2], r1[2], r2[2], pole[2];
  int n, cap;
  double *traj;
} orc_test_ctx;
double complex orc_test_fn(void *ctx, double complex om);
int orc_solve_test(int method, orc_test_ctx *t, const orc_root_ctl *c, double *om);

/* Map search (ALPS_fns.f90:3595-3966).  cal/val are [nr][ni] row-major.
   iroots is [numroots][2] (ir, ii, 1-based).  Returns the number of minima recorded. */
int orc_map_search(orc_plasma *P, double omi, double omf, double gami, double gamf,
                   int nr, int ni, int loggridw, int loggridg,
                   double complex *om_out, double complex *cal, double *val,
                   int numroots, int *iroots);
int orc_find_minima(const double *val, int nr, int ni, int numroots, int *iroots);

/* om_scan (ALPS_fns.f90:2198-2600) without eigenfunction/heating output.
   roots[nroots] in/out.  Output arrays sized [nt+1]: row 0 is the initial state.
   scan_type 0..4; diff, diff2 as computed by ALPS_io.f90:482-548.
   Returns 0, or 9 when every root has been dropped (alps_error(9)). */
int orc_om_scan(orc_plasma *P, int scan_type, int log_scan, double diff, double diff2,
                int n_out, int n_res, int nroots, double complex *roots,
                double *kperp_out, double *kpar_out, double complex *roots_out);

/* One wavevector scan of &scan_input (ALPS_var.f90 scanner type): scan_type
   0..4, log or linear steps, diff/diff2 as scan_read computes them
   (ALPS_io.f90:482-548), n_out output steps of n_res sub-steps each. */
typedef struct {
  int type, log_scan;
  double diff, diff2;
  int n_out, n_res;
} orc_scan;

/* om_double_scan (ALPS_fns.f90:2900-3591): scan 1 outer, scan 2 inner from
   every output point of scan 1.  roots[nroots] in/out (omlast at the end).
   Output point (i1, i2), i1 = 0..n_out1, i2 = 0..n_out2, is row
   i1*(n_out2+1)+i2 of kperp_out/kpar_out, and row*nroots+in of roots_out
   (wroots(in) as written to the .scan file) and written (1 when the reference
   writes that row for root in, i.e. the root was still followed).  With eigen
   or heat, calc_eigen's outputs for each written (row, root) go to ef[3],
   bf[3], Us[3*nspec], ds[nspec], Ps[nspec], Ps_split[4*nspec], W_EM[1] blocks
   (may be NULL otherwise).  Returns 0, or 9 for alps_error(9). */
int orc_om_double_scan(orc_plasma *P, const orc_scan *s1, const orc_scan *s2, int nroots,
                       double complex *roots, int eigen, int heat, double *kperp_out,
                       double *kpar_out, double complex *roots_out, int *written,
                       double complex *ef, double complex *bf, double complex *Us,
                       double complex *ds, double *Ps, double *Ps_split, double *W_EM);

/* determine_param_fit (ALPS_analyt.f90:365-728): the Levenberg-Marquardt fits
   of every ACmethod-1 species, row by row from the initial guesses in
   param_fit[s][0] (each row starts from the previous row's result), the
   Chebyshev coefficients of ACmethod-2 species, and the reset of use_bM,
   ACmethod-0 and ACmethod-2 species (zero param_fit, fit_type 1,
   perp_correction 1).  Updates P in place; qualitytotal and steps_total
   (both may be NULL) get the summed least-squares quality and LM steps.
   Returns 0, 2 for alps_error(2) (fewer points than parameters), 3 for a
   singular fit matrix (the reference stops), 10 for alps_error(10). */
int orc_determine_param_fit(orc_plasma *P, int maxsteps, double lambda0, double lambdafac, double eps,
                            double *qualitytotal, int *steps_total);

#ifdef __cplusplus
}
#endif
#endif
