// Host-side tables of a relativistic species and the Gamma(m) table of the
// relativistic Landau term, shared by a handle (alpsg_create) and the root
// server (its plasma blocks).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/alpsg.h"

struct RelHost {
  // the (Gamma, pparbar) grid of derivative_f0_rel, point ig*(npb+1)+ip
  std::vector<double> G, PB, F;
  std::vector<double2> DF;  // (df/dGamma, df/dpparbar)
  std::vector<int2> lim;    // first and last index of each row inside the subluminal cone
  std::vector<double> gpp;  // Gamma at the (pperp, ppar) points, [iperp*(npar+1)+ipar]
  double gmin = 0.0, gmax = 0.0;  // the smallest and largest of them
  bool gmono = false;             // non-decreasing in iperp along every ppar column
  double dg = 0.0, dpb = 0.0;
  double ee = 0.0;  // int_ee_rel (ALPS_fns_rel.f90:1097-1215), omega-independent
};

// Tables of relativistic species s, index r among the relativistic species,
// with pe[0..nperp], pa[0..npar] its (pperp, ppar) grid.  Returns ALPSG_OK or
// ALPSG_E_ARG (pparbar rows that do not increase).
int rel_host_tables(const alpsg_plasma_desc *d, int s, int r, int nspec_rel, const double *pe,
                    const double *pa, RelHost &out);

// Juettner fits in Gamma (types 4, 5) on the Gamma rows 0..ngamma of species
// s in the Maxwell form c0 exp(-c1 (p - c2)^2) of the fit table: c[row][fit][FIT_STRIDE].
void rel_fit_rows(const alpsg_plasma_desc *d, int s, int r, int nspec_rel, double *c);

// Gamma(m), m = 0..GTAB_N-1, then Fact(0..20) (ALPS_fns_rel.f90:1738-1791, :1557-1574),
// then the CBESSJ series coefficients 1/(Fact(k) Gamma(nu+k+1)), nu = 0..RFG_NU-1.
void rel_gamma_table(std::vector<double> &gt);
