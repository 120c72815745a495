/* CPU oracle for the finite-rate chemistry step -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C FP64 restatement of the reference path, used by tests/, by
 * __graft_entry__.smoke() and by bench.py's CPU-baseline leg as the checker
 * and as the timed CPU stand-in for the reference.  Nothing in the product
 * (tiger_nga_amd/) links or calls it.
 *
 * Parts and the reference they restate (paths relative to the reference root):
 *   orc_prodrates   combustion/HYDROGENF.f90:4-277 (PRODRATES), 280-316
 *                   (GETLINDRATECOEFF); same routines in grimech30F.f90 /
 *                   DME39F.f90.  Pinned by golden vectors of the reference
 *                   mechanism files compiled with flang (tests/golden).
 *   orc_thermo      combustion/HYDROGENF.f90:514-673 (COMPTHERMODATA). Pinned
 *                   likewise, plus dh/dT = cp.
 *   orc_rhs         combustion/finitechem.f90:93-175 (finitechem_compute_rhs),
 *                   79-88 (finitechem_Cp).  Pinned by composition with the
 *                   golden production rates and by invariants (sum of mass
 *                   rates = 0 up to binary32 molar masses, enthalpy balance).
 *   orc_integrate   the SUNDIALS CVODE 2.5.0 BDF/Newton/dense solver behind
 *                   FCVODE (finitechem.f90:1014-1034), restated from the
 *                   published CVODE algorithm (Hindmarsh et al., ACM TOMS
 *                   31(3), 2005; CVODE 2.5 user guide).  SUNDIALS is not in the
 *                   reference tree: pinned against scipy's Radau/BDF at tight
 *                   tolerance, closed forms (dt -> 0, frozen cold cells) and
 *                   invariants, not against CVODE's own step sequence.
 *   orc_source_chemistry  finitechem.f90:962-1071 (pack, integrate, unpack,
 *                   heat release and H2O/CO/CO2 source diagnostics).
 *   orc_properties  finitechem.f90:65-74, 324-486 (W_mix, density, Wilke
 *                   viscosity, conductivity / diffusivities, omegamu).  Pinned
 *                   by the golden GETMUCOEFF / GETKOVEREPS values and closed
 *                   forms (pure species, ideal gas).
 *   orc_temperature enthalpy -> temperature Newton inversion (no reference
 *                   routine), pinned by round trips.
 *   orc_jacobian    analytic Jacobian of finitechem_compute_rhs (no reference
 *                   routine; CVODE differences the RHS, finitechem.f90:1028),
 *                   pinned by central differences of orc_rhs.
 *   orc_diffusion   finitechem_diffusion (finitechem.f90:538-923), the
 *                   molar-diffusion correction and enthalpy flux; pinned by
 *                   flux conservation, closed forms on linear fields and
 *                   planted mistakes (tests/test_diffusion.py).
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_H2 = 0, ORC_GRI30 = 1, ORC_DME39 = 2, ORC_ETHYRED = 3 };

typedef struct {
  int nsp, nrxn, nslot, n_third;
} orc_info;

typedef struct {
  double rtol, atol;  /* scalar tolerances (finitechem.f90:266-267 defaults 1e-6, 1e-12) */
  int max_steps;      /* MAX_NSTEPS (finitechem.f90:1025: 10000) */
  int max_order;      /* BDF maximum order (CVODE default 5) */
  int jac;            /* 0: difference-quotient Jacobian (CVODE FCVDENSE, the reference's);
                         1: the analytic Jacobian (orc_jacobian) */
} orc_opts;

/* per-cell integrator counters (same meaning as CVODE's) */
typedef struct {
  int64_t nst, nfe, nfe_dq, nje, nsetups, nni, netf, ncfn, nlu;
  int64_t status; /* 0 ok, <0 failure code (see oracle/cvode_bdf.c) */
} orc_stats;

int orc_mech_info(int mech, orc_info *info);
int orc_molar_mass(int mech, double *W);
const char *orc_species_name(int mech, int k);

/* PRODRATES: C[nsp] kmol/m3, T K, P Pa -> cdot[nsp]; optional per-slot K, W
 * (reference slot order) and third-body sums M. */
int orc_prodrates(int mech, const double *C, double T, double P, double *cdot,
                  double *K, double *W, double *M);
/* COMPTHERMODATA: h[nsp] J/kg, cp[nsp] J/kg/K */
int orc_thermo(int mech, double T, double *h, double *cp);
/* temperature from mixture enthalpy (Newton; no reference routine): returns
 * the iteration count, -1 if not converged, -2 on a bad mechanism id */
#define ORC_T_LOW 50.0
#define ORC_T_HIGH 5000.0
#define ORC_T_MAXIT 50
int orc_temperature(int mech, const double *Y, double hmix, double T0, double *T);
/* finitechem_density / _viscosity / _diffusivity of one cell (Lewis may be NULL) */
int orc_properties(int mech, const double *Y, double Tcell, double P_thermo, const double *Lewis, double *visc,
                   double *diff_sp, double *diff_T, double *rho);
/* GETMUCOEFF / GETKOVEREPS tables */
int orc_transport_fits(int mech, double *mucoeff, double *koveps);
/* finitechem_compute_rhs: sol = [rho*Y (nsp), rho*T], p_gp = 0.5(P_th+P_th_old) */
int orc_rhs(int mech, double p_gp, const double *sol, double *rhs);
/* analytic derivatives (no reference routine: CVODE differences the RHS):
 * dcdC[j*nsp + k] = d cdot_k / d C_j, dcdT[k] = d cdot_k / dT at fixed C;
 * J[j*neq + i] = d rhs_i / d sol_j of finitechem_compute_rhs, f = rhs */
int orc_dprodrates(int mech, const double *C, double T, double P, double *cdot, double *dcdC, double *dcdT);
int orc_jacobian(int mech, double p_gp, const double *sol, double *f, double *J);
/* one cell, t: 0 -> dt, sol in/out (neq = nsp+1) */
int orc_integrate(int mech, const orc_opts *o, double dt, double p_gp, double *sol, orc_stats *st);
/* batch of packed cells: sol[cell*neq + i]; threads >= 1 */
int orc_integrate_batch(int mech, const orc_opts *o, double dt, double p_gp, int64_t ncell,
                        double *sol, orc_stats *per_cell, int nthreads);

/* Fortran-layout grid: arrays RHO_(imino_:imaxo_, jmino_:jmaxo_, kmino_:kmaxo_),
 * SC_(..., nscalar), mask(imino:imaxo, jmino:jmaxo); isc_* are 1-based. */
typedef struct {
  int imin_, imax_, jmin_, jmax_, kmin_, kmax_;
  int imino_, imaxo_, jmino_, jmaxo_, kmino_, kmaxo_;
  int imino, imaxo, jmino, jmaxo;
  int nscalar, isc_sc, isc_T;
} orc_grid;

int orc_source_chemistry(int mech, const orc_opts *o, double dt, double p_thermo, double p_thermo_old,
                         const orc_grid *g, double *RHO_, double *SC_, const int *mask,
                         double *HR_gp, double *chemsrc_gp, orc_stats *total, int nthreads);

/* Metric operators of metric_generic (metric/metric_generic.f90:46-84), Fortran
 * arrays with the reference's extents: interp_sc_x(imin_-st1:imax_+st2,
 * jmino_:jmaxo_, -st2:st1); interp_sc_y, interp_sc_z(imino_:imaxo_,
 * jmin_-st1:jmax_+st2, -st2:st1); grad_x/y/z(imin_-st1:imax_+st2,
 * jmin_-st1:jmax_+st2, -st2:st1); div_u/v/w and interp_u_xm/v_ym/w_zm
 * (imin_:imax_, jmin_:jmax_, -st1:st2); grad_xm/ym/zm(imin_:imax_, jmin_:jmax_,
 * -stp:stp), stp = st1 + st2. */
typedef struct {
  int st1, st2;
  const double *interp_sc_x, *interp_sc_y, *interp_sc_z;
  const double *grad_x, *grad_y, *grad_z;
  const double *div_u, *div_v, *div_w;
  const double *interp_u_xm, *interp_v_ym, *interp_w_zm;
  const double *grad_xm, *grad_ym, *grad_zm;
} orc_metric;

/* finitechem_diffusion (finitechem.f90:538-923): diff_corr(..., nsp+1) and
 * diff_gp(..., 3) over the ghosted block (interior cells written, the rest 0),
 * from SC_, DIFF_ (both (..., nscalar)), RHO_ and mask; molar / gasrad / cv:
 * the 'Molar diffusion' and 'Gas radiation' parser keys and compressible or
 * constant volume. */
int orc_diffusion(int mech, const orc_grid *g, const orc_metric *m, double dt, int molar, int gasrad, int cv,
                  const double *RHO_, const double *SC_, const double *DIFF_, const int *mask, double *diff_corr,
                  double *diff_gp);

#ifdef __cplusplus
}
#endif
