/* tiger_chem: MI355X-native finite-rate chemistry step (C ABI).
 *
 * Drop-in for the chemistry step of NGA / Tiger_NGA: the body of
 * finitechem_source_chemistry (reference combustion/finitechem.f90:962-1071),
 * the SUNDIALS CVODE integration it drives (finitechem.f90:987-1068) and the
 * link-time mechanism plugin (combustion/mechanism.f90:17-108, one of
 * HYDROGENF.f90 / grimech30F.f90 / DME39F.f90 selected in
 * combustion/Makefile:3).  Plain pointers and sizes only; a Fortran host binds
 * these with ISO_C_BINDING (tiger_nga_amd/fortran/tiger_chem_mod.f90).
 *
 * Conventions: every function returns 0 on success and a negative TGR_E_*
 * code on failure (tgr_chem_last_error() gives the message).  Arrays use the
 * Fortran layout of the reference: RHO_(imino_:imaxo_, jmino_:jmaxo_,
 * kmino_:kmaxo_), SC_(..., nscalar) with i fastest, mask(imino:imaxo,
 * jmino:jmaxo); scalar indices are 1-based as in the reference.
 */
#ifndef TIGER_CHEM_H
#define TIGER_CHEM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TGR_E_ARG (-1)      /* invalid argument */
#define TGR_E_HIP (-2)      /* HIP runtime error */
#define TGR_E_MECH (-3)     /* unknown or unsupported mechanism */
#define TGR_E_NOMEM (-4)    /* device allocation failed */

typedef struct tgr_chem tgr_chem;

/* Extents of the calling rank's block (reference geometry/partition.f90:53-90,
 * geometry/masks.f90:55) and the scalar slots of the species and T
 * (finitechem.f90:19 isc_sc, combustion isc_T). */
typedef struct {
  int imin_, imax_, jmin_, jmax_, kmin_, kmax_;       /* interior cells, inclusive */
  int imino_, imaxo_, jmino_, jmaxo_, kmino_, kmaxo_; /* extents of RHO_ / SC_ */
  int imino, imaxo, jmino, jmaxo;                     /* extents of mask */
  int nscalar, isc_sc, isc_T;                         /* 1-based */
} tgr_grid;

/* Per-call integrator statistics, summed over the cells of the call (the
 * reference discards CVODE's iout counters and return code). */
typedef struct {
  int64_t cells;                  /* cells integrated (interior, mask /= 1) */
  int64_t nst, nfe, nfe_dq, nje;  /* steps, RHS evals, DQ-Jacobian RHS evals, Jacobians */
  int64_t nsetups, nni, netf, ncfn, nlu; /* setups, Newton iters, error/conv failures, LU factorisations */
  int64_t failed;                 /* cells whose integration stopped early (state = last step) */
  double kernel_ms;               /* device time of the step kernel */
} tgr_stats;

/* Mechanism sizes (reference getters GETNSPECIES/GETNSPECS/GETNREACTIONS,
 * HYDROGENF.f90:408-450) plus the table sizes the kernels use. */
typedef struct {
  int nsp, nrxn, nslot, n_third, n_falloff, n_third_body_rxn, n_reversible, n_nu;
  int n_fc_terms;       /* sum of Fc exponential terms over fall-off reactions */
  int n_conc_factors;   /* concentration factors over all rates of progress (forward + backward) */
  int n_eff_corr;       /* third-body efficiencies different from 1 (terms beyond the total concentration) */
  int n_rev_falloff;    /* reversible fall-off reactions */
} tgr_mech_info;

/* Create a context: mechanism tables on the current HIP device, solver
 * options, work buffers.  Replaces finitechem_init (finitechem.f90:184-319:
 * GETNSPECIES, GETMOLARMASS, rtol/atol parser keys, defaults 1e-6 / 1e-12)
 * and the CVODE set-up of finitechem.f90:1014-1028 (FCVMALLOC, STOP_TIME,
 * MAX_NSTEPS = 10000, FCVDENSE).  mech: "h2" | "gri30" | "dme39" | "ethyred" (EthyRedF, the
 * mechanism combustion/Makefile:3 links by default). */
int tgr_chem_create(const char *mech, double rtol, double atol, int max_steps, tgr_chem **out);

/* Release the context (FCVFREE, finitechem.f90:1068). */
int tgr_chem_destroy(tgr_chem *ctx);

/* GETNSPECIES / GETNREACTIONS and table sizes. */
int tgr_chem_mech_info(const tgr_chem *ctx, tgr_mech_info *info);

/* GETSPECIESNAMES (HYDROGENF.f90:389-405): name of species k (0-based). */
int tgr_chem_species_name(const tgr_chem *ctx, int k, char *buf, int len);

/* GETMOLARMASS (HYDROGENF.f90:366-386): W[nsp] in kg/kmol. */
int tgr_chem_molar_mass(const tgr_chem *ctx, double *W);

/* finitechem_source_chemistry(RHO_, SC_) (finitechem.f90:962-1071) on HOST
 * arrays: advances species and T of every interior cell with mask /= 1 over
 * dt at P = (p_thermo + p_thermo_old)/2, in place, and writes HR_gp and
 * chemsrc_gp(..., 1:3) = d(rho Y)/dt of H2O, CO, CO2 (0 where the mechanism
 * lacks the species).  mask, HR_gp, chemsrc_gp may be NULL.  Synchronous.
 * stats may be NULL. */
int tgr_chem_source_chemistry(tgr_chem *ctx, double dt, double p_thermo, double p_thermo_old,
                              const tgr_grid *g, double *RHO_, double *SC_, const int *mask,
                              double *HR_gp, double *chemsrc_gp, tgr_stats *stats);

/* Same on DEVICE arrays (HBM-resident fields), enqueued on `stream` (a
 * hipStream_t, NULL = default).  With stats != NULL the call waits for the
 * step and fills them; with NULL it returns after enqueueing. */
int tgr_chem_source_chemistry_device(tgr_chem *ctx, double dt, double p_thermo, double p_thermo_old,
                                     const tgr_grid *g, double *RHO_, double *SC_, const int *mask,
                                     double *HR_gp, double *chemsrc_gp, void *stream, tgr_stats *stats);

/* The same step on DEVICE arrays, enqueued on `stream`, allocating nothing.
 * No counterpart in the reference: for callers that cannot allocate device
 * memory during the call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The work counter, the counter totals and
 * the per-wave Jacobian workspace come from a 48 MiB file-scope device array
 * (one per GPU, held until the process ends) instead of context buffers; the
 * call makes no hipMalloc / hipFree and no host-device copy, returns after
 * enqueueing two kernels and reports no statistics.  It returns TGR_E_ARG when
 * the persistent grid's workspace (workgroups x 8 waves x neq x 512 bytes)
 * exceeds 48 MiB - 256 bytes: the default grid on MI355X fits every compiled
 * mechanism (EthyRedF needs 46 MiB); a larger "grid" option may not.  Every
 * call on a GPU shares that GPU's array: calls that overlap on one GPU (two
 * streams, two contexts, two host threads) race and must be serialised by the
 * caller.  The counters are cleared by a kernel at the start of each call. */
int tgr_chem_source_chemistry_device_noalloc(tgr_chem *ctx, double dt, double p_thermo, double p_thermo_old,
                                             const tgr_grid *g, double *RHO_, double *SC_, const int *mask,
                                             double *HR_gp, double *chemsrc_gp, void *stream);

/* PRODRATES (HYDROGENF.f90:4-277) for n states, device arrays, SoA:
 * C[nsp][n] kmol/m3, T[n] K, P[n] Pa -> cdot[nsp][n] kmol/m3/s. */
int tgr_chem_prodrates_device(tgr_chem *ctx, int64_t n, const double *C, const double *T, const double *P,
                              double *cdot, void *stream);

/* finitechem_compute_rhs (finitechem.f90:93-175) for n states, device arrays:
 * sol[neq][n] = [rho*Y, rho*T] -> rhs[neq][n] at pressure p_gp. */
int tgr_chem_rhs_device(tgr_chem *ctx, int64_t n, double p_gp, const double *sol, double *rhs, void *stream);

/* Transport and equation-of-state properties of a block, DEVICE arrays in the
 * layout of the step (RHO_ / SC_ Fortran order): for every cell of the array
 * (ghosts included) with mask /= 1,
 *   RHO_  = P_thermo W_mix / (R_cst T)                    finitechem_density (finitechem.f90:324-345)
 *   VISC_ = Wilke mixture viscosity                        finitechem_viscosity (351-411)
 *   DIFF_(..., species) = lambda / cp / Lewis_k, DIFF_(..., isc_T) = lambda / cp and, when
 *   isc_zmix > 0, DIFF_(..., isc_zmix) = lambda / cp       finitechem_diffusivity (414-461)
 * with GETMUCOEFF / GETKOVEREPS (HYDROGENF.f90:453-496) and finitechem_omegamu
 * (466-486).  Lewis: device array [nsp] or NULL (all 1, finitechem.f90:295);
 * any of RHO_, VISC_, DIFF_ may be NULL.  The border exchange that follows in
 * the reference (boundary_update_border) stays with the caller. */
int tgr_chem_properties_device(tgr_chem *ctx, double p_thermo, const tgr_grid *g, const double *SC_,
                               const int *mask, const double *Lewis, int isc_zmix, double *RHO_, double *VISC_,
                               double *DIFF_, void *stream);

/* Metric operators of the flow solver (metric_generic, reference
 * metric/metric_generic.f90:46-84), DEVICE arrays in Fortran layout with the
 * reference's extents (st1 >= 0, st2 >= 1: stencil widths, order = 2 st2 and
 * st1 = st2 - 1 in the reference; stp = st1 + st2):
 *   interp_sc_x(imin_-st1:imax_+st2, jmino_:jmaxo_, -st2:st1)
 *   interp_sc_y, interp_sc_z(imino_:imaxo_, jmin_-st1:jmax_+st2, -st2:st1)
 *   grad_x, grad_y, grad_z(imin_-st1:imax_+st2, jmin_-st1:jmax_+st2, -st2:st1)
 *   div_u/v/w, interp_u_xm / v_ym / w_zm(imin_:imax_, jmin_:jmax_, -st1:st2)
 *   grad_xm / ym / zm(imin_:imax_, jmin_:jmax_, -stp:stp) */
typedef struct {
  int st1, st2;
  const double *interp_sc_x, *interp_sc_y, *interp_sc_z;
  const double *grad_x, *grad_y, *grad_z;
  const double *div_u, *div_v, *div_w;
  const double *interp_u_xm, *interp_v_ym, *interp_w_zm;
  const double *grad_xm, *grad_ym, *grad_zm;
} tgr_metric;

#define TGR_DIFF_MOLAR 1   /* 'Molar diffusion' (finitechem.f90:254, default on) */
#define TGR_DIFF_GASRAD 2  /* 'Gas radiation' (finitechem.f90:216) */
#define TGR_DIFF_CV 4      /* compressible or constant volume: cv instead of cp */

/* finitechem_diffusion (finitechem.f90:538-923) on DEVICE arrays, and with
 * src != NULL the update of finitechem_source_transport (492-528): for every
 * interior cell with mask /= 1 writes diff_corr(..., 1:nsp) = dt div F_k (the
 * species' corrected molar-diffusion flux), diff_corr(..., nsp+1) = dt (sum_k
 * cp_k interp(F_k total) grad T [+ gas radiation] + DIFF_T grad cp grad T) /
 * cp, and diff_gp(..., 1:3) = div(Fickian flux) + the correction of H2O, CO,
 * CO2 (0 where the mechanism lacks the species is left untouched), and adds
 * diff_corr to src's species block and T (same layout as SC_).  RHO_(...),
 * SC_ and DIFF_(..., nscalar) as in the step (DIFF_ holds rho D per species and
 * lambda / cp at isc_T); mask(imino:imaxo, jmino:jmaxo) must cover the
 * array extents; the block needs st1 + st2 ghost cells on every side.  Any of
 * diff_corr, diff_gp, src may be NULL; cells outside the interior (and wall
 * cells) are not written.  Readings of the reference's non-conforming lines:
 * SC(i-st2:j+st1) at 652 as i+st1; z sections k-st2:k-st1 at 634-635 and 706
 * as k-st2:k+st1 (identical at second order); Cp_sp(isc) at 720-730 as the
 * species' own cp.  The burner heat loss (796-802) and the external ignition
 * source (842-859), fixed to single cases' meshes, stay with the caller. */
int tgr_chem_diffusion_device(tgr_chem *ctx, double dt, const tgr_grid *g, const tgr_metric *m, int flags,
                              const double *RHO_, const double *SC_, const double *DIFF_, const int *mask,
                              double *diff_corr, double *diff_gp, double *src, void *stream);

/* Temperature from mixture enthalpy for n states, device arrays, SoA:
 * Y[nsp][n] mass fractions, hmix[n] J/kg, T0[n] K starting guesses (NULL or
 * <= 0: 1500 K) -> T[n] K, iters[n] (may be NULL): Newton iterations, -1 when
 * not converged in 50 or for a non-finite h.  No counterpart in the reference:
 * the FP64 inversion of COMPTHERMODATA's h(T) (HYDROGENF.f90:514-673) used to
 * check a predicted state's enthalpy; Newton on sum_k Y_k h_k(T) = h with cp as
 * derivative, iterates clamped to the transport range [50, 5000] K
 * (finitechem.f90:7-8), converged at |dT| <= 1e-10 T. */
int tgr_chem_temperature_device(tgr_chem *ctx, int64_t n, const double *Y, const double *hmix, const double *T0,
                                double *T, int *iters, void *stream);

/* ---- surrogate-guided step (no counterpart in the reference) -------------
 * Every cell is integrated over dt by a one-step, L-stable, stiffly accurate
 * SDIRK3(2) method (Alexander's three stages, gamma = 0.4358665215) with one
 * difference-quotient Jacobian per cell ("surr_split_jac"), frozen over the
 * call, and FP64 simplified Newton per stage.  Two schedules: "surr_adaptive" 1 (default)
 * controls the step size by the embedded error estimate (filtered by
 * (I - gamma h J)^-1, accepted when |e_k| <= surr_tol_y in Y and surr_tol_t in
 * T; next step x min(surr_grow, max(0.2, 0.9 e^-1/3)), where a factor in
 * [1, surr_keep) keeps h and reuses the inverse of I - gamma h J; a stage
 * whose Newton iteration does not converge retries the step at h/4) from a
 * first step of
 * dt (r - 1) / (r^n - 1), r = "surr_grade", n = "surr_nsub", for at most
 * "surr_max_att" attempted steps; "surr_adaptive" 0 takes exactly the n
 * geometric sub-steps and accepts the cell when every sub-step's estimate
 * passes.  The learned part: a perceptron (include/tiger_mlp.h) trained on
 * (state, dt) -> increment pairs that the plain FP64 step generates predicts
 * the state at the schedule's stage (fixed) or sub-step-end (adaptive)
 * times; with "surr_guess" 1 each stage's Newton iteration starts from the
 * previous stage value moved by the prediction's change, with 0 (default)
 * from the stage's known part plus gamma h times the previous derivative
 * (adaptive) or from the previous stage value (fixed) -- on GRI-3.0 64^3 the
 * derivative start needs 13 % fewer iterations than the network's, so the
 * network is used where the tuning record says it pays.  An accepted cell's
 * species are projected back onto exact conservation of each element's mass
 * and the temperature implied by conservation of the mixture enthalpy is
 * written when it agrees with the solved one within surr_tol_t.  Rejected
 * cells are integrated by the plain FP64 step on the same call, so every cell
 * ends either accepted or exactly as tgr_chem_source_chemistry_device leaves
 * it.  Adaptive use: while the accepted fraction of the last guided call is
 * below "surr_min_accept" (default 0.15), calls run the plain step only,
 * except one guided probe every "surr_probe" calls (default 8; 0 = always
 * guided).  With surr_probe > 0 a guided call reads the accepted count back
 * and is therefore synchronous. */
typedef struct tgr_mlp tgr_mlp;

typedef struct {
  int64_t accepted, rejected;     /* cells accepted by the guided solve / handed to the plain step */
  int64_t newton_iters, rhs_evals; /* of the guided solve (Jacobian columns included) */
  int64_t rej_newton, rej_error, rej_enthalpy; /* why cells were rejected */
  int64_t substeps;               /* sub-steps taken by the guided solve (accepted ones, until a cell stops) */
  int64_t factorizations;         /* factorisations (Gauss-Jordan inverses) of I - gamma h J in the guided solve */
  int64_t jacobians;              /* Jacobian evaluations of the guided solve (1 per cell, more with refreshes) */
  double surrogate_ms, fallback_ms; /* device time of features + network + guided solve, of the plain step */
  int guided;                     /* 1: the guided solve ran; 0: the adaptive policy sent the call to the plain step */
} tgr_surrogate_stats;

/* Attach a trained network (borrowed; NULL detaches).  Inputs: nsp + 2
 * (log10 max(Y_k, 1e-30), T, log10 dt), each scaled as (x - in_shift) *
 * in_scale; outputs: nsp + 1, the increments of Y_k and T over dt divided by
 * out_scale. */
int tgr_chem_set_surrogate(tgr_chem *ctx, tgr_mlp *net, const float *in_shift, const float *in_scale,
                           const double *out_scale);

/* Network inputs of the interior cells of a block (i fastest) as bfloat16 bit
 * patterns X[rows][in_pad], rows >= cells and a multiple of 64 (the rows past
 * the cells are zero): the features the surrogate step feeds the network. */
int tgr_chem_surrogate_features_device(tgr_chem *ctx, double dt, const tgr_grid *g, const double *SC_, uint16_t *X,
                                       int64_t rows, void *stream);

/* finitechem_source_chemistry (finitechem.f90:962-1071) on DEVICE arrays by
 * the surrogate-guided path described above.  stats: the whole call (counters
 * of the plain step on the rejected cells, kernel_ms = total); sstats: the
 * guided solve.  Either may be NULL (then the call only enqueues). */
int tgr_chem_source_chemistry_surrogate_device(tgr_chem *ctx, double dt, double p_thermo, double p_thermo_old,
                                               const tgr_grid *g, double *RHO_, double *SC_, const int *mask,
                                               double *HR_gp, double *chemsrc_gp, void *stream, tgr_stats *stats,
                                               tgr_surrogate_stats *sstats);

/* The same on HOST arrays (staged through device buffers; synchronous), the
 * counterpart of tgr_chem_source_chemistry for a Fortran caller. */
int tgr_chem_source_chemistry_surrogate(tgr_chem *ctx, double dt, double p_thermo, double p_thermo_old,
                                        const tgr_grid *g, double *RHO_, double *SC_, const int *mask, double *HR_gp,
                                        double *chemsrc_gp, tgr_stats *stats, tgr_surrogate_stats *sstats);

/* Load a trained network file (tiger_nga_amd/models/<mech>.bin: "TGRMLP01",
 * int32 nin nout hidden layers, int64 nparam, float32 in_shift[nin]
 * in_scale[nin], float64 out_scale[nout], float32 params[nparam]) and attach
 * it; the context owns that network and frees it in tgr_chem_destroy. */
int tgr_chem_load_surrogate(tgr_chem *ctx, const char *path);

/* Real-valued knobs of the surrogate step: "surr_tol_y" (default 2.5e-7: the
 * adaptive schedule's error grows with its number of steps, each within the
 * threshold, mostly through the stage solves' Newton convergence error; at
 * the defaults its largest species error against a 1e-10 tolerance solve on
 * 2 x 32^3 benchmark states is 1.0e-7 (GRI-3.0), 6.6e-8 (EthyRedF), 2.6e-7
 * (DME39) and 8.3e-7 (H2) -- the tuning record holds the fastest safe values
 * per mechanism, for GRI-3.0 5e-7 with a Newton stop of 0.03),
 * "surr_tol_t" (K, default 1e-2), "surr_newton_tol" (Newton stops when the
 * correction is below this fraction of the acceptance tolerance, default 0.03;
 * at 0.1 the convergence error of the stage solves adds up over the steps to
 * two or three times the error at 0.03),
 * "surr_grade" (ratio of consecutive sub-step sizes, default 2),
 * "surr_min_accept" (adaptive-use threshold, default 0.15), "surr_h0_fac"
 * (adaptive schedule: first step h0_fac / max_k |f_k(y0)| / tol_k within
 * [dt/1024, dt] instead of the first sub-step; default 0 = off), "surr_grow"
 * (adaptive: largest step growth, 1.5..100, default 10), "surr_keep" (adaptive:
 * growth factors below this keep the step and its factors, 1..4, default
 * 1.5; on GRI-3.0 64^3 this halves the factorisation count and makes the guided step 9 %
 * faster at unchanged error), "surr_jac_theta" (adaptive: a step attempt whose
 * Newton iterations contract slower than this rate, or fail, with a Jacobian
 * of an earlier state re-evaluates the Jacobian at the current state before
 * the next attempt, while "surr_max_jac" allows; default 0.3).  Every stage's
 * Newton iteration stops when max(1, theta / (1 - theta)) |correction| is below
 * surr_newton_tol acceptance tolerances, theta its contraction rate. */
int tgr_chem_set_option_real(tgr_chem *ctx, const char *key, double value);

/* Tuning knobs (no counterpart in the reference):
 *   "grid"        size of the persistent grid in workgroups of 8 waves (0: occupancy x CUs);
 *                 every wave pulls cells
 *                 from a device work counter
 *   "rhs_mapping" tgr_chem_rhs_device: 1 (default) one state per lane (generated
 *                 straight-line code of the mechanism), 0 one state per wavefront
 *   "step_mapping"  0 (default): a wavefront per cell (the species in lanes);
 *                 1: a thread per cell -- the same CVODE-semantics BDF with the
 *                 cell's vectors and Nordsieck history in the thread's registers
 *                 and its iteration matrix in LDS -- for the mechanisms small
 *                 enough for it (H2, neq 10; others, and "jacobian" 1, keep the
 *                 wave per cell); the autotuner times both
 *   "step_waves"  wave per cell: resident waves per SIMD, 2 (default; up to 256
 *                 VGPRs per lane) or 4 (at most 128, two workgroups per CU; for
 *                 neq 10 only, whose tables leave LDS room for the second; with
 *                 "jacobian" 1 the two-wave kernel runs); the results are the
 *                 same, the autotuner times both
 *   "props_mapping"  tgr_chem_properties_device: 1 (default) a thread per
 *                 cell (the Wilke double sum in every lane), 0 a wavefront per
 *                 cell (species in lanes); the same results to rounding
 *   "jacobian"    0 (default): the Jacobian by difference quotients, one
 *                 right-hand side per column, as CVODE's FCVDENSE builds it for
 *                 the reference (finitechem.f90:1028); 1: the analytic Jacobian
 *                 of the same right-hand side (every PRODRATES / GETLINDRATECOEFF
 *                 / COMPTHERMODATA term differentiated, from the mechanism
 *                 tables), in the plain step and in the guided solve
 *   "max_order"   BDF maximum order 1..5 (CVODE default 5)
 *   "max_steps"   MAX_NSTEPS
 *   "surr_nsub"   sub-steps of the surrogate-guided schedule, 1..8 (default 8)
 *   "surr_guess"  1: Newton from the network's prediction; 0 (default): without
 *                 the network (derivative or previous-stage start, see above)
 *   "surr_adaptive"  1 (default): error-controlled step sizes; 0: the fixed schedule
 *   "surr_max_att"   adaptive schedule: attempted steps before the cell is
 *                    rejected, 1..1000 (default 48)
 *   "surr_split_jac" 1 (default): the guided solve's Jacobian takes its species
 *                    columns as difference quotients at fixed temperature that
 *                    reuse the rate constants and thermo of the base evaluation,
 *                    plus df/dT times dT/dy (T = rho T / rho); 0: every column a
 *                    full difference quotient as in the plain step
 *   "surr_newton_max"  Newton iterations per stage before the stage fails (default 8)
 *   "surr_mode"   0 (default): the guided solve above; 1: standalone surrogate --
 *                 the network's predictions at the stage times of the fixed
 *                 schedule (surr_nsub sub-steps of ratio surr_grade) are taken
 *                 as the stage values, each checked by ONE FP64 Newton
 *                 correction of its stage equation z = base + gamma h f(z) with
 *                 the frozen Jacobian (the correction is the prediction's
 *                 distance from the equation's solution, and is applied); a
 *                 stage whose correction exceeds "surr_sa_tol" (real, default
 *                 1) acceptance tolerances, or a sub-step whose embedded error
 *                 estimate exceeds 1, rejects the cell; accepted cells then pass
 *                 the element projection and enthalpy check as above, rejected
 *                 ones go to the plain step.  With "surr_guess" 0 the same
 *                 check runs on the zero-change prediction (each stage starts
 *                 from the previous stage value): the method without the network
 *   "surr_max_jac"   adaptive schedule: Jacobians per cell, 1..64 (default 4; 1: the
 *                    initial one only; see surr_jac_theta)
 *   "surr_probe"  adaptive use: a guided probe every this many calls while the
 *                 guided solve is not paying (default 8; 0 = always guided) */
int tgr_chem_set_option(tgr_chem *ctx, const char *key, int64_t value);

/* Message of the last failing call on this thread. */
const char *tgr_chem_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* TIGER_CHEM_H */
