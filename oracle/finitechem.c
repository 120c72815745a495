/* CPU oracle, part 3: the per-cell driver -- TEST INFRASTRUCTURE ONLY.
 *
 * Restates finitechem_source_chemistry (combustion/finitechem.f90:962-1071):
 * for every interior cell with mask /= 1 (996) pack sol = [max(RHO*Y, 0),
 * RHO*T] (1002-1003), integrate 0 -> dt (1014-1034), unpack by dividing by
 * RHO_ (1041-1042), then HR_gp = cp(Y_new, T_new) * (rhoT_new - rhoT_old)/dt
 * and chemsrc_gp(:,:,:,1:3) = d(rho Y)/dt of H2O, CO, CO2 (1047-1058).  A
 * species absent from the mechanism (CO/CO2 in H2) gets a source of 0; the
 * reference indexes out of bounds there (finitechem.f90:245-248, 1055-1057).
 *
 * Cells are independent; the loop is split over POSIX threads, one solver
 * workspace per thread (the reference's module scratch is not reentrant, so
 * it parallelises only across MPI ranks).
 */
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "oracle_int.h"


typedef struct {
  int mech;
  const orc_opts *o;
  double dt, p_gp;
  int64_t lo, hi;
  double *sol;
  orc_stats *per_cell;
  int neq;
} batch_job;

static void *batch_worker(void *arg) {
  batch_job *j = (batch_job *)arg;
  for (int64_t c = j->lo; c < j->hi; ++c) {
    orc_stats st;
    orc_integrate(j->mech, j->o, j->dt, j->p_gp, j->sol + c * j->neq, &st);
    if (j->per_cell) j->per_cell[c] = st;
  }
  return 0;
}

int orc_integrate_batch(int mech, const orc_opts *o, double dt, double p_gp, int64_t ncell, double *sol,
                        orc_stats *per_cell, int nthreads) {
  const tgr_mech_tables *t = orc_tables(mech);
  if (!t || ncell < 0) return -1;
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  pthread_t th[256];
  batch_job jobs[256];
  int64_t chunk = (ncell + nthreads - 1) / nthreads;
  int started = 0;
  for (int k = 0; k < nthreads; ++k) {
    jobs[k].mech = mech;
    jobs[k].o = o;
    jobs[k].dt = dt;
    jobs[k].p_gp = p_gp;
    jobs[k].lo = k * chunk < ncell ? k * chunk : ncell;
    jobs[k].hi = (k + 1) * chunk < ncell ? (k + 1) * chunk : ncell;
    jobs[k].sol = sol;
    jobs[k].per_cell = per_cell;
    jobs[k].neq = t->nsp + 1;
    if (nthreads == 1) {
      batch_worker(&jobs[k]);
    } else if (pthread_create(&th[k], 0, batch_worker, &jobs[k]) == 0) {
      started = k + 1;
    } else {
      batch_worker(&jobs[k]);
      started = k + 1;
    }
  }
  for (int k = 0; k < started && nthreads > 1; ++k) pthread_join(th[k], 0);
  return 0;
}

static int find_species(const tgr_mech_tables *t, const char *name) {
  for (int k = 0; k < t->nsp; ++k)
    if (strcmp(t->species[k], name) == 0) return k;
  return -1;
}

int orc_source_chemistry(int mech, const orc_opts *o, double dt, double p_thermo, double p_thermo_old,
                         const orc_grid *g, double *RHO_, double *SC_, const int *mask, double *HR_gp,
                         double *chemsrc_gp, orc_stats *total, int nthreads) {
  const tgr_mech_tables *t = orc_tables(mech);
  if (!t || !g || !RHO_ || !SC_) return -1;
  const int nsp = t->nsp, neq = nsp + 1;
  const int64_t nxo = g->imaxo_ - g->imino_ + 1, nyo = g->jmaxo_ - g->jmino_ + 1, nzo = g->kmaxo_ - g->kmino_ + 1;
  const int64_t plane = nxo * nyo * nzo;
  const int64_t mx = g->imaxo - g->imino + 1;
  const double p_gp = 0.5 * (p_thermo + p_thermo_old);
  /* gather the active cells (k, j, i order as the reference loop) */
  int64_t ncell_max = (int64_t)(g->imax_ - g->imin_ + 1) * (g->jmax_ - g->jmin_ + 1) * (g->kmax_ - g->kmin_ + 1);
  if (ncell_max <= 0) return 0;
  int64_t *idx = (int64_t *)malloc(sizeof(int64_t) * ncell_max);
  double *sol = (double *)malloc(sizeof(double) * ncell_max * neq);
  double *old = (double *)malloc(sizeof(double) * ncell_max * neq);
  orc_stats *st = (orc_stats *)calloc(ncell_max, sizeof(orc_stats));
  if (!idx || !sol || !old || !st) {
    free(idx); free(sol); free(old); free(st);
    return -1;
  }
  int64_t n = 0;
  for (int k = g->kmin_; k <= g->kmax_; ++k)
    for (int j = g->jmin_; j <= g->jmax_; ++j)
      for (int i = g->imin_; i <= g->imax_; ++i) {
        if (mask && mask[(int64_t)(j - g->jmino) * mx + (i - g->imino)] == 1) continue;
        int64_t c = (int64_t)(k - g->kmino_) * nxo * nyo + (int64_t)(j - g->jmino_) * nxo + (i - g->imino_);
        double rho = RHO_[c];
        double *s = sol + n * neq;
        for (int q = 0; q < nsp; ++q) s[q] = fmax(rho * SC_[(int64_t)(g->isc_sc - 1 + q) * plane + c], 0.0);
        s[nsp] = rho * SC_[(int64_t)(g->isc_T - 1) * plane + c];
        memcpy(old + n * neq, s, sizeof(double) * neq);
        idx[n++] = c;
      }
  orc_integrate_batch(mech, o, dt, p_gp, n, sol, st, nthreads);
  int iH2O = find_species(t, "H2O"), iCO = find_species(t, "CO"), iCO2 = find_species(t, "CO2");
  double h[64], cp[64];
  if (total) memset(total, 0, sizeof(*total));
  for (int64_t m = 0; m < n; ++m) {
    int64_t c = idx[m];
    double rho = RHO_[c];
    double *s = sol + m * neq, *s0 = old + m * neq;
    for (int q = 0; q < nsp; ++q) SC_[(int64_t)(g->isc_sc - 1 + q) * plane + c] = s[q] / rho;
    double Tn = s[nsp] / rho;
    SC_[(int64_t)(g->isc_T - 1) * plane + c] = Tn;
    orc_thermo_t(t, Tn, h, cp);
    double cpm = 0.0;
    for (int q = 0; q < nsp; ++q) cpm += SC_[(int64_t)(g->isc_sc - 1 + q) * plane + c] * cp[q];
    if (HR_gp) HR_gp[c] = cpm * (s[nsp] - s0[nsp]) / dt;
    if (chemsrc_gp) {
      chemsrc_gp[c] = iH2O >= 0 ? (s[iH2O] - s0[iH2O]) / dt : 0.0;
      chemsrc_gp[plane + c] = iCO >= 0 ? (s[iCO] - s0[iCO]) / dt : 0.0;
      chemsrc_gp[2 * plane + c] = iCO2 >= 0 ? (s[iCO2] - s0[iCO2]) / dt : 0.0;
    }
    if (total) {
      total->nst += st[m].nst;
      total->nfe += st[m].nfe;
      total->nfe_dq += st[m].nfe_dq;
      total->nje += st[m].nje;
      total->nsetups += st[m].nsetups;
      total->nni += st[m].nni;
      total->netf += st[m].netf;
      total->ncfn += st[m].ncfn;
      total->nlu += st[m].nlu;
      if (st[m].status != 0) total->status++;
    }
  }
  free(idx); free(sol); free(old); free(st);
  return 0;
}
