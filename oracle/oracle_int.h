/* Internal declarations shared by the oracle's translation units (test
 * infrastructure only). */
#pragma once
#include "oracle.h"
#include "../tiger_nga_amd/csrc/mech_tables.h"

const tgr_mech_tables *orc_tables(int mech);
int orc_prodrates_t(const tgr_mech_tables *t, const double *C, double T, double P, double *cdot, double *K,
                    double *W, double *M);
void orc_thermo_t(const tgr_mech_tables *t, double T, double *h, double *cp);
int orc_temperature_t(const tgr_mech_tables *t, const double *Y, double hmix, double T0, double *T);
void orc_rhs_t(const tgr_mech_tables *t, double p_gp, const double *sol, double *rhs);
int orc_dprodrates_t(const tgr_mech_tables *t, const double *C, double T, double P, double *cdot, double *dcdC,
                     double *dcdT);
void orc_jacobian_t(const tgr_mech_tables *t, double p_gp, const double *sol, double *f, double *J);
