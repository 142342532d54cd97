/*
 * budget_oracle.h -- CPU restatement of the reference's per-snapshot turbulence
 * statistics and kinetic-energy budget, TKE-budget/computations/budget2D.c
 * (TEST INFRASTRUCTURE: the checker of the HIP budget, never the product path).
 *
 * Inputs are one snapshot in the solver's vel-* / den-* file layouts
 * ([X][Y][Z][3] and [X][Y][Z], coarse grid and both bands with their wall layers).
 * Outputs are the profiles the reference tool writes, NZ = 2(zds+1) + zdv - 4
 * values each:
 *   turb[19]: UAVG VAVG WAVG UTIN VTIN WTIN uw DAVG DTIN SKEWu SKEWv SKEWw SKEWuw
 *             SKEWp KURTu KURTv KURTw KURTuw KURTp          (turb-fld-2d.000.NNNN)
 *   vort[3]:  wx wy wz                                       (vort-fluc-2d.000.NNNN)
 *   ke[30]:   Pii Puu Pvv Pww Puw Tiir Tuur Tvvr Twwr Tuwr Tiipai Tuupai Tvvpai
 *             Twwpai Tuwpai Paiuu Paivv Paiww Paiuw Tiis Tuus Tvvs Twws Tuws Epii
 *             Epuu Epvv Epww Epuw KE                         (KE-Budget-2d.000.NNNN)
 * and, if ufl/vfl are given, the lower band's u' and v' at k = 4 ([Xf][Yf],
 * uv_fluc.c: ufl4/vfl4.NNNN).  The viscosity comes from Rebs and ub as in the
 * tool (budget2D.c:55-58).
 */
#ifndef DNSG_BUDGET_ORACLE_H
#define DNSG_BUDGET_ORACLE_H

#include "../include/dnsg.h"

#ifdef __cplusplus
extern "C" {
#endif

int  orc_budget_profile_len(const dnsg_params *p);
int  orc_budget2d(const dnsg_params *p, const double *velc, const double *denc, const double *velfl,
                  const double *denfl, const double *velfu, const double *denfu, double *turb, double *vort,
                  double *ke, double *ufl, double *vfl);
/* the O(NZ) tail of the budget: normalisation and derived profiles from the layer
 * means Um[4][NZ] (U, V, W, rho/3), dU/dz[NZ] and the 44 raw per-layer sums */
void orc_budget_finish(const dnsg_params *p, const double *Um, const double *dUdz, const double *acc,
                       double *turb, double *vort, double *ke);

#ifdef __cplusplus
}
#endif

#endif
