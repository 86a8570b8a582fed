/* The FP64 oracle's entry points under the sanitizers (called by
 * sanitize_driver.c): a small batch with lattice images, a spline tube and
 * the reference clamps, the spline set-up, EVMean and the closed forms. */
#include <stdlib.h>
#include <string.h>
#include "wfl_oracle.h"

int exercise_oracle(const float *x, const float *y, int Nl, int N, const float *kn, const float *c4,
                    const double *fli, int ns) {
  orc_geom og[2];
  memset(og, 0, sizeof og);
  og[0].profile = ORC_GAUSSIAN; og[0].l2 = 0.5; og[0].F = 1.0; og[0].lattice = 1.5; og[0].images = 1;
  og[1].profile = ORC_SPLINE; og[1].l2 = 1.0; og[1].F = 1.0; og[1].tubedist = 1.0;
  og[1].nspline = ns; og[1].rho2sp = kn; og[1].coefs = c4;
  int tg[4] = {0, 0, 1, 1};
  double tcx[4] = {0.2, 0.9, 0.0, 1.3}, tcy[4] = {0.1, 0.0, 0.0, 0.0}, trt[4] = {0.3, 1.2, 0.5, 2.0};
  double *pl = malloc(sizeof(double) * 4 * Nl * 6), *gs = malloc(sizeof(double) * 4 * ((Nl + 1) / 2) * 3);
  for (int mode = 0; mode < 2; ++mode)
    for (int spinor = 0; spinor < 2; ++spinor)
      orc_eval_batch(og, mode, mode ? 2 : 4, spinor, x, y, Nl, N, 2, 4, tg, tcx, tcy, trt, pl, gs, 1);
  float *oc4 = malloc(sizeof(float) * 4 * ns);
  double *ocd = malloc(sizeof(double) * 4 * ns);
  orc_spline(fli, kn, ns, 1, oc4, ocd);
  float ws[32], wf[32];
  int WL[8];
  for (int i = 0; i < 32; ++i) { ws[i] = 0.01f * i; wf[i] = 0.02f * i; }
  for (int i = 0; i < 8; ++i) WL[i] = i;
  double T[3] = {0.1, 1.0, 10.0}, f[3] = {1.0, 2.0, 3.0}, rn = 0.0;
  volatile double v = orc_logsimpson(T, f, 3) + orc_evmean(ws, wf, 8, WL, 1) + orc_expint(-3.0) +
                      orc_field_at(&og[0], 0.3, 0.2, &rn) + orc_exact(0.5, 1.0, 1) + orc_magfield(&og[0], 0.2);
  (void)v;
  free(pl); free(gs); free(oc4); free(ocd);
  return 0;
}
