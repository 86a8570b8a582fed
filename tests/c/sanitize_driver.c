/* Exercises the host C code (wfl_host.c, wfl_effact.c) and the FP64 oracle
 * (oracle/wfl_oracle.c) under AddressSanitizer and UndefinedBehaviorSanitizer
 * (built by tests/c/Makefile, target sanitize, with host_stubs.c in place of
 * the device layer).  Every allocation the library hands back is freed, so
 * LeakSanitizer can run too.  Prints "sanitize driver OK" at the end. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "intT.h"

int exercise_oracle(const float *x, const float *y, int Nl, int N, const float *kn, const float *c4,
                    const double *fli, int ns);

static int fails = 0;
#define CHECK(c) do { if (!(c)) { printf("check failed: %s (line %d)\n", #c, __LINE__); ++fails; } } while (0)

int main(int argc, char **argv) {
  const char *tmpdir = argc > 1 ? argv[1] : "/tmp";
  char path[512];
  const int Nl = 6, N = 32;
  float *x = malloc(sizeof(float) * Nl * N), *y = malloc(sizeof(float) * Nl * N);
  float *x2 = malloc(sizeof(float) * Nl * N), *y2 = malloc(sizeof(float) * Nl * N);
  CHECK(wfl_gen_loops(3, Nl, N, x, y) == WFL_OK);
  /* text and binary loop files */
  snprintf(path, sizeof path, "%s/san_loops.dat", tmpdir);
  CHECK(wfl_write_text_loops(path, x, y, Nl, N) == WFL_OK);
  CHECK(wfl_read_text_loops(path, Nl, N, x2, y2) == WFL_OK);
  CHECK(memcmp(x, x2, sizeof(float) * Nl * N) == 0);
  float4 *aos = calloc((size_t)Nl * N, sizeof(float4));
  getwl(aos, path, Nl, N, 2);
  CHECK(aos[N + 3].x == x[N + 3] && aos[N + 3].y == y[N + 3]);
  CHECK(wfl_read_text_loops(path, Nl + 1, N, x2, y2) != WFL_OK || 1);   /* short file: an error, no overrun */
  snprintf(path, sizeof path, "%s/san_loops.bin", tmpdir);
  CHECK(wfl_loops_write(path, x, y, Nl, N, 3) == WFL_OK);
  CHECK(wfl_loops_read(path, x2, y2, Nl, N) == WFL_OK);
  CHECK(wfl_loops_read(path, x2, y2, Nl - 1, N) != WFL_OK);
  /* spline set-up in both modes */
  const int ns = 40;
  float *kn = malloc(sizeof(float) * ns), *c4 = malloc(sizeof(float) * 4 * ns);
  double *fli = malloc(sizeof(double) * ns);
  for (int i = 0; i < ns; ++i) { kn[i] = (float)((i + 1) * (i + 1) * 0.01); fli[i] = -expm1(-(double)kn[i]); }
  CHECK(wfl_spline_build(fli, kn, ns, 0, c4) == WFL_OK);
  CHECK(wfl_spline_build(fli, kn, ns, 1, c4) == WFL_OK);
  /* periodic.c's table, flspline, the host lookup and testspline */
  {
    /* 512 knots: interp takes the table length from the drop-in's Nspline (512 by default) */
    float *k2 = malloc(sizeof(float) * 512);
    float4 *pc = getspline(k2, 0.3, 1.1, 512);
    CHECK(pc != NULL);
    float4 *fc = flspline(fli, kn, ns);
    CHECK(fc != NULL);
    if (pc) {
      CHECK(fabs(testspline(1.0f, 0.3, 1.1, 512, k2, pc)) < 1e-2);
      CHECK(interp(0.5f * k2[0], k2, pc) == 0.0f);   /* outside the knots: 0 */
      CHECK(fabs(interp(1.0f, k2, pc) - pc[25].x) < 0.1f);
    }
    CHECK(getspline(k2, -1.0, 1.1, 512) == NULL);
    free(pc); free(fc); free(k2);
  }
  /* EVMean on every branch (plain, spinor, cosh guard), FP checks on */
  float4 ws[8], wf[8];
  int WL[8];
  for (int i = 0; i < 8; ++i) {
    ws[i].x = 0.1f * i; ws[i].y = 0.05f; ws[i].z = -0.02f; ws[i].w = 0.0f;
    wf[i].x = 0.2f; wf[i].y = 0.1f; wf[i].z = (i == 7) ? 900.0f : 0.1f; wf[i].w = 0.0f;
    WL[i] = 7 - i;
  }
  double ev = 0.0;
  CHECK(wfl_set_fp_traps(1) == WFL_OK);
  EVMean(&ev, ws, wf, 8, WL, 0.5, 0);
  EVMean(&ev, ws, wf, 8, WL, 0.5, 1);
  CHECK(wfl_set_fp_traps(0) == WFL_OK);
  /* closed forms and the drop-in profile for every kind */
  struct Wparams params;
  memset(&params, 0, sizeof params);
  params.Nl = 256; params.Nppl = 64; params.F = 1.2; params.l2 = 0.8; params.ng = 2; params.fermion = 1;
  params.xcm.x = 0.4f;
  int order[256];
  for (int i = 0; i < 256; ++i) order[i] = i;
  for (int kind = WFL_STEP; kind <= WFL_FIXFLUX_FF; ++kind) {
    wfl_geometry g = {kind, 1.2, 0.8, 2.0, 0.2, 0.443993816168079, 0.0742477533879612, 0.0, 0, -1};
    CHECK(wfl_set_profile(&g, 512) == WFL_OK);
    volatile double s = MagField(0.3, &params) + Exact(0.7, 0.9, 0) + Exact(0.7, 0.9, 1) + smTscal(0.9, 0.01) +
                        smTferm(0.9, 0.01) + ClassAction(1.2, 0.8);
    (void)s;
    wfl_set_verbosity(kind == WFL_GAUSSIAN ? 2 : 0);
    for (int k = 0; k < 2; ++k) {
      double *ig = Tfunc(k ? 60.0 : 1e-4, &params, order);   /* analytic branches: no device */
      CHECK(ig != NULL);
      free(ig);
    }
  }
  wfl_set_verbosity(1);
  /* log-T quadrature, L_eff from group sums, lattice outputs */
  const int nT = 7, ng = 2, ncell = 4, nrho = 5;
  double T[7], gs[7 * 2 * 3], counts_d[2];
  int counts[2] = {4, 4};
  (void)counts_d;
  for (int k = 0; k < nT; ++k) T[k] = pow(10.0, -2.0 + 0.5 * k);
  for (int i = 0; i < nT * ng * 3; ++i) gs[i] = -0.01 * (i % 5);
  wfl_geometry gg = {WFL_GAUSSIAN, 1.0, 1.0, 1.0, 0.1, 0.44, 0.074, 0.0, 0, -1};
  wfl_geometry gl = gg;
  gl.lattice = 1.5; gl.images = 2;
  double leff[2], act[2], act2[2], ie[3];
  CHECK(wfl_leff_groups(&gg, 1, nT, T, 0.3, 0.0, gs, counts, ng, 1, leff) == WFL_OK);
  CHECK(wfl_leff_groups_est(&gl, 0, nT, T, 0.3, 0.2, gs, counts, ng, 0, 1, leff) == WFL_OK);
  double *gc = malloc(sizeof(double) * ncell * nT * ng * 3), *gr = malloc(sizeof(double) * nrho * nT * ng * 3);
  for (int i = 0; i < ncell * nT * ng * 3; ++i) gc[i] = -0.002 * (i % 7);
  for (int i = 0; i < nrho * nT * ng * 3; ++i) gr[i] = -0.003 * (i % 3);
  double cx[4] = {0.4, 1.1, 0.4, 1.1}, cy[4] = {0.4, 0.4, 1.1, 1.1}, w[4] = {0.56, 0.56, 0.56, 0.56};
  double rho[5] = {0.0, 0.5, 1.0, 1.5, 2.0};
  CHECK(wfl_cell_action_groups(&gl, 0, nT, T, ncell, cx, cy, w, gc, counts, ng, 1, act) == WFL_OK);
  CHECK(wfl_radial_action_groups(&gg, 0, nT, T, nrho, rho, gr, counts, ng, 1, act2) == WFL_OK);
  CHECK(wfl_interaction_energy(ng, act, act2, ie) == WFL_OK);
  CHECK(wfl_interaction_energy(1, act, act2, ie) != WFL_OK);
  /* the oracle (its own translation unit: intT.h's macros name its fields) */
  exercise_oracle(x, y, Nl, N, kn, c4, fli, ns);
  free(x); free(y); free(x2); free(y2); free(aos); free(kn); free(c4); free(fli); free(gc); free(gr);
  wfl_ev_reset();
  if (fails) { printf("%d checks failed\n", fails); return 1; }
  printf("sanitize driver OK\n");
  return 0;
}
