/* Host-only stand-ins for the device layer (wfl_api.hip) so that the C host
 * code (wfl_host.c, wfl_effact.c) and the oracle can be built and run under
 * AddressSanitizer/UndefinedBehaviorSanitizer with gcc, without HIP.  Test
 * infrastructure only: the device entry points report "no device", and the
 * effective tuple is the nominal one (the sanitized run checks memory and
 * undefined behaviour, not numbers; those are checked against the real
 * library elsewhere). */
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include "wfl_internal.h"

static char g_err[512];

void wfl_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
const char *wfl_last_error(void) { return g_err; }

int wfl_effective_tuple(const wfl_geometry *g, double cx, double cy, double T, double out[6]) {
  if (!g || !out || !(g->l2 > 0.0) || !(T > 0.0)) { wfl_set_error("wfl_effective_tuple: bad arguments"); return WFL_ERR_ARG; }
  out[0] = g->l2; out[1] = g->F; out[2] = cx; out[3] = cy; out[4] = sqrt(T); out[5] = g->lattice > 0.0 ? g->lattice : 0.0;
  return WFL_OK;
}

int wfl_internal_ev_launch(const void *a, int b, int c, const wfl_geometry *d, const float *e, const float *f, int g,
                           int h, int i, int j, int k, float l, float m, float n, void *o, void *p) {
  /* the signature of wfl_internal.h */
  (void)a; (void)b; (void)c; (void)d; (void)e; (void)f; (void)g; (void)h; (void)i; (void)j; (void)k;
  (void)l; (void)m; (void)n; (void)o; (void)p;
  wfl_set_error("no HIP device in the sanitized host build");
  return WFL_ERR_DEVICE;
}
int wfl_internal_d2h(void *dst, const void *src, size_t bytes) { (void)dst; (void)src; (void)bytes; return WFL_ERR_DEVICE; }
void wfl_internal_ev_release(void) {}
int wfl_loops_shape(const wfl_loops *L, int *Nl, int *N) { (void)L; (void)Nl; (void)N; return WFL_ERR_DEVICE; }
int wfl_plan_create(wfl_plan **out, const wfl_batch *b, wfl_loops *L) { (void)b; (void)L; if (out) *out = NULL; return WFL_ERR_DEVICE; }
int wfl_plan_run(wfl_plan *p, void *s) { (void)p; (void)s; return WFL_ERR_DEVICE; }
int wfl_plan_read(wfl_plan *p, double *g, double *q) { (void)p; (void)g; (void)q; return WFL_ERR_DEVICE; }
void wfl_plan_destroy(wfl_plan *p) { (void)p; }
