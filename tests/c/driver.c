/* A C driver written the way the reference's EffAct.c and testT.c are: it
 * includes intT.h, fills struct Wparams field by field and calls the host
 * functions of the library.  Built with gcc as C11 and linked against
 * libwfl.so by tests/test_abi.py; it prints name=value lines that the test
 * checks against the same functions called through ctypes. */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include "intT.h"

int main(void) {
  struct Wparams params;
  memset(&params, 0, sizeof params);
  params.Nl = 256;
  params.Nppl = 64;
  params.F = 1.5;
  params.l2 = 0.64;
  params.xcm.x = 0.3f;
  params.ng = 2;
  params.fermion = 0;
  if (wfl_sync_intT(&params) != WFL_OK) {
    printf("sync failed: %s\n", wfl_last_error());
    return 1;
  }
  printf("sizeof_Wparams=%zu\n", sizeof(struct Wparams));
  printf("MagField=%.17g\n", MagField(0.25, &params));
  printf("Exact_scalar=%.17g\n", Exact(0.7, 0.9, 0));
  printf("Exact_spinor=%.17g\n", Exact(0.7, 0.9, 1));
  printf("smTscal=%.17g\n", smTscal(0.9, 0.01));
  printf("ClassAction=%.17g\n", ClassAction(params.F, params.l2));
  int count = 0;
  if (wfl_device_count(&count) != WFL_OK || count == 0) {
    int wl[128];
    for (int i = 0; i < 128; ++i) wl[i] = i;
    double v = EV(1.0, &params, wl);   /* no device: reported, NaN */
    printf("EV_without_device=%s\n", isnan(v) ? "nan" : "number");
  }
  return 0;
}
