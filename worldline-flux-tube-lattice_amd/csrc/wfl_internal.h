// wfl_internal.h -- declarations shared by the C host code (wfl_host.c) and
// the HIP side (wfl_api.hip, wfl_kernels_*.hip).  Not part of the public ABI.
#ifndef WFL_INTERNAL_H
#define WFL_INTERNAL_H
#include <stdint.h>
#include "../../include/wfl.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Error string of the calling thread (set by both sides). */
void wfl_set_error(const char *fmt, ...);

/* One EV() evaluation on the device (used by the C drop-in EV()):
 * loops given as the caller's device float4 AoS [Nl*N]; one tuple
 * (cx, cy, rtT); writes the reference's float4 per-loop outputs into the
 * caller's device buffers ws_scal_d / ws_ferm_d ([Nl], ws_ferm_d may be NULL
 * when fermion == 0) and returns WFL_OK or an error code. */
int wfl_internal_ev_launch(const void *dev_aos, int Nl, int N, const wfl_geometry *g,
                           const float *knots_d, const float *coefs_d, int nspline,
                           int mode, int precision, int nsub, int fermion,
                           float cx, float cy, float rtT, void *ws_scal_d, void *ws_ferm_d);
/* Copy device -> host (synchronous). */
int wfl_internal_d2h(void *dst, const void *src, size_t bytes);
/* Drop the loop copies cached for EV(). */
void wfl_internal_ev_release(void);

/* The drop-in profile set by wfl_set_profile / wfl_sync_intT. */
const wfl_geometry *wfl_internal_profile(void);
/* Tfunc's integrand for an isolated tube (Integrand.c:252-305), with the
 * field it used in *B. */
double wfl_internal_igrand(const wfl_geometry *g, int fermion, double T, double rho, double rho2, double evm1,
                           int *branch, double *B);

/* Output level of the drivers (wfl_set_verbosity; the reference's verbosity). */
int wfl_internal_verbosity(void);

/* FP32 quantization used for every parameter the kernels see. */
static inline double wfl_q32(double v) { return (double)(float)v; }

#ifdef __cplusplus
}
#endif
#endif
