// wfl_api.hip -- device side of the C ABI (include/wfl.h): loop ensembles,
// batched plans, the EV() launch helper, and the small data-movement and
// reduction kernels.  All functions exported here are extern "C".
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>
#include "wfl_internal.h"
#include "wfl_kernels.h"

using namespace wfl;

extern "C" int wfl_launch_tuple_f32(const WflLaunch* a);
extern "C" int wfl_launch_loop_f32(const WflLaunch* a);
extern "C" int wfl_launch_tuple_f64(const WflLaunch* a);
extern "C" int wfl_launch_loop_f64(const WflLaunch* a);

// ---------------------------------------------------------------- errors
static thread_local char g_err[512] = "";

extern "C" void wfl_set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof g_err, fmt, ap);
  va_end(ap);
}
extern "C" const char* wfl_last_error(void) { return g_err; }

#define HIP_TRY(expr)                                                                       \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      wfl_set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
      return (e_ == hipErrorOutOfMemory) ? WFL_ERR_NOMEM : WFL_ERR_DEVICE;                  \
    }                                                                                       \
  } while (0)

extern "C" int wfl_device_count(int* count) {
  if (!count) { wfl_set_error("wfl_device_count: NULL"); return WFL_ERR_ARG; }
  *count = 0;
  hipError_t e = hipGetDeviceCount(count);
  if (e != hipSuccess) { *count = 0; wfl_set_error("hipGetDeviceCount: %s", hipGetErrorString(e)); return WFL_ERR_DEVICE; }
  return WFL_OK;
}

extern "C" int wfl_device_sync(void) {
  HIP_TRY(hipDeviceSynchronize());
  return WFL_OK;
}

extern "C" int wfl_internal_d2h(void* dst, const void* src, size_t bytes) {
  HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return WFL_OK;
}

static int require_device() {
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) {
    wfl_set_error("no HIP device available (%s)", e != hipSuccess ? hipGetErrorString(e) : "count = 0");
    return WFL_ERR_DEVICE;
  }
  return WFL_OK;
}

// ------------------------------------------------------ data kernels
// Loop records (see wfl_kernels.h): w, 4 dw (dw = w_succ - w), e = w_succ - w_pred.
// Differences of FP32 inputs are exact in double and rounded once to R.
template <class R>
__global__ void k_records(const float* __restrict__ x, const float* __restrict__ y, int Nl, int N,
                          R* __restrict__ rec) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)Nl * N) return;
  int l = (int)(idx / N), i = (int)(idx % N);
  const float* xl = x + (size_t)l * N;
  const float* yl = y + (size_t)l * N;
  int s = succ_idx(i, N), p = pred_idx(i, N);
  double wx = xl[i], wy = yl[i];
  R* r = rec + (size_t)idx * REC;
  r[0] = (R)wx;
  r[1] = (R)wy;
  r[2] = (R)(4.0 * ((double)xl[s] - wx));   // 4 dw: exact scaling, carries Simpson's midpoint weight
  r[3] = (R)(4.0 * ((double)yl[s] - wy));
  r[4] = (R)((double)xl[s] - (double)xl[p]);
  r[5] = (R)((double)yl[s] - (double)yl[p]);
}

// Axis-form records (see wfl_kernels.h): w_x, w_y^2, 4 dw_x, m_y^2 with
// m_y = (w_y + w_y,succ)/2, 4 dw_y, e_y, w_y 4 dw_x, w_y e_x; differences
// of FP32 inputs (exact in double) and products rounded once.
__global__ void k_records_axis(const float* __restrict__ x, const float* __restrict__ y, int Nl, int N,
                               float* __restrict__ rec) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)Nl * N) return;
  int l = (int)(idx / N), i = (int)(idx % N);
  const float* xl = x + (size_t)l * N;
  const float* yl = y + (size_t)l * N;
  int s = succ_idx(i, N), p = pred_idx(i, N);
  const double wx = xl[i], wy = yl[i];
  const double dx4 = 4.0 * ((double)xl[s] - wx), my = 0.5 * (wy + (double)yl[s]);
  const double ex = (double)xl[s] - (double)xl[p];
  float* r = rec + (size_t)idx * RECA;
  r[0] = (float)wx;
  r[1] = (float)(wy * wy);
  r[2] = (float)dx4;
  r[3] = (float)(my * my);
  r[4] = (float)(4.0 * ((double)yl[s] - wy));
  r[5] = (float)((double)yl[s] - (double)yl[p]);
  r[6] = (float)(wy * dx4);
  r[7] = (float)(wy * ex);
}

// Largest |w|^2 over all points: block maximum, then one atomic per block
// on the float's bits (non-negative floats order like their bit patterns).
__global__ __launch_bounds__(256) void k_max_radius2(const float* __restrict__ x, const float* __restrict__ y,
                                                     long long n, unsigned int* __restrict__ out) {
  __shared__ float red[256];
  float m = 0.0f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    m = fmaxf(m, fmaf(x[i], x[i], y[i] * y[i]));
  red[threadIdx.x] = m;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + k]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, __float_as_uint(red[0]));
}

// Per-loop constants of the local Taylor node: class sums
// Q_c = sum_{i in c} w_i x (e_i + 4 dw_i) and the largest radius, in FP64
// from the FP32 coordinates (products of differences of FP32 values).  One
// block per loop; each thread sums a fixed stride of points and the block
// adds the partial sums in a fixed order.
__global__ __launch_bounds__(256) void k_loop_info(const float* __restrict__ x, const float* __restrict__ y, int N,
                                                   LoopInfo* __restrict__ info) {
  __shared__ double red[4][256];
  const int l = blockIdx.x;
  const float* xl = x + (size_t)l * N;
  const float* yl = y + (size_t)l * N;
  double q[3] = {0.0, 0.0, 0.0}, r2 = 0.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    const int sI = succ_idx(i, N), pI = pred_idx(i, N);
    const double wx = xl[i], wy = yl[i];
    const double vx = ((double)xl[sI] - (double)xl[pI]) + 4.0 * ((double)xl[sI] - wx);
    const double vy = ((double)yl[sI] - (double)yl[pI]) + 4.0 * ((double)yl[sI] - wy);
    const double a = wx * vy - wy * vx;
    const int c = cls_of(i);
    q[0] += c == 0 ? a : 0.0;
    q[1] += c == 1 ? a : 0.0;
    q[2] += c == 2 ? a : 0.0;
    r2 = fmax(r2, wx * wx + wy * wy);
  }
  red[0][threadIdx.x] = q[0]; red[1][threadIdx.x] = q[1]; red[2][threadIdx.x] = q[2]; red[3][threadIdx.x] = r2;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) {
      for (int c = 0; c < 3; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + k];
      red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    LoopInfo o;
    o.q[0] = red[0][0]; o.q[1] = red[1][0]; o.q[2] = red[2][0];
    o.r = sqrt(red[3][0]) * (1.0 + 1e-6);
    info[l] = o;
  }
}

__global__ void k_aos_to_soa(const float4* __restrict__ aos, long long n, float* __restrict__ x, float* __restrict__ y) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float4 v = aos[i];
  x[i] = v.x;
  y[i] = v.y;
}

// Clears the per-run counters (non-finite count, work-queue ticket).  A
// kernel rather than a memset so that the run can be captured in a graph.
__global__ void k_zero_counters(unsigned long long* nonfinite, int* queue) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    *nonfinite = 0ull;
    if (queue) *queue = 0;
  }
}

// Group sums of per-loop W - 1 (wave-per-loop mapping): one thread per
// (tuple, group), sequential FP64 sum in loop order.
__global__ void k_group_sums(const double* __restrict__ wm1, const DevTuple* __restrict__ tups, int ntup, int Nl,
                             int group, int ngroups, double* __restrict__ gsum) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)ntup * ngroups) return;
  int t = (int)(idx / ngroups), q = (int)(idx % ngroups);
  int nl = tups[t].nloops;
  int l0 = q * group, l1 = std::min(l0 + group, nl);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int l = l0; l < l1; ++l) {
    const double* p = wm1 + ((size_t)t * Nl + l) * 3;
    s0 += p[0]; s1 += p[1]; s2 += p[2];
  }
  double* o = gsum + idx * 3;
  o[0] = s0; o[1] = s1; o[2] = s2;
}

// Group sums from per-loop (Wsscal, Wsferm) (split finalize of the
// lane-per-tuple mapping): one thread per (tuple, group), W - 1 in FP64 and
// a sequential sum in loop order.
__global__ void k_finalize_groups(const double* __restrict__ perloop, const DevTuple* __restrict__ tups, int ntup,
                                  int Nl, int group, int ngroups, int mode, int spinor, double* __restrict__ gsum,
                                  unsigned long long* __restrict__ nonfinite) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)ntup * ngroups) return;
  int t = (int)(idx / ngroups), q = (int)(idx % ngroups);
  int nl = tups[t].nloops;
  int l0 = q * group, l1 = std::min(l0 + group, nl);
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  unsigned long long bad = 0;
  for (int l = l0; l < l1; ++l) {
    const double* p = perloop + ((size_t)t * Nl + l) * 6;
    double sc[3] = {p[0], p[1], p[2]}, fe[3] = {p[3], p[4], p[5]}, wm[3];
    wminus1(mode, spinor != 0, sc, fe, wm);
    if (!(finite_r(wm[0]) && finite_r(wm[1]) && finite_r(wm[2]))) ++bad;
    s0 += wm[0]; s1 += wm[1]; s2 += wm[2];
  }
  if (bad) atomicAdd(nonfinite, bad);
  if (l1 > l0) {
    double* o = gsum + idx * 3;
    o[0] = s0; o[1] = s1; o[2] = s2;
  }
}

// Group sums from chunk sums (lane-per-tuple mapping with each loop group
// split over cpg tasks): one thread per (tuple, group), chunks in order.
__global__ void k_sum_slots(const double* __restrict__ part, long long n, int cpg, double* __restrict__ gsum) {
  long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const double* p = part + idx * cpg * 3;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0;
  for (int c = 0; c < cpg; ++c) { s0 += p[3 * c]; s1 += p[3 * c + 1]; s2 += p[3 * c + 2]; }
  double* o = gsum + idx * 3;
  o[0] = s0; o[1] = s1; o[2] = s2;
}

// ------------------------------------------------------ loop ensembles
struct wfl_loops_s {
  int Nl = 0, N = 0;
  float* x = nullptr;      // SoA [Nl*N]
  float* y = nullptr;
  float* rec32 = nullptr;  // [Nl][N][REC]
  double* rec64 = nullptr;
  float* reca = nullptr;   // axis form [Nl][N][RECA]
  float rmax = -1.0f;      // largest |w| over the ensemble (< 0: not yet measured)
  LoopInfo* info = nullptr;          // per-loop constants of the local Taylor node [Nl]
  std::vector<double> radius;        // host copy of the per-loop largest radii
};

static int loops_alloc(wfl_loops** out, int Nl, int N) {
  if (!out) { wfl_set_error("loops: NULL output"); return WFL_ERR_ARG; }
  *out = nullptr;
  if (Nl < 1 || N < 4 || N % 4 != 0) {
    wfl_set_error("loops: need Nl >= 1 and N a positive multiple of 4 (got Nl=%d, N=%d)", Nl, N);
    return WFL_ERR_ARG;
  }
  int rc = require_device();
  if (rc) return rc;
  wfl_loops* L = new wfl_loops();
  L->Nl = Nl; L->N = N;
  size_t bytes = sizeof(float) * (size_t)Nl * N;
  hipError_t e1 = hipMalloc(&L->x, bytes);
  hipError_t e2 = (e1 == hipSuccess) ? hipMalloc(&L->y, bytes) : e1;
  if (e1 != hipSuccess || e2 != hipSuccess) {
    if (L->x) (void)hipFree(L->x);
    delete L;
    wfl_set_error("loops: hipMalloc of %zu bytes failed", 2 * bytes);
    return WFL_ERR_NOMEM;
  }
  *out = L;
  return WFL_OK;
}

extern "C" void wfl_loops_destroy(wfl_loops* L) {
  if (!L) return;
  if (L->x) (void)hipFree(L->x);
  if (L->y) (void)hipFree(L->y);
  if (L->rec32) (void)hipFree(L->rec32);
  if (L->rec64) (void)hipFree(L->rec64);
  if (L->reca) (void)hipFree(L->reca);
  if (L->info) (void)hipFree(L->info);
  delete L;
}

extern "C" int wfl_loops_create(wfl_loops** out, const float* x, const float* y, int Nl, int N) {
  if (!x || !y) { wfl_set_error("wfl_loops_create: NULL coordinates"); return WFL_ERR_ARG; }
  int rc = loops_alloc(out, Nl, N);
  if (rc) return rc;
  size_t bytes = sizeof(float) * (size_t)Nl * N;
  hipError_t e = hipMemcpy((*out)->x, x, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy((*out)->y, y, bytes, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    wfl_loops_destroy(*out); *out = nullptr;
    wfl_set_error("wfl_loops_create: upload failed: %s", hipGetErrorString(e));
    return WFL_ERR_DEVICE;
  }
  return WFL_OK;
}

extern "C" int wfl_loops_from_aos_device(wfl_loops** out, const void* aos, int Nl, int N) {
  if (!aos) { wfl_set_error("wfl_loops_from_aos_device: NULL"); return WFL_ERR_ARG; }
  int rc = loops_alloc(out, Nl, N);
  if (rc) return rc;
  long long n = (long long)Nl * N;
  hipLaunchKernelGGL(k_aos_to_soa, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0,
                     (const float4*)aos, n, (*out)->x, (*out)->y);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    wfl_loops_destroy(*out); *out = nullptr;
    wfl_set_error("AoS->SoA conversion failed: %s", hipGetErrorString(e));
    return WFL_ERR_DEVICE;
  }
  return WFL_OK;
}

extern "C" int wfl_gen_loops(uint32_t config_id, int Nl, int N, float* x, float* y);

extern "C" int wfl_loops_generate(wfl_loops** out, uint32_t config_id, int Nl, int N) {
  if (Nl < 1 || N < 4 || N % 4 != 0) { wfl_set_error("wfl_loops_generate: bad shape"); return WFL_ERR_ARG; }
  size_t n = (size_t)Nl * N;
  float* x = (float*)malloc(n * sizeof(float));
  float* y = (float*)malloc(n * sizeof(float));
  if (!x || !y) { free(x); free(y); wfl_set_error("wfl_loops_generate: host malloc"); return WFL_ERR_NOMEM; }
  int rc = wfl_gen_loops(config_id, Nl, N, x, y);
  if (rc == WFL_OK) rc = wfl_loops_create(out, x, y, Nl, N);
  free(x); free(y);
  return rc;
}

extern "C" int wfl_loops_shape(const wfl_loops* L, int* Nl, int* N) {
  if (!L) { wfl_set_error("wfl_loops_shape: NULL"); return WFL_ERR_ARG; }
  if (Nl) *Nl = L->Nl;
  if (N) *N = L->N;
  return WFL_OK;
}

// Device data derived from the loop set is built (and allocated) on first
// use; with build = false (a plan in caller-owned memory) data not yet built
// is an error.
static int loops_absent(const char* what) {
  wfl_set_error("wfl_eval_batch_noalloc: the loop set has no %s yet (an evaluation of a batch of the same "
                "precision and form with wfl_eval_batch or wfl_plan_create builds it)", what);
  return WFL_ERR_STATE;
}

// The ensemble's largest loop radius max |w| (measured once, on the device).
static int loops_rmax(wfl_loops* L, float* rmax, bool build = true) {
  if (L->rmax < 0.0f) {
    if (!build) return loops_absent("radius");
    unsigned int* d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(unsigned int)));
    unsigned int h = 0;
    hipError_t e = hipMemcpy(d, &h, sizeof h, hipMemcpyHostToDevice);
    if (e == hipSuccess) {
      const long long n = (long long)L->Nl * L->N;
      const unsigned grid = (unsigned)std::min<long long>((n + 255) / 256, 4096);
      hipLaunchKernelGGL(k_max_radius2, dim3(grid), dim3(256), 0, 0, L->x, L->y, n, d);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpy(&h, d, sizeof h, hipMemcpyDeviceToHost);
    (void)hipFree(d);
    if (e != hipSuccess) { wfl_set_error("loop radius: %s", hipGetErrorString(e)); return WFL_ERR_DEVICE; }
    float r2;
    memcpy(&r2, &h, sizeof r2);
    L->rmax = sqrtf(r2) * (1.0f + 1e-6f);
  }
  *rmax = L->rmax;
  return WFL_OK;
}

static int loops_records(wfl_loops* L, int precision, bool build = true) {
  if (!build && precision == WFL_MIXED && !L->rec32) return loops_absent("FP32 point records");
  if (!build && precision == WFL_FP64 && !L->rec64) return loops_absent("FP64 point records");
  long long n = (long long)L->Nl * L->N;
  dim3 grid((unsigned)((n + 255) / 256)), block(256);
  if (precision == WFL_MIXED && !L->rec32) {
    HIP_TRY(hipMalloc(&L->rec32, sizeof(float) * REC * (size_t)n));
    hipLaunchKernelGGL(k_records<float>, grid, block, 0, 0, L->x, L->y, L->Nl, L->N, L->rec32);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
  }
  if (precision == WFL_FP64 && !L->rec64) {
    HIP_TRY(hipMalloc(&L->rec64, sizeof(double) * REC * (size_t)n));
    hipLaunchKernelGGL(k_records<double>, grid, block, 0, 0, L->x, L->y, L->Nl, L->N, L->rec64);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipDeviceSynchronize());
  }
  return WFL_OK;
}

static int loops_records_axis(wfl_loops* L, bool build = true) {
  if (L->reca) return WFL_OK;
  if (!build) return loops_absent("axis-form records");
  long long n = (long long)L->Nl * L->N;
  HIP_TRY(hipMalloc(&L->reca, sizeof(float) * RECA * (size_t)n));
  hipLaunchKernelGGL(k_records_axis, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, L->x, L->y, L->Nl, L->N,
                     L->reca);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipDeviceSynchronize());
  return WFL_OK;
}

static int loops_info(wfl_loops* L, bool build = true) {
  if (L->info) return WFL_OK;
  if (!build) return loops_absent("per-loop sums");
  HIP_TRY(hipMalloc(&L->info, sizeof(LoopInfo) * (size_t)L->Nl));
  hipLaunchKernelGGL(k_loop_info, dim3((unsigned)L->Nl), dim3(256), 0, 0, L->x, L->y, L->N, L->info);
  HIP_TRY(hipGetLastError());
  std::vector<LoopInfo> h((size_t)L->Nl);
  HIP_TRY(hipMemcpy(h.data(), L->info, sizeof(LoopInfo) * h.size(), hipMemcpyDeviceToHost));
  L->radius.resize(h.size());
  for (size_t l = 0; l < h.size(); ++l) L->radius[l] = h[l].r;
  return WFL_OK;
}

// ------------------------------------------------------ local Taylor node
// a[k] = g^(k)(uc)/k!, k = 0..7, of g(u) = (1 - 2^-u)/u (the FP32 Gaussian
// node in scaled units): g^(k)(u) = ln2 (-ln2)^k I_k(ln2 u) with
// I_k(b) = int_0^1 t^k e^(-b t) dt.  I_7 from its series
// sum_j (-b)^j / (j! (8 + j)), the others by the backward recurrence
// I_(k-1) = (b I_k + e^-b)/k (integration by parts; stable downwards).
static const int kLocalK = 7;
static void local_taylor(double uc, double a[kLocalK + 1]) {
  const double L = 0.69314718055994530942, b = L * uc;
  double I[kLocalK + 1];
  double term = 1.0, sum = 0.0;
  for (int j = 0; j < 1000; ++j) {
    sum += term / (double)(kLocalK + j + 1);
    term *= -b / (double)(j + 1);
    if ((double)j > b && fabs(term) < 1e-20 * fabs(sum)) break;
  }
  I[kLocalK] = sum;
  const double eb = exp(-b);
  for (int k = kLocalK; k >= 1; --k) I[k - 1] = (b * I[k] + eb) / (double)k;
  double pw = L, fact = 1.0;
  for (int k = 0; k <= kLocalK; ++k) {
    a[k] = pw * I[k] / fact;
    pw *= -L;
    fact *= (double)(k + 1);
  }
}

extern "C" int wfl_local_taylor(double uc, double* a) {
  if (!a || !(uc >= 0.0) || !isfinite(uc)) { wfl_set_error("wfl_local_taylor: need uc >= 0 and an output"); return WFL_ERR_ARG; }
  local_taylor(uc, a);
  return WFL_OK;
}

// Tuples of the FP32 Gaussian axis form whose loops (up to radius R) keep
// every node within kLocalDmax of uc = c'^2 take the local node; |c'| is
// kept at most 3 (uc <= 9).
static const double kLocalCmax = 3.0;
static double local_dmax(const DevTuple& d, double R) {
  const double c = fabs((double)d.cx), s = d.s;
  return s * R * (s * R + 2.0 * c);
}
static void set_local(DevTuple& d, double R) {
  d.loc = 0;
  const double c = fabs((double)d.cx);
  if (!(c <= kLocalCmax) || !(local_dmax(d, R) <= (double)kLocalDmax)) return;
  double a[kLocalK + 1];
  const double uc = (double)d.cx * (double)d.cx;
  local_taylor(uc, a);
  d.loc = 1;
  d.ta0 = a[0];
  d.tec = exp2(-uc);
  const float a1h = (float)a[1];
  d.ta[0] = a1h;
  d.ta[1] = (float)(a[1] - (double)a1h);
  for (int k = 2; k <= kLocalK; ++k) d.ta[k] = (float)a[k];
}

// ------------------------------------------------------ geometry upload
static const double WFL_LOG2E = 1.4426950408889634074;

// Unit scale of the FP32 kernel for geometry g: sigma = sqrt(log2(e)/l2)
// rounded to FP32 on the Gaussian fast path (node = exp2(-u'), 1/u'), 1 for
// the generic profiles.  The effective problem parameters are defined from
// the FP32 values so that the FP32 and FP64 kernels evaluate the same problem.
static double unit_scale(const wfl_geometry& g, bool scaled) {
  if (!scaled || g.kind != WFL_GAUSSIAN) return 1.0;
  return (double)(float)sqrt(WFL_LOG2E / g.l2);
}

// ------------------------------------------------------ spline blob
// All spline tables of a launch in one float4 buffer:
//   [coefficient rows x (n + 1) per table][bucket records x btot][knots (float), padded]
// Rows: 0 = at or below the first knot, j + 1 = interval j, n = at or beyond
// the last knot; the outer rows depend on the segment mode (zero in
// reference mode; in physical mode the line s u with s = fl32(c0/k0), or c1
// if k0 = 0, and the constant c0 of the last knot).
// Table t has nb equal buckets over r = sqrt(rho^2) in [sqrt(k0), sqrt(k_{n-1})].
// Bucket b holds (s, k[s], k[s+1], k[s+2]) (entries past n - 2 are +inf),
// s = the interval of the lowest u that the kernel's FP32 bucket arithmetic
// can place in b.  The table is "fast" when the interval of the highest such
// u is at most s + 2 for every bucket; nb is the smallest of n/2, n, 2n, 4n,
// 8n (at most 16384) for which it is, else 2n with the scan.
// Uniform-grid rows (FP32 physical mode, tables whose knots lie on
// r_j = (m + j) h, m an integer), at the front of the blob: at most kUniRows
// float4 of rows (each table's guard row with its row 0's polynomial, then m
// copies of the row below the first knot, its n - 1 intervals, and copies of
// the row past the last knot up to the largest u its tuples reach), each
// cubic re-expanded about the point the kernel computes for its row,
// kappa_i = fl32(i h)^2 (kappa = (h/2)^2 for the guard row).
// uni_f4 is the size of that prefix, the part the uniform-grid kernels
// stage in LDS.  A table whose padded rows do not fit is not uniform.
struct SplineTab {
  int cf, rec, kn, n, nb, fast;
  float r0, inv_w, k0, kl;
  int uni, ucf, ulo, urows;   // uniform rows: offset (after the guard row), last row below the first knot, rows
  float uinv_h, uh;
};
struct SplineBlob {
  std::vector<float4> data;
  std::vector<SplineTab> tab;
  int uni_f4 = 0;        // uniform-grid rows of all tables (a prefix of data)
  bool all_uni = false;  // every table has them
  bool uni_origin = false;  // they hold their cubics about u = 0
};

// Knots on a uniform grid in r = sqrt(rho^2) that starts on the grid:
// r_j = (m + j) h with an integer m >= 0 (periodic.c:129-151 puts the first
// knot at rho = h, m = 1).  With inv_h = fl32(1/h) every knot k_j has
// sqrt(k_j) inv_h within 1e-4 of m + j, so the kernel's row
// floor(sqrt(u) inv_h) is interval row - m; p = m - 1 is the last row
// below the first knot (-1 when the first knot is rho = 0).
static bool spline_uniform(const float* k, int n, SplineTab& st) {
  if (n < 3 || !(k[0] >= 0.0f)) return false;
  const double r0 = sqrt((double)k[0]), r1 = sqrt((double)k[n - 1]);
  const double h = (r1 - r0) / (double)(n - 1);
  if (!(h > 0.0)) return false;
  const float inv_h = (float)(1.0 / h);
  const double m = nearbyint(r0 * (double)inv_h);
  if (!(m >= 0.0 && m < (double)kUniRows)) return false;
  for (int j = 0; j < n; ++j) {
    const double t = sqrt((double)k[j]) * (double)inv_h;
    if (!(fabs(t - (m + (double)j)) <= 1e-4)) return false;
  }
  st.uinv_h = inv_h;
  st.uh = (float)h;
  st.ulo = (int)m - 1;
  return true;
}

// Rows a uniform table needs for nodes up to u = umax: the kernel's row
// floor(sqrt(u) inv_h), FP32-rounded, plus two rows of margin.
static double spline_uniform_rows(const SplineTab& st, double umax) {
  if (!(umax >= 0.0) || !std::isfinite(umax)) return HUGE_VAL;
  const double t = sqrt(umax * (1.0 + 1e-5)) * (double)st.uinv_h * (1.0 + 1e-6);
  return floor(t) + 3.0;
}

// The exact lookup's row (0 the line below the first knot, n the constant
// past the last) whose cubic uniform row i holds (i = -1: the guard row).
static int uniform_row_interval(const SplineTab& st, int i) {
  return std::min(std::max(std::max(i, 0) - st.ulo, 0), st.n);
}

// Whether the cubics of a uniform table's rows, expanded about u = 0 as
// P(u) = d0 + d1 u + d2 u^2 + d3 u^3, keep the accuracy of the expanded
// rows in FP32: on every row's range of r = sqrt(u) (9 points),
// sum_k |d_k| u^k <= 16 |P(u)| (the relative rounding error of f stays
// within 16 times that of its value), and for spin factors (spin) also
// sum_k k |d_k| u^(k-1) <= 16 |P'(u)|, pointwise: the spin sum of a small
// loop takes f' at one u, and the spinor weight cancels it against the
// phase at leading order, so a relative error there is not averaged away.
// Smooth profiles of a few knots' width or more pass the first; a table
// whose cubics cancel about 0 (a nearly linear f far out, a wide tube's last
// rows) does not.  A Gaussian tube fails the second wherever f' ~ e^-u is
// small against d1 ~ e^-u (1 + u + u^2/2): spinor plans over such tables
// keep the rows about their start.  rows: the table's exact-lookup rows.
static bool spline_origin_ok(const float4* rows, const float* k, const SplineTab& st, bool spin) {
  for (int i = -1; i < st.urows; ++i) {
    const int r = uniform_row_interval(st, i);
    const float4 cr = rows[r];
    const double kr = r == 0 ? 0.0 : (double)k[r - 1];
    const double c0 = cr.x, c1 = cr.y, c2 = cr.z, c3 = cr.w;
    const double d3 = c3, d2 = c2 - 3.0 * c3 * kr, d1 = c1 - kr * (2.0 * c2 - 3.0 * c3 * kr);
    const double d0 = c0 - kr * (c1 - kr * (c2 - c3 * kr));
    const double lo = i < 0 ? 0.0 : i * (double)st.uh, hi = i < 0 ? 0.5 * st.uh : (i + 1) * (double)st.uh;
    for (int s = 0; s <= 8; ++s) {
      const double x = lo + (hi - lo) * (s / 8.0), u = x * x;
      const double P = d0 + u * (d1 + u * (d2 + u * d3));
      const double S = fabs(d0) + u * (fabs(d1) + u * (fabs(d2) + u * fabs(d3)));
      if (!(S <= 16.0 * fabs(P))) return false;
      if (!spin) continue;
      const double P1 = d1 + u * (2.0 * d2 + 3.0 * d3 * u);
      const double S1 = fabs(d1) + u * (2.0 * fabs(d2) + 3.0 * fabs(d3) * u);
      if (!(S1 <= 16.0 * fabs(P1))) return false;
    }
  }
  return true;
}

// FP32 coefficients (d0, d1, d2, d3) of P(u) = sum_k d_k u^k for a row
// whose cubic is cr about kr, taken near u = uc (the row's centre): written
// about uc (b_k, exact in double to rounding), d3 is rounded first and
// d3 u^3 = d3 (uc + v)^3 subtracted from the b_k, then likewise d2, d1 and
// d0, so that each rounding error is left as (b_k - d_k) v^k with
// |v| <= the row's width instead of times u^k.  What remains are two
// constant offsets per row: d1's rounding in f' = P', and d0's in P (half
// an FP32 ulp of d1 / P' and of d0 / P at most).  Each is rounded up or
// down so that these offsets, relative to P'(uc) and P(uc), sum to near
// zero over consecutive rows (carry[1], carry[0]: the running sums); d1's
// choice is compensated in d0 before d0 is rounded.
// Exactly representable rows (the line below the first knot, the constant
// past the last) come out unchanged, and a cubic through u = 0 (c0 = 0 about
// kr = 0) keeps d0 = 0, so that g = d0/u + ... stays finite at the centre.
static float round_spread(double b, double scale, double& carry) {
  float d = (float)b;
  if ((double)d == b || !(scale > 0.0) || !std::isfinite(b)) return d;
  const float other = (double)d > b ? nextafterf(d, -HUGE_VALF) : nextafterf(d, HUGE_VALF);
  if (fabs(carry + ((double)other - b) / scale) < fabs(carry + ((double)d - b) / scale)) d = other;
  carry += ((double)d - b) / scale;
  return d;
}
static float4 origin_row(const float4& cr, double kr, double uc, double carry[2]) {
  const double e = uc - kr;
  const double c0 = cr.x, c1 = cr.y, c2 = cr.z, c3 = cr.w;
  const double P = c0 + e * (c1 + e * (c2 + e * c3));
  double b0 = P, b1 = c1 + e * (2.0 * c2 + 3.0 * c3 * e), b2 = c2 + 3.0 * c3 * e;
  const double P1 = b1;   // P'(uc)
  const float d3 = (float)c3;
  b2 -= 3.0 * uc * (double)d3; b1 -= 3.0 * uc * uc * (double)d3; b0 -= uc * uc * uc * (double)d3;
  const float d2 = (float)b2;
  b1 -= 2.0 * uc * (double)d2; b0 -= uc * uc * (double)d2;
  const float d1 = round_spread(b1, fabs(P1), carry[1]);
  b0 -= uc * (double)d1;
  if (kr == 0.0 && cr.x == 0.0f) return make_float4(0.0f, d1, d2, d3);
  return make_float4(round_spread(b0, fabs(P), carry[0]), d1, d2, d3);
}

// largest j <= n - 2 with k[j] <= v, 0 if none
static int spline_interval_host(const float* k, int n, double v) {
  int lo = 0, hi = n - 1;
  if (!((double)k[0] <= v)) return 0;
  if ((double)k[n - 1] <= v) return n - 2;
  while (hi - lo > 1) {
    int mid = (lo + hi) >> 1;
    if ((double)k[mid] <= v) lo = mid; else hi = mid;
  }
  return lo;
}

static bool spline_buckets(const float* k, int n, int nb, float r0, float inv_w, double margin, std::vector<int>& start) {
  start.assign(nb, 0);
  bool fast = true;
  for (int b = 0; b < nb; ++b) {
    double e0 = (double)r0 + (b - margin) / (double)inv_w, e1 = (double)r0 + (b + 1 + margin) / (double)inv_w;
    e0 = std::max(e0, 0.0);
    const int s = spline_interval_host(k, n, e0 * e0), h = spline_interval_host(k, n, e1 * e1);
    start[b] = s;
    if (h - s > 2) fast = false;
  }
  return fast;
}

// umax[t] (FP32 physical mode): the largest u = rho^2 that a node of table
// t can see (table t gets uniform rows only if it is given and they fit);
// spin: the kernels take f' as well (spinor).
static void build_spline_blob(int ntab, const float* const* rho2, const float* const* coefs, const int* n, int mode,
                              SplineBlob& sb, const double* umax = nullptr, bool spin = false) {
  sb = SplineBlob();
  int rows = 0, btot = 0, ktot = 0, urows = 0;
  std::vector<std::vector<int>> starts(ntab);
  for (int t = 0; t < ntab; ++t) {
    const int nt = n[t];
    const float* k = rho2[t];
    SplineTab st{};
    st.n = nt;
    st.k0 = k[0]; st.kl = k[nt - 1];
    st.r0 = sqrtf(std::max(k[0], 0.0f));
    const float r1 = sqrtf(std::max(k[nt - 1], 0.0f));
    const double span = std::max((double)r1 - (double)st.r0, 1e-30);
    // bucket position error of the kernel (FP32 sqrt, subtract, multiply), in bucket widths
    auto margin_of = [&](int nb) { return std::max(0.01, 4e-7 * nb * (1.0 + (double)r1 / span)); };
    st.fast = 0;
    for (int nb : {std::max(16, nt / 2), nt, 2 * nt, 4 * nt, 8 * nt}) {
      nb = std::max(16, std::min(16384, nb));
      const float inv_w = (float)(nb / span);
      if (spline_buckets(k, nt, nb, st.r0, inv_w, margin_of(nb), starts[t])) {
        st.nb = nb; st.inv_w = inv_w; st.fast = 1;
        break;
      }
      if (nb == 16384) break;
    }
    if (!st.fast) {
      st.nb = std::max(16, std::min(16384, 2 * nt));
      st.inv_w = (float)(st.nb / span);
      spline_buckets(k, nt, st.nb, st.r0, st.inv_w, margin_of(st.nb), starts[t]);
    }
    st.uni = (umax && mode == WFL_MODE_PHYSICAL && spline_uniform(k, nt, st)) ? 1 : 0;
    st.ucf = -1;
    if (st.uni) {
      const double need = std::max(spline_uniform_rows(st, umax[t]), (double)(st.ulo + nt + 1));
      if (need + 1.0 <= (double)(kUniRows - urows)) {   // the rows and the guard row before them
        st.ucf = urows + 1;
        st.urows = (int)need;
        urows += st.urows + 1;
      } else {
        st.uni = 0;
      }
    }
    st.cf = rows;
    rows += nt + 1;
    st.rec = btot;
    btot += st.nb;
    st.kn = ktot;
    ktot += nt;
    sb.tab.push_back(st);
  }
  const size_t kn_f4 = ((size_t)ktot + 3) / 4;
  const int uni_f4 = urows;
  sb.data.assign((size_t)uni_f4 + rows + btot + kn_f4 + 1, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
  sb.uni_f4 = uni_f4;
  sb.all_uni = ntab > 0;
  float4* d = sb.data.data() + uni_f4;     // the exact-lookup part follows the uniform rows
  float* kn = (float*)(d + rows + btot);
  const float inf = __builtin_inff();
  for (int t = 0; t < ntab; ++t) {
    SplineTab& st = sb.tab[t];
    const int nt = st.n;
    const float* k = rho2[t];
    const float* c = coefs[t];
    for (int j = 0; j + 1 < nt; ++j) d[st.cf + j + 1] = make_float4(c[4 * j], c[4 * j + 1], c[4 * j + 2], c[4 * j + 3]);
    if (mode == WFL_MODE_PHYSICAL) {
      const float slope = k[0] > 0.0f ? (float)((double)c[0] / (double)k[0]) : c[1];
      d[st.cf] = make_float4(0.0f, slope, 0.0f, 0.0f);
      d[st.cf + nt] = make_float4(c[4 * (nt - 1)], 0.0f, 0.0f, 0.0f);
    }
    sb.all_uni = sb.all_uni && st.uni;
    for (int i = 0; i < nt; ++i) kn[st.kn + i] = k[i];
    for (int b = 0; b < st.nb; ++b) {
      const int s0 = starts[t][b];
      float x;
      memcpy(&x, &s0, sizeof x);
      d[rows + st.rec + b] = make_float4(x, k[s0], s0 + 1 <= nt - 2 ? k[s0 + 1] : inf, s0 + 2 <= nt - 2 ? k[s0 + 2] : inf);
    }
    st.cf += uni_f4;                       // float4 offsets in the blob
    st.rec += uni_f4 + rows;
    st.kn += 4 * (uni_f4 + rows + btot);   // float offset in the blob
  }
  // Row i of a uniform table is the exact lookup's row r (expanded about its
  // left knot: 0 for the line, k[r - 1] otherwise) shifted exactly, in
  // double, to kappa_i = fl32(i h)^2, what the kernel forms from the row
  // index (i = -1/2 for the guard row, the kernel's row for
  // sqrt(u) inv_h < 1/4), or to u = 0 when every uniform table is well
  // conditioned in that form.
  sb.uni_origin = true;
  for (int t = 0; t < ntab; ++t)
    if (sb.tab[t].uni && !spline_origin_ok(sb.data.data() + sb.tab[t].cf, rho2[t], sb.tab[t], spin))
      sb.uni_origin = false;
  for (int t = 0; t < ntab; ++t) {
    const SplineTab& st = sb.tab[t];
    if (!st.uni) continue;
    const float* k = rho2[t];
    float4* u = sb.data.data() + st.ucf;
    double carry[2] = {0.0, 0.0};
    for (int i = -1; i < st.urows; ++i) {
      const int r = uniform_row_interval(st, i);
      const float4 cr = sb.data[st.cf + r];
      const double kr = r == 0 ? 0.0 : (double)k[r - 1];
      if (sb.uni_origin) {
        // the row's centre in r: (h/8)^2 for the guard row (r < h/4)
        const double rc = (i < 0 ? 0.125 : i + 0.5) * (double)st.uh;
        u[i] = origin_row(cr, kr, rc * rc, carry);
        continue;
      }
      const double ri = (double)(float)((i < 0 ? -0.5 : (double)i) * (double)st.uh);
      const double e = ri * ri - kr;
      const double c0 = cr.x, c1 = cr.y, c2 = cr.z, c3 = cr.w;
      u[i] = make_float4((float)(c0 + e * (c1 + e * (c2 + e * c3))), (float)(c1 + e * (2.0 * c2 + 3.0 * c3 * e)),
                         (float)(c2 + 3.0 * c3 * e), (float)c3);
    }
  }
}

// Largest rho^2 that a node of a tuple can reach: the farthest tube of its
// image window (one image more on each side than the kernels sum, against
// rounding of the window's centre) plus sqrt(T) times the ensemble's
// largest loop radius (a midpoint lies within the radius of its ends).
static double node_u_bound(const wfl_geometry& G, double cx, double cy, double rtT, float rmax) {
  double ax = fabs(cx), ay = fabs(cy);
  if (G.lattice > 0.0) {
    const double L = G.lattice, M = (double)G.images + 1.0;
    const double m0 = floor(cx / L + 0.5), n0 = floor(cy / L + 0.5);
    ax = std::max(fabs(cx - (m0 - M) * L), fabs(cx - (m0 + M) * L));
    ay = std::max(fabs(cy - (n0 - M) * L), fabs(cy - (n0 + M) * L));
  }
  const double r = sqrt(ax * ax + ay * ay) + rtT * (double)rmax;
  return r * r;
}

// Largest spline blob staged in LDS (spline-only launches whose tables are all fast).
static const size_t kSplineLdsBytes = 64 * 1024;

static void fill_geom(const wfl_geometry& g, const SplineBlob* sb, bool scaled, DevGeom& d) {
  memset(&d, 0, sizeof d);
  const double sg = unit_scale(g, scaled);
  d.kind = g.kind;
  d.images = g.images;
  if (g.kind == WFL_SPLINE && sb) {
    const SplineTab& t = sb->tab[sb->tab.size() == 1 ? 0 : g.spline_table];
    d.spline_cf = t.cf; d.spline_rec = t.rec; d.spline_kn = t.kn;
    d.spline_n = t.n; d.spline_nb = t.nb; d.spline_fast = t.fast;
    d.spline_r0 = t.r0; d.spline_inv_w = t.inv_w;
    d.spline_k0 = t.k0; d.spline_kl = t.kl;
    d.spline_ucf = t.ucf;
    d.spline_uinv_h = t.uinv_h;
    d.spline_uh = t.uh;
  } else {
    d.spline_ucf = -1;
  }
  d.Ff = (float)g.F;       d.F = d.Ff;
  d.l2 = (sg != 1.0) ? WFL_LOG2E / (sg * sg) : (double)(float)g.l2;
  d.l2f = (float)d.l2;
  d.inv_l2 = 1.0 / d.l2;   d.inv_l2f = (float)d.inv_l2;
  d.af = (float)g.a;       d.a = d.af;
  d.lminf = (float)g.lam_min; d.lmin = d.lminf;
  d.q1f = (float)g.fix_q1; d.q1 = d.q1f;
  d.q2f = (float)g.fix_q2; d.q2 = d.q2f;
  if (g.lattice > 0.0) {
    d.latticef = (float)(sg * g.lattice);
    d.lattice = (double)d.latticef / sg;
  }
  d.c8f = (float)(1.0e-8 * sg * sg);
  d.c10f = (float)(1.0e-10 * sg * sg);
  // fixflux phi(0) = 1 - q2*/q2, q2* = (e^-1 - E1(1))/2 (E1(1) = 0.219383934395520273677...)
  d.fix_c0f = (float)(1.0 - 0.5 * (0.36787944117144232160 - 0.21938393439552027368) / d.q2);
}

static void fill_tuple(double sg, const DevGeom& dg, int gi, double cx, double cy, double rtT, int nloops,
                       DevTuple& d) {
  memset(&d, 0, sizeof d);
  d.geom = gi;
  d.nloops = nloops;
  d.cx = (float)(sg * cx);
  d.cy = (float)(sg * cy);
  d.s = (float)(sg * rtT);
  d.cxd = (double)d.cx / sg;
  d.cyd = (double)d.cy / sg;
  d.sd = (double)d.s / sg;
  if (dg.lattice > 0.0) {
    d.m0 = (int)floor(d.cxd / dg.lattice + 0.5);
    d.n0 = (int)floor(d.cyd / dg.lattice + 0.5);
  }
}

// Tuples whose loops can reach a tube core with the accurate FP32 Gaussian
// node (scaled units, sigma != 1): the loops are small (s' = sigma sqrt(T)
// < 1/2) and some point can come within |z'| < 0.75 (u' < 0.56) of the
// nearest tube axis, given the ensemble's largest radius rmax.  The plan
// puts such tuples in lane groups of their own; the kernels decide per image
// tube and wavefront (the same test on each lane).  The plain quotient's phase error for a loop in the
// core falls roughly as 1/s'^2: ~4e-7 relative in W - 1 at s' = 0.19, ~1e-8
// at s' = 0.6 (5120 loops x 4096 points, lambda = 2); from s' = 1/2 up it is
// below 1e-7, and loops away from the core have no nodes there.
static int tuple_needs_accurate(const DevGeom& dg, const DevTuple& d, double sg, float rmax) {
  if (dg.kind != WFL_GAUSSIAN || sg == 1.0) return 0;
  const double s = d.s;
  if (!(s < 0.5)) return 0;
  double dx = d.cx, dy = d.cy;
  if (dg.lattice > 0.0) {
    const double L = dg.latticef;
    dx -= L * floor(dx / L + 0.5);
    dy -= L * floor(dy / L + 0.5);
  }
  return (sqrt(dx * dx + dy * dy) - s * (double)rmax < 0.75) ? 1 : 0;
}

extern "C" int wfl_effective_tuple(const wfl_geometry* g, double cx, double cy, double T, double out[6]) {
  if (!g || !out || !(g->l2 > 0.0) || !(T > 0.0)) { wfl_set_error("wfl_effective_tuple: bad arguments"); return WFL_ERR_ARG; }
  DevGeom dg;
  DevTuple dt;
  fill_geom(*g, nullptr, true, dg);
  fill_tuple(unit_scale(*g, true), dg, 0, cx, cy, sqrt(T), 0, dt);
  out[0] = dg.l2; out[1] = dg.F; out[2] = dt.cxd; out[3] = dt.cyd; out[4] = dt.sd; out[5] = dg.lattice;
  return WFL_OK;
}

static int check_geom(const wfl_geometry& g, int nspline_tables, const wfl_spline_table* tabs) {
  if (g.kind < WFL_STEP || g.kind > WFL_FIXFLUX_FF) { wfl_set_error("geometry: unknown profile %d", g.kind); return WFL_ERR_ARG; }
  if (!(g.l2 > 0.0) || !isfinite(g.F)) { wfl_set_error("geometry: need l2 > 0 and finite F"); return WFL_ERR_ARG; }
  if (g.lattice > 0.0 && (g.images < 0 || g.images > 16)) { wfl_set_error("geometry: images must be in [0, 16]"); return WFL_ERR_ARG; }
  if ((g.kind == WFL_PERIODIC || g.kind == WFL_FIXFLUX || g.kind == WFL_FIXFLUX_FF) && !(g.a > 0.0)) {
    wfl_set_error("geometry: profile %d needs tubedist a > 0", g.kind); return WFL_ERR_ARG;
  }
  if (g.kind == WFL_SPLINE && (g.spline_table < 0 || g.spline_table >= nspline_tables || !tabs)) {
    wfl_set_error("geometry: bad spline table index %d", g.spline_table);
    return WFL_ERR_ARG;
  }
  return WFL_OK;
}

// Every table of a batch goes into the plan's blob, used or not.
static int check_spline_tables(int nspline_tables, const wfl_spline_table* tabs) {
  if (nspline_tables < 0 || (nspline_tables > 0 && !tabs)) { wfl_set_error("spline tables missing"); return WFL_ERR_ARG; }
  for (int s = 0; s < nspline_tables; ++s) {
    const wfl_spline_table& t = tabs[s];
    if (t.n < 2 || !t.rho2 || !t.coefs) { wfl_set_error("spline table %d: need n >= 2 and data", s); return WFL_ERR_ARG; }
    if (!isfinite(t.rho2[0])) { wfl_set_error("spline table %d: knot 0 not finite", s); return WFL_ERR_ARG; }
    for (int i = 1; i < t.n; ++i)
      if (!(t.rho2[i] > t.rho2[i - 1]) || !isfinite(t.rho2[i])) {
        wfl_set_error("spline table %d: knots not increasing at %d", s, i);
        return WFL_ERR_ARG;
      }
  }
  return WFL_OK;
}

// ------------------------------------------------------ plans
struct wfl_plan_s {
  wfl_batch b{};
  wfl_loops* L = nullptr;
  int ngroups = 0, mapping = 1, prof_gauss = 1, ntasks = 0;
  int cpg = 1;                   // tasks per loop group (lane-per-tuple mappings)
  int ntasks1 = 0;               // tasks with one task per loop group
  std::vector<int> lg_nloops, lg_images;   // per lane group: loops, image tubes
  size_t tasks_cap = 0;          // capacity of d_tasks
  double* d_gpart = nullptr;     // chunk sums [ntup][ngroups][cpg][3] when cpg > 1
  size_t gpart_cap = 0;
  std::vector<DevTuple> tups;
  DevTuple* d_tups = nullptr;
  DevGeom* d_geoms = nullptr;
  int* d_lanes = nullptr;
  Task* d_tasks = nullptr;
  float4* d_sblob = nullptr;     // spline blob
  int sblob_f4 = 0, lds_splines = 0;
  int uni_splines = 0, sblob_uni_f4 = 0;   // uniform-grid spline rows (FP32 physical)
  int uni_origin = 0;                      // those rows expanded about u = 0
  int axis = 0;                            // axis-form records (see wfl_plan_axis_form)
  bool has_local = false;                  // some tuples take the local Taylor node (lane-per-tuple kernels)
  double* d_gsum = nullptr;
  double* d_perloop = nullptr;
  double* d_wm1 = nullptr;
  unsigned long long* d_nonfinite = nullptr;
  int* d_queue = nullptr;
  int wpb = 4, persistent = 0, grid = 0;   // launch configuration (autotuner)
  bool tuned = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  hipStream_t gstream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;   // bracket the main kernel (timed runs)
  bool timed = false;
  hipStream_t last = nullptr;              // stream of the latest run (wfl_plan_read waits for it)
  // device buffers carved from caller-owned memory instead of hipMalloc
  // (wfl_eval_batch_noalloc): never freed or grown by the plan
  char* arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
};

extern "C" void wfl_plan_destroy(wfl_plan* p) {
  if (!p) return;
  if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
  if (p->graph) (void)hipGraphDestroy(p->graph);
  if (p->gstream) (void)hipStreamDestroy(p->gstream);
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  void* bufs[] = {p->d_tups, p->d_geoms, p->d_lanes, p->d_tasks, p->d_sblob, p->d_gpart,
                  p->d_gsum, p->d_perloop, p->d_wm1, p->d_nonfinite, p->d_queue};
  if (!p->arena)
    for (void* b : bufs) if (b) (void)hipFree(b);
  delete p;
}

// A device buffer of the plan: hipMalloc, or the next 256-byte aligned
// piece of the plan's arena (an error when the arena is too small).
static int plan_alloc(wfl_plan* p, void** dst, size_t bytes) {
  if (!p->arena) {
    HIP_TRY(hipMalloc(dst, bytes));
    return WFL_OK;
  }
  const size_t at = (p->arena_used + 255) & ~(size_t)255;
  if (at > p->arena_cap || bytes > p->arena_cap - at) {
    wfl_set_error("wfl_eval_batch_noalloc: the batch needs more than the %zu-byte scratch "
                  "(%zu bytes in use, %zu more asked for)", p->arena_cap, at, bytes);
    return WFL_ERR_ARG;
  }
  *dst = p->arena + at;
  p->arena_used = at + bytes;
  return WFL_OK;
}

template <class T>
static int upload(wfl_plan* p, T** dst, const T* src, size_t n) {
  if (n == 0) n = 1;
  int rc = plan_alloc(p, (void**)dst, sizeof(T) * n);
  if (rc) return rc;
  if (src) HIP_TRY(hipMemcpy(*dst, src, sizeof(T) * n, hipMemcpyHostToDevice));
  return WFL_OK;
}

// Compute units of the current device.
static int device_cus() {
  int dev = 0, ncu = 0;
  (void)hipGetDevice(&dev);
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
  return ncu;
}

// Tasks of the lane-per-tuple mapping with each loop group split into cpg
// chunks of ceil(group / cpg) loops (cpg = 1: one task per lane group and
// loop group), longest first (loops x image tubes) so the tail of the launch
// is short.
static std::vector<Task> plan_task_list(const wfl_plan* p, int cpg) {
  const int group = p->b.group, chunk = (group + cpg - 1) / cpg;
  const int nlg = (int)p->lg_nloops.size();
  std::vector<Task> tasks;
  for (int q = 0; q < p->ngroups; ++q)
    for (int c = 0; c < cpg; ++c)
      for (int lg = 0; lg < nlg; ++lg) {
        int l0 = q * group + c * chunk;
        int l1 = std::min(std::min(l0 + chunk, (q + 1) * group), p->lg_nloops[lg]);
        if (l1 > l0) tasks.push_back(Task{lg, l0, l1 - l0, q * cpg + c});
      }
  std::stable_sort(tasks.begin(), tasks.end(), [&](const Task& u, const Task& v) {
    return (long long)u.nloop * p->lg_images[u.lanegroup] > (long long)v.nloop * p->lg_images[v.lanegroup];
  });
  return tasks;
}

// Installs the task list for cpg chunks per group (and the chunk-sum buffer,
// cleared: slots that no task writes must read as zero).
static int plan_set_tasks(wfl_plan* p, int cpg) {
  std::vector<Task> tasks = plan_task_list(p, cpg);
  if (tasks.size() > p->tasks_cap || !p->d_tasks) {
    if (p->d_tasks && !p->arena) (void)hipFree(p->d_tasks);
    p->d_tasks = nullptr; p->tasks_cap = 0;
    int rc = plan_alloc(p, (void**)&p->d_tasks, sizeof(Task) * std::max<size_t>(tasks.size(), 1));
    if (rc) return rc;
    p->tasks_cap = std::max<size_t>(tasks.size(), 1);
  }
  if (!tasks.empty()) HIP_TRY(hipMemcpy(p->d_tasks, tasks.data(), sizeof(Task) * tasks.size(), hipMemcpyHostToDevice));
  p->ntasks = (int)tasks.size();
  if (cpg > 1) {
    const size_t n = (size_t)std::max(p->b.ntup, 1) * p->ngroups * cpg * 3;
    if (n > p->gpart_cap) {
      if (p->d_gpart && !p->arena) (void)hipFree(p->d_gpart);
      p->d_gpart = nullptr; p->gpart_cap = 0;
      int rc = plan_alloc(p, (void**)&p->d_gpart, sizeof(double) * n);
      if (rc) return rc;
      p->gpart_cap = n;
    }
    HIP_TRY(hipMemset(p->d_gpart, 0, sizeof(double) * n));
  }
  p->cpg = cpg;
  return WFL_OK;
}

// Tasks per loop group for the lane-per-tuple mappings.  Small batches:
// enough tasks for about 32 wavefronts per CU.  Every batch: tasks of at
// most kTaskLoops loops, so that the wavefronts sharing a loop group start
// it close together and stay within a few loops of each other, and its
// point records are reused from L2 instead of being fetched once per lane
// group (C2, 128-loop groups: 700 ms with one task per group, 649 ms with
// 16-loop tasks, 647 to 649 ms with 8, 646 ms with 4).  The chunk sums cost
// 24 B per tuple, group and chunk; they are kept under 2 GB (under half the
// arena for a plan in caller-owned memory).
static const int kTaskLoops = 8;

static int plan_default_cpg(const wfl_plan* p, int mapping) {
  if (mapping == 2) return 1;
  const int group = p->b.group;
  const double slot_bytes = 24.0 * (double)std::max(p->b.ntup, 1) * (double)p->ngroups;
  const double slot_cap = p->arena ? 0.5 * (double)p->arena_cap : 2.0e9;
  auto fits = [&](int c) { return c <= 32 && c <= group && slot_bytes * c <= slot_cap; };
  const long long want = 32LL * device_cus();
  int cpg = 1;
  while (fits(2 * cpg) && (long long)p->ntasks1 * cpg < want) cpg *= 2;
  while (fits(2 * cpg) && (group + cpg - 1) / cpg > kTaskLoops) cpg *= 2;
  return cpg;
}

#define TRY(expr) do { int rc_ = (expr); if (rc_) { wfl_plan_destroy(p); return rc_; } } while (0)

static int plan_create_impl(wfl_plan** out, const wfl_batch* b, wfl_loops* L, char* arena, size_t arena_cap);

extern "C" int wfl_plan_create(wfl_plan** out, const wfl_batch* b, wfl_loops* L) {
  return plan_create_impl(out, b, L, nullptr, 0);
}

static int plan_create_impl(wfl_plan** out, const wfl_batch* b, wfl_loops* L, char* arena, size_t arena_cap) {
  if (!out || !b || !L) { wfl_set_error("wfl_plan_create: NULL argument"); return WFL_ERR_ARG; }
  *out = nullptr;
  if (b->mode != WFL_MODE_REFERENCE && b->mode != WFL_MODE_PHYSICAL) { wfl_set_error("bad mode %d", b->mode); return WFL_ERR_ARG; }
  if (b->precision != WFL_FP64 && b->precision != WFL_MIXED) { wfl_set_error("bad precision %d", b->precision); return WFL_ERR_ARG; }
  if (b->nsub < 2 || b->nsub % 2 != 0 || b->nsub > 4096) { wfl_set_error("nsub must be even and in [2, 4096] (got %d)", b->nsub); return WFL_ERR_ARG; }
  if (b->group < 1) { wfl_set_error("group must be >= 1"); return WFL_ERR_ARG; }
  if (b->ngeom < 1 || !b->geoms) { wfl_set_error("need at least one geometry"); return WFL_ERR_ARG; }
  if (b->ntup < 0 || (b->ntup > 0 && (!b->tup_geom || !b->tup_cx || !b->tup_cy || !b->tup_T))) {
    wfl_set_error("tuple arrays missing"); return WFL_ERR_ARG;
  }
  if (b->mapping < 0 || b->mapping > 3) { wfl_set_error("bad mapping %d", b->mapping); return WFL_ERR_ARG; }
  { int rc = check_spline_tables(b->nspline, b->splines); if (rc) return rc; }
  for (int g = 0; g < b->ngeom; ++g) { int rc = check_geom(b->geoms[g], b->nspline, b->splines); if (rc) return rc; }
  for (int t = 0; t < b->ntup; ++t) {
    if (b->tup_geom[t] < 0 || b->tup_geom[t] >= b->ngeom) { wfl_set_error("tuple %d: geometry index out of range", t); return WFL_ERR_ARG; }
    if (!(b->tup_T[t] > 0.0) || !isfinite(b->tup_T[t]) || !isfinite(b->tup_cx[t]) || !isfinite(b->tup_cy[t])) {
      wfl_set_error("tuple %d: need finite x_cm and T > 0", t); return WFL_ERR_ARG;
    }
  }
  int rc0 = require_device();
  if (rc0) return rc0;
  float rmax = 0.0f;
  if (b->precision == WFL_MIXED) {
    rc0 = loops_rmax(L, &rmax, !arena);
    if (rc0) return rc0;
  }

  wfl_plan* p = new wfl_plan();
  p->arena = arena;
  p->arena_cap = arena ? arena_cap : 0;
  p->b = *b;
  p->b.geoms = nullptr; p->b.splines = nullptr; p->b.tup_geom = nullptr;
  p->b.tup_cx = p->b.tup_cy = p->b.tup_T = nullptr; p->b.geom_nloops = nullptr;
  p->L = L;
  const int Nl = L->Nl;
  p->ngroups = (Nl + b->group - 1) / b->group;

  // splines: one blob of every table
  SplineBlob sblob;
  {
    std::vector<const float*> kp, cp;
    std::vector<int> np;
    for (int s = 0; s < b->nspline; ++s) {
      kp.push_back(b->splines[s].rho2); cp.push_back(b->splines[s].coefs); np.push_back(b->splines[s].n);
    }
    // the uniform-grid rows reach the largest u of each table's tuples
    std::vector<double> umax((size_t)std::max(b->nspline, 1), 0.0);
    for (int t = 0; t < b->ntup; ++t) {
      const wfl_geometry& G = b->geoms[b->tup_geom[t]];
      if (G.kind != WFL_SPLINE) continue;
      const int tab = b->nspline == 1 ? 0 : G.spline_table;
      umax[tab] = std::max(umax[tab], node_u_bound(G, b->tup_cx[t], b->tup_cy[t], sqrt(b->tup_T[t]), rmax));
    }
    build_spline_blob(b->nspline, kp.data(), cp.data(), np.data(), b->mode, sblob,
                      b->precision == WFL_MIXED ? umax.data() : nullptr, b->spinor != 0);
  }
  std::vector<DevGeom> geoms(b->ngeom);
  p->prof_gauss = 1;
  bool all_spline = true;
  for (int g = 0; g < b->ngeom; ++g) {
    const wfl_geometry& G = b->geoms[g];
    fill_geom(G, &sblob, true, geoms[g]);
    if (b->precision == WFL_MIXED && G.kind == WFL_GAUSSIAN) { geoms[g].acc_on = 1; geoms[g].acc_rmax = rmax; }
    if (G.kind != WFL_GAUSSIAN) p->prof_gauss = 0;
    if (G.kind != WFL_SPLINE) all_spline = false;
  }
  p->sblob_f4 = (int)sblob.data.size();
  bool all_fast = true;
  for (const SplineTab& t : sblob.tab) all_fast = all_fast && t.fast;
  p->lds_splines = all_spline && all_fast && sblob.data.size() * sizeof(float4) <= kSplineLdsBytes;
  // FP32 physical spline-only plans on uniform-grid tables: direct-index rows in LDS
  p->uni_splines = all_spline && sblob.all_uni && (size_t)sblob.uni_f4 * sizeof(float4) <= kSplineLdsBytes;
  p->sblob_uni_f4 = sblob.uni_f4;
  p->uni_origin = p->uni_splines && sblob.uni_origin;
  if (p->uni_splines) p->lds_splines = 1;
  // tuples: FP32 kernel-unit values and the effective doubles
  p->tups.resize(b->ntup);
  for (int t = 0; t < b->ntup; ++t) {
    const int gi = b->tup_geom[t];
    int nl = Nl;
    if (b->geom_nloops) nl = std::max(0, std::min(Nl, b->geom_nloops[gi]));
    const double sg = unit_scale(b->geoms[gi], true);
    fill_tuple(sg, geoms[gi], gi, b->tup_cx[t], b->tup_cy[t], sqrt(b->tup_T[t]), nl, p->tups[t]);
    if (b->precision == WFL_MIXED) p->tups[t].acc = tuple_needs_accurate(geoms[gi], p->tups[t], sg, rmax);
  }
  // the axis form: mixed physical n = 2 batches whose tuples all have
  // x_cm on the x axis and no lattice geometry
  {
    bool ax = b->precision == WFL_MIXED && b->mode == WFL_MODE_PHYSICAL && b->nsub == 2 && b->ntup > 0;
    for (int g = 0; g < b->ngeom && ax; ++g) ax = !(b->geoms[g].lattice > 0.0);
    for (int t = 0; t < b->ntup && ax; ++t) ax = p->tups[t].cy == 0.0f;
    p->axis = ax ? 1 : 0;
  }
  // local Taylor node (Gaussian axis form): tuples whose nodes stay near the
  // tube-relative centre for loops up to the 90th percentile of the
  // ensemble's loop radii (the kernel checks each loop's own radius)
  std::vector<double> ldmax(b->ntup, 0.0);
  if (p->axis && p->prof_gauss) {
    TRY(loops_info(L, !arena));
    std::vector<double> rs = L->radius;
    std::sort(rs.begin(), rs.end());
    const double R = rs[std::min(rs.size() - 1, (size_t)(0.9 * (double)rs.size()))];
    for (int t = 0; t < b->ntup; ++t) {
      set_local(p->tups[t], R);
      ldmax[t] = p->tups[t].loc ? local_dmax(p->tups[t], R) : 0.0;
    }
  }
  // lane groups: tuples sorted by node kind (local Taylor node, accurate
  // Gaussian node, plain node) and geometry, each such set padded to 64
  // lanes.  A wavefront holds one node kind and one loop count (the kernels
  // run every lane over its group's loops); it holds one geometry except in
  // the Gaussian axis form, whose kernels read every geometry parameter per
  // lane and use no image window, so that there only the node kinds and loop
  // counts are padded, not each geometry's.  Local-node tuples take the
  // local path whatever their accurate flag; within a kind they are ordered
  // by geometry and node range.
  const bool cross_geom = p->axis && p->prof_gauss;
  std::vector<int> order(b->ntup);
  for (int t = 0; t < b->ntup; ++t) order[t] = t;
  auto kind_of = [&](int t) { return p->tups[t].loc ? 2LL : (long long)p->tups[t].acc; };
  auto key = [&](int t) {
    return cross_geom ? 4LL * p->tups[t].nloops + kind_of(t) : 4LL * p->tups[t].geom + kind_of(t);
  };
  std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
    if (key(a) != key(c)) return key(a) < key(c);
    if (p->tups[a].geom != p->tups[c].geom) return p->tups[a].geom < p->tups[c].geom;
    return ldmax[a] < ldmax[c];
  });
  std::vector<int> lanes;
  std::vector<int> lg_nloops;
  for (size_t k = 0; k < order.size();) {
    const long long g = key(order[k]);
    size_t e = k;
    while (e < order.size() && key(order[e]) == g) ++e;
    for (size_t s = k; s < e; s += 64) {
      int nl = 0;
      for (size_t q = 0; q < 64; ++q) {
        int v = (s + q < e) ? order[s + q] : -1;
        lanes.push_back(v);
        if (v >= 0) nl = std::max(nl, p->tups[v].nloops);
      }
      lg_nloops.push_back(nl);
    }
    k = e;
  }
  int nlg = (int)lg_nloops.size();
  std::vector<int> lg_images(nlg, 1);
  for (int lg = 0; lg < nlg; ++lg) {
    int t0 = lanes[(size_t)lg * 64];
    const wfl_geometry& G = b->geoms[p->tups[t0].geom];
    if (G.lattice > 0.0) lg_images[lg] = (2 * G.images + 1) * (2 * G.images + 1);
  }
  p->lg_nloops = lg_nloops;
  p->lg_images = lg_images;
  p->ntasks = p->ntasks1 = (int)plan_task_list(p, 1).size();
  // mapping (autotune heuristic: lane-per-tuple needs enough lane groups)
  // (mapping 3 is opt-in or autotuned only: its per-loop buffer takes 48 B per
  // tuple and loop, 13.6 GB on the 55296-tuple, 5120-loop spinor workload,
  // where it runs 630 against 648 ms for mapping 1)
  int mapping = b->mapping;
  int nloc = 0;
  for (const DevTuple& t : p->tups) nloc += t.loc;
  p->has_local = nloc > 0;
  if (mapping == 0) {
    double util = b->ntup > 0 ? (double)b->ntup / (64.0 * std::max(nlg, 1)) : 0.0;
    mapping = (util >= 0.5 && p->ntasks >= 512) ? 1 : 2;
    // the local Taylor node exists in the lane-per-tuple kernels only
    if (p->has_local) mapping = 1;
  }
  if (mapping == 3 && !p->prof_gauss) mapping = 1;
  p->mapping = mapping;

  TRY(loops_records(L, b->precision, !arena));
  if (p->axis) TRY(loops_records_axis(L, !arena));
  TRY(upload(p, &p->d_tups, p->tups.data(), p->tups.size()));
  TRY(upload(p, &p->d_geoms, geoms.data(), geoms.size()));
  TRY(upload(p, &p->d_lanes, lanes.data(), lanes.size()));
  // loop groups split into chunk tasks (the autotuner measures alternatives)
  TRY(plan_set_tasks(p, plan_default_cpg(p, mapping)));
  TRY(upload(p, &p->d_sblob, sblob.data.data(), sblob.data.size()));
  size_t ng3 = (size_t)std::max(b->ntup, 1) * p->ngroups * 3;
  TRY(upload<double>(p, &p->d_gsum, nullptr, ng3));
  {
    hipError_t e = hipMemset(p->d_gsum, 0, sizeof(double) * ng3);
    if (e != hipSuccess) { wfl_set_error("hipMemset: %s", hipGetErrorString(e)); wfl_plan_destroy(p); return WFL_ERR_DEVICE; }
  }
  if (b->want_perloop || mapping == 3) {
    size_t n6 = (size_t)std::max(b->ntup, 1) * Nl * 6;
    TRY(upload<double>(p, &p->d_perloop, nullptr, n6));
    hipError_t e = hipMemset(p->d_perloop, 0, sizeof(double) * n6);
    if (e != hipSuccess) { wfl_set_error("hipMemset: %s", hipGetErrorString(e)); wfl_plan_destroy(p); return WFL_ERR_DEVICE; }
  }
  if (mapping == 2) TRY(upload<double>(p, &p->d_wm1, nullptr, (size_t)std::max(b->ntup, 1) * Nl * 3));
  TRY(upload<unsigned long long>(p, &p->d_nonfinite, nullptr, 1));
  TRY(upload<int>(p, &p->d_queue, nullptr, 1));
  {
    int dev = 0, ncu = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu <= 0) ncu = 256;
    p->grid = ncu * 32 / p->wpb;     // persistent grid: 32 wavefronts per CU
  }
  {
    hipError_t e = hipMemset(p->d_nonfinite, 0, sizeof(unsigned long long));
    if (e != hipSuccess) { wfl_set_error("hipMemset: %s", hipGetErrorString(e)); wfl_plan_destroy(p); return WFL_ERR_DEVICE; }
  }
  *out = p;
  return WFL_OK;
}

static void fill_launch(const wfl_plan* p, WflLaunch& a, hipStream_t s) {
  memset(&a, 0, sizeof a);
  a.rec = p->b.precision == WFL_MIXED ? (const void*)p->L->rec32 : (const void*)p->L->rec64;
  a.N = p->L->N; a.Nl = p->L->Nl;
  a.tups = p->d_tups; a.geoms = p->d_geoms; a.ntup = p->b.ntup;
  a.lanes = p->d_lanes; a.tasks = p->d_tasks; a.ntasks = p->ntasks;
  a.nslots = p->mapping == 1 ? p->ngroups * p->cpg : p->ngroups;
  a.sblob = p->d_sblob; a.sblob_f4 = p->sblob_f4; a.lds_splines = p->lds_splines;
  a.uni_splines = p->uni_splines; a.sblob_uni_f4 = p->sblob_uni_f4; a.uni_origin = p->uni_origin;
  a.axis = p->axis; a.rec_axis = p->L->reca;
  a.linfo = p->L->info;
  a.prof_gauss = p->prof_gauss; a.spinor = p->b.spinor ? 1 : 0; a.mode = p->b.mode; a.nsub = p->b.nsub;
  a.split = p->mapping == 3;
  a.gsum = (p->mapping == 1 && p->cpg > 1) ? p->d_gpart : p->d_gsum;
  // per-loop outputs only when asked for or when the split finalize reads
  // them; W - 1 per loop only for the wave-per-loop group sums
  a.perloop = (p->b.want_perloop || p->mapping == 3) ? p->d_perloop : nullptr;
  a.wm1 = p->mapping == 2 ? p->d_wm1 : nullptr;
  a.nonfinite = p->d_nonfinite;
  a.wpb = p->wpb;
  a.persistent = p->persistent;
  a.grid = p->grid;
  a.queue = p->d_queue;
  a.stream = s;
}

// A stream under graph capture records work rather than running it.
static bool hipStreamIsCapturing_safe(hipStream_t s) {
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (s == nullptr) return false;
  return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

extern "C" int wfl_plan_run(wfl_plan* p, void* stream) {
  if (!p) { wfl_set_error("wfl_plan_run: NULL plan"); return WFL_ERR_STATE; }
  if (p->b.ntup == 0) return WFL_OK;
  hipStream_t s = (hipStream_t)stream;
  if (!hipStreamIsCapturing_safe(s)) p->last = s;
  hipLaunchKernelGGL(k_zero_counters, dim3(1), dim3(64), 0, s, p->d_nonfinite, p->d_queue);
  HIP_TRY(hipGetLastError());
  WflLaunch a;
  fill_launch(p, a, s);
  int e;
  if (p->timed) HIP_TRY(hipEventRecord(p->ev0, s));
  if (p->mapping == 1 || p->mapping == 3) {
    e = p->b.precision == WFL_MIXED ? wfl_launch_tuple_f32(&a) : wfl_launch_tuple_f64(&a);
    if (p->timed && e == 0) e = (int)hipEventRecord(p->ev1, s);
    if (e == 0 && p->mapping == 1 && p->cpg > 1) {
      long long n = (long long)p->b.ntup * p->ngroups;
      hipLaunchKernelGGL(k_sum_slots, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, p->d_gpart, n, p->cpg,
                         p->d_gsum);
      e = (int)hipGetLastError();
    }
    if (e == 0 && p->mapping == 3) {
      long long n = (long long)p->b.ntup * p->ngroups;
      hipLaunchKernelGGL(k_finalize_groups, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, p->d_perloop,
                         p->d_tups, p->b.ntup, p->L->Nl, p->b.group, p->ngroups, p->b.mode, p->b.spinor,
                         p->d_gsum, p->d_nonfinite);
      e = (int)hipGetLastError();
    }
  } else {
    e = p->b.precision == WFL_MIXED ? wfl_launch_loop_f32(&a) : wfl_launch_loop_f64(&a);
    if (p->timed && e == 0) e = (int)hipEventRecord(p->ev1, s);
    if (e == 0) {
      long long n = (long long)p->b.ntup * p->ngroups;
      hipLaunchKernelGGL(k_group_sums, dim3((unsigned)((n + 127) / 128)), dim3(128), 0, s, p->d_wm1, p->d_tups,
                         p->b.ntup, p->L->Nl, p->b.group, p->ngroups, p->d_gsum);
      e = (int)hipGetLastError();
    }
  }
  if (e) { wfl_set_error("kernel launch failed: %s", hipGetErrorString((hipError_t)e)); return WFL_ERR_DEVICE; }
  return WFL_OK;
}

extern "C" int wfl_plan_run_timed(wfl_plan* p, void* stream, float* kernel_ms) {
  if (!p || !kernel_ms) { wfl_set_error("wfl_plan_run_timed: NULL argument"); return WFL_ERR_ARG; }
  *kernel_ms = 0.0f;
  if (!p->ev0) HIP_TRY(hipEventCreate(&p->ev0));
  if (!p->ev1) HIP_TRY(hipEventCreate(&p->ev1));
  p->timed = true;
  int rc = wfl_plan_run(p, stream);
  p->timed = false;
  if (rc) return rc;
  HIP_TRY(hipEventSynchronize(p->ev1));
  HIP_TRY(hipEventElapsedTime(kernel_ms, p->ev0, p->ev1));
  return WFL_OK;
}

// ------------------------------------------------------ autotuner
// Measured choice of the launch configuration: mapping (lane per tuple
// inline / split finalize, wave per loop for small batches), static or
// persistent grid, wavefronts per block, and for the lane-per-tuple mappings
// the number of tasks each loop group is split into (small batches: more,
// shorter tasks fill the machine).  Every candidate run is a full, valid
// evaluation; the fastest configuration is kept on the plan and remembered
// for later plans of the same shape in this process.
struct TuneKey {
  int ntup, Nl, N, precision, spinor, mode, nsub, prof_gauss, lds_splines, ntasks1;
  bool operator==(const TuneKey& o) const { return memcmp(this, &o, sizeof o) == 0; }
};
struct TuneVal { int mapping, persistent, wpb, cpg; };
static std::vector<std::pair<TuneKey, TuneVal>> g_tuned;   // guarded by g_tuned_mu
static std::mutex g_tuned_mu;

static int plan_apply(wfl_plan* p, int mapping, int persistent, int wpb, int cpg) {
  if (mapping == 3 && !p->d_perloop) {
    size_t n6 = (size_t)std::max(p->b.ntup, 1) * p->L->Nl * 6;
    HIP_TRY(hipMalloc(&p->d_perloop, sizeof(double) * n6));
  }
  if (mapping == 2 && !p->d_wm1) HIP_TRY(hipMalloc(&p->d_wm1, sizeof(double) * (size_t)std::max(p->b.ntup, 1) * p->L->Nl * 3));
  if (mapping == 2) cpg = 1;
  if (cpg != p->cpg) {
    int rc = plan_set_tasks(p, cpg);
    if (rc) return rc;
  }
  // buffers the configuration no longer reads
  if (mapping != 3 && !p->b.want_perloop && p->d_perloop) { (void)hipFree(p->d_perloop); p->d_perloop = nullptr; }
  if (mapping != 2 && p->d_wm1) { (void)hipFree(p->d_wm1); p->d_wm1 = nullptr; }
  p->mapping = mapping;
  p->persistent = persistent;
  p->wpb = wpb;
  p->grid = device_cus() * 32 / wpb;
  if (p->gexec) { (void)hipGraphExecDestroy(p->gexec); p->gexec = nullptr; }
  if (p->graph) { (void)hipGraphDestroy(p->graph); p->graph = nullptr; }
  return WFL_OK;
}

static int tune_code(const TuneVal& v) { return v.mapping * 1000 + v.persistent * 100 + v.wpb * 10 + v.cpg; }

extern "C" int wfl_plan_autotune(wfl_plan* p, void* stream, int* chosen) {
  if (!p) { wfl_set_error("wfl_plan_autotune: NULL plan"); return WFL_ERR_STATE; }
  TuneKey key;
  memset(&key, 0, sizeof key);
  key.ntup = p->b.ntup; key.Nl = p->L->Nl; key.N = p->L->N; key.precision = p->b.precision;
  key.spinor = p->b.spinor; key.mode = p->b.mode; key.nsub = p->b.nsub; key.prof_gauss = p->prof_gauss;
  key.lds_splines = p->lds_splines + 2 * p->uni_splines + 4 * p->axis + 8 * p->uni_origin;
  key.ntasks1 = p->ntasks1;
  {
    bool hit = false;
    TuneVal v{};
    {
      std::lock_guard<std::mutex> lk(g_tuned_mu);
      for (auto& kv : g_tuned)
        if (kv.first == key) { v = kv.second; hit = true; break; }
    }
    if (hit) {
      int rc = plan_apply(p, v.mapping, v.persistent, v.wpb, v.cpg);
      if (rc) return rc;
      if (chosen) *chosen = tune_code(v);
      p->tuned = true;
      return WFL_OK;
    }
  }
  // every chunk count when one task per (lane group, loop group) gives
  // fewer than 4 wavefronts per SIMD; else one task per group and the
  // creation-time choice (short tasks for L2 reuse)
  std::vector<int> cpgs{1};
  const bool few = (long long)p->ntasks1 < 16LL * device_cus();
  const int cpg_lane = plan_default_cpg(p, 1);
  for (int c : {2, 4, 8, 16, 32})
    if ((few || c == cpg_lane) && c <= p->b.group &&
        24.0 * c * (double)std::max(p->b.ntup, 1) * (double)p->ngroups <= 2.0e9) cpgs.push_back(c);
  const double perloop_bytes = 48.0 * (double)std::max(p->b.ntup, 1) * (double)p->L->Nl;
  std::vector<TuneVal> cand;
  for (int c : cpgs)
    for (int pers = 0; pers <= 1; ++pers)
      for (int wpb : {2, 4}) {
        cand.push_back(TuneVal{1, pers, wpb, c});
        if (p->prof_gauss && perloop_bytes <= 24.0e9) cand.push_back(TuneVal{3, pers, wpb, c});
      }
  // (wave per loop has no local Taylor node: it would change the numbers)
  if ((p->b.ntup < 2048 || few) && !p->has_local) cand.push_back(TuneVal{2, 0, 4, 1});
  TuneVal best = cand[0];
  float best_ms = 1e30f;
  int timed = 0;
  hipEvent_t e0, e1;
  HIP_TRY(hipEventCreate(&e0));
  HIP_TRY(hipEventCreate(&e1));
  int rc = WFL_OK;
  // one untimed run first (first-touch costs are not charged to a candidate)
  rc = plan_apply(p, cand[0].mapping, cand[0].persistent, cand[0].wpb, cand[0].cpg);
  if (!rc) rc = wfl_plan_run(p, stream);
  if (!rc && hipStreamSynchronize((hipStream_t)stream) != hipSuccess) rc = WFL_ERR_DEVICE;
  for (size_t k = 0; k < cand.size() && !rc; ++k) {
    const TuneVal& c = cand[k];
    // a candidate that cannot be set up or launched (allocation, launch
    // limits) is skipped, not fatal: the next one is measured
    if (plan_apply(p, c.mapping, c.persistent, c.wpb, c.cpg) != WFL_OK) continue;
    float ms = 0.0f;
    (void)hipEventRecord(e0, (hipStream_t)stream);
    if (wfl_plan_run(p, stream) != WFL_OK) continue;  // the whole run, every kernel of it
    (void)hipEventRecord(e1, (hipStream_t)stream);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
      wfl_set_error("wfl_plan_autotune: event timing failed");
      rc = WFL_ERR_DEVICE;
      break;
    }
    ++timed;
    if (ms < best_ms) { best_ms = ms; best = c; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if (!timed) { wfl_set_error("wfl_plan_autotune: no launch configuration could run"); return WFL_ERR_DEVICE; }
  rc = plan_apply(p, best.mapping, best.persistent, best.wpb, best.cpg);
  if (rc) return rc;
  rc = wfl_plan_run(p, stream);      // leave the plan's outputs from the chosen configuration
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    g_tuned.push_back({key, best});
  }
  p->tuned = true;
  if (chosen) *chosen = tune_code(best);
  return WFL_OK;
}

extern "C" int wfl_plan_config(const wfl_plan* p, int* mapping, int* persistent, int* wpb, int* chunks) {
  if (!p) { wfl_set_error("wfl_plan_config: NULL plan"); return WFL_ERR_ARG; }
  if (mapping) *mapping = p->mapping;
  if (persistent) *persistent = p->persistent;
  if (wpb) *wpb = p->wpb;
  if (chunks) *chunks = p->mapping == 2 ? 1 : p->cpg;
  return WFL_OK;
}

extern "C" int wfl_plan_set_config(wfl_plan* p, int mapping, int persistent, int wpb, int chunks) {
  if (!p || mapping < 1 || mapping > 3 || (persistent != 0 && persistent != 1) || (wpb != 1 && wpb != 2 && wpb != 4) ||
      chunks < 1 || chunks > 64 || chunks > p->b.group) {
    wfl_set_error("wfl_plan_set_config: bad configuration"); return WFL_ERR_ARG;
  }
  if (mapping == 3 && !p->prof_gauss) { wfl_set_error("wfl_plan_set_config: split finalize needs the Gaussian fast path"); return WFL_ERR_ARG; }
  return plan_apply(p, mapping, persistent, wpb, chunks);
}

extern "C" int wfl_plan_graph_run(wfl_plan* p, void* stream) {
  if (!p) { wfl_set_error("wfl_plan_graph_run: NULL plan"); return WFL_ERR_STATE; }
  if (!p->gexec) {
    if (!p->gstream) HIP_TRY(hipStreamCreateWithFlags(&p->gstream, hipStreamNonBlocking));
    HIP_TRY(hipStreamBeginCapture(p->gstream, hipStreamCaptureModeThreadLocal));
    int rc = wfl_plan_run(p, p->gstream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(p->gstream, &g);
    if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) { wfl_set_error("graph capture failed: %s", hipGetErrorString(e)); return WFL_ERR_DEVICE; }
    p->graph = g;
    HIP_TRY(hipGraphInstantiate(&p->gexec, p->graph, nullptr, nullptr, 0));
  }
  HIP_TRY(hipGraphLaunch(p->gexec, (hipStream_t)stream));
  p->last = (hipStream_t)stream;
  return WFL_OK;
}

extern "C" int wfl_plan_read(wfl_plan* p, double* gsum, double* perloop) {
  if (!p) { wfl_set_error("wfl_plan_read: NULL plan"); return WFL_ERR_STATE; }
  if (perloop && !(p->b.want_perloop && p->d_perloop)) {
    wfl_set_error("wfl_plan_read: per-loop outputs were not requested");
    return WFL_ERR_STATE;
  }
  // the copies are ordered after the latest run on its stream
  hipStream_t s = p->last;
  if (gsum && p->b.ntup > 0)
    HIP_TRY(hipMemcpyAsync(gsum, p->d_gsum, sizeof(double) * (size_t)p->b.ntup * p->ngroups * 3,
                           hipMemcpyDeviceToHost, s));
  if (perloop)
    HIP_TRY(hipMemcpyAsync(perloop, p->d_perloop, sizeof(double) * (size_t)p->b.ntup * p->L->Nl * 6,
                           hipMemcpyDeviceToHost, s));
  HIP_TRY(hipStreamSynchronize(s));
  return WFL_OK;
}

extern "C" int wfl_plan_ngroups(const wfl_plan* p) { return p ? p->ngroups : 0; }
extern "C" int wfl_plan_accurate_tuples(const wfl_plan* p) {
  if (!p) return 0;
  int n = 0;
  for (const DevTuple& t : p->tups) n += t.acc ? 1 : 0;
  return n;
}
extern "C" int wfl_plan_uniform_splines(const wfl_plan* p) {
  return p && p->uni_splines ? (p->uni_origin ? 2 : 1) : 0;
}

extern "C" int wfl_spline_uniform_form(const float* rho2, const float* coefs, int n, double umax, int spinor,
                                       float* rows, int max_rows, int* nrows) {
  if (nrows) *nrows = 0;
  if (!rho2 || !coefs || n < 3 || !(umax >= 0.0)) return 0;
  SplineBlob sb;
  const float* k[1] = {rho2};
  const float* c[1] = {coefs};
  const int nn[1] = {n};
  build_spline_blob(1, k, c, nn, WFL_MODE_PHYSICAL, sb, &umax, spinor != 0);
  if (!sb.all_uni) return 0;
  if (rows && max_rows > 0) {
    const SplineTab& st = sb.tab[0];
    const int m = std::min(max_rows, st.urows + 1);   // the guard row, then rows 0, 1, ...
    memcpy(rows, sb.data.data() + st.ucf - 1, sizeof(float4) * (size_t)m);
    if (nrows) *nrows = m;
  }
  return sb.uni_origin ? 2 : 1;
}

extern "C" int wfl_plan_axis_form(const wfl_plan* p) { return p ? p->axis : 0; }
extern "C" int wfl_plan_local_tuples(const wfl_plan* p) {
  if (!p) return 0;
  int n = 0;
  for (const DevTuple& t : p->tups) n += t.loc ? 1 : 0;
  return n;
}
extern "C" const double* wfl_plan_gsum_device(const wfl_plan* p) { return p ? p->d_gsum : nullptr; }
extern "C" int wfl_plan_mapping(const wfl_plan* p) { return p ? p->mapping : 0; }

extern "C" int wfl_plan_nonfinite(wfl_plan* p, long long* count) {
  if (!p || !count) { wfl_set_error("wfl_plan_nonfinite: NULL"); return WFL_ERR_ARG; }
  unsigned long long v = 0;
  HIP_TRY(hipMemcpy(&v, p->d_nonfinite, sizeof v, hipMemcpyDeviceToHost));
  *count = (long long)v;
  return WFL_OK;
}

extern "C" int wfl_eval_batch(const wfl_batch* b, wfl_loops* L, double* gsum, double* perloop) {
  wfl_plan* p = nullptr;
  int rc = wfl_plan_create(&p, b, L);
  if (rc) return rc;
  rc = wfl_plan_run(p, nullptr);
  if (!rc) rc = wfl_plan_read(p, gsum, perloop);
  wfl_plan_destroy(p);
  return rc;
}

// ------------------------------------------------------ no-allocation evaluation
// wfl_eval_batch for callers that cannot allocate device memory: the plan's
// device buffers (tuples, geometries, lane groups, tasks, spline blob, chunk
// and group sums, per-loop outputs, counters) are carved from one static
// device array per GPU.  Each GPU keeps its copy of the array until the
// process ends, and the array keeps the last call's contents: every buffer
// that must start at zero is cleared by the call.  Calls on one GPU share the
// array, so they are serialised by a lock (the call returns after its
// results are on the host).
static constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static constexpr int kScratchDevices = 64;
static void* g_scratch_ptr[kScratchDevices];   // device address of scratch, per device
static std::mutex g_scratch_mu;                // guards g_scratch_ptr and the arrays

// Address of the current device's copy of scratch (looked up once per
// device).  Host code uses this address, never the array's name, which on
// the host is only a stand-in for the symbol.
static int scratch_address(void** out) {
  int dev = 0;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kScratchDevices) {
    wfl_set_error("wfl_eval_batch_noalloc: device %d is beyond the %d devices served", dev, kScratchDevices);
    return WFL_ERR_DEVICE;
  }
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
      wfl_set_error("wfl_eval_batch_noalloc: scratch address on device %d: %s", dev, hipGetErrorString(e));
      return WFL_ERR_DEVICE;
    }
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return WFL_OK;
}

extern "C" int wfl_eval_batch_noalloc(const wfl_batch* b, wfl_loops* L, double* gsum, double* perloop) {
  if (!b || !L) { wfl_set_error("wfl_eval_batch_noalloc: NULL argument"); return WFL_ERR_ARG; }
  int rc = require_device();
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  void* base = nullptr;
  rc = scratch_address(&base);
  if (rc) return rc;
  wfl_plan* p = nullptr;
  rc = plan_create_impl(&p, b, L, (char*)base, kScratchBytes);
  if (rc) return rc;
  rc = wfl_plan_run(p, nullptr);
  if (!rc) rc = wfl_plan_read(p, gsum, perloop);
  wfl_plan_destroy(p);
  return rc;
}

// ------------------------------------------------------ Simpson sub-steps
// The smallest n = 2, 4, 8, ... whose group sums on a sample of the batch
// (every stride-th tuple, at most max_tuples; the first max_loops loops of
// each geometry, one group) differ from those at 2n by at most tol relative
// to the sample's largest |W - 1| sum.  Composite Simpson converges as n^-4,
// so that difference is 15/16 of the error of the n-point sums.  When no
// n <= nmax/2 passes, *nsub = nmax and *est_err is the last difference
// (the estimate for nmax/2; nmax itself is about 16 times closer).
extern "C" int wfl_autotune_nsub(const wfl_batch* b, wfl_loops* L, double tol, int nmax, int max_tuples,
                                 int max_loops, int* nsub, double* est_err) {
  if (!b || !L || !nsub || !(tol > 0.0) || nmax < 2 || max_tuples < 1 || max_loops < 1 || b->ntup < 1) {
    wfl_set_error("wfl_autotune_nsub: bad arguments");
    return WFL_ERR_ARG;
  }
  int Nl = 0, N = 0;
  int rc = wfl_loops_shape(L, &Nl, &N);
  if (rc) return rc;
  const int nl = std::min(Nl, max_loops);
  const int stride = (b->ntup + max_tuples - 1) / max_tuples;
  std::vector<int> tg, gl(b->ngeom);
  std::vector<double> cx, cy, tT;
  for (int t = 0; t < b->ntup; t += stride) {
    tg.push_back(b->tup_geom[t]);
    cx.push_back(b->tup_cx[t]);
    cy.push_back(b->tup_cy[t]);
    tT.push_back(b->tup_T[t]);
  }
  for (int g = 0; g < b->ngeom; ++g) gl[g] = b->geom_nloops ? std::min(nl, b->geom_nloops[g]) : nl;
  wfl_batch s = *b;
  s.ntup = (int)tg.size();
  s.tup_geom = tg.data(); s.tup_cx = cx.data(); s.tup_cy = cy.data(); s.tup_T = tT.data();
  s.geom_nloops = gl.data();
  s.group = nl;
  s.want_perloop = 0;
  s.mapping = 0;
  std::vector<double> prev, cur, all;
  double err = HUGE_VAL;
  int n = 2;
  for (;;) {
    s.nsub = n;
    wfl_plan* p = nullptr;
    rc = wfl_plan_create(&p, &s, L);
    if (rc) return rc;
    // group sums [ntup][ngroups][3]: the sample's loops all sit in group 0
    const int ng = wfl_plan_ngroups(p);
    all.assign((size_t)s.ntup * ng * 3, 0.0);
    rc = wfl_plan_run(p, nullptr);
    if (!rc) rc = wfl_plan_read(p, all.data(), nullptr);
    wfl_plan_destroy(p);
    if (rc) return rc;
    cur.assign((size_t)s.ntup * 3, 0.0);
    for (int t = 0; t < s.ntup; ++t)
      for (int k = 0; k < 3; ++k) cur[(size_t)t * 3 + k] = all[((size_t)t * ng) * 3 + k];
    if (!prev.empty()) {
      double dmax = 0.0, vmax = 0.0;
      for (size_t i = 0; i < cur.size(); ++i) {
        dmax = std::max(dmax, fabs(cur[i] - prev[i]));
        vmax = std::max(vmax, fabs(cur[i]));
      }
      err = vmax > 0.0 ? dmax / vmax : dmax;
      if (err <= tol) {
        *nsub = n / 2;
        if (est_err) *est_err = err;
        return WFL_OK;
      }
    }
    if (2 * n > nmax) break;
    prev = cur;
    n *= 2;
  }
  *nsub = n;
  if (est_err) *est_err = err;
  return WFL_OK;
}

// ------------------------------------------------------ EV() helper
// Loop copies of the caller's AoS device buffer, reused while the pointer
// and shape stay the same.
// The launch's geometry and tuple travel in one block: DevGeom at 0,
// DevTuple at kEvTupOff; staged in pinned host memory and sent with one
// asynchronous copy on the launch's stream.  Every EV() call waits for its
// results (the device-to-host copy of Wsscal) before it returns, so the
// next call may rewrite the staging block.
static constexpr size_t kEvTupOff = (sizeof(DevGeom) + 255) / 256 * 256;
static constexpr size_t kEvBlock = kEvTupOff + sizeof(DevTuple);
static struct {
  const void* aos = nullptr;
  int Nl = 0, N = 0;
  wfl_loops* L = nullptr;
  void* d_block = nullptr;       // DevGeom, then DevTuple at kEvTupOff
  void* h_stage = nullptr;       // pinned copy source of the same layout
  DevTuple* d_tup = nullptr;
  DevGeom* d_geom = nullptr;
  unsigned long long* d_nf = nullptr;
  float4* d_sblob = nullptr;     // spline blob built from the caller's tables
  size_t sblob_cap = 0;          // in float4
} g_ev;

extern "C" void wfl_internal_ev_release(void) {
  if (g_ev.L) wfl_loops_destroy(g_ev.L);
  if (g_ev.d_block) (void)hipFree(g_ev.d_block);
  if (g_ev.h_stage) (void)hipHostFree(g_ev.h_stage);
  if (g_ev.d_nf) (void)hipFree(g_ev.d_nf);
  if (g_ev.d_sblob) (void)hipFree(g_ev.d_sblob);
  g_ev.L = nullptr; g_ev.d_block = nullptr; g_ev.h_stage = nullptr;
  g_ev.d_tup = nullptr; g_ev.d_geom = nullptr; g_ev.d_nf = nullptr;
  g_ev.d_sblob = nullptr; g_ev.sblob_cap = 0;
  g_ev.aos = nullptr; g_ev.Nl = g_ev.N = 0;
}

extern "C" int wfl_internal_ev_launch(const void* dev_aos, int Nl, int N, const wfl_geometry* g,
                                      const float* knots_d, const float* coefs_d, int nspline,
                                      int mode, int precision, int nsub, int fermion,
                                      float cx, float cy, float rtT, void* ws_scal_d, void* ws_ferm_d) {
  if (!dev_aos || !g || !ws_scal_d || (fermion == 1 && !ws_ferm_d)) { wfl_set_error("EV: NULL device buffer in Wparams"); return WFL_ERR_ARG; }
  if (Nl < 1 || N < 4 || N % 4 != 0) { wfl_set_error("EV: need Nl >= 1 and Nppl a multiple of 4 (got %d, %d)", Nl, N); return WFL_ERR_ARG; }
  if (g->kind == WFL_SPLINE && (!knots_d || !coefs_d || nspline < 2)) { wfl_set_error("EV: spline profile needs rho2sp and flcoefs"); return WFL_ERR_ARG; }
  if (nsub < 2 || nsub % 2) { wfl_set_error("EV: nsub must be even"); return WFL_ERR_ARG; }
  if (!(g_ev.aos == dev_aos && g_ev.Nl == Nl && g_ev.N == N && g_ev.L)) {
    wfl_internal_ev_release();
    int rc = wfl_loops_from_aos_device(&g_ev.L, dev_aos, Nl, N);
    if (rc) return rc;
    HIP_TRY(hipMalloc(&g_ev.d_block, kEvBlock));
    HIP_TRY(hipHostMalloc(&g_ev.h_stage, kEvBlock, hipHostMallocDefault));
    g_ev.d_geom = (DevGeom*)g_ev.d_block;
    g_ev.d_tup = (DevTuple*)((char*)g_ev.d_block + kEvTupOff);
    HIP_TRY(hipMalloc(&g_ev.d_nf, sizeof(unsigned long long)));
    // the key last: a failure above leaves it clear, and the next call sets up again
    g_ev.aos = dev_aos; g_ev.Nl = Nl; g_ev.N = N;
  }
  int rc = loops_records(g_ev.L, precision);
  if (rc) return rc;
  // FP64: the reference's float parameters as they are; FP32 Gaussian:
  // the same parameters in the scaled units of the fast path.
  wfl_geometry gg = *g;
  gg.lattice = 0.0;
  gg.images = 0;
  const bool scaled = precision == WFL_MIXED;
  // spline: the caller's device tables, read back at every launch, as a
  // blob with its bucket table (EV() skips the launch, and this read, when
  // its inputs equal the previous launch's: see EV in intT.h)
  SplineBlob sb;
  if (g->kind == WFL_SPLINE) {
    std::vector<float> kn((size_t)nspline), cf((size_t)nspline * 4);
    HIP_TRY(hipMemcpy(kn.data(), knots_d, sizeof(float) * kn.size(), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(cf.data(), coefs_d, sizeof(float) * cf.size(), hipMemcpyDeviceToHost));
    for (int i = 1; i < nspline; ++i)
      if (!(kn[i] > kn[i - 1])) { wfl_set_error("EV: spline knots not increasing at %d", i); return WFL_ERR_ARG; }
    const float* kp = kn.data();
    const float* cp = cf.data();
    double umax = 0.0;
    if (precision == WFL_MIXED) {
      float rmax = 0.0f;
      rc = loops_rmax(g_ev.L, &rmax);
      if (rc) return rc;
      umax = node_u_bound(gg, (double)cx, (double)cy, (double)rtT, rmax);
    }
    build_spline_blob(1, &kp, &cp, &nspline, mode, sb, precision == WFL_MIXED ? &umax : nullptr, fermion == 1);
    if (sb.data.size() > g_ev.sblob_cap) {
      if (g_ev.d_sblob) (void)hipFree(g_ev.d_sblob);
      g_ev.d_sblob = nullptr; g_ev.sblob_cap = 0;
      HIP_TRY(hipMalloc(&g_ev.d_sblob, sizeof(float4) * sb.data.size()));
      g_ev.sblob_cap = sb.data.size();
    }
    HIP_TRY(hipMemcpy(g_ev.d_sblob, sb.data.data(), sizeof(float4) * sb.data.size(), hipMemcpyHostToDevice));
  }
  DevGeom dg;
  fill_geom(gg, g->kind == WFL_SPLINE ? &sb : nullptr, scaled, dg);
  DevTuple dt;
  fill_tuple(unit_scale(gg, scaled), dg, 0, (double)cx, (double)cy, (double)rtT, Nl, dt);
  if (scaled && g->kind == WFL_GAUSSIAN) {
    float rmax = 0.0f;
    rc = loops_rmax(g_ev.L, &rmax);
    if (rc) return rc;
    dg.acc_on = 1;
    dg.acc_rmax = rmax;
    dt.acc = tuple_needs_accurate(dg, dt, unit_scale(gg, scaled), rmax);
  }
  memcpy(g_ev.h_stage, &dg, sizeof dg);
  memcpy((char*)g_ev.h_stage + kEvTupOff, &dt, sizeof dt);
  HIP_TRY(hipMemcpyAsync(g_ev.d_block, g_ev.h_stage, kEvBlock, hipMemcpyHostToDevice, 0));
  WflLaunch a;
  memset(&a, 0, sizeof a);
  a.rec = precision == WFL_MIXED ? (const void*)g_ev.L->rec32 : (const void*)g_ev.L->rec64;
  a.N = N; a.Nl = Nl; a.tups = g_ev.d_tup; a.geoms = g_ev.d_geom; a.ntup = 1;
  if (g->kind == WFL_SPLINE) {
    a.sblob = g_ev.d_sblob; a.sblob_f4 = (int)sb.data.size();
    a.lds_splines = sb.tab[0].fast && sb.data.size() * sizeof(float4) <= kSplineLdsBytes;
    a.uni_splines = sb.all_uni && (size_t)sb.uni_f4 * sizeof(float4) <= kSplineLdsBytes;
    a.sblob_uni_f4 = sb.uni_f4;
    a.uni_origin = a.uni_splines && sb.uni_origin;
    if (a.uni_splines) a.lds_splines = 1;
  }
  a.prof_gauss = g->kind == WFL_GAUSSIAN; a.spinor = fermion == 1; a.mode = mode; a.nsub = nsub;
  a.ws_scal = (float4*)ws_scal_d; a.ws_ferm = fermion == 1 ? (float4*)ws_ferm_d : nullptr;
  a.nonfinite = g_ev.d_nf;
  a.wpb = 4;
  a.stream = 0;
  int e = precision == WFL_MIXED ? wfl_launch_loop_f32(&a) : wfl_launch_loop_f64(&a);
  if (e) { wfl_set_error("EV kernel launch failed: %s", hipGetErrorString((hipError_t)e)); return WFL_ERR_DEVICE; }
  return WFL_OK;
}
