// wfl_profiles.h -- device-side field profiles f(rho^2), f'(rho^2) and the
// segment integrand g(u) = f(u)/u for gfx950, templated on the arithmetic
// type (float for the mixed kernels, double for the FP64 kernels).
//
// Profiles follow intEV.cu:29-406 (flambda/fplambda and helpers) and the
// intEVff.cu:134-243 variant; parameters are the FP32-quantized values the
// host stores in DevGeom, so both precisions evaluate the same problem.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace wfl {

enum { P_STEP = 0, P_SMOOTH = 1, P_QUADRATIC = 2, P_GAUSSIAN = 3,
       P_PERIODIC = 4, P_SPLINE = 5, P_FIXFLUX = 6, P_FIXFLUX_FF = 7 };
enum { M_REFERENCE = 0, M_PHYSICAL = 1 };
// Kernel-only segment mode: physical mode for launches whose tuples all
// have x_cm on the x axis and no lattice images (FP32, Simpson n = 2), with
// the loop-only point records of the axis form (wfl_kernels.h).
enum { M_AXIS = 2 };
// kernel tag for "profile chosen at run time from DevGeom::kind"
enum { P_ANY = 99 };
// Kernel-only profiles: a spline whose knots lie on a uniform grid in
// r = sqrt(rho^2), evaluated in FP32 physical mode by direct indexing, with
// each row's cubic expanded about the row's start (P_SPLINE_U) or about
// u = 0 (P_SPLINE_U0, for tables where that form loses no accuracy).
enum { P_SPLINE_U = 98, P_SPLINE_U0 = 97 };
constexpr bool uniform_spline(int prof) { return prof == P_SPLINE_U || prof == P_SPLINE_U0; }
// Uniform-grid spline rows in LDS: the rows of every table from byte 0, at
// most kUniRows rows of 16 B.
constexpr int kUniRows = 1024;

// One geometry as the kernels see it.  Every double field holds a value
// that is exactly representable in FP32.
//
// Coordinates.  The FP64 kernels and the FP32 kernels of the generic
// profiles work in the problem's own units.  The FP32 Gaussian kernel works
// in scaled units z' = sigma z with sigma = sqrt(log2(e)/l2) rounded to FP32,
// so that a node needs only exp2(-u') and 1/u'; the host then defines the
// problem's effective parameters from the FP32 values (l2 = log2(e)/sigma^2,
// x_cm = c'/sigma, sqrt(T) = s'/sigma, lattice = L'/sigma), which is what the
// FP64 kernels and the oracle evaluate.
struct DevGeom {
  int kind, images;
  int spline_cf, spline_rec, spline_kn;             // spline blob offsets: coefs, buckets (float4), knots (float)
  int spline_n, spline_nb, spline_fast;             // knots, buckets, 1: every interval within 2 of its bucket start
  float spline_r0, spline_inv_w;                    // first bucket edge and 1/width in sqrt(rho^2)
  float spline_k0, spline_kl;                       // first and last knot
  int spline_ucf;                                   // uniform-grid rows (float4 offset, -1: none)
  float spline_uinv_h, spline_uh;                   // row of u: floor(sqrt(u) uinv_h); h = fl32 grid step
  double F, l2, inv_l2, a, lmin, q1, q2, lattice;   // effective values (FP64 kernels)
  float Ff, l2f, inv_l2f, af, lminf, q1f, q2f;      // FP32 generic-profile values
  float latticef;                                   // lattice spacing in FP32 kernel units
  float c8f, c10f;                                  // reference clamps 1e-8, 1e-10 in FP32 kernel units
  int acc_on;                                       // FP32 Gaussian: accurate node where loops reach a core
  float acc_rmax;                                   // largest loop radius of the ensemble (unit loops)
  float fix_c0f;                                    // fixflux: phi(0) of the reference's form, 1 - q2*/q2
};

// One evaluation tuple; the image-window centre (m0, n0) is computed on the
// host from the effective doubles.
struct DevTuple {
  int geom, m0, n0, nloops;
  float cx, cy, s;              // FP32 kernel units
  int acc;                      // 1: the loops can reach a tube core with the accurate node (groups lanes)
  double cxd, cyd, sd, padd;    // effective values (FP64 kernels)
  // Local Taylor node of the FP32 Gaussian axis form (loc = 1; see
  // wfl_kernels.h): g(uc) and 2^-uc in FP64 at uc = cx^2, and the FP32
  // coefficients a1 (split in two floats), a2 ... a7.
  double ta0, tec;
  float ta[8];
  int loc, pad_loc;
};

// ---------------------------------------------------------------- math
template <class R> struct Math;
template <> struct Math<float> {
  static __device__ __forceinline__ float exp2(float y) { return __builtin_amdgcn_exp2f(y); }
  static __device__ __forceinline__ float rcp(float x) { return __builtin_amdgcn_rcpf(x); }
  static __device__ __forceinline__ float exp(float x) { return __builtin_amdgcn_exp2f(x * 1.4426950408889634f); }
  static __device__ __forceinline__ float log(float x) { return __logf(x); }
  static __device__ __forceinline__ float sqrt(float x) { return __builtin_sqrtf(x); }
  static __device__ __forceinline__ float floor(float x) { return __builtin_floorf(x); }
  static __device__ __forceinline__ float sin(float x) { return __sinf(x); }
  static __device__ __forceinline__ float cos(float x) { return __cosf(x); }
  static __device__ __forceinline__ float cosh(float x) { return ::coshf(x); }
  static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
};
template <> struct Math<double> {
  static __device__ __forceinline__ double exp(double x) { return ::exp(x); }
  static __device__ __forceinline__ double log(double x) { return ::log(x); }
  static __device__ __forceinline__ double sqrt(double x) { return ::sqrt(x); }
  static __device__ __forceinline__ double floor(double x) { return ::floor(x); }
  static __device__ __forceinline__ double sin(double x) { return ::sin(x); }
  static __device__ __forceinline__ double cos(double x) { return ::cos(x); }
  static __device__ __forceinline__ double cosh(double x) { return ::cosh(x); }
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
};

template <class R> struct GP;   // geometry parameters in the kernel's type
template <> struct GP<float> {
  float F, l2, inv_l2, a, lmin, q1, q2, c0;
  __device__ GP(const DevGeom& g)
      : F(g.Ff), l2(g.l2f), inv_l2(g.inv_l2f), a(g.af), lmin(g.lminf), q1(g.q1f), q2(g.q2f), c0(g.fix_c0f) {}
};
template <> struct GP<double> {
  double F, l2, inv_l2, a, lmin, q1, q2;
  __device__ GP(const DevGeom& g) : F(g.F), l2(g.l2), inv_l2(g.inv_l2), a(g.a), lmin(g.lmin), q1(g.q1), q2(g.q2) {}
};

// -------------------------------------------- Gaussian tube (hot profile)
// FP32, scaled units (u' = log2(e) rho^2 / l2): returns g' = (1 - e)/u' and
// e = 2^-u' = exp(-rho^2/l2).  The caller adds 1e-30 to u'.
//
// ACC = 0: the plain quotient.  For small u' the difference 1 - e carries
// the absolute rounding of e, a relative error ~6e-8/u' in the quotient; a
// loop of scaled size s' lying in the tube core then gets a relative phase
// error that grows like 1/s'^2.
// ACC = 1, 2: the argument v~ = -log2(e) is the one whose power of two e is
// exactly, and 1 - e is exact (Sterbenz), so (1 - e)/v~ is g'(v~), which is
// off from g'(u') by O(|v~ - u'|) = O(1e-7) absolutely: ~1e-7 relative at
// every u' (the device of Kahan's expm1 from exp and log).  u' is held at or
// above 1e-6 in the exponential (g' there is ln 2 to 3.5e-7; e for such a
// node, within 1e-3 of the axis, is off by < 7e-7).  ACC = 1 needs u' < 100
// (e stays a normal number), which the caller establishes for the whole
// wavefront; ACC = 2 also clamps u' at 100 in the exponential and adds
// u' - 100 back to the denominator (g' = 1/u' there to 2^-100).
template <int ACC>
__device__ __forceinline__ float gauss_g(float u, double, float& e) {
  if constexpr (ACC == 0) {
    e = __builtin_amdgcn_exp2f(-u);
    const float r = __builtin_amdgcn_rcpf(u);
    return fmaf(-e, r, r);          // (1 - e)/u in one FMA
  } else if constexpr (ACC == 1) {
    e = __builtin_amdgcn_exp2f(-fmaxf(u, 1e-6f));
    const float vt = -__builtin_amdgcn_logf(e);
    return (1.0f - e) * __builtin_amdgcn_rcpf(vt);
  } else {
    const float v = __builtin_amdgcn_fmed3f(u, 1e-6f, 100.0f);
    e = __builtin_amdgcn_exp2f(-v);
    const float d = fmaxf(u - v, 0.0f) - __builtin_amdgcn_logf(e);
    return (1.0f - e) * __builtin_amdgcn_rcpf(d);
  }
}
// FP64, problem units: g(u) = -expm1(-u/l2)/u (limit 1/l2 at u = 0),
// e = 1 + expm1(-u/l2).
template <int ACC>
__device__ __forceinline__ double gauss_g(double u, double inv_l2, double& e) {
  double x = u * inv_l2;
  double em1 = ::expm1(-x);
  e = 1.0 + em1;
  return (u > 0.0) ? -em1 / u : inv_l2;
}

// ------------------------------------------------ remaining profiles
template <class R>
__device__ R expint_neg(R x) {   // Ei(x) = -E1(-x), x < -1 (intEV.cu:29-61), converged to R
  const R EPS = sizeof(R) == 4 ? (R)2.0e-7 : (R)1.0e-15;
  R b = -x + (R)1, c = (R)1e30, d = Math<R>::div((R)1, b), h = d;
  for (int i = 1; i <= 1000; ++i) {
    R a = -(R)i * (R)i;
    b += (R)2;
    d = Math<R>::div((R)1, a * d + b);
    c = b + Math<R>::div(a, c);
    R del = c * d;
    h *= del;
    if (fabs(del - (R)1) <= EPS) return -h * Math<R>::exp(x);
  }
  return (R)0;
}

template <class R> __device__ R bump(R x) {
  return (x * x < (R)0.999) ? Math<R>::exp(Math<R>::div((R)-1, (R)1 - x * x)) : (R)0;
}

// The fixflux core profile, phi(x) = 1 - (om e^{-1/om} + Ei(-1/om)) / (2 q2)
// with om = 1 - x^2, is phi(0) plus (1/(2 q2)) times the integral of
// e^{-1/t} over [1 - x^2, 1] (d/dom of om e^{-1/om} + Ei(-1/om) is e^{-1/om}).
// phi(0) = 1 - q2*/q2 with q2* = (e^{-1} - E1(1))/2, the exact value of the
// constant: zero up to the rounding of the q2 the profile is given.  Near
// the axis the first form is 1 minus nearly 1; the FP32 kernels use the
// integral there, by 6-point Gauss-Legendre (error below 1e-9 for
// x^2 <= 1/2: the integrand's singularity at t = 0 lies 3 half-widths from
// the centre of the interval).
__device__ __forceinline__ float phi_fix_core_f32(float q2, float c0, float x) {
  const float gx[3] = {0.2386191861f, 0.6612093865f, 0.9324695142f};
  const float gw[3] = {0.4679139346f, 0.3607615730f, 0.1713244924f};
  const float h = 0.5f * x * x, c = 1.0f - h;   // [1 - x^2, 1] = c -+ h
  float sum = 0.0f;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const float t1 = fmaf(-h, gx[k], c), t2 = fmaf(h, gx[k], c);
    sum = fmaf(gw[k], Math<float>::exp(-1.0f / t1) + Math<float>::exp(-1.0f / t2), sum);
  }
  return fmaf(sum, h / (2.0f * q2), c0);
}

template <class R> __device__ R phi_fix(const GP<R>& g, R x, bool ff) {
  R om = (R)1 - x * x;
  if constexpr (sizeof(R) == 4)
    if (x * x <= 0.5f) return phi_fix_core_f32(g.q2, g.c0, x);
  if (x < (ff ? (R)1 : (R)0.999))
    return (R)1 - Math<R>::div((R)0.5, g.q2) * (om * Math<R>::exp(Math<R>::div((R)-1, om)) + expint_neg<R>(Math<R>::div((R)-1, om)));
  return (R)1;
}

template <class R> __device__ R chi_fix(const GP<R>& g, R x, R n, R lam, bool ff) {
  R om = (R)1 - x * x, x2 = x * x, x4 = x2 * x2;
  R thr = ff ? (R)1 : (R)0.999;
  if (x <= -thr) return (R)0;
  if (x2 < thr) {
    R ans = (R)0.5 * (-om * Math<R>::exp(Math<R>::div((R)-1, om)) - expint_neg<R>(Math<R>::div((R)-1, om)));
    if (!ff) {
      R arg = Math<R>::div((R)3.31 * x + (R)5.25 * x * x2 * Math<R>::sin(x) * Math<R>::cos((R)-0.907 * x2 - (R)1.29 * x4 * x4), Math<R>::cos(x));
      ans += Math<R>::div((R)2 * n * g.a, lam) * Math<R>::div((R)0.444, (R)1 + Math<R>::exp(-arg));
    } else {
      ans += Math<R>::div((R)2 * n * g.a, lam) *
             ((R)0.218 + Math<R>::div((R)0.393 * x * Math<R>::cosh((R)0.806 * x4 - (R)0.696 * x2 + (R)0.0902),
                                       Math<R>::cosh((R)0.825 * x2 - (R)0.0234 * x + (R)0.375)));
    }
    return ans;
  }
  return Math<R>::div((R)2 * n * g.a * g.q1, lam);
}

// With the i = 0 ring in the window, the reference's sum
//   a/(lam ln2) * [-(rho/a) e/(1+e) - (lam/a) log(1+e)] + 1,  e = exp(-t), t = rho/lam,
// is 1 minus nearly 1 near the axis (f ~ t^2/(8 ln2)).  It equals
// [(t/2) tanh(t/2) - log cosh(t/2)] / ln2 (both vanish at t = 0 and have the
// derivative (t/4) sech^2(t/2) / ln2), whose two terms differ by a factor
// of about 2 for t < 2; the FP32 kernels use that form there and the sum
// itself (1 minus at most 0.53) beyond.
__device__ __forceinline__ float peri_core_f32(float t) {
  if (t >= 2.0f) {
    const float e = Math<float>::exp(-t);
    return 1.0f - (t * e / (1.0f + e) + log1pf(e)) * 1.44269504f;
  }
  // with m = e^y - 1: tanh y = m(m + 2)/(m(m + 2) + 2), cosh y - 1 = m^2/(2(1 + m))
  const float y = 0.5f * t, m = expm1f(y), mm = m * (m + 2.0f);
  const float lc = log1pf(m * m / (2.0f * (1.0f + m)));
  return (y * (mm / (mm + 2.0f)) - lc) * 1.44269504f;
}

template <class R> __device__ R flperi(const GP<R>& g, R rho2) {   // intEV.cu:202-232
  R rho = Math<R>::sqrt(rho2), lam = Math<R>::sqrt(g.l2);
  R flraf = Math<R>::floor(Math<R>::div(rho, g.a)), f = (R)0;
  int fl = (int)flraf;
  if constexpr (sizeof(R) == 4) {
    if (fl <= 2) {              // the i = 0 ring is in the window: stable core form plus the rings i > 0
      for (int i = 1; i <= fl + 2; ++i)
        f += Math<R>::div((R)i, (R)1 + Math<R>::exp(-Math<R>::div(rho - (R)i * g.a, lam)));
      return Math<R>::div(g.a, lam * (R)0.69314718055994530942) * f + peri_core_f32(Math<R>::div(rho, lam));
    }
  }
  if (fl - 2 > 0) f = (R)0.5 * (flraf - (R)3) * (flraf - (R)2);
  for (int i = fl - 2; i <= fl + 2; ++i) {
    if (i == 0) {
      R ex = Math<R>::exp(-Math<R>::div(rho, lam)), ep1 = (R)1 + ex;
      f = -Math<R>::div(rho, g.a) * Math<R>::div(ex, ep1) - Math<R>::div(lam, g.a) * Math<R>::log(ep1);
    } else if (i > 0) {
      f += Math<R>::div((R)i, (R)1 + Math<R>::exp(-Math<R>::div(rho - (R)i * g.a, lam)));
    }
  }
  return Math<R>::div(g.a, lam * (R)0.69314718055994530942) * f + (R)1;
}

template <class R> __device__ R fplperi(const GP<R>& g, R rho2) {  // intEV.cu:301-328
  R rho = Math<R>::sqrt(rho2), lam = Math<R>::sqrt(g.l2);
  R flraf = Math<R>::floor(Math<R>::div(rho, g.a)), f = (R)0;
  int fl = (int)flraf;
  for (int i = fl - 2; i <= fl + 2; ++i) {
    if (i == 0) {
      R ex = Math<R>::exp(-Math<R>::div(rho, lam)), ep1 = (R)1 + ex;
      f = Math<R>::div(ex, g.a * ep1 * ep1);
    } else if (i > 0) {
      R ea = Math<R>::exp(-Math<R>::div(rho - (R)i * g.a, lam)), eap1 = (R)1 + ea;
      f += Math<R>::div((R)i, rho) * Math<R>::div(ea, eap1 * eap1);
    }
  }
  return Math<R>::div(g.a, (R)2 * g.l2 * (R)0.69314718055994530942) * f;
}

template <class R> __device__ R ffixflux(const GP<R>& g, R rho2, bool ff) {  // intEV.cu:237-257
  R lam = Math<R>::sqrt(g.l2);
  R lml = Math<R>::div(lam - g.lmin, g.a - g.lmin), aml = Math<R>::div(g.a - lam, g.a - g.lmin);
  R n = Math<R>::floor(Math<R>::div(Math<R>::sqrt(rho2) + g.a * (R)0.5, g.a));
  if (rho2 <= g.a * g.a * (R)0.25)
    return ((R)1 - (R)0.75 * lml) * phi_fix<R>(g, (R)2 * Math<R>::sqrt(Math<R>::div(rho2, g.l2)), ff) +
           Math<R>::div((R)3 * rho2, g.a * g.a) * lml;
  R ans = (R)1 + (R)0.75 * (Math<R>::div((R)4 * rho2, g.a * g.a) - (R)1) * lml + (R)3 * n * (n - (R)1) * aml;
  ans += Math<R>::div((R)3 * lam, g.q1 * g.a) * aml *
         chi_fix<R>(g, Math<R>::div((R)2 * (Math<R>::sqrt(rho2) - n * g.a), lam), n, lam, ff);
  return ans;
}

template <class R> __device__ R fpfixflux(const GP<R>& g, R rho2) {   // intEV.cu:334-357
  R lam = Math<R>::sqrt(g.l2), rho = Math<R>::sqrt(rho2);
  R lml = Math<R>::div(lam - g.lmin, g.a - g.lmin), aml = Math<R>::div(g.a - lam, g.a - g.lmin);
  R n = Math<R>::floor(Math<R>::div(rho + g.a * (R)0.5, g.a));
  if (rho <= g.a * (R)0.5)
    return Math<R>::div((R)2, g.l2 * g.q2) * ((R)1 - (R)0.75 * lml) * bump<R>(Math<R>::div((R)2 * rho, lam)) +
           Math<R>::div((R)3, g.a * g.a) * lml;
  return Math<R>::div((R)3, g.a * g.a) * lml +
         Math<R>::div((R)6, g.q1 * lam * g.a) * aml * bump<R>(Math<R>::div((R)2 * (rho - n * g.a), lam));
}

// Spline tables.  A plan stores all of its tables in one blob of float4:
//   [coefs x (n + 1) per table][bucket records x btot][knots (float) x ctot]
// read from global memory or from a copy in LDS.  Coefficient row 0 is the
// part at or below the first knot and row n the part at or beyond the last
// (for the segment mode of the launch: zero in reference mode, intEV.cu:96-99;
// in physical mode the line s u, s = fl32(c0/k0), from u = 0, and the
// constant f(k_{n-1}) from k_{n-1}); row j + 1 is interval j.  The interval
// of k0 < u < k_{n-1} is the largest j <= n - 2 with knot[j] <= u
// (intEV.cu:81-89, a bisection there).  Here nb equal buckets over
// r = sqrt(u) each hold a record (s, knot[s], knot[s+1], knot[s+2]) with s at
// or below the interval of any u that can land in the bucket (entries past
// n - 2 are +inf).  When the host has checked that the interval is always one
// of s, s+1, s+2 ("fast"), two comparisons give it with no further reads;
// otherwise a scan from s ends on it.  Either way the result is the
// bisection's, whatever the rounding of r, and nothing branches on u in the
// fast case.
struct SplineRef {
  const float4* cf;       // coefficient rows (c0, c1, c2, c3)
  const float4* rec;      // bucket records
  const float* kn;        // knots rho^2
  int n, nb, fast;
  float r0, inv_w, k0, kl;
  const float4* ucf;      // uniform-grid rows
  float uinv_h, uh, uhoff;  // 1/h, h and -2^23 h (exact)
  unsigned ubias;         // LDS byte address of ucf less 0x4B000000 * 16 (mod 2^32)
};

__device__ __forceinline__ SplineRef spline_ref(const float4* blob, const DevGeom& g) {
  SplineRef s;
  s.cf = blob + g.spline_cf;
  s.rec = blob + g.spline_rec;
  s.kn = (const float*)blob + g.spline_kn;
  s.n = g.spline_n;
  s.nb = g.spline_nb;
  s.fast = g.spline_fast;
  s.r0 = g.spline_r0;
  s.inv_w = g.spline_inv_w;
  s.k0 = g.spline_k0;
  s.kl = g.spline_kl;
  s.ucf = blob + (g.spline_ucf > 0 ? g.spline_ucf : 0);
  s.uinv_h = g.spline_uinv_h;
  s.uh = g.spline_uh;
  s.uhoff = -8388608.0f * g.spline_uh;
  // only the uniform-grid kernels read it, and they stage the rows in LDS;
  // the empty asm keeps the bias out of the compiler's reach, so that each
  // node adds the shifted row index to it in one instruction
  s.ubias = (unsigned)(uintptr_t)(const __attribute__((address_space(3))) float4*)s.ucf + 0x50000000u;
  asm("" : "+v"(s.ubias));
  return s;
}

// Uniform-grid spline, FP32 physical mode: g(u) = f(u)/u and f'(u) from one
// row in LDS, with no search and no clamp.  The knots lie on
// r_j = (m + j) h with an integer m >= 0 (checked on the host to 1e-4 h), so
// row i = floor(sqrt(u) uinv_h) is interval i - m; rows 0..m-1 are the line
// s u below the first knot, and the rows from n - 1 + m on the constant past
// the last knot, as many as the largest u that the plan's tuples can reach
// needs (the host bounds it from the loops' largest radius), so every index
// stays in the table.
// The floor comes from the rounding of one FMA: y + (2^23 - 1/2) with
// y = sqrt(u) uinv_h >= 1/4 lies in [2^23, 2^24), where FP32 steps by 1, so
// it rounds to t = 2^23 + floor(y) (round to nearest of y - 1/2; a tie, y on
// a knot, may give the row below, which meets the row above there), and the
// row's byte offset is t's bit pattern times 16 less 0x4B000000 * 16;
// y < 1/4 rounds to t = 2^23 - 1/2, the bit pattern one below, which reads
// the guard row -1 (row 0's polynomial).
// Row i holds the cubic (c0, c1, c2, c3) of the exact lookup's row
// re-expanded on the host about kappa_i = r_i^2, r_i = fl32(i h) (i = -1/2
// for the guard row): the kernel forms r_i = fma(t, h, -2^23 h) (the product
// and the difference exact, one rounding) while the row is read and
// dx = u - r_i^2 in one FMA, so no knot is read.  Near a knot the FP32
// sqrt(u) may pick the neighbouring interval: interior intervals join with
// two continuous derivatives and the line with a continuous value, so the
// difference is at the FP32 rounding level.  u > 0 always holds here (the
// FP32 kernels add 1e-30).
// ORIGIN: the rows hold their cubics P(u) = d0 + u (d1 + u (d2 + u d3))
// about u = 0 instead, and the node forms g = d0 (1/u) + (d1 + u (d2 + u d3))
// directly: no expansion point and no final product, three instructions
// fewer per node.  The host takes this form only for tables where it is
// well conditioned: over every row, sum_k |d_k| u^k <= 16 |P(u)| and
// sum_k k |d_k| u^(k-1) <= 16 max |P'|; it rounds the d_k to FP32 so that
// each rounding error is carried by the lower coefficients about the row's
// centre, and spreads d0's rounding over consecutive rows with alternating
// signs.
template <bool ORIGIN>
__device__ __forceinline__ float spline_u_g(float u, const SplineRef& sp, bool need_fp, float& fp) {
  typedef const __attribute__((address_space(3))) float4 LdsRow;
  const float t = fmaf(__builtin_amdgcn_sqrtf(u), sp.uinv_h, 8388607.5f);
  const float4* row = (const float4*)(LdsRow*)(uintptr_t)((__float_as_uint(t) << 4) + sp.ubias);
  const float4 c = row[0];
  // f' = c1 + 2 c2 x + 3 c3 x^2 = q + x (i + c3 x) from the Horner
  // partials i = c2 + c3 x and q = c1 + x i: two FMAs
  if constexpr (ORIGIN) {
    const float i = fmaf(u, c.w, c.z), q = fmaf(u, i, c.y);
    fp = need_fp ? fmaf(u, fmaf(u, c.w, i), q) : 0.0f;
    return fmaf(c.x, __builtin_amdgcn_rcpf(u), q);
  }
  const float rk = fmaf(t, sp.uh, sp.uhoff);
  const float dx = fmaf(-rk, rk, u);
  const float i = fmaf(dx, c.w, c.z), q = fmaf(dx, i, c.y);
  const float f = fmaf(dx, q, c.x);
  fp = need_fp ? fmaf(dx, fmaf(dx, c.w, i), q) : 0.0f;
  return f * __builtin_amdgcn_rcpf(u);
}

// Interval j (0 <= j <= n - 2) of u and its left knot; exact for
// k0 < u < k_{n-1}, an in-bounds index for any u.  FAST: every table of the
// launch is fast (no scan code at all).
template <class R, bool FAST>
__device__ __forceinline__ int spline_find(R u, const SplineRef& sp, R& klo) {
  const float r = __builtin_amdgcn_sqrtf((float)u);
  int b = (int)((r - sp.r0) * sp.inv_w);
  b = b < 0 ? 0 : (b >= sp.nb ? sp.nb - 1 : b);
  const float4 q = sp.rec[b];
  int j = __float_as_int(q.x);
  if (FAST || sp.fast) {
    const bool a1 = u >= (R)q.z, a2 = u >= (R)q.w;
    klo = a2 ? (R)q.w : (a1 ? (R)q.z : (R)q.y);
    return j + (int)a1 + (int)a2;
  }
  klo = (R)q.y;
  while (j > 0 && u < klo) { --j; klo = (R)sp.kn[j]; }
  R khi = (R)sp.kn[j + 1];
  while (j + 2 < sp.n && khi <= u) { ++j; klo = khi; khi = (R)sp.kn[j + 1]; }
  return j;
}

// f(u) and, if need_fp, f'(u) from one lookup and one coefficient row: the
// cubic of the row and its derivative (the outer rows are polynomials of the
// same form).  Reference mode: f' from the analytic sin^2 form
// (intEV.cu:389-398).
template <class R, bool FAST = false>
__device__ __forceinline__ void spline_eval(R u, const SplineRef& sp, int mode, bool need_fp, const GP<R>& g,
                                            R& f, R& fp) {
  R klo;
  const int j = spline_find<R, FAST>(u, sp, klo);
  const bool lo = !(u > (R)sp.k0), hi = !(u < (R)sp.kl);
  const int row = lo ? 0 : (hi ? sp.n : j + 1);
  klo = lo ? (R)0 : (hi ? (R)sp.kl : klo);
  const float4 c = sp.cf[row];
  const R dx = u - klo;
  const R in = (R)c.z + dx * (R)c.w, q = (R)c.y + dx * in;
  f = (R)c.x + dx * q;
  fp = (R)0;
  if (need_fp) {
    if (mode == M_REFERENCE) {
      R s = Math<R>::sin(Math<R>::sqrt(u) * (R)3.14159265358979323846);
      fp = (s > (R)100) ? (R)0 : Math<R>::exp(-s * s * g.inv_l2) * g.inv_l2;
    } else {
      fp = q + dx * (in + dx * (R)c.w);   // c1 + 2 c2 dx + 3 c3 dx^2
    }
  }
}

// Generic profile evaluation: f(u) and f'(u).
template <class R>
__device__ void profile_eval(int kind, const GP<R>& g, R u, int mode, bool need_fp, const SplineRef& sp,
                             R& f, R& fp) {
  switch (kind) {
    case P_STEP:
      f = (u < g.l2) ? u * g.inv_l2 : (R)1;
      fp = (u < g.l2) ? g.inv_l2 : (R)0;
      break;
    case P_SMOOTH: {
      R den = g.l2 + u;
      f = Math<R>::div(u, den);
      fp = need_fp ? Math<R>::div(g.l2, den * den) : (R)0;
    } break;
    case P_QUADRATIC:
      f = (u < g.l2) ? u * g.inv_l2 * ((R)2 - u * g.inv_l2) : (R)1;
      fp = (u < g.l2) ? (R)2 * g.inv_l2 * ((R)1 - u * g.inv_l2) : (R)0;
      break;
    case P_GAUSSIAN: {
      // FP32: u is in the scaled units of the Gaussian kernels, u' = log2(e) rho^2/l2
      R e;
      if constexpr (sizeof(R) == 4) e = __builtin_amdgcn_exp2f(-u);
      else e = Math<R>::exp(-u * g.inv_l2);
      f = (R)1 - e;
      fp = e * g.inv_l2;
    } break;
    case P_PERIODIC:
      f = flperi<R>(g, u);
      fp = need_fp ? fplperi<R>(g, u) : (R)0;
      break;
    case P_SPLINE:
      spline_eval<R>(u, sp, mode, need_fp, g, f, fp);
      break;
    case P_FIXFLUX:
    case P_FIXFLUX_FF:
      f = ffixflux<R>(g, u, kind == P_FIXFLUX_FF);
      fp = need_fp ? fpfixflux<R>(g, u) : (R)0;
      break;
    default:
      f = (R)0; fp = (R)0;
  }
}

// g(u) = f(u)/u for the generic profiles.  Reference mode: plain quotient
// (the caller has clamped u).  Physical mode: analytic limits at u -> 0.
// PROF is the profile when it is fixed at compile time, P_ANY to take it
// from kind at run time.
template <class R, int PROF>
__device__ __forceinline__ R generic_g(int kind_rt, const GP<R>& g, R u, int mode, bool need_fp,
                                       const SplineRef& sp, R& fp) {
  if constexpr (uniform_spline(PROF) && sizeof(R) == 4) return spline_u_g<PROF == P_SPLINE_U0>(u, sp, need_fp, fp);
  const int kind = PROF == P_ANY ? kind_rt : PROF;
  R f;
  if (mode == M_PHYSICAL) {
    switch (kind) {
      case P_STEP:
        fp = (u < g.l2) ? g.inv_l2 : (R)0;
        return (u < g.l2) ? g.inv_l2 : Math<R>::div((R)1, u);
      case P_SMOOTH: {
        R den = g.l2 + u;
        fp = need_fp ? Math<R>::div(g.l2, den * den) : (R)0;
        return Math<R>::div((R)1, den);
      }
      case P_QUADRATIC:
        fp = (u < g.l2) ? (R)2 * g.inv_l2 * ((R)1 - u * g.inv_l2) : (R)0;
        return (u < g.l2) ? ((R)2 - u * g.inv_l2) * g.inv_l2 : Math<R>::div((R)1, u);
      case P_GAUSSIAN:
        if constexpr (sizeof(R) == 4) {   // scaled units, as in the Gaussian kernels
          R e;
          const R gv = gauss_g<2>(u, 0.0, e);
          fp = e * g.inv_l2;
          return gv;
        }
        break;
      default:
        break;
    }
    const bool pos = u > (R)0;
    if (PROF == P_SPLINE) spline_eval<R, true>(u, sp, mode, need_fp || !pos, g, f, fp);
    else profile_eval<R>(kind, g, u, mode, need_fp || !pos, sp, f, fp);
    if (pos) return Math<R>::div(f, u);
    return fp;   // f(0) = 0: the limit of f(u)/u is f'(0)
  }
  if (PROF == P_SPLINE) spline_eval<R, true>(u, sp, mode, need_fp, g, f, fp);
  else profile_eval<R>(kind, g, u, mode, need_fp, sp, f, fp);
  return Math<R>::div(f, u);
}

}  // namespace wfl
