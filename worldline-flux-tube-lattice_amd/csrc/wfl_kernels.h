// wfl_kernels.h -- the Wilson-loop kernels for gfx950 (templated on the
// arithmetic type; instantiated in wfl_kernels_f32.hip / wfl_kernels_f64.hip).
//
// Data: each loop point i carries a record of six values
//   (w_x, w_y, 4 dw_x, 4 dw_y, e_x, e_y),  dw = w_succ(i) - w_i,  e = w_succ(i) - w_pred(i)
// where succ/pred follow the sub-loop map of getzp1 (intEV.cu:462-506).
// Records are stored point-major inside each loop ([Nl][N][REC], REC = 6 values, 24 B in FP32) so that one
// wavefront walking one loop reads them with scalar loads.
//
// Two mappings of work to the machine:
//  * k_tuple (lane per tuple): the 64 lanes of a wavefront evaluate 64
//    (geometry, x_cm, T) tuples on the SAME loop point at a time; the point
//    record is wave-uniform (SGPRs), nothing is shared between lanes, and
//    each lane reduces its own tuple's group sums in FP64 registers.
//  * k_loop (wave per loop): one wavefront evaluates one (tuple, loop) pair,
//    lanes stride over points; sub-loop sums are combined with a 4-step
//    cross-lane butterfly.  Used for EV() and for batches with few tuples.
//
// Per point i (sub-loop class c: odd -> x, 2 mod 4 -> y, 0 mod 4 -> z) with
// z = c_img + s w_i, s = sqrt(T):
//   PHYSICAL:  P_c += (z x e) g(|z|^2) + (z x dw) sum_k w_k g(|z + t_k s dw|^2)
//   REFERENCE: P_c += (z x dw) [g(A) + g(A1) + sum_k w_k g(rho^2(t_k))]  (clamped A, B, C)
// with Simpson weights w_k = 4 (odd k), 2 (even k), t_k = k/n, and
// Phi_c = F s (h/3) P_c, h = 1/n.  The spin factor uses f' the same way.
#pragma once
#include "wfl_profiles.h"

namespace wfl {

// One wavefront's work: a lane group (64 tuples of one geometry) on loops
// [loop0, loop0 + nloop), summed into output slot `slot` (the loop group, or
// one of its chunks when groups are split into several tasks).
struct Task { int lanegroup, loop0, nloop, slot; };

// values per point record (w, 4 dw, e: 24 B in FP32)
constexpr int REC = 6;

__device__ __forceinline__ int succ_idx(int i, int N) {
  if (i & 1) return (i == N - 1) ? 1 : i + 2;
  if ((i & 3) == 0) return (i == N - 4) ? 0 : i + 4;
  return (i == N - 2) ? 2 : i + 4;
}
__device__ __forceinline__ int pred_idx(int i, int N) {
  if (i & 1) return (i == 1) ? N - 1 : i - 2;
  if ((i & 3) == 0) return (i == 0) ? N - 4 : i - 4;
  return (i == 2) ? N - 2 : i - 4;
}
// accumulator slot of point i: 0 = x (odd), 1 = y (2 mod 4), 2 = z (0 mod 4)
__device__ __forceinline__ int cls_of(int i) { return (i & 1) ? 0 : ((i & 3) == 0 ? 2 : 1); }

// Axis-form records (M_AXIS), 8 floats per point, loop-only:
//   (w_x, w_y^2, 4 dw_x, m_y^2, 4 dw_y, e_y, w_y 4 dw_x, w_y e_x),  m = w + dw/2.
// With x_cm = (c, 0) only the x coordinate mixes c and s w, so a tuple
// (s = sqrt(T)) forms z_x = c + s w_x and the midpoint's x as in point_eval
// (the same roundings) and takes the rest from the record:
//   rho^2 = z_x^2 + s^2 w_y^2                       (midpoint: m_x, m_y)
//   z x v = z_x v_y - s (w_y v_x)                   (v = 4 dw, e)
// 11 VALU per point besides the nodes (point_eval: 14).
constexpr int RECA = 8;

// Per-loop constants of the local Taylor node (axis form): the class sums
// Q_c = sum_{i in c} w_i x (e_i + 4 dw_i) in FP64 and the loop's largest
// radius max |w_i|.
struct LoopInfo { double q[3]; double r; };

// Local Taylor node (FP32 Gaussian axis form, scaled units).  For a loop
// whose nodes all lie within |u - uc| <= kLocalDmax of uc = c'^2 (c' the
// tube-relative x_cm), g(u) = (1 - 2^-u)/u is written as
//   g(u) = a0 + h(d),  h(d) = sum_{k=1..7} a_k d^k,  d = u - uc,
// with a_k = g^(k)(uc)/k!, and the spin node as 2^-u = 2^-uc 2^-d.  Since
// the cross products of a sub-loop's segments with a fixed point sum to
// zero, sum_i (z x v_i) a0 = a0 s' sum_i w_i x v_i = a0 s' Q_c, which is
// added in FP64 once per loop; the FP32 sums carry only (z x v) h, whose
// terms have no cancellation.  The phase then needs no transcendental:
// the gfx950 exp2/log2/rcp results carry small biases that vary smoothly
// with the argument, and in sum_i (c' x v_i) g(u_i) (terms ~ c'/(s' |w|)
// times the result) such a bias does not average out.  Truncation of the
// series at k = 7 is below 4e-9 relative in h for |d| <= 0.5 at any uc
// >= 0 (|g^(k)| <= ln2^(k+1)/(k+1)).
constexpr float kLocalDmax = 0.5f;
struct LocalC {
  float a1h, a1l, a2, a3, a4, a5, a6, a7;   // a1 = a1h + a1l
};

__device__ __forceinline__ float local_h(float d, const LocalC& t) {
  float D = fmaf(t.a7, d, t.a6);
  D = fmaf(D, d, t.a5);
  D = fmaf(D, d, t.a4);
  D = fmaf(D, d, t.a3);
  D = fmaf(D, d, t.a2);
  const float lo = fmaf(D, d, t.a1l);
  return fmaf(d, t.a1h, d * lo);
}

// Per-lane constants of one tuple in the kernel's type and units.
template <class R> struct LaneC {
  R s;          // sqrt(T)
  R m8;         // s/8: midpoint offset per unit of the stored 4 dw
  R s4;         // s/4: segment vector per unit of the stored 4 dw
  R inv_l2;     // 1/l2 (FP64 Gaussian node)
  R c8, c10;    // the reference clamps 1e-8, 1e-10 in kernel units
};

// --------------------------------------------------------- node evaluation
template <class R, int PROF, int ACC>
__device__ __forceinline__ R node_g(R u, const LaneC<R>& lc, int kind, const GP<R>& gp, int mode,
                                    const SplineRef& sp, bool spinor, R& fp) {
  if (PROF == P_GAUSSIAN) {
    R e;
    R gval = gauss_g<ACC>(u, (double)lc.inv_l2, e);
    fp = e;
    return gval;
  } else {
    return generic_g<R, PROF>(kind, gp, u, mode, spinor, sp, fp);
  }
}

// One point (= one segment start) for one image centre.  The record holds
// 4 dw (an exact power-of-two scaling), so X = 4 z x dw carries the Simpson
// midpoint weight.  Adds to P (phase) and S (half the spin sum) of the
// point's class, before the common factors F s h/3 (phase) and
// 2 h/3 spin_scale (spin), h = 1/n.
template <class R, int PROF, int SPINOR, int MODE, int NSUB, int ACC>
__device__ __forceinline__ void point_eval(R wx, R wy, R dx4, R dy4, R ex, R ey,
                                           R ccx, R ccy, const LaneC<R>& lc, int kind, const GP<R>& gp,
                                           const SplineRef& sp, int nsub_rt, R& P, R& S) {
  const int n = NSUB ? NSUB : nsub_rt;
  const R tiny = sizeof(R) == 4 ? (R)1e-30 : (R)0;   // keeps 1/u finite on the axis (FP32 node)
  R zx = fma(lc.s, wx, ccx);
  R zy = fma(lc.s, wy, ccy);
  R X = fma(zx, dy4, -(zy * dx4));        // 4 z x dw
  if (MODE == M_PHYSICAL) {
    R E = fma(zx, ey, -(zy * ex));        // z x e: both endpoint weights of this point
    R A = fma(zx, zx, fma(zy, zy, tiny));
    R fA, fk;
    R gA = node_g<R, PROF, ACC>(A, lc, kind, gp, MODE, sp, SPINOR, fA);
    R gin, fin;
    if (n == 2) {
      R mx = fma(lc.m8, dx4, zx), my = fma(lc.m8, dy4, zy);
      R Am = fma(mx, mx, fma(my, my, tiny));
      gin = node_g<R, PROF, ACC>(Am, lc, kind, gp, MODE, sp, SPINOR, fk);
      fin = fk + fk;
    } else {
      const R h4 = lc.s4 / (R)n;          // node k sits at z + (k h s/4)(4 dw)
      gin = (R)0; fin = (R)0;
      for (int k = 1; k < n; ++k) {
        R t = (R)k * h4;
        R px = fma(t, dx4, zx), py = fma(t, dy4, zy);
        R u = fma(px, px, fma(py, py, tiny));
        R gk = node_g<R, PROF, ACC>(u, lc, kind, gp, MODE, sp, SPINOR, fk);
        gin = (k & 1) ? gin + gk : fma((R)0.5, gk, gin);      // Simpson weights 4, 2 over 4
        if (SPINOR) fin = (k & 1) ? fma((R)2, fk, fin) : fin + fk;   // and over 2
      }
    }
    P = fma(E, gA, fma(X, gin, P));
    if (SPINOR) S += fA + fin;
  } else {
    // Reference Idt: rho^2(t) = A + 2Bt + Ct^2 with the clamps of intEV.cu:418-422, 439.
    R dsx = lc.s4 * dx4, dsy = lc.s4 * dy4;   // z' - z
    R A = fma(zx, zx, zy * zy);
    R B = fma(zx, dsx, zy * dsy);
    R C = fma(dsx, dsx, dsy * dsy);
    A = A < lc.c8 ? lc.c8 : A;
    B = B < lc.c8 ? lc.c8 : B;
    C = C < lc.c8 ? lc.c8 : C;
    R A1 = A + (R)2 * B + C;
    A1 = A1 < lc.c8 ? lc.c8 : A1;
    R f0, f1, fk;
    R g0 = node_g<R, PROF, ACC>(A, lc, kind, gp, MODE, sp, SPINOR, f0);
    R g1 = node_g<R, PROF, ACC>(A1, lc, kind, gp, MODE, sp, SPINOR, f1);
    R gs = (R)0.25 * (g0 + g1), fs = SPINOR ? (R)0.5 * (f0 + f1) : (R)0;
    if (n == 2) {
      R u = A + B + (R)0.25 * C;
      u = u < lc.c10 ? lc.c10 : u;
      gs += node_g<R, PROF, ACC>(u, lc, kind, gp, MODE, sp, SPINOR, fk);
      if (SPINOR) fs = fma((R)2, fk, fs);
    } else {
      const R h = (R)1 / (R)n;
      for (int k = 1; k < n; ++k) {
        R t = (R)k * h;
        R u = A + (R)2 * B * t + C * t * t;
        u = u < lc.c10 ? lc.c10 : u;
        R gk = node_g<R, PROF, ACC>(u, lc, kind, gp, MODE, sp, SPINOR, fk);
        gs = (k & 1) ? gs + gk : fma((R)0.5, gk, gs);
        if (SPINOR) fs = (k & 1) ? fma((R)2, fk, fs) : fs + fk;
      }
    }
    P = fma(X, gs, P);
    if (SPINOR) S += fs;
  }
}

// Lane constants of the axis form: c and s^2 (FP32, s^2 rounded once).
struct LaneA {
  float c, a2;
};

// One point of the axis form (physical mode, n = 2): the same nodes, weights
// and sums as point_eval, with z_x and the midpoint's x rounded as there (so
// rho^2 and the cross products stay consistent where a loop passes the
// axis), the y parts from the axis record r.  Pz collects z_x (e_y g_A +
// 4 dw_y g_mid), Ps the w_y parts; the caller adds Pz - s Ps in FP64.
template <int PROF, int SPINOR, int ACC>
__device__ __forceinline__ void point_eval_axis(const float* __restrict__ r, const LaneA& la,
                                                const LaneC<float>& lc, int kind, const GP<float>& gp,
                                                const SplineRef& sp, float& Pz, float& Ps, float& S) {
  const float zx = fmaf(lc.s, r[0], la.c);
  const float zm = fmaf(lc.m8, r[2], zx);
  const float A = fmaf(zx, zx, fmaf(la.a2, r[1], 1e-30f));
  const float Am = fmaf(zm, zm, fmaf(la.a2, r[3], 1e-30f));
  float fA, fk;
  const float gA = node_g<float, PROF, ACC>(A, lc, kind, gp, M_PHYSICAL, sp, SPINOR, fA);
  const float gm = node_g<float, PROF, ACC>(Am, lc, kind, gp, M_PHYSICAL, sp, SPINOR, fk);
  Pz = fmaf(zx, fmaf(r[5], gA, r[4] * gm), Pz);
  Ps = fmaf(r[7], gA, fmaf(r[6], gm, Ps));
  if (SPINOR) S += fA + (fk + fk);
}

// (a library test: v - v == 0 is not one once the compiler contracts v = a b
// into fma(a, b, -v), the rounding error of the product)
template <class R> __device__ __forceinline__ bool finite_r(R v) { return __builtin_isfinite(v); }

// W - 1 of the three Richardson estimates from the per-loop sums
// (Integrand.c:352-390), in FP64; the reference's cosh guard
// (Integrand.c:371-377, |trunc(arg)| > 600) in reference mode.
__device__ __forceinline__ bool guard600(double v) {
  if (!(fabs(v) < 2147483648.0)) return false;
  return fabs(trunc(v)) > 600.0;
}
__device__ __forceinline__ void wminus1(int mode, bool spinor, const double sc[3], const double fe[3], double wm[3]) {
  // The three estimates take the angles a1 = sc0, a2 = (sc1 + sc2)/2 and
  // their mean a0 = (a1 + a2)/2 (likewise b from fe), so every half-angle
  // sine follows from one sincos of a1/4 and one of a2/4 by the double-angle
  // and addition formulas (bounded terms: absolute error of a few ulp, as
  // for the direct evaluation), and the hyperbolic values from expm1 of
  // b1/4, b2/4 and b0/2 (b0/2 directly: b1 and b2 of opposite signs would
  // cancel in an addition formula): five transcendental calls for nine.
  double s1, c1, s2, c2;
  sincos(0.25 * sc[0], &s1, &c1);
  sincos(0.125 * (sc[1] + sc[2]), &s2, &c2);
  const double sh[3] = {fma(s1, c2, c1 * s2), 2.0 * s1 * c1, 2.0 * s2 * c2};   // sin(a_k / 2)
  if (!spinor) {
#pragma unroll
    for (int k = 0; k < 3; ++k) wm[k] = -2.0 * sh[k] * sh[k];
    return;
  }
  double fyz = fe[1] + fe[2];
  if (mode == M_REFERENCE && (guard600(fyz) || guard600(fe[0]))) {
#pragma unroll
    for (int k = 0; k < 3; ++k) wm[k] = fma(-2.0 * sh[k], sh[k], 1.0) * 1.0e10 - 1.0;   // cos(a_k) 1e10 - 1
    return;
  }
  // sinh and cosh of p = b1/4 and q = b2/4 from e = expm1(|p|) >= 0: with
  // E = e + 1 >= 1, sinh|p| = e (1 + 1/E)/2 (no cancellation for small
  // arguments, and inf, not NaN, when E overflows) and cosh = (E + 1/E)/2;
  // sinh takes the sign of p
  const double p = 0.25 * fe[0], q = 0.125 * fyz;
  const double ep = expm1(fabs(p)), eq = expm1(fabs(q));
  const double ip = 1.0 / (ep + 1.0), iq = 1.0 / (eq + 1.0);
  const double shp = copysign(0.5 * ep * (1.0 + ip), p), chp = 0.5 * ((ep + 1.0) + ip);
  const double shq = copysign(0.5 * eq * (1.0 + iq), q), chq = 0.5 * ((eq + 1.0) + iq);
  const double r = p + q;                                   // b0 / 2
  const double er = expm1(fabs(r));
  const double shr = copysign(0.5 * er * (1.0 + 1.0 / (er + 1.0)), r);
  const double hh[3] = {shr, 2.0 * shp * chp, 2.0 * shq * chq};   // sinh(b_k / 2)
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    const double h2 = 2.0 * hh[k] * hh[k];          // cosh(b_k) - 1
    wm[k] = h2 - 2.0 * sh[k] * sh[k] * (1.0 + h2);   // 2 sinh^2(b/2) - 2 sin^2(a/2) cosh(b)
  }
}

template <class R>
__device__ __forceinline__ void lane_setup(const DevTuple& tp, const DevGeom& gm, LaneC<R>& lc) {
  const bool f32 = sizeof(R) == 4;
  lc.s = f32 ? (R)tp.s : (R)tp.sd;
  lc.m8 = (R)0.125 * lc.s;
  lc.s4 = (R)0.25 * lc.s;
  lc.inv_l2 = (R)gm.inv_l2;
  lc.c8 = f32 ? (R)gm.c8f : (R)1.0e-8;
  lc.c10 = f32 ? (R)gm.c10f : (R)1.0e-10;
}

// Per-loop outputs from the class sums (FP64): Wsscal = F s (h/3) P and
// Wsferm = 2 F T (2 h/3 spin_scale S) / N_class (intEV.cu:553-560), where s
// is sqrt(T) in the kernel's units and T the effective proper time.
__device__ __forceinline__ void finalize_loop(const double P[3], const double S[3], double F, double s,
                                              double T, double spin_scale, int n, int N, bool spinor,
                                              double sc[3], double fe[3]) {
  const double h3 = 1.0 / (3.0 * (double)n);
  sc[0] = F * s * h3 * P[0];
  sc[1] = F * s * h3 * P[1];
  sc[2] = F * s * h3 * P[2];
  if (spinor) {
    const double k = 4.0 * F * T * h3 * spin_scale;
    fe[0] = k * S[0] / (0.5 * N);
    fe[1] = k * S[1] / (0.25 * N);
    fe[2] = k * S[2] / (0.25 * N);
  } else {
    fe[0] = fe[1] = fe[2] = 0.0;
  }
}

// Spline-only kernels copy the plan's spline blob to LDS once per block
// (all threads, before any work); other kernels read it from global memory.
template <int PROF>
__device__ __forceinline__ const float4* stage_splines(const float4* __restrict__ sblob, int sblob_f4) {
  if constexpr (PROF == P_SPLINE || uniform_spline(PROF)) {
    extern __shared__ float4 wfl_lds_splines[];
    for (int i = threadIdx.x; i < sblob_f4; i += blockDim.x) wfl_lds_splines[i] = sblob[i];
    __syncthreads();
    return wfl_lds_splines;
  } else {
    return sblob;
  }
}

// ------------------------------------------------------ lane per tuple
// Class sums of one loop around one image tube for one lane's tuple: the
// points in 32-point blocks (FP32 partial sums, added to the FP64 P and S
// after each block), four points per step (wave-uniform records).  ACC
// selects the FP32 Gaussian node (0 plain, 1 and 2 accurate).
template <class R, int PROF, int SPINOR, int MODE, int NSUB, int ACC>
__device__ __forceinline__ void image_sums(const R* __restrict__ rl, int N, R ccx, R ccy, const LaneC<R>& lc,
                                           int kind, const GP<R>& gp, const SplineRef& sp, int nsub_rt,
                                           double P[3], double S[3]) {
  if constexpr (MODE == M_AXIS) {
    static_assert(sizeof(R) == 4 && NSUB == 2, "the axis form is FP32 with n = 2");
    const LaneA la = {ccx, lc.s * lc.s};
    for (int i0 = 0; i0 < N; i0 += 32) {
      float pz[3] = {0.0f, 0.0f, 0.0f}, ps[3] = {0.0f, 0.0f, 0.0f}, q[3] = {0.0f, 0.0f, 0.0f};
      const int i1 = i0 + 32 < N ? i0 + 32 : N;
      for (int i = i0; i < i1; i += 4) {
        const float* r = rl + (size_t)i * RECA;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int c = (u & 1) ? 0 : (u == 0 ? 2 : 1);
          point_eval_axis<PROF, SPINOR, ACC>(r + RECA * u, la, lc, kind, gp, sp, pz[c], ps[c], q[c]);
        }
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        P[c] += fma(-(double)lc.s, (double)ps[c], (double)pz[c]);
        if (SPINOR) S[c] += (double)q[c];
      }
    }
    return;
  }
  for (int i0 = 0; i0 < N; i0 += 32) {
    R p[3] = {(R)0, (R)0, (R)0}, q[3] = {(R)0, (R)0, (R)0};
    const int i1 = i0 + 32 < N ? i0 + 32 : N;
    for (int i = i0; i < i1; i += 4) {
      const R* r = rl + (size_t)i * REC;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int c = (u & 1) ? 0 : (u == 0 ? 2 : 1);
        point_eval<R, PROF, SPINOR, MODE, NSUB, ACC>(r[REC * u + 0], r[REC * u + 1], r[REC * u + 2],
                                                     r[REC * u + 3], r[REC * u + 4], r[REC * u + 5], ccx, ccy,
                                                     lc, kind, gp, sp, nsub_rt, p[c], q[c]);
      }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) { P[c] += (double)p[c]; if (SPINOR) S[c] += (double)q[c]; }
  }
}

// Class sums of one loop in the axis form with the local Taylor node (see
// LocalC): per point d = u - c^2 from the small terms t = s w_x and the
// y parts, t (t + 2c) + s^2 w_y^2 (and the same at the midpoint), and h(d)
// for the phase.  The spin node 2^-u is 2^-uc - D with
// D = uc h + d (a0 + h) (from 1 - 2^-u = u g(u)), so the FP32 sums carry
// only D, which is small against 2^-u, and no exp2 (whose results just
// below 1 are biased by about -0.12 ulp on gfx950).  The FP32 partial sums
// are flushed into P and S every 32 points, S as 3 n 2^-uc minus the sum of
// D (n = the block's points of the class).  The caller adds a0 s Q_c to P.
template <int SPINOR>
__device__ __forceinline__ void axis_local_sums(const float* __restrict__ rl, int N, float c, float s,
                                                const LocalC& tc, float a0f, float ucf, double ec, double P[3],
                                                double S[3]) {
  const float c2 = c + c, m8 = 0.125f * s, a2 = s * s;
  for (int i0 = 0; i0 < N; i0 += 32) {
    float pz[3] = {0.0f, 0.0f, 0.0f}, ps[3] = {0.0f, 0.0f, 0.0f}, q[3] = {0.0f, 0.0f, 0.0f};
    const int i1 = i0 + 32 < N ? i0 + 32 : N;
    for (int i = i0; i < i1; i += 4) {
      const float* r4 = rl + (size_t)i * RECA;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int k = (u & 1) ? 0 : (u == 0 ? 2 : 1);
        const float* r = r4 + RECA * u;
        const float t = s * r[0];
        const float zx = t + c;
        const float dA = fmaf(t, t + c2, a2 * r[1]);
        const float tm = fmaf(m8, r[2], t);
        const float dm = fmaf(tm, tm + c2, a2 * r[3]);
        const float hA = local_h(dA, tc), hm = local_h(dm, tc);
        pz[k] = fmaf(zx, fmaf(r[5], hA, r[4] * hm), pz[k]);
        ps[k] = fmaf(r[7], hA, fmaf(r[6], hm, ps[k]));
        if (SPINOR) {
          const float DA = fmaf(dA, a0f + hA, ucf * hA);
          const float Dm = fmaf(dm, a0f + hm, ucf * hm);
          q[k] += fmaf(2.0f, Dm, DA);
        }
      }
    }
    const double nb = (double)(i1 - i0);
    const double cnt[3] = {0.5 * nb, 0.25 * nb, 0.25 * nb};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      P[k] += fma(-(double)s, (double)ps[k], (double)pz[k]);
      if (SPINOR) S[k] += fma(3.0 * ec, cnt[k], -(double)q[k]);
    }
  }
}

// The FP32 Gaussian node for image tube c (centre-relative offset c' in
// kernel units), one choice for the whole wavefront (ballots), so every lane
// runs the same code: the accurate node when some lane's loops are small
// (s' < 1/2) and can reach that tube's core (|c'| - s' rmax < 0.75, i.e.
// u' < 0.56), the bounded form (1) when every lane's points stay within
// u' < 90 of the tube (|c'| + s' rmax < 9.5), else the clamped form (2).
template <class R>
__device__ __forceinline__ int image_acc(const DevGeom& gm, const LaneC<R>& lc, R ccx, R ccy) {
  if constexpr (sizeof(R) != 4) {
    return 0;
  } else {
    if (!gm.acc_on) return 0;
    const float d = __builtin_sqrtf(fmaf(ccx, ccx, ccy * ccy));
    const float reach = lc.s * gm.acc_rmax;
    if (__builtin_amdgcn_ballot_w64(lc.s < 0.5f && d - reach < 0.75f) == 0) return 0;
    return __builtin_amdgcn_ballot_w64(!(d + reach < 9.5f)) == 0 ? 1 : 2;
  }
}

// Class sums of one loop over all image tubes for one lane's tuple.
template <class R, int PROF, int SPINOR, int MODE, int NSUB>
__device__ __forceinline__ void loop_sums(const R* __restrict__ rl, int N, int M, bool lattice, const DevTuple& tp,
                                          const DevGeom& gm, R cx, R cy, R lat, const LaneC<R>& lc, int kind,
                                          const GP<R>& gp, const SplineRef& sp, int nsub_rt, double P[3],
                                          double S[3]) {
  for (int ia = -M; ia <= M; ++ia) {
    for (int ib = -M; ib <= M; ++ib) {
      R ccx = cx, ccy = cy;
      if (M > 0 || lattice) {
        ccx = fma(-(R)(tp.m0 + ia), lat, cx);
        ccy = fma(-(R)(tp.n0 + ib), lat, cy);
      }
      const int acc = (PROF == P_GAUSSIAN) ? image_acc<R>(gm, lc, ccx, ccy) : 0;
      if (acc == 1)
        image_sums<R, PROF, SPINOR, MODE, NSUB, 1>(rl, N, ccx, ccy, lc, kind, gp, sp, nsub_rt, P, S);
      else if (acc == 2)
        image_sums<R, PROF, SPINOR, MODE, NSUB, 2>(rl, N, ccx, ccy, lc, kind, gp, sp, nsub_rt, P, S);
      else
        image_sums<R, PROF, SPINOR, MODE, NSUB, 0>(rl, N, ccx, ccy, lc, kind, gp, sp, nsub_rt, P, S);
    }
  }
}

// Class sums of one loop over the (2M+1)^2 image tubes of a lattice
// window, points outer and image tubes inner.  In a square window z's x
// part depends only on the window column and its y part only on the row,
// so per point each column costs z_x and the midpoint's x (2 VALU), each
// row z_y, the midpoint's y, their squares and z_y e_x, z_y 4 dw_x (6), and
// each image tube its two rho^2 FMAs, two nodes, two cross products and the
// two accumulations, with the same roundings as point_eval.  Physical
// mode, n = 2, the plain node (the caller has checked that no image tube
// of the wavefront asks for the accurate Gaussian node).  The FP32 sums
// over one window row's image tubes are added to P and S in FP64 (one sum
// over all of a point's image tubes doubled the rounding of the group sums
// on dense lattices).
template <class R, int PROF, int SPINOR, int M>
__device__ __forceinline__ void lattice_sums(const R* __restrict__ rl, int N, const R* ccx, const R* ccy,
                                             const LaneC<R>& lc, int kind, const GP<R>& gp, const SplineRef& sp,
                                             double P[3], double S[3]) {
  constexpr int W = 2 * M + 1;
  const R tiny = sizeof(R) == 4 ? (R)1e-30 : (R)0;   // keeps 1/u finite on the axis (FP32 node)
  for (int i = 0; i < N; i += 4) {
    const R* r4 = rl + (size_t)i * REC;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int c = (u & 1) ? 0 : (u == 0 ? 2 : 1);
      const R* r = r4 + REC * u;
      const R wx = r[0], wy = r[1], dx4 = r[2], dy4 = r[3], ex = r[4], ey = r[5];
      R zx[W], mx[W];
#pragma unroll
      for (int a = 0; a < W; ++a) {
        zx[a] = fma(lc.s, wx, ccx[a]);
        mx[a] = fma(lc.m8, dx4, zx[a]);
      }
#pragma unroll
      for (int b = 0; b < W; ++b) {
        R p = (R)0, q = (R)-0.0;   // -0: the first spin term's add folds away
        const R zy = fma(lc.s, wy, ccy[b]);
        const R my = fma(lc.m8, dy4, zy);
        const R zy2 = fma(zy, zy, tiny), my2 = fma(my, my, tiny);
        const R zyex = zy * ex, zydx = zy * dx4;
#pragma unroll
        for (int a = 0; a < W; ++a) {
          R fA, fk;
          const R gA = node_g<R, PROF, 0>(fma(zx[a], zx[a], zy2), lc, kind, gp, M_PHYSICAL, sp, SPINOR, fA);
          const R gm = node_g<R, PROF, 0>(fma(mx[a], mx[a], my2), lc, kind, gp, M_PHYSICAL, sp, SPINOR, fk);
          const R E = fma(zx[a], ey, -zyex);    // z x e
          const R X = fma(zx[a], dy4, -zydx);   // 4 z x dw
          p = fma(E, gA, fma(X, gm, p));
          if (SPINOR) q += fA + (fk + fk);
        }
        P[c] += (double)p;
        if (SPINOR) S[c] += (double)q;
      }
    }
  }
}

// SPLIT = 1 writes each loop's (Wsscal, Wsferm) to perloop and leaves W - 1
// and the group sums to k_finalize_groups: the FP64 trigonometry then does
// not set this kernel's register allocation (more resident wavefronts).
template <class R, int PROF, int SPINOR, int MODE, int NSUB, int SPLIT>
__device__ __forceinline__ void tuple_task(
    int task_id, int lane, const R* __restrict__ rec, int N, int Nl, const DevTuple* __restrict__ tups,
    const DevGeom* __restrict__ geoms, const int* __restrict__ lanes,
    const float4* __restrict__ sblob,
    const Task* __restrict__ tasks, int nslots, int nsub_rt,
    double* __restrict__ gsum, double* __restrict__ perloop,
    unsigned long long* __restrict__ nonfinite, const LoopInfo* __restrict__ linfo) {
  const Task tk = tasks[task_id];
  const int lg = __builtin_amdgcn_readfirstlane(tk.lanegroup);
  const int loop0 = __builtin_amdgcn_readfirstlane(tk.loop0);
  const int nloop = __builtin_amdgcn_readfirstlane(tk.nloop);
  const int slot = __builtin_amdgcn_readfirstlane(tk.slot);
  const int tid_raw = lanes[lg * 64 + lane];
  const bool valid = tid_raw >= 0;
  const int tid = valid ? tid_raw : lanes[lg * 64];
  const DevTuple tp = tups[tid];
  const DevGeom gm = geoms[tp.geom];
  const GP<R> gp(gm);
  LaneC<R> lc;
  lane_setup<R>(tp, gm, lc);
  const double s_phase = sizeof(R) == 4 ? (double)tp.s : tp.sd;
  const double spin_scale = (PROF == P_GAUSSIAN) ? gm.inv_l2 : 1.0;
  const int kind = gm.kind;
  const SplineRef sp = spline_ref(sblob, gm);
  const int n = NSUB ? NSUB : nsub_rt;
  // image window (uniform over the lane group: one geometry per group)
  const int M = MODE == M_AXIS ? 0 : __builtin_amdgcn_readfirstlane(gm.lattice > 0.0 ? gm.images : 0);
  const R lat = sizeof(R) == 4 ? (R)gm.latticef : (R)gm.lattice;
  const R cx = sizeof(R) == 4 ? (R)tp.cx : (R)tp.cxd, cy = sizeof(R) == 4 ? (R)tp.cy : (R)tp.cyd;

  const bool lattice = MODE != M_AXIS && gm.lattice > 0.0;
  constexpr int RS = MODE == M_AXIS ? RECA : REC;   // record stride

  // local Taylor node: every lane of the group has one (the host groups them)
  constexpr bool kLocal = MODE == M_AXIS && PROF == P_GAUSSIAN && sizeof(R) == 4;
  const bool loc_group = kLocal && linfo && __builtin_amdgcn_readfirstlane(tp.loc) != 0;

  // FP32, 3 x 3 and 5 x 5 windows, physical mode, n = 2: image tubes inner
  // (lattice_sums) unless an image tube of the wavefront asks for the
  // accurate Gaussian node (a choice that depends on the tuples only).  The
  // FP64 kernels keep the image tubes outer: unrolled over the window, the
  // software exp and divide of the FP64 node take 256 VGPRs and spill.
  constexpr bool kLatInner = MODE == M_PHYSICAL && NSUB == 2 && sizeof(R) == 4;
  bool lat_inner = false;
  if constexpr (kLatInner) {
    if (lattice && (M == 1 || M == 2)) {
      lat_inner = true;
      if constexpr (PROF == P_GAUSSIAN)
        for (int ia = -M; ia <= M; ++ia)
          for (int ib = -M; ib <= M; ++ib)
            if (image_acc<R>(gm, lc, fma(-(R)(tp.m0 + ia), lat, cx), fma(-(R)(tp.n0 + ib), lat, cy)) != 0)
              lat_inner = false;
    }
  }

  double G[3] = {0.0, 0.0, 0.0};
  for (int l = loop0; l < loop0 + nloop; ++l) {
    const R* __restrict__ rl = (const R*)__builtin_assume_aligned(rec + (size_t)l * N * RS, 2 * sizeof(R));
    double P[3] = {0.0, 0.0, 0.0}, S[3] = {0.0, 0.0, 0.0};
    bool done = false;
    if constexpr (kLocal) {
      if (loc_group) {
        // this loop's nodes stay within kLocalDmax of c^2 for this lane's
        // tuple (lanes that differ take the two paths one after the other)
        const double rlw = linfo[l].r;
        const float r = (float)rlw;
        const float dmax = lc.s * r * fmaf(lc.s, r, 2.0f * fabsf((float)cx));
        if (dmax <= kLocalDmax) {
          // the coefficients are read here, per loop, so that they occupy
          // registers only on this path (hoisted, they would hold 12 VGPRs
          // through the other path and cost a wavefront per SIMD)
          int tl = tid;
          asm volatile("" : "+v"(tl));
          const DevTuple* __restrict__ tq = tups + tl;
          const LocalC tc = {tq->ta[0], tq->ta[1], tq->ta[2], tq->ta[3], tq->ta[4], tq->ta[5], tq->ta[6],
                             tq->ta[7]};
          axis_local_sums<SPINOR>((const float*)rl, N, (float)cx, (float)lc.s, tc, (float)tq->ta0,
                                  (float)((double)cx * (double)cx), tq->tec, P, S);
          const double a0s = tq->ta0 * (double)lc.s;
          P[0] = fma(a0s, linfo[l].q[0], P[0]);
          P[1] = fma(a0s, linfo[l].q[1], P[1]);
          P[2] = fma(a0s, linfo[l].q[2], P[2]);
          done = true;
        }
      }
    }
    if constexpr (kLatInner) {
      if (lat_inner) {
        R ccx[5], ccy[5];
#pragma unroll
        for (int a = 0; a < 5; ++a) {
          ccx[a] = fma(-(R)(tp.m0 + a - M), lat, cx);
          ccy[a] = fma(-(R)(tp.n0 + a - M), lat, cy);
        }
        if (M == 1)
          lattice_sums<R, PROF, SPINOR, 1>(rl, N, ccx, ccy, lc, kind, gp, sp, P, S);
        else
          lattice_sums<R, PROF, SPINOR, 2>(rl, N, ccx, ccy, lc, kind, gp, sp, P, S);
        done = true;
      }
    }
    if (!done)
      loop_sums<R, PROF, SPINOR, MODE, NSUB>(rl, N, M, lattice, tp, gm, cx, cy, lat, lc, kind, gp, sp, nsub_rt, P, S);
    double sc[3], fe[3];
    finalize_loop(P, S, gm.F, s_phase, tp.sd * tp.sd, spin_scale, n, N, SPINOR, sc, fe);
    if (!SPLIT) {
      double wm[3];
      wminus1(MODE, SPINOR, sc, fe, wm);
      bool ok = finite_r(wm[0]) && finite_r(wm[1]) && finite_r(wm[2]);
      if (valid && !ok) atomicAdd(nonfinite, 1ull);
      G[0] += wm[0]; G[1] += wm[1]; G[2] += wm[2];
    }
    if (perloop && valid) {
      double* o = perloop + ((size_t)tid * Nl + l) * 6;
      o[0] = sc[0]; o[1] = sc[1]; o[2] = sc[2]; o[3] = fe[0]; o[4] = fe[1]; o[5] = fe[2];
    }
  }
  if (!SPLIT && valid) {
    double* o = gsum + ((size_t)tid * nslots + slot) * 3;
    o[0] = G[0]; o[1] = G[1]; o[2] = G[2];
  }
}

// Next task of a persistent wavefront: one lane takes a ticket (a vector
// atomic), the wavefront reads it from that lane.
__device__ __forceinline__ int next_task(int* __restrict__ queue, int lane) {
  int t = 0;
  if (lane == 0) t = atomicAdd(queue, 1);
  return __shfl(t, 0, 64);
}

// wpb wavefronts per block.  Static: wavefront w of the grid runs task w.
// Persistent: a grid sized to the machine; each wavefront keeps taking the
// next task from *queue (zeroed before the launch) until none is left, so
// uneven tasks (lattice images, partial loop groups) balance themselves.
#define WFL_K_TUPLE_PARAMS                                                                   \
    const R* __restrict__ rec, int N, int Nl, const DevTuple* __restrict__ tups,             \
    const DevGeom* __restrict__ geoms, const int* __restrict__ lanes,                        \
    const float4* __restrict__ sblob, int sblob_f4,                                          \
    const Task* __restrict__ tasks, int ntasks, int nslots, int nsub_rt,                     \
    double* __restrict__ gsum, double* __restrict__ perloop,                                 \
    unsigned long long* __restrict__ nonfinite, int wpb, int persistent, int* __restrict__ queue, \
    const LoopInfo* __restrict__ linfo

template <class R, int PROF, int SPINOR, int MODE, int NSUB, int SPLIT>
__device__ __forceinline__ void k_tuple_body(WFL_K_TUPLE_PARAMS) {
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const float4* sb = stage_splines<PROF>(sblob, sblob_f4);
  int task_id = persistent ? next_task(queue, lane) : (int)blockIdx.x * wpb + wv;
  while (task_id < ntasks) {
    tuple_task<R, PROF, SPINOR, MODE, NSUB, SPLIT>(task_id, lane, rec, N, Nl, tups, geoms, lanes, sb,
                                                   tasks, nslots, nsub_rt, gsum, perloop, nonfinite, linfo);
    if (!persistent) break;
    task_id = next_task(queue, lane);
  }
}

template <class R, int PROF, int SPINOR, int MODE, int NSUB, int SPLIT>
__global__ __launch_bounds__(256) void k_tuple(WFL_K_TUPLE_PARAMS) {
  k_tuple_body<R, PROF, SPINOR, MODE, NSUB, SPLIT>(rec, N, Nl, tups, geoms, lanes, sblob, sblob_f4, tasks, ntasks,
                                                   nslots, nsub_rt, gsum, perloop, nonfinite, wpb, persistent, queue,
                                                   linfo);
}

// ------------------------------------------------------- wave per loop
// One loop's sums for one tuple around one image tube, lanes striding over
// points (lane l takes points l + 64k, all of class l & 3), in blocks of 8
// points per lane.
template <class R, int PROF, int SPINOR, int MODE, int NSUB, int ACC>
__device__ __forceinline__ void wave_image_sums(const R* __restrict__ rl, int N, R ccx, R ccy, const LaneC<R>& lc,
                                                int kind, const GP<R>& gp, const SplineRef& sp, int nsub_rt,
                                                int lane, double& P, double& S) {
  for (int i0 = lane; i0 < N; i0 += 64 * 8) {
    R p = (R)0, q = (R)0;
    const int i1 = i0 + 64 * 8 < N ? i0 + 64 * 8 : N;
    for (int i = i0; i < i1; i += 64) {
      const R* r = rl + (size_t)i * REC;
      point_eval<R, PROF, SPINOR, MODE, NSUB, ACC>(r[0], r[1], r[2], r[3], r[4], r[5], ccx, ccy, lc, kind, gp, sp,
                                                   nsub_rt, p, q);
    }
    P += (double)p;
    S += (double)q;
  }
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) {
  return __shfl_xor(v, m, 64);
}

// One (tuple, loop) pair w = t Nl + l of the wave-per-loop mapping.
template <class R, int PROF, int SPINOR, int MODE, int NSUB>
__device__ __forceinline__ void loop_pair(
    long long w, const R* __restrict__ rec, int N, int Nl, const DevTuple* __restrict__ tups,
    const DevGeom* __restrict__ geoms, const float4* __restrict__ sb, int nsub_rt,
    double* __restrict__ perloop, double* __restrict__ wm1,
    float4* __restrict__ ws_scal, float4* __restrict__ ws_ferm,
    unsigned long long* __restrict__ nonfinite) {
  const int t = (int)(w / Nl), l = (int)(w % Nl);
  const int lane = threadIdx.x & 63;
  const DevTuple tp = tups[t];
  if (l >= tp.nloops) return;
  const DevGeom gm = geoms[tp.geom];
  const GP<R> gp(gm);
  LaneC<R> lc;
  lane_setup<R>(tp, gm, lc);
  const double s_phase = sizeof(R) == 4 ? (double)tp.s : tp.sd;
  const double spin_scale = (PROF == P_GAUSSIAN) ? gm.inv_l2 : 1.0;
  const int kind = gm.kind;
  const SplineRef sp = spline_ref(sb, gm);
  const int n = NSUB ? NSUB : nsub_rt;
  const int M = gm.lattice > 0.0 ? gm.images : 0;
  const R lat = sizeof(R) == 4 ? (R)gm.latticef : (R)gm.lattice;
  const R cx = sizeof(R) == 4 ? (R)tp.cx : (R)tp.cxd, cy = sizeof(R) == 4 ? (R)tp.cy : (R)tp.cyd;
  const R* __restrict__ rl = rec + (size_t)l * N * REC;

  double P = 0.0, S = 0.0;      // this lane's class: lane & 3 (points i = lane + 64k)
  for (int ia = -M; ia <= M; ++ia) {
    for (int ib = -M; ib <= M; ++ib) {
      R ccx = cx, ccy = cy;
      if (gm.lattice > 0.0) {
        ccx = fma(-(R)(tp.m0 + ia), lat, cx);
        ccy = fma(-(R)(tp.n0 + ib), lat, cy);
      }
      const int acc = (PROF == P_GAUSSIAN) ? image_acc<R>(gm, lc, ccx, ccy) : 0;
      if (acc == 1)
        wave_image_sums<R, PROF, SPINOR, MODE, NSUB, 1>(rl, N, ccx, ccy, lc, kind, gp, sp, nsub_rt, lane, P, S);
      else if (acc == 2)
        wave_image_sums<R, PROF, SPINOR, MODE, NSUB, 2>(rl, N, ccx, ccy, lc, kind, gp, sp, nsub_rt, lane, P, S);
      else
        wave_image_sums<R, PROF, SPINOR, MODE, NSUB, 0>(rl, N, ccx, ccy, lc, kind, gp, sp, nsub_rt, lane, P, S);
    }
  }
  // lanes with equal (lane & 3) hold the same class: reduce over them
#pragma unroll
  for (int m = 4; m < 64; m <<= 1) {
    P += shfl_xor_d(P, m);
    if (SPINOR) S += shfl_xor_d(S, m);
  }
  // lane 0: class z (0 mod 4), lanes 1 and 3: class x (odd), lane 2: class y
  double P1 = __shfl(P, 1, 64), P2 = __shfl(P, 2, 64), P3 = __shfl(P, 3, 64);
  double S1 = __shfl(S, 1, 64), S2 = __shfl(S, 2, 64), S3 = __shfl(S, 3, 64);
  if (lane == 0) {
    double Pc[3] = {P1 + P3, P2, P};
    double Sc[3] = {S1 + S3, S2, S};
    double sc[3], fe[3], wm[3];
    finalize_loop(Pc, Sc, gm.F, s_phase, tp.sd * tp.sd, spin_scale, n, N, SPINOR, sc, fe);
    wminus1(MODE, SPINOR, sc, fe, wm);
    if (!(finite_r(wm[0]) && finite_r(wm[1]) && finite_r(wm[2]))) atomicAdd(nonfinite, 1ull);
    const size_t o = (size_t)t * Nl + l;
    if (perloop) {
      double* po = perloop + o * 6;
      po[0] = sc[0]; po[1] = sc[1]; po[2] = sc[2]; po[3] = fe[0]; po[4] = fe[1]; po[5] = fe[2];
    }
    if (wm1) { wm1[o * 3] = wm[0]; wm1[o * 3 + 1] = wm[1]; wm1[o * 3 + 2] = wm[2]; }
    if (ws_scal) ws_scal[l] = make_float4((float)sc[0], (float)sc[1], (float)sc[2], 0.0f);
    if (ws_ferm && SPINOR) ws_ferm[l] = make_float4((float)fe[0], (float)fe[1], (float)fe[2], 0.0f);
  }
}

// Wavefront w of the grid takes the pairs w, w + W, w + 2W, ... (W = the
// grid's wavefronts), so a capped grid covers any ntup x Nl.
template <class R, int PROF, int SPINOR, int MODE, int NSUB>
__global__ __launch_bounds__(256) void k_loop(
    const R* __restrict__ rec, int N, int Nl, const DevTuple* __restrict__ tups,
    const DevGeom* __restrict__ geoms, int ntup,
    const float4* __restrict__ sblob, int sblob_f4, int nsub_rt,
    double* __restrict__ perloop, double* __restrict__ wm1,
    float4* __restrict__ ws_scal, float4* __restrict__ ws_ferm,
    unsigned long long* __restrict__ nonfinite, int wpb) {
  const float4* sb = stage_splines<PROF>(sblob, sblob_f4);   // before any wavefront leaves
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long long total = (long long)ntup * Nl, stride = (long long)gridDim.x * wpb;
  for (long long w = (long long)blockIdx.x * wpb + wv; w < total; w += stride)
    loop_pair<R, PROF, SPINOR, MODE, NSUB>(w, rec, N, Nl, tups, geoms, sb, nsub_rt, perloop, wm1, ws_scal, ws_ferm,
                                           nonfinite);
}

}  // namespace wfl

// Launch interface of one precision's kernel instantiations.
struct WflLaunch {
  const void* rec; int N, Nl;
  const float* rec_axis;                      // axis-form records (FP32, lane-per-tuple kernels)
  int axis;                                   // 1: every tuple on the x axis, no lattice, n = 2
  const wfl::LoopInfo* linfo;                 // per-loop constants of the local Taylor node (axis form)
  const wfl::DevTuple* tups; const wfl::DevGeom* geoms; int ntup;
  const int* lanes; const wfl::Task* tasks; int ntasks; int nslots;
  const float4* sblob; int sblob_f4;          // spline blob and its size in float4
  int lds_splines;                            // 1: spline-only kernels with the blob staged in LDS
  int uni_splines, sblob_uni_f4;              // 1: uniform-grid rows only (FP32 physical); their size
  int uni_origin;                             // 1: those rows are expanded about u = 0
  int prof_gauss, spinor, mode, nsub, split;
  int wpb, persistent, grid;   // wavefronts per block; persistent grid size in blocks
  int* queue;
  double* gsum; double* perloop; double* wm1;
  float4* ws_scal; float4* ws_ferm;
  unsigned long long* nonfinite;
  hipStream_t stream;
};
