/*
 * wfl_oracle.c -- FP64 CPU oracle (TEST INFRASTRUCTURE ONLY; never linked
 * into the product library).  See wfl_oracle.h for the reference map.
 *
 * Pins (each function is checked by a test that a plausible mistake fails):
 *   orc_philox4x32_10   Random123 known-answer vectors            tests/test_oracle_loops.py
 *   orc_gen_loops       numpy restatement, closure, CM = 0, increment variance,
 *                       Levy-area law eBT/sinh(eBT) (through orc_eval_batch),
 *                       committed SHA-256 of the FP32 buffers      test_oracle_loops.py
 *   orc_succ (getzp1)   brute-force orbit enumeration N = 4 ... 64   test_oracle_wilson.py
 *   orc_wilson_loop     shoelace identity Phi = (2F/l2) Area (constant field, both
 *                       modes); Simpson exact for quadratic integrands; arctan
 *                       closed form for the smooth profile with O(h^4) convergence;
 *                       hand-built B<0 segments for the reference clamps; rotation,
 *                       reflection and reversal invariances; sqrt(T) w scaling;
 *                       lattice K=1 limit, far images, lattice-vector translation
 *                                                                  test_oracle_wilson.py
 *   orc_eval_batch      group sums of orc_wilson_loop (Levy-area law)  test_oracle_loops.py
 *   orc_f / orc_fp      closed forms, f' by finite differences, periodic rings   test_oracle_math.py
 *   orc_expint          scipy.special.expi                          test_oracle_math.py
 *   orc_spline,         scipy CubicSpline(bc_type=clamped, 0 slopes); the
 *   orc_interp,         Gaussian f reproduced; zero outside the knots in
 *   orc_spline_low_slope  reference mode; the line s u below the first knot  test_oracle_math.py
 *   orc_wminus1,        hand-computed vectors, Richardson coefficients cancel
 *   orc_evmean          1/N and 1/N^2 terms, the cosh guard        test_oracle_math.py
 *   orc_exact, smT*     mpmath; small-T series = Taylor series of Exact   test_oracle_math.py
 *   orc_igrand          the branch policy's closed-form branches    test_oracle_math.py
 *   orc_logsimpson      exact for cubics in ln T                    test_oracle_math.py
 *   orc_field_at,       single-tube window = isolated, periodic in x_cm, cell
 *   orc_igrand_B        average 2 pi F/(e l^2)                      test_oracle_math.py
 *   orc_magfield        -1/2 int B^2 = the reference's closed-form ClassAction
 *                       (step, smooth, quadratic, Gaussian)         test_abi.py
 */
#include "wfl_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static const double ORC_E = 0.30282212;   /* Integrand.c:14 */
static const double ORC_M2 = 1.0;         /* Integrand.c:12 */

/* ------------------------------------------------------------------ */
/* Philox4x32-10                                                        */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r > 0) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Closed unit loop: increments N(0, 2/N) per component from Philox +
 * Box-Muller (counter = (point pair j, loop l, 0, 0), key = (0x5EED0000 +
 * config_id, 0)), mean removed, inclusive cumulative sum, centre of mass
 * removed, rounded to FP32.  All sums sequential in index order. */
void orc_gen_loops(uint32_t config_id, int Nl, int N, float *x, float *y) {
  const double sigma = sqrt(2.0 / (double)N);
  const double two_pi = 6.283185307179586476925286766559;
  const double inv32 = 1.0 / 4294967296.0;
  double *dx = (double *)malloc(sizeof(double) * (size_t)N);
  double *dy = (double *)malloc(sizeof(double) * (size_t)N);
  for (int l = 0; l < Nl; ++l) {
    for (int j = 0; j < N / 2; ++j) {
      uint32_t ctr[4] = {(uint32_t)j, (uint32_t)l, 0u, 0u};
      uint32_t key[2] = {0x5EED0000u + config_id, 0u};
      uint32_t r[4];
      orc_philox4x32_10(ctr, key, r);
      double u0 = ((double)r[0] + 0.5) * inv32, u1 = ((double)r[1] + 0.5) * inv32;
      double u2 = ((double)r[2] + 0.5) * inv32, u3 = ((double)r[3] + 0.5) * inv32;
      double ra = sqrt(-2.0 * log(u0)), aa = two_pi * u1;
      double rb = sqrt(-2.0 * log(u2)), ab = two_pi * u3;
      dx[2 * j] = sigma * ra * cos(aa);
      dy[2 * j] = sigma * ra * sin(aa);
      dx[2 * j + 1] = sigma * rb * cos(ab);
      dy[2 * j + 1] = sigma * rb * sin(ab);
    }
    double mx = 0.0, my = 0.0;
    for (int i = 0; i < N; ++i) { mx += dx[i]; my += dy[i]; }
    mx /= N; my /= N;
    double px = 0.0, py = 0.0, cx = 0.0, cy = 0.0;
    for (int i = 0; i < N; ++i) {
      px += dx[i] - mx; py += dy[i] - my;
      dx[i] = px; dy[i] = py;
      cx += px; cy += py;
    }
    cx /= N; cy /= N;
    for (int i = 0; i < N; ++i) {
      x[(size_t)l * N + i] = (float)(dx[i] - cx);
      y[(size_t)l * N + i] = (float)(dy[i] - cy);
    }
  }
  free(dx); free(dy);
}

/* ------------------------------------------------------------------ */
/* Profiles (intEV.cu:29-406; intEVff.cu:134-243)                      */
/* ------------------------------------------------------------------ */
/* E1 continued fraction (modified Lentz), Ei(x) = -E1(-x) for x < -1;
 * intEV.cu:29-61 follows Numerical Recipes' expint.  Converged to FP64. */
double orc_expint(double x) {
  const int MAXIT = 1000;
  const double EPS = 1e-15, BIG = 1e300;
  double b = -x + 1.0, c = BIG, d = 1.0 / b, h = d;
  for (int i = 1; i <= MAXIT; ++i) {
    double a = -(double)i * (double)i;
    b += 2.0;
    d = 1.0 / (a * d + b);
    c = b + a / c;
    double del = c * d;
    h *= del;
    if (fabs(del - 1.0) <= EPS) return -h * exp(x);
  }
  return 0.0;
}

static double orc_bump(double x, double thr) {   /* intEV.cu:152-159 */
  return (x * x < thr) ? exp(-1.0 / (1.0 - x * x)) : 0.0;
}

static double orc_phi(const orc_geom *g, double x, int ff) {   /* intEV.cu:162-174 */
  double om = 1.0 - x * x;
  if (x < (ff ? 1.0 : 0.999))
    return 1.0 - (0.5 / g->q2) * (om * exp(-1.0 / om) + orc_expint(-1.0 / om));
  return 1.0;
}

static double orc_chi(const orc_geom *g, double x, double n, double lam, int ff) {
  /* intEV.cu:177-198 (fit A); intEVff.cu:160-172 (fit B) */
  double om = 1.0 - x * x, x2 = x * x, x4 = x2 * x2, a = g->tubedist;
  double thr = ff ? 1.0 : 0.999;
  if (x <= -thr) return 0.0;
  if (x2 < thr) {
    double ans = 0.5 * (-om * exp(-1.0 / om) - orc_expint(-1.0 / om));
    if (!ff)
      ans += 2.0 * n * a / lam * 0.444 /
             (1.0 + exp(-(3.31 * x + 5.25 * x * x2 * sin(x) * cos(-0.907 * x2 - 1.29 * x4 * x4)) / cos(x)));
    else
      ans += 2.0 * n * a / lam *
             (0.218 + 0.393 * x * cosh(0.806 * x4 - 0.696 * x2 + 0.0902) / cosh(0.825 * x2 - 0.0234 * x + 0.375));
    return ans;
  }
  return 2.0 * n * a * g->q1 / lam;
}

static double orc_flperi(const orc_geom *g, double rho2) {   /* intEV.cu:202-232 */
  const int N = 2;
  double a = g->tubedist, rho = sqrt(rho2), lam = sqrt(g->l2);
  double flraf = floor(rho / a), f = 0.0;
  int fl = (int)flraf;
  if (fl - N > 0) f = 0.5 * (flraf - N - 1.0) * (flraf - N);
  for (int i = fl - N; i <= fl + N; ++i) {
    if (i == 0) {
      double ex = exp(-rho / lam), ep1 = 1.0 + ex;
      f = -rho / a * (ex / ep1) - lam / a * log(ep1);
    } else if (i > 0) {
      f += (double)i / (1.0 + exp(-(rho - (double)i * a) / lam));
    }
  }
  return a / (lam * log(2.0)) * f + 1.0;
}

static double orc_fplperi(const orc_geom *g, double rho2) {  /* intEV.cu:301-328 */
  const int N = 2;
  double a = g->tubedist, rho = sqrt(rho2), lam = sqrt(g->l2);
  double flraf = floor(rho / a), f = 0.0;
  int fl = (int)flraf;
  for (int i = fl - N; i <= fl + N; ++i) {
    if (i == 0) {
      double ex = exp(-rho / lam), ep1 = 1.0 + ex;
      f = ex / (a * ep1 * ep1);
    } else if (i > 0) {
      double ea = exp(-(rho - (double)i * a) / lam), eap1 = 1.0 + ea;
      f += ((double)i / rho) * ea / (eap1 * eap1);
    }
  }
  return a / (2.0 * g->l2 * log(2.0)) * f;
}

static double orc_ffixflux(const orc_geom *g, double rho2, int ff) {  /* intEV.cu:237-257 */
  double a = g->tubedist, lam = sqrt(g->l2);
  double lml = (lam - g->lmin) / (a - g->lmin), aml = (a - lam) / (a - g->lmin);
  double n = floor((sqrt(rho2) + a / 2.0) / a);
  if (rho2 <= a * a / 4.0)
    return (1.0 - 0.75 * lml) * orc_phi(g, 2.0 * sqrt(rho2 / g->l2), ff) + 3.0 * rho2 / (a * a) * lml;
  double ans = 1.0 + 0.75 * (4.0 * rho2 / (a * a) - 1.0) * lml + 3.0 * n * (n - 1.0) * aml;
  ans += 3.0 * lam / (g->q1 * a) * aml * orc_chi(g, 2.0 * (sqrt(rho2) - n * a) / lam, n, lam, ff);
  return ans;
}

static double orc_fpfixflux(const orc_geom *g, double rho2) {  /* intEV.cu:334-357 */
  double a = g->tubedist, lam = sqrt(g->l2), rho = sqrt(rho2);
  double lml = (lam - g->lmin) / (a - g->lmin), aml = (a - lam) / (a - g->lmin);
  double n = floor((rho + a / 2.0) / a);
  if (rho <= a / 2.0)
    return 2.0 / (g->l2 * g->q2) * (1.0 - 0.75 * lml) * orc_bump(2.0 * rho / lam, 0.999) + 3.0 / (a * a) * lml;
  return 3.0 / (a * a) * lml + 6.0 / (g->q1 * lam * a) * aml * orc_bump(2.0 * (rho - n * a) / lam, 0.999);
}

double orc_interp(double rho2, const float *k, const float *c, int n) {  /* intEV.cu:64-107 */
  if (rho2 < (double)k[n - 1] && rho2 > (double)k[0]) {
    int lo = 0, hi = n - 1;
    while (hi - lo > 1) {
      int mid = (hi + lo) / 2;
      if (rho2 >= (double)k[mid]) lo = mid; else hi = mid;
    }
    double dx = rho2 - (double)k[lo];
    const float *q = c + 4 * lo;
    return (double)q[0] + dx * ((double)q[1] + dx * ((double)q[2] + dx * (double)q[3]));
  }
  return 0.0;
}

/* Spline in the physical mode (no counterpart in the reference, whose interp
 * returns 0 outside the knots): at or below the first knot the straight line
 * f = s u through f(0) = 0 with the FP32 slope s = fl32(a_0 / u_0) (a_0 when
 * the first knot is at 0), so that the line is one more table interval;
 * f = f(last knot) at or beyond the last knot. */
double orc_spline_low_slope(const float *k, const float *c) {
  return k[0] > 0.0f ? (double)(float)((double)c[0] / (double)k[0]) : (double)c[1];
}
static double orc_interp_phys(double rho2, const float *k, const float *c, int n) {
  if (rho2 <= (double)k[0]) return orc_spline_low_slope(k, c) * rho2;
  if (rho2 >= (double)k[n - 1]) return (double)c[4 * (n - 1)];
  return orc_interp(rho2, k, c, n);
}

static double orc_f_mode(const orc_geom *g, int mode, double u) {
  double l2 = g->l2;
  switch (g->profile) {
    case ORC_STEP: return u < l2 ? u / l2 : 1.0;
    case ORC_SMOOTH: return u / (l2 + u);
    case ORC_QUADRATIC: return u < l2 ? u / l2 * (2.0 - u / l2) : 1.0;
    case ORC_GAUSSIAN: return -expm1(-u / l2);
    case ORC_PERIODIC: return orc_flperi(g, u);
    case ORC_SPLINE:
      return mode == ORC_MODE_PHYSICAL ? orc_interp_phys(u, g->rho2sp, g->coefs, g->nspline)
                                       : orc_interp(u, g->rho2sp, g->coefs, g->nspline);
    case ORC_FIXFLUX: return orc_ffixflux(g, u, 0);
    case ORC_FIXFLUX_FF: return orc_ffixflux(g, u, 1);
  }
  return 0.0;
}

/* f(rho^2) in a given segment mode (physical mode extends the spline table). */
double orc_f_in_mode(const orc_geom *g, int mode, double u) { return orc_f_mode(g, mode, u); }

static double orc_fp_mode(const orc_geom *g, int mode, double u) {
  double l2 = g->l2;
  switch (g->profile) {
    case ORC_STEP: return u < l2 ? 1.0 / l2 : 0.0;
    case ORC_SMOOTH: return l2 / ((l2 + u) * (l2 + u));
    case ORC_QUADRATIC: return u < l2 ? 2.0 / l2 * (1.0 - u / l2) : 0.0;
    case ORC_GAUSSIAN: return exp(-u / l2) / l2;
    case ORC_PERIODIC: return orc_fplperi(g, u);
    case ORC_SPLINE:
      if (mode == ORC_MODE_PHYSICAL) {
        const float *k = g->rho2sp; int n = g->nspline;
        if (u <= (double)k[0]) return orc_spline_low_slope(k, g->coefs);
        if (u >= (double)k[n - 1]) return 0.0;
        int lo = 0, hi = n - 1;
        while (hi - lo > 1) { int mid = (hi + lo) / 2; if (u >= (double)k[mid]) lo = mid; else hi = mid; }
        double dx = u - (double)k[lo];
        const float *q = g->coefs + 4 * lo;
        return (double)q[1] + dx * (2.0 * (double)q[2] + dx * 3.0 * (double)q[3]);
      } else {
        /* intEV.cu:389-398: analytic sin^2 form, no division by tubedist */
        double s = sin(sqrt(u) * 3.14159265358979323846);
        return s > 100.0 ? 0.0 : exp(-s * s / l2) / l2;
      }
    case ORC_FIXFLUX:
    case ORC_FIXFLUX_FF: return orc_fpfixflux(g, u);
  }
  return 0.0;
}

double orc_f(const orc_geom *g, double rho2) { return orc_f_mode(g, ORC_MODE_REFERENCE, rho2); }
double orc_fp(const orc_geom *g, double rho2) { return orc_fp_mode(g, ORC_MODE_REFERENCE, rho2); }

/* g(u) = f(u)/u.  Reference mode: plain quotient (u is clamped by Idt).
 * Physical mode: the removable singularity at u = 0 is taken analytically. */
static double orc_g(const orc_geom *g, int mode, double u) {
  if (mode == ORC_MODE_REFERENCE) return orc_f_mode(g, mode, u) / u;
  double l2 = g->l2;
  switch (g->profile) {
    case ORC_STEP: return u < l2 ? 1.0 / l2 : 1.0 / u;
    case ORC_SMOOTH: return 1.0 / (l2 + u);
    case ORC_QUADRATIC: return u < l2 ? (2.0 - u / l2) / l2 : 1.0 / u;
    case ORC_GAUSSIAN: {
      double x = u / l2;
      return x > 0.0 ? -expm1(-x) / x / l2 : 1.0 / l2;
    }
    default:
      if (u > 0.0) return orc_f_mode(g, mode, u) / u;
      return orc_fp_mode(g, mode, 0.0);
  }
}

/* ------------------------------------------------------------------ */
/* Segment integral (Idt, intEV.cu:409-459) and Wilson loop             */
/* ------------------------------------------------------------------ */
static void orc_segment(const orc_geom *g, int mode, int nsub, int spinor,
                        double zx, double zy, double dx, double dy,
                        double *I, double *If) {
  const double h = 1.0 / (double)nsub;
  double s0, s2 = 0.0, s4 = 0.0, f0 = 0.0, f2 = 0.0, f4 = 0.0;
  if (mode == ORC_MODE_REFERENCE) {
    double A = zx * zx + zy * zy, B = zx * dx + zy * dy, C = dx * dx + dy * dy;
    if (A < 1.0e-8) A = 1.0e-8;
    if (B < 1.0e-8) B = 1.0e-8;      /* the reference's linear-coefficient clamp */
    if (C < 1.0e-8) C = 1.0e-8;
    double A1 = A + 2.0 * B + C;
    if (A1 < 1.0e-8) A1 = 1.0e-8;
    s0 = orc_g(g, mode, A) + orc_g(g, mode, A1);
    if (spinor) f0 = orc_fp_mode(g, mode, A) + orc_fp_mode(g, mode, A1);
    for (int k = 1; k < nsub; ++k) {
      double t = (double)k * h;
      double r = A + 2.0 * B * t + C * t * t;
      if (r < 1.0e-10) r = 1.0e-10;
      if (k % 2 == 0) { s2 += orc_g(g, mode, r); if (spinor) f2 += orc_fp_mode(g, mode, r); }
      else            { s4 += orc_g(g, mode, r); if (spinor) f4 += orc_fp_mode(g, mode, r); }
    }
  } else {
    double r0 = zx * zx + zy * zy;
    double ex = zx + dx, ey = zy + dy, r1 = ex * ex + ey * ey;
    s0 = orc_g(g, mode, r0) + orc_g(g, mode, r1);
    if (spinor) f0 = orc_fp_mode(g, mode, r0) + orc_fp_mode(g, mode, r1);
    for (int k = 1; k < nsub; ++k) {
      double t = (double)k * h;
      double px = zx + t * dx, py = zy + t * dy, r = px * px + py * py;
      if (k % 2 == 0) { s2 += orc_g(g, mode, r); if (spinor) f2 += orc_fp_mode(g, mode, r); }
      else            { s4 += orc_g(g, mode, r); if (spinor) f4 += orc_fp_mode(g, mode, r); }
    }
  }
  *I = (s0 + 2.0 * s2 + 4.0 * s4) * h / 3.0;
  *If = spinor ? (f0 + 2.0 * f2 + 4.0 * f4) * h / 3.0 : 0.0;
}

/* getzp1 successor map (intEV.cu:462-506). */
int orc_succ(int i, int N) {
  if (i % 2 == 1) return (i == N - 1) ? 1 : i + 2;
  if (i % 4 == 0) return (i == N - 4) ? 0 : i + 4;
  return (i == N - 2) ? 2 : i + 4;
}

static double orc_round_half_up(double v) { return floor(v + 0.5); }

void orc_wilson_loop(const orc_geom *g, int mode, int nsub, int spinor,
                     const float *wx, const float *wy, int N,
                     double cx, double cy, double rtT,
                     double scal[3], double ferm[3]) {
  double S[3] = {0, 0, 0}, Sf[3] = {0, 0, 0};   /* x: odd, y: 2 mod 4, z: 0 mod 4 */
  int M = 0;
  double l = g->lattice, m0 = 0.0, n0 = 0.0;
  if (l > 0.0) { M = g->images; m0 = orc_round_half_up(cx / l); n0 = orc_round_half_up(cy / l); }
  for (int ia = -M; ia <= M; ++ia) {
    for (int ib = -M; ib <= M; ++ib) {
      double ccx = cx, ccy = cy;
      if (l > 0.0) { ccx = cx - (m0 + ia) * l; ccy = cy - (n0 + ib) * l; }
      for (int i = 0; i < N; ++i) {
        int j = orc_succ(i, N);
        double zx = ccx + rtT * (double)wx[i], zy = ccy + rtT * (double)wy[i];
        double qx = ccx + rtT * (double)wx[j], qy = ccy + rtT * (double)wy[j];
        double dx = qx - zx, dy = qy - zy;
        double I, If;
        orc_segment(g, mode, nsub, spinor, zx, zy, dx, dy, &I, &If);
        double X = (mode == ORC_MODE_REFERENCE) ? (zx * qy - zy * qx) : (zx * dy - zy * dx);
        int c = (i % 2 == 1) ? 0 : ((i % 4 == 0) ? 2 : 1);
        S[c] += X * I;
        Sf[c] += If;
      }
    }
  }
  double F = g->F;
  scal[0] = F * S[0]; scal[1] = F * S[1]; scal[2] = F * S[2];
  ferm[0] = ferm[1] = ferm[2] = 0.0;
  if (spinor) {
    ferm[0] = 2.0 * F * Sf[0] * rtT * rtT / (N / 2.0);
    ferm[1] = 2.0 * F * Sf[1] * rtT * rtT / (N / 4.0);
    ferm[2] = 2.0 * F * Sf[2] * rtT * rtT / (N / 4.0);
  }
}

/* ------------------------------------------------------------------ */
/* Ensemble average (EVMean, Integrand.c:325-407)                       */
/* ------------------------------------------------------------------ */
/* `abs(double)` in Integrand.c:371 converts to int: |trunc(v)| > 600; on
 * x86-64 an out-of-range conversion yields INT_MIN, whose abs stays
 * negative, so the guard is off there. */
static int orc_guard600(double v) {
  if (!(fabs(v) < 2147483648.0)) return 0;
  return fabs(trunc(v)) > 600.0;
}

void orc_wminus1(int mode, int spinor, const double scal[3], const double ferm[3], double wm1[3]) {
  double yz = scal[1] + scal[2];
  double a[3] = {0.5 * scal[0] + 0.25 * yz, scal[0], 0.5 * yz};
  if (!spinor) {
    for (int k = 0; k < 3; ++k) { double s = sin(0.5 * a[k]); wm1[k] = -2.0 * s * s; }
    return;
  }
  double fyz = ferm[1] + ferm[2];
  double b[3] = {0.5 * ferm[0] + 0.25 * fyz, ferm[0], 0.5 * fyz};
  if (mode == ORC_MODE_REFERENCE && (orc_guard600(fyz) || orc_guard600(ferm[0]))) {
    for (int k = 0; k < 3; ++k) wm1[k] = cos(a[k]) * 1.0e10 - 1.0;
    return;
  }
  for (int k = 0; k < 3; ++k) {
    double s = sin(0.5 * a[k]), sh = sinh(0.5 * b[k]);
    wm1[k] = 2.0 * sh * sh - 2.0 * s * s * cosh(b[k]);
  }
}

double orc_evmean(const float *ws, const float *wf, int n, const int *WL, int fermion) {
  double e1 = 0.0, e2 = 0.0, e4 = 0.0;
  for (int i = 0; i < n; ++i) {
    const float *s = ws + 4 * (size_t)WL[i];
    float nb4f = s[1] + s[2];
    double nb4 = (double)nb4f;
    double p1 = cos(0.5 * (double)s[0] + 0.25 * nb4);
    double p2 = cos((double)s[0]);
    double p4 = cos(0.5 * nb4);
    if (fermion == 1) {
      const float *f = wf + 4 * (size_t)WL[i];
      float fb4f = f[1] + f[2];
      double fb4 = (double)fb4f;
      if (orc_guard600(fb4) || orc_guard600((double)f[0])) {
        p1 *= 1.0e10; p2 *= 1.0e10; p4 *= 1.0e10;
      } else {
        p1 *= cosh(0.5 * (double)f[0] + 0.25 * fb4);
        p2 *= cosh((double)f[0]);
        p4 *= cosh(0.5 * fb4);
      }
    }
    e4 += p4; e2 += p2; e1 += p1;
  }
  e4 /= n; e2 /= n; e1 /= n;
  return 8.0 / 3.0 * e1 - 2.0 * e2 + 1.0 / 3.0 * e4;
}

void orc_eval_batch(const orc_geom *geoms, int mode, int nsub, int spinor,
                    const float *x, const float *y, int Nl, int N, int group,
                    int ntup, const int *tup_geom, const double *tup_cx,
                    const double *tup_cy, const double *tup_rtT,
                    double *perloop, double *gsum, int nthreads) {
  int ng = (Nl + group - 1) / group;
  double *wm = (double *)malloc(sizeof(double) * 3 * (size_t)ntup * Nl);
  long total = (long)ntup * Nl;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
  for (long w = 0; w < total; ++w) {
    int t = (int)(w / Nl), l = (int)(w % Nl);
    double sc[3], fe[3], wm1[3];
    orc_wilson_loop(&geoms[tup_geom[t]], mode, nsub, spinor, x + (size_t)l * N, y + (size_t)l * N, N,
                    tup_cx[t], tup_cy[t], tup_rtT[t], sc, fe);
    orc_wminus1(mode, spinor, sc, fe, wm1);
    if (perloop) {
      double *p = perloop + 6 * (size_t)w;
      p[0] = sc[0]; p[1] = sc[1]; p[2] = sc[2]; p[3] = fe[0]; p[4] = fe[1]; p[5] = fe[2];
    }
    wm[3 * w] = wm1[0]; wm[3 * w + 1] = wm1[1]; wm[3 * w + 2] = wm1[2];
  }
  if (gsum) {
    for (int t = 0; t < ntup; ++t)
      for (int q = 0; q < ng; ++q) {
        double s0 = 0, s1 = 0, s2 = 0;
        int l1 = (q + 1) * group < Nl ? (q + 1) * group : Nl;
        for (int l = q * group; l < l1; ++l) {
          const double *p = wm + 3 * ((size_t)t * Nl + l);
          s0 += p[0]; s1 += p[1]; s2 += p[2];
        }
        double *o = gsum + 3 * ((size_t)t * ng + q);
        o[0] = s0; o[1] = s1; o[2] = s2;
      }
  }
  free(wm);
}

/* ------------------------------------------------------------------ */
/* Proper-time integrand policy (Integrand.c:20-316)                    */
/* ------------------------------------------------------------------ */
double orc_exact(double T, double B, int fermion) {
  double TB = ORC_E * T * B;
  if (B < 1.0e-40) return 0.0;
  if (fermion == 1) return exp(-ORC_M2 * T) / (T * T * T) * (TB / tanh(TB) - 1.0 - TB * TB / 3.0);
  if (TB < 50.0) return exp(-ORC_M2 * T) / (T * T * T) * (TB / sinh(TB) - 1.0 + TB * TB / 6.0);
  return exp(-ORC_M2 * T) / (T * T * T) * (TB * TB / 6.0 - 1.0);
}

double orc_smTscal(double B, double T) {
  double B2 = ORC_E * ORC_E * B * B, B4 = B2 * B2, B6 = B2 * B4, m2 = ORC_M2;
  double T3 = T * T * T, T4 = T3 * T;
  return 7.0 / 360.0 * B4 * T * (1.0 - m2 * T) + (147.0 * B4 * m2 * m2 - 31.0 * B6) / 15120.0 * T3 +
         (31.0 * B6 * m2 - 49.0 * B4 * m2 * m2 * m2) / 15120.0 * T4;
}

double orc_smTferm(double B, double T) {
  double B2 = ORC_E * ORC_E * B * B, B4 = B2 * B2, B6 = B2 * B4, m2 = ORC_M2;
  double T3 = T * T * T, T4 = T3 * T;
  return 1.0 / 45.0 * B4 * T * (m2 * T - 1.0) + (2.0 * B6 / 945.0 - B4 * m2 * m2 / 90.0) * T3 +
         (7.0 * B4 * m2 * m2 * m2 - 4.0 * B6 * m2) / 1890.0 * T4;
}

static double orc_periodicB(const orc_geom *g, double rho2) {   /* Integrand.c:58-96 */
  double a = g->tubedist, rho = sqrt(rho2), lam = sqrt(g->l2);
  double ex = exp(-rho / lam), flraf = floor(rho / a), ceilraf = ceil(rho / a), f;
  if (rho < a) {
    double eal = 0.0;
    if ((a - rho) / lam < 10.0) eal = exp(-(rho - a) / lam);
    f = ex / (a * (1.0 + ex) * (1.0 + ex));
    if (rho > 0.0) f += eal / (rho * (1.0 + eal) * (1.0 + eal));
  } else {
    double fm = fmod(rho, a);
    /* abs() of a double in the reference truncates to int first */
    if (fabs(trunc(fm)) < 1.0e-6) {
      f = flraf / (4.0 * rho);
    } else {
      double e1 = exp(-(rho - flraf * a) / lam), e2 = exp(-(rho - ceilraf * a) / lam);
      f = flraf * e1 / (rho * (1.0 + e1) * (1.0 + e1)) + ceilraf * e2 / (rho * (1.0 + e2) * (1.0 + e2));
    }
  }
  return a / (2.0 * log(2.0) * g->l2) * f;
}

static double orc_fixfluxB(const orc_geom *g, double rho2) {   /* Integrand.c:48-56 */
  float l = (float)sqrt(g->l2);
  float lml = (float)(((double)l - g->lmin) / (g->tubedist - g->lmin));
  double x = 2.0 * sqrt(rho2 / g->l2);
  double bd = (fabs(x) < 1.0) ? exp(-1.0 / (1.0 - x * x)) : 0.0;
  return 2.0 / (g->l2 * g->q2) * (1.0 - 0.75 * (double)lml) * bd + 3.0 / (g->tubedist * g->tubedist) * (double)lml;
}

double orc_magfield(const orc_geom *g, double rho2) {
  double B = 0.0, l2 = g->l2;
  switch (g->profile) {
    case ORC_STEP: B = rho2 < l2 ? 1.0 / l2 : 0.0; break;
    case ORC_SMOOTH: B = l2 / ((l2 + rho2) * (l2 + rho2)); break;
    case ORC_QUADRATIC: B = rho2 < l2 ? 2.0 / l2 * (1.0 - rho2 / l2) : 0.0; break;
    case ORC_GAUSSIAN: B = rho2 < 100.0 * l2 ? 1.0 / (l2 * exp(rho2 / l2)) : 0.0; break;
    case ORC_PERIODIC: B = orc_periodicB(g, rho2); break;
    case ORC_SPLINE: {
      double s = sin(3.14159265358979323846 * sqrt(rho2) / g->tubedist);
      B = 1.0 / (l2 * exp(s * s / l2));
    } break;
    case ORC_FIXFLUX:
    case ORC_FIXFLUX_FF: B = orc_fixfluxB(g, rho2); break;
  }
  return 2.0 * g->F / ORC_E * B;
}

double orc_smallT(const orc_geom *g, double rho, double rho2) {   /* Integrand.c:219-248 */
  double l2 = g->l2, a = g->tubedist, pi = 3.14159265358979323846;
  (void)rho2;
  switch (g->profile) {
    case ORC_STEP: return fabs(trunc(rho * rho - l2));
    case ORC_SMOOTH:
    case ORC_QUADRATIC:
    case ORC_PERIODIC: return 0.5 * (l2 + rho * rho);
    case ORC_GAUSSIAN: return l2;
    case ORC_SPLINE: {
      double s = sin(2.0 * pi * rho / a);
      return rho / a > 1.0e-8 ? a * a * l2 / (pi * pi * s * s) : 1.0e12;
    }
    case ORC_FIXFLUX:
    case ORC_FIXFLUX_FF:
      if (rho * rho < l2 / 4.0) return 0.25 * l2;
      return 0.25 * l2 + (rho - sqrt(l2) / 2.0) * (rho - sqrt(l2) / 2.0);
  }
  return l2;
}

/* Field at x_cm (eB, Integrand.c:98-140 per tube).  Lattice geometries (an
 * extension: the reference has isolated tubes only) sum the tubes of the
 * image window that the Wilson loop sees, (2M+1)^2 tubes centred on the
 * nearest one (window centre floor(c/l + 1/2)); *rho_near is the distance to
 * that nearest tube (the radius the branch policy uses). */
double orc_field_at(const orc_geom *g, double cx, double cy, double *rho_near) {
  if (!(g->lattice > 0.0)) {
    *rho_near = sqrt(cx * cx + cy * cy);
    return orc_magfield(g, cx * cx + cy * cy);
  }
  const double L = g->lattice, m0 = floor(cx / L + 0.5), n0 = floor(cy / L + 0.5);
  double B = 0.0;
  for (int ia = -g->images; ia <= g->images; ++ia)
    for (int ib = -g->images; ib <= g->images; ++ib) {
      const double dx = cx - (m0 + ia) * L, dy = cy - (n0 + ib) * L;
      B += orc_magfield(g, dx * dx + dy * dy);
    }
  const double dx = cx - m0 * L, dy = cy - n0 * L;
  *rho_near = sqrt(dx * dx + dy * dy);
  return B;
}

/* Tfunc's integrand (Integrand.c:252-305) at a given field B and radius rho. */
double orc_igrand_B(const orc_geom *g, int fermion, double T, double B, double rho, double EV, int *branch) {
  double TB = ORC_E * T * B, sT = orc_smallT(g, rho, rho * rho);
  if (T < 0.1 && T < 0.01 * sT) { *branch = 0; return fermion == 1 ? orc_smTferm(B, T) : orc_smTscal(B, T); }
  if (T < 0.01 * sT) { *branch = 1; return TB > 1.0e-6 ? orc_exact(T, B, fermion) : 0.0; }
  if (T > 49.0) {
    *branch = 2;
    if (fermion == 1) return -1.0 * exp(-ORC_M2 * T) / (3.0 * T) * ORC_E * ORC_E * B * B;
    return exp(-ORC_M2 * T) / (6.0 * T) * ORC_E * ORC_E * B * B;
  }
  *branch = 3;
  if (fermion == 1) return exp(-ORC_M2 * T) / (T * T * T) * (EV - 1.0 - 1.0 / 3.0 * TB * TB);
  return exp(-ORC_M2 * T) / (T * T * T) * (EV - 1.0 + 1.0 / 6.0 * TB * TB);
}

double orc_igrand(const orc_geom *g, int fermion, double T, double rho, double rho2,
                  double EV, int *branch) {
  double B = orc_magfield(g, rho2), TB = ORC_E * T * B, sT = orc_smallT(g, rho, rho2);
  if (T < 0.1 && T < 0.01 * sT) { *branch = 0; return fermion == 1 ? orc_smTferm(B, T) : orc_smTscal(B, T); }
  if (T < 0.01 * sT) { *branch = 1; return TB > 1.0e-6 ? orc_exact(T, B, fermion) : 0.0; }
  if (T > 49.0) {
    *branch = 2;
    if (fermion == 1) return -1.0 * exp(-ORC_M2 * T) / (3.0 * T) * ORC_E * ORC_E * B * B;
    return exp(-ORC_M2 * T) / (6.0 * T) * ORC_E * ORC_E * B * B;
  }
  *branch = 3;
  if (fermion == 1) return exp(-ORC_M2 * T) / (T * T * T) * (EV - 1.0 - 1.0 / 3.0 * TB * TB);
  return exp(-ORC_M2 * T) / (T * T * T) * (EV - 1.0 + 1.0 / 6.0 * TB * TB);
}

/* ------------------------------------------------------------------ */
/* Clamped cubic spline (Burden & Faires Alg. 3.5, zero end slopes)     */
/* ------------------------------------------------------------------ */
void orc_spline(const double *fli, const float *x, int n, int quirk, float *c4, double *cd) {
  double *h = (double *)malloc(sizeof(double) * n), *l = (double *)malloc(sizeof(double) * n);
  double *mu = (double *)malloc(sizeof(double) * n), *z = (double *)malloc(sizeof(double) * n);
  double *al = (double *)malloc(sizeof(double) * n), *c = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n - 1; ++i) h[i] = quirk ? (double)(x[i + 1] - x[i]) : (double)x[i + 1] - (double)x[i];
  al[0] = 3.0 * (fli[1] - fli[0]) / h[0];
  al[n - 1] = -3.0 * (fli[n - 1] - fli[n - 2]) / h[n - 2];
  for (int i = 1; i < n - 1; ++i) al[i] = 3.0 / h[i] * (fli[i + 1] - fli[i]) - 3.0 / h[i - 1] * (fli[i] - fli[i - 1]);
  l[0] = 2.0 * h[0]; mu[0] = 0.5; z[0] = al[0] / l[0];
  for (int i = 1; i < n - 1; ++i) {
    double span = quirk ? (double)(x[i + 1] - x[i - 1]) : (double)x[i + 1] - (double)x[i - 1];
    l[i] = 2.0 * span - h[i - 1] * mu[i - 1];
    mu[i] = h[i] / l[i];
    z[i] = (al[i] - h[i - 1] * z[i - 1]) / l[i];
  }
  l[n - 1] = h[n - 2] * (2.0 - mu[n - 2]);
  z[n - 1] = (al[n - 1] - h[n - 2] * z[n - 2]) / l[n - 1];
  c[n - 1] = quirk ? (double)(float)z[n - 1] : z[n - 1];
  for (int i = 0; i < 4 * n; ++i) { if (c4) c4[i] = 0.0f; if (cd) cd[i] = 0.0; }
  for (int i = n - 2; i >= 0; --i) {
    double ci = z[i] - mu[i] * c[i + 1];
    if (quirk) ci = (double)(float)ci;
    c[i] = ci;
    double bi = (fli[i + 1] - fli[i]) / h[i] - h[i] * (c[i + 1] + 2.0 * c[i]) / 3.0;
    double di = quirk ? (double)((float)c[i + 1] - (float)c[i]) / (3.0 * h[i]) : (c[i + 1] - c[i]) / (3.0 * h[i]);
    if (c4) { c4[4 * i] = (float)fli[i]; c4[4 * i + 1] = (float)bi; c4[4 * i + 2] = (float)c[i]; c4[4 * i + 3] = (float)di; }
    if (cd) { cd[4 * i] = fli[i]; cd[4 * i + 1] = bi; cd[4 * i + 2] = c[i]; cd[4 * i + 3] = di; }
  }
  if (c4) { c4[4 * (n - 1)] = (float)fli[n - 1]; c4[4 * (n - 1) + 2] = (float)c[n - 1]; }
  if (cd) { cd[4 * (n - 1)] = fli[n - 1]; cd[4 * (n - 1) + 2] = c[n - 1]; }
  free(h); free(l); free(mu); free(z); free(al); free(c);
}

double orc_logsimpson(const double *T, const double *f, int n) {
  if (n < 2) return 0.0;
  double u0 = log(T[0]), u1 = log(T[n - 1]);
  double h = (u1 - u0) / (double)(n - 1);
  double *g = (double *)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) g[i] = T[i] * f[i];
  double s = 0.0;
  int m = n - 1;                 /* intervals */
  if (m == 1) { s = 0.5 * h * (g[0] + g[1]); free(g); return s; }
  int ms = (m % 2 == 0) ? m : m - 3;
  for (int i = 0; i + 2 <= ms; i += 2) s += h / 3.0 * (g[i] + 4.0 * g[i + 1] + g[i + 2]);
  if (ms != m) s += 3.0 * h / 8.0 * (g[ms] + 3.0 * g[ms + 1] + 3.0 * g[ms + 2] + g[ms + 3]);
  free(g);
  return s;
}
