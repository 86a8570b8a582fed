/*
 * wfl_oracle.h -- FP64 CPU oracle for the worldline Wilson-loop hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it, as the checker.
 *
 * It restates, in plain C and double precision, the reference's algorithm:
 *   getzp1            intEV.cu:462-506   (sub-loop successor map)
 *   Idt               intEV.cu:409-459   (Simpson rule per segment, clamps)
 *   WilsonLoop        intEV.cu:509-571   (per-loop sub-loop sums)
 *   flambda/fplambda  intEV.cu:261-297, 361-406 and the profile helpers
 *   interp            intEV.cu:64-107
 *   EVMean            Integrand.c:325-407
 *   Tfunc policy      Integrand.c:20-37, 98-172, 195-316
 *   integrateT        Tint.c:15-33, 100-175
 *   flspline          periodic.c:45-110 (clamped cubic spline)
 * plus the pieces the reference does not ship (seeded loop generator,
 * lattice image sums), written from their mathematical definition.
 *
 * Parity pins (which test checks which function) are listed at the top of
 * wfl_oracle.c.
 */
#ifndef WFL_ORACLE_H
#define WFL_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_STEP = 0, ORC_SMOOTH = 1, ORC_QUADRATIC = 2, ORC_GAUSSIAN = 3,
       ORC_PERIODIC = 4, ORC_SPLINE = 5, ORC_FIXFLUX = 6, ORC_FIXFLUX_FF = 7 };
enum { ORC_MODE_REFERENCE = 0, ORC_MODE_PHYSICAL = 1 };

typedef struct {
  int profile;
  double l2, F;             /* lambda^2 and flux F (as the kernels receive them) */
  double tubedist, lmin, q1, q2;
  int nspline;              /* spline profile: knots (rho^2, float) and coefs */
  const float *rho2sp;      /* [nspline] */
  const float *coefs;       /* [nspline][4] = (a, b, c, d) */
  double lattice;           /* lattice spacing; <= 0 means an isolated tube */
  int images;               /* image window half-width M: (2M+1)^2 tubes */
} orc_geom;

/* Philox4x32-10 (Salmon et al., SC'11). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Unit loops, SoA loop-major x[Nl*N], y[Nl*N]. */
void orc_gen_loops(uint32_t config_id, int Nl, int N, float *x, float *y);

/* Successor of point i in its sub-loop (getzp1, intEV.cu:462-506). */
int orc_succ(int i, int N);
double orc_f(const orc_geom *g, double rho2);
double orc_fp(const orc_geom *g, double rho2);
double orc_expint(double x);
double orc_interp(double rho2, const float *rho2sp, const float *coefs, int n);
/* slope of the physical-mode spline line below the first knot */
double orc_spline_low_slope(const float *rho2sp, const float *coefs);
double orc_f_in_mode(const orc_geom *g, int mode, double u);

/* Per-loop sub-loop sums: scal[3] = F*(x, y, z) and ferm[3] as WilsonLoop
 * writes them (intEV.cu:553-560); lattice images summed. */
void orc_wilson_loop(const orc_geom *g, int mode, int nsub, int spinor,
                     const float *wx, const float *wy, int N,
                     double cx, double cy, double rtT,
                     double scal[3], double ferm[3]);

/* Batched evaluation: for each tuple t (geometry index, cx, cy, rtT) and loop
 * l in [0, Nl): per-loop outputs (optional, [ntup][Nl][6]) and per-group
 * sums of (W_N - 1, W_{N/2} - 1, W_{N/4} - 1) ([ntup][ngroups][3]). */
void orc_eval_batch(const orc_geom *geoms, int mode, int nsub, int spinor,
                    const float *x, const float *y, int Nl, int N, int group,
                    int ntup, const int *tup_geom, const double *tup_cx,
                    const double *tup_cy, const double *tup_rtT,
                    double *perloop, double *gsum, int nthreads);

/* W - 1 for the three Richardson estimates of one loop (Integrand.c:352-390). */
void orc_wminus1(int mode, int spinor, const double scal[3], const double ferm[3],
                 double wm1[3]);

/* EVMean restated on the reference's float4 per-loop interface. */
double orc_evmean(const float *wsscal4, const float *wsferm4, int n, const int *WL,
                  int fermion);

/* Tfunc policy and closed forms. */
double orc_exact(double T, double B, int fermion);
double orc_smTscal(double B, double T);
double orc_smTferm(double B, double T);
double orc_magfield(const orc_geom *g, double rho2);
double orc_smallT(const orc_geom *g, double rho, double rho2);
/* One group's integrand at T given the worldline estimate EV (used only on
 * the worldline branch); returns the branch id through *branch. */
/* eB at x_cm summed over a lattice geometry's image window (isolated: the
 * tube's own field); *rho_near = distance to the nearest tube. */
double orc_field_at(const orc_geom *g, double cx, double cy, double *rho_near);
double orc_igrand_B(const orc_geom *g, int fermion, double T, double B, double rho, double EV, int *branch);
double orc_igrand(const orc_geom *g, int fermion, double T, double rho, double rho2,
                  double EV, int *branch);

/* Clamped cubic spline (zero end slopes) of fli over knots rho2 (FP64).
 * quirk=1 reproduces the float round trip of c (periodic.c:99-101). */
void orc_spline(const double *fli, const float *rho2, int n, int quirk, float *coefs4,
                double *coefs_d);

/* Composite Simpson in u = ln T on a log-spaced grid (odd count: Simpson;
 * even count: Simpson on the first n-3 intervals + 3/8 rule on the last 3). */
double orc_logsimpson(const double *T, const double *f, int n);

#ifdef __cplusplus
}
#endif
#endif
