/*
 * wfl.h -- C ABI of the MI355X worldline Wilson-loop engine (libwfl.so).
 *
 * Plain C: pointers, sizes and status codes only.  Device buffers are
 * ordinary HIP device pointers passed as void*; streams are hipStream_t
 * passed as void* (NULL = the default stream).
 *
 * Two entry levels:
 *   1. The reference's own call, `double EV(double T, void *p, int *WLlist)`
 *      (declared in intT.h, which reconstructs the reference's missing
 *      header), together with EVMean, Tfunc, integrateT and the profile
 *      closed forms.
 *   2. The batched engine below: many (geometry, x_cm, T) tuples over one
 *      device-resident loop ensemble in one launch sequence, reduced on the
 *      device to FP64 per-group sums.
 *
 * Status codes: WFL_OK (0) on success, negative on error; wfl_last_error()
 * returns a message for the calling thread's most recent failure.
 */
#ifndef WFL_H
#define WFL_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define WFL_OK 0
#define WFL_ERR_ARG (-1)      /* invalid argument or shape */
#define WFL_ERR_DEVICE (-2)   /* HIP runtime error (no device, launch failure) */
#define WFL_ERR_NOMEM (-3)    /* host or device allocation failed */
#define WFL_ERR_STATE (-4)    /* call out of order (e.g. run before create) */

/* Field profiles: the cases of flambda/fplambda (intEV.cu:261-406).
 * WFL_FIXFLUX_FF is the intEVff.cu variant (phi threshold 1, chi fit B). */
enum { WFL_STEP = 0, WFL_SMOOTH = 1, WFL_QUADRATIC = 2, WFL_GAUSSIAN = 3,
       WFL_PERIODIC = 4, WFL_SPLINE = 5, WFL_FIXFLUX = 6, WFL_FIXFLUX_FF = 7 };

/* Numerical definition of the segment integral.
 * REFERENCE: Idt exactly as intEV.cu:409-459 defines it, including the clamps
 *            of A, B, C and rho^2 (the linear coefficient B = z.(z'-z) is
 *            clamped to >= 1e-8) and both endpoints evaluated per segment.
 * PHYSICAL:  the same composite Simpson rule along the true path
 *            rho^2(t) = |z + t(z'-z)|^2 with no clamps; the removable
 *            singularity of f(u)/u at u = 0 is taken analytically and each
 *            point's endpoint value is shared by its two segments. */
enum { WFL_MODE_REFERENCE = 0, WFL_MODE_PHYSICAL = 1 };

/* Arithmetic of the segment phases and Simpson sums.
 * FP64:  everything in double.
 * MIXED: segment geometry and profile evaluations in FP32, per-loop sums
 *        carried in FP64 blocks, W-1 and the ensemble sums in FP64. */
enum { WFL_FP64 = 0, WFL_MIXED = 1 };

/* One flux-tube geometry (a field profile, optionally repeated on a square
 * lattice).  F and l2 are quantized to FP32 by the engine so that both
 * precisions evaluate the same problem (the reference narrows them to
 * float at intEV.cu:124). */
typedef struct {
  int kind;             /* WFL_* profile */
  double F;             /* flux parameter (Integrand.c:137: eB = 2 F f'(rho^2)) */
  double l2;            /* lambda^2 */
  double a;             /* tubedist: ring/lattice scale of periodic, spline, fixflux */
  double lam_min;       /* lmin (fixflux) */
  double fix_q1, fix_q2;/* q1, q2 (fixflux normalisations) */
  double lattice;       /* square-lattice spacing of image tubes; <= 0: isolated */
  int images;           /* image window half-width M: (2M+1)^2 tubes around x_cm's cell */
  int spline_table;     /* index of the spline table (WFL_SPLINE), else -1 */
} wfl_geometry;

/* A clamped cubic spline of f(rho^2) on sorted knots (periodic.c:45-110
 * layout: knots as float, coefficients as float4 (a, b, c, d)). */
typedef struct {
  int n;
  const float *rho2;    /* host [n] */
  const float *coefs;   /* host [n][4] */
} wfl_spline_table;

/* Batched workload description (host arrays; copied by wfl_plan_create). */
typedef struct {
  int mode;             /* WFL_MODE_* */
  int precision;        /* WFL_FP64 or WFL_MIXED */
  int nsub;             /* Simpson sub-intervals per segment (even, >= 2) */
  int spinor;           /* 0: scalar QED; 1: spinor QED (spin factor) */
  int group;            /* loops per group (the reference uses 128) */
  int ngeom;
  const wfl_geometry *geoms;
  int nspline;
  const wfl_spline_table *splines;
  int ntup;             /* evaluation tuples */
  const int *tup_geom;  /* [ntup] geometry index */
  const double *tup_cx; /* [ntup] x_cm components */
  const double *tup_cy;
  const double *tup_T;  /* [ntup] proper time T (quantized: sqrt(T) to FP32) */
  const int *geom_nloops; /* [ngeom] loops used per geometry (NULL: all Nl) */
  int want_perloop;     /* also keep per-loop (scal[3], ferm[3]) FP64 outputs */
  int mapping;          /* 0: chosen at creation from the batch's size (small
                           batches split loop groups into several tasks;
                           wfl_plan_autotune measures the alternatives);
                           1: lane per tuple; 2: wave per loop;
                           3: lane per tuple with W-1 in a second kernel */
} wfl_batch;

typedef struct wfl_loops_s wfl_loops;   /* device-resident loop ensemble */
typedef struct wfl_plan_s wfl_plan;     /* uploaded batched workload */

/* ---- device and errors ---- */
const char *wfl_last_error(void);
int wfl_device_count(int *count);
int wfl_device_sync(void);

/* ---- loop ensembles ----
 * Host SoA, loop-major: x[l*N + i], y[l*N + i]; N must be a multiple of 4. */
int wfl_loops_create(wfl_loops **out, const float *x, const float *y, int Nl, int N);
/* From the reference's device layout: float4 AoS [Nl*N] (getwl.c:18-35). */
int wfl_loops_from_aos_device(wfl_loops **out, const void *dev_float4, int Nl, int N);
/* Seeded unit loops generated on the host (Philox4x32-10, key
 * 0x5EED0000 + config_id) and uploaded; bit-identical to wfl_gen_loops. */
int wfl_loops_generate(wfl_loops **out, uint32_t config_id, int Nl, int N);
int wfl_loops_shape(const wfl_loops *L, int *Nl, int *N);
void wfl_loops_destroy(wfl_loops *L);

/* Host generator used by wfl_loops_generate (x, y: [Nl*N]). */
int wfl_gen_loops(uint32_t config_id, int Nl, int N, float *x, float *y);
/* Text reader for the reference's worldline files (getwl.c:6-40): per loop
 * N x values then N y values.  Writes SoA x, y. */
int wfl_read_text_loops(const char *path, int Nl, int N, float *x, float *y);
/* Writes the same text layout (getwl.c:18-28), 9 significant digits, so
 * that wfl_read_text_loops and getwl read the floats back exactly. */
int wfl_write_text_loops(const char *path, const float *x, const float *y, int Nl, int N);
/* Binary loop files (SoA float32 with a header: magic "WFLLOOP1", version,
 * Nl, N, generator config id, FNV-1a 64 hash of the data).  read verifies
 * the shape and the hash (WFL_ERR_ARG on a mismatch). */
int wfl_loops_write(const char *path, const float *x, const float *y, int Nl, int N, uint32_t config_id);
int wfl_loops_read_header(const char *path, int *Nl, int *N, uint32_t *config_id);
int wfl_loops_read(const char *path, float *x, float *y, int Nl, int N);

/* ---- batched engine ---- */
int wfl_plan_create(wfl_plan **out, const wfl_batch *b, wfl_loops *L);
/* Enqueue the whole evaluation on `stream` (asynchronous). */
int wfl_plan_run(wfl_plan *p, void *stream);
/* wfl_plan_run, then wait for it; *kernel_ms is the main Wilson-loop
 * kernel's duration measured with HIP events on `stream`. */
int wfl_plan_run_timed(wfl_plan *p, void *stream, float *kernel_ms);
/* Measured autotuning of the launch configuration (mapping, static or
 * persistent grid, wavefronts per block, tasks per loop group); runs the plan
 * once per candidate and keeps the fastest.  *chosen (may be NULL) =
 * 1000*mapping + 100*persistent + 10*wavefronts per block + tasks per loop
 * group.  Outputs afterwards hold a valid run. */
int wfl_plan_autotune(wfl_plan *p, void *stream, int *chosen);
/* Current launch configuration, and setting one explicitly: mapping 1..3,
 * persistent 0/1, wavefronts per block 1, 2 or 4, and `chunks` tasks per loop
 * group (1..min(64, group); lane-per-tuple mappings: each task sums
 * ceil(group/chunks) loops and the chunk sums are added in order, so small
 * batches still fill the device). */
int wfl_plan_config(const wfl_plan *p, int *mapping, int *persistent, int *wpb, int *chunks);
int wfl_plan_set_config(wfl_plan *p, int mapping, int persistent, int wpb, int chunks);
/* Capture one wfl_plan_run into a hipGraph and replay it (T-sweep reuse). */
int wfl_plan_graph_run(wfl_plan *p, void *stream);
/* Copy results to the host (synchronizes): gsum [ntup][ngroups][3] sums of
 * (W_N - 1, W_{N/2} - 1, W_{N/4} - 1); perloop [ntup][Nl][6] if kept. */
int wfl_plan_read(wfl_plan *p, double *gsum, double *perloop);
int wfl_plan_ngroups(const wfl_plan *p);
/* Tuples routed to the accurate FP32 Gaussian node (mixed plans): their loops
 * are small (sigma sqrt(T) < 1/2, sigma^2 = log2(e)/l2) and can reach the tube
 * core, where the plain quotient (1 - 2^-u)/u loses relative accuracy. */
int wfl_plan_accurate_tuples(const wfl_plan *p);
/* Nonzero when a mixed, physical-mode, spline-only plan evaluates its tables
 * by direct indexing: every table's knots lie on a uniform grid in
 * r = sqrt(rho^2) (the layout of periodic.c:129-151), checked to 1e-4 of
 * the spacing; 0 when the kernels use the exact interval search.  2 when the
 * rows hold their cubics expanded about u = 0 (one FMA per Horner step and
 * no expansion point per node), taken when every table is well conditioned
 * in that form (see wfl_spline_uniform_form), 1 when each row is expanded
 * about its own start. */
int wfl_plan_uniform_splines(const wfl_plan *p);
/* The form wfl_plan_uniform_splines would report for a plan over one table
 * (n knots rho2, coefficients as wfl_spline_build writes them) whose nodes
 * reach u = rho^2 <= umax, scalar (spinor = 0) or spinor: 0, 1 or 2.  The
 * rows are expanded about u = 0 when, at every point of every row,
 * sum_k |d_k| u^k <= 16 |P(u)| for the row's cubic P(u) = sum_k d_k u^k,
 * and, for spinor plans, sum_k k |d_k| u^(k-1) <= 16 |P'(u)| (the spin
 * factor's f'; Gaussian tubes fail it where f' is small, far out).  With
 * rows != NULL and a nonzero form, also writes
 * up to max_rows rows of 4 floats as the kernels read them: the guard row
 * (for sqrt(u) < h/4), then row i for sqrt(u) in [i h, (i + 1) h), with
 * h = (sqrt(rho2[n-1]) - sqrt(rho2[0])) / (n - 1); *nrows = rows written.
 * Form 1 rows hold the cubic in u - fl32(i h)^2 (guard row: u - (h/2)^2),
 * form 2 rows the cubic in u.  Host only: makes no HIP call.  No
 * counterpart in the reference (periodic.c:129-151 defines the table,
 * intEV.cu:64-107 evaluates it). */
int wfl_spline_uniform_form(const float *rho2, const float *coefs, int n, double umax, int spinor, float *rows,
                            int max_rows, int *nrows);
/* 1 when the lane-per-tuple kernels of a mixed plan use the axis form:
 * physical mode, Simpson n = 2, no lattice geometry and every tuple's x_cm
 * on the x axis (cy = 0), so that only z_x = c + s w_x is formed per
 * point and rho^2 = z_x^2 + s^2 w_y^2, z x v = z_x v_y - s (w_y v_x) take
 * their y parts from loop-only point records; the same nodes and sums as
 * intEV.cu:509-571 (the reference forms z and its products per point). */
int wfl_plan_axis_form(const wfl_plan *p);
/* Tuples of a Gaussian axis-form plan that take the local Taylor node:
 * loops small against their distance from the tube centre (every node
 * within 0.5 of uc = c'^2 in the scaled units u' = log2(e) rho^2 / l2, for
 * loops up to the 90th percentile of the ensemble's loop radii; the kernel
 * checks each loop) evaluate (1 - 2^-u')/u' as g(uc) + sum_k a_k (u' - uc)^k,
 * k = 1..7, with the g(uc) part of every sub-loop sum taken exactly from
 * per-loop FP64 area sums.  A mixed-precision accuracy measure with no
 * counterpart in the reference (intEV.cu evaluates the node directly).
 * The lane-per-tuple mappings (1, 3) carry it; mapping 0 and the autotuner
 * keep to them when a plan has such tuples, and an explicit mapping 2
 * evaluates those tuples with the direct node. */
int wfl_plan_local_tuples(const wfl_plan *p);
/* The local node's coefficients a[k] = g^(k)(uc)/k!, k = 0..7, of
 * g(u) = (1 - 2^-u)/u, in FP64. */
int wfl_local_taylor(double uc, double *a);
/* Simpson sub-steps chosen by measured accuracy (the batch's own nsub is
 * ignored): the smallest n in 2, 4, 8, ... <= nmax whose group sums on a
 * sample of the batch (every k-th tuple, at most max_tuples; the first
 * max_loops loops, one group) change by at most tol, relative to the
 * sample's largest |sum of W - 1|, when n is doubled.  *est_err = that
 * relative change (about the n-point error, Simpson converging as n^-4);
 * nmax when none passes.  Runs a few small plans; precision and mode are
 * the batch's. */
int wfl_autotune_nsub(const wfl_batch *b, wfl_loops *L, double tol, int nmax, int max_tuples, int max_loops,
                      int *nsub, double *est_err);
/* Device pointer of the group sums (for callers that keep results on device). */
const double *wfl_plan_gsum_device(const wfl_plan *p);
/* Number of non-finite per-loop values met in the last run (diagnostic). */
int wfl_plan_nonfinite(wfl_plan *p, long long *count);
/* Mapping the plan uses (1, 2 or 3 as in wfl_batch.mapping). */
int wfl_plan_mapping(const wfl_plan *p);
void wfl_plan_destroy(wfl_plan *p);
/* create + run + read + destroy. */
int wfl_eval_batch(const wfl_batch *b, wfl_loops *L, double *gsum, double *perloop);
/* wfl_eval_batch without device allocation.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (serving loops,
 * harnesses that forbid it).  The plan's device buffers live in one static
 * 64 MiB device array per GPU, held until the process ends; a batch that
 * needs more fails with WFL_ERR_ARG and a message naming the sizes.  The
 * loop set's device records must already exist: the first wfl_eval_batch or
 * wfl_plan_create of a batch of the same precision and form on L builds them
 * (otherwise WFL_ERR_STATE).  The default launch configuration is used (no
 * autotuning); results equal wfl_eval_batch's for batches whose chunk sums
 * take at most half the array.  Calls on one GPU share the array and are
 * serialised by a lock; the call returns when gsum (and perloop) are
 * written.  Host-side bookkeeping uses the C++ heap. */
int wfl_eval_batch_noalloc(const wfl_batch *b, wfl_loops *L, double *gsum, double *perloop);

/* Richardson-extrapolated <W> - 1 of one group from its sums
 * (Integrand.c:400): 8/3 m_N - 2 m_{N/2} + 1/3 m_{N/4}. */
double wfl_group_ev_minus1(const double gsum3[3], int count);

/* ---- spline set-up (periodic.c:45-110), FP64 ----
 * quirk = 1 reproduces the reference's float round trip of c. */
int wfl_spline_build(const double *fli, const float *rho2, int n, int quirk, float *coefs4);

/* ---- drop-in configuration for EV() (what intT.h fixes at compile time) ---- */
int wfl_set_profile(const wfl_geometry *g, int nspline);
int wfl_set_numerics(int mode, int precision, int nsub);
/* Forget the cached per-loop results and loop copies held for EV(). */
void wfl_ev_reset(void);
/* Kernel launches EV() made since the last reset (cache diagnostics). */
long long wfl_ev_launches(void);

/* ---- proper-time integration on a log-spaced grid ----
 * Composite Simpson in ln T (3/8 rule on the last three intervals when the
 * interval count is odd). */
double wfl_logsimpson(const double *T, const double *f, int n);
/* Effective parameters a batch evaluates for geometry g and a tuple
 * (cx, cy, T): out = {l2, F, cx, cy, sqrt(T), lattice}.  Every kernel
 * parameter is narrowed to FP32 (as the reference does at intEV.cu:124); for
 * the Gaussian fast path the narrowing happens in the FP32 kernel's scaled
 * units (sigma = sqrt(log2 e / l2) in FP32), so both precisions evaluate
 * exactly these doubles. */
int wfl_effective_tuple(const wfl_geometry *g, double cx, double cy, double T, double out[6]);
/* The effective proper time of a requested T for geometry g. */
double wfl_effective_T(const wfl_geometry *g, double T);
/* L_eff(rho) per loop group from batched sums on a log T grid.
 * gsum [nT][ngroups][3] holds the group sums of the nT tuples of one
 * (geometry, x_cm) in T order, counts [ngroups] the loops per group, T the
 * effective proper times.  policy 1 applies the reference's analytic small-
 * and large-T branches (Integrand.c:252-288), policy 0 uses the worldline
 * estimate at every node.  Normalised as Tint.c:69-72: /(-2 pi) scalar,
 * /(4 pi) spinor. */
int wfl_leff_groups(const wfl_geometry *g, int fermion, int nT, const double *T, double cx, double cy,
                    const double *gsum, const int *counts, int ngroups, int policy, double *leff);
/* The same with a choice of the N -> infinity estimate of <W> - 1 per
 * group: estimator 0 the reference's Richardson combination
 * 8/3 W_N - 2 W_{N/2} + 1/3 W_{N/4} (Integrand.c:400, as wfl_leff_groups);
 * 1 the two-level extrapolation 2 W_{N/2} - W_{N/4}.  The reference's W_N
 * is cos(x/2 + (y + z)/4), a blend of the sub-loop phases, whose
 * discretization error is not half of W_{N/2}'s, so its weights leave a
 * 1/N term (about +6.5 % in a constant-field scalar L_eff at N = 1024);
 * estimator 1 cancels it.  (An extension: no counterpart in the reference.) */
int wfl_leff_groups_est(const wfl_geometry *g, int fermion, int nT, const double *T, double cx, double cy,
                        const double *gsum, const int *counts, int ngroups, int policy, int estimator,
                        double *leff);
/* <W> - 1 of one group by estimator (0 reference, 1 two-level). */
double wfl_group_ev_minus1_est(const double gsum3[3], int count, int estimator);

/* Effective action of an isolated analytic tube with the batched engine:
 * EAction's rho quadrature (EffAct.c:265-306, integrateRho with Npoints
 * intervals, or integrateRhotoa with 2 Npoints for periodic/fixflux) over
 * integrateT's nT-interval rule in t (Tint.c:100-175), Tfunc's branch policy
 * at every node; the worldline nodes of the whole (rho, T) grid run as one
 * plan.  groups_out [ceil(Nl/group)] (may be NULL) gets each group's action,
 * *result their mean and *se its standard error (the reference's EAvsLambda:
 * and ActRatio: lines are printed), *classical ClassAction (EffAct.c:231-263).
 * Not for spline or lattice geometries (WFL_ERR_ARG). */
int wfl_eaction(const wfl_geometry *g, wfl_loops *L, int fermion, int mode, int precision, int nsub, int group,
                int Npoints, int nT, double *groups_out, double *result, double *se, double *classical);
/* Output level of the drivers (the reference's verbosity; default 1): at 2
 * and above Tfunc prints its branch traces and the WLvconstExpression:
 * line (rho^2, T, Exact, smTscal, mean and spread of the integrand over the
 * groups; Integrand.c:255-313), at 3 the EV= line. */
int wfl_set_verbosity(int v);

/* Floating-point exception checks in EVMean (Integrand.c:318-342, where
 * they are always on and process-wide): 0 off (default); 1 report (the
 * four flags divide-by-zero, invalid, overflow, underflow are cleared on
 * entry and tested on exit; a raised flag prints the reference's message
 * and is counted); 2 strict (the reference's traps and SIGFPE handler:
 * message and exit status 1). */
int wfl_set_fp_traps(int mode);
long long wfl_fp_trap_count(void);

/* ---- lattice physics outputs (analysis-scripts/interaction.sage:10-25) ----
 * Per-group action of one tube of a lattice geometry: sum_c w_c L_eff(x_c)
 * over ncell x_cm points of the unit cell (midpoint rule on an n x n grid:
 * w_c = l^2/n^2).  gsum [ncell][nT][ngroups][3] holds the group sums of the
 * points' tuples in T order (wfl_leff_groups for each point). */
int wfl_cell_action_groups(const wfl_geometry *g, int fermion, int nT, const double *T, int ncell, const double *cx,
                           const double *cy, const double *w, const double *gsum, const int *counts, int ngroups,
                           int policy, double *action);
/* Per-group action of an isolated tube: 2 pi int rho L_eff(rho) drho over
 * nrho uniformly spaced rho >= 0 (composite Simpson, 3/8 rule on the last
 * three intervals for an odd interval count).  gsum [nrho][nT][ngroups][3]. */
int wfl_radial_action_groups(const wfl_geometry *g, int fermion, int nT, const double *T, int nrho, const double *rho,
                             const double *gsum, const int *counts, int ngroups, int policy, double *action);
/* Interaction energy per tube from per-group actions of the lattice tube
 * and of the isolated tube on the same loop groups, as interaction.sage
 * forms it: E = -action, out[0] = mean of E_cell - E_iso, out[1] = the
 * script's error sqrt(var E_cell + var E_iso - 2 cov) (the spread of one
 * group's difference), out[2] = out[1]/sqrt(ngroups) (standard error). */
int wfl_interaction_energy(int ngroups, const double *cell_action, const double *iso_action, double *out);

#ifdef __cplusplus
}
#endif
#endif
