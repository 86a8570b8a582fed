/* wfl_effact.c -- the rho-integral drivers above the Wilson-loop path.
 *
 *   rhogrand, testvalrho, transformRho      EffAct.c:15-81
 *   integrateRhotoa                         EffAct.c:83-135 (periodic tubes, 0..a/2)
 *   integrateRho                            EffAct.c:137-199 (isolated tubes, 0..inf)
 *   PeriodicCA, FixFluxCA, ClassAction      EffAct.c:201-263
 *   EAction                                 EffAct.c:265-306 (EAvsLambda:, ActRatio:)
 *
 * These keep the reference's names, signatures, quadrature nodes and output
 * lines, and drive EV() through integrateT like the reference (one launch
 * per distinct (x_cm, T)).  wfl_eaction below computes the same effective
 * action with the batched engine: every (rho, T) node whose integrand needs
 * a worldline estimate goes into one plan and one launch.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "../../include/intT.h"
#include "wfl_internal.h"

static const double WFL_E = 0.30282212;   /* Integrand.c:14 */
static const double WFL_PI = 3.14159265358979323846;

static int g_verbosity = 1;

int wfl_set_verbosity(int v) {
  g_verbosity = v;
  return WFL_OK;
}

int wfl_internal_verbosity(void) { return g_verbosity; }

/* ------------------------------------------------------------------ */
/* Drop-in drivers (reference names)                                    */
/* ------------------------------------------------------------------ */
double *rhogrand(double rho, void *p) {   /* EffAct.c:15-39 */
  struct Wparams *params = (struct Wparams *)p;
  params->xcm.x = (float)rho;
  params->xcm.y = 0.0f;
  params->xcm.z = 0.0f;
  if (g_verbosity >= 2) printf("Beginning T integration\n");
  double *intout = integrateT(40, p);
  if (g_verbosity >= 2) printf("Returning from T integration\n");
  if (!intout) return NULL;
  for (int i = 0; i < params->ng; ++i) intout[i] = rho * intout[i];
  return intout;
}

void testvalrho(double *val, int n, double t) {   /* EffAct.c:41-62 */
  double avg = 0.0, sem = 0.0;
  for (int i = 0; i < n; ++i) {
    if (isnan(val[i])) printf("for rho=%f val[%d]=nan\n", (1.0 - t) / t, i);
    avg += val[i];
  }
  avg /= n;
  for (int i = 0; i < n; ++i) sem += (val[i] - avg) * (val[i] - avg);
  sem = sqrt(sem / ((double)n * ((double)n - 1.0)));
  printf("testvalrho: %f %f %f\n", (1.0 - t) / t, avg, sem);
}

double *transformRho(double rhot, void *p) {   /* EffAct.c:64-81: rho = (1 - t)/t */
  struct Wparams *params = (struct Wparams *)p;
  double *f = rhogrand((1.0 - rhot) / rhot, p);
  if (!f) return NULL;
  for (int i = 0; i < params->ng; ++i) f[i] = (f[i] / rhot) / rhot;
  return f;
}

/* Weights of integrateRhotoa's sum over its n interior nodes rho_k = k h,
 * h = a/(2n), k = 1..n: h/3 times 2 for odd k and 4 for even k (the
 * reference's accumulators, EffAct.c:113-126), plus h/3 for each of the two
 * end values rho = 0 and rho = a/2 (a/2 is also the node k = n). */
static double rhotoa_weight(int k, int n) {
  (void)n;
  return (k % 2 == 1) ? 2.0 : 4.0;
}

double *integrateRhotoa(int n, void *p) {   /* EffAct.c:83-135 */
  struct Wparams *params = (struct Wparams *)p;
  const int ng = params->ng;
  const wfl_geometry *g = wfl_internal_profile();
  double *val = (double *)calloc((size_t)ng, sizeof(double));
  double *x0 = (double *)calloc((size_t)ng, sizeof(double));
  if (!val || !x0) { printf("integrate(): malloc failed\n"); free(val); free(x0); return NULL; }
  const double h = g->a / (double)(2 * n);
  double *rg = rhogrand(0.0, p), *rg2 = rhogrand(g->a / 2.0, p);
  if (!rg || !rg2) { free(rg); free(rg2); free(val); free(x0); return NULL; }
  for (int ig = 0; ig < ng; ++ig) x0[ig] = rg[ig] + rg2[ig];
  free(rg);
  free(rg2);
  for (int k = 1; k <= n; ++k) {
    printf("rho integrate inx=%d\n", k - 1);
    double *r = rhogrand(h * (double)k, p);
    if (!r) { free(val); free(x0); return NULL; }
    for (int ig = 0; ig < ng; ++ig) val[ig] += rhotoa_weight(k, n) * r[ig];
    free(r);
  }
  for (int ig = 0; ig < ng; ++ig) val[ig] = h / 3.0 * (x0[ig] + val[ig]);
  free(x0);
  return val;
}

double *integrateRho(int n, void *p) {   /* EffAct.c:137-199 */
  struct Wparams *params = (struct Wparams *)p;
  const int ng = params->ng;
  double *val = (double *)calloc((size_t)ng, sizeof(double));
  if (!val) { printf("integrate(): malloc failed\n"); return NULL; }
  const double h = 1.0 / (double)n;
  for (int inx = 0; inx < n; ++inx) {
    printf("rho integrate inx=%d\n", inx);
    /* the reference evaluates t = h(inx + 1) twice (as the left and the right
     * end): the rule equals composite Simpson because the integrand vanishes
     * at t = 0 (rho = inf) and t = 1 (rho = 0) */
    double *a = transformRho(h * ((double)inx + 1.0), p);
    double *b = transformRho(h * ((double)inx + 0.5), p);
    double *c = transformRho(h * ((double)inx + 1.0), p);
    if (!a || !b || !c) { free(a); free(b); free(c); free(val); return NULL; }
    for (int ig = 0; ig < ng; ++ig) {
      if (isinf(a[ig])) { printf("Warning: rhort[%d]=%f.  Setting to zero.\n", ig, a[ig]); a[ig] = 0.0; }
      if (isinf(b[ig])) { printf("Warning: rhortphalf[%d]=%f.  Setting to zero.\n", ig, b[ig]); b[ig] = 0.0; }
      if (isinf(c[ig])) { printf("Warning: rhortp1[%d]=%f.  Setting to zero.\n", ig, c[ig]); c[ig] = 0.0; }
      val[ig] += h / 6.0 * (a[ig] + 4.0 * b[ig] + c[ig]);
    }
    free(a);
    free(b);
    free(c);
  }
  return val;
}

static double periodic_ca(double a, double l2) {   /* EffAct.c:201-215 */
  const double l = sqrt(l2), ep = 1.0 + exp(a / l);
  return 2.0 * a + 4.0 * a / (ep * ep * ep) - l + 4.0 * l / ep - 2.0 * (3.0 * a + 2.0 * l) / (ep * ep) +
         l * log(16.0) - 4.0 * l * log(ep);
}

static double fixflux_ca(const wfl_geometry *g, double l2) {   /* EffAct.c:217-229 */
  const double q3 = 0.0187671, a = g->a, l = sqrt(l2);
  const double A = q3 * (a * a / (l2 * g->fix_q2 * g->fix_q2));
  const double lm = (l - g->lam_min) / (a - g->lam_min);
  return 4.0 * A + lm * ((9.0 / 4.0 * A - 9.0 / 2.0) * lm - 6.0 * A + 12.0);
}

/* Classical action of profile g->kind (EffAct.c:231-263); the spline's is 1. */
static double class_action_geom(const wfl_geometry *g, double F, double l2) {
  const double e2 = WFL_E * WFL_E;
  switch (g->kind) {
    case WFL_STEP: return -2.0 * WFL_PI * F * F / (e2 * l2);
    case WFL_SMOOTH: return -2.0 * WFL_PI * F * F / (3.0 * l2 * e2);
    case WFL_QUADRATIC: return -8.0 * WFL_PI * F * F / (3.0 * l2 * e2);
    case WFL_GAUSSIAN: return -WFL_PI * F * F / (l2 * e2);
    case WFL_PERIODIC: return -WFL_PI * sqrt(l2) / (24.0 * l2 * l2 * log(2.0) * log(2.0)) * periodic_ca(g->a, l2);
    case WFL_SPLINE: return 1.0;
    case WFL_FIXFLUX:
    case WFL_FIXFLUX_FF: return -WFL_PI * F * F / (e2 * g->a * g->a) * fixflux_ca(g, l2);
  }
  return 1.0;
}

double PeriodicCA(double l2) { return periodic_ca(wfl_internal_profile()->a, l2); }
double FixFluxCA(double l2) { return fixflux_ca(wfl_internal_profile(), l2); }
double ClassAction(double F, double l2) { return class_action_geom(wfl_internal_profile(), F, l2); }

/* Mean and standard error over ng groups of the normalised rho integral,
 * and the reference's two summary lines. */
static void eaction_report(double *intout, int ng, int fermion, double l2, double ca, double *result, double *se) {
  double r = 0.0, s = 0.0;
  for (int i = 0; i < ng; ++i) {
    intout[i] /= (fermion == 1) ? 4.0 * WFL_PI : -2.0 * WFL_PI;
    r += intout[i];
    printf("rho: intout[%d]=%e\n", i, intout[i]);
  }
  r /= ng;
  for (int i = 0; i < ng; ++i) s += (intout[i] - r) * (intout[i] - r);
  s = sqrt(s / ((double)ng * ((double)ng - 1.0)));
  if (g_verbosity >= 1) {
    printf("EAvsLambda: %e %e %e %e \n", sqrt(l2), ca, r, s);
    printf("ActRatio: %e %e %e\n", sqrt(l2), r / ca, s / ca);
  }
  if (result) *result = r;
  if (se) *se = s;
}

double *EAction(int Npoints, void *p) {   /* EffAct.c:265-306 */
  struct Wparams *params = (struct Wparams *)p;
  const int kind = wfl_internal_profile()->kind;
  double *intout = (kind == WFL_PERIODIC || kind == WFL_SPLINE || kind == WFL_FIXFLUX || kind == WFL_FIXFLUX_FF)
                       ? integrateRhotoa(2 * Npoints, p)
                       : integrateRho(Npoints, p);
  if (!intout) return NULL;
  eaction_report(intout, params->ng, params->fermion, params->l2, ClassAction(params->F, params->l2), NULL, NULL);
  return intout;
}

/* ------------------------------------------------------------------ */
/* Batched effective action                                             */
/* ------------------------------------------------------------------ */
/* The rho nodes of EAction: integrateRho's t = h(i+1) (twice) and h(i+1/2),
 * rho = (1 - t)/t, with weights h/6 (1, 4, 1) times 1/t^2; or
 * integrateRhotoa's rho = 0, a/2 and k a/(4N), k = 1..2N, with its weights.
 * Returns the node count; rho[], w[] hold nodes and weights (the rho factor
 * of rhogrand included). */
static int eaction_rho_nodes(int kind, double a, int Npoints, double *rho, double *w) {
  int m = 0;
  if (kind == WFL_PERIODIC || kind == WFL_SPLINE || kind == WFL_FIXFLUX || kind == WFL_FIXFLUX_FF) {
    const int n = 2 * Npoints;
    const double h = a / (double)(2 * n);
    rho[m] = 0.0; w[m++] = 0.0;                       /* rhogrand(0) = 0 */
    rho[m] = a / 2.0; w[m++] = h / 3.0 * (a / 2.0);
    for (int k = 1; k <= n; ++k) { rho[m] = h * k; w[m++] = h / 3.0 * rhotoa_weight(k, n) * (h * k); }
    return m;
  }
  const double h = 1.0 / (double)Npoints;
  for (int inx = 0; inx < Npoints; ++inx) {
    const double t1 = h * ((double)inx + 1.0), th = h * ((double)inx + 0.5);
    const double r1 = (1.0 - t1) / t1, rh = (1.0 - th) / th;
    rho[m] = r1; w[m++] = h / 6.0 * 2.0 * r1 / (t1 * t1);
    rho[m] = rh; w[m++] = h / 6.0 * 4.0 * rh / (th * th);
  }
  return m;
}

int wfl_eaction(const wfl_geometry *g, wfl_loops *L, int fermion, int mode, int precision, int nsub, int group,
                int Npoints, int nT, double *groups_out, double *result, double *se, double *classical) {
  if (!g || !L || Npoints < 1 || nT < 1 || group < 1 || (fermion != 0 && fermion != 1)) {
    wfl_set_error("wfl_eaction: bad arguments");
    return WFL_ERR_ARG;
  }
  if (g->lattice > 0.0 || g->kind == WFL_SPLINE) {
    wfl_set_error("wfl_eaction: isolated analytic profiles only (the spline and lattice geometries need tables "
                  "and cell integrals)");
    return WFL_ERR_ARG;
  }
  int Nl = 0, N = 0;
  int rc = wfl_loops_shape(L, &Nl, &N);
  if (rc) return rc;
  const int ng = (Nl + group - 1) / group;
  const int nr_max = 4 * Npoints + 4;
  double *rho = (double *)malloc(sizeof(double) * nr_max), *wr = (double *)malloc(sizeof(double) * nr_max);
  /* T nodes of integrateT(nT): t = h(i+1) (twice) and h(i+1/2), weights h/6 (2, 4)/t^2 */
  double *tT = (double *)malloc(sizeof(double) * 2 * nT), *wT = (double *)malloc(sizeof(double) * 2 * nT);
  if (!rho || !wr || !tT || !wT) { free(rho); free(wr); free(tT); free(wT); wfl_set_error("wfl_eaction: malloc"); return WFL_ERR_NOMEM; }
  const int nr = eaction_rho_nodes(g->kind, g->a, Npoints, rho, wr);
  const double hT = 1.0 / (double)nT;
  for (int i = 0; i < nT; ++i) {
    const double t1 = hT * ((double)i + 1.0), th = hT * ((double)i + 0.5);
    tT[2 * i] = (1.0 - t1) / t1;  wT[2 * i] = hT / 6.0 * 2.0 / (t1 * t1);
    tT[2 * i + 1] = (1.0 - th) / th;  wT[2 * i + 1] = hT / 6.0 * 4.0 / (th * th);
  }
  const int nTn = 2 * nT;
  /* tuples: the (rho, T) nodes whose integrand takes the worldline branch */
  const size_t nn = (size_t)nr * nTn;
  int *tg = (int *)calloc(nn, sizeof(int)), *slot = (int *)malloc(sizeof(int) * nn);
  double *tcx = (double *)malloc(sizeof(double) * nn), *tcy = (double *)calloc(nn, sizeof(double));
  double *tTv = (double *)malloc(sizeof(double) * nn), *fixed = (double *)calloc(nn, sizeof(double));
  double *Bv = (double *)malloc(sizeof(double) * nn), *Te = (double *)malloc(sizeof(double) * nn);
  double *acc = (double *)calloc((size_t)ng, sizeof(double));
  double *gsum = NULL;
  wfl_plan *plan = NULL;
  if (!tg || !slot || !tcx || !tcy || !tTv || !fixed || !Bv || !Te || !acc) { rc = WFL_ERR_NOMEM; wfl_set_error("wfl_eaction: malloc"); goto done; }
  int ntup = 0;
  for (int ir = 0; ir < nr; ++ir)
    for (int k = 0; k < nTn; ++k) {
      const size_t idx = (size_t)ir * nTn + k;
      slot[idx] = -1;
      if (wr[ir] == 0.0) continue;                   /* rhogrand(0) = 0 */
      const double T = tT[k];
      wfl_geometry gq = *g;
      double cx = rho[ir];
      if (T > 0.0) {
        double e[6];
        rc = wfl_effective_tuple(g, rho[ir], 0.0, T, e);
        if (rc) goto done;
        gq.l2 = e[0];
        gq.F = e[1];
        Te[idx] = e[4] * e[4];
        cx = e[2];
      } else {
        Te[idx] = 0.0;
      }
      int branch = 3;
      fixed[idx] = wfl_internal_igrand(&gq, fermion, Te[idx], fabs(cx), cx * cx, 0.0, &branch, &Bv[idx]);
      /* T = 0 (t = 1) in the worldline branch: a loop of zero size, W - 1 = 0,
       * and the integrand 0/0; taken as 0 */
      if (branch != 3 || !(T > 0.0)) {
        if (branch == 3) fixed[idx] = 0.0;
        continue;
      }
      slot[idx] = ntup;
      tcx[ntup] = rho[ir];
      tTv[ntup] = T;
      ++ntup;
    }
  gsum = (double *)malloc(sizeof(double) * (size_t)(ntup > 0 ? ntup : 1) * ng * 3);
  if (!gsum) { rc = WFL_ERR_NOMEM; wfl_set_error("wfl_eaction: malloc"); goto done; }
  if (ntup > 0) {
    wfl_batch b;
    memset(&b, 0, sizeof b);
    b.mode = mode; b.precision = precision; b.nsub = nsub; b.spinor = fermion; b.group = group;
    b.ngeom = 1; b.geoms = g; b.ntup = ntup;
    b.tup_geom = tg; b.tup_cx = tcx; b.tup_cy = tcy; b.tup_T = tTv;
    rc = wfl_plan_create(&plan, &b, L);
    if (!rc) rc = wfl_plan_run(plan, NULL);
    if (!rc) rc = wfl_plan_read(plan, gsum, NULL);
    if (rc) goto done;
  }
  for (int q = 0; q < ng; ++q) {
    const int cnt = (q + 1) * group <= Nl ? group : Nl - q * group;
    double S = 0.0;
    for (int ir = 0; ir < nr; ++ir) {
      double I = 0.0;
      for (int k = 0; k < nTn; ++k) {
        const size_t idx = (size_t)ir * nTn + k;
        double v;
        if (slot[idx] < 0) {
          v = fixed[idx];
        } else {
          const double evm1 = wfl_group_ev_minus1(gsum + ((size_t)slot[idx] * ng + q) * 3, cnt);
          const double TB = WFL_E * Te[idx] * Bv[idx], t = Te[idx];
          v = (fermion == 1) ? exp(-t) / (t * t * t) * (evm1 - TB * TB / 3.0)
                             : exp(-t) / (t * t * t) * (evm1 + TB * TB / 6.0);
        }
        if (isinf(v)) v = 0.0;   /* integrateT sets infinite node values to zero */
        I += wT[k] * v;
      }
      S += wr[ir] * I;
    }
    acc[q] = S;
  }
  {
    double r = 0.0, s = 0.0;
    const double ca = class_action_geom(g, g->F, g->l2);
    eaction_report(acc, ng, fermion, g->l2, ca, &r, &s);
    if (groups_out) memcpy(groups_out, acc, sizeof(double) * ng);
    if (result) *result = r;
    if (se) *se = s;
    if (classical) *classical = ca;
  }
done:
  if (plan) wfl_plan_destroy(plan);
  free(rho); free(wr); free(tT); free(wT); free(tg); free(slot); free(tcx); free(tcy); free(tTv);
  free(fixed); free(Bv); free(Te); free(acc); free(gsum);
  return rc;
}

/* ------------------------------------------------------------------ */
/* Lattice physics outputs (analysis-scripts/interaction.sage:10-25)     */
/* ------------------------------------------------------------------ */
/* Per-group action of one lattice tube: the L_eff of each x_cm point of
 * the unit cell weighted by w_c (the midpoint rule over an n x n grid has
 * w_c = l^2 / n^2).  gsum [ncell][nT][ngroups][3]: the group sums of the
 * cell points' tuples, each in T order. */
int wfl_cell_action_groups(const wfl_geometry *g, int fermion, int nT, const double *T, int ncell, const double *cx,
                           const double *cy, const double *w, const double *gsum, const int *counts, int ngroups,
                           int policy, double *action) {
  if (!g || !T || !cx || !cy || !w || !gsum || !counts || !action || nT < 2 || ncell < 1 || ngroups < 1) {
    wfl_set_error("wfl_cell_action_groups: bad arguments");
    return WFL_ERR_ARG;
  }
  double *leff = (double *)malloc(sizeof(double) * (size_t)ngroups);
  if (!leff) { wfl_set_error("wfl_cell_action_groups: malloc"); return WFL_ERR_NOMEM; }
  for (int q = 0; q < ngroups; ++q) action[q] = 0.0;
  for (int c = 0; c < ncell; ++c) {
    int rc = wfl_leff_groups(g, fermion, nT, T, cx[c], cy[c], gsum + (size_t)c * nT * ngroups * 3, counts, ngroups,
                             policy, leff);
    if (rc) { free(leff); return rc; }
    for (int q = 0; q < ngroups; ++q) action[q] += w[c] * leff[q];
  }
  free(leff);
  return WFL_OK;
}

/* Composite Simpson on n uniformly spaced values (the 3/8 rule on the last
 * three intervals when the interval count is odd; trapezoid for n = 2). */
static double simpson_uniform(double h, const double *f, int n) {
  const int m = n - 1;
  if (m < 1) return 0.0;
  if (m == 1) return 0.5 * h * (f[0] + f[1]);
  const int ms = (m % 2 == 0) ? m : m - 3;
  double s = 0.0;
  for (int i = 0; i + 2 <= ms; i += 2) s += h / 3.0 * (f[i] + 4.0 * f[i + 1] + f[i + 2]);
  if (ms != m) s += 3.0 * h / 8.0 * (f[ms] + 3.0 * f[ms + 1] + 3.0 * f[ms + 2] + f[ms + 3]);
  return s;
}

/* Per-group action of an isolated tube, 2 pi int rho L_eff(rho) drho over
 * nrho uniformly spaced rho (composite Simpson).  gsum [nrho][nT][ngroups][3]. */
int wfl_radial_action_groups(const wfl_geometry *g, int fermion, int nT, const double *T, int nrho, const double *rho,
                             const double *gsum, const int *counts, int ngroups, int policy, double *action) {
  if (!g || !T || !rho || !gsum || !counts || !action || nT < 2 || nrho < 2 || ngroups < 1) {
    wfl_set_error("wfl_radial_action_groups: bad arguments");
    return WFL_ERR_ARG;
  }
  const double h = (rho[nrho - 1] - rho[0]) / (double)(nrho - 1);
  for (int r = 1; r < nrho; ++r)
    if (!(fabs(rho[r] - rho[0] - h * r) <= 1e-9 * (fabs(rho[nrho - 1]) + h))) {
      wfl_set_error("wfl_radial_action_groups: rho must be uniformly spaced");
      return WFL_ERR_ARG;
    }
  double *f = (double *)malloc(sizeof(double) * ((size_t)nrho * ngroups + ngroups));
  if (!f) { wfl_set_error("wfl_radial_action_groups: malloc"); return WFL_ERR_NOMEM; }
  double *leff = f + (size_t)nrho * ngroups;
  for (int r = 0; r < nrho; ++r) {
    int rc = wfl_leff_groups(g, fermion, nT, T, rho[r], 0.0, gsum + (size_t)r * nT * ngroups * 3, counts, ngroups,
                             policy, leff);
    if (rc) { free(f); return rc; }
    for (int q = 0; q < ngroups; ++q) f[(size_t)q * nrho + r] = 2.0 * WFL_PI * rho[r] * leff[q];
  }
  for (int q = 0; q < ngroups; ++q) action[q] = simpson_uniform(h, f + (size_t)q * nrho, nrho);
  free(f);
  return WFL_OK;
}

/* Interaction energy per tube as interaction.sage:10-25 forms it from
 * per-group results: energies E = -action (the script negates the action
 * column, :13-16), the group-wise difference E_cell - E_iso, its mean, and
 * the script's error sqrt(var(E_cell) + var(E_iso) - 2 cov(E_cell, E_iso))
 * (sample variance and covariance, R's var/cov, :23-24), which is the
 * spread of one group's difference; out = {mean, that spread, spread /
 * sqrt(ngroups) (the standard error of the mean)}.  Energies in units of
 * the electron mass (the script's toMeV = 510.9989 converts to keV). */
int wfl_interaction_energy(int ngroups, const double *cell_action, const double *iso_action, double *out) {
  if (ngroups < 2 || !cell_action || !iso_action || !out) {
    wfl_set_error("wfl_interaction_energy: need >= 2 groups and buffers");
    return WFL_ERR_ARG;
  }
  double mc = 0.0, mi = 0.0;
  for (int q = 0; q < ngroups; ++q) { mc += -cell_action[q]; mi += -iso_action[q]; }
  mc /= ngroups; mi /= ngroups;
  double vc = 0.0, vi = 0.0, cv = 0.0;
  for (int q = 0; q < ngroups; ++q) {
    const double dc = -cell_action[q] - mc, di = -iso_action[q] - mi;
    vc += dc * dc; vi += di * di; cv += dc * di;
  }
  const double n1 = (double)(ngroups - 1);
  const double var = vc / n1 + vi / n1 - 2.0 * cv / n1;
  out[0] = mc - mi;
  out[1] = sqrt(var > 0.0 ? var : 0.0);
  out[2] = out[1] / sqrt((double)ngroups);
  return WFL_OK;
}
