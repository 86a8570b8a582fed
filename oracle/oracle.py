"""ctypes binding of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker, never by the product.
See wfl_oracle.h for the reference file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

STEP, SMOOTH, QUADRATIC, GAUSSIAN, PERIODIC, SPLINE, FIXFLUX, FIXFLUX_FF = range(8)
MODE_REFERENCE, MODE_PHYSICAL = 0, 1


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


if not os.path.exists(LIB_PATH):
    build()


class OrcGeom(C.Structure):
    _fields_ = [
        ("profile", C.c_int), ("l2", C.c_double), ("F", C.c_double),
        ("tubedist", C.c_double), ("lmin", C.c_double), ("q1", C.c_double), ("q2", C.c_double),
        ("nspline", C.c_int), ("rho2sp", C.c_void_p), ("coefs", C.c_void_p),
        ("lattice", C.c_double), ("images", C.c_int),
    ]


lib = C.CDLL(LIB_PATH)
_P, _I, _D = C.c_void_p, C.c_int, C.c_double
for _n, _r, _a in [
    ("orc_philox4x32_10", None, [_P, _P, _P]),
    ("orc_gen_loops", None, [C.c_uint32, _I, _I, _P, _P]),
    ("orc_succ", _I, [_I, _I]),
    ("orc_f", _D, [_P, _D]),
    ("orc_f_in_mode", _D, [_P, _I, _D]),
    ("orc_spline_low_slope", _D, [_P, _P]),
    ("orc_fp", _D, [_P, _D]),
    ("orc_expint", _D, [_D]),
    ("orc_interp", _D, [_D, _P, _P, _I]),
    ("orc_wilson_loop", None, [_P, _I, _I, _I, _P, _P, _I, _D, _D, _D, _P, _P]),
    ("orc_eval_batch", None, [_P, _I, _I, _I, _P, _P, _I, _I, _I, _I, _P, _P, _P, _P, _P, _P, _I]),
    ("orc_wminus1", None, [_I, _I, _P, _P, _P]),
    ("orc_evmean", _D, [_P, _P, _I, _P, _I]),
    ("orc_exact", _D, [_D, _D, _I]),
    ("orc_smTscal", _D, [_D, _D]),
    ("orc_smTferm", _D, [_D, _D]),
    ("orc_magfield", _D, [_P, _D]),
    ("orc_smallT", _D, [_P, _D, _D]),
    ("orc_igrand", _D, [_P, _I, _D, _D, _D, _D, _P]),
    ("orc_field_at", _D, [_P, _D, _D, _P]),
    ("orc_igrand_B", _D, [_P, _I, _D, _D, _D, _D, _P]),
    ("orc_spline", None, [_P, _P, _I, _I, _P, _P]),
    ("orc_logsimpson", _D, [_P, _P, _I]),
]:
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a


def _p(a):
    return None if a is None else a.ctypes.data


LOG2E = 1.4426950408889634074


def q32(v):
    """FP32 narrowing of one parameter."""
    return float(np.float32(v))


def unit_scale(profile, l2):
    """The engine's FP32 unit scale: sigma = sqrt(log2 e / l2) in FP32 for
    the Gaussian fast path, 1 for the other profiles."""
    return q32(np.sqrt(LOG2E / l2)) if profile == GAUSSIAN else 1.0


class Geom:
    """Oracle geometry.

    quantize=True: the batched engine's effective parameters (every kernel
    parameter narrowed to FP32, in the scaled units sigma z for a Gaussian
    tube: l2 = log2(e)/sigma^2, lattice = fl32(sigma a)/sigma);
    quantize="float": the reference's narrowing of F and l2 to float
    (intEV.cu:124), which EV() keeps; quantize=False: the values as given."""

    def __init__(self, profile=GAUSSIAN, F=1.0, l2=1.0, tubedist=1.0, lmin=0.1,
                 q1=0.443993816168079, q2=0.0742477533879612, lattice=0.0, images=0,
                 knots=None, coefs=None, quantize=True):
        qz = q32 if quantize else float
        self.sg = unit_scale(profile, l2) if quantize is True else 1.0
        l2e = LOG2E / (self.sg * self.sg) if self.sg != 1.0 else qz(l2)
        late = (q32(self.sg * lattice) / self.sg if quantize else float(lattice)) if lattice > 0 else 0.0
        self.knots = None if knots is None else np.ascontiguousarray(knots, np.float32)
        self.coefs = None if coefs is None else np.ascontiguousarray(coefs, np.float32)
        self.c = OrcGeom(profile, l2e, qz(F), qz(tubedist), qz(lmin), qz(q1), qz(q2),
                         0 if knots is None else len(self.knots), _p(self.knots), _p(self.coefs),
                         late, images)
        self.quantize = quantize

    def tuple(self, cx, cy, T):
        """Effective (cx, cy, sqrt(T)) of a batch tuple."""
        if self.quantize is True:
            sg = self.sg
            return (q32(sg * cx) / sg, q32(sg * cy) / sg, q32(sg * np.sqrt(T)) / sg)
        if self.quantize:
            return (q32(cx), q32(cy), q32(np.sqrt(T)))
        return (float(cx), float(cy), float(np.sqrt(T)))

    @property
    def ref(self):
        return C.byref(self.c)


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib.orc_philox4x32_10(c, k, o)
    return list(o)


def gen_loops(config_id, Nl, N):
    x = np.empty((Nl, N), np.float32)
    y = np.empty((Nl, N), np.float32)
    lib.orc_gen_loops(config_id, Nl, N, _p(x), _p(y))
    return x, y


def wilson_loop(geom, mode, nsub, spinor, wx, wy, cx, cy, rtT):
    wx = np.ascontiguousarray(wx, np.float32)
    wy = np.ascontiguousarray(wy, np.float32)
    sc = np.zeros(3)
    fe = np.zeros(3)
    lib.orc_wilson_loop(geom.ref, mode, nsub, spinor, _p(wx), _p(wy), len(wx), float(cx), float(cy),
                        float(rtT), _p(sc), _p(fe))
    return sc, fe


def eval_batch(geoms, mode, nsub, spinor, x, y, group, tup_geom, tup_cx, tup_cy, tup_T,
               perloop=True, nthreads=0):
    """Same tuple semantics as the engine (Geom.tuple).  Returns
    (gsum [ntup, ng, 3], perloop [ntup, Nl, 6])."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    Nl, N = x.shape
    ntup = len(tup_geom)
    ng = (Nl + group - 1) // group
    garr = (OrcGeom * len(geoms))(*[g.c for g in geoms])
    tg = np.ascontiguousarray(tup_geom, np.int32)
    eff = np.array([geoms[g].tuple(a, b, t) for g, a, b, t in zip(tg, tup_cx, tup_cy, tup_T)], np.float64)
    eff = eff.reshape(-1, 3)
    cx = np.ascontiguousarray(eff[:, 0])
    cy = np.ascontiguousarray(eff[:, 1])
    rt = np.ascontiguousarray(eff[:, 2])
    gs = np.zeros((ntup, ng, 3))
    pl = np.zeros((ntup, Nl, 6)) if perloop else None
    lib.orc_eval_batch(C.cast(garr, C.c_void_p), mode, nsub, spinor, _p(x), _p(y), Nl, N, group, ntup,
                       _p(tg), _p(cx), _p(cy), _p(rt), _p(pl), _p(gs), nthreads)
    return gs, pl


def wminus1(mode, spinor, sc, fe):
    sc = np.ascontiguousarray(sc, np.float64)
    fe = np.ascontiguousarray(fe, np.float64)
    o = np.zeros(3)
    lib.orc_wminus1(mode, spinor, _p(sc), _p(fe), _p(o))
    return o


def evmean(ws4, wf4, WL, fermion):
    ws4 = np.ascontiguousarray(ws4, np.float32)
    wf4 = np.ascontiguousarray(wf4 if wf4 is not None else np.zeros_like(ws4), np.float32)
    WL = np.ascontiguousarray(WL, np.int32)
    return lib.orc_evmean(_p(ws4), _p(wf4), len(WL), _p(WL), fermion)


def igrand(geom, fermion, T, rho, rho2, EV):
    b = C.c_int(-1)
    v = lib.orc_igrand(geom.ref, fermion, T, rho, rho2, EV, C.byref(b))
    return v, b.value


def spline(fli, rho2, quirk=False):
    fli = np.ascontiguousarray(fli, np.float64)
    rho2 = np.ascontiguousarray(rho2, np.float32)
    n = len(rho2)
    c4 = np.zeros((n, 4), np.float32)
    cd = np.zeros((n, 4), np.float64)
    lib.orc_spline(_p(fli), _p(rho2), n, int(quirk), _p(c4), _p(cd))
    return c4, cd


def logsimpson(T, f):
    T = np.ascontiguousarray(T, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    return lib.orc_logsimpson(_p(T), _p(f), len(T))


def field_at(geom, cx, cy):
    """(eB at x_cm over the lattice image window, distance to the nearest tube)."""
    r = C.c_double(0.0)
    B = lib.orc_field_at(geom.ref, cx, cy, C.byref(r))
    return B, r.value


def leff_groups(geom, fermion, T, cx, cy, gsum, counts, policy=1, estimator=0):
    """Oracle L_eff per group: Tfunc's branch policy at each node's effective
    parameters, composite Simpson in ln T over the nominal nodes, /(-2pi)
    scalar or /(4pi) spinor (Tint.c:69-72).  The field is the image-window
    sum for lattice geometries."""
    ng = gsum.shape[1]
    out = np.zeros(ng)
    f = np.zeros(len(T))
    for q in range(ng):
        for k, Tn in enumerate(T):
            ex, ey, rt = geom.tuple(cx, cy, Tn)
            t = rt * rt
            B, rho = field_at(geom, ex, ey)
            g3 = gsum[k, q]
            if estimator == 1:   # two-level extrapolation 2 W_{N/2} - W_{N/4}
                evm1 = (2.0 * g3[1] - g3[2]) / counts[q]
            else:                # Integrand.c:400
                evm1 = (8.0 / 3.0 * g3[0] - 2.0 * g3[1] + 1.0 / 3.0 * g3[2]) / counts[q]
            b = C.c_int(-1)
            v = lib.orc_igrand_B(geom.ref, fermion, t, B, rho, 1.0 + evm1, C.byref(b))
            if b.value == 3 or not policy:
                TB = 0.30282212 * t * B
                v = np.exp(-t) / t**3 * (evm1 - TB * TB / 3.0 if fermion else evm1 + TB * TB / 6.0)
            f[k] = v
        I = logsimpson(np.asarray(T, np.float64), f)
        out[q] = I / (4 * np.pi) if fermion else I / (-2 * np.pi)
    return out


def simpson_uniform(h, f):
    """Composite Simpson on uniformly spaced values; the 3/8 rule on the last
    three intervals when the interval count is odd (trapezoid for two
    values)."""
    f = np.asarray(f, np.float64)
    m = len(f) - 1
    if m < 1:
        return 0.0
    if m == 1:
        return 0.5 * h * (f[0] + f[1])
    ms = m if m % 2 == 0 else m - 3
    s = sum(h / 3.0 * (f[i] + 4.0 * f[i + 1] + f[i + 2]) for i in range(0, ms, 2))
    if ms != m:
        s += 3.0 * h / 8.0 * (f[ms] + 3.0 * f[ms + 1] + 3.0 * f[ms + 2] + f[ms + 3])
    return s


def cell_action_groups(geom, fermion, T, cx, cy, w, gsum, counts, policy=1):
    """Per-group action of a lattice tube: sum_c w_c L_eff(x_c), gsum
    [ncell, nT, ngroups, 3]."""
    return sum(wc * leff_groups(geom, fermion, T, a, b, gsum[c], counts, policy)
               for c, (a, b, wc) in enumerate(zip(cx, cy, w)))


def radial_action_groups(geom, fermion, T, rho, gsum, counts, policy=1):
    """Per-group action of an isolated tube: 2 pi int rho L_eff drho,
    composite Simpson on the uniform rho grid; gsum [nrho, nT, ngroups, 3]."""
    L = np.array([leff_groups(geom, fermion, T, r, 0.0, gsum[k], counts, policy) for k, r in enumerate(rho)])
    f = 2.0 * np.pi * np.asarray(rho)[:, None] * L
    h = (rho[-1] - rho[0]) / (len(rho) - 1)
    return np.array([simpson_uniform(h, f[:, q]) for q in range(f.shape[1])])


def interaction_energy(cell_action, iso_action):
    """analysis-scripts/interaction.sage:10-25: energies E = -action, the
    mean of E_cell - E_iso over groups and the script's error
    sqrt(var(E_cell) + var(E_iso) - 2 cov(E_cell, E_iso)) with R's sample
    (n - 1) var and cov; also that error / sqrt(n)."""
    ec = -np.asarray(cell_action, np.float64)
    ei = -np.asarray(iso_action, np.float64)
    n = len(ec)
    var = np.var(ec, ddof=1) + np.var(ei, ddof=1) - 2.0 * np.cov(ec, ei, ddof=1)[0, 1]
    err = float(np.sqrt(max(var, 0.0)))
    return float(np.mean(ec - ei)), err, err / np.sqrt(n)


def _rule_nodes(n):
    """integrateT's / integrateRho's rule on t in (0, 1] (Tint.c:100-175,
    EffAct.c:137-199): per interval the right end twice and the midpoint,
    h/6 (2 f(t_(i+1)) + 4 f(t_(i+1/2))); returns (t, weight)."""
    h = 1.0 / n
    t, w = [], []
    for i in range(n):
        t += [h * (i + 1.0), h * (i + 0.5)]
        w += [h / 6.0 * 2.0, h / 6.0 * 4.0]
    return np.array(t), np.array(w)


def eaction_groups(geom, x, y, fermion, mode, nsub, group, npoints, nT, lcf=False, nthreads=0):
    """Per-group effective action of an isolated tube as EffAct.c:265-306
    computes it: integrateRho(npoints) over rho = (1 - t)/t with
    rhogrand's factor rho and 1/t^2, integrateT(nT) over T = (1 - t)/t with
    1/t^2, Tfunc's branch policy (Integrand.c:195-316) at each node's
    effective parameters (Geom.tuple), the worldline estimate from the
    oracle's group sums, and the /(4 pi) (spinor) or /(-2 pi) (scalar)
    normalisation.  lcf=True takes the constant-field closed form Exact at
    every node (T > 0) instead: the locally-constant-field action on the
    same rules."""
    tr, wr = _rule_nodes(npoints)
    tt, wt = _rule_nodes(nT)
    rho = (1.0 - tr) / tr
    T = (1.0 - tt) / tt
    Nl = x.shape[0]
    ng = (Nl + group - 1) // group
    counts = np.array([min(group, Nl - q * group) for q in range(ng)])
    fixed = np.zeros((len(rho), len(T)))
    need = []
    for i, r in enumerate(rho):
        for k, t in enumerate(T):
            if r == 0.0:
                continue
            ex, ey, rt = geom.tuple(r, 0.0, t)
            te = rt * rt
            B = lib.orc_magfield(geom.ref, ex * ex)
            if lcf:
                fixed[i, k] = exact(te, B, fermion) if te > 0 else 0.0
                continue
            v, b = igrand(geom, fermion, te, abs(ex), ex * ex, 1.0)
            if b == 3 and t > 0.0:
                need.append((i, k, te, B))
            else:
                fixed[i, k] = 0.0 if b == 3 else v
    gs = None
    if need:
        gs, _ = eval_batch([geom], mode, nsub, fermion, x, y, group, np.zeros(len(need), np.int32),
                           np.array([rho[i] for i, _, _, _ in need]), np.zeros(len(need)),
                           np.array([T[k] for _, k, _, _ in need]), perloop=False, nthreads=nthreads)
    out = np.zeros(ng)
    for q in range(ng):
        f = fixed.copy()
        for n, (i, k, te, B) in enumerate(need):
            g3 = gs[n, q]
            evm1 = (8.0 / 3.0 * g3[0] - 2.0 * g3[1] + 1.0 / 3.0 * g3[2]) / counts[q]
            TB = 0.30282212 * te * B
            f[i, k] = np.exp(-te) / te ** 3 * (evm1 - TB * TB / 3.0 if fermion else evm1 + TB * TB / 6.0)
        f[np.isinf(f)] = 0.0
        I = f @ (wt / (tt * tt))
        out[q] = np.sum(wr * rho / (tr * tr) * I)
    return out / (4.0 * np.pi) if fermion else out / (-2.0 * np.pi)


def exact(T, B, fermion):
    return lib.orc_exact(T, B, fermion)
