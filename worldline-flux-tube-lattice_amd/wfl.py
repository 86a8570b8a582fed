"""ctypes binding of libwfl.so, the MI355X worldline Wilson-loop engine.

Python mirror of the C ABI in include/wfl.h.  The reference's interface for
this path is the C call ``double EV(double T, void *p, int *WLlist)``
(intEV.cu:111) fed by ``struct Wparams``; this module exposes that call, the
host C functions around it (EVMean, Tfunc-style policy, closed forms) and the
batched engine (plans over many (geometry, x_cm, T) tuples).

The library is loaded from this directory; importing fails loudly when it has
not been built (run ``python -c "import __graft_entry__ as g; g.build()"``).
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libwfl.so")

WFL_OK = 0
STEP, SMOOTH, QUADRATIC, GAUSSIAN, PERIODIC, SPLINE, FIXFLUX, FIXFLUX_FF = range(8)
MODE_REFERENCE, MODE_PHYSICAL = 0, 1
FP64, MIXED = 0, 1


class WflError(RuntimeError):
    pass


class float4(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float), ("w", C.c_float)]


class Wparams(C.Structure):
    """struct Wparams as the reconstructed intT.h declares it."""

    _fields_ = [
        ("Nl", C.c_int), ("Nppl", C.c_int),
        ("F", C.c_double), ("l2", C.c_double),
        ("xcm", float4),
        ("Wsscal_h", C.c_void_p), ("Wsscal_d", C.c_void_p),
        ("Wsferm_h", C.c_void_p), ("Wsferm_d", C.c_void_p),
        ("worldlines", C.c_void_p),
        ("SE", C.c_double),
        ("nBlocks", C.c_int), ("nThreads", C.c_int), ("ng", C.c_int), ("fermion", C.c_int),
        ("flcoefs", C.c_void_p), ("rho2sp", C.c_void_p),
    ]


class _Geometry(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("F", C.c_double), ("l2", C.c_double), ("a", C.c_double),
        ("lam_min", C.c_double), ("fix_q1", C.c_double), ("fix_q2", C.c_double),
        ("lattice", C.c_double), ("images", C.c_int), ("spline_table", C.c_int),
    ]


class _SplineTable(C.Structure):
    _fields_ = [("n", C.c_int), ("rho2", C.c_void_p), ("coefs", C.c_void_p)]


class _Batch(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("precision", C.c_int), ("nsub", C.c_int), ("spinor", C.c_int),
        ("group", C.c_int), ("ngeom", C.c_int), ("geoms", C.c_void_p),
        ("nspline", C.c_int), ("splines", C.c_void_p),
        ("ntup", C.c_int), ("tup_geom", C.c_void_p), ("tup_cx", C.c_void_p),
        ("tup_cy", C.c_void_p), ("tup_T", C.c_void_p), ("geom_nloops", C.c_void_p),
        ("want_perloop", C.c_int), ("mapping", C.c_int),
    ]


# Every symbol include/wfl.h and include/intT.h declare (checked by tests).
EXPORTS = [
    "wfl_last_error", "wfl_device_count", "wfl_device_sync",
    "wfl_loops_create", "wfl_loops_from_aos_device", "wfl_loops_generate", "wfl_loops_shape",
    "wfl_loops_destroy", "wfl_gen_loops", "wfl_read_text_loops",
    "wfl_plan_create", "wfl_plan_run", "wfl_plan_run_timed", "wfl_plan_autotune", "wfl_plan_config",
    "wfl_plan_set_config", "wfl_plan_graph_run", "wfl_plan_read", "wfl_plan_ngroups", "wfl_plan_accurate_tuples",
    "wfl_plan_uniform_splines", "wfl_spline_uniform_form", "wfl_autotune_nsub", "wfl_plan_axis_form", "wfl_plan_local_tuples",
    "wfl_local_taylor",
    "wfl_plan_gsum_device", "wfl_plan_nonfinite", "wfl_plan_mapping", "wfl_plan_destroy",
    "wfl_eval_batch", "wfl_eval_batch_noalloc", "wfl_group_ev_minus1", "wfl_spline_build", "wfl_set_profile",
    "wfl_set_numerics", "wfl_ev_reset", "wfl_ev_launches", "wfl_logsimpson", "wfl_effective_T",
    "wfl_effective_tuple",
    "wfl_leff_groups", "wfl_leff_groups_est", "wfl_group_ev_minus1_est",
    "EV", "EVMean", "Tfunc", "integrateT", "getwl", "Exact", "MagField", "smTscal", "smTferm",
    "flspline", "getspline", "testspline", "interp",
    "rhogrand", "testvalrho", "transformRho", "integrateRhotoa", "integrateRho", "PeriodicCA", "FixFluxCA",
    "ClassAction", "EAction", "wfl_eaction", "wfl_set_verbosity",
    "wfl_cell_action_groups", "wfl_radial_action_groups", "wfl_interaction_energy",
    "wfl_set_fp_traps", "wfl_fp_trap_count",
    "wfl_loops_write", "wfl_loops_read_header", "wfl_loops_read", "wfl_write_text_loops",
]


def _load():
    if not os.path.exists(LIB_PATH):
        raise WflError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
    lib = C.CDLL(LIB_PATH)
    P, I, D = C.c_void_p, C.c_int, C.c_double
    sig = {
        "wfl_last_error": (C.c_char_p, []),
        "wfl_device_count": (I, [P]),
        "wfl_device_sync": (I, []),
        "wfl_loops_create": (I, [P, P, P, I, I]),
        "wfl_loops_from_aos_device": (I, [P, P, I, I]),
        "wfl_loops_generate": (I, [P, C.c_uint32, I, I]),
        "wfl_loops_shape": (I, [P, P, P]),
        "wfl_loops_destroy": (None, [P]),
        "wfl_gen_loops": (I, [C.c_uint32, I, I, P, P]),
        "wfl_read_text_loops": (I, [C.c_char_p, I, I, P, P]),
        "wfl_loops_write": (I, [C.c_char_p, P, P, I, I, C.c_uint32]),
        "wfl_write_text_loops": (I, [C.c_char_p, P, P, I, I]),
        "wfl_loops_read_header": (I, [C.c_char_p, P, P, P]),
        "wfl_loops_read": (I, [C.c_char_p, P, P, I, I]),
        "wfl_plan_create": (I, [P, P, P]),
        "wfl_plan_run": (I, [P, P]),
        "wfl_plan_graph_run": (I, [P, P]),
        "wfl_plan_run_timed": (I, [P, P, P]),
        "wfl_plan_autotune": (I, [P, P, P]),
        "wfl_plan_config": (I, [P, P, P, P, P]),
        "wfl_plan_set_config": (I, [P, I, I, I, I]),
        "wfl_plan_read": (I, [P, P, P]),
        "wfl_plan_ngroups": (I, [P]),
        "wfl_plan_accurate_tuples": (I, [P]),
        "wfl_plan_uniform_splines": (I, [P]),
        "wfl_spline_uniform_form": (I, [P, P, I, D, I, P, I, P]),
        "wfl_plan_axis_form": (I, [P]),
        "wfl_plan_local_tuples": (I, [P]),
        "wfl_local_taylor": (I, [D, P]),
        "wfl_autotune_nsub": (I, [P, P, C.c_double, I, I, I, P, P]),
        "wfl_plan_gsum_device": (P, [P]),
        "wfl_plan_nonfinite": (I, [P, P]),
        "wfl_plan_mapping": (I, [P]),
        "wfl_plan_destroy": (None, [P]),
        "wfl_eval_batch": (I, [P, P, P, P]),
        "wfl_eval_batch_noalloc": (I, [P, P, P, P]),
        "wfl_group_ev_minus1": (D, [P, I]),
        "wfl_spline_build": (I, [P, P, I, I, P]),
        "wfl_set_profile": (I, [P, I]),
        "wfl_set_numerics": (I, [I, I, I]),
        "wfl_ev_reset": (None, []),
        "wfl_ev_launches": (C.c_longlong, []),
        "wfl_logsimpson": (D, [P, P, I]),
        "wfl_effective_T": (D, [P, D]),
        "wfl_effective_tuple": (I, [P, D, D, D, P]),
        "wfl_leff_groups": (I, [P, I, I, P, D, D, P, P, I, I, P]),
        "wfl_leff_groups_est": (I, [P, I, I, P, D, D, P, P, I, I, I, P]),
        "wfl_group_ev_minus1_est": (D, [P, I, I]),
        "EV": (D, [D, P, P]),
        "EVMean": (None, [P, P, P, I, P, D, I]),
        "Tfunc": (P, [D, P, P]),
        "integrateT": (P, [I, P]),
        "rhogrand": (P, [D, P]),
        "transformRho": (P, [D, P]),
        "integrateRhotoa": (P, [I, P]),
        "integrateRho": (P, [I, P]),
        "PeriodicCA": (D, [D]),
        "FixFluxCA": (D, [D]),
        "ClassAction": (D, [D, D]),
        "EAction": (P, [I, P]),
        "wfl_eaction": (I, [P, P, I, I, I, I, I, I, I, P, P, P, P]),
        "wfl_set_verbosity": (I, [I]),
        "wfl_cell_action_groups": (I, [P, I, I, P, I, P, P, P, P, P, I, I, P]),
        "wfl_radial_action_groups": (I, [P, I, I, P, I, P, P, P, I, I, P]),
        "wfl_interaction_energy": (I, [I, P, P, P]),
        "wfl_set_fp_traps": (I, [I]),
        "wfl_fp_trap_count": (C.c_longlong, []),
        "getwl": (I, [P, C.c_char_p, I, I, I]),
        "flspline": (P, [P, P, I]),
        "getspline": (P, [P, D, D, I]),
        "testspline": (D, [C.c_float, D, D, I, P, P]),
        "interp": (C.c_float, [C.c_float, P, P]),
        "Exact": (D, [D, D, I]),
        "MagField": (D, [D, P]),
        "smTscal": (D, [D, D]),
        "smTferm": (D, [D, D]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()
_libc = C.CDLL(None)
_libc.free.restype = None
_libc.free.argtypes = [C.c_void_p]


def _ptr(a):
    return None if a is None else a.ctypes.data


def check(rc):
    if rc != WFL_OK:
        raise WflError(f"wfl error {rc}: {lib.wfl_last_error().decode()}")
    return rc


def device_count() -> int:
    n = C.c_int(0)
    rc = lib.wfl_device_count(C.byref(n))
    return n.value if rc == WFL_OK else 0


def gen_loops(config_id: int, Nl: int, N: int):
    """Seeded unit loops (host generator), SoA float32 arrays [Nl, N]."""
    x = np.empty((Nl, N), np.float32)
    y = np.empty((Nl, N), np.float32)
    check(lib.wfl_gen_loops(config_id, Nl, N, _ptr(x), _ptr(y)))
    return x, y


def write_loops(path, x, y, config_id=0):
    """Store an SoA ensemble as a binary loop file (header + hash)."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    check(lib.wfl_loops_write(str(path).encode(), _ptr(x), _ptr(y), x.shape[0], x.shape[1], config_id))


def write_text_loops(path, x, y):
    """Store an SoA ensemble in the text layout getwl.c reads."""
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    check(lib.wfl_write_text_loops(str(path).encode(), _ptr(x), _ptr(y), x.shape[0], x.shape[1]))


def read_text_loops(path, Nl, N):
    """Read Nl loops of N points from the getwl.c text layout."""
    x = np.empty((Nl, N), np.float32)
    y = np.empty((Nl, N), np.float32)
    check(lib.wfl_read_text_loops(str(path).encode(), Nl, N, _ptr(x), _ptr(y)))
    return x, y


def read_loops(path):
    """Read a binary loop file; returns (x, y, config_id), hash-verified."""
    Nl, N, cid = C.c_int(), C.c_int(), C.c_uint32()
    check(lib.wfl_loops_read_header(str(path).encode(), C.addressof(Nl), C.addressof(N), C.addressof(cid)))
    x = np.empty((Nl.value, N.value), np.float32)
    y = np.empty((Nl.value, N.value), np.float32)
    check(lib.wfl_loops_read(str(path).encode(), _ptr(x), _ptr(y), Nl.value, N.value))
    return x, y, cid.value


def spline_build(fli, rho2, quirk=False):
    fli = np.ascontiguousarray(fli, np.float64)
    rho2 = np.ascontiguousarray(rho2, np.float32)
    c4 = np.zeros((len(rho2), 4), np.float32)
    check(lib.wfl_spline_build(_ptr(fli), _ptr(rho2), len(rho2), int(quirk), _ptr(c4)))
    return c4


def spline_uniform_form(rho2, coefs, umax, spinor=0, rows=False):
    """The form of the direct-index rows a mixed physical-mode plan over this
    one table would use for nodes up to u = umax, scalar or spinor: 0 none
    (exact search), 1 rows expanded about their start, 2 rows expanded about
    u = 0.  With
    rows=True, (form, rows) with rows[0] the guard row and rows[1 + i] row i
    (float32, shape (nrows, 4))."""
    rho2 = np.ascontiguousarray(rho2, np.float32)
    coefs = np.ascontiguousarray(coefs, np.float32).reshape(-1)
    if not rows:
        return lib.wfl_spline_uniform_form(_ptr(rho2), _ptr(coefs), len(rho2), float(umax), int(spinor), None, 0,
                                           None)
    out = np.zeros((1025, 4), np.float32)
    nr = C.c_int(0)
    form = lib.wfl_spline_uniform_form(_ptr(rho2), _ptr(coefs), len(rho2), float(umax), int(spinor), _ptr(out), len(out),
                                       C.byref(nr))
    return form, out[:nr.value].copy()


def _take_rows(ptr, n):
    """Copy n float4 rows the library malloc'd at ptr, then free them."""
    if not ptr:
        raise WflError(lib.wfl_last_error().decode())
    try:
        return np.ctypeslib.as_array((C.c_float * (4 * n)).from_address(ptr)).reshape(n, 4).copy()
    finally:
        _libc.free(ptr)


def getspline(lambda2, a, n=512):
    """(rho2, coefs) of the periodic profile's table, as getspline
    (periodic.c:112-162) builds it: knots (i + 1) 20/n, f by quadrature."""
    rho2 = np.zeros(n, np.float32)
    return rho2, _take_rows(lib.getspline(_ptr(rho2), float(lambda2), float(a), int(n)), n)


def flspline(fli, rho2):
    """flspline (periodic.c:45-110): rows with the reference's float arithmetic."""
    fli = np.ascontiguousarray(fli, np.float64)
    rho2 = np.ascontiguousarray(rho2, np.float32)
    return _take_rows(lib.flspline(_ptr(fli), _ptr(rho2), len(rho2)), len(rho2))


def testspline(rho2, lambda2, a, knots, coefs):
    """testspline (periodic.c:204-245): (f - s)/f at rho2."""
    knots = np.ascontiguousarray(knots, np.float32)
    coefs = np.ascontiguousarray(coefs, np.float32)
    return lib.testspline(float(rho2), float(lambda2), float(a), len(knots), _ptr(knots), _ptr(coefs))


def interp(rho2, knots, coefs):
    """interp (intEV.cu:64-107) on the host, over the Nspline knots that
    set_profile last gave (512 unless set)."""
    knots = np.ascontiguousarray(knots, np.float32)
    coefs = np.ascontiguousarray(coefs, np.float32)
    return lib.interp(float(rho2), _ptr(knots), _ptr(coefs))


def effective_T(geom, T):
    """Effective proper times the kernels evaluate for requested T values."""
    g = geom.c()
    return np.array([lib.wfl_effective_T(C.byref(g), float(t)) for t in np.atleast_1d(T)])


def effective_tuple(geom, cx, cy, T):
    """(l2, F, cx, cy, sqrt(T), lattice) a batch evaluates for one tuple."""
    g = geom.c()
    out = np.zeros(6)
    check(lib.wfl_effective_tuple(C.byref(g), float(cx), float(cy), float(T), _ptr(out)))
    return out


class Loops:
    """A device-resident loop ensemble."""

    def __init__(self, x=None, y=None, *, config_id=None, Nl=None, N=None):
        self.h = C.c_void_p()
        if config_id is not None:
            check(lib.wfl_loops_generate(C.byref(self.h), config_id, Nl, N))
        else:
            x = np.ascontiguousarray(x, np.float32)
            y = np.ascontiguousarray(y, np.float32)
            Nl, N = x.shape
            check(lib.wfl_loops_create(C.byref(self.h), _ptr(x), _ptr(y), Nl, N))
        self.Nl, self.N = Nl, N

    def close(self):
        if self.h:
            lib.wfl_loops_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Geometry:
    kind: int = GAUSSIAN
    F: float = 1.0
    l2: float = 1.0
    a: float = 1.0
    lam_min: float = 0.1
    fix_q1: float = 0.443993816168079
    fix_q2: float = 0.0742477533879612
    lattice: float = 0.0
    images: int = 0
    spline_table: int = -1

    def c(self):
        return _Geometry(self.kind, self.F, self.l2, self.a, self.lam_min, self.fix_q1, self.fix_q2,
                         self.lattice, self.images, self.spline_table)


@dataclass
class Batch:
    geoms: list
    tup_geom: np.ndarray
    tup_cx: np.ndarray
    tup_cy: np.ndarray
    tup_T: np.ndarray
    mode: int = MODE_PHYSICAL
    precision: int = MIXED
    nsub: int = 2
    spinor: int = 0
    group: int = 128
    splines: list = field(default_factory=list)   # [(rho2 float32[n], coefs float32[n,4])]
    geom_nloops: np.ndarray | None = None
    want_perloop: bool = False
    mapping: int = 0

    def c(self):
        keep = []
        g = (_Geometry * len(self.geoms))(*[gg.c() for gg in self.geoms])
        keep.append(g)
        sp = None
        if self.splines:
            arrs = []
            tabs = []
            for k, cf in self.splines:
                k = np.ascontiguousarray(k, np.float32)
                cf = np.ascontiguousarray(cf, np.float32)
                arrs += [k, cf]
                tabs.append(_SplineTable(len(k), k.ctypes.data, cf.ctypes.data))
            sp = (_SplineTable * len(tabs))(*tabs)
            keep += arrs + [sp]
        tg = np.ascontiguousarray(self.tup_geom, np.int32)
        cx = np.ascontiguousarray(self.tup_cx, np.float64)
        cy = np.ascontiguousarray(self.tup_cy, np.float64)
        T = np.ascontiguousarray(self.tup_T, np.float64)
        gn = None if self.geom_nloops is None else np.ascontiguousarray(self.geom_nloops, np.int32)
        keep += [tg, cx, cy, T, gn]
        b = _Batch(self.mode, self.precision, self.nsub, self.spinor, self.group, len(self.geoms),
                   C.cast(g, C.c_void_p), len(self.splines), None if sp is None else C.cast(sp, C.c_void_p),
                   len(tg), _ptr(tg), _ptr(cx), _ptr(cy), _ptr(T), _ptr(gn),
                   int(self.want_perloop), self.mapping)
        return b, keep


class Plan:
    """An uploaded batched workload (wfl_plan)."""

    def __init__(self, batch: Batch, loops: Loops):
        self.batch = batch
        self.loops = loops
        self._b, self._keep = batch.c()
        self.h = C.c_void_p()
        check(lib.wfl_plan_create(C.byref(self.h), C.byref(self._b), loops.h))
        self.ngroups = lib.wfl_plan_ngroups(self.h)
        self.accurate_tuples = lib.wfl_plan_accurate_tuples(self.h)
        self.uniform_splines = lib.wfl_plan_uniform_splines(self.h)
        self.axis_form = lib.wfl_plan_axis_form(self.h)
        self.local_tuples = lib.wfl_plan_local_tuples(self.h)
        self.mapping = lib.wfl_plan_mapping(self.h)
        self.ntup = len(batch.tup_geom)

    def run(self, stream=None):
        check(lib.wfl_plan_run(self.h, stream))

    def run_timed(self, stream=None):
        """Run and wait; returns the main kernel's duration (ms, HIP events)."""
        ms = C.c_float(0.0)
        check(lib.wfl_plan_run_timed(self.h, stream, C.byref(ms)))
        return ms.value

    def autotune(self, stream=None):
        """Measure the candidate launch configurations and keep the fastest."""
        c = C.c_int(0)
        check(lib.wfl_plan_autotune(self.h, stream, C.byref(c)))
        self.mapping = lib.wfl_plan_mapping(self.h)
        return c.value

    def config(self):
        """(mapping, persistent, wavefronts per block, tasks per loop group)"""
        m, pe, w, c = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        check(lib.wfl_plan_config(self.h, C.byref(m), C.byref(pe), C.byref(w), C.byref(c)))
        return m.value, pe.value, w.value, c.value

    def set_config(self, mapping, persistent, wpb, chunks=1):
        check(lib.wfl_plan_set_config(self.h, mapping, persistent, wpb, chunks))
        self.mapping = mapping

    def graph_run(self, stream=None):
        check(lib.wfl_plan_graph_run(self.h, stream))

    def read(self):
        gs = np.zeros((self.ntup, self.ngroups, 3), np.float64)
        pl = None
        if self.batch.want_perloop:
            pl = np.zeros((self.ntup, self.loops.Nl, 6), np.float64)
        check(lib.wfl_plan_read(self.h, _ptr(gs), _ptr(pl)))
        return gs, pl

    def nonfinite(self):
        v = C.c_longlong(0)
        check(lib.wfl_plan_nonfinite(self.h, C.byref(v)))
        return v.value

    def gsum_device_ptr(self):
        return lib.wfl_plan_gsum_device(self.h)

    def close(self):
        if self.h:
            lib.wfl_plan_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def autotune_nsub(batch: Batch, loops: Loops, tol, nmax=64, max_tuples=256, max_loops=256):
    """Smallest Simpson n (2, 4, 8, ...) whose sample group sums change by at
    most tol (relative) when n doubles; returns (n, relative change)."""
    cb, keep = batch.c()
    n, e = C.c_int(), C.c_double()
    check(lib.wfl_autotune_nsub(C.byref(cb), loops.h, float(tol), int(nmax), int(max_tuples), int(max_loops),
                                C.byref(n), C.byref(e)))
    return n.value, e.value


def eval_batch(batch: Batch, loops: Loops):
    p = Plan(batch, loops)
    try:
        p.run()
        return p.read()
    finally:
        p.close()


def eval_batch_noalloc(batch: Batch, loops: Loops):
    """eval_batch through wfl_eval_batch_noalloc: no device allocation (the
    plan lives in the library's static per-GPU scratch array).  The loop set's
    records must exist already (an eval_batch of the same precision and form
    on these loops builds them)."""
    cb, keep = batch.c()
    ntup = len(batch.tup_geom)
    ngroups = (loops.Nl + batch.group - 1) // batch.group
    gs = np.zeros((ntup, ngroups, 3), np.float64)
    pl = np.zeros((ntup, loops.Nl, 6), np.float64) if batch.want_perloop else None
    check(lib.wfl_eval_batch_noalloc(C.byref(cb), loops.h, _ptr(gs), _ptr(pl)))
    return gs, pl


def group_ev_minus1(gsum3, count):
    g = np.ascontiguousarray(gsum3, np.float64)
    return lib.wfl_group_ev_minus1(_ptr(g), int(count))


def logsimpson(T, f):
    T = np.ascontiguousarray(T, np.float64)
    f = np.ascontiguousarray(f, np.float64)
    return lib.wfl_logsimpson(_ptr(T), _ptr(f), len(T))


def leff_groups(geom: Geometry, fermion, T, cx, cy, gsum, counts, policy=1, estimator=0):
    """L_eff per group; gsum [nT, ngroups, 3] for one (geometry, x_cm); T,
    cx, cy as the batch was created with (nominal, log-spaced T);
    estimator 0 the reference's Richardson combination, 1 the two-level
    extrapolation 2 W_{N/2} - W_{N/4}."""
    T_eff = np.ascontiguousarray(T, np.float64)
    gsum = np.ascontiguousarray(gsum, np.float64)
    counts = np.ascontiguousarray(counts, np.int32)
    ng = gsum.shape[1]
    out = np.zeros(ng, np.float64)
    g = geom.c()
    check(lib.wfl_leff_groups_est(C.byref(g), int(fermion), len(T_eff), _ptr(T_eff), float(cx), float(cy),
                                  _ptr(gsum), _ptr(counts), ng, int(policy), int(estimator), _ptr(out)))
    return out


def cell_action_groups(geom: Geometry, fermion, T, cx, cy, w, gsum, counts, policy=1):
    """Per-group action of one lattice tube: sum_c w_c L_eff(x_c); gsum
    [ncell, nT, ngroups, 3] (the cell points' tuples in T order)."""
    T = np.ascontiguousarray(T, np.float64)
    cx, cy, w = (np.ascontiguousarray(a, np.float64) for a in (cx, cy, w))
    gsum = np.ascontiguousarray(gsum, np.float64)
    counts = np.ascontiguousarray(counts, np.int32)
    ng = gsum.shape[2]
    out = np.zeros(ng, np.float64)
    g = geom.c()
    check(lib.wfl_cell_action_groups(C.byref(g), int(fermion), len(T), _ptr(T), len(cx), _ptr(cx), _ptr(cy), _ptr(w),
                                     _ptr(gsum), _ptr(counts), ng, int(policy), _ptr(out)))
    return out


def radial_action_groups(geom: Geometry, fermion, T, rho, gsum, counts, policy=1):
    """Per-group action of an isolated tube, 2 pi int rho L_eff drho on a
    uniform rho grid; gsum [nrho, nT, ngroups, 3]."""
    T = np.ascontiguousarray(T, np.float64)
    rho = np.ascontiguousarray(rho, np.float64)
    gsum = np.ascontiguousarray(gsum, np.float64)
    counts = np.ascontiguousarray(counts, np.int32)
    ng = gsum.shape[2]
    out = np.zeros(ng, np.float64)
    g = geom.c()
    check(lib.wfl_radial_action_groups(C.byref(g), int(fermion), len(T), _ptr(T), len(rho), _ptr(rho), _ptr(gsum),
                                       _ptr(counts), ng, int(policy), _ptr(out)))
    return out


def interaction_energy(cell_action, iso_action):
    """(mean E_cell - E_iso with E = -action, the spread sqrt(var + var -
    2 cov) of interaction.sage, its standard error)."""
    a = np.ascontiguousarray(cell_action, np.float64)
    b = np.ascontiguousarray(iso_action, np.float64)
    out = np.zeros(3, np.float64)
    check(lib.wfl_interaction_energy(len(a), _ptr(a), _ptr(b), _ptr(out)))
    return tuple(float(v) for v in out)


def eaction(geom: Geometry, loops: "Loops", fermion, mode=MODE_PHYSICAL, precision=MIXED, nsub=2, group=128,
            npoints=20, nT=40):
    """Effective action of an isolated tube with the batched engine
    (EffAct.c's rho rule over Tint.c's T rule): (per-group actions, mean,
    standard error, classical action)."""
    ng = (loops.Nl + group - 1) // group
    out = np.zeros(ng, np.float64)
    r, se, ca = C.c_double(), C.c_double(), C.c_double()
    g = geom.c()
    check(lib.wfl_eaction(C.byref(g), loops.h, int(fermion), mode, precision, nsub, group, npoints, nT, _ptr(out),
                          C.byref(r), C.byref(se), C.byref(ca)))
    return out, r.value, se.value, ca.value


def set_profile(geom: Geometry, nspline=512):
    g = geom.c()
    check(lib.wfl_set_profile(C.byref(g), nspline))


def set_numerics(mode, precision, nsub):
    check(lib.wfl_set_numerics(mode, precision, nsub))
