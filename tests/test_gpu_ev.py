"""The drop-in boundary on the GPU: EV(), Tfunc() and integrateT() called the
way the reference's C drivers call them (struct Wparams with device
buffers), checked against the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NL, N = 256, 256


class Hip:
    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]

    def malloc(self, n):
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), n) == 0
        return p

    def h2d(self, dst, src):
        assert self.lib.hipMemcpy(dst, src.ctypes.data, src.nbytes, 1) == 0

    def free(self, p):
        self.lib.hipFree(p)


@pytest.fixture(scope="module")
def setup(gpu):
    hip = Hip()
    x, y = gpu.gen_loops(8, NL, N)
    aos = np.zeros((NL * N, 4), np.float32)
    aos[:, 0] = x.ravel()
    aos[:, 1] = y.ravel()
    d_wl = hip.malloc(aos.nbytes)
    hip.h2d(d_wl, aos)
    d_s = hip.malloc(NL * 16)
    d_f = hip.malloc(NL * 16)
    hs = np.zeros((NL, 4), np.float32)
    hf = np.zeros((NL, 4), np.float32)
    yield gpu, hip, x, y, d_wl, d_s, d_f, hs, hf
    for p in (d_wl, d_s, d_f):
        hip.free(p)
    gpu.lib.wfl_ev_reset()


def make_params(wfl, d_wl, d_s, d_f, hs, hf, F, l2, rho, fermion, ng=2):
    p = wfl.Wparams()
    p.Nl, p.Nppl, p.F, p.l2 = NL, N, F, l2
    p.xcm.x, p.xcm.y = rho, 0.0
    p.Wsscal_h, p.Wsscal_d = hs.ctypes.data, d_s.value
    p.Wsferm_h, p.Wsferm_d = hf.ctypes.data, d_f.value
    p.worldlines = d_wl.value
    p.nBlocks, p.nThreads, p.ng, p.fermion = 1, 256, ng, fermion
    return p


def oracle_float4(orc, x, y, F, l2, rho, T, fermion, mode=0, nsub=50):
    g = orc.Geom(orc.GAUSSIAN, F=F, l2=l2, quantize="float")   # F, l2 narrowed to float (intEV.cu:124)
    rt = float(np.float32(np.sqrt(T)))
    ws = np.zeros((NL, 4), np.float32)
    wf = np.zeros((NL, 4), np.float32)
    for l in range(NL):
        sc, fe = orc.wilson_loop(g, mode, nsub, fermion, x[l], y[l], float(np.float32(rho)), 0.0, rt)
        ws[l, :3] = sc
        wf[l, :3] = fe
    return ws, wf


@pytest.mark.parametrize("fermion", [0, 1])
def test_ev_matches_oracle_and_reuses_results(setup, orc, fermion):
    wfl, hip, x, y, d_wl, d_s, d_f, hs, hf = setup
    wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN))
    wfl.set_numerics(wfl.MODE_REFERENCE, wfl.FP64, 50)
    wfl.lib.wfl_ev_reset()
    F, l2, rho, T = 1.0, 1.0, 0.5, 0.7
    p = make_params(wfl, d_wl, d_s, d_f, hs, hf, F, l2, rho, fermion)
    WL = np.arange(128, dtype=np.int32)
    print(f"EV fermion={fermion}", flush=True)
    ev = wfl.lib.EV(T, C.byref(p), WL.ctypes.data)
    ws, wf = oracle_float4(orc, x, y, F, l2, rho, T, fermion)
    # per-loop float4 outputs agree to the last float bit or one ulp
    assert np.allclose(hs[:, :3], ws[:, :3], rtol=3e-7, atol=1e-12)
    if fermion:
        assert np.allclose(hf[:, :3], wf[:, :3], rtol=3e-7)
    assert ev == orc.evmean(hs, hf, WL, fermion)          # EV = EVMean over WLlist
    assert ev == pytest.approx(orc.evmean(ws, wf, WL, fermion), rel=1e-6)
    assert wfl.lib.wfl_ev_launches() == 1
    WL2 = np.arange(128, 256, dtype=np.int32)
    keep_s, keep_f = hs.copy(), hf.copy()
    hs[:] = -7.0                                            # EV overwrites the host buffers every call
    hf[:] = -7.0
    ev2 = wfl.lib.EV(T, C.byref(p), WL2.ctypes.data)        # second group: no relaunch
    assert wfl.lib.wfl_ev_launches() == 1
    assert np.array_equal(hs, keep_s) and (not fermion or np.array_equal(hf, keep_f))
    assert ev2 == orc.evmean(hs, hf, WL2, fermion)
    wfl.lib.EV(1.3, C.byref(p), WL.ctypes.data)            # new T: relaunch
    assert wfl.lib.wfl_ev_launches() == 2


def test_ev_mixed_precision(setup, orc):
    wfl, hip, x, y, d_wl, d_s, d_f, hs, hf = setup
    wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN))
    wfl.set_numerics(wfl.MODE_REFERENCE, wfl.MIXED, 50)
    p = make_params(wfl, d_wl, d_s, d_f, hs, hf, 1.0, 1.0, 0.3, 1)
    WL = np.arange(128, dtype=np.int32)
    print("EV mixed", flush=True)
    ev = wfl.lib.EV(0.4, C.byref(p), WL.ctypes.data)
    ws, wf = oracle_float4(orc, x, y, 1.0, 1.0, 0.3, 0.4, 1)
    assert np.allclose(hs[:, :3], ws[:, :3], rtol=2e-5, atol=1e-7)
    assert ev == pytest.approx(orc.evmean(ws, wf, WL, 1), rel=1e-5)
    wfl.set_numerics(wfl.MODE_REFERENCE, wfl.FP64, 50)


@pytest.mark.parametrize("mode,prec,nsub,tol", [(0, 0, 50, 3e-7), (1, 1, 2, 2e-5)])
def test_ev_spline_profile(setup, orc, mode, prec, nsub, tol):
    """profile == spline through EV(): the caller's device tables
    (params.flcoefs, params.rho2sp, Nspline knots as periodic.c:129-151 lays
    them out), in the reference's numerics (FP64, n = 50; exact interval
    search) and in mixed physical mode (the direct-index node on these
    uniform knots), against the oracle's bisection."""
    wfl, hip, x, y, d_wl, d_s, d_f, hs, hf = setup
    n = 512
    rho = (np.arange(n) + 1) * (20.0 / n)
    knots = (rho * rho).astype(np.float32)
    coefs = wfl.spline_build(-np.expm1(-knots.astype(np.float64)), knots)
    d_kn, d_cf = hip.malloc(knots.nbytes), hip.malloc(coefs.nbytes)
    hip.h2d(d_kn, knots)
    hip.h2d(d_cf, np.ascontiguousarray(coefs, np.float32))
    wfl.set_profile(wfl.Geometry(kind=wfl.SPLINE, F=1.0, l2=1.0, a=1.0, spline_table=0), n)
    wfl.set_numerics(mode, prec, nsub)
    wfl.lib.wfl_ev_reset()
    WL = np.arange(128, dtype=np.int32)
    g = orc.Geom(orc.SPLINE, F=1.0, l2=1.0, tubedist=1.0, knots=knots, coefs=coefs, quantize="float")
    try:
        for fermion, rho_cm, T in ((0, 0.4, 0.3), (1, 1.1, 0.05)):
            p = make_params(wfl, d_wl, d_s, d_f, hs, hf, 1.0, 1.0, rho_cm, fermion)
            p.flcoefs, p.rho2sp = d_cf.value, d_kn.value
            print(f"EV spline mode={mode} prec={prec} fermion={fermion}", flush=True)
            ev = wfl.lib.EV(T, C.byref(p), WL.ctypes.data)
            assert np.isfinite(ev)
            rt = float(np.float32(np.sqrt(T)))
            for l in range(0, NL, 7):
                sc, fe = orc.wilson_loop(g, mode, nsub, fermion, x[l], y[l], float(np.float32(rho_cm)), 0.0, rt)
                scale = max(np.max(np.abs(sc)), 1e-300)
                assert np.max(np.abs(hs[l, :3] - sc)) <= tol * scale, (l, hs[l, :3], sc)
                if fermion:
                    scf = max(np.max(np.abs(fe)), 1e-300)
                    assert np.max(np.abs(hf[l, :3] - fe)) <= tol * scf, (l, hf[l, :3], fe)
    finally:
        hip.free(d_kn)
        hip.free(d_cf)
        wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN))
        wfl.set_numerics(wfl.MODE_REFERENCE, wfl.FP64, 50)
        wfl.lib.wfl_ev_reset()


@pytest.mark.parametrize("kind", ["STEP", "SMOOTH", "QUADRATIC", "PERIODIC", "FIXFLUX", "FIXFLUX_FF"])
def test_ev_other_profiles(setup, orc, kind):
    """The remaining profiles through EV() (the intT.h profile switch,
    tubedist a = 2, lmin = 0.2) in the reference's numerics, FP64, against
    the oracle: per-loop float4 outputs to their float rounding."""
    wfl, hip, x, y, d_wl, d_s, d_f, hs, hf = setup
    wfl.set_profile(wfl.Geometry(kind=getattr(wfl, kind), F=0.8, l2=0.64, a=2.0, lam_min=0.2))
    wfl.set_numerics(wfl.MODE_REFERENCE, wfl.FP64, 50)
    wfl.lib.wfl_ev_reset()
    WL = np.arange(128, dtype=np.int32)
    g = orc.Geom(getattr(orc, kind), F=0.8, l2=0.64, tubedist=2.0, lmin=0.2, quantize="float")
    try:
        for fermion, rho_cm, T in ((0, 0.3, 0.2), (1, 1.4, 0.6)):
            p = make_params(wfl, d_wl, d_s, d_f, hs, hf, 0.8, 0.64, rho_cm, fermion)
            print(f"EV {kind} fermion={fermion}", flush=True)
            ev = wfl.lib.EV(T, C.byref(p), WL.ctypes.data)
            assert np.isfinite(ev)
            rt = float(np.float32(np.sqrt(T)))
            for l in range(0, NL, 9):
                sc, fe = orc.wilson_loop(g, 0, 50, fermion, x[l], y[l], float(np.float32(rho_cm)), 0.0, rt)
                assert np.max(np.abs(hs[l, :3] - sc)) <= 3e-7 * max(np.max(np.abs(sc)), 1e-300), (l, hs[l, :3], sc)
                if fermion:
                    assert np.max(np.abs(hf[l, :3] - fe)) <= 3e-7 * max(np.max(np.abs(fe)), 1e-300), (l, hf[l], fe)
    finally:
        wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN))
        wfl.lib.wfl_ev_reset()


def test_rho_drivers_compose(setup):
    """The rho-integral drivers of EffAct.c:15-135 built on integrateT:
    rhogrand(rho) = rho * integrateT(40) at x_cm = (rho, 0); transformRho(t)
    = rhogrand((1 - t)/t) / t^2; integrateRhotoa(n) = h/3 [rg(0) + rg(a/2)
    + sum_k w_k rg(k h)], h = a/(2n), with the reference's weights (2 for
    odd k, 4 for even k, EffAct.c:113-126)."""
    wfl, hip, x, y, d_wl, d_s, d_f, hs, hf = setup
    ng = 2

    def arr(ptr):
        return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(C.c_double)), shape=(ng,)).copy()

    a = 1.5
    wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN, F=1.0, l2=1.0, a=a))
    wfl.set_numerics(wfl.MODE_REFERENCE, wfl.MIXED, 50)
    wfl.lib.wfl_ev_reset()
    wfl.lib.wfl_set_verbosity(0)
    try:
        p = make_params(wfl, d_wl, d_s, d_f, hs, hf, 1.0, 1.0, 0.0, 0, ng=ng)
        print("rho drivers", flush=True)

        def rg(rho):
            return arr(wfl.lib.rhogrand(rho, C.byref(p)))

        r = rg(0.3)
        p.xcm.x = 0.3
        assert np.array_equal(r, 0.3 * arr(wfl.lib.integrateT(40, C.byref(p))))
        assert np.allclose(arr(wfl.lib.transformRho(0.4, C.byref(p))), rg(1.5) / 0.16, rtol=1e-14)
        n = 3
        h = a / (2 * n)
        want = rg(0.0) + rg(a / 2)
        for k in range(1, n + 1):
            want = want + (2.0 if k % 2 else 4.0) * rg(h * k)
        got = arr(wfl.lib.integrateRhotoa(n, C.byref(p)))
        assert np.allclose(got, h / 3 * want, rtol=1e-13), (got, h / 3 * want)
    finally:
        wfl.lib.wfl_set_verbosity(1)
        wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN))
        wfl.set_numerics(wfl.MODE_REFERENCE, wfl.FP64, 50)
        wfl.lib.wfl_ev_reset()


def test_integrateT_matches_oracle(setup, orc):
    """integrateT (Tint.c:100-175) through Tfunc and EV against the oracle's
    restatement of the same t-grid Simpson rule and branch policy."""
    wfl, hip, x, y, d_wl, d_s, d_f, hs, hf = setup
    F, l2, rho, fermion, n = 1.0, 1.0, 0.4, 0, 4
    wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN))
    wfl.set_numerics(wfl.MODE_REFERENCE, wfl.FP64, 50)
    p = make_params(wfl, d_wl, d_s, d_f, hs, hf, F, l2, rho, fermion, ng=2)
    print("integrateT", flush=True)
    res = wfl.lib.integrateT(n, C.byref(p))
    got = np.ctypeslib.as_array(C.cast(res, C.POINTER(C.c_double)), shape=(2,)).copy()
    g = orc.Geom(orc.GAUSSIAN, F=F, l2=l2, quantize=False)
    rho_f = float(np.float32(rho))
    rho2 = float(np.float32(np.float32(rho) * np.float32(rho)))

    def tfunc(T):
        out = np.zeros(2)
        for q in range(2):
            v, br = orc.igrand(g, fermion, T, rho_f, rho2, 1.0)
            if br == 3:
                ws, wf = oracle_float4(orc, x, y, F, l2, rho, T, fermion)
                ev = orc.evmean(ws, wf, np.arange(128 * q, 128 * q + 128, dtype=np.int32), fermion)
                v, _ = orc.igrand(g, fermion, T, rho_f, rho2, ev)
            out[q] = v
        return out

    h = 1.0 / n
    want = np.zeros(2)
    for inx in range(n):
        t1, th = h * (inx + 1.0), h * (inx + 0.5)
        a = tfunc((1 - t1) / t1) / t1 / t1
        b = tfunc((1 - th) / th) / th / th
        want += h / 6.0 * (2 * a + 4 * b)
    assert np.allclose(got, want, rtol=1e-6, atol=1e-12)


def test_batched_eaction_matches_drop_in(setup, orc):
    """The effective action two ways on the same loops: EAction driving EV()
    node by node (EffAct.c:265-306 -> integrateRho -> integrateT -> Tfunc ->
    EV), and wfl_eaction with all (rho, T) worldline nodes in one plan.  The
    drop-in narrows F and l2 to float where the batch uses its own FP32
    parameter grid, a difference of ~1e-7 in the inputs."""
    wfl, hip, x, y, d_wl, d_s, d_f, hs, hf = setup
    F, l2, npts = 1.0, 1.0, 2
    wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN, F=F, l2=l2))
    wfl.set_numerics(wfl.MODE_REFERENCE, wfl.FP64, 50)
    wfl.lib.wfl_ev_reset()
    p = make_params(wfl, d_wl, d_s, d_f, hs, hf, F, l2, 0.0, 0, ng=2)
    print("EAction drop-in", flush=True)
    res = wfl.lib.EAction(npts, C.byref(p))
    drop = np.ctypeslib.as_array(C.cast(res, C.POINTER(C.c_double)), shape=(2,)).copy()
    L = wfl.Loops(x, y)
    print("EAction batched", flush=True)
    groups, mean, se, ca = wfl.eaction(wfl.Geometry(kind=wfl.GAUSSIAN, F=F, l2=l2), L, 0, wfl.MODE_REFERENCE,
                                       wfl.FP64, 50, 128, npts, 40)
    L.close()
    assert ca == pytest.approx(wfl.lib.ClassAction(F, l2), rel=1e-12)
    assert mean == pytest.approx(groups.mean(), rel=1e-12)
    assert np.all(np.isfinite(groups))
    assert np.allclose(groups, drop, rtol=1e-4, atol=0.0)
