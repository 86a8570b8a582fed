"""Oracle pins for the profiles (intEV.cu:29-406), the ensemble average
(Integrand.c:325-407), the proper-time closed forms (Integrand.c:20-172),
the spline set-up (periodic.c:45-110) and the log-T quadrature."""
import numpy as np
import pytest
import scipy.integrate as si
import scipy.interpolate as sint
import scipy.special as sp


# ---------------------------------------------------------------- profiles
def test_expint_matches_scipy(orc):
    for x in [-1.0001, -1.5, -3.0, -10.0, -50.0, -300.0]:
        assert orc.lib.orc_expint(x) == pytest.approx(sp.expi(x), rel=1e-13)


def test_fixflux_constants():
    """q1 = int_-1^1 e^{-1/(1-x^2)} dx and q2 = int_0^1 x e^{-1/(1-x^2)} dx
    = (e^{-1} - E1(1))/2: the values the engine uses by default."""
    bump = lambda x: np.exp(-1.0 / (1.0 - x * x)) if abs(x) < 1 else 0.0
    q1 = si.quad(bump, -1, 1, epsabs=1e-14)[0]
    q2 = si.quad(lambda x: x * bump(x), 0, 1, epsabs=1e-14)[0]
    assert q1 == pytest.approx(0.443993816168079, rel=1e-12)
    assert q2 == pytest.approx(0.0742477533879612, rel=1e-12)
    assert q2 == pytest.approx(0.5 * (np.exp(-1) - sp.exp1(1.0)), rel=1e-12)


def test_gaussian_and_simple_profiles_closed_forms(orc):
    l2 = 0.7
    u = np.array([0.0, 1e-6, 0.3, 0.7, 2.0, 30.0])
    g = orc.Geom(orc.GAUSSIAN, l2=l2, quantize=False)
    for v in u:
        assert orc.lib.orc_f(g.ref, v) == pytest.approx(-np.expm1(-v / l2), rel=1e-14, abs=1e-300)
        assert orc.lib.orc_fp(g.ref, v) == pytest.approx(np.exp(-v / l2) / l2, rel=1e-14)
    s = orc.Geom(orc.SMOOTH, l2=l2, quantize=False)
    q = orc.Geom(orc.QUADRATIC, l2=l2, quantize=False)
    st = orc.Geom(orc.STEP, l2=l2, quantize=False)
    for v in u:
        assert orc.lib.orc_f(s.ref, v) == pytest.approx(v / (l2 + v), rel=1e-14, abs=1e-300)
        w = min(v / l2, 1.0)
        assert orc.lib.orc_f(q.ref, v) == pytest.approx(w * (2 - w), rel=1e-14, abs=1e-300)
        assert orc.lib.orc_f(st.ref, v) == pytest.approx(w, rel=1e-14, abs=1e-300)


@pytest.mark.parametrize("prof", ["SMOOTH", "QUADRATIC", "GAUSSIAN", "PERIODIC", "FIXFLUX"])
def test_fp_is_derivative_of_f(orc, prof):
    """f'(rho^2) = df/d(rho^2) wherever the profile is smooth."""
    g = orc.Geom(getattr(orc, prof), F=1.0, l2=0.36, tubedist=2.0, lmin=0.2, quantize=False)
    for u in [0.05, 0.2, 0.31, 0.8, 2.5, 5.0]:
        if prof == "QUADRATIC" and abs(u - 0.36) < 0.05:
            continue
        if prof == "FIXFLUX" and np.sqrt(u) > 0.95:
            # outside rho = a/2 the reference defines f through a fitted chi
            # (intEV.cu:188-190) whose derivative is not fpfixflux
            continue
        h = 1e-6 * max(u, 1.0)
        fd = (orc.lib.orc_f(g.ref, u + h) - orc.lib.orc_f(g.ref, u - h)) / (2 * h)
        assert orc.lib.orc_fp(g.ref, u) == pytest.approx(fd, rel=2e-6, abs=1e-8), u


def test_periodic_profile_rings(orc):
    """flperi (intEV.cu:202-232): f(0) = 0 and f steps up across each ring of
    radius i*a (logistic steps of width lambda)."""
    a = 1.0
    g = orc.Geom(orc.PERIODIC, l2=0.0025, tubedist=a, quantize=False)
    assert abs(orc.lib.orc_f(g.ref, 1e-12)) < 1e-3
    for k in (1, 2, 3):
        below = orc.lib.orc_f(g.ref, ((k - 0.3) * a) ** 2)
        above = orc.lib.orc_f(g.ref, ((k + 0.3) * a) ** 2)
        assert above > below


def test_interp_outside_knots_is_zero(orc):
    """intEV.cu:96-99: interp returns 0 outside (rho2sp[0], rho2sp[N-1])."""
    knots = np.array([0.1, 0.5, 1.0, 2.0], np.float32)
    coefs = np.array([[0.2, 1.0, 0.0, 0.0]] * 4, np.float32)
    assert orc.lib.orc_interp(0.05, knots.ctypes.data, coefs.ctypes.data, 4) == 0.0
    assert orc.lib.orc_interp(2.0, knots.ctypes.data, coefs.ctypes.data, 4) == 0.0
    assert orc.lib.orc_interp(0.5, knots.ctypes.data, coefs.ctypes.data, 4) == pytest.approx(0.2)
    assert orc.lib.orc_interp(0.7, knots.ctypes.data, coefs.ctypes.data, 4) == pytest.approx(0.2 + 0.2)


def test_spline_physical_extension(orc):
    """Physical mode extends the spline table: f = s u below the first knot
    with s = fl32(c0/k0) (c1 when the first knot is 0), meeting the first
    interval's value at k0 to FP32 rounding; f = f(last knot) from the last
    knot on; the table's own cubic in between (reference mode: 0 outside)."""
    rho = (np.arange(12) + 1) * 0.25
    knots = (rho * rho).astype(np.float32)
    coefs, _ = orc.spline(-np.expm1(-knots.astype(np.float64)), knots)
    g = orc.Geom(orc.SPLINE, knots=knots, coefs=coefs)
    k0, kl = float(knots[0]), float(knots[-1])
    s = float(np.float32(float(coefs[0, 0]) / k0))
    assert orc.lib.orc_spline_low_slope(knots.ctypes.data, coefs.ctypes.data) == s
    for u in (1e-6, 0.3 * k0, 0.999 * k0):
        assert orc.lib.orc_f_in_mode(g.ref, orc.MODE_PHYSICAL, u) == pytest.approx(s * u, rel=1e-15)
        assert orc.lib.orc_f_in_mode(g.ref, orc.MODE_REFERENCE, u) == 0.0
    assert s * k0 == pytest.approx(float(coefs[0, 0]), rel=2e-7)            # continuous at k0
    for u in (kl, 1.5 * kl, 100.0):
        assert orc.lib.orc_f_in_mode(g.ref, orc.MODE_PHYSICAL, u) == float(coefs[-1, 0])
    u = 0.5 * (float(knots[4]) + float(knots[5]))
    assert orc.lib.orc_f_in_mode(g.ref, orc.MODE_PHYSICAL, u) == orc.lib.orc_interp(
        u, knots.ctypes.data, coefs.ctypes.data, len(knots))
    zero = np.concatenate([[0.0], knots[1:]]).astype(np.float32)
    cz = coefs.copy()
    assert orc.lib.orc_spline_low_slope(zero.ctypes.data, cz.ctypes.data) == float(cz[0, 1])


# ---------------------------------------------------------------- spline
def test_spline_matches_scipy_clamped(orc, wfl):
    n = 64
    rho = (np.arange(n) + 1) * (6.0 / n)
    knots = (rho * rho).astype(np.float32)
    fli = -np.expm1(-knots.astype(np.float64) / 1.3)
    c4, cd = orc.spline(fli, knots)
    cs = sint.CubicSpline(knots.astype(np.float64), fli, bc_type=((1, 0.0), (1, 0.0)))
    for k, col in enumerate([3, 2, 1, 0]):      # scipy stores the cubic from the highest power down
        assert np.allclose(cd[:-1, k], cs.c[col], rtol=1e-9, atol=1e-12)
    # the product's host spline builder gives the same float4 table
    assert np.array_equal(wfl.spline_build(fli, knots), c4)
    q4, _ = orc.spline(fli, knots, quirk=True)
    assert np.array_equal(wfl.spline_build(fli, knots, quirk=True), q4)
    assert np.allclose(q4, c4, rtol=1e-4, atol=1e-6)


def test_spline_interpolates_gaussian(orc):
    n = 512
    rho = (np.arange(n) + 1) * (20.0 / n)
    knots = (rho * rho).astype(np.float32)
    fli = -np.expm1(-knots.astype(np.float64))
    c4, _ = orc.spline(fli, knots)
    # the zero end slope at the first knot (periodic.c:72) bends the spline
    # within a few knot spacings of the origin; test beyond that
    for u in [0.1, 0.3, 1.0, 2.7, 9.0]:
        assert orc.lib.orc_interp(u, knots.ctypes.data, c4.ctypes.data, n) == pytest.approx(-np.expm1(-u), abs=2e-6)


# ---------------------------------------------------------------- EVMean
def test_wminus1_scalar_and_spinor(orc):
    sc = np.array([0.3, -0.1, 0.25])
    yz = sc[1] + sc[2]
    a = [0.5 * sc[0] + 0.25 * yz, sc[0], 0.5 * yz]
    assert np.allclose(orc.wminus1(1, 0, sc, np.zeros(3)), np.cos(a) - 1.0, rtol=1e-13)
    fe = np.array([0.8, 0.35, 0.45])
    fyz = fe[1] + fe[2]
    b = [0.5 * fe[0] + 0.25 * fyz, fe[0], 0.5 * fyz]
    assert np.allclose(orc.wminus1(1, 1, sc, fe), np.cos(a) * np.cosh(b) - 1.0, rtol=1e-13)
    # the reference's cosh guard: |trunc(arg)| > 600 multiplies by 1e10 instead
    big = np.array([800.0, 0.0, 0.0])
    assert np.allclose(orc.wminus1(0, 1, sc, big), np.cos(a) * 1e10 - 1.0)
    assert not np.all(np.isfinite(orc.wminus1(1, 1, sc, big)))      # physical mode: no guard


def test_evmean_richardson_cancels_discretisation(orc):
    """When the three sub-loop estimates describe the same phase (x = phi,
    y = z = phi) all three W agree and EV = W: the coefficients 8/3, -2, 1/3
    sum to one; they also cancel W = 1 + a/N + b/N^2 corrections."""
    phi = np.linspace(-2, 2, 16)
    ws = np.zeros((16, 4), np.float32)
    ws[:, 0] = phi; ws[:, 1] = phi; ws[:, 2] = phi
    ev = orc.evmean(ws, None, np.arange(16), 0)
    assert ev == pytest.approx(np.mean(np.cos(phi.astype(np.float32).astype(np.float64))), rel=1e-12)
    c = np.array([8 / 3, -2.0, 1 / 3])
    Ns = np.array([1.0, 0.5, 0.25])          # N, N/2, N/4 in units of N
    assert c.sum() == pytest.approx(1.0)
    assert (c / Ns).sum() == pytest.approx(0.0, abs=1e-14)
    assert (c / Ns**2).sum() == pytest.approx(0.0, abs=1e-14)


def test_evmean_spinor_hand_vector(orc, wfl):
    rng = np.random.default_rng(0)
    ws = rng.normal(size=(10, 4)).astype(np.float32)
    wf = np.abs(rng.normal(size=(10, 4))).astype(np.float32)
    WL = np.array([3, 1, 4, 1, 5, 9, 2, 6], np.int32)
    want = 0.0
    e = np.zeros(3)
    for w in WL:
        s, f = ws[w].astype(np.float64), wf[w].astype(np.float64)
        nb, fb = float(np.float32(ws[w, 1] + ws[w, 2])), float(np.float32(wf[w, 1] + wf[w, 2]))
        e += [np.cos(0.5 * s[0] + 0.25 * nb) * np.cosh(0.5 * f[0] + 0.25 * fb),
              np.cos(s[0]) * np.cosh(f[0]), np.cos(0.5 * nb) * np.cosh(0.5 * fb)]
    e /= len(WL)
    want = 8 / 3 * e[0] - 2 * e[1] + 1 / 3 * e[2]
    assert orc.evmean(ws, wf, WL, 1) == pytest.approx(want, rel=1e-13)
    # the product's host EVMean (C) computes the same value bit for bit
    import ctypes as C
    out = C.c_double(0.0)
    wfl.lib.EVMean(C.byref(out), ws.ctypes.data, wf.ctypes.data, len(WL), WL.ctypes.data, 1.0, 1)
    assert out.value == orc.evmean(ws, wf, WL, 1)


# ---------------------------------------------------------------- Tfunc
def test_small_T_series_match_exact_taylor(orc):
    """smTscal / smTferm are the small-T Taylor series of Exact (checked in
    high precision with mpmath): the relative gap shrinks like T^3."""
    import mpmath as mp

    mp.mp.dps = 50
    e = mp.mpf("0.30282212")
    B = 3.0

    def exact(T, ferm):
        T = mp.mpf(T)
        x = e * T * B
        if ferm:
            return mp.e**(-T) / T**3 * (x / mp.tanh(x) - 1 - x * x / 3)
        return mp.e**(-T) / T**3 * (x / mp.sinh(x) - 1 + x * x / 6)

    for ferm, fn in ((0, orc.lib.orc_smTscal), (1, orc.lib.orc_smTferm)):
        gaps = []
        for T in (1e-2, 1e-3):
            ex = float(exact(T, ferm))
            gaps.append(abs(fn(B, T) - ex) / abs(ex))
        assert gaps[1] < gaps[0] / 500
        assert gaps[1] < 1e-8


def test_exact_matches_mpmath(orc):
    import mpmath as mp

    mp.mp.dps = 40
    e = mp.mpf("0.30282212")
    for T, B, ferm in [(0.5, 2.0, 0), (3.0, 1.0, 1), (10.0, 0.3, 0), (1.0, 40.0, 0)]:
        x = e * mp.mpf(T) * B
        if ferm:
            v = mp.e**(-T) / mp.mpf(T)**3 * (x / mp.tanh(x) - 1 - x * x / 3)
        elif x < 50:
            v = mp.e**(-T) / mp.mpf(T)**3 * (x / mp.sinh(x) - 1 + x * x / 6)
        else:
            v = mp.e**(-T) / mp.mpf(T)**3 * (x * x / 6 - 1)
        assert orc.lib.orc_exact(T, B, ferm) == pytest.approx(float(v), rel=1e-11)


def test_igrand_branches(orc):
    g = orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0, quantize=False)
    assert orc.igrand(g, 0, 0.005, 0.0, 0.0, 1.0)[1] == 0     # T < 0.1 and T < 0.01 l2: series
    assert orc.igrand(g, 0, 0.05, 0.0, 0.0, 1.0)[1] == 3      # worldline estimate
    assert orc.igrand(g, 1, 60.0, 0.0, 0.0, 1.0)[1] == 2      # large-T asymptote
    wide = orc.Geom(orc.GAUSSIAN, F=1.0, l2=100.0, quantize=False)
    assert orc.igrand(wide, 0, 0.2, 0.0, 0.0, 1.0)[1] == 1    # 0.1 <= T < 0.01 l2: closed form


def test_logsimpson_exact_for_cubics_in_lnT(orc, wfl):
    for n in (5, 6, 64, 96):
        T = np.logspace(-3, 2, n)
        u = np.log(T)
        p = lambda v: 0.3 - 1.2 * v + 0.05 * v**2 + 0.01 * v**3
        P = lambda v: 0.3 * v - 0.6 * v**2 + 0.05 * v**3 / 3 + 0.0025 * v**4
        f = p(u) / T                      # int f dT = int p(u) du
        want = P(u[-1]) - P(u[0])
        assert orc.logsimpson(T, f) == pytest.approx(want, rel=1e-11)
        assert wfl.logsimpson(T, f) == pytest.approx(want, rel=1e-11)


def test_host_closed_forms_match_oracle(orc, wfl):
    for T, B, ferm in [(0.01, 3.0, 0), (0.7, 1.5, 1), (5.0, 0.2, 0)]:
        assert wfl.lib.Exact(T, B, ferm) == orc.lib.orc_exact(T, B, ferm)
        assert wfl.lib.smTscal(B, T) == orc.lib.orc_smTscal(B, T)
        assert wfl.lib.smTferm(B, T) == orc.lib.orc_smTferm(B, T)


def test_lattice_field_window(orc):
    """The image-window field: one tube's field when the window is a single
    tube or the lattice is wide; periodic in x_cm; and its cell average is
    the flux per cell, eB = 2 F f'(rho^2) integrating to 2 pi F / e over the
    plane (Integrand.c:137), when the window holds all the flux."""
    iso = orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.25)
    one = orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.25, lattice=1.0, images=0)
    B1, r1 = orc.field_at(one, 0.3, -0.2)
    Bi, ri = orc.field_at(iso, 0.3, -0.2)
    assert B1 == pytest.approx(Bi, rel=1e-15) and r1 == pytest.approx(ri, rel=1e-15)
    # x_cm beyond the first cell: the window follows it (nearest tube at (2, -1))
    B2, r2 = orc.field_at(one, 2.3, -1.2)
    assert B2 == pytest.approx(Bi, rel=1e-12) and r2 == pytest.approx(ri, rel=1e-12)
    wide = orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.25, lattice=50.0, images=2)
    assert orc.field_at(wide, 0.3, -0.2)[0] == pytest.approx(Bi, rel=1e-12)
    # cell average: lattice l = 2, lambda = 0.5, 5 x 5 window
    lat = orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.25, lattice=2.0, images=2)
    n = 64
    c = (np.arange(n) + 0.5) * (2.0 / n)
    avg = np.mean([orc.field_at(lat, x, y)[0] for x in c for y in c])
    e = 0.30282212
    assert avg == pytest.approx(2 * np.pi * 1.0 / e / 4.0, rel=1e-6)


def test_fp32_profile_identities():
    """The identities behind the FP32 kernels' forms of two profiles, checked
    against the reference's expressions (intEV.cu:162-174, 202-232) in
    50-digit arithmetic:
    * the periodic sum's central ring plus 1 equals
      [(t/2) tanh(t/2) - log cosh(t/2)] / ln 2 (t = rho/lambda), with
      tanh y = m(m+2)/(m(m+2)+2) and cosh y - 1 = m^2/(2(1+m)), m = e^y - 1;
    * the fixflux phi(x) equals phi(0) + (1/(2 q2)) int_{1-x^2}^1 e^{-1/t} dt,
      phi(0) = 1 - q2*/q2, q2* = (e^-1 - E1(1))/2."""
    import mpmath as mp
    mp.mp.dps = 50
    for t in [1e-3, 0.05, 0.4, 1.0, 1.9]:
        t = mp.mpf(t)
        e = mp.exp(-t)
        ref = 1 - (t * e / (1 + e) + mp.log(1 + e)) / mp.log(2)
        y = t / 2
        m = mp.expm1(y)
        th = m * (m + 2) / (m * (m + 2) + 2)
        lc = mp.log(1 + m * m / (2 * (1 + m)))
        assert abs((y * th - lc) / mp.log(2) - ref) <= 1e-30 * abs(ref)
        assert abs(th - mp.tanh(y)) < 1e-28 and abs(lc - mp.log(mp.cosh(y))) < 1e-28
    q2_star = (mp.exp(-1) - mp.e1(1)) / 2
    assert abs(q2_star - mp.mpf("0.0742477533879612")) < 1e-15
    for q2 in (float(q2_star), 0.0742478, float(np.float32(0.0742477533879612))):
        q2 = mp.mpf(q2)
        phi0 = 1 - q2_star / q2
        for x in [1e-3, 0.1, 0.5, 0.7]:
            x = mp.mpf(x)
            om = 1 - x * x
            ref = 1 - (om * mp.exp(-1 / om) + mp.ei(-1 / om)) / (2 * q2)
            integral = mp.quad(lambda s: mp.exp(-1 / s), [om, 1])
            assert abs(phi0 + integral / (2 * q2) - ref) < 1e-25


def test_eaction_rules_converge_to_lcf_quadrature(orc):
    """The oracle's restatement of EAction's rho and T rules (integrateRho,
    integrateT: the right end twice and the midpoint per interval, rho and
    T = (1 - t)/t with 1/t^2, rhogrand's factor rho, /(4 pi)), run on the
    constant-field closed form Exact at every node, converges to an
    independent adaptive quadrature of the locally-constant-field action
    int rho drho int dT Exact(T, B(rho))/(4 pi) (scipy; x coth x - 1 - x^2/3
    by its series below x = 0.05) for a wide Gaussian tube (F = 50,
    lambda^2 = 100, the testT.c:10-20 tube): within 1e-6 at 80 x 160
    intervals, and the reference's 20 x 40 within 1e-2."""
    from scipy import integrate
    g = orc.Geom(orc.GAUSSIAN, F=50.0, l2=100.0, quantize=False)
    e = 0.30282212

    def gfun(xv):
        if abs(xv) < 0.05:
            x2 = xv * xv
            return x2 * x2 * (-1 / 45 + x2 * (2 / 945 - x2 / 4725))
        return xv / np.tanh(xv) - 1 - xv * xv / 3

    def inner(rho):
        B = orc.lib.orc_magfield(g.ref, rho * rho)
        return integrate.quad(lambda T: np.exp(-T) / T ** 3 * gfun(e * T * B), 0, 60, limit=400, epsabs=0,
                              epsrel=1e-12)[0]

    ref = integrate.quad(lambda r: r * inner(r), 0, 60, limit=400, epsabs=0, epsrel=1e-11)[0] / (4 * np.pi)
    x = np.zeros((2, 8), np.float32)
    fine = orc.eaction_groups(g, x, x, 1, 1, 2, 2, 80, 160, lcf=True)[0]
    coarse = orc.eaction_groups(g, x, x, 1, 1, 2, 2, 20, 40, lcf=True)[0]
    assert abs(fine - ref) <= 1e-6 * abs(ref)
    assert abs(coarse - ref) <= 1e-2 * abs(ref)


def _bump_integral(x):
    """int_-1^x e^{-1/(1-t^2)} dt by adaptive quadrature."""
    bump = lambda t: np.exp(-1.0 / (1.0 - t * t)) if abs(t) < 1 else 0.0
    return si.quad(bump, -1.0, x, epsabs=1e-14, limit=200)[0]


# The fitted part of the fixflux chi (intEV.cu:188-190, fit A; intEVff.cu:
# 170-171, fit B) approximates int_-1^x bump; its error against the integral
# over x in (-1, 1) (801 points), measured once with scipy: rms, mean, max.
CHI_FIT_ACCURACY = {"FIXFLUX": (0.0220420, 0.0117647, 0.0893153),
                    "FIXFLUX_FF": (0.00287788, -0.00253405, 0.00401459)}


@pytest.mark.parametrize("prof", ["FIXFLUX", "FIXFLUX_FF"])
def test_fixflux_outer_branch_vs_quadrature(orc, prof):
    """Outside rho = a/2 the fixflux f is 1 + 0.75 (4 rho^2/a^2 - 1) lml +
    3 n (n - 1) aml + (3 lambda aml/(q1 a)) chi(x), x = 2 (rho - n a)/lambda
    (intEV.cu:247-252), with chi = int_-1^x (t + 2 n a/lambda) bump(t) dt:
    the t bump part in closed form (E1), the rest a fit of int bump.  So f
    must equal f(a/2) plus the integral of f' = fpfixflux (intEV.cu:334-357)
    up to 6 n aml/q1 times the fit's error at x, and the fit must match the
    integral with its recorded accuracy (rms within 1 %, mean within 1 % of
    the rms): a wrong constant, sign or term in either restatement fails
    one of the two."""
    g = orc.Geom(getattr(orc, prof), F=1.0, l2=0.36, tubedist=2.0, lmin=0.2, quantize=False)
    a, lam, lmin, q1 = 2.0, 0.6, 0.2, 0.443993816168079
    aml = (a - lam) / (a - lmin)
    f0 = orc.lib.orc_f(g.ref, (a / 2) ** 2)
    rms0, mean0, max0 = CHI_FIT_ACCURACY[prof]
    fit_err = []
    for rho in np.linspace(a / 2 + 1e-3, 2.9 * a, 120):
        n = np.floor((rho + a / 2) / a)
        x = 2 * (rho - n * a) / lam
        fex = f0 + si.quad(lambda r: orc.lib.orc_fp(g.ref, r * r) * 2 * r, a / 2, rho, limit=400, epsabs=1e-12,
                           points=[k * a + s * lam / 2 for k in (1, 2) for s in (-1, 0, 1)])[0]
        d = orc.lib.orc_f(g.ref, rho * rho) - fex
        coef = 6 * n * aml / q1
        if abs(x) < 0.999:
            fit_err.append(d / coef)
        assert abs(d) <= coef * max0 * 1.02 + 1e-9, (rho, d)
    # the fit itself, against the integral
    xs = np.linspace(-0.9995, 0.9995, 801)
    ex = np.array([_bump_integral(x) for x in xs])
    x2 = xs * xs
    x4 = x2 * x2
    if prof == "FIXFLUX":
        fit = 0.444 / (1 + np.exp(-(3.31 * xs + 5.25 * xs * x2 * np.sin(xs) * np.cos(-0.907 * x2 - 1.29 * x4 * x4))
                                  / np.cos(xs)))
    else:
        fit = 0.218 + 0.393 * xs * np.cosh(0.806 * x4 - 0.696 * x2 + 0.0902) / np.cosh(0.825 * x2 - 0.0234 * xs + 0.375)
    e = fit - ex
    assert abs(np.sqrt(np.mean(e * e)) / rms0 - 1) < 0.01 and abs(e.mean() - mean0) < 0.01 * rms0
    # the oracle's outer branch carries exactly that fit error
    assert len(fit_err) > 20 and max(abs(v) for v in fit_err) <= max0 * 1.02
