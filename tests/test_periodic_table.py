"""The spline profile's table (periodic.c:45-245) and the host lookup
(intEV.cu:64-107), checked on the CPU.

getspline tabulates f(rho) = int_0^rho (2/l2) t exp(-sin^2(pi t/a)/l2) dt
(the ODE of periodic.c:18-24, whose right-hand side does not depend on y) at
rho_i = (i + 1) 20/N.  Pins: scipy's adaptive quadrature at sample knots;
the closed form at whole periods, where the integrand's symmetry
p(t) = p(a - t) of p(t) = exp(-sin^2(pi t/a)/l2) gives
int_0^{na} t p = (na)^2/2 * (1/a) int_0^a p and
(1/a) int_0^a p = e^{-1/(2 l2)} I0(1/(2 l2)) (the generating function of the
modified Bessel functions), so f(na) = (na)^2/l2 * ive(0, 1/(2 l2));
flspline against wfl_spline_build in its reference-arithmetic mode; and
testspline against the same ratio formed here with interp emulated in FP32.
"""
import numpy as np
import pytest

scipy = pytest.importorskip("scipy")
from scipy import integrate, special  # noqa: E402


def ulp32(x):
    x = np.abs(np.asarray(x, np.float32))
    return (np.nextafter(x, np.float32(np.inf)) - x).astype(np.float64)


def df(l2, a):
    return lambda t: 2.0 / l2 * t * np.exp(-np.sin(np.pi * t / a) ** 2 / l2)


@pytest.mark.parametrize("l2,a", [(1.0, 1.0), (0.25, 1.25), (0.05, 1.0), (4.0, 2.0)])
def test_getspline_values_against_quadrature(wfl, l2, a):
    n = 512
    knots, coefs = wfl.getspline(l2, a, n)
    rho = (np.arange(n) + 1) * (20.0 / n)
    assert np.array_equal(knots, (rho * rho).astype(np.float32))
    idx = [0, 1, 10, 57, 100, 255, 300, 511]
    q = np.array([integrate.quad(df(l2, a), 0.0, rho[i], limit=4000, epsabs=0.0, epsrel=1e-13)[0] for i in idx])
    c0 = coefs[idx, 0].astype(np.float64)
    # the table holds fl32(f): within half an FP32 ulp of the quadrature (plus its 1e-13)
    assert np.all(np.abs(c0 - q) <= 0.5 * ulp32(coefs[idx, 0]) + 1e-12 * q), (c0, q)


@pytest.mark.parametrize("l2", [0.1, 0.5, 2.0])
def test_getspline_whole_periods_closed_form(wfl, l2):
    a, n = 1.25, 512               # rho = k a falls on knot 32 k - 1
    knots, coefs = wfl.getspline(l2, a, n)
    k = np.arange(1, 17)
    i = 32 * k - 1
    assert np.allclose(np.sqrt(knots[i].astype(np.float64)), k * a, rtol=1e-7)
    want = (k * a) ** 2 / l2 * special.ive(0, 1.0 / (2.0 * l2))
    got = coefs[i, 0].astype(np.float64)
    assert np.all(np.abs(got - want) <= 0.5 * ulp32(coefs[i, 0]) + 1e-13 * want), (got, want)


def test_flspline_is_the_reference_arithmetic(wfl):
    n = 200
    rho = (np.arange(n) + 1) * 0.05
    knots = (rho * rho).astype(np.float32)
    fli = -np.expm1(-knots.astype(np.float64) / 0.7)
    rows = wfl.flspline(fli, knots)
    ref = wfl.spline_build(fli, knots, quirk=True)
    assert np.array_equal(rows[:-1], ref[:-1])
    assert rows[-1, 0] == ref[-1, 0] and rows[-1, 2] == ref[-1, 2]


def interp32(rho2, knots, coefs):
    """intEV.cu:64-107 in FP32, operation by operation."""
    rho2 = np.float32(rho2)
    n = len(knots)
    if not (rho2 < knots[-1] and rho2 > knots[0]):
        return np.float32(0.0)
    lo, hi = 0, n - 1
    while hi - lo > 1:
        mid = (hi + lo) // 2
        if rho2 >= knots[mid]:
            lo = mid
        else:
            hi = mid
    c = coefs[lo].astype(np.float32)
    d = np.float32(rho2 - knots[lo])
    inner = np.float32(c[2] + np.float32(d * c[3]))
    mid = np.float32(c[1] + np.float32(d * inner))
    return np.float32(c[0] + np.float32(d * mid))


@pytest.mark.parametrize("l2,a", [(1.0, 1.0), (0.05, 1.0)])
def test_testspline_and_interp(wfl, l2, a):
    wfl.set_profile(wfl.Geometry(kind=wfl.GAUSSIAN), 512)   # interp's table length: the drop-in's Nspline
    knots, coefs = wfl.getspline(l2, a, 512)
    for r2 in (1.0, 0.37, 7.3):
        s = interp32(r2, knots, coefs)
        assert wfl.interp(r2, knots, coefs) == s
        f = integrate.quad(df(l2, a), 0.0, np.sqrt(np.float64(np.float32(r2))), limit=4000, epsabs=0.0,
                           epsrel=1e-13)[0]
        assert abs(wfl.testspline(r2, l2, a, knots, coefs) - (f - float(s)) / f) < 1e-12
    # outside the knots the reference's lookup returns 0
    assert wfl.interp(float(knots[0]) * 0.5, knots, coefs) == 0.0
    assert wfl.interp(float(knots[-1]) * 1.01, knots, coefs) == 0.0


def test_periodic_table_keeps_rows_about_their_start(wfl):
    """f of the periodic table grows like rho^2 (the flux of every tube
    within rho), nearly linear in u far out, where its cubics cancel about
    u = 0: the direct-index rows stay expanded about their start for both
    species (the GPU parity test on this table checks the plan's form)."""
    knots, coefs = wfl.getspline(0.3, 1.1, 512)
    assert wfl.spline_uniform_form(knots, coefs, 400.0) == 1
    assert wfl.spline_uniform_form(knots, coefs, 400.0, spinor=1) == 1


def test_table_builders_reject_bad_arguments(wfl):
    with pytest.raises(wfl.WflError, match="flspline"):
        wfl.flspline(np.zeros(2), np.array([0.1, 0.2], np.float32))
    with pytest.raises(wfl.WflError, match="getspline"):
        wfl.getspline(-1.0, 1.0, 64)
    with pytest.raises(wfl.WflError, match="knots must increase"):
        wfl.flspline(np.zeros(4), np.array([0.1, 0.3, 0.2, 0.4], np.float32))
