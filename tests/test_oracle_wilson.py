"""Oracle pins for getzp1, Idt and WilsonLoop (intEV.cu:409-571) against
closed forms and invariants that a dropped term, wrong sign or wrong index
would break."""
import numpy as np
import pytest


def ref_succ(i, N):
    """getzp1's successor rule, enumerated from intEV.cu:469-500."""
    if i % 2 == 1:
        return 1 if i == N - 1 else i + 2
    if i % 4 == 0:
        return 0 if i == N - 4 else i + 4
    return 2 if i == N - 2 else i + 4


@pytest.mark.parametrize("N", [4, 8, 16, 64])
def test_succ_orbits(orc, N):
    for i in range(N):
        assert orc.lib.orc_succ(i, N) == ref_succ(i, N)
    # three cycles: odd points (N/2), 0 mod 4 (N/4), 2 mod 4 (N/4), each visited in index order
    for start, length, step in [(1, N // 2, 2), (0, N // 4, 4), (2, N // 4, 4)]:
        seen, i = [], start
        for _ in range(length):
            seen.append(i)
            i = orc.lib.orc_succ(i, N)
        assert i == start and seen == list(range(start, N, step))


def shoelace(px, py):
    return 0.5 * np.sum(px * np.roll(py, -1) - np.roll(px, -1) * py)


@pytest.mark.parametrize("mode", [0, 1])
def test_shoelace_constant_field(orc, mode):
    """Step profile with the whole loop inside the tube: g = 1/l2 exactly, so
    each sub-loop phase is (F/l2) * sum z_i x z_succ(i) = (2F/l2) * area."""
    x, y = orc.gen_loops(21, 4, 64)
    F, l2, rtT, cx, cy = 0.7, 100.0, 0.9, 0.3, -0.2
    g = orc.Geom(orc.STEP, F=F, l2=l2, quantize=False)
    for l in range(4):
        sc, fe = orc.wilson_loop(g, mode, 2, 1, x[l], y[l], cx, cy, rtT)
        zx = cx + rtT * x[l].astype(np.float64)
        zy = cy + rtT * y[l].astype(np.float64)
        areas = [shoelace(zx[1::2], zy[1::2]), shoelace(zx[2::4], zy[2::4]), shoelace(zx[0::4], zy[0::4])]
        for k in range(3):
            assert sc[k] == pytest.approx(2 * F / l2 * areas[k], rel=1e-12, abs=1e-15)
        # spin factor in a constant field: 2 F T f' = eB T for every sub-loop
        assert np.allclose(fe, 2 * F * rtT**2 / l2, rtol=1e-12)


def smooth_segment_exact(zx, zy, dx, dy, l2):
    """Closed form of int_0^1 dt / (l2 + |z + t d|^2) (arctan)."""
    A, B, C = zx * zx + zy * zy, zx * dx + zy * dy, dx * dx + dy * dy
    D = C * (l2 + A) - B * B
    s = np.sqrt(D)
    return (np.arctan((C + B) / s) - np.arctan(B / s)) / s


def test_simpson_converges_h4_smooth_profile(orc):
    """Physical mode: the smooth profile has g(u) = 1/(l2 + u); the Simpson
    phase converges to the exact line integral as h^4."""
    N = 64
    x, y = orc.gen_loops(4, 1, N)
    F, l2, rtT, cx, cy = 1.0, 0.5, 1.0, 0.4, 0.1
    g = orc.Geom(orc.SMOOTH, F=F, l2=l2, quantize=False)
    zx = cx + rtT * x[0].astype(np.float64)
    zy = cy + rtT * y[0].astype(np.float64)
    exact = np.zeros(3)
    for i in range(N):
        j = orc.lib.orc_succ(i, N)
        dx, dy = zx[j] - zx[i], zy[j] - zy[i]
        X = zx[i] * dy - zy[i] * dx
        c = 0 if i % 2 else (2 if i % 4 == 0 else 1)
        exact[c] += F * X * smooth_segment_exact(zx[i], zy[i], dx, dy, l2)
    errs = []
    for n in (4, 8, 16, 32):
        sc, _ = orc.wilson_loop(g, orc.MODE_PHYSICAL, n, 0, x[0], y[0], cx, cy, rtT)
        errs.append(np.sum(np.abs(sc - exact)))
    ratios = [errs[k] / errs[k + 1] for k in range(3)]
    assert all(14.0 < r < 18.0 for r in ratios), (errs, ratios)
    sc, _ = orc.wilson_loop(g, orc.MODE_PHYSICAL, 50, 0, x[0], y[0], cx, cy, rtT)
    assert np.sum(np.abs(sc - exact)) < errs[-1] * (32.0 / 50.0) ** 4 * 1.2


def test_simpson_exact_for_quadratic_integrand(orc):
    """Quadratic profile inside the tube: g = (2 - u/l2)/l2 is quadratic in t
    along a segment, so n = 2 Simpson is already exact (n = 50 agrees)."""
    x, y = orc.gen_loops(8, 2, 32)
    g = orc.Geom(orc.QUADRATIC, F=1.3, l2=50.0, quantize=False)
    for l in range(2):
        a, _ = orc.wilson_loop(g, orc.MODE_PHYSICAL, 2, 0, x[l], y[l], 0.5, 0.2, 1.1)
        b, _ = orc.wilson_loop(g, orc.MODE_PHYSICAL, 50, 0, x[l], y[l], 0.5, 0.2, 1.1)
        assert np.allclose(a, b, rtol=1e-12, atol=1e-15)


def reference_idt_smooth(A, B, C, l2, n):
    """intEV.cu:409-459 for g(u) = f(u)/u with the smooth profile f = u/(l2+u)."""
    A, B, C = max(A, 1e-8), max(B, 1e-8), max(C, 1e-8)
    A1 = max(A + 2 * B + C, 1e-8)
    g = lambda u: 1.0 / (l2 + u)
    s = g(A) + g(A1)
    h = 1.0 / n
    for k in range(1, n):
        t = k * h
        u = max(A + 2 * B * t + C * t * t, 1e-10)
        s += (4.0 if k % 2 else 2.0) * g(u)
    return s * h / 3.0


def test_reference_clamps_on_inward_segments(orc):
    """Reference mode clamps the linear coefficient B = z.(z'-z) at 1e-8: on a
    closed loop some segments move inward (B < 0) and are integrated along
    A + 2e-8 t + C t^2 instead of the true path."""
    x, y = orc.gen_loops(9, 1, 16)
    F, l2, rtT, cx, cy = 1.0, 0.3, 0.8, 0.2, 0.0
    g = orc.Geom(orc.SMOOTH, F=F, l2=l2, quantize=False)
    zx = cx + rtT * x[0].astype(np.float64)
    zy = cy + rtT * y[0].astype(np.float64)
    want = np.zeros(3)
    nneg = 0
    for i in range(16):
        j = orc.lib.orc_succ(i, 16)
        dx, dy = zx[j] - zx[i], zy[j] - zy[i]
        A, B, C = zx[i] ** 2 + zy[i] ** 2, zx[i] * dx + zy[i] * dy, dx * dx + dy * dy
        nneg += B < 0
        c = 0 if i % 2 else (2 if i % 4 == 0 else 1)
        want[c] += F * (zx[i] * zy[j] - zy[i] * zx[j]) * reference_idt_smooth(A, B, C, l2, 50)
    assert nneg >= 3
    got, _ = orc.wilson_loop(g, orc.MODE_REFERENCE, 50, 0, x[0], y[0], cx, cy, rtT)
    phys, _ = orc.wilson_loop(g, orc.MODE_PHYSICAL, 50, 0, x[0], y[0], cx, cy, rtT)
    assert np.allclose(got, want, rtol=1e-12)
    assert np.max(np.abs(got - phys)) > 1e-4 * np.max(np.abs(phys))


@pytest.mark.parametrize("mode", [0, 1])
def test_rotation_and_reflection(orc, mode):
    """At x_cm = 0 the phase is rotation invariant; y -> -y flips the phases
    and leaves the spin factor unchanged."""
    x, y = orc.gen_loops(13, 3, 64)
    g = orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.8, quantize=False)
    th = 0.7
    for l in range(3):
        sc, fe = orc.wilson_loop(g, mode, 2, 1, x[l], y[l], 0.0, 0.0, 1.2)
        xr = (np.cos(th) * x[l] - np.sin(th) * y[l]).astype(np.float32)
        yr = (np.sin(th) * x[l] + np.cos(th) * y[l]).astype(np.float32)
        sr, fr = orc.wilson_loop(g, mode, 2, 1, xr, yr, 0.0, 0.0, 1.2)
        assert np.allclose(sr, sc, rtol=2e-6, atol=1e-9) and np.allclose(fr, fe, rtol=2e-6)
        sm, fm = orc.wilson_loop(g, mode, 2, 1, x[l], -y[l], 0.0, 0.0, 1.2)
        assert np.allclose(sm, -sc, rtol=1e-12, atol=1e-15) and np.allclose(fm, fe, rtol=1e-12)


def test_scalar_phase_depends_on_sqrtT_w_only(orc):
    x, y = orc.gen_loops(2, 2, 32)
    g = orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0)
    for mode in (0, 1):
        a, _ = orc.wilson_loop(g, mode, 2, 0, x[0], y[0], 0.4, 0.0, 0.5)
        b, _ = orc.wilson_loop(g, mode, 2, 0, 2 * x[0], 2 * y[0], 0.4, 0.0, 0.25)
        assert np.array_equal(a, b)


def test_lattice_limits(orc):
    x, y = orc.gen_loops(6, 2, 64)
    iso = orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.5)
    one = orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.5, lattice=2.0, images=0)   # K = 1
    far = orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.5, lattice=200.0, images=1)
    win = orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.5, lattice=2.0, images=2)
    for l in range(2):
        a, fa = orc.wilson_loop(iso, 1, 2, 1, x[l], y[l], 0.3, 0.1, 0.9)
        b, fb = orc.wilson_loop(one, 1, 2, 1, x[l], y[l], 0.3, 0.1, 0.9)
        assert np.allclose(a, b, rtol=1e-14) and np.allclose(fa, fb, rtol=1e-14)
        c, fc = orc.wilson_loop(far, 1, 2, 1, x[l], y[l], 0.3, 0.1, 0.9)
        assert np.allclose(c, a, rtol=1e-7, atol=1e-9) and np.allclose(fc, fa, rtol=1e-12)
        # translation by a lattice vector moves the window with x_cm
        d, fd = orc.wilson_loop(win, 1, 2, 1, x[l], y[l], 0.3, 0.1, 0.9)
        e, fe = orc.wilson_loop(win, 1, 2, 1, x[l], y[l], 0.3 + 2.0, 0.1 - 4.0, 0.9)
        assert np.allclose(d, e, rtol=1e-10, atol=1e-13) and np.allclose(fd, fe, rtol=1e-10)
