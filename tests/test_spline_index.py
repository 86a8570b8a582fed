"""The direct index of the uniform-grid spline node, checked on the CPU.

For tables whose knots lie on r_j = (m + j) h (the periodic.c:129-151 layout
has m = 1), the FP32 kernels take row i = floor(sqrt(u)/h) from one FMA:
t = fl32(sqrt(u) * (1/h) + (2^23 - 1/2)); for y = sqrt(u)/h >= 1/4 the
exact sum lies in [2^23, 2^24), where FP32 steps by 1, so t = 2^23 +
floor(y) (a tie, y on a knot, may give the row below); y < 1/4 gives
t = 2^23 - 1/2, one bit pattern below 2^23, the guard row -1.  Row i's
expansion point is fl32(i h)^2 with fl32(i h) = fma(t, h, -2^23 h).

The FMA is emulated exactly: the product of two FP32 values and the sum
with 2^23 - 1/2 fit the 64-bit significand of x87 long double, rounded
once to FP32 afterwards.  Pins: the row equals the interval an exact search
finds, except within rounding of a knot, where the neighbour is taken; the
guard row exactly for y < 1/4; the expansion point is i h to FP32 rounding.
"""
import numpy as np
import pytest

LD = np.longdouble


def fl32(x):
    return np.asarray(x, dtype=np.float32)


def fma32(a, b, c):
    """fl32(a * b + c) for FP32 a, b, c: exact in long double, one rounding."""
    return fl32(LD(1) * np.asarray(a, LD) * np.asarray(b, LD) + np.asarray(c, LD))


def kernel_row(u, inv_h):
    t = fma32(np.sqrt(fl32(u)), inv_h, np.float32(8388607.5))
    bits = t.view(np.uint32).astype(np.int64)
    return bits - 0x4B000000, t


@pytest.fixture(scope="module")
def table():
    if np.finfo(LD).nmant < 60:
        pytest.skip("long double lacks the 64-bit significand the exact FMA emulation needs")
    n, rho_max = 512, 20.0
    h = rho_max / n
    rho = (np.arange(n) + 1) * h                      # periodic.c: first knot at rho = h (m = 1)
    knots = (rho * rho).astype(np.float32)
    inv_h = np.float32(1.0 / h)
    return knots, inv_h, np.float32(h)


def test_row_is_the_interval(table):
    knots, inv_h, _ = table
    rng = np.random.default_rng(7)
    rmax = float(np.sqrt(knots[-1])) * 1.3
    u = fl32(rng.uniform(0.0, rmax, 400000) ** 2)
    u = u[u > 0]
    row, _ = kernel_row(u, inv_h)
    # exact: the row whose r-interval [i h, (i + 1) h) holds sqrt(u) in the
    # table's own knots (row i = interval i - 1 for m = 1; row 0 below k0)
    exact = np.searchsorted(knots.astype(np.float64), u.astype(np.float64), side="right")
    # past the last knot the rows go on (copies of the constant row): there
    # the row is the grid index itself
    past = u >= knots[-1]
    exact[past] = np.floor(np.sqrt(u[past].astype(LD)) * LD(len(knots)) / LD(20.0)).astype(np.int64)
    row = np.maximum(row, 0)          # the guard row holds row 0's polynomial
    ok = row == exact
    # where they differ, u must sit within FP32 rounding of the knot between
    near = np.zeros_like(ok)
    bad = ~ok
    kb = np.maximum(np.minimum(row[bad], exact[bad]), 0)
    kb = np.minimum(kb, len(knots) - 1)
    rel = np.abs(u[bad].astype(np.float64) - knots[kb].astype(np.float64)) / knots[kb].astype(np.float64)
    near[bad] = (np.abs(row[bad] - exact[bad]) == 1) & (rel < 4e-7)
    assert np.all(ok | near), (u[~(ok | near)][:5], row[~(ok | near)][:5], exact[~(ok | near)][:5])
    assert ok.mean() > 0.999


def test_guard_row_below_a_quarter(table):
    _, inv_h, h = table
    y = np.array([1e-30, 1e-12, 1e-6, 0.1, 0.2, 0.2499, 0.2501, 0.3, 0.49, 0.51, 0.99, 1.01])
    u = fl32((y * float(h)) ** 2)
    row, _ = kernel_row(u, inv_h)
    yk = np.sqrt(u.astype(np.float64)) * float(inv_h)
    want = np.where(yk < 0.25, -1, np.floor(yk)).astype(np.int64)
    # rounding of sqrt and of y * h may move a sample across 1/4 or 1
    edge = (np.abs(yk - 0.25) < 1e-6) | (np.abs(yk - 1.0) < 1e-6)
    assert np.all((row == want) | edge), (y, row, want)
    assert row[0] == -1 and row[3] == -1 and row[7] == 0 and row[-1] == 1


def test_expansion_point_from_the_row(table):
    knots, inv_h, h = table
    u = fl32(np.linspace(1e-4, float(knots[-1]), 20001))
    row, t = kernel_row(u, inv_h)
    hoff = np.float32(-8388608.0) * h                  # exact: a power-of-two multiple
    rk = fma32(t, h, hoff)
    want = fl32(np.where(row < 0, -0.5, row).astype(np.float64) * float(h))
    assert np.array_equal(rk, want)
    # the point is the row's interval start to FP32 rounding, so
    # dx = u - rk^2 stays within the interval's width, or a few ulp of u
    # below it where a node on a knot takes the row above (the guard row's
    # point, (h/2)^2, lies above its nodes: dx < 0 there)
    rows = row >= 0
    uu = u.astype(np.float64)[rows]
    dx = uu - rk.astype(np.float64)[rows] ** 2
    width = (2.0 * row[rows] + 1.0) * float(h) ** 2
    assert np.all(dx > -4e-7 * uu) and np.all(dx < width * (1 + 1e-5))


def gaussian_table(wfl, l2, n=512, rho_max=20.0):
    rho = (np.arange(n) + 1) * (rho_max / n)
    knots = (rho * rho).astype(np.float32)
    return knots, wfl.spline_build(-np.expm1(-knots.astype(np.float64) / l2), knots)


def test_origin_form_gate(wfl):
    """Rows expanded about u = 0 (form 2) for tables well conditioned in that
    form, rows about their start (form 1) otherwise, the exact search (0)
    off the grid.  A Gaussian tube of lambda^2 <= 10 on the 512-knot grid to
    rho = 20 is smooth on the scale of its rows; at lambda^2 = 100 the last
    rows, and a linear f everywhere far out, cancel in the expansion about 0
    (the FP32 coefficients' c3 noise times kappa^3).  Spinor plans also need
    f' well conditioned at every u, which a Gaussian's f' ~ e^-u is not far
    out (against d1 ~ e^-u (1 + u + u^2/2)): form 1."""
    for l2 in (0.01, 0.1, 1.0, 10.0):
        assert wfl.spline_uniform_form(*gaussian_table(wfl, l2), 400.0) == 2, l2
        assert wfl.spline_uniform_form(*gaussian_table(wfl, l2), 400.0, spinor=1) == 1, l2
    assert wfl.spline_uniform_form(*gaussian_table(wfl, 100.0), 400.0) == 1
    knots, _ = gaussian_table(wfl, 1.0)
    assert wfl.spline_uniform_form(knots, wfl.spline_build(knots.astype(np.float64) / 3.0, knots), 400.0) == 1
    off = knots.copy()
    off[3] = np.float32((np.sqrt(float(off[3])) + 1e-3 * 20.0 / 512) ** 2)
    assert wfl.spline_uniform_form(off, wfl.spline_build(-np.expm1(-off.astype(np.float64)), off), 400.0) == 0
    # more rows than the LDS layout holds: no direct index at all
    assert wfl.spline_uniform_form(*gaussian_table(wfl, 1.0), 1e4) == 0


def kernel_g_origin(u, rows, inv_h):
    """The origin-form node as the FP32 kernel forms it, FMAs exact with one
    rounding; the reciprocal correctly rounded (the hardware's is within one
    ulp of it)."""
    row, _ = kernel_row(u, inv_h)
    c = rows[row + 1]
    rcp = fl32(1.0 / u.astype(np.float64))
    q = fma32(u, fma32(u, c[:, 3], c[:, 2]), c[:, 1])
    return fma32(c[:, 0], rcp, q)


def test_origin_form_accuracy(wfl):
    """The library's rows of form 2 (expanded about u = 0) for the 512-knot
    lambda = 1 table, evaluated as the FP32 kernel evaluates them, against
    the table's cubic in FP64 divided by u: the error of g at the level of
    FP32 rounding (rms 2e-8), with no coherent offset over ranges of rows
    (mean within 2e-9 on each of [k0, 1), [1, 4), [4, 16), [16, 64)), which
    the host's compensation of the coefficient roundings and its spreading
    of d0's rounding over rows are for: plain rounding of the expansion
    about 0 leaves offsets of 1e-8.  Rows past the last knot hold the
    constant; below the first knot, the line."""
    if np.finfo(LD).nmant < 60:
        pytest.skip("long double lacks the 64-bit significand the exact FMA emulation needs")
    knots, coefs = gaussian_table(wfl, 1.0)
    form, rows = wfl.spline_uniform_form(knots, coefs, 500.0, rows=True)
    assert form == 2 and len(rows) > 520
    inv_h = np.float32(512 / 20.0)
    k = knots.astype(np.float64)
    c = coefs.astype(np.float64).reshape(-1, 4)
    rng = np.random.default_rng(11)
    u = fl32(rng.uniform(0.0, 21.0, 400000) ** 2)
    u = u[u > 1e-6]
    u64 = u.astype(np.float64)
    j = np.searchsorted(k, u64, side="right") - 1
    jj = np.clip(j, 0, len(k) - 1)
    x = u64 - k[jj]
    f = c[jj, 0] + x * (c[jj, 1] + x * (c[jj, 2] + x * c[jj, 3]))
    f = np.where(j < 0, u64 * c[0, 0] / k[0], np.where(j >= len(k) - 1, c[-1, 0], f))
    ref = f / u64
    rel = kernel_g_origin(u, rows, inv_h).astype(np.float64) / ref - 1.0
    print(f"max {np.abs(rel).max():.2e} rms {np.sqrt((rel ** 2).mean()):.2e} mean {rel.mean():+.2e}")
    assert np.abs(rel).max() < 4e-7 and np.sqrt((rel ** 2).mean()) < 3.5e-8
    for lo, hi in ((float(k[0]), 1.0), (1.0, 4.0), (4.0, 16.0), (16.0, 64.0)):
        m = (u64 >= lo) & (u64 < hi)
        print(f"  [{lo:.4f}, {hi}): rms {np.sqrt((rel[m] ** 2).mean()):.2e} mean {rel[m].mean():+.2e}")
        assert abs(rel[m].mean()) < 2e-9


def test_row_start_form_accuracy(wfl):
    """The library's rows of form 1 (each cubic expanded about its row's
    start fl32(i h)^2, the guard row about (h/2)^2), for the lambda^2 = 100
    table that keeps them, evaluated as the FP32 kernel does (row from the
    FMA's bits, r = fma(t, h, -2^23 h), dx = fma(-r, r, u), Horner, times
    the reciprocal) against the table's cubic in FP64 divided by u: within
    FP32 rounding, no offset coherent over rows, and f' (the spinor's
    derivative from the Horner partials) at FP32 rounding too (rms below
    1e-7 relative) wherever it is above 1e-3 of its largest value."""
    if np.finfo(LD).nmant < 60:
        pytest.skip("long double lacks the 64-bit significand the exact FMA emulation needs")
    knots, coefs = gaussian_table(wfl, 100.0)
    form, rows = wfl.spline_uniform_form(knots, coefs, 500.0, rows=True)
    assert form == 1 and len(rows) > 520
    h = np.float32(20.0 / 512)
    inv_h = np.float32(512 / 20.0)
    k = knots.astype(np.float64)
    c = coefs.astype(np.float64).reshape(-1, 4)
    rng = np.random.default_rng(13)
    u = fl32(rng.uniform(float(np.sqrt(k[0])), 19.9, 300000) ** 2)
    u64 = u.astype(np.float64)
    j = np.clip(np.searchsorted(k, u64, side="right") - 1, 0, len(k) - 2)
    x = u64 - k[j]
    f = c[j, 0] + x * (c[j, 1] + x * (c[j, 2] + x * c[j, 3]))
    fp = c[j, 1] + x * (2.0 * c[j, 2] + 3.0 * c[j, 3] * x)
    row, t = kernel_row(u, inv_h)
    cc = rows[row + 1]
    rk = fma32(t, h, np.float32(-8388608.0) * h)
    dx = fma32(-rk, rk, u)
    i = fma32(dx, cc[:, 3], cc[:, 2])
    q = fma32(dx, i, cc[:, 1])
    fk = fma32(dx, q, cc[:, 0])
    g = (fk * fl32(1.0 / u64)).astype(np.float64)
    rel = g / (f / u64) - 1.0
    print(f"g: max {np.abs(rel).max():.2e} rms {np.sqrt((rel ** 2).mean()):.2e} mean {rel.mean():+.2e}")
    assert np.abs(rel).max() < 3e-7 and abs(rel.mean()) < 2e-9
    d = fma32(dx, fma32(dx, cc[:, 3], i), q).astype(np.float64)
    big = np.abs(fp) > 1e-3 * np.abs(fp).max()
    relp = d[big] / fp[big] - 1.0
    print(f"f': max {np.abs(relp).max():.2e} rms {np.sqrt((relp ** 2).mean()):.2e}")
    assert np.abs(relp).max() < 1e-6 and np.sqrt((relp ** 2).mean()) < 1e-7


def test_origin_form_on_random_tables(wfl):
    """Random smooth tables (f a sum of Gaussian tubes of random widths and
    fluxes, or of tanh steps, 128 knots to rho = 8): every row the library
    writes is finite, and wherever the gate takes the expansion about u = 0
    the FP32 node (emulated exactly) stays within 16 (the gate's factor)
    times 8 roundings of f, relative (f floored at 1e-6 of its largest
    value), of the table's cubic; the rows about their start meet the same
    bound."""
    if np.finfo(LD).nmant < 60:
        pytest.skip("long double lacks the 64-bit significand the exact FMA emulation needs")
    rng = np.random.default_rng(17)
    n, rho_max = 128, 8.0
    h = np.float32(rho_max / n)
    inv_h = np.float32(n / rho_max)
    rho = (np.arange(n) + 1) * (rho_max / n)
    knots = (rho * rho).astype(np.float32)
    k = knots.astype(np.float64)
    seen = set()
    for trial in range(24):
        m = int(rng.integers(1, 4))
        w = rng.uniform(0.05, 6.0, m)
        a = rng.uniform(0.2, 2.0, m)
        if trial % 2:
            c0 = rng.uniform(0.5, 20.0, m)
            fli = sum(a[i] * (np.tanh((k - c0[i]) / w[i]) + np.tanh(c0[i] / w[i])) for i in range(m))
        else:
            fli = sum(-a[i] * np.expm1(-k / w[i]) for i in range(m))
        coefs = wfl.spline_build(fli, knots)
        form, rows = wfl.spline_uniform_form(knots, coefs, 100.0, rows=True)
        assert form in (1, 2)
        seen.add(form)
        assert np.all(np.isfinite(rows))
        c = coefs.astype(np.float64).reshape(-1, 4)
        u = fl32(rng.uniform(float(np.sqrt(k[0])), float(np.sqrt(k[-1])) * 0.999, 20000) ** 2)
        u64 = u.astype(np.float64)
        j = np.clip(np.searchsorted(k, u64, side="right") - 1, 0, n - 2)
        x = u64 - k[j]
        f = c[j, 0] + x * (c[j, 1] + x * (c[j, 2] + x * c[j, 3]))
        row, t = kernel_row(u, inv_h)
        cc = rows[row + 1]
        rcp = fl32(1.0 / u64)
        if form == 2:
            g = fma32(cc[:, 0], rcp, fma32(u, fma32(u, cc[:, 3], cc[:, 2]), cc[:, 1])).astype(np.float64)
        else:
            rk = fma32(t, h, np.float32(-8388608.0) * h)
            dx = fma32(-rk, rk, u)
            g = (fma32(dx, fma32(dx, fma32(dx, cc[:, 3], cc[:, 2]), cc[:, 1]), cc[:, 0]) * rcp).astype(np.float64)
        # relative to f, floored at 1e-6 of the table's largest value (a
        # tanh step far out leaves f at underflow level near the axis)
        rel = np.abs(g * u64 - f) / np.maximum(np.abs(f), 1e-6 * np.abs(fli).max())
        assert rel.max() < 16 * 8 * 2.0 ** -24, (trial, form, rel.max())
        if form == 2:                   # measured: at most 3.4e-7 over these tables
            assert rel.max() < 1e-6, (trial, rel.max())
    assert seen == {1, 2}
