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
