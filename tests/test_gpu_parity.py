"""GPU parity: the HIP kernels (through the C ABI) against the FP64 oracle.

Bars (per-loop sub-loop phases and spin sums, and per-group sums of W - 1):
  FP64 kernels   relative error <= 1e-10 of the largest magnitude in the case
  mixed kernels  relative error <= 2e-5 per loop and <= 1e-5 in group sums
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NL, N = 96, 256


@pytest.fixture(scope="module")
def loops(gpu):
    x, y = gpu.gen_loops(3, NL, N)
    L = gpu.Loops(x, y)
    yield x, y, L
    L.close()


def tuples(cxs, Ts, cy=0.0, geom=0):
    tg, tcx, tcy, tT = [], [], [], []
    for c in cxs:
        for t in Ts:
            tg.append(geom); tcx.append(c); tcy.append(cy); tT.append(t)
    return np.array(tg), np.array(tcx), np.array(tcy), np.array(tT)


def relerr(a, b):
    scale = max(np.max(np.abs(b)), 1e-300)
    return np.max(np.abs(a - b)) / scale


def run_case(wfl, orc, loops, geoms, ogeoms, mode, nsub, spinor, prec, mapping, tup, splines=(), group=32,
             on_plan=None):
    x, y, L = loops
    tg, tcx, tcy, tT = tup
    b = wfl.Batch(list(geoms), tg, tcx, tcy, tT, mode=mode, precision=prec, nsub=nsub, spinor=spinor,
                  group=group, splines=list(splines), want_perloop=True, mapping=mapping)
    print(f"case mode={mode} nsub={nsub} spinor={spinor} prec={prec} mapping={mapping} ntup={len(tg)}", flush=True)
    plan = wfl.Plan(b, L)
    try:
        plan.run()
        gs, pl = plan.read()
        nonfinite = plan.nonfinite()
        if on_plan is not None:
            on_plan(plan)
    finally:
        plan.close()
    rgs, rpl = orc.eval_batch(ogeoms, mode, nsub, spinor, x, y, group, tg, tcx, tcy, tT)
    assert np.all(np.isfinite(gs)) and np.all(np.isfinite(pl))
    assert nonfinite == 0      # the kernels' own count of non-finite W - 1
    return gs, pl, rgs, rpl


GAUSS_T = [0.01, 0.1, 0.7, 3.0]
GAUSS_RHO = [0.0, 0.6, 1.7]


@pytest.mark.parametrize("mapping", [1, 2, 3])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("spinor", [0, 1])
def test_gaussian_fp64(gpu, orc, loops, mapping, mode, spinor):
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=1.0)]
    og = [orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0)]
    gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, mode, 2, spinor, gpu.FP64, mapping,
                                tuples(GAUSS_RHO, GAUSS_T))
    e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
    print(f"  perloop {e1:.2e} gsum {e2:.2e}")
    assert e1 < 1e-10 and e2 < 1e-10


@pytest.mark.parametrize("mapping", [1, 2, 3])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("spinor", [0, 1])
def test_gaussian_mixed(gpu, orc, loops, mapping, mode, spinor):
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=1.0)]
    og = [orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0)]
    gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, mode, 2, spinor, gpu.MIXED, mapping,
                                tuples(GAUSS_RHO, GAUSS_T))
    e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
    print(f"  perloop {e1:.2e} gsum {e2:.2e}")
    assert e1 < 2e-5 and e2 < 1e-5


@pytest.mark.parametrize("spinor", [0, 1])
def test_axis_form(gpu, orc, loops, spinor):
    """Mixed physical n = 2 batches whose tuples all have x_cm on the x axis
    take the axis form (rho^2 and the cross products from loop-only
    records); the same tuples with cy = 1e-30 (the same problem to FP32)
    take the general form.  Both meet the mixed bars against the oracle, and
    the axis form's error stays within a small multiple of the general
    form's: the Gaussian tube (with small loops in the core, the accurate
    node), a uniform-grid spline table, and the run-time profile kernels
    (smooth), under each mapping.  A lattice geometry, a tuple off the axis,
    n = 4 and reference mode keep the general form."""
    knots, coefs = gaussian_spline(gpu)
    cases = [
        ([gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=1.0)], [orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0)], ()),
        ([gpu.Geometry(kind=gpu.SPLINE, F=1.0, l2=1.0, a=1.0, spline_table=0)],
         [orc.Geom(orc.SPLINE, F=1.0, l2=1.0, tubedist=1.0, knots=knots, coefs=coefs)], [(knots, coefs)]),
        ([gpu.Geometry(kind=gpu.SMOOTH, F=0.7, l2=0.5)], [orc.Geom(orc.SMOOTH, F=0.7, l2=0.5)], ()),
    ]
    for geoms, og, spl in cases:
        for mapping in (1, 2, 3):
            res = {}
            for cy in (0.0, 1e-30):
                seen = []
                gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, spinor, gpu.MIXED, mapping,
                                            tuples(GAUSS_RHO, GAUSS_T, cy=cy), splines=spl,
                                            on_plan=lambda p: seen.append(p.axis_form))
                res[cy] = (relerr(pl, rpl), relerr(gs, rgs), seen[0])
            (e1a, e2a, ax), (e1g, e2g, gen) = res[0.0], res[1e-30]
            print(f"  profile {geoms[0].kind} mapping {mapping}: axis perloop {e1a:.2e} gsum {e2a:.2e}; "
                  f"general perloop {e1g:.2e} gsum {e2g:.2e}")
            assert ax == 1 and gen == 0
            assert e1a < 2e-5 and e2a < 1e-5 and e1g < 2e-5 and e2g < 1e-5
            assert e1a < 4.0 * e1g + 1e-7 and e2a < 4.0 * e2g + 1e-7
    # batches that keep the general form
    g0 = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=1.0)]
    o0 = [orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0)]
    tg, tcx, tcy, tT = tuples(GAUSS_RHO, GAUSS_T)
    tcy_off = tcy.copy()
    tcy_off[-1] = 0.3
    lat = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=0.25, lattice=1.5, images=1)]
    olat = [orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.25, lattice=1.5, images=1)]
    for geoms, og, mode, nsub, tup in ((g0, o0, 1, 2, (tg, tcx, tcy_off, tT)), (g0, o0, 1, 4, (tg, tcx, tcy, tT)),
                                       (g0, o0, 0, 2, (tg, tcx, tcy, tT)), (lat, olat, 1, 2, (tg, tcx, tcy, tT))):
        seen = []
        gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, mode, nsub, spinor, gpu.MIXED, 1, tup,
                                    on_plan=lambda p: seen.append(p.axis_form))
        assert seen == [0]
        assert relerr(pl, rpl) < 2e-5 and relerr(gs, rgs) < 1e-5


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 2e-5)])
@pytest.mark.parametrize("mode", [0, 1])
def test_nsub50_and_multi_geometry(gpu, orc, loops, prec, tol, mode):
    """n = 50 (the reference's Simpson count) and several geometries per launch."""
    specs = [(0.1, 0.25), (1.0, 1.0), (10.0, 4.0)]
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=F, l2=l2) for F, l2 in specs]
    og = [orc.Geom(orc.GAUSSIAN, F=F, l2=l2) for F, l2 in specs]
    parts = [tuples([0.0, 1.1], [0.05, 0.9], geom=g) for g in range(3)]
    tup = tuple(np.concatenate([p[k] for p in parts]) for k in range(4))
    for mapping in (1, 2, 3):
        gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, mode, 50, 1, prec, mapping, tup)
        e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
        print(f"  mapping {mapping}: perloop {e1:.2e} gsum {e2:.2e}")
        assert e1 < tol and e2 < tol


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 2e-5)])
def test_lattice_images(gpu, orc, loops, prec, tol):
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=0.25, lattice=1.5, images=1)]
    og = [orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.25, lattice=1.5, images=1)]
    tg, tcx, tcy, tT = tuples([0.2, 0.75, 1.3], [0.05, 0.8, 4.0])
    tcy = tcx * 0.5
    for spinor in (0, 1):
        gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, spinor, prec, 1, (tg, tcx, tcy, tT))
        e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
        print(f"  spinor {spinor}: perloop {e1:.2e} gsum {e2:.2e}")
        assert e1 < tol and e2 < tol


def gaussian_spline(wfl, n=512, rho_max=20.0, l2=1.0, rng=None):
    """Knots rho^2 of a Gaussian tube: the periodic.c:129-151 layout (rho
    uniform: every interval within two of its bucket start), or, with rng,
    knots at random positions with a tight cluster (the lookup's scan)."""
    if rng is None:
        rho = (np.arange(n) + 1) * (rho_max / n)
        knots = (rho * rho).astype(np.float32)
    else:   # random positions and a cluster of 8 knots 2e-4 apart at rho^2 = 0.7 rho_max^2
        rho = rng.uniform(0.0, rho_max, n)
        c = 0.7 * rho_max * rho_max
        knots = np.unique(np.concatenate([(rho * rho), c + 2e-4 * np.arange(8)]).astype(np.float32))
    fli = -np.expm1(-knots.astype(np.float64) / l2)
    return knots, wfl.spline_build(fli, knots)


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 5e-5)])
@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("mapping", [1, 2])
def test_spline_profile(gpu, orc, loops, prec, tol, mode, mapping):
    """Spline-only batch: the tables are staged in LDS."""
    knots, coefs = gaussian_spline(gpu)
    geoms = [gpu.Geometry(kind=gpu.SPLINE, F=1.0, l2=1.0, a=1.0, spline_table=0)]
    og = [orc.Geom(orc.SPLINE, F=1.0, l2=1.0, tubedist=1.0, knots=knots, coefs=coefs)]
    for spinor in (0, 1):
        gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, mode, 2, spinor, prec, mapping,
                                    tuples([0.0, 1.2], [0.02, 0.5, 3.0]), splines=[(knots, coefs)])
        e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
        print(f"  spinor {spinor}: perloop {e1:.2e} gsum {e2:.2e}")
        assert e1 < tol and e2 < tol


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 2e-5)])
@pytest.mark.parametrize("mode", [0, 1])
def test_spline_periodic_table(gpu, orc, loops, prec, tol, mode):
    """The table that periodic.c's getspline builds for the spline profile
    (tubes exp(-sin^2(pi t/a)/l2) in every direction, l2 = 0.3, a = 1.1, 512
    knots to rho = 20; here by the library's getspline), scalar and spinor,
    against the oracle.  f grows like rho^2 far out, so the expansion about
    u = 0 cancels and the mixed physical-mode plan keeps the rows about their
    start (form 1); reference mode and FP64 take the exact search."""
    knots, coefs = gpu.getspline(0.3, 1.1, 512)
    geoms = [gpu.Geometry(kind=gpu.SPLINE, F=1.0, l2=0.3, a=1.1, spline_table=0)]
    og = [orc.Geom(orc.SPLINE, F=1.0, l2=0.3, tubedist=1.1, knots=knots, coefs=coefs)]
    for spinor in (0, 1):
        seen = []
        gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, mode, 2, spinor, prec, 1,
                                    tuples([0.0, 0.45, 2.0], [0.01, 0.3, 3.0]), splines=[(knots, coefs)],
                                    on_plan=lambda p: seen.append(p.uniform_splines))
        e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
        print(f"  spinor {spinor}: uniform {seen[0]} perloop {e1:.2e} gsum {e2:.2e}")
        assert seen == [1 if (prec == gpu.MIXED and mode == 1) else 0]
        assert e1 < tol and e2 < tol / 2


@pytest.mark.parametrize("mapping", [1, 2])
@pytest.mark.parametrize("nsub", [2, 4])
def test_spline_uniform_grid(gpu, orc, loops, mapping, nsub):
    """Mixed physical-mode spline-only plans: two tables with knots uniform
    in rho (512 knots to rho 20 and 64 to rho 8) take the direct-index
    kernels (scalar: rows expanded about u = 0, both tables being well
    conditioned in that form; spinor: rows about their start, f' being
    small against the expansion's terms far out); adding a table with
    irregular knots sends the plan to the exact interval search.  Both match the oracle (which
    bisects); x_cm on the
    axis puts nodes below the first knot (the line row) and large T beyond
    the last (the constant row)."""
    k1, c1 = gaussian_spline(gpu)
    k2, c2 = gaussian_spline(gpu, n=64, rho_max=8.0, l2=2.0)
    k3, c3 = gaussian_spline(gpu, n=300, rho_max=12.0, l2=0.5, rng=np.random.default_rng(3))
    parts = [tuples([0.0, 0.9], [0.004, 0.3, 2.0, 40.0], geom=g) for g in range(3)]
    tabs = [(k1, c1), (k2, c2), (k3, c3)]
    geoms = [gpu.Geometry(kind=gpu.SPLINE, F=1.0, l2=1.0, a=1.0, spline_table=0),
             gpu.Geometry(kind=gpu.SPLINE, F=2.0, l2=2.0, a=1.0, spline_table=1),
             gpu.Geometry(kind=gpu.SPLINE, F=0.7, l2=0.5, a=1.0, spline_table=2)]
    og = [orc.Geom(orc.SPLINE, F=1.0, l2=1.0, tubedist=1.0, knots=k1, coefs=c1),
          orc.Geom(orc.SPLINE, F=2.0, l2=2.0, tubedist=1.0, knots=k2, coefs=c2),
          orc.Geom(orc.SPLINE, F=0.7, l2=0.5, tubedist=1.0, knots=k3, coefs=c3)]
    for ngeom, want_uni in ((2, [2, 1]), (3, [0, 0])):
        tup = tuple(np.concatenate([p[k] for p in parts[:ngeom]]) for k in range(4))
        seen = []
        for spinor in (0, 1):
            gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms[:ngeom], og[:ngeom], 1, nsub, spinor, 1, mapping,
                                        tup, splines=tabs[:ngeom], on_plan=lambda p: seen.append(p.uniform_splines))
            e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
            print(f"  tables {ngeom} spinor {spinor}: perloop {e1:.2e} gsum {e2:.2e}")
            assert e1 < 2e-5 and e2 < 1e-5
        assert seen == want_uni


@pytest.mark.parametrize("spinor", [0, 1])
def test_spline_uniform_grid_at_knots(gpu, orc, loops, spinor):
    """Loops centred on knots, so that nodes fall on both sides of them,
    where the FP32 square root can pick the neighbouring row.
    * Small loops across the first knots of a lambda = 1 table, where the
      line row meets the first cubic with a continuous value only: within
      1e-5 of the oracle's bisection (scalar: rows expanded about u = 0).
    * Loops across the last knot (the constant row) of a lambda = 10 table,
      whose f still varies there.  At rho = 20 the symmetric gauge sums the
      loop's small flux from terms ~1e4 larger, so FP32 kernels of either
      lookup sit near 1e-4; the direct-index kernel is held to the exact
      search's own error on the same table with one knot moved by 1e-3 of
      the spacing (not uniform, so the plan takes the exact search).  Its
      last rows cancel when expanded about u = 0, so the plan expands each
      row about its start."""
    def case(knots, coefs, l2, idx, Ts):
        r = np.sqrt(knots.astype(np.float64))
        geoms = [gpu.Geometry(kind=gpu.SPLINE, F=1.0, l2=l2, a=1.0, spline_table=0)]
        og = [orc.Geom(orc.SPLINE, F=1.0, l2=l2, tubedist=1.0, knots=knots, coefs=coefs)]
        tg, tcx, tcy, tT = tuples([r[j] for j in idx], Ts)
        seen = []
        gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, spinor, 1, 1, (tg, tcx, tcy, tT),
                                    splines=[(knots, coefs)], on_plan=lambda p: seen.append(p.uniform_splines))
        errs = np.array([relerr(pl[t], rpl[t]) for t in range(len(tg))])
        for t in range(len(tg)):
            print(f"  l2 {l2} uniform {seen[0]} x_cm {tcx[t]:.4f} T {tT[t]:.0e}: perloop {errs[t]:.2e}")
        return seen[0], errs

    k1, c1 = gaussian_spline(gpu, l2=1.0)
    uni, e = case(k1, c1, 1.0, [0, 1, 2, 5], [1e-4, 1e-3])
    assert uni == (1 if spinor else 2) and np.all(e < 1e-5)
    k2, c2 = gaussian_spline(gpu, l2=100.0)
    r2 = np.sqrt(k2.astype(np.float64))
    k3 = k2.copy()
    k3[3] = np.float32((r2[3] + 1e-3 * (r2[1] - r2[0])) ** 2)
    c3 = gpu.spline_build(-np.expm1(-k3.astype(np.float64) / 100.0), k3)
    uni, e_uni = case(k2, c2, 100.0, [508, 511], [0.25, 1.0])
    exact, e_exact = case(k3, c3, 100.0, [508, 511], [0.25, 1.0])
    assert uni == 1 and exact == 0
    assert np.all(e_uni < 2.0 * e_exact + 1e-6), (e_uni, e_exact)


@pytest.mark.parametrize("spinor", [0, 1])
def test_spline_small_tables(gpu, orc, loops, spinor):
    """The smallest tables the spline set-up builds (3 and 5 knots), a
    uniform table whose first knot is rho = 0 (no line rows; its guard row
    is interval 0), one whose first knot lies three spacings out (rows below
    it repeat the line row), one evenly spaced but starting 3.17 spacings
    out (off the grid that the direct index needs: the exact search), and one
    with a knot off the grid (the exact search): mixed physical mode,
    against the oracle.  The 3-knot table and the one from 0.9 keep rows
    expanded about their start, the 5- and 9-knot tables take rows expanded
    about u = 0 when scalar."""
    off_grid = np.arange(8) * 0.3 + 0.9
    off_grid[4] += 0.05
    cases = [((np.arange(3) + 1) * 0.7, 1), ((np.arange(5) + 1) * 0.5, 2 - spinor), (np.arange(9) * 0.4, 2 - spinor),
             (np.arange(8) * 0.3 + 0.9, 1), (np.arange(8) * 0.3 + 0.95, 0), (off_grid, 0)]
    for rho_k, want_uni in cases:
        knots = (rho_k * rho_k).astype(np.float32)
        coefs = gpu.spline_build(-np.expm1(-knots.astype(np.float64)), knots)
        geoms = [gpu.Geometry(kind=gpu.SPLINE, F=1.0, l2=1.0, a=1.0, spline_table=0)]
        og = [orc.Geom(orc.SPLINE, F=1.0, l2=1.0, tubedist=1.0, knots=knots, coefs=coefs)]
        seen = []
        gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, spinor, 1, 1,
                                    tuples([0.0, 0.6, 2.5], [0.01, 0.3, 3.0]), splines=[(knots, coefs)],
                                    on_plan=lambda p: seen.append(p.uniform_splines))
        e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
        print(f"  {len(knots)} knots from rho {rho_k[0]:.1f}: uniform {seen[0]} perloop {e1:.2e} gsum {e2:.2e}")
        assert seen == [want_uni]
        assert e1 < 2e-5 and e2 < 1e-5


@pytest.mark.parametrize("spinor", [0, 1])
def test_spline_uniform_grid_rows_past_the_table(gpu, orc, loops, spinor):
    """The direct-index kernels read one row per node with no clamp: the
    plan repeats the constant row past the last knot as far as its tuples'
    nodes can reach (the largest loop radius times sqrt(T) beyond x_cm).  At
    T = 150 the nodes reach rho ~ 25 against the table's 20, which the
    repeated rows cover; at T = 1e4 they would need more rows than the LDS
    layout holds, and the plan takes the exact interval search.  Both
    against the oracle."""
    knots, coefs = gaussian_spline(gpu)
    geoms = [gpu.Geometry(kind=gpu.SPLINE, F=1.0, l2=1.0, a=1.0, spline_table=0)]
    og = [orc.Geom(orc.SPLINE, F=1.0, l2=1.0, tubedist=1.0, knots=knots, coefs=coefs)]
    x, y, _ = loops
    rmax = float(np.sqrt(x.astype(np.float64) ** 2 + y.astype(np.float64) ** 2).max())
    for Ts, want_uni in (([0.3, 150.0], 2 - spinor), ([0.3, 1.0e4], 0)):
        assert (0.9 + np.sqrt(Ts[1]) * rmax > 20.0) and (want_uni == 0 or 0.9 + np.sqrt(Ts[1]) * rmax < 30.0)
        seen = []
        gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, spinor, 1, 1, tuples([0.0, 0.9], Ts),
                                    splines=[(knots, coefs)], on_plan=lambda p: seen.append(p.uniform_splines))
        e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
        print(f"  T {Ts}: uniform {seen[0]} perloop {e1:.2e} gsum {e2:.2e}")
        assert seen == [want_uni]
        assert e1 < 2e-5 and e2 < 1e-5


@pytest.mark.parametrize("images", [1, 2])
@pytest.mark.parametrize("spinor", [0, 1])
def test_spline_uniform_grid_on_a_lattice(gpu, orc, loops, spinor, images):
    """A uniform-grid spline tube on a square lattice (spacing 1.5, 3 x 3
    and 5 x 5 image windows, whose tubes the mixed kernels sum in the inner
    loop): the rows past the last knot must reach the farthest image tube of
    the window plus the loops' extent, which the plan's bound takes from the
    window; x_cm off the axes inside the cell.  Mixed physical mode (the
    direct-index kernels) and FP64 (the exact search), against the oracle."""
    knots, coefs = gaussian_spline(gpu, n=128, rho_max=6.0, l2=0.25)
    geoms = [gpu.Geometry(kind=gpu.SPLINE, F=1.0, l2=0.25, a=1.0, lattice=1.5, images=images, spline_table=0)]
    og = [orc.Geom(orc.SPLINE, F=1.0, l2=0.25, tubedist=1.0, lattice=1.5, images=images, knots=knots,
                   coefs=coefs)]
    tg, tcx, tcy, tT = tuples([0.1, 0.6, 1.2], [0.01, 0.2, 1.5])
    tcy = 0.4 * tcx + 0.05
    for prec, bars in ((gpu.MIXED, (2e-5, 1e-5)), (gpu.FP64, (1e-10, 1e-10))):
        seen = []
        gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, spinor, prec, 1, (tg, tcx, tcy, tT),
                                    splines=[(knots, coefs)], on_plan=lambda p: seen.append(p.uniform_splines))
        e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
        print(f"  precision {prec}: uniform {seen[0]} perloop {e1:.2e} gsum {e2:.2e}")
        assert seen == [(2 - spinor) if prec == gpu.MIXED else 0]
        assert e1 < bars[0] and e2 < bars[1]


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 5e-5)])
@pytest.mark.parametrize("mode", [0, 1])
def test_spline_irregular_knots_and_mixed_profiles(gpu, orc, loops, prec, tol, mode):
    """Two spline tables (one with knots at random positions, so the bucket
    scans move) beside a Gaussian geometry: the run-time-profile kernels,
    tables read from global memory; nsub = 4."""
    rng = np.random.default_rng(7)
    k1, c1 = gaussian_spline(gpu, n=300, rho_max=12.0, l2=0.5, rng=rng)
    k2, c2 = gaussian_spline(gpu, n=64, rho_max=8.0, l2=2.0)
    geoms = [gpu.Geometry(kind=gpu.SPLINE, F=1.0, l2=0.5, a=1.0, spline_table=0),
             gpu.Geometry(kind=gpu.GAUSSIAN, F=0.5, l2=1.0),
             gpu.Geometry(kind=gpu.SPLINE, F=2.0, l2=2.0, a=1.0, spline_table=1)]
    og = [orc.Geom(orc.SPLINE, F=1.0, l2=0.5, tubedist=1.0, knots=k1, coefs=c1),
          orc.Geom(orc.GAUSSIAN, F=0.5, l2=1.0),
          orc.Geom(orc.SPLINE, F=2.0, l2=2.0, tubedist=1.0, knots=k2, coefs=c2)]
    parts = [tuples([0.0, 0.8], [0.03, 0.7, 5.0], geom=g) for g in range(3)]
    tup = tuple(np.concatenate([p[k] for p in parts]) for k in range(4))
    for spinor in (0, 1):
        for mapping in (1, 2):
            gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, mode, 4, spinor, prec, mapping, tup,
                                        splines=[(k1, c1), (k2, c2)])
            e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
            print(f"  spinor {spinor} mapping {mapping}: perloop {e1:.2e} gsum {e2:.2e}")
            assert e1 < tol and e2 < tol


@pytest.mark.parametrize("kind", ["STEP", "SMOOTH", "QUADRATIC", "PERIODIC", "FIXFLUX", "FIXFLUX_FF"])
@pytest.mark.parametrize("mode", [0, 1])
def test_other_profiles_fp64(gpu, orc, loops, kind, mode):
    k = getattr(gpu, kind)
    geoms = [gpu.Geometry(kind=k, F=0.5, l2=0.64, a=2.0, lam_min=0.2)]
    og = [orc.Geom(getattr(orc, kind), F=0.5, l2=0.64, tubedist=2.0, lmin=0.2)]
    gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, mode, 2, 1, gpu.FP64, 2,
                                tuples([0.0, 0.9], [0.03, 0.6]))
    e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
    print(f"  {kind}: perloop {e1:.2e} gsum {e2:.2e}")
    assert e1 < 1e-9 and e2 < 1e-9


@pytest.mark.parametrize("kind", ["STEP", "SMOOTH", "QUADRATIC", "PERIODIC", "FIXFLUX", "FIXFLUX_FF"])
def test_other_profiles_mixed(gpu, orc, loops, kind):
    """FP32 node arithmetic of the run-time-profile kernels (physical
    segments).  Bar: 1e-4 of the case's largest per-loop value (the FP32
    special-function forms of these profiles carry ~1e-6 relative error per
    node; the sums are FP64)."""
    k = getattr(gpu, kind)
    geoms = [gpu.Geometry(kind=k, F=0.5, l2=0.64, a=2.0, lam_min=0.2)]
    og = [orc.Geom(getattr(orc, kind), F=0.5, l2=0.64, tubedist=2.0, lmin=0.2)]
    gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, 1, gpu.MIXED, 1,
                                tuples([0.0, 0.9], [0.03, 0.6]))
    e1, e2 = relerr(pl, rpl), relerr(gs, rgs)
    print(f"  {kind}: perloop {e1:.2e} gsum {e2:.2e}")
    assert e1 < 1e-4 and e2 < 1e-4


@pytest.mark.parametrize("kind", ["PERIODIC", "FIXFLUX"])
@pytest.mark.parametrize("spinor", [0, 1])
def test_other_profiles_mixed_small_loops_at_centre(gpu, orc, loops, kind, spinor):
    """Small loops at and near a tube centre, where f(rho^2) ~ rho^2: the
    reference's forms of the periodic and fixflux profiles are 1 minus nearly
    1 there (relative error ~1e-7 / f in FP32, ~1e-3 in the phase at
    T = 1e-3); the FP32 kernels use equivalent forms without the
    cancellation and keep the per-loop phases within 4e-6 (for the fixflux
    form, the reference's phi(0) = 1 - q2*/q2 of the rounded q2 included).
    The periodic rings sit at a = 10 lambda: nearer rings leave f(0) > 0 (the
    sum of their tails), a point flux whose pure-gauge part small loops
    beside the axis cancel in FP32 like a distant tube's.  Loops further out
    (|x_cm|/sqrt(T) ~ 10) carry the symmetric gauge's ~1e-7 |x_cm|/sqrt(T).
    These tuples lie on the x axis (the axis form); nodes next to the axis
    set the error through the FP32 roundings of rho^2 and the cross products
    (2e-7 to 2e-6 with either form's roundings)."""
    k = getattr(gpu, kind)
    a = 8.0 if kind == "PERIODIC" else 2.0
    geoms = [gpu.Geometry(kind=k, F=0.5, l2=0.64, a=a, lam_min=0.2)]
    og = [orc.Geom(getattr(orc, kind), F=0.5, l2=0.64, tubedist=a, lmin=0.2)]
    gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, spinor, gpu.MIXED, 1,
                                tuples([0.0, 0.05, 0.15], [1e-3, 1e-2]))
    errs = []
    for t in range(pl.shape[0]):
        e_phase = relerr(pl[t, :, :3], rpl[t, :, :3])
        e_spin = relerr(pl[t, :, 3:], rpl[t, :, 3:]) if spinor else 0.0
        print(f"  {kind} tuple {t}: phase {e_phase:.2e} spin {e_spin:.2e}")
        errs.append(max(e_phase, e_spin))
    assert max(errs) < 4e-6, errs


@pytest.mark.parametrize("spinor", [0, 1])
def test_gaussian_mixed_small_loops_in_core(gpu, orc, loops, spinor):
    """Loops far smaller than the tube and inside its core (lambda = 2,
    T = 1e-3 .. 1e-2 at and near the axis): the plain FP32 quotient
    (1 - 2^-u)/u would carry a relative error ~6e-8/u per node, ~1e-4 in the
    phase here; the accurate nodes keep the per-loop phases within 2e-6 of
    the oracle (relative to the largest) and the spin sums within 1e-6."""
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=4.0)]
    og = [orc.Geom(orc.GAUSSIAN, F=1.0, l2=4.0)]
    gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, spinor, gpu.MIXED, 1,
                                tuples([0.0, 0.05, 0.3], [1e-3, 3e-3, 1e-2]))
    for t in range(pl.shape[0]):
        e_phase = relerr(pl[t, :, :3], rpl[t, :, :3])
        assert e_phase < 2e-6, (t, e_phase)
        if spinor:
            assert relerr(pl[t, :, 3:], rpl[t, :, 3:]) < 1e-6
    print(f"  max phase err {max(relerr(pl[t, :, :3], rpl[t, :, :3]) for t in range(pl.shape[0])):.2e}")


def test_gaussian_mixed_small_loops_lattice(gpu, orc, loops):
    """The same in a lattice (spacing 1.5 lambda, 3 x 3 tubes): the kernel
    picks the node per image tube and wavefront.  Bar 2e-5: a small loop
    (scaled size s') at scaled distance |c'| from a neighbouring tube sums
    that tube's gauge potential over terms ~|c'|/s' times its flux through
    the loop, so the FP32 rounding of the coordinates (not the node) gives
    a relative error ~1e-7 |c'|/s' in that tube's share of the phase:
    |c'|/s' ~ 100 here at T = 1e-3."""
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=4.0, lattice=3.0, images=1)]
    og = [orc.Geom(orc.GAUSSIAN, F=1.0, l2=4.0, lattice=3.0, images=1)]
    tg, tcx, tcy, tT = tuples([0.02, 0.4, 1.5, 2.9], [1e-3, 1e-2, 0.3])
    tcy = 0.5 * tcx
    gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, 1, gpu.MIXED, 1, (tg, tcx, tcy, tT))
    for t in range(pl.shape[0]):
        assert relerr(pl[t, :, :3], rpl[t, :, :3]) < 2e-5, t
        assert relerr(pl[t, :, 3:], rpl[t, :, 3:]) < 2e-6, t
    print(f"  max phase err {max(relerr(pl[t, :, :3], rpl[t, :, :3]) for t in range(pl.shape[0])):.2e}")
    print(f"  max spin err {max(relerr(pl[t, :, 3:], rpl[t, :, 3:]) for t in range(pl.shape[0])):.2e}")


@pytest.mark.parametrize("spinor", [0, 1])
def test_gaussian_mixed_local_node(gpu, orc, loops, spinor):
    """Small loops away from the axis of a wide tube (lambda = 2, x_cm up
    to 1.5 lambda, T = 1e-3 .. 3e-2) take the local Taylor node of the axis
    form: g(u) expanded about the loop centre's u, the pure-gauge part of the
    sub-loop sums from per-loop FP64 area sums, the spin node as 2^-uc minus
    a small deficit.  Per-loop phases within 2e-6 of the oracle (relative to
    the tuple's largest; the direct node reaches 1.4e-5 on these tuples, and
    a lane whose own loop is too large for the series takes it for that
    loop), spin sums within 1e-7, and the local node serves every tuple of
    the case."""
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=4.0)]
    og = [orc.Geom(orc.GAUSSIAN, F=1.0, l2=4.0)]
    seen = []
    tg, tcx, tcy, tT = tuples([0.0, 0.4, 1.14, 2.0], [1e-3, 3e-3, 1e-2, 3e-2])
    t3 = tuples([3.0], [1e-3, 3e-3, 1e-2])
    tup = tuple(np.concatenate([a, b]) for a, b in zip((tg, tcx, tcy, tT), t3))
    gs, pl, rgs, rpl = run_case(gpu, orc, loops, geoms, og, 1, 2, spinor, gpu.MIXED, 1, tup,
                                on_plan=lambda p: seen.append(p.local_tuples))
    assert seen == [19]
    e_ph = max(relerr(pl[t, :, :3], rpl[t, :, :3]) for t in range(pl.shape[0]))
    print(f"  max phase err {e_ph:.2e}")
    assert e_ph < 2e-6
    if spinor:
        e_sp = max(relerr(pl[t, :, 3:], rpl[t, :, 3:]) for t in range(pl.shape[0]))
        print(f"  max spin err {e_sp:.2e}")
        assert e_sp < 1e-7
    assert relerr(gs, rgs) < 1e-6
