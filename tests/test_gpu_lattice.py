"""Dense square lattices of Gaussian tubes with the 5 x 5 image window
(images = 2), on the GPU against the FP64 oracle.

The lattice workload sums A_phi over the (2M+1)^2 = 25 tubes of the window
centred on x_cm's nearest tube (an extension of the reference, whose
"periodic" profiles are rings, intEV.cu:202-257).  Cases: spacing
l = 0.1, 0.3 and 3 lambda (lambda = 1, F = 1), x_cm on a 3 x 3 grid over
one unit cell (off the axes), T in {0.01, 0.3, 3}, scalar and spinor
QED, physical mode, Simpson n = 2, lane-per-tuple and wave-per-loop
mappings.

Bars: FP64 kernels within 1e-10 of the largest per-loop value and group
sum of the case; mixed kernels within 1e-6 per loop and 5e-6 in the group
sums, relative to the case's largest (measured: at most 2.6e-7 and 2.6e-6,
the latter for spinor QED at l = 0.1 lambda with both mappings; one FP32
sum over a point's 25 image tubes in the lane-per-tuple inner loop gave
7.4e-6 there)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NL, N, GROUP = 48, 256, 16
SPACINGS = (0.1, 0.3, 3.0)
TS = (0.01, 0.3, 3.0)


@pytest.fixture(scope="module")
def loops(gpu):
    x, y = gpu.gen_loops(3, NL, N)
    L = gpu.Loops(x, y)
    yield x, y, L
    L.close()


def relerr(a, b):
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def cell_tuples():
    tg, tcx, tcy, tT = [], [], [], []
    for gi, l in enumerate(SPACINGS):
        for fx in (0.13, 0.5, 0.81):
            for fy in (0.07, 0.44, 0.9):
                for t in TS:
                    tg.append(gi); tcx.append(fx * l); tcy.append(fy * l); tT.append(t)
    return np.array(tg), np.array(tcx), np.array(tcy), np.array(tT)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("spinor", [0, 1])
@pytest.mark.parametrize("mapping", [1, 2])
def test_lattice_5x5_window(gpu, orc, loops, prec, spinor, mapping):
    x, y, L = loops
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=1.0, lattice=l, images=2) for l in SPACINGS]
    og = [orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0, lattice=l, images=2) for l in SPACINGS]
    tg, tcx, tcy, tT = cell_tuples()
    print(f"case prec={prec} spinor={spinor} mapping={mapping} ntup={len(tg)}", flush=True)
    b = gpu.Batch(geoms, tg, tcx, tcy, tT, mode=gpu.MODE_PHYSICAL, precision=prec, nsub=2, spinor=spinor,
                  group=GROUP, want_perloop=True, mapping=mapping)
    plan = gpu.Plan(b, L)
    try:
        assert plan.axis_form == 0
        plan.run()
        gs, pl = plan.read()
        assert plan.nonfinite() == 0
    finally:
        plan.close()
    rgs, rpl = orc.eval_batch(og, orc.MODE_PHYSICAL, 2, spinor, x, y, GROUP, tg, tcx, tcy, tT)
    for gi, l in enumerate(SPACINGS):
        sel = tg == gi
        e1, e2 = relerr(pl[sel], rpl[sel]), relerr(gs[sel], rgs[sel])
        print(f"  spacing {l}: perloop {e1:.2e} gsum {e2:.2e}")
        if prec == 0:
            assert e1 <= 1e-10 and e2 <= 1e-10
        else:
            assert e1 <= 1e-6 and e2 <= 5e-6


def test_interaction_energy_small_lattice_vs_oracle(gpu, orc, loops):
    """The interaction energy per tube through the library (FP64 engine
    group sums, wfl_cell_action_groups over a 2 x 2 cell grid,
    wfl_radial_action_groups, wfl_interaction_energy) against the oracle's
    group sums and its restatement of the same quadratures and of
    interaction.sage (scalar, spacing 1.5 lambda, 3 x 3 images): within
    1e-9 of the isolated tube's action."""
    x, y, L = loops
    T = np.logspace(-2, 1, 9)
    nT = len(T)
    counts = np.full(NL // GROUP, GROUP, np.int32)
    lat = gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=1.0, lattice=1.5, images=1)
    iso = gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=1.0)
    olat = orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0, lattice=1.5, images=1)
    oiso = orc.Geom(orc.GAUSSIAN, F=1.0, l2=1.0)
    c = (np.arange(2) + 0.5) * 0.75
    cx, cy = np.meshgrid(c, c, indexing="ij")
    cx, cy = cx.ravel(), cy.ravel()
    w = np.full(4, 1.5 * 1.5 / 4)
    rho = np.linspace(0.0, 3.0, 9)
    tg = np.concatenate([np.zeros(4 * nT, int), np.ones(len(rho) * nT, int)])
    tcx = np.concatenate([np.repeat(cx, nT), np.repeat(rho, nT)])
    tcy = np.concatenate([np.repeat(cy, nT), np.zeros(len(rho) * nT)])
    tT = np.tile(T, 4 + len(rho))
    b = gpu.Batch([lat, iso], tg, tcx, tcy, tT, mode=gpu.MODE_PHYSICAL, precision=gpu.FP64, nsub=2, spinor=0,
                  group=GROUP)
    gs, _ = gpu.eval_batch(b, L)
    rgs, _ = orc.eval_batch([olat, oiso], orc.MODE_PHYSICAL, 2, 0, x, y, GROUP, tg, tcx, tcy, tT, perloop=False)
    ng = len(counts)
    gc, gr = gs[:4 * nT].reshape(4, nT, ng, 3), gs[4 * nT:].reshape(len(rho), nT, ng, 3)
    rc, rr = rgs[:4 * nT].reshape(4, nT, ng, 3), rgs[4 * nT:].reshape(len(rho), nT, ng, 3)
    cell = gpu.cell_action_groups(lat, 0, T, cx, cy, w, gc, counts)
    isoa = gpu.radial_action_groups(iso, 0, T, rho, gr, counts)
    got = gpu.interaction_energy(cell, isoa)
    ocell = orc.cell_action_groups(olat, 0, T, cx, cy, w, rc, counts)
    oisoa = orc.radial_action_groups(oiso, 0, T, rho, rr, counts)
    ref = orc.interaction_energy(ocell, oisoa)
    scale = np.abs(oisoa).max()
    print(f"  E_int {got[0]:.6e} +- {got[1]:.2e} (oracle {ref[0]:.6e} +- {ref[1]:.2e}); |action_iso| {scale:.3e}")
    assert abs(got[0] - ref[0]) <= 1e-9 * scale and abs(got[1] - ref[1]) <= 1e-9 * scale
    assert np.max(np.abs(cell - ocell)) <= 1e-9 * scale


def test_interaction_energy_vanishes_for_distant_tubes(gpu):
    """Tubes 10 lambda apart do not interact: the cell action (midpoint rule
    on a 20 x 20 grid) equals the isolated tube's (Simpson in rho to
    5 lambda), so E_int is zero within three standard errors, on the mixed
    engine, scalar and spinor (1024 loops x 256 points in 8 groups,
    physical mode, 16 T in [1e-2, 30])."""
    nl, grp = 1024, 128
    x, y = gpu.gen_loops(4, nl, 256)
    L = gpu.Loops(x, y)
    T = np.logspace(-2, np.log10(30.0), 16)
    nT = len(T)
    counts = np.full(nl // grp, grp, np.int32)
    l = 10.0
    lat = gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=1.0, lattice=l, images=1)
    iso = gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=1.0)
    n = 20
    c = (np.arange(n) + 0.5) * (l / n) - 0.5 * l
    cx, cy = np.meshgrid(c, c, indexing="ij")
    cx, cy = cx.ravel(), cy.ravel()
    w = np.full(n * n, l * l / (n * n))
    rho = np.linspace(0.0, 5.0, 21)
    tg = np.concatenate([np.zeros(n * n * nT, int), np.ones(len(rho) * nT, int)])
    tcx = np.concatenate([np.repeat(cx, nT), np.repeat(rho, nT)])
    tcy = np.concatenate([np.repeat(cy, nT), np.zeros(len(rho) * nT)])
    tT = np.tile(T, n * n + len(rho))
    ng = len(counts)
    for spinor in (0, 1):
        b = gpu.Batch([lat, iso], tg, tcx, tcy, tT, mode=gpu.MODE_PHYSICAL, precision=gpu.MIXED, nsub=2,
                      spinor=spinor, group=grp)
        gs, _ = gpu.eval_batch(b, L)
        cell = gpu.cell_action_groups(lat, spinor, T, cx, cy, w, gs[:n * n * nT].reshape(n * n, nT, ng, 3), counts)
        isoa = gpu.radial_action_groups(iso, spinor, T, rho, gs[n * n * nT:].reshape(len(rho), nT, ng, 3), counts)
        m, spread, sem = gpu.interaction_energy(cell, isoa)
        print(f"  spinor {spinor}: E_int {m:.3e} +- {sem:.2e} (spread {spread:.2e}); action_iso {isoa.mean():.4e}")
        assert abs(m) <= 3.0 * sem + 1e-4 * abs(isoa.mean())
    L.close()
