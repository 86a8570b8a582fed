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
sum of the case; mixed kernels within 2e-5 per loop and 1e-5 in the group
sums, relative to the case's largest (the single-image bars of the other
parity tests)."""
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
            assert e1 <= 2e-5 and e2 <= 1e-5
