"""wfl_eval_batch_noalloc (the entry point that allocates no device memory)
against wfl_eval_batch through the C ABI.

Both run the same kernels in the same launch configuration on the same
inputs, so their group sums and per-loop values are compared for equality
(bit for bit).  Two calls in a row, the second on a smaller batch and loop
set, check that what the shared scratch array keeps from the first call does
not reach the second.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

T_GRID = [0.01, 0.2, 3.0, 40.0]


def axis_batch(wfl, cxs, Ts, precision, spinor, group=32, want_perloop=True):
    geoms = [wfl.Geometry(kind=wfl.GAUSSIAN, F=1.0, l2=1.0), wfl.Geometry(kind=wfl.GAUSSIAN, F=10.0, l2=4.0)]
    tg, tcx, tT = [], [], []
    for g in range(len(geoms)):
        for c in cxs:
            for t in Ts:
                tg.append(g); tcx.append(c); tT.append(t)
    tg, tcx, tT = np.array(tg), np.array(tcx), np.array(tT)
    return wfl.Batch(geoms, tg, tcx, 0.0 * tcx, tT, mode=wfl.MODE_PHYSICAL, precision=precision, nsub=2,
                     spinor=spinor, group=group, want_perloop=want_perloop)


def lattice_batch(wfl, precision):
    geoms = [wfl.Geometry(kind=wfl.GAUSSIAN, F=1.0, l2=1.0, lattice=1.0, images=1)]
    c = np.array([0.1, 0.45, 0.8])
    tg, tcx, tcy, tT = [], [], [], []
    for cx in c:
        for t in [0.05, 0.5, 5.0]:
            tg.append(0); tcx.append(cx); tcy.append(0.5 * cx); tT.append(t)
    return wfl.Batch(geoms, np.array(tg), np.array(tcx), np.array(tcy), np.array(tT), mode=wfl.MODE_PHYSICAL,
                     precision=precision, nsub=4, spinor=1, group=16, want_perloop=True)


def same(a, b):
    assert a.shape == b.shape
    np.testing.assert_array_equal(a, b)


def compare(wfl, b, L, label):
    print(f"no-alloc case {label}: ntup={len(b.tup_geom)} Nl={L.Nl}", flush=True)
    gs, pl = wfl.eval_batch(b, L)
    gs2, pl2 = wfl.eval_batch_noalloc(b, L)
    same(gs2, gs)
    if pl is not None:
        same(pl2, pl)
    assert np.all(np.isfinite(gs))


@pytest.mark.parametrize("precision,spinor", [("mixed", 1), ("mixed", 0), ("fp64", 1)])
def test_noalloc_axis_two_calls(gpu, precision, spinor):
    wfl = gpu
    prec = wfl.MIXED if precision == "mixed" else wfl.FP64
    xa, ya = wfl.gen_loops(11, 512, 256)
    xb, yb = wfl.gen_loops(12, 96, 128)
    La, Lb = wfl.Loops(xa, ya), wfl.Loops(xb, yb)
    try:
        compare(wfl, axis_batch(wfl, [0.0, 0.3, 1.2, 2.5], T_GRID, prec, spinor), La, f"{precision} large")
        compare(wfl, axis_batch(wfl, [0.7], T_GRID[:2], prec, spinor, group=16), Lb, f"{precision} small")
    finally:
        La.close(); Lb.close()


def test_noalloc_axis_uses_local_node(gpu):
    """The mixed Gaussian axis batch above takes the local Taylor node for
    some tuples, so that path (and its per-loop sums) runs from the scratch."""
    wfl = gpu
    x, y = wfl.gen_loops(11, 512, 256)
    L = wfl.Loops(x, y)
    try:
        p = wfl.Plan(axis_batch(wfl, [0.0, 0.3, 1.2, 2.5], T_GRID, wfl.MIXED, 1), L)
        assert p.local_tuples > 0 and p.axis_form == 1
        p.close()
    finally:
        L.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_noalloc_lattice_two_calls(gpu, precision):
    wfl = gpu
    prec = wfl.MIXED if precision == "mixed" else wfl.FP64
    x, y = wfl.gen_loops(13, 128, 256)
    L = wfl.Loops(x, y)
    try:
        compare(wfl, lattice_batch(wfl, prec), L, f"lattice {precision}")
        b = lattice_batch(wfl, prec)
        keep = slice(0, 4)
        small = wfl.Batch(b.geoms, b.tup_geom[keep], b.tup_cx[keep], b.tup_cy[keep], b.tup_T[keep],
                          mode=b.mode, precision=prec, nsub=4, spinor=1, group=16, want_perloop=True)
        compare(wfl, small, L, f"lattice {precision} small")
    finally:
        L.close()


def test_noalloc_rejects_oversize_batch(gpu):
    """A plan larger than the 64 MiB scratch fails with WFL_ERR_ARG before
    any launch; the next call of a batch that fits is unaffected."""
    wfl = gpu
    x, y = wfl.gen_loops(14, 256, 128)
    L = wfl.Loops(x, y)
    try:
        ok = axis_batch(wfl, [0.2], T_GRID, wfl.MIXED, 1)
        wfl.eval_batch(ok, L)                       # builds the loop set's records
        big = axis_batch(wfl, np.linspace(0.0, 3.0, 512), T_GRID, wfl.MIXED, 1, group=1)
        assert len(big.tup_geom) == 4096            # per-loop outputs alone: 4096 x 256 x 48 B = 50 MB
        with pytest.raises(wfl.WflError, match=r"error -1: .*scratch"):
            wfl.eval_batch_noalloc(big, L)
        compare(wfl, ok, L, "after an oversize batch")
    finally:
        L.close()


def test_noalloc_needs_built_records(gpu):
    """On a loop set whose records were never built the call fails with
    WFL_ERR_STATE instead of allocating them."""
    wfl = gpu
    x, y = wfl.gen_loops(15, 64, 128)
    L = wfl.Loops(x, y)
    try:
        b = axis_batch(wfl, [0.2], T_GRID, wfl.MIXED, 1)
        with pytest.raises(wfl.WflError, match=r"error -4: .*loop set has no"):
            wfl.eval_batch_noalloc(b, L)
        compare(wfl, b, L, "after building the records")
    finally:
        L.close()
