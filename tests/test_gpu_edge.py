"""Edge cases of the batched engine on the GPU, checked against the FP64
oracle: an empty batch, the smallest loops (N = 4) and a single loop, loop
counts that are not a multiple of the group size, geometries limited to
fewer loops (or none), extreme proper times and far-away x_cm."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


def oracle_groups(orc, og, x, y, group, tup, nloops, spinor, mode=1, nsub=2):
    tg, tcx, tcy, tT = tup
    _, pl = orc.eval_batch(og, mode, nsub, spinor, x, y, group, tg, tcx, tcy, tT)
    Nl = x.shape[0]
    ng = (Nl + group - 1) // group
    gs = np.zeros((len(tg), ng, 3))
    for t in range(len(tg)):
        for l in range(nloops[tg[t]]):
            gs[t, l // group] += orc.wminus1(mode, spinor, pl[t, l, :3], pl[t, l, 3:])
    return gs


def test_empty_batch(gpu):
    x, y = gpu.gen_loops(11, 8, 16)
    L = gpu.Loops(x, y)
    e = np.zeros(0)
    b = gpu.Batch([gpu.Geometry(kind=gpu.GAUSSIAN)], e.astype(np.int32), e, e, e, precision=gpu.MIXED)
    plan = gpu.Plan(b, L)
    print("empty batch", flush=True)
    plan.run()
    gs, _ = plan.read()
    assert gs.shape[0] == 0 and plan.nonfinite() == 0
    plan.close()
    L.close()


@pytest.mark.parametrize("Nl,N", [(1, 4), (3, 8), (100, 64)])
@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 2e-5)])
def test_small_and_ragged_ensembles(gpu, orc, Nl, N, prec, tol):
    """One loop of four points, three loops, and 100 loops in groups of 32
    (a last group of 4)."""
    x, y = gpu.gen_loops(12, Nl, N)
    L = gpu.Loops(x, y)
    group = 32
    tg = np.zeros(6, np.int32)
    tcx = np.array([0.0, 0.3, 1.1, 0.0, 0.7, 2.0])
    tcy = np.array([0.0, 0.1, -0.4, 0.0, 0.2, 0.0])
    tT = np.array([0.01, 0.2, 1.0, 3.0, 0.05, 0.5])
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=0.5)]
    og = [orc.Geom(orc.GAUSSIAN, F=1.0, l2=0.5)]
    want = oracle_groups(orc, og, x, y, group, (tg, tcx, tcy, tT), [Nl], 1)
    for mapping in (1, 2):
        b = gpu.Batch(geoms, tg, tcx, tcy, tT, precision=prec, spinor=1, group=group, mapping=mapping)
        plan = gpu.Plan(b, L)
        print(f"Nl={Nl} N={N} prec={prec} mapping={mapping}", flush=True)
        plan.run()
        gs, _ = plan.read()
        assert plan.nonfinite() == 0
        assert relerr(gs, want) < tol
        plan.close()
    L.close()


def test_geometry_loop_counts(gpu, orc):
    """geom_nloops: a geometry on 40 of 96 loops, one on none (its tuples
    read zero), one asking for more than there are (all 96)."""
    x, y = gpu.gen_loops(13, 96, 32)
    L = gpu.Loops(x, y)
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=0.5, l2=1.0), gpu.Geometry(kind=gpu.GAUSSIAN, F=2.0, l2=0.3),
             gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=2.0)]
    og = [orc.Geom(orc.GAUSSIAN, F=0.5, l2=1.0), orc.Geom(orc.GAUSSIAN, F=2.0, l2=0.3),
          orc.Geom(orc.GAUSSIAN, F=1.0, l2=2.0)]
    nl = np.array([40, 0, 500], np.int32)
    tg = np.repeat(np.arange(3), 4).astype(np.int32)
    tcx = np.tile([0.0, 0.4, 0.9, 1.5], 3)
    tcy = np.zeros(12)
    tT = np.tile([0.05, 0.3, 1.0, 2.0], 3)
    want = oracle_groups(orc, og, x, y, 32, (tg, tcx, tcy, tT), [40, 0, 96], 0)
    for prec, tol in ((gpu.FP64, 1e-10), (gpu.MIXED, 2e-5)):
        for mapping in (1, 2, 3):
            b = gpu.Batch(geoms, tg, tcx, tcy, tT, precision=prec, spinor=0, group=32, geom_nloops=nl,
                          mapping=mapping)
            plan = gpu.Plan(b, L)
            print(f"geom_nloops prec={prec} mapping={mapping}", flush=True)
            plan.run()
            gs, _ = plan.read()
            assert np.all(gs[4:8] == 0.0)
            assert relerr(gs, want) < tol
            plan.close()
    L.close()


@pytest.mark.parametrize("prec,tol", [(0, 1e-10), (1, 2e-5)])
def test_extreme_T_and_far_xcm(gpu, orc, prec, tol):
    """T from 1e-8 to 1e4 and x_cm up to 100 lambda away: finite results.
    On the axis each tuple matches the oracle to the precision's bar.  Off
    the axis a loop of size sqrt(T) at distance rho sums its small flux from
    terms ~rho/sqrt(T) larger (the symmetric gauge), so FP32 coordinates add
    ~1e-7 rho/sqrt(T) relative; 100 lambda away no loop holds flux, W - 1
    is the rounding residual of a pure gauge (~1e-10 and below) and both
    precisions are held to absolute bars (1e-9 mixed, 1e-20 FP64).  F = 0.3 keeps the large loops, which wind around the tube, away
    from the Aharonov-Bohm cancellation of an integer flux."""
    x, y = gpu.gen_loops(14, 64, 64)
    L = gpu.Loops(x, y)
    T = np.array([1e-8, 1e-5, 1e-2, 10.0, 1e3, 1e4])
    tcx = np.array([0.0, 0.5, 100.0])
    tg = np.zeros(len(T) * len(tcx), np.int32)
    cx = np.repeat(tcx, len(T))
    tT = np.tile(T, len(tcx))
    cy = np.zeros_like(cx)
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=0.3, l2=1.0)]
    og = [orc.Geom(orc.GAUSSIAN, F=0.3, l2=1.0)]
    for spinor in (0, 1):
        want = oracle_groups(orc, og, x, y, 32, (tg, cx, cy, tT), [64], spinor)
        b = gpu.Batch(geoms, tg, cx, cy, tT, precision=prec, spinor=spinor, group=32, mapping=1)
        plan = gpu.Plan(b, L)
        print(f"extreme T prec={prec} spinor={spinor}", flush=True)
        plan.run()
        gs, _ = plan.read()
        assert np.all(np.isfinite(gs)) and plan.nonfinite() == 0
        for t in range(len(tg)):
            if cx[t] > 50.0 and tT[t] <= 1e3:
                # no flux inside these loops (at most 2.4 sqrt(T) < 100 across): W - 1 is the quadrature's own
                # residual of a closed-path integral of a pure gauge, ~1e-10
                assert np.max(np.abs(want[t])) < 1e-6
                bar_abs = 1e-9 if prec == gpu.MIXED else 1e-20
                assert np.max(np.abs(gs[t] - want[t])) < bar_abs, (t, tT[t], gs[t], want[t])
                continue
            bar = tol + (3e-7 * cx[t] / np.sqrt(tT[t]) if prec == gpu.MIXED else 0.0)
            sc = max(np.max(np.abs(want[t])), 1e-300)
            if prec == gpu.MIXED and spinor:
                # cosh(S) of a loop carries S's relative error times S (S ~ ln|W|)
                bar *= max(1.0, np.log(sc))
            assert np.max(np.abs(gs[t] - want[t])) <= bar * sc, (t, cx[t], tT[t], gs[t], want[t])
        plan.close()
    L.close()
