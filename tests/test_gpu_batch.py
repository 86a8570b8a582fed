"""Batched uneven workloads on the GPU: several geometries (isolated and
lattice) with their own rho/T grids and loop counts in one launch, every
launch configuration (lane per tuple inline / split, static / persistent
grid, 1-4 wavefronts per block, 1-4 tasks per loop group), the measured
autotuner, and hipGraph replay.
Checked against the FP64 oracle (per-loop W - 1 summed per group)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NL, N, GROUP = 96, 128, 32


@pytest.fixture(scope="module")
def uneven(gpu, orc):
    rng = np.random.default_rng(42)
    x, y = gpu.gen_loops(4, NL, N)
    L = gpu.Loops(x, y)
    specs = []   # (F, l2, lattice, images, nloops)
    for g in range(5):
        lam = rng.uniform(0.3, 3.0)
        F = float(np.exp(rng.uniform(np.log(0.01), np.log(10.0))))
        lat = 0.0 if g % 3 == 0 else rng.uniform(0.5, 10.0) * lam
        specs.append((F, lam * lam, lat, 1 if lat > 0 else 0, int(rng.choice([32, 64, 80, 96]))))
    tg, tcx, tcy, tT = [], [], [], []
    for g, (F, l2, lat, M, nl) in enumerate(specs):
        nr, nt = int(rng.integers(2, 5)), int(rng.integers(3, 7))
        for rho in np.linspace(0.0, 2.0 * np.sqrt(l2), nr):
            for t in np.logspace(-2, 1, nt):
                tg.append(g); tcx.append(rho); tcy.append(0.3 * rho if lat else 0.0); tT.append(t)
    tup = tuple(np.array(v) for v in (tg, tcx, tcy, tT))
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=F, l2=l2, lattice=lat, images=M) for F, l2, lat, M, _ in specs]
    ogeoms = [orc.Geom(orc.GAUSSIAN, F=F, l2=l2, lattice=lat, images=M) for F, l2, lat, M, _ in specs]
    nloops = np.array([s[4] for s in specs])
    yield x, y, L, geoms, ogeoms, tup, nloops
    L.close()


def oracle_groups(orc, ogeoms, x, y, tup, nloops, spinor, mode=1):
    tg, tcx, tcy, tT = tup
    _, pl = orc.eval_batch(ogeoms, mode, 2, spinor, x, y, GROUP, tg, tcx, tcy, tT)
    ng = (NL + GROUP - 1) // GROUP
    gs = np.zeros((len(tg), ng, 3))
    for t in range(len(tg)):
        nl = nloops[tg[t]]
        for l in range(nl):
            gs[t, l // GROUP] += orc.wminus1(mode, spinor, pl[t, l, :3], pl[t, l, 3:])
    return gs


def relerr(a, b):
    return np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300)


@pytest.mark.parametrize("spinor", [0, 1])
def test_uneven_batch_all_configurations(gpu, orc, uneven, spinor):
    x, y, L, geoms, ogeoms, tup, nloops = uneven
    want = oracle_groups(orc, ogeoms, x, y, tup, nloops, spinor)
    for prec, tol in ((gpu.FP64, 1e-10), (gpu.MIXED, 2e-5)):
        b = gpu.Batch(geoms, *tup, precision=prec, spinor=spinor, group=GROUP, geom_nloops=nloops, mapping=1)
        plan = gpu.Plan(b, L)
        ref = None
        for mapping in (1, 3):
            for persistent in (0, 1):
                for wpb in (1, 2, 4):
                    for chunks in (1, 3, 4):
                        print(f"config prec={prec} spinor={spinor} mapping={mapping} persistent={persistent} "
                              f"wpb={wpb} chunks={chunks}", flush=True)
                        plan.set_config(mapping, persistent, wpb, chunks)
                        assert plan.config() == (mapping, persistent, wpb, chunks)
                        plan.run()
                        gs, _ = plan.read()
                        assert relerr(gs, want) < tol
                        if ref is None:
                            ref = gs
                        elif chunks == 1 or mapping == 3:   # whole tasks reordered: same sums
                            assert relerr(gs, ref) < 1e-14
                        else:   # chunk sums added in order: rounding of the FP64 sums only
                            assert relerr(gs, ref) < 1e-12
        assert plan.nonfinite() == 0
        plan.close()


def test_autotune_and_graph_replay(gpu, orc, uneven):
    x, y, L, geoms, ogeoms, tup, nloops = uneven
    b = gpu.Batch(geoms, *tup, precision=gpu.MIXED, spinor=1, group=GROUP, geom_nloops=nloops)
    plan = gpu.Plan(b, L)
    print("autotune", flush=True)
    choice = plan.autotune()
    m, pe, w, c = plan.config()
    assert choice == 1000 * m + 100 * pe + 10 * w + c and m in (1, 2, 3)
    tuned, _ = plan.read()
    plan.run()
    eager, _ = plan.read()
    assert relerr(tuned, eager) < 1e-14
    want = oracle_groups(orc, ogeoms, x, y, tup, nloops, 1)
    assert relerr(eager, want) < 2e-5
    # hipGraph: replays, with eager work between them, reproduce the eager result
    for k in range(3):
        print(f"graph replay {k}", flush=True)
        plan.graph_run()
        g, _ = plan.read()
        assert relerr(g, eager) < 1e-14
        plan.run()
        e2, _ = plan.read()
        assert relerr(e2, eager) < 1e-14
    plan.close()


def test_autotune_nsub(gpu):
    """Simpson sub-steps by measured accuracy: tighter tolerances ask for more
    sub-steps; the chosen n's group sums are within tol of those at 2n and
    n/2's are not (checked with separate batches).  Loops of size ~sqrt(T)
    cross a narrow tube's core, where few sub-steps are least accurate."""
    x, y = gpu.gen_loops(21, 128, 128)
    L = gpu.Loops(x, y)
    geoms = [gpu.Geometry(kind=gpu.GAUSSIAN, F=1.0, l2=0.25)]
    tg = np.zeros(8, np.int32)
    tcx = np.linspace(0.0, 1.0, 8)
    tcy = np.zeros(8)
    tT = np.full(8, 2.0)

    def sums(n):
        b = gpu.Batch(geoms, tg, tcx, tcy, tT, precision=gpu.FP64, spinor=1, nsub=n, group=128)
        gs, _ = gpu.eval_batch(b, L)
        return gs[:, 0, :]

    def rel(a, b):
        return np.max(np.abs(a - b)) / np.max(np.abs(b))

    chosen = []
    for tol in (1e-2, 1e-4, 1e-6, 1e-8):
        b = gpu.Batch(geoms, tg, tcx, tcy, tT, precision=gpu.FP64, spinor=1, nsub=2, group=128)
        print(f"autotune_nsub tol={tol}", flush=True)
        n, e = gpu.autotune_nsub(b, L, tol, nmax=256, max_tuples=8, max_loops=128)
        chosen.append(n)
        assert n < 256 and e <= tol
        assert rel(sums(n), sums(2 * n)) == pytest.approx(e, rel=1e-9)
        if n > 2:
            assert rel(sums(n // 2), sums(n)) > tol
    assert chosen == sorted(chosen) and chosen[-1] > chosen[0]
    # a sample smaller than the ensemble (the first 64 of 128 loops)
    b = gpu.Batch(geoms, tg, tcx, tcy, tT, precision=gpu.FP64, spinor=1, nsub=2, group=128)
    n, e = gpu.autotune_nsub(b, L, 1e-6, nmax=256, max_tuples=4, max_loops=64)
    b64 = [gpu.Batch(geoms, tg[::2], tcx[::2], tcy[::2], tT[::2], precision=gpu.FP64, spinor=1, nsub=m, group=64,
                     geom_nloops=np.array([64], np.int32)) for m in (n, 2 * n)]
    g1, g2 = (gpu.eval_batch(bb, L)[0][:, 0, :] for bb in b64)
    assert rel(g1, g2) == pytest.approx(e, rel=1e-9) and e <= 1e-6
    L.close()
