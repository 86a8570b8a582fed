"""Benchmark of the worldline Wilson-loop hot path on one MI355X.

One step = one batched evaluation of the spinor-QED workload (the engine's
multi-geometry configuration): 5120 unit loops x 4096 points, 9 Gaussian
tubes (lambda in {0.5, 1, 2} x F in {0.1, 1, 10}), 64 x_cm values
rho in [0, 4 lambda], 96 log-spaced T in [1e-3, 1e2], Simpson n = 2,
physical segment mode, reduced on the device to FP64 sums of W - 1 per
group of 128 loops.

Prints one JSON line: throughput of the mixed-precision engine (point-evals/s,
one point-eval = one (loop, point, x_cm, T, geometry) tuple), the FP64 engine
on the same workload, the L_eff / <W> error of mixed against FP64, the
roofline of the dominant kernel and a CPU baseline (the FP64 C oracle on a
bounded sample).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Algorithmic FLOPs per point-eval (Gaussian tube, n Simpson nodes per segment,
# K images): scalar 5 + K(21 + 9n), spinor 5 + K(27 + 11n).
def flops_per_point_eval(spinor, n, K=1):
    return 5 + K * ((27 + 11 * n) if spinor else (21 + 9 * n))


PEAK_FP32_VECTOR = 157.3e12   # MI355X FP32 vector (packed) FLOP/s
PEAK_FP64_VECTOR = 78.6e12
PEAK_HBM = 8.0e12


def config_c2():
    lams = [0.5, 1.0, 2.0]
    Fs = [0.1, 1.0, 10.0]
    geoms = [(F, lam * lam) for lam in lams for F in Fs]
    nrho, nT = 64, 96
    T = np.logspace(-3, 2, nT)
    tg, tcx, tT = [], [], []
    for g, (F, l2) in enumerate(geoms):
        lam = np.sqrt(l2)
        for rho in np.linspace(0.0, 4.0 * lam, nrho):
            for t in T:
                tg.append(g); tcx.append(rho); tT.append(t)
    return dict(name="spinor_gaussian_5120x4096_9geom_64rho_96T", Nl=5120, N=4096, config_id=2, spinor=1,
                nsub=2, geoms=geoms, nrho=nrho, nT=nT, T=T,
                tup=(np.array(tg), np.array(tcx), np.zeros(len(tcx)), np.array(tT)))


def config_c1():
    geoms = [(1.0, 1.0)]
    nrho, nT = 32, 64
    T = np.logspace(-3, 2, nT)
    tg, tcx, tT = [], [], []
    for rho in np.linspace(0.0, 3.0, nrho):
        for t in T:
            tg.append(0); tcx.append(rho); tT.append(t)
    return dict(name="scalar_gaussian_1024x1024_32rho_64T", Nl=1024, N=1024, config_id=1, spinor=0, nsub=2,
                geoms=geoms, nrho=nrho, nT=nT, T=T,
                tup=(np.array(tg), np.array(tcx), np.zeros(len(tcx)), np.array(tT)))


def leff_table(wfl, cfg, gsum, group, policy):
    """L_eff per (geometry, rho) and per group, from group sums."""
    nrho, nT = cfg["nrho"], cfg["nT"]
    ng = gsum.shape[1]
    counts = np.full(ng, group, np.int32)
    tg, tcx = cfg["tup"][0], cfg["tup"][1]
    out = np.zeros((len(cfg["geoms"]), nrho, ng))
    for g, (F, l2) in enumerate(cfg["geoms"]):
        geom = wfl.Geometry(kind=wfl.GAUSSIAN, F=F, l2=l2)
        for r in range(nrho):
            base = (g * nrho + r) * nT
            out[g, r] = wfl.leff_groups(geom, cfg["spinor"], cfg["T"], tcx[base], 0.0, gsum[base:base + nT],
                                        counts, policy)
    return out


def cpu_baseline(cfg, x, y, seconds=12.0):
    """FP64 C oracle (OpenMP over loops) on a bounded sample of the workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc

    cores = int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count() or 1
    tg, tcx, tcy, tT = cfg["tup"]
    geoms = [orc.Geom(orc.GAUSSIAN, F=F, l2=l2) for F, l2 in cfg["geoms"]]
    nl = 64
    stride = max(1, len(tg) // 64)
    idx = np.arange(0, len(tg), stride)
    t0 = time.perf_counter()
    orc.eval_batch(geoms, orc.MODE_PHYSICAL, cfg["nsub"], cfg["spinor"], x[:nl], y[:nl], nl, tg[idx], tcx[idx],
                   tcy[idx], tT[idx], perloop=False, nthreads=cores)
    dt = time.perf_counter() - t0
    # scale the tuple sample to roughly `seconds` of CPU work
    scale = max(1, int(seconds / max(dt, 1e-3)))
    stride2 = max(1, stride // scale)
    idx = np.arange(0, len(tg), stride2)
    t0 = time.perf_counter()
    orc.eval_batch(geoms, orc.MODE_PHYSICAL, cfg["nsub"], cfg["spinor"], x[:nl], y[:nl], nl, tg[idx], tcx[idx],
                   tcy[idx], tT[idx], perloop=False, nthreads=cores)
    dt = time.perf_counter() - t0
    pe = nl * cfg["N"] * len(idx)
    return {"value": pe / dt, "unit": "point-evals/s", "cores": cores, "kind": "port",
            "sample": f"first {nl} loops x {cfg['N']} points x every {stride2}th of {len(tg)} tuples "
                      f"({len(idx)} tuples), FP64 C oracle, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp64-steps", type=int, default=1)
    ap.add_argument("--config", default="c2", choices=["c1", "c2"])
    ap.add_argument("--mode", default="physical", choices=["physical", "reference"])
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-ref-mode", action="store_true")
    ap.add_argument("--only-mixed", action="store_true", help="time the mixed engine only (profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.only_mixed:
        args.no_fp64 = args.no_cpu = args.no_ref_mode = True
    if args.copy_probe:
        import torch

        n = args.copy_probe * (1 << 20) // 4
        a = torch.ones(n, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes": 4 * n, "dispatches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 (the multi-GPU sharding "
                                   "is described, not built)"}))
        sys.exit(2)

    import __graft_entry__ as ge

    wfl = ge._import_pkg()
    if wfl.device_count() < 1:
        raise SystemExit("bench: no HIP device")
    cfg = config_c2() if args.config == "c2" else config_c1()
    mode = wfl.MODE_PHYSICAL if args.mode == "physical" else wfl.MODE_REFERENCE
    t0 = time.perf_counter()
    x, y = wfl.gen_loops(cfg["config_id"], cfg["Nl"], cfg["N"])
    loops = wfl.Loops(x, y)
    print(f"[bench] loops {cfg['Nl']}x{cfg['N']} generated+uploaded in {time.perf_counter() - t0:.1f}s",
          file=sys.stderr, flush=True)
    group = 128
    geoms = [wfl.Geometry(kind=wfl.GAUSSIAN, F=F, l2=l2) for F, l2 in cfg["geoms"]]
    tg, tcx, tcy, tT = cfg["tup"]
    ntup = len(tg)
    pe_step = cfg["Nl"] * cfg["N"] * ntup

    def make(prec, m=mode):
        return wfl.Plan(wfl.Batch(geoms, tg, tcx, tcy, tT, mode=m, precision=prec, nsub=cfg["nsub"],
                                  spinor=cfg["spinor"], group=group, mapping=1), loops)

    def timed(plan, steps, warmup, label):
        for _ in range(warmup):
            plan.run()
        wfl.check(wfl.lib.wfl_device_sync())
        kms = []
        t0 = time.perf_counter()
        for s in range(steps):
            kms.append(plan.run_timed())
            print(f"[bench] {label} step {s}: kernel {kms[-1]:.1f} ms", file=sys.stderr, flush=True)
        wfl.check(wfl.lib.wfl_device_sync())
        wall = time.perf_counter() - t0
        return wall / steps, float(np.mean(kms)) / 1e3

    mixed = make(wfl.MIXED)
    ms_step, k_mixed = timed(mixed, args.steps, args.warmup, "mixed")
    gs_mixed, _ = mixed.read()
    nonfinite = mixed.nonfinite()
    mixed.close()
    value = pe_step / ms_step
    result = {
        "metric": "worldline point-evals/s, mixed vs FP64 on 1 GPU, at <=1e-6 rel. err in L_eff",
        "value": value,
        "unit": "point-evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32 segment math, fp64 sums",
        "data": "synthetic (seeded Philox unit loops)",
        "config": {"workload": cfg["name"], "loops": cfg["Nl"], "points": cfg["N"], "geometries": len(geoms),
                   "tuples": ntup, "simpson_n": cfg["nsub"], "segment_mode": args.mode, "group": group,
                   "point_evals_per_step": pe_step, "parallelism": "single GPU"},
    }
    fl = flops_per_point_eval(cfg["spinor"], cfg["nsub"])
    achieved = fl * pe_step / k_mixed
    result["roofline"] = {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32_VECTOR / 1e12,
                          "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_VECTOR, "traffic": None,
                          "kernel": "k_tuple<float,GAUSSIAN,spinor,physical,n=2>",
                          "kernel_ms": k_mixed * 1e3, "flops_per_point_eval": fl}
    tr = os.path.join(ROOT, "profiles", f"traffic_{cfg['name']}.json")
    if os.path.exists(tr):
        with open(tr) as fh:
            result["roofline"]["traffic"] = json.load(fh).get("hbm_bytes_per_launch")
    result["nonfinite_loops"] = nonfinite

    if not args.no_fp64:
        fp64 = make(wfl.FP64)
        ms64, k64 = timed(fp64, args.fp64_steps, 0, "fp64")
        gs64, _ = fp64.read()
        fp64.close()
        result["fp64_value"] = pe_step / ms64
        result["mixed_over_fp64"] = value / (pe_step / ms64)
        result["fp64_kernel_ms"] = k64 * 1e3
        # accuracy of mixed against FP64 on the full workload
        ev_m = gs_mixed.sum(axis=1)
        ev_d = gs64.sum(axis=1)
        w_m = 1.0 + (8 / 3 * ev_m[:, 0] - 2 * ev_m[:, 1] + 1 / 3 * ev_m[:, 2]) / cfg["Nl"]
        w_d = 1.0 + (8 / 3 * ev_d[:, 0] - 2 * ev_d[:, 1] + 1 / 3 * ev_d[:, 2]) / cfg["Nl"]
        acc = {"W_max_rel_err": float(np.max(np.abs(w_m - w_d) / np.abs(w_d)))}
        for policy, tag in ((1, "tfunc_policy"), (0, "worldline_all_T")):
            Lm = leff_table(wfl, cfg, gs_mixed, group, policy)
            Ld = leff_table(wfl, cfg, gs64, group, policy)
            mean_m, mean_d = Lm.mean(-1), Ld.mean(-1)
            sem_d = Ld.std(-1, ddof=1) / np.sqrt(Ld.shape[-1])
            rel = np.abs(mean_m - mean_d) / np.maximum(np.abs(mean_d), 1e-300)
            acc[f"Leff_{tag}_max_rel_err"] = float(rel.max())
            acc[f"Leff_{tag}_median_rel_err"] = float(np.median(rel))
            acc[f"Leff_{tag}_max_err_over_SEM"] = float(np.max(np.abs(mean_m - mean_d) / np.maximum(sem_d, 1e-300)))
            acc[f"action_{tag}_rel_err"] = [float(abs(Lm[g].mean(-1).sum() - Ld[g].mean(-1).sum()) /
                                                  max(abs(Ld[g].mean(-1).sum()), 1e-300)) for g in range(len(geoms))]
        result["accuracy_mixed_vs_fp64"] = acc
    if not args.no_ref_mode and mode == wfl.MODE_PHYSICAL:
        ref = make(wfl.MIXED, wfl.MODE_REFERENCE)
        msr, kr = timed(ref, 1, 0, "mixed-reference-mode")
        ref.close()
        result["reference_mode_mixed_value"] = pe_step / msr
    loops.close()
    if not args.no_cpu:
        print("[bench] cpu baseline...", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(cfg, x, y)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
