#!/usr/bin/env python3
"""Throughput of the solve -> sensitivity -> update loop on one MI355X.

Default workload (the largest configured cell, 256^3 = 16.8M elements): the
precision form of the design loop. Density from the seeded cos-mode
generator at volume fraction 0.25, Spline4 r=1.5 density filter,
kappa = 1e-3 + (1 - 1e-3) rho^3; MMA (m = 3) maximising C00 subject to
C11, C22 <= 0.05 and volume <= 0.25. Every step is one MMA design
iteration = filter + SIMP + Galerkin hierarchy rebuild + three-load-case
MG-PCG solve with FP64 iterative refinement to relative residual 1e-8
(FP32 smoothing and transfers, FP16-stored coarse stencils with a per-level
scale, FP64 dots, kappa^H and coarsest inverse) + kappa^H + three
sensitivity contractions through the adjoint filter + the MMA update. One
step is one homogenization + sensitivity solve, so value = solves/s. The
same loop on the plain FP32 path (relative residual 1e-6) is timed beside
it, and the 128^3 OC loop (the plain optimisation configuration) as extra
keys.

Prints one JSON line. --gpus above 1 is refused: this build runs on one GPU.

Other workloads: --workload oc (128^3 OC loop alone), batch (96 cells),
serve (Poisson arrivals), copy (calibration of the byte counters).
--oracle adds the FP64 oracle's kappa^H and sensitivity errors on the first
256^3 design (MG-preconditioned CG on the host, minutes).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "homog. solves/s + p50/p99 latency at κᴴ rel err ≤1e-5; stencil HBM GB/s"
HBM_PEAK_GBS = 8000.0
FINE_KERNEL = "k_fine_smooth_brick<32>"
GRAPH = True  # V-cycles replayed from captured HIP graphs (--no-graph: eager launches)
AUTOTUNE = True  # the tuner picks the fine tile shape and coarse sweeps first (--no-autotune: defaults)


def measured_traffic(kernel, reso):
    """HBM bytes per launch of `kernel` at `reso` from the committed counter
    summary profiles/traffic_fine_brick_<reso>.json (rocprofv3 FETCH_SIZE and
    WRITE_SIZE runs scaled by a copy-kernel calibration; the matching .md
    gives the commands), or None when there is none for this kernel/size."""
    path = os.path.join(ROOT, "profiles", f"traffic_fine_brick_{reso}.json")
    try:
        with open(path) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    if t.get("kernel") != kernel or t.get("reso") != reso:
        return None
    return t.get("hbm_bytes_per_launch")


def host_description():
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "omp_threads": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline_mg(n_full, seed, rtol, mean, fp32=False):
    """The FP64 MG-preconditioned CG of the oracle (OpenMP on the host cores;
    same method: Galerkin V(2,2) multicolour Gauss-Seidel in the reference's
    colour order) on a bounded sample: one full homogenization + sensitivity
    of the same generator's design at half the edge (1/8 of the elements),
    scaled by 8 to the configured cell (assembly, Galerkin setup, smoothing
    and element energies all scale with the element count, and MG-PCG
    iteration counts do not depend on the size: 7 at 64^3 and at 128^3)."""
    import oracle.oracle as orc
    from opentm_amd import workloads as wl
    ns = n_full // 2
    rho = wl.plain_density(ns, seed, mean=mean).astype(np.float64)
    t0 = time.perf_counter()
    rf = orc.filter_fwd(rho, 0, 1.5, False)
    kappa = 1e-3 + (1 - 1e-3) * rf ** 3
    C, _, S, its, rr = orc.mg_homogenize(kappa, rtol=rtol, want_S=True, fp32=fp32)
    for dC in (-np.diag([1.0, 0, 0]), np.diag([0, 1.0, 0]), np.diag([0, 0, 1.0])):
        s = orc.sensitivity(S, dC).reshape(rf.shape)
        orc.filter_adj(s * 3 * (1 - 1e-3) * rf ** 2, 0, 1.5, False)
    dt = time.perf_counter() - t0
    scale = (n_full / ns) ** 3
    host = host_description()
    cores = int(host["omp_threads"] or host["nproc"] or 1)
    return {"value": 1.0 / (dt * scale), "unit": "solves/s", "cores": cores, "kind": "port",
            "sample": f"{'FP32' if fp32 else 'FP64'} MG-PCG (oracle, OpenMP) on the same generator's {ns}^3 design "
                      f"to rtol {rtol:g}: "
                      f"filter + homogenization ({max(its)} PCG iterations) + element energies + 3 sensitivity "
                      f"contractions in {dt:.1f} s, scaled x{scale:g} to {n_full}^3 (linear in the element count)",
            "host": host}


def timed_loop(step, steps, warmup, sync, label, on_timed_start=None):
    """W untimed steps, then exactly K steps bracketed by synchronisations;
    returns (total seconds, per-step seconds, per-step results)."""
    for i in range(warmup):
        step()
        print(f"[bench] {label} warmup {i + 1}/{warmup}", file=sys.stderr, flush=True)
    sync()
    if on_timed_start:
        on_timed_start()
    lat, out = [], []
    t0 = time.perf_counter()
    for i in range(steps):
        ts = time.perf_counter()
        out.append(step())
        lat.append(time.perf_counter() - ts)
        print(f"[bench] {label} step {i + 1}/{steps} {1e3 * lat[-1]:.2f} ms", file=sys.stderr, flush=True)
    sync()
    return time.perf_counter() - t0, lat, out


def roofline_from(prof, reso):
    """Fine brick smoother: algorithmic bytes per launch (28 B/node for the
    zero-start pre-smoother, 40 B/node for the post-smoother, averaged over
    the launches) over the mean launch time from HIP events on the engine's
    stream (profiled_steps)."""
    nl, ms, by = prof["fine_gs"]
    if not nl or ms <= 0:
        return None
    avg_s = ms * 1e-3 / nl
    achieved = (by / nl) / avg_s / 1e9
    return {"bound": "hbm", "kernel": FINE_KERNEL, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "bytes_per_launch": by / nl,
            "avg_launch_us": round(avg_s * 1e6, 3), "launches": nl, "timed_over": f"{PROFILE_STEPS} steps after the timed region",
            "traffic": measured_traffic(FINE_KERNEL, reso)}


MMA_KERNELS = {"trial": "k_mma_point_wide<3, true, false>", "bound": "k_mma_point_wide<3, true, true>",
               "point": "k_mma_point<3, false, false>"}


def measured_mma_traffic(name, reso):
    """HBM bytes per launch of an MMA pass at `reso` from the committed
    counter summary profiles/traffic_mma_<reso>.json (FETCH_SIZE and
    WRITE_SIZE runs scaled by the copy calibration; traffic_256.md gives the
    commands), or None."""
    try:
        with open(os.path.join(ROOT, "profiles", f"traffic_mma_{reso}.json")) as fh:
            t = json.load(fh)
    except (OSError, ValueError):
        return None
    return t.get("passes", {}).get(name, {}).get("hbm_bytes_per_launch")


def mma_rooflines(mprof, reso=None):
    """The MMA subsolver passes: algorithmic bytes per launch (x 4, asymptotes
    16, the FP32 gradients 4 (1 + m), the FP64 interior point 24 bytes per
    element; the trial passes also write the 24-byte trial point) over the
    mean launch time from HIP events on the optimiser's stream."""
    out = {}
    for name, (nl, ms, by) in mprof.items():
        if not nl or ms <= 0:
            continue
        avg_s = ms * 1e-3 / nl
        achieved = (by / nl) / avg_s / 1e9
        out[name] = {"bound": "hbm", "kernel": MMA_KERNELS[name], "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "bytes_per_launch": by / nl, "avg_launch_us": round(avg_s * 1e6, 3), "launches": nl,
                     "total_ms": round(ms, 3), "timed_over": f"{PROFILE_STEPS} steps after the timed region",
                     "traffic": measured_mma_traffic(name, reso)}
    return out


def dominant_roofline(prof, mprof, reso):
    """The roofline of the kernel with the most time among the live-timed
    ones (the fine smoother and the MMA passes)."""
    cands = []
    st = roofline_from(prof, reso)
    if st:
        cands.append((st["avg_launch_us"] * st["launches"], st))
    for r in mma_rooflines(mprof, reso).values():
        cands.append((r["avg_launch_us"] * r["launches"], r))
    return max(cands, key=lambda c: c[0])[1] if cands else None


def class_rows(prof):
    rows = {}
    for name, (nl, ms, by) in prof.items():
        if nl:
            rows[name] = {"launches": nl, "ms": round(ms, 3), "avg_us": round(1e3 * ms / nl, 3),
                          "GB/s": round(by / (ms * 1e-3) / 1e9, 1) if ms > 0 and by > 0 else None}
    return rows


PROFILE_STEPS = 3


def profiled_steps(hom, step, classes, label, mma=None):
    """Kernel-class timing with HIP events on the engine's stream over a few
    steps run right after the timed region (events are recorded around
    eager launches; the timed region replays captured V-cycle graphs), and
    of the MMA passes on the optimiser's stream when mma is given."""
    hom.synchronize()
    hom.profile_enable(True, None if classes == "all" else classes)
    if mma is not None:
        mma.profile_enable(True)
    for i in range(PROFILE_STEPS):
        step()
        print(f"[bench] {label} profiled step {i + 1}/{PROFILE_STEPS}", file=sys.stderr, flush=True)
    prof = hom.profile_read()
    hom.profile_enable(False)
    if mma is not None:
        prof = dict(prof)
        prof["mma"] = mma.profile_read()
        mma.profile_enable(False)
    return prof


def precision_run(n, rho0, refine, steps, warmup, rtol, label, profile_classes=None):
    """The precision-form MMA loop (or, refine=False, the same loop on the
    plain FP32 path) on one cell."""
    from opentm_amd import precision as pf
    loop = pf.PrecisionLoop(n, rho0, rtol=rtol, refine=refine, fp16=refine, graph=GRAPH)
    tuned = None
    if AUTOTUNE:
        # after one design iteration the context holds the next design and
        # the previous solution: the tuner times every (fine tile shape,
        # coarse sweeps) candidate on that design from that warm start, as
        # the loop's next solve would run
        loop.step()
        loop.hom.apply_design()
        rows, chosen = loop.hom.autotune(warm=True)
        tuned = {"chosen": rows[chosen], "candidates": rows}
        print(f"[bench] {label} autotune chose {rows[chosen]}", file=sys.stderr, flush=True)

    total, lat, hist = timed_loop(loop.step, steps, warmup, loop.hom.synchronize, label, loop.history.clear)
    hist = list(hist)
    prof = profiled_steps(loop.hom, loop.step, profile_classes, label, loop.mma) if profile_classes is not None else None
    res = {"solves_per_s": steps / total, "ms_per_step": 1e3 * total / steps,
           "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
           "pcg_iters_mean": float(np.mean([max(h["pcg_iters"]) for h in hist])),
           "relres_max": float(max(max(h["relres"]) for h in hist)), "rtol": rtol,
           "C_final": hist[-1]["C"].reshape(-1).tolist(), "g_final": hist[-1]["g"].tolist()}
    if tuned:
        res["autotune"] = tuned
        res["coarse_sweeps"] = tuned["chosen"]["coarse_sweeps"]
    else:
        res["coarse_sweeps"] = 2  # the library default: as many as on the fine level
    res["fine_tile"] = loop.hom.fine_tile()[1]
    loop.mma.close()
    loop.hom.close()
    for b in [loop.x, loop.df] + loop.dg:
        b.free()
    return res, prof


def oc_loop(n, seed, steps, warmup, rtol, coarse_sweeps, profile_classes=None, verbose=False, min_coarse=None):
    """The plain optimisation loop: OC maximising trace(kappa^H) at volume 0.3."""
    import opentm_amd as otm
    from opentm_amd import workloads as wl
    cfg = otm.HomoConfig(n, (1.0, 1e-3))
    cfg.femRelThres = rtol
    cfg.coarse_sweeps = coarse_sweeps
    cfg.graph_vcycle = 1 if GRAPH else 0
    if min_coarse is not None:
        cfg.min_coarse = min_coarse
    cfg.verbose = 1 if verbose else 0
    hom = otm.Homogenization_H(cfg)
    hom.set_design(wl.oc_start_density(n, seed))
    oc = otm.OCOptimizer(n ** 3, 0.001, 0.2, 0.5, bisect_steps=60, bisect_tol=1e-6)
    dC = -np.eye(3)
    params = otm.DensityParams(0, 1.5, False)

    def step():
        C = hom.eval(params)
        st = hom.solve_stats()
        hom.backward(dC, want=False)
        oc.step(hom, 2.0, 0.3)
        return C, st

    total, lat, out = timed_loop(step, steps, warmup, hom.synchronize, f"oc{n}")
    prof = profiled_steps(hom, step, profile_classes, f"oc{n}") if profile_classes is not None else None
    res = {"solves_per_s": steps / total, "ms_per_step": 1e3 * total / steps,
           "p50_ms": 1e3 * float(np.percentile(lat, 50)), "p99_ms": 1e3 * float(np.percentile(lat, 99)),
           "pcg_iters_mean": float(np.mean([max(st[0]) for _, st in out])),
           "relres_max": float(max(max(st[1]) for _, st in out)),
           "kappaH_final": [float(v) for v in out[-1][0].reshape(-1)]}
    oc.close()
    hom.close()
    return res, prof


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30, help="timed design iterations")
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--reso", type=int, default=None, help="cell edge (default: 256 precision, 128 oc)")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--rtol", type=float, default=1e-6, help="relative residual of the plain FP32 path")
    ap.add_argument("--coarse-sweeps", type=int, default=1,
                    help="Gauss-Seidel sweeps per smoothing step on the coarse levels (fine level: 2)")
    ap.add_argument("--min-coarse", type=int, default=None,
                    help="coarsest level edge floor (otm_config.min_coarse); default: the library's")
    ap.add_argument("--verbose", action="store_true", help="library prints per-solve PCG progress")
    ap.add_argument("--profile-all", action="store_true",
                    help="time every kernel class with HIP events (adds event overhead to the step time)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true", help="precision workload: skip the plain-path and 128^3 runs")
    ap.add_argument("--workload", choices=["precision", "oc", "batch", "serve", "copy"], default="precision",
                    help="copy: three 256 MiB device copies, the calibration run for rocprofv3 byte counters")
    ap.add_argument("--qps", default="100,200,300,400", help="serve workload: arrival rates (requests/s)")
    ap.add_argument("--requests", type=int, default=2000, help="serve workload: requests per arrival rate")
    ap.add_argument("--plain-requests", type=int, default=1000,
                    help="serve workload: requests of the same mix run through the plain single-solve path")
    ap.add_argument("--hbm-gb", type=float, default=64.0, help="serve workload: HBM budget of the server")
    ap.add_argument("--max-active", type=int, default=96, help="serve workload: requests on the GPU at once")
    ap.add_argument("--cells", type=int, default=96, help="batch workload: number of cells")
    ap.add_argument("--oracle", action="store_true", help="precision workload: FP64 oracle errors (slow)")
    ap.add_argument("--no-graph", action="store_true", help="launch every V-cycle kernel eagerly (no HIP graph replay)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="precision workload: keep the default fine tile shape and coarse sweeps (no autotuner run "
                         "on the first design before the timed loop)")
    args = ap.parse_args()
    global GRAPH, AUTOTUNE
    GRAPH = not args.no_graph
    AUTOTUNE = not args.no_autotune
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (no multi-GPU path is built)"}))
        sys.exit(2)
    if args.workload == "precision":
        precision_main(args)
    elif args.workload == "oc":
        oc_main(args)
    elif args.workload == "batch":
        batch_main(args)
    elif args.workload == "serve":
        serve_main(args)
    else:
        import opentm_amd as otm
        from opentm_amd import _capi
        n = 64 * 1024 * 1024  # floats: 256 MiB per buffer
        a, b = otm.DeviceBuffer(n), otm.DeviceBuffer(n)
        for _ in range(3):
            _capi.check(_capi.load().otm_memcpy(b.ptr, a.ptr, 4 * n, 1, 1), "otm_memcpy")
        a.free()
        b.free()
        print(json.dumps({"workload": "copy", "bytes_read_per_copy": 4 * n, "bytes_written_per_copy": 4 * n,
                          "copies": 3}))


def precision_main(args):
    from opentm_amd import workloads as wl
    n = args.reso or 256
    rho0 = wl.plain_density(n, args.seed, mean=0.25)
    classes = "all" if args.profile_all else ["fine_gs"]
    prec, prof = precision_run(n, rho0, True, args.steps, args.warmup, 1e-8, f"precision{n}", classes)
    out = {
        "metric": METRIC, "value": prec["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": prec["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32 smoother/transfers, fp16 coarse stencils, fp64 refinement",
        "data": "synthetic",
        "config": {"workload": f"{n}^3 precision form: MMA (m=3) max C00 s.t. C11, C22 <= 0.05, volume <= 0.25; "
                               f"one step = one homogenization (3 load cases, FP64 refinement to 1e-8) + "
                               f"3 sensitivities + MMA update",
                   "reso": n, "elements": n ** 3, "seed": args.seed, "rhs_per_solve": 3,
                   "mg": f"V(2,2) fine in {prec['fine_tile'][0]}x{prec['fine_tile'][1]}x{prec['fine_tile'][2]} LDS "
                         f"tiles, {prec['coarse_sweeps']} sweep(s) on the coarse levels, 8-colour GS in LDS bricks, "
                         f"Galerkin coarse levels",
                   "parallelism": "single GPU"},
        "p50_ms": prec["p50_ms"], "p99_ms": prec["p99_ms"], "pcg_iters_mean": prec["pcg_iters_mean"],
        "relres_max": prec["relres_max"], "C_final": prec["C_final"], "g_final": prec["g_final"],
        "roofline": dominant_roofline(prof, prof["mma"], n), "stencil_roofline": roofline_from(prof, n),
        "mma_passes": mma_rooflines(prof["mma"], n),
        "kernel_classes": class_rows({k: v for k, v in prof.items() if k != "mma"}),
    }
    if "autotune" in prec:
        out["autotune"] = prec["autotune"]
    if not args.no_extras:
        plain, _ = precision_run(n, rho0, False, args.steps, args.warmup, args.rtol, f"plain{n}")
        out["plain_fp32"] = {k: plain[k] for k in ("solves_per_s", "ms_per_step", "p50_ms", "p99_ms",
                                                   "pcg_iters_mean", "relres_max", "rtol")}
        out["time_vs_plain"] = prec["ms_per_step"] / plain["ms_per_step"]
        oc, ocprof = oc_loop(128, args.seed, 20, 5, args.rtol, args.coarse_sweeps, ["fine_gs"])
        oc["roofline"] = roofline_from(ocprof, 128)
        out["config1_oc_loop_128"] = oc
    if args.oracle:
        import oracle.oracle as orc
        print("[bench] FP64 oracle (MG-PCG) on the first design ...", file=sys.stderr, flush=True)
        out["oracle"] = oracle_errors(orc, n, rho0)
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline (FP64 MG-PCG oracle, bounded sample) ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline_mg(n, args.seed, 1e-8, 0.25)
        # the plain path's method on the host: FP32 MG-PCG to the plain tolerance
        out["cpu_baseline_fp32"] = cpu_baseline_mg(n, args.seed, args.rtol, 0.25, fp32=True)
    print(json.dumps(out))


def oracle_errors(orc, n, rho0):
    """kappa^H and per-element sensitivity errors of the first precision-form
    solve (and of the plain FP32 path) against the FP64 oracle on the same
    design."""
    import opentm_amd as otm
    t0 = time.perf_counter()
    rf = orc.filter_fwd(rho0.astype(np.float64), 0, 1.5, False)
    kap = 1e-3 + (1 - 1e-3) * rf ** 3
    Co, _, So, its, rr = orc.mg_homogenize(kap, rtol=1e-11, want_S=True)
    dt = time.perf_counter() - t0
    res = {"solver": "FP64 MG-PCG", "rtol": 1e-11, "relres": rr, "iterations": its, "seconds": dt,
           "C": Co.reshape(-1).tolist()}
    dCs = {"00": np.diag([1.0, 0, 0]), "11": np.diag([0, 1.0, 0]), "22": np.diag([0, 0, 1.0])}
    for label, refine, rtol in (("precision", 1, 1e-8), ("plain", 0, 1e-6)):
        cfg = otm.HomoConfig(n, (1.0, 1e-3))
        cfg.femRelThres = rtol
        cfg.fp64_refine = refine
        cfg.fp16_stencils = refine
        hom = otm.Homogenization_H(cfg)
        hom.update(kap.astype(np.float32))
        C = hom.heatMatrix()
        e = {"kappaH_rel_err": float(np.abs(C - Co).max() / np.abs(Co).max())}
        for name, dC in dCs.items():
            s = hom.Sensitivity(dC).astype(np.float64)
            so = orc.sensitivity(So, dC)
            floor = 1e-3 * np.abs(so).max()
            e[f"sens{name}_max_elementwise_rel_err"] = float((np.abs(s - so) / np.maximum(np.abs(so), floor)).max())
            e[f"sens{name}_normwise_rel_err"] = float(np.abs(s - so).max() / np.abs(so).max())
        e["relres"] = hom.solve_stats()[1]
        res[label] = e
        hom.close()
    return res


def oc_main(args):
    n = args.reso or 128
    classes = "all" if args.profile_all else ["fine_gs"]
    res, prof = oc_loop(n, args.seed, args.steps, args.warmup, args.rtol, args.coarse_sweeps, classes,
                        args.verbose, args.min_coarse)
    out = {"metric": METRIC, "value": res["solves_per_s"], "unit": "solves/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"{n}^3 OC loop, max trace(kappaH) at volume 0.3 (plain optimisation loop)",
                      "reso": n, "rhs_per_solve": 3, "rtol": args.rtol, "seed": args.seed,
                      "parallelism": "single GPU"},
           "roofline": roofline_from(prof, n), "kernel_classes": class_rows(prof)}
    out.update({k: res[k] for k in ("p50_ms", "p99_ms", "pcg_iters_mean", "relres_max", "kappaH_final")})
    print(json.dumps(out))


def batch_cells(count, seed):
    """Batched-form cells: edge from {32, 48, 64, 96, 128} with probabilities
    {.3, .25, .2, .15, .1}, cos-mode density (plain generator) at a volume
    fraction drawn from U[0.15, 0.5], all from one Philox seed."""
    rng = np.random.Generator(np.random.Philox(int(seed) + 7919))
    sizes = rng.choice([32, 48, 64, 96, 128], size=count, p=[0.3, 0.25, 0.2, 0.15, 0.1])
    vols = rng.uniform(0.15, 0.5, size=count)
    seeds = rng.integers(0, 2 ** 31 - 1, size=count)
    return [(int(n), float(v), int(sd)) for n, v, sd in zip(sizes, vols, seeds)]


def batch_main(args):
    import opentm_amd as otm
    from opentm_amd import workloads as wl
    cells = batch_cells(args.cells, args.seed)
    print(f"[bench] batch of {len(cells)} cells, sizes {sorted(set(c[0] for c in cells))}", file=sys.stderr, flush=True)
    homs, kbufs, sbufs = [], [], []
    for n, vol, sd in cells:
        cfg = otm.HomoConfig(n, (1.0, 1e-3))
        cfg.femRelThres = args.rtol
        cfg.coarse_sweeps = args.coarse_sweeps
        h = otm.Homogenization_H(cfg)
        k = wl.simp(wl.plain_density(n, sd, mean=vol))
        homs.append(h)
        kbufs.append(otm.DeviceBuffer.from_numpy(k))
        sbufs.append(otm.DeviceBuffer(n ** 3))
    dC = -np.eye(3)
    dofs = sum(n ** 3 for n, _, _ in cells)

    def looped():
        for h, kb, sb in zip(homs, kbufs, sbufs):
            h.update(kb)
            h.heatMatrix()
            h.Sensitivity(dC, sens=sb)

    batch = otm.Batch()

    def batched():
        batch.update(homs, kbufs)
        batch.heat_matrix(homs)
        batch.sensitivity(homs, [dC] * len(homs), outs=sbufs)

    def timed(fn, label):
        for h in homs:
            h.reset_displacement()
        for h in homs:
            h.synchronize()
        t0 = time.perf_counter()
        fn()
        for h in homs:
            h.synchronize()
        dt = time.perf_counter() - t0
        its = [max(h.solve_stats()[0]) for h in homs]
        print(f"[bench] {label}: {1e3 * dt:.1f} ms, mean PCG its {np.mean(its):.2f}", file=sys.stderr, flush=True)
        return dt, its

    for i in range(max(args.warmup, 1)):
        timed(looped, "warmup looped")
        timed(batched, "warmup batched")
    t_loop, its_loop = [], None
    t_bat, its_bat = [], None
    for i in range(args.steps if args.steps != 10 else 3):
        dt, its_loop = timed(looped, "looped")
        t_loop.append(dt)
        dt, its_bat = timed(batched, "batched")
        t_bat.append(dt)
    C_b = batch.heat_matrix(homs)
    C_l = np.stack([h.heatMatrix() for h in homs])
    launches, steps = batch.stats()
    tb, tl = float(np.median(t_bat)), float(np.median(t_loop))
    out = {"metric": "homog. solves/s (kappaH + sensitivity per cell), batched form", "value": len(cells) / tb,
           "unit": "solves/s", "n_gpus": 1, "steps": len(t_bat), "warmup": max(args.warmup, 1),
           "ms_per_step": 1e3 * tb, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32", "data": "synthetic",
           "config": {"workload": f"{len(cells)} cells, edges from {{32,48,64,96,128}} (p .3/.25/.2/.15/.1), "
                                  f"vol U[0.15,0.5], one kappaH + sensitivity solve each, cold start",
                      "dof_total": dofs, "rtol": args.rtol, "seed": args.seed,
                      "sizes": {str(n): sum(1 for c in cells if c[0] == n) for n in (32, 48, 64, 96, 128)}},
           "looped_plain_solves_per_s": len(cells) / tl, "looped_ms": 1e3 * tl,
           "speedup_vs_looped": tl / tb,
           "pcg_iters_mean": {"batched": float(np.mean(its_bat)), "looped": float(np.mean(its_loop))},
           "max_rel_diff_batched_vs_looped": float(np.max(np.abs(C_b - C_l)) / np.max(np.abs(C_l))),
           "batch_launches_total": launches, "batch_steps_total": steps}
    batch.close()
    for h in homs:
        h.close()
    for b in kbufs + sbufs:
        b.free()
    print(json.dumps(out))


def serve_main(args):
    """Open-loop Poisson arrivals of homogenization + sensitivity requests
    (edges log-uniform over 32..192 in multiples of 16, thresholded periodic
    Gaussian random fields with correlation length U[4,12] voxels, volume
    U[0.1,0.5], projection beta 8), FP16 coarse stencils and FP64
    refinement; p50/p99 latency (arrival -> result) and solves/s per rate."""
    import threading
    import opentm_amd as otm
    from opentm_amd import workloads as wl
    rng = np.random.Generator(np.random.Philox(int(args.seed) + 104729))
    edges = list(range(32, 193, 16))
    pool_per_edge = 2
    t0 = time.perf_counter()
    pool = {}
    for e in edges:
        pool[e] = []
        for j in range(pool_per_edge):
            corr = float(rng.uniform(4.0, 12.0))
            vol = float(rng.uniform(0.1, 0.5))
            rho = wl.grf_density(e, int(rng.integers(0, 2 ** 31 - 1)), corr, vol)
            pool[e].append(wl.simp(rho))
    print(f"[bench] density pool ({len(edges)} edges x {pool_per_edge}) in {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)

    def draw_edge():
        v = float(np.exp(rng.uniform(np.log(32.0), np.log(192.0))))
        return int(min(192, max(32, 16 * round(v / 16.0))))

    def cfg_for(e):
        c = otm.HomoConfig(e, (1.0, 1e-3))
        c.femRelThres = args.rtol
        c.coarse_sweeps = args.coarse_sweeps
        c.fp16_stencils = 1
        c.fp64_refine = 1
        return c

    cfgs = {e: cfg_for(e) for e in edges}
    dC = -np.eye(3)
    sv = otm.Server(hbm_budget_gb=args.hbm_gb, max_active=args.max_active)
    # warm the hierarchy cache with one request per shape (not timed)
    for r in [sv.submit(pool[e][0], cfg=cfgs[e], dC=dC) for e in edges]:
        sv.wait(r)
    levels = []
    for qps in [float(q) for q in args.qps.split(",") if q]:
        n = args.requests
        gaps = rng.exponential(1.0 / qps, size=n)
        picks = [(draw_edge(), int(rng.integers(0, pool_per_edge))) for _ in range(n)]
        ids = [None] * n
        t_start = time.perf_counter()

        def submitter():
            t_next = t_start
            for i in range(n):
                t_next += gaps[i]
                dt = t_next - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                e, j = picks[i]
                ids[i] = sv.submit(pool[e][j], cfg=cfgs[e], dC=dC)

        th = threading.Thread(target=submitter)
        th.start()
        lat, worst_rr = [], 0.0
        for i in range(n):
            while ids[i] is None:
                time.sleep(0.0005)
            C, _, info = sv.wait(ids[i])
            lat.append(info["latency_ms"])
            worst_rr = max(worst_rr, max(info["relres"]))
        t_end = time.perf_counter()
        th.join()
        lv = {"qps_offered": qps, "requests": n, "solves_per_s": n / (t_end - t_start),
              "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
              "mean_ms": float(np.mean(lat)), "relres_max": worst_rr}
        levels.append(lv)
        print(f"[bench] qps {qps:g}: {lv['solves_per_s']:.1f} solves/s p50 {lv['p50_ms']:.1f} ms "
              f"p99 {lv['p99_ms']:.1f} ms", file=sys.stderr, flush=True)
    st = sv.stats()
    sv.close()
    # the same request mix through the plain single-solve FP32 path, one
    # request after another (contexts cached per shape, as the server does)
    plain_cfg = {}
    for e in edges:
        c = otm.HomoConfig(e, (1.0, 1e-3))
        c.femRelThres = args.rtol
        plain_cfg[e] = c
    homs = {}
    npl = min(args.requests, args.plain_requests)
    picks = [(draw_edge(), int(rng.integers(0, pool_per_edge))) for _ in range(npl)]
    for e in edges:  # hierarchies built before the clock starts, as for the server
        homs[e] = otm.Homogenization_H(plain_cfg[e])
    lat = []
    t_start = time.perf_counter()
    for e, j in picks:
        ts = time.perf_counter()
        h = homs[e]
        h.reset_displacement()
        h.update(pool[e][j])
        h.heatMatrix()
        h.Sensitivity(dC)
        lat.append(1e3 * (time.perf_counter() - ts))
    t_plain = time.perf_counter() - t_start
    for h in homs.values():
        h.close()
    plain = {"requests": npl, "solves_per_s": npl / t_plain, "p50_ms": float(np.percentile(lat, 50)),
             "p99_ms": float(np.percentile(lat, 99)), "path": "plain single-solve FP32 (rtol %g), looped" % args.rtol}
    print(f"[bench] plain single-solve loop: {plain['solves_per_s']:.1f} solves/s", file=sys.stderr, flush=True)
    best = max(levels, key=lambda l: l["solves_per_s"])
    out = {"metric": "homog. solves/s + p50/p99 latency (arrival -> result), serving form", "value": best["solves_per_s"],
           "unit": "solves/s", "n_gpus": 1, "steps": sum(l["requests"] for l in levels), "warmup": len(edges),
           "ms_per_step": 1e3 / best["solves_per_s"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "fp32+fp16 stencils+fp64 refinement", "data": "synthetic",
           "config": {"workload": "open-loop Poisson arrivals, edges 32..192 step 16 log-uniform, thresholded periodic "
                                  "GRF densities (corr U[4,12], vol U[0.1,0.5], beta 8), kappaH + sensitivity",
                      "rtol": args.rtol, "seed": args.seed, "hbm_budget_gb": args.hbm_gb,
                      "max_active": args.max_active},
           "levels": levels, "server": st, "plain_single_solve": plain,
           "served_over_plain": best["solves_per_s"] / plain["solves_per_s"]}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
