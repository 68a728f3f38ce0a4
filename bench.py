#!/usr/bin/env python3
"""Benchmark: factor + solve of the 3D P1 heat system on 61^3 nodes (60^3 hexes x 6 Kuhn tets, log-uniform
conductivity in [1, 100], Dirichlet on z = 0, 223,260 free DOFs) on one MI355X, to a relative residual <= 1e-10.

Two paths, each timed as one step = numeric factorization of values already resident in HBM + solve of a resident
right-hand side down to |b - A x| / |b| <= 1e-10:
  fp64   FP64 factor on the matrix cores, one FP64 solve (no refinement needed);
  mixed  FP32 factor on the matrix cores, FP64 residual / update with GMRES-IR around the FP32 solves.
value = algorithmic factor flops (sum of colCount^2, the figure the reference divides by its factor time) / step time
of the faster path. Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "applied-multiphysics_amd"))

METRIC = "factor+solve s & TFLOP/s vs peak, 3D P1 heat 61³ nodes, FP64 vs mixed @1e-10 resid"
PEAK = {"fp64": 78.6, "mixed": 157.3}  # MI355X dense matrix peaks (TFLOP/s): FP64, FP32
PEAK_HBM = 8000.0  # MI355X HBM3E (GB/s)
KERNEL_CLASSES = ("zero", "assemble", "small_front", "extend_add", "diag_block", "panel_step", "gemm_tile", "split_panel",
                  "gemm_bf16x3", "solve_small_fwd", "solve_gather", "solve_fwd_flow", "solve_small_bwd", "solve_bwd_gather",
                  "solve_bwd_init", "solve_bwd_flow", "solve_fwd_fused", "solve_bwd_fused", "solve_total")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(pat, values, b, target_flops):
    """The oracle (the C restatement of the reference's factor and solve, bitwise the same for any thread count) on
    the host cores: the complete factor + solve of the same system (timed as real-front-gate.mjs:68-114 times its
    legs: ordering + symbolic, factor, solve; one run, all cores), and a single-thread sample (the reference's factor
    is single-threaded) on one complete subtree of about target_flops. Returns (record, oracle solution)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    model, ncpu = cpu_info()
    t0 = time.perf_counter()
    S = O.Sym(pat)
    t_sym = time.perf_counter() - t0
    O.set_threads(ncpu)
    t0 = time.perf_counter()
    ok, Lx = O.chol_factor(S, values)
    t_fac = time.perf_counter() - t0
    t0 = time.perf_counter()
    x = O.chol_solve(S, Lx, b) if ok else None
    t_sol = time.perf_counter() - t0
    del Lx
    log(f"cpu baseline: oracle analyze {t_sym:.2f} s, factor {t_fac:.1f} s ({S.flops / t_fac / 1e9:.1f} GFLOP/s on "
        f"{ncpu} threads), solve {t_sol:.2f} s")
    # single-thread sample: one complete subtree of the same elimination tree
    O.set_threads(1)
    colflops = np.concatenate([[0.0], np.cumsum(S.colCount.astype(np.float64) ** 2)])
    first = np.arange(S.nsuper)
    for s in range(S.nsuper):  # subtree of s = postorder range [first[s], s]
        p = S.superParent[s]
        if p >= 0:
            first[p] = min(first[p], first[s])
    sub = colflops[S.superStart[1:]] - colflops[S.superStart[first]]
    root = int(np.argmin(np.abs(sub - target_flops)))
    t0 = time.perf_counter()
    ok1, fl, _ = O.chol_factor_subtree(S, values, root)
    dt = time.perf_counter() - t0
    m = int(S.Sp[root + 1] - S.Sp[root])
    rec = {"value": S.flops / (t_fac + t_sol) / 1e12, "unit": "TFLOP/s", "cores": ncpu, "kind": "port",
           "cpu": model, "nproc": ncpu, "ok": bool(ok),
           "factor_s": t_fac, "solve_s": t_sol, "e2e_s": t_sym + t_fac + t_sol, "ordering_symbolic_s": t_sym,
           "runs": 1,
           "sample": f"oracle FP64 factor + solve of the complete 61^3 system ({S.flops:.3e} factor flops) on {ncpu} "
                     f"threads, one run: factor {t_fac:.1f} s, solve {t_sol:.2f} s; value = factor flops / (factor + "
                     f"solve)",
           "single_thread": {"value": fl / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "ok": bool(ok1),
                             "sample": f"oracle FP64 factor of the complete subtree of supernode {root} (root front "
                                       f"m={m}, supernodes {first[root]}..{root}): {fl:.3e} flops in {dt:.1f} s",
                             "full_factor_at_that_rate_s": S.flops / (fl / dt)}}
    return rec, x


def gemm_traffic(path, summary="gemm_tile_heat3d61_hbm_traffic.txt"):
    """HBM bytes per launch of the product kernel of the given path, from a committed counter summary (rocprofv3
    --pmc FETCH_SIZE / WRITE_SIZE runs of this benchmark, scaled by the same counters on a plain copy), or None."""
    f = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(f):
        return None
    for line in open(f):
        k, _, v = line.partition(":")
        if k.strip() == f"bytes_per_launch_{path}":
            return float(v)
    return None


def cpu_info():
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, len(os.sched_getaffinity(0))


def run_path(name, sym, dv, db, dx, steps, warmup, tol, log_prefix, method="ir", graph=False, ir_steps=2):
    """One path: `steps` timed factor + solve steps (no per-launch events in the timed region), then one profiled
    step for the kernel breakdown and the k_gemm_tile figures. graph: the step is one replay of the captured factor +
    solve (+ ir_steps refinement steps for the mixed path) sequence (amp_graph_*); the residual of its result is
    checked by the caller against the tolerance like every path's."""
    import numpy as np

    import ampsolve as amp
    from ampsolve import device as D

    kind = "hip" if name.startswith("fp64") else "hip-bf16x3" if name.startswith("bf16x3") else "hip-f32"
    mixed = kind != "hip"
    num = amp.NumericFactor(sym, 0, kind)
    g = num.graph(dv, db, dx, 1, ir_steps if mixed else 0) if graph else None

    def step():
        if g is not None:
            g.launch()
            st = g.sync()
            if not st["ok"]:
                raise RuntimeError(f"{name}: factorization failed ({st})")
            return None
        ok = num.factor(dv)
        if not ok:
            raise RuntimeError(f"{name}: factorization failed ({num.info})")
        if not mixed:
            num.solve(db, dx)
            return None
        if method == "ir":
            _, info = num.refine(dv, db, "ir", maxSteps=30, tol=tol, out=dx)
        else:
            _, info = num.refine(dv, db, "gmres", maxSteps=200, restart=30, tol=tol, out=dx)
        return info

    for i in range(warmup):
        info = step()
        log(f"{log_prefix} warmup {i}" + (f": refine {info}" if info else ""))
    D.synchronize()
    rinfo = []
    t0 = time.perf_counter()
    for _ in range(steps):
        info = step()
        if info:
            rinfo.append(info)
    D.synchronize()
    dt = (time.perf_counter() - t0) / steps
    res = {"ms_per_step": 1e3 * dt, "tflops": sym.flops / dt / 1e12, "graph": bool(graph)}
    if rinfo:
        res["refine_method"] = method
        res["refine_iterations"] = [r["steps"] for r in rinfo]
        res["refine_ms"] = float(np.mean([r["ms"] for r in rinfo]))
    if graph:
        res["refine_method"] = "ir" if mixed else None
        res["refine_iterations"] = [ir_steps] if mixed else None
    if not graph:
        # profiled step: HIP events around every launch, on the stream each launch runs on
        num.set_profiling(True)
        step()
        D.synchronize()
        kinds = {k: num.kernel_ms(k) for k in KERNEL_CLASSES}
        num.set_profiling(False)
        gm, gn = kinds["gemm_tile"]
        bm, bn = kinds["gemm_bf16x3"]
        st_ms, st_n = kinds["solve_total"]
        res["solve_ms_per_pass"] = st_ms / st_n if st_n else 0.0
        res.update({"factor_ms_profiled": float(num.info["ms"]), "gemm_ms_per_step": gm, "gemm_launches_per_step": gn,
                    "bf16x3_ms_per_step": bm, "bf16x3_launches_per_step": bn,
                    "kernel_ms_profiled_step": {k: v[0] for k, v in kinds.items()}})
        log(f"{log_prefix}: {1e3 * dt:.2f} ms/step; profiled step by class: " +
            ", ".join(f"{k} {v[0]:.2f} ms/{v[1]}" for k, v in kinds.items()))
    else:
        log(f"{log_prefix}: {1e3 * dt:.2f} ms/step (captured graph)")
    del g
    return res, num


def hbm_bytes(sym, w, wv, q=1, small_max=96):
    """Algorithmic bytes of one solve pass pair (forward + backward) and of the extend-adds into tiled fronts:
    solve B = 2 w sum_s (nc m - nc (nc - 1) / 2) (L read once per direction) + 2 (4 sum_s mu) (row indices)
    + q 4 wv n (right-hand side in, solution out, two sweeps); extend-add sum over children of tiled fronts
    (m > small_max) of w 3 mu (mu + 1) / 2 (read the update, read and write the parent entries). w: factor bytes
    per entry, wv: solve-vector bytes per entry."""
    import numpy as np

    ss = np.asarray(sym.superStart, np.int64)
    Sp = np.asarray(sym.Sp, np.int64)
    nc, m = np.diff(ss), np.diff(Sp)
    mu = m - nc
    solve = 2 * w * int(np.sum(nc * m - nc * (nc - 1) // 2)) + 2 * 4 * int(mu.sum()) + q * 4 * wv * int(sym.n)
    par = np.asarray(sym.superParent, np.int64)
    has = par >= 0
    big_parent = np.zeros(len(m), bool)
    big_parent[has] = m[par[has]] > small_max
    ea = int(np.sum(w * 3 * mu[big_parent] * (mu[big_parent] + 1) // 2))
    return solve, ea


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ncell", type=int, default=60)
    ap.add_argument("--paths", default="fp64,mixed,bf16x3,fp64-graph,mixed-graph,bf16x3-graph")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--cpu-flops", type=float, default=6e10,
                    help="work of the single-thread CPU sample (0: skip the CPU baseline)")
    ap.add_argument("--refine", default="ir", choices=("ir", "gmres"), help="refinement of the mixed path")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="first run 3 plain device copies of this many MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path is single-device by design",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import scipy.sparse as sp

    import ampsolve as amp
    from ampsolve import device as D
    from ampsolve import workloads as W

    if not D.available():
        raise SystemExit("bench.py needs a HIP device")
    if args.copy_mb > 0:
        src = D.DeviceArray.from_numpy(np.ones(args.copy_mb << 17))
        dst = D.DeviceArray(src.shape)
        for _ in range(3):
            dst.copy_(src)
        D.synchronize()
        log(f"copy calibration: 3 copies of {args.copy_mb} MiB")
        del src, dst
    t0 = time.perf_counter()
    pat, values, b = W.heat3d(args.ncell)
    log(f"workload heat3d ncell={args.ncell}: n={pat['n']} nnz={pat['nnz']} ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    sym = amp.analyze(pat)
    log(f"analyze: nsuper={sym.nsuper} levels={sym.nlevels} nnzL={sym.nnzL} stored={sym.nnzLStored} "
        f"flops={sym.flops:.3e} gemm flops={sym.gemmFlops:.3e} maxFront={sym.maxFront} "
        f"({time.perf_counter() - t0:.1f} s)")
    A = sp.csr_matrix((values, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))
    dv = D.DeviceArray.from_numpy(values)
    db = D.DeviceArray.from_numpy(b)
    paths = {}
    for name in args.paths.split(","):
        dx = D.DeviceArray(db.shape)
        # a captured reduced-precision sequence runs the refinement steps its eager path needed
        eager = paths.get(name[:-len("-graph")], (None,))[0] if name.endswith("-graph") else None
        ir_steps = max(eager["refine_iterations"]) if eager and eager.get("refine_iterations") else 2
        res, num = run_path(name, sym, dv, db, dx, args.steps, args.warmup, args.tol, name, args.refine,
                            graph=name.endswith("-graph"), ir_steps=ir_steps)
        x = dx.numpy()
        res["relres"] = float(np.linalg.norm(A @ x - b) / np.linalg.norm(b))
        paths[name] = (res, x)
        log(f"{name}: relative residual {res['relres']:.3e}")
        del num
    if "fp64" in paths:
        xf = paths["fp64"][1]
        for k in paths:
            if k != "fp64":
                paths[k][0]["rel_diff_vs_fp64"] = float(np.abs(paths[k][1] - xf).max() / np.abs(xf).max())
    ok = {k: v[0] for k, v in paths.items() if v[0]["relres"] <= args.tol}
    best = min(ok, key=lambda k: ok[k]["ms_per_step"]) if ok else next(iter(paths))
    r = paths[best][0]
    # roofline of the dominant kernel, from the profiled step of the same precision's eager path: k_gemm_tile (matrix
    # cores) for the FP64 and FP32 paths; for the bf16 path its bulk products k_gemm_bf16x3, which are bound by their
    # HBM traffic (the trailing matrix read and written once per super-panel)
    eager_of = best[:-len("-graph")] if best.endswith("-graph") else best
    prof = paths.get(eager_of, (r, None))[0]
    gms = prof.get("gemm_ms_per_step", 0.0) + prof.get("bf16x3_ms_per_step", 0.0)
    ach = sym.gemmFlops / (gms / 1e3) / 1e12 if gms > 0 else 0.0
    prec = "fp64" if best.startswith("fp64") else "mixed"
    result = {
        "metric": METRIC,
        "value": r["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else ("f32 factor (bulk updates as 3 bf16 products) + f64 refinement"
                                               if best.startswith("bf16x3") else "f32 factor + f64 refinement"),
        "data": "synthetic",
        "config": {"workload": f"heat3d P1 tets {args.ncell}^3 hexes x 6 Kuhn, {pat['n']} free DOFs, log-uniform k in "
                               f"[1,100], Dirichlet z=0; factor + solve to relres <= {args.tol:g}",
                   "path": best, "n": pat["n"], "nnzA": sym.nnzA, "nnzL": sym.nnzL, "factor_flops": sym.flops},
        "factor_plus_solve_s": r["ms_per_step"] / 1e3,
        "relres": r["relres"],
        "paths": {k: v[0] for k, v in paths.items()},
        "roofline": {"bound": "mfma", "kernel": "k_gemm_tile", "achieved": ach, "peak": PEAK[prec],
                     "unit": "TFLOP/s", "frac": ach / PEAK[prec], "traffic": gemm_traffic(prec),
                     "launches_per_step": prof.get("gemm_launches_per_step"), "flops_per_step": sym.gemmFlops,
                     "measured_on": f"profiled step of the {prec} path (HIP events around each launch)"},
    }
    bms, bn = prof.get("bf16x3_ms_per_step", 0.0), prof.get("bf16x3_launches_per_step", 0)
    if best.startswith("bf16x3") and bms > 0 and bn > 0:
        per = sym.bulkBytes / sym.bulkLaunches  # algorithmic bytes per launch
        gbs = per / (bms / bn / 1e3) / 1e9
        result["roofline"] = {
            "bound": "hbm", "kernel": "k_gemm_bf16x3", "achieved": gbs, "peak": PEAK_HBM, "unit": "GB/s",
            "frac": gbs / PEAK_HBM, "traffic": gemm_traffic("bf16x3", "gemm_bf16x3_heat3d61_hbm_traffic.txt"),
            "bytes_per_launch": per, "launches_per_step": bn, "avg_launch_ms": bms / bn,
            "measured_on": "profiled step of the bf16x3 path (HIP events around each launch)",
            "products_fp32_equivalent": {"flops_per_step": sym.gemmFlops, "ms": gms, "tflops": ach,
                                         "note": "all product launches of the step (bf16 bulk and FP32 lookahead)"}}
    # achieved HBM rates of the scatter (extend-add into tiled fronts) and of the solve pass, same profiled step
    w = 8 if prec == "fp64" else 4
    ks = prof.get("kernel_ms_profiled_step", {})
    solve_b, ea_b = hbm_bytes(sym, w, w)
    hbm = {"peak": PEAK_HBM, "unit": "GB/s", "measured_on": result["roofline"]["measured_on"]}
    if ks.get("extend_add", 0) > 0:
        gbs = ea_b / (ks["extend_add"] / 1e3) / 1e9
        hbm["extend_add"] = {"bytes": ea_b, "ms": ks["extend_add"], "achieved": gbs, "frac": gbs / PEAK_HBM}
    sm = prof.get("solve_ms_per_pass", 0.0)
    if sm > 0:
        gbs = solve_b / (sm / 1e3) / 1e9
        hbm["solve"] = {"bytes": solve_b, "ms": sm, "achieved": gbs, "frac": gbs / PEAK_HBM}
    result["hbm"] = hbm
    if args.cpu_flops > 0:
        log("cpu baseline: oracle factor + solve of the same system on the host cores ...")
        rec, xo = cpu_baseline(pat, values, b, args.cpu_flops)
        result["cpu_baseline"] = rec
        if xo is not None:  # the error of every path's solution against the FP64 oracle
            for k, (res, x) in paths.items():
                res["rel_diff_vs_oracle"] = float(np.abs(x - xo).max() / np.abs(xo).max())
            result["rel_diff_vs_oracle"] = r["rel_diff_vs_oracle"]
            rec["gpu_speedup_factor_plus_solve"] = (rec["factor_s"] + rec["solve_s"]) / (r["ms_per_step"] / 1e3)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
