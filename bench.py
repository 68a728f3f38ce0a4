#!/usr/bin/env python3
"""Benchmark: factor + solve of the 3D P1 heat system on 61^3 nodes (60^3 hexes x 6 Kuhn tets, log-uniform
conductivity in [1, 100], Dirichlet on z = 0, 223,260 free DOFs) on one MI355X.

One step = device numeric factorization of the values already resident in HBM + one triangular solve of a resident
right-hand side. value = algorithmic factor flops (sum of colCount^2, the reference's own figure) / step time.
Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "applied-multiphysics_amd"))

METRIC = "factor+solve s & TFLOP/s vs peak, 3D P1 heat 61³ nodes, FP64 vs mixed @1e-10 resid"
PEAK_FP64_TFLOPS = 78.6   # MI355X dense FP64 matrix
PEAK_FP32_TFLOPS = 157.3  # MI355X dense FP32 matrix


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(pat, values, target_flops):
    """The oracle (single-threaded C restatement of the reference factor) on one complete subtree of the same
    elimination tree, chosen with about target_flops of work; returns the rate and the sample description."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import oracle as O

    S = O.Sym(pat)
    cc2 = S.colCount.astype(np.float64) ** 2
    colflops = np.concatenate([[0.0], np.cumsum(cc2)])
    # subtree of s = postorder range [first[s], s]
    first = np.arange(S.nsuper)
    for s in range(S.nsuper):
        p = S.superParent[s]
        if p >= 0:
            first[p] = min(first[p], first[s])
    sub = colflops[S.superStart[1:]] - colflops[S.superStart[first]]
    root = int(np.argmin(np.abs(sub - target_flops)))
    t0 = time.perf_counter()
    ok, fl, _ = O.chol_factor_subtree(S, values, root)
    dt = time.perf_counter() - t0
    m = int(S.Sp[root + 1] - S.Sp[root])
    return {"value": fl / dt / 1e12, "unit": "TFLOP/s", "cores": 1, "kind": "port",
            "sample": f"oracle factor of the subtree of supernode {root} (front m={m}, {first[root]}..{root}) of the "
                      f"same 61^3 elimination tree: {fl:.3e} flops in {dt:.1f} s; full factor "
                      f"({S.flops:.3e} flops) at that rate: {S.flops / (fl / dt):.0f} s",
            "ok": bool(ok), "est_full_factor_s": S.flops / (fl / dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ncell", type=int, default=60)
    ap.add_argument("--cpu-flops", type=float, default=6e10, help="work of the CPU-baseline sample (0: skip)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}): the path is single-device by design",
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import scipy.sparse as sp
    import torch

    import ampsolve as amp
    from ampsolve import workloads as W

    if not torch.cuda.is_available():
        raise SystemExit("bench.py needs a HIP device")
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    pat, values, b = W.heat3d(args.ncell)
    log(f"workload heat3d ncell={args.ncell}: n={pat['n']} nnz={pat['nnz']} ({time.perf_counter() - t0:.1f} s)")
    t0 = time.perf_counter()
    sym = amp.analyze(pat)
    log(f"analyze: nsuper={sym.nsuper} levels={sym.nlevels} nnzL={sym.nnzL} stored={sym.nnzLStored} "
        f"flops={sym.flops:.3e} maxFront={sym.maxFront} ({time.perf_counter() - t0:.1f} s)")
    num = amp.NumericFactor(sym, 0, "hip")
    dv = torch.from_numpy(values).to(dev)
    db = torch.from_numpy(b).to(dev)
    dx = torch.empty_like(db)
    for i in range(args.warmup):
        ok = num.factor(dv)
        num.solve(db, dx)
        log(f"warmup {i}: ok={ok} factor {num.info['ms']:.2f} ms")
    num.set_profiling(True)
    torch.cuda.synchronize()
    fms, gemm_ms, gemm_n = [], 0.0, 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        num.factor(dv)
        num.solve(db, dx)
        fms.append(num.info["ms"])
        gm, gn = num.kernel_ms("gemm_tile")
        gemm_ms += gm
        gemm_n += gn
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    log(f"timed: {1e3 * dt:.2f} ms/step (factor {np.mean(fms):.2f} ms)")
    kinds = {k: num.kernel_ms(k) for k in ("zero", "assemble", "small_front", "extend_add", "diag_block", "gemm_tile")}
    log("kernel classes (last step): " + ", ".join(f"{k} {v[0]:.2f} ms/{v[1]}" for k, v in kinds.items()))
    x = dx.cpu().numpy()
    A = sp.csr_matrix((values, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))
    relres = float(np.linalg.norm(A @ x - b) / np.linalg.norm(b))
    log(f"relative residual {relres:.3e}")

    gemm_flops = sym.gemmFlops
    achieved = gemm_flops * args.steps / (gemm_ms / 1e3) / 1e12 if gemm_ms > 0 else 0.0
    result = {
        "metric": METRIC,
        "value": sym.flops / dt / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * dt,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"heat3d P1 tets {args.ncell}^3 hexes x 6, {pat['n']} free DOFs, log-uniform k, "
                               f"Dirichlet z=0, FP64 factor + 1 solve", "n": pat["n"], "nnzA": sym.nnzA,
                   "nnzL": sym.nnzL, "factor_flops": sym.flops, "path": "fp64"},
        "factor_ms": float(np.mean(fms)),
        "relres": relres,
        "roofline": {"bound": "mfma", "kernel": "k_gemm_tile", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "launches_per_step": gemm_n // max(args.steps, 1),
                     "flops_per_step": gemm_flops},
    }
    if args.cpu_flops > 0:
        log("cpu baseline sample ...")
        result["cpu_baseline"] = cpu_baseline(pat, values, args.cpu_flops)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
