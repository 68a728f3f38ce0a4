"""Fused solve launches (one launch per direction over a run of levels of tiled fronts, the gather and initialisation
tasks among the flow tasks, handed over through counters) against the per-level launches (fusedSolve = 0) on the 3D
heat system of 20^3 nodes, whose nine top levels form a run: the same tasks in the same arithmetic order, so the
solutions are bitwise equal, for one and three right-hand sides (the kernels' single- and multi-right-hand-side
instances), Cholesky and LDL^T, FP64 and FP32 factors; and the oracle's solution within 1e-12 (FP64) or 1e-3 (FP32
factor), relative in the max norm."""
import numpy as np
import pytest

import oracle as O
import ampsolve as amp
from ampsolve import workloads as W

pytestmark = pytest.mark.gpu

PRM = {"smallMax": 96, "smallNarrow": 32, "wideLaunch": 512, "superPanel": 256, "flowWideMin": 160, "fusedPanel": 1}


def relmax(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("kind", ["hip", "hip-f32"])
@pytest.mark.parametrize("mode", [0, 1])
def test_fused_solve_matches_per_level_launches(kind, mode):
    pat, v = W.heat3d(20)[:2]
    n = pat["n"]
    B = np.concatenate([2 * W.lcg_uniform(21 + k, n) - 1 for k in range(3)])
    out = {}
    for fs in (3, 0):
        print(f"fused solve: kind {kind} mode {mode} n {n} fusedSolve {fs}", flush=True)
        sym = amp.analyze(pat)
        sym.set_params(dict(PRM, fusedSolve=fs))
        num = amp.NumericFactor(sym, mode, kind)
        assert num.factor(v)
        out[fs] = (num.solve(B[:n]), num.solve(B))
        del num
    assert np.array_equal(out[3][0], out[0][0]) and np.array_equal(out[3][1], out[0][1])
    S = O.Sym(pat)
    tol = 1e-12 if kind == "hip" else 1e-3
    if mode == 0:
        _, L = O.chol_factor(S, v)
        xo = [O.chol_solve(S, L, B[k * n:(k + 1) * n]) for k in range(3)]
    else:
        _, L, D, _, _ = O.ldlt_factor(S, v)
        xo = [O.ldlt_solve(S, L, D, B[k * n:(k + 1) * n]) for k in range(3)]
    assert relmax(out[3][0], xo[0]) <= tol
    for k in range(3):
        assert relmax(out[3][1][k * n:(k + 1) * n], xo[k]) <= tol
