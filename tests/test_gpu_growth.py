"""The pivot-growth FP64 fallback of the FP32 factor (amp_num_set_growth_limit): a diagonal block of a tiled front whose
largest elimination multiplier crosses the limit is factored again in FP64 arithmetic with FP64 pivots. On a nearly
singular indefinite grid (the 27-point grid shifted to diagonal 0.01, LDL^T) the limit is crossed; the pivots D of
the FP32 factor with the fallback must be closer to the oracle's FP64 pivots than without it, the inertia must equal
the oracle's, and GMRES-IR must reach 1e-10 from both factors. SPD Cholesky factors never cross it."""
import numpy as np
import pytest
import scipy.sparse as sp

import ampsolve as amp
import oracle as O
from ampsolve import workloads as W
from fixtures import grid27

pytestmark = pytest.mark.gpu


def test_growth_fallback_on_nearly_singular_ldlt():
    pat, v, n = grid27(14, 14, 14, shift=26.49, symmetric=True)
    A = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(n, n))
    b = np.sin(np.arange(n, dtype=np.float64))
    sym = amp.analyze(pat)
    So = O.Sym(pat)
    ok, Lo, Do, nneg, npert = O.ldlt_factor(So, v)
    assert ok and npert == 0
    res = {}
    for lim in (0.0, 100.0):
        num = amp.NumericFactor(sym, 1, "hip-f32")
        num.set_growth_limit(lim)
        assert num.factor(v)
        redo = num.kernel_ms("diag_fp64_redo")[1]
        _, Dg = num.download()
        errD = np.abs(Dg - Do) / np.abs(Do)
        x, info = num.refine(v, b, "gmres", maxSteps=200, restart=30, tol=1e-10)
        rel = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
        res[lim] = (redo, errD, num.info["nNeg"], rel, info["steps"])
        print(f"growth limit {lim:g}: {redo} blocks redone in FP64, D error vs oracle max {errD.max():.2e} median "
              f"{np.median(errD):.2e}, nNeg {num.info['nNeg']} (oracle {nneg}), GMRES {info['steps']} -> {rel:.2e}",
              flush=True)
    assert res[0.0][0] == 0 and res[100.0][0] > 0
    assert res[100.0][1].max() < res[0.0][1].max()
    for lim in res:
        assert res[lim][2] == nneg and res[lim][3] <= 1e-10


def test_spd_cholesky_never_redoes():
    pat, v, _ = W.heat3d(20)
    num = amp.NumericFactor(amp.analyze(pat), 0, "hip-f32")
    num.set_growth_limit(100.0)
    assert num.factor(v)
    assert num.kernel_ms("diag_fp64_redo")[1] == 0
