"""GPU refinement (FP64 residual / update in HBM): solveRefined semantics (ldlt.mjs:253-275) and GMRES-IR
(gmres-real.mjs:31-104) around FP64 and FP32 device factors, judged against the oracle's FP64 direct solution and the
exact solutions of constructed systems."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import ampsolve as amp
from ampsolve import workloads as W
from fixtures import grid27, ldlt_system

pytestmark = pytest.mark.gpu


def relres(pat, v, x, b):
    A = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))
    return np.linalg.norm(A @ x - b) / np.linalg.norm(b)


def test_ldlt_perturbed_pivots_refined_to_exact_solution():
    """la-ldlt.test.mjs:59-68 on the device: 1x1 pivots hit exact zeros (identity order), the perturbed factor is
    refined 30 steps and must recover the exact x to 1e-10."""
    pat, v, x, b = ldlt_system(40)
    sym = amp.analyze(pat, {"perm": np.arange(pat["n"], dtype=np.int32)})
    f = amp.SupernodalLDLT(sym)
    r = f.factor(v)
    assert r["ok"] and r["nPerturbed"] > 0
    out = f.solveRefined(pat, v, b, maxSteps=30)
    assert np.abs(out["x"] - x).max() / np.abs(x).max() < 1e-10
    plain = f.solve(b)
    assert np.abs(plain - x).max() / np.abs(x).max() > 1e-6  # the refinement is exercised
    So = O.Sym(pat, perm=np.arange(pat["n"]))
    ok, Lo, Do, nneg, npert = O.ldlt_factor(So, v)
    assert npert == r["nPerturbed"] and nneg == r["nNeg"]


@pytest.mark.parametrize("kind", ["hip", "hip-f32"])
def test_ir_reaches_fp64_accuracy(kind):
    pat, v, b = W.heat3d(12)
    sym = amp.analyze(pat)
    num = amp.NumericFactor(sym, 0, kind)
    assert num.factor(v)
    x, info = num.refine(v, b, "ir", maxSteps=20, tol=1e-12)
    assert info["converged"], info
    assert relres(pat, v, x, b) < 1e-12
    So = O.Sym(pat)
    _, Lo = O.chol_factor(So, v)
    xo = O.chol_solve(So, Lo, b)
    assert np.abs(x - xo).max() / np.abs(xo).max() < 1e-10
    if kind == "hip-f32":
        x0 = num.solve(b)
        assert relres(pat, v, x0, b) > 1e-9  # the FP32 factor alone is visibly less accurate


def test_gmres_ir_fp32_factor():
    pat, v, b = W.heat3d(16)
    sym = amp.analyze(pat)
    num = amp.NumericFactor(sym, 0, "hip-f32")
    assert num.factor(v)
    x, info = num.refine(v, b, "gmres", maxSteps=100, restart=30, tol=1e-10)
    assert info["converged"], info
    assert relres(pat, v, x, b) < 1e-10


def test_gmres_ir_indefinite_ldlt_fp32():
    pat, v, n = grid27(10, 10, 10, shift=26, symmetric=True)
    b = np.cos(np.arange(n, dtype=np.float64))
    f = amp.SupernodalLDLT(amp.analyze(pat), factor="hip-f32")
    r = f.factor(v)
    assert r["ok"] and r["nNeg"] > 0
    x, info = f.num.refine(v, b, "gmres", maxSteps=200, restart=40, tol=1e-10)
    assert info["converged"], info
    assert relres(pat, v, x, b) < 1e-10
