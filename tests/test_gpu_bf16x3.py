"""The FP32 factor with bf16 bulk trailing updates ('hip-bf16x3': side-stream product launches as three bf16 products
per FP32 product, k_gemm_bf16x3) against the oracle. The bulk launches are forced onto small systems by lowering the
wide-launch threshold. Bars: the factor within 1e-4 (max-norm relative) of the oracle's (a relative error of about
2^-16 per product), the bulk kernel's result equal to what the FP32 kernel gives to that level, and FP64 refinement
of its solves to |b - A x| / |b| <= 1e-10 within 1e-9 of the oracle's solution (Cholesky: IR; LDL^T on a
quasi-definite matrix: GMRES-IR)."""
import numpy as np
import pytest
import scipy.sparse as sp

import ampsolve as amp
import oracle as O
from ampsolve import workloads as W

pytestmark = pytest.mark.gpu

WIDE = {"smallMax": 96, "smallNarrow": 32, "wideLaunch": 4, "superPanel": 256}


def relmax(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def lower_mask(S):
    mask = np.zeros(S.nnzLStored, dtype=bool)
    for s in range(S.nsuper):
        nc = S.superStart[s + 1] - S.superStart[s]
        m = S.Sp[s + 1] - S.Sp[s]
        mask[S.Lxp[s]:S.Lxp[s + 1]] = np.tril(np.ones((m, nc), dtype=bool)).T.ravel()
    return mask


@pytest.mark.parametrize("ncell", [24, 32])
def test_bf16x3_cholesky_factor_and_refined_solution(ncell):
    pat, v, b = W.heat3d(ncell)
    A = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))
    sym = amp.analyze(pat)
    sym.set_params(WIDE)
    So = O.Sym(pat)
    ok, Lo = O.chol_factor(So, v)
    xo = O.chol_solve(So, Lo, b)
    m = lower_mask(So)
    f32 = amp.NumericFactor(sym, 0, "hip-f32")
    bf = amp.NumericFactor(sym, 0, "hip-bf16x3")
    bf.set_profiling(True)
    assert f32.factor(v) and bf.factor(v)
    ms, launches = bf.kernel_ms("gemm_tile")
    L32, _ = f32.download()
    Lbf, _ = bf.download()
    e32, ebf = relmax(L32[m], Lo[m]), relmax(Lbf[m], Lo[m])
    print(f"bf16x3 heat3d({ncell}) n={pat['n']} maxFront={sym.maxFront}: factor error vs oracle FP32 {e32:.2e}, "
          f"bf16x3 {ebf:.2e}", flush=True)
    assert ebf < 1e-4 and ebf > e32  # the bulk kernel ran (its error is above FP32's) and stayed at its level
    x, info = bf.refine(v, b, "ir", maxSteps=20, tol=1e-10)
    rel = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
    err = relmax(x, xo)
    print(f"bf16x3 heat3d({ncell}): {info['steps']} IR steps, relres {rel:.2e}, error vs oracle {err:.2e}", flush=True)
    assert info["converged"] and rel <= 1e-10 and err < 1e-9


def test_bf16x3_ldlt_gmres_ir():
    n_ = W.heat3d(16)[0]["n"]
    pat, v, _ = W.heat3d(16)
    K = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(n_, n_))
    I = sp.identity(n_, format="csr")
    A = sp.bmat([[K, 0.5 * I], [0.5 * I, -K]], format="csr")
    A.sort_indices()
    pq, vq = W.csr_dict(A), A.data.astype(np.float64)
    N = pq["n"]
    b = 2 * W.lcg_uniform(9, N) - 1
    sym = amp.analyze(pq)
    sym.set_params(WIDE)
    So = O.Sym(pq)
    ok, Lo, Do, nneg, _ = O.ldlt_factor(So, vq)
    xo = O.ldlt_solve(So, Lo, Do, b)
    f = amp.SupernodalLDLT(sym, factor="hip-bf16x3")
    r = f.factor(vq)
    assert r["ok"] and r["nNeg"] == nneg == n_
    x, info = f.num.refine(vq, b, "gmres", maxSteps=100, restart=30, tol=1e-10)
    rel = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
    print(f"bf16x3 LDL^T N={N}: {info['steps']} GMRES iterations, relres {rel:.2e}, error vs oracle "
          f"{relmax(x, xo):.2e}", flush=True)
    assert info["converged"] and rel <= 1e-10 and relmax(x, xo) < 1e-9
