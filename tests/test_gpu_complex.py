"""Complex-symmetric LDL^T on the device (ampsolve.ComplexSupernodalLDLT: the real 2 x 2-block system on the device
LDL^T) against the oracle's restatement of ComplexSupernodalLDLT (ldlt.mjs:648-866). Bars: the solution within 1e-11
(max-norm relative) of the oracle's and a complex residual at the oracle's level on the la-complex-solve grids (real
part definite and indefinite, exact zeros in the off-diagonal, tiled fronts); la-ldlt's complex system with exact zero
pivots: perturbed pivots, a visibly wrong plain solve and 30 refinement steps to the exact solution within 1e-10; the
FP32 factor refined to 1e-10; a system whose imaginary parts dominate the diagonal (multiplied by -i first)."""
import numpy as np
import pytest
import scipy.sparse as sp

import ampsolve as amp
import oracle as O
from fixtures import complex_grid27, complex_ldlt_system

pytestmark = pytest.mark.gpu


def cres(pat, re, im, xr, xi, br, bi):
    A = sp.csr_matrix((re + 1j * im, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))
    b = br + 1j * bi
    return np.linalg.norm(A @ (xr + 1j * xi) - b) / np.linalg.norm(b)


def crel(xr, xi, yr, yi):
    return np.abs((xr - yr) + 1j * (xi - yi)).max() / np.abs(yr + 1j * yi).max()


@pytest.mark.parametrize("N,shift,zeros", [(8, 0.0, 0), (8, 13.0, 3), (14, 0.0, 0), (14, 13.0, 0)])
def test_complex_ldlt_matches_oracle(N, shift, zeros):
    pat, re, im = complex_grid27(N, shift=shift, zeroEvery=zeros)
    n = pat["n"]
    print(f"complex grid N={N} shift={shift} n={n}", flush=True)
    br, bi = np.sin(0.3 * np.arange(n)), 0.5 * np.cos(0.7 * np.arange(n))
    S = O.Sym(pat)
    ok, Lr, Li, Dr, Di, npert = O.cldlt_factor(S, re, im)
    xo_r, xo_i = O.cldlt_solve(S, Lr, Li, Dr, Di, br, bi)
    sym = amp.analyze(pat)
    f = amp.ComplexSupernodalLDLT(sym)
    r = f.factor(re, im)
    assert ok and r["ok"] and npert == 0 and r["nPerturbed"] == 0
    if N == 14:
        assert f.sym2.maxFront > 96  # tiled (matrix-core) fronts
    x = f.solve(br, bi)
    err = crel(x["re"], x["im"], xo_r, xo_i)
    ro, rg = cres(pat, re, im, xo_r, xo_i, br, bi), cres(pat, re, im, x["re"], x["im"], br, bi)
    print(f"complex: error vs oracle {err:.2e}, residual {rg:.2e} (oracle {ro:.2e})", flush=True)
    assert err < 1e-11 and rg <= 10 * ro + 1e-15


def test_complex_perturbed_pivots_refined_to_exact_solution():
    pat, re, im, xr, xi, br, bi = complex_ldlt_system(40)
    sym = amp.analyze(pat, {"perm": np.arange(pat["n"], dtype=np.int32)})
    f = amp.ComplexSupernodalLDLT(sym)
    r = f.factor(re, im)
    assert r["ok"] and r["nPerturbed"] > 0
    plain = f.solve(br, bi)
    assert np.abs(plain["re"] - xr).max() / np.abs(xr).max() > 1e-6
    ref = f.solveRefined(pat, re, im, br, bi, maxSteps=30)
    assert np.abs(ref["re"] - xr).max() / np.abs(xr).max() < 1e-10
    assert np.abs(ref["im"] - xi).max() / np.abs(xi).max() < 1e-10


def test_complex_fp32_factor_refined():
    pat, re, im = complex_grid27(12)
    n = pat["n"]
    br, bi = np.sin(0.3 * np.arange(n)), 0.5 * np.cos(0.7 * np.arange(n))
    f = amp.ComplexSupernodalLDLT(amp.analyze(pat), factor="hip-f32")
    assert f.factor(re, im)["ok"]
    ref = f.solveRefined(pat, re, im, br, bi, maxSteps=10, tol=1e-10)
    assert ref["relResidual"] <= 1e-10
    S = O.Sym(pat)
    _, Lr, Li, Dr, Di, _ = O.cldlt_factor(S, re, im)
    xo_r, xo_i = O.cldlt_solve(S, Lr, Li, Dr, Di, br, bi)
    assert crel(ref["re"], ref["im"], xo_r, xo_i) < 1e-9


def test_complex_dominant_imaginary_part():
    """Diagonal imaginary parts larger than the real ones (a strongly lossy system): the real system is formed for
    -i A, whose real part is the larger one, and the solution still matches the oracle."""
    pat, re, im = complex_grid27(8)
    re2, im2 = 0.05 * re, im + 3.0 * np.where(pat["colidx"] == np.repeat(np.arange(pat["n"]), np.diff(pat["rowptr"])),
                                              re, 0.0)
    n = pat["n"]
    br, bi = np.sin(0.3 * np.arange(n)), 0.5 * np.cos(0.7 * np.arange(n))
    f = amp.ComplexSupernodalLDLT(amp.analyze(pat))
    assert f.factor(re2, im2)["ok"] and f.rotated
    x = f.solve(br, bi)
    S = O.Sym(pat)
    _, Lr, Li, Dr, Di, _ = O.cldlt_factor(S, re2, im2)
    xo_r, xo_i = O.cldlt_solve(S, Lr, Li, Dr, Di, br, bi)
    assert crel(x["re"], x["im"], xo_r, xo_i) < 1e-11
