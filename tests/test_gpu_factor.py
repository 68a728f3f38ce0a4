"""GPU parity of the device Cholesky / LDL^T factor and solves against the oracle (the C restatement of the reference
algorithm in oracle/). Tolerances: the device factor accumulates with fused multiply-adds on the matrix cores, so it
is not bitwise equal to the reference's unfused arithmetic; the solution must agree with the oracle's to 1e-12
relative (max norm) and reach the same residual level (< 1e-12) on these well-conditioned cases."""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from ampsolve import workloads as W
import ampsolve as amp
from fixtures import chain, grid27, hex_matrix

pytestmark = pytest.mark.gpu


def lower_mask(S):
    mask = np.zeros(S.nnzLStored, dtype=bool)
    for s in range(S.nsuper):
        nc = S.superStart[s + 1] - S.superStart[s]
        m = S.Sp[s + 1] - S.Sp[s]
        blk = np.tril(np.ones((m, nc), dtype=bool))
        mask[S.Lxp[s]:S.Lxp[s + 1]] = blk.T.ravel()
    return mask


def relmax(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def residual(pat, v, x, b):
    A = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))
    return np.linalg.norm(A @ x - b) / np.linalg.norm(b)


# (pattern, values, symmetric values?) -- the grid() fixtures of la-real-front.test.mjs are deliberately
# non-symmetric: the factor reads only the entries analyze() maps through lowerSrc, so no residual against the CSR
CASES = {
    "poisson2d_64": lambda: (*W.poisson2d(64)[:2], True),
    "grid27_11": lambda: (*grid27(11, 11, 11)[:2], False),
    "grid27_23x9x5": lambda: (*grid27(23, 9, 5)[:2], False),
    "grid27_10_zeros": lambda: (*grid27(10, 10, 10, zeroEvery=3)[:2], False),
    "grid27_12_sym": lambda: (*grid27(12, 12, 12, symmetric=True)[:2], True),
    "heat3d_14": lambda: (*W.heat3d(14)[:2], True),
    "heat3d_20": lambda: (*W.heat3d(20)[:2], True),
    "hex_matrix": lambda: (*hex_matrix(), True),
}


@pytest.mark.parametrize("name", list(CASES))
def test_cholesky_matches_oracle(name):
    print(f"case {name}", flush=True)
    pat, v, symmetric = CASES[name]()
    n = pat["n"]
    b = 2 * W.lcg_uniform(7, n) - 1
    sym = amp.analyze(pat)
    So = O.Sym(pat)
    assert np.array_equal(sym.perm, So.perm) and np.array_equal(sym.Si, So.Si)  # identical structure
    ok_o, Lo = O.chol_factor(So, v)
    xo = O.chol_solve(So, Lo, b)
    num = amp.factorNumeric(sym, v)
    assert ok_o and num.ok
    Lg, _ = num.download()
    m = lower_mask(So)
    assert relmax(Lg[m], Lo[m]) < 1e-11
    xg = amp.solveNumeric(num, b)
    assert relmax(xg, xo) < 1e-12
    if symmetric:
        assert residual(pat, v, xg, b) < 1e-12


def test_tridiagonal_closed_form():
    """tridiag(-1, 2, -1) in natural order: L_kk = sqrt((k+2)/(k+1)), L_k+1,k = -sqrt((k+1)/(k+2)) (0-based)."""
    n = 300
    pat, v = chain(n)
    sym = amp.analyze(pat, {"perm": np.arange(n, dtype=np.int32), "relax": False})
    num = amp.factorNumeric(sym, v)
    assert num.ok
    Lx, _ = num.download()
    k = np.arange(n, dtype=np.float64)
    diag = np.empty(n)
    sub = np.empty(n - 1)
    for s in range(sym.nsuper):
        f, nc = sym.superStart[s], sym.superStart[s + 1] - sym.superStart[s]
        m = sym.Sp[s + 1] - sym.Sp[s]
        blk = Lx[sym.Lxp[s]:sym.Lxp[s + 1]].reshape(nc, m).T
        for c in range(nc):
            j = f + c
            diag[j] = blk[c, c]
            if j + 1 < n:
                sub[j] = blk[c + 1, c]
    assert relmax(diag, np.sqrt((k + 2) / (k + 1))) < 1e-14
    assert relmax(sub, -np.sqrt((k[:-1] + 1) / (k[:-1] + 2))) < 1e-14


@pytest.mark.parametrize("dims,opt", [((11, 11, 11), {}), ((23, 9, 5), {}), ((10, 10, 10), {"zeroEvery": 3})])
@pytest.mark.parametrize("pivot_tol", [1e-13, 0.05])
def test_ldlt_indefinite_matches_oracle(dims, opt, pivot_tol):
    """The shifted grids are nearly singular (diagonal 26.5 - 26), so the two factors are compared through the inertia
    (exact), the pivots and a residual of the same level as the oracle's, against the symmetric values."""
    print(f"ldlt {dims} {opt} {pivot_tol}", flush=True)
    pat, v, n = grid27(*dims, shift=26, symmetric=True, **opt)
    b = np.sin(np.arange(n, dtype=np.float64))
    sym = amp.analyze(pat)
    So = O.Sym(pat)
    ok_o, Lo, Do, nneg_o, npert_o = O.ldlt_factor(So, v, pivot_tol)
    f = amp.SupernodalLDLT(sym)
    r = f.factor(v, pivotTol=pivot_tol)
    assert ok_o and r["ok"]
    assert r["nNeg"] == nneg_o > 0
    assert r["nPerturbed"] == npert_o
    if pivot_tol > 1e-3:
        assert npert_o > 0
    Lg, Dg = f.num.download()
    # pivots of a nearly singular matrix factored without pivoting (|D| spans 1e-3 .. 7e3 here): their sensitivity to
    # rounding is measured on the oracle itself, as the spread of D when every input value moves by one ulp; the
    # device factor (another summation order) must stay within 100 times that spread
    xo = O.ldlt_solve(So, Lo, Do, b)
    rng = np.random.default_rng(7)
    spreadD = spreadX = 0.0
    for _ in range(2):
        vp = v * (1 + np.finfo(np.float64).eps * rng.choice([-1.0, 1.0], size=v.shape))
        _, Lp, Dp, _, _ = O.ldlt_factor(So, vp, pivot_tol)
        spreadD = max(spreadD, relmax(Dp, Do))
        spreadX = max(spreadX, relmax(O.ldlt_solve(So, Lp, Dp, b), xo))
    assert relmax(Dg, Do) < max(1e-7, 100 * spreadD), (relmax(Dg, Do), spreadD)
    xg = f.solve(b)
    ro, rg = residual(pat, v, xo, b), residual(pat, v, xg, b)
    assert rg <= 10 * ro + 1e-13, (rg, ro)
    if pivot_tol < 1e-3:
        assert relmax(xg, xo) < max(1e-7, 100 * spreadX), (relmax(xg, xo), spreadX)


def quasi_definite(ncell, c=0.5):
    """[[K, c I], [c I, -K]] with K the SPD 3D heat matrix: symmetric quasi-definite, so LDL^T exists in any order
    without perturbation, its inertia is (n, n) (Sylvester), and its conditioning is that of K (well conditioned)."""
    return quasi_definite_of(*W.heat3d(ncell)[:2], c)


def quasi_definite_of(pat, v, c=0.5):
    """[[K, c I], [c I, -K]] for the SPD matrix K given by (pat, v)."""
    n = pat["n"]
    K = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(n, n))
    I = sp.identity(n, format="csr")
    A = sp.bmat([[K, c * I], [c * I, -K]], format="csr")
    A.sort_indices()
    return W.csr_dict(A), A.data.astype(np.float64), A, n


@pytest.mark.parametrize("ncell", [8, 14])
def test_ldlt_well_conditioned_indefinite_matches_oracle(ncell):
    """LDL^T factor + solves (ldlt.mjs:450-643) held to 1e-12 against the oracle on a well-conditioned indefinite
    matrix: nNeg = n exactly (Sylvester), no perturbed pivot, single and multi right-hand-side solves (q = 1 divides by
    D, q > 1 multiplies by 1/D) within 1e-12 of the oracle's, residual at the oracle's level."""
    pat, v, A, nhalf = quasi_definite(ncell)
    N = pat["n"]
    print(f"quasi-definite ncell={ncell} N={N}", flush=True)
    sym = amp.analyze(pat)
    So = O.Sym(pat)
    ok_o, Lo, Do, nneg_o, npert_o = O.ldlt_factor(So, v)
    assert ok_o and nneg_o == nhalf and npert_o == 0
    f = amp.SupernodalLDLT(sym)
    r = f.factor(v)
    assert r["ok"] and r["nNeg"] == nhalf and r["nPerturbed"] == 0
    if ncell == 14:
        assert sym.maxFront > 96  # the tiled (matrix-core) LDL^T kernels run, not only the whole-front ones
    Lg, Dg = f.num.download()
    m = lower_mask(So)
    assert relmax(Dg, Do) < 1e-12
    assert relmax(Lg[m], Lo[m]) < 1e-11
    b = 2 * W.lcg_uniform(11, N) - 1
    xo = O.ldlt_solve(So, Lo, Do, b)
    xg = f.solve(b)
    assert relmax(xg, xo) < 1e-12
    ro, rg = residual(pat, v, xo, b), residual(pat, v, xg, b)
    assert ro < 1e-13 and rg < 1e-13
    B = np.stack([b, W.lcg_normal(12, N), np.sin(np.arange(N, dtype=np.float64))])
    Xo = O.ldlt_solve_many(So, Lo, Do, B)
    Xg = f.solveMany(B.reshape(-1), 3).reshape(3, N)
    for k in range(3):
        assert relmax(Xg[k], Xo[k]) < 1e-12


@pytest.mark.parametrize("flow_wide", [0, 2**30])
@pytest.mark.parametrize("kind", ["hip", "hip-f32"])
def test_solve_task_widths_match_oracle(kind, flow_wide):
    """Forward solve launches with multi-block tasks everywhere (flowWideMin = 0: a workgroup solves several 64-row
    blocks and hands each to the next inside LDS) and with one block per task: Cholesky and LDL^T (quasi-definite)
    solves of one and three right-hand sides. FP64 factors: within 1e-12 of the oracle's solutions. FP32 factors:
    within 1e-5 of the oracle's FP64 solves with the same (downloaded) FP32 factor, and refined to 1e-12 of the
    oracle's solution."""
    prm = {"smallMax": 96, "smallNarrow": 32, "wideLaunch": 512, "superPanel": 256, "flowWideMin": flow_wide}
    for mode, (pat, v) in ((0, W.heat3d(20)[:2]), (1, quasi_definite(14)[:2])):
        n = pat["n"]
        print(f"kind {kind} flowWideMin {flow_wide} mode {mode} n {n}", flush=True)
        sym = amp.analyze(pat)
        sym.set_params(prm)
        So = O.Sym(pat)
        if mode == 0:
            _, Lo = O.chol_factor(So, v)
            ref_solve = lambda L, D, b: O.chol_solve(So, L, b)  # noqa: E731
            Do = None
        else:
            _, Lo, Do, _, _ = O.ldlt_factor(So, v)
            ref_solve = lambda L, D, b: O.ldlt_solve(So, L, D, b)  # noqa: E731
        num = amp.NumericFactor(sym, mode, kind)
        assert num.factor(v)
        B = np.stack([2 * W.lcg_uniform(31 + k, n) - 1 for k in range(3)])
        exact = [ref_solve(Lo, Do, B[k]) for k in range(3)]
        if kind == "hip":
            ref, tol = exact, 1e-12
        else:
            Lg, Dg = num.download()
            ref, tol = [ref_solve(Lg, Dg if mode else None, B[k]) for k in range(3)], 1e-5
        assert relmax(num.solve(B[0]), ref[0]) < tol
        X = num.solve(B.ravel(), nrhs=3).reshape(3, n)
        for k in range(3):
            assert relmax(X[k], ref[k]) < tol
        if kind == "hip-f32":
            x, info = num.refine(v, B[0], maxSteps=8, tol=1e-15)
            assert relmax(x, exact[0]) < 1e-12


@pytest.mark.parametrize("pack", [0, 2**30])
@pytest.mark.parametrize("kind", ["hip", "hip-f32"])
def test_packed_small_fronts_match_oracle(kind, pack):
    """Whole fronts of at most 8 / 16 rows factored and solved 8 / 4 to a wave (packMin = 0: packed on every level)
    and one per wave (packMin = 2^30): Cholesky of a batch of 2D heat systems and LDL^T of a quasi-definite system
    built on it, one, three and eight right-hand sides. FP64 factors: within 1e-12 of the oracle's solutions. FP32
    factors: within 1e-5 of the oracle's FP64 solves with the same (downloaded) FP32 factor. FP64 LDL^T of the
    quasi-definite system with pivotTol 0.05 (1,583 perturbed pivots): nNeg and nPerturbed equal the oracle's."""
    prm = {"smallMax": 96, "smallNarrow": 32, "wideLaunch": 512, "superPanel": 256, "flowWideMin": 160,
           "packMin": pack}
    K2 = W.batch_heat2d(nsys=8, lo=30, hi=90, nrhs=1)[:2]
    for mode, (pat, v) in ((0, K2), (1, quasi_definite_of(*K2)[:2])):
        n = pat["n"]
        print(f"packed kind {kind} packMin {pack} mode {mode} n {n}", flush=True)
        sym = amp.analyze(pat)
        sym.set_params(prm)
        So = O.Sym(pat)
        if mode == 0:
            _, Lo = O.chol_factor(So, v)
            ref_solve = lambda L, D, b: O.chol_solve(So, L, b)  # noqa: E731
            Do = None
        else:
            _, Lo, Do, _, _ = O.ldlt_factor(So, v)
            ref_solve = lambda L, D, b: O.ldlt_solve(So, L, D, b)  # noqa: E731
        num = amp.NumericFactor(sym, mode, kind)
        assert num.factor(v)
        B = np.stack([2 * W.lcg_uniform(41 + k, n) - 1 for k in range(8)])
        if kind == "hip":
            Lr, Dr, tol = Lo, Do, 1e-12
        else:
            Lr, Dr = num.download()
            tol = 1e-5
        ref = [ref_solve(Lr, Dr if mode else None, B[k]) for k in range(8)]
        assert relmax(num.solve(B[0]), ref[0]) < tol
        for q in (3, 8):
            X = num.solve(B[:q].ravel(), nrhs=q).reshape(q, n)
            for k in range(q):
                assert relmax(X[k], ref[k]) < tol
    if kind == "hip":
        pat, v = quasi_definite_of(*K2)[:2]
        print(f"packed ldlt perturbed packMin {pack}", flush=True)
        So = O.Sym(pat)
        ok_o, _, _, nneg_o, npert_o = O.ldlt_factor(So, v, 0.05)
        sym = amp.analyze(pat)
        sym.set_params(prm)
        num = amp.NumericFactor(sym, 1, kind)
        assert ok_o and num.factor(v, pivotTol=0.05)
        assert num.info["nNeg"] == nneg_o > 0 and num.info["nPerturbed"] == npert_o > 0


def test_cholesky_reports_indefinite_and_solve_refuses():
    pat, v, n = grid27(11, 11, 11, shift=26)
    num = amp.factorNumeric(amp.analyze(pat), v)
    assert num.ok is False
    with pytest.raises(RuntimeError, match="no successful numeric factorization"):
        amp.solveNumeric(num, np.zeros(n))
    ch = amp.SparseCholesky(pat)
    assert ch.factor(v) is False
    with pytest.raises(RuntimeError, match="without a successful factorization"):
        ch.solve(np.zeros(n))


def test_failure_in_root_front():
    """Non-positive Cholesky pivot and non-finite LDL^T pivot in the last-eliminated column (root front)."""
    pat, v, n = grid27(11, 11, 11)
    sym = amp.analyze(pat)
    last = sym.perm[n - 1]

    def with_diag(x):
        w = v.copy()
        for p in range(pat["rowptr"][last], pat["rowptr"][last + 1]):
            if pat["colidx"][p] == last:
                w[p] = x
        return w

    assert amp.factorNumeric(sym, with_diag(-1e6)).ok is False
    So = O.Sym(pat)
    ok_o, _, _, nneg_o, npert_o = O.ldlt_factor(So, with_diag(np.nan))
    r = amp.SupernodalLDLT(sym).factor(with_diag(np.nan))
    assert r["ok"] is False and ok_o is False
    assert r["nNeg"] == nneg_o and r["nPerturbed"] == npert_o


@pytest.mark.parametrize("q", [1, 2, 3, 4, 11])
def test_multi_rhs(q):
    pat, v, n = grid27(12, 12, 12, symmetric=True)
    So = O.Sym(pat)
    _, Lo = O.chol_factor(So, v)
    B = np.stack([2 * W.lcg_uniform(100 + k, n) - 1 for k in range(q)])
    B[0, ::5] = 0.0
    ch = amp.SparseCholesky(pat)
    assert ch.factor(v)
    X = ch.solveMany(list(B))
    for k in range(q):
        assert relmax(X[k], O.chol_solve(So, Lo, B[k])) < 1e-12
    f = amp.SupernodalLDLT(ch.sym)
    assert f.factor(v)["ok"]
    Y = f.solveMany(B.reshape(-1), q).reshape(q, n)
    for k in range(q):
        assert relmax(Y[k], X[k]) < 1e-11


def test_pattern_and_count_checks():
    pat, v, n = grid27(6, 6, 6)
    sym = amp.analyze(pat)
    other = dict(pat, colidx=pat["colidx"].copy())
    last = other["rowptr"][1] - 1
    other["colidx"][last] = (other["colidx"][last] + 1) % n
    with pytest.raises(ValueError, match="different sparsity pattern"):
        amp.SparseCholesky(other, sym=sym)
    with pytest.raises(ValueError, match="different sparsity pattern"):
        amp.factorNumeric(sym, v, None, other)
    with pytest.raises(ValueError, match="values for a pattern"):
        amp.factorNumeric(sym, v[1:])


def test_empty_pattern():
    empty = {"n": 0, "rowptr": np.zeros(1, dtype=np.int32), "colidx": np.zeros(0, dtype=np.int32), "nnz": 0}
    ch = amp.SparseCholesky(empty)
    assert ch.factor(np.zeros(0))
    assert len(ch.solve(np.zeros(0))) == 0


def test_device_pointers_and_determinism(device):
    pat, v, b = W.heat3d(10)
    sym = amp.analyze(pat)
    num = amp.NumericFactor(sym, 0, "hip")
    dv = device.DeviceArray.from_numpy(v)
    db = device.DeviceArray.from_numpy(b)
    assert num.factor(dv)
    L1, _ = num.download()
    x1 = num.solve(db).numpy()
    assert num.factor(dv)
    L2, _ = num.download()
    x2 = num.solve(db).numpy()
    assert np.array_equal(L1, L2) and np.array_equal(x1, x2)  # run-to-run bitwise (no atomics in the arithmetic)
    So = O.Sym(pat)
    _, Lo = O.chol_factor(So, v)
    assert relmax(x1, O.chol_solve(So, Lo, b)) < 1e-12


def test_autotune_keeps_parity(device):
    """The tuner rebuilds the plans with other shape thresholds: the factor must stay the oracle's to round-off, the
    chosen thresholds must be no slower than the defaults on the measured runs, and plans cannot change under a
    live numeric factor."""
    pat, v, b = W.heat3d(16)
    sym = amp.analyze(pat)
    dv = device.DeviceArray.from_numpy(v)
    t = sym.autotune(dv)
    assert t["candidates"] >= 2 and t["bestMs"] <= t["defaultMs"]
    assert 1 <= t["smallNarrow"] <= t["smallMax"] <= 96 and t["superPanel"] % 64 == 0
    So = O.Sym(pat)
    _, Lo = O.chol_factor(So, v)
    xo = O.chol_solve(So, Lo, b)
    num = amp.factorNumeric(sym, v)
    assert num.ok and relmax(amp.solveNumeric(num, b), xo) < 1e-12
    with pytest.raises(RuntimeError):
        sym.set_params({"smallMax": 64, "smallNarrow": 32, "wideLaunch": 256, "superPanel": 128})
    del num
    sym.set_params({"smallMax": 64, "smallNarrow": 32, "wideLaunch": 256, "superPanel": 128})
    num = amp.factorNumeric(sym, v)
    assert num.ok and relmax(amp.solveNumeric(num, b), xo) < 1e-12


def test_inputs_ordered_after_the_callers_stream(device):
    """Inputs whose producing copies are still queued on the null stream when the call is made (behind ~8 GB of
    unrelated device copies) must be read only after they are written: same bits as with completed inputs."""
    pat, v, b = W.heat3d(12)
    sym = amp.analyze(pat)
    num = amp.NumericFactor(sym, 0, "hip")
    dv = device.DeviceArray.from_numpy(v)
    db = device.DeviceArray.from_numpy(b)
    assert num.factor(dv)
    L1, _ = num.download()
    x1 = num.solve(db).numpy()
    # the same values and right-hand side staged elsewhere; the inputs are scrambled, then restored by copies queued
    # on the null stream behind a chain of large copies
    sv, sb = dv.clone(), db.clone()
    dv.upload(np.full(len(v), np.nan))
    db.upload(np.full(len(b), np.nan))
    big = device.DeviceArray.zeros(1 << 27)  # 1 GiB
    big2 = device.DeviceArray(1 << 27)
    device.synchronize()
    for _ in range(4):
        big2.copy_async_(big)
        big.copy_async_(big2)
    dv.copy_async_(sv)
    db.copy_async_(sb)
    assert num.factor(dv)
    x2 = num.solve(db).numpy()
    L2, _ = num.download()
    assert np.array_equal(L1, L2) and np.array_equal(x1, x2)


@pytest.mark.parametrize("kind", ["hip", "hip-f32", "hip-bf16x3"])
def test_fused_panel_steps_match_the_three_launch_form(kind):
    """A panel step of the tiled fronts as one fused launch (every workgroup eliminates the diagonal block itself,
    forms its L rows and updates its rows of the super-panel's later columns as (L W) A_c^T on the matrix cores)
    against the same step as three launches (diagonal block, L = A W^T, C -= L D L_c^T): the same factor up to the
    rounding of the reordered update product. FP32 factors: within 1e-5 (relative, max norm), equal pivot records
    (inertia, perturbations). FP64 factors run the three-launch form of the plan whatever fusedPanel says (the fused
    kernel's FP64 instance cannot share a CU with the FP64 bulk updates), so theirs are bitwise equal.
    Cholesky of the 3D heat system (several super-panels per front) and LDL^T of a quasi-definite system. For FP32
    also the growth-checked instance with a limit that redoes blocks in FP64, on the nearly singular shifted 27-point
    grid, where rounding differences are amplified and the two factors are not close: there the fused factor must
    redo blocks, have the oracle's inertia, and bring GMRES-IR to 1e-10."""
    heat = W.heat3d(20)[:2]
    qd = quasi_definite(14)[:2]
    near = grid27(14, 14, 14, shift=26.49, symmetric=True)[:2]
    cases = [(0, heat, None), (1, qd, None)]
    if kind == "hip-f32":
        cases.append((1, near, 100.0))
    tol = 0.0 if kind == "hip" else 1e-5
    for mode, (pat, v), limit in cases:
        n = pat["n"]
        b = 2 * W.lcg_uniform(5, n) - 1
        out = {}
        for fused in (1, 0):
            print(f"fused panel steps: kind {kind} mode {mode} n {n} growth limit {limit} fused {fused}", flush=True)
            sym = amp.analyze(pat)
            sym.set_params({"smallMax": 96, "smallNarrow": 32, "wideLaunch": 512, "superPanel": 256,
                            "flowWideMin": 160, "fusedPanel": fused})
            num = amp.NumericFactor(sym, mode, kind)
            if limit is not None:
                num.set_growth_limit(limit)
            assert num.factor(v)
            L, D = num.download()
            out[fused] = (L, D, num.info["nNeg"], num.info["nPerturbed"], num.kernel_ms("diag_fp64_redo")[1],
                          num.solve(b), lower_mask(sym))
            if limit is not None:
                _, info = num.refine(v, b, "gmres", maxSteps=200, restart=30, tol=1e-10)
                assert info["converged"] and info["relResidual"] <= 1e-10
            del num
        (L1, D1, nn1, np1, r1, x1, m), (L0, D0, nn0, np0, r0, x0, _) = out[1], out[0]
        print(f"  L {relmax(L1[m], L0[m]):.2e}  x {relmax(x1, x0):.2e}  redone {r1} / {r0}", flush=True)
        if limit is not None:
            _, _, _, nneg, _ = O.ldlt_factor(O.Sym(pat), v)
            assert r1 > 0 and nn1 == nneg
            continue
        assert relmax(L1[m], L0[m]) <= tol
        if mode == 1:
            assert relmax(D1, D0) <= tol and nn1 == nn0 and np1 == np0
        if kind == "hip":
            assert np.array_equal(x1, x0)
