"""Pins of the oracle (oracle/, the C restatement of the reference algorithm) against facts it could get wrong:
closed forms, dense LAPACK factors, the up-looking structure, brute-force symbolic elimination, Sylvester inertia, the
reference's own test properties (la-sparse, la-ldlt, core-solvers tests) and an independent fingerprint."""
import numpy as np
import pytest
import scipy.linalg as sl
import scipy.sparse as sp

import oracle as O
from ampsolve import workloads as W
from fixtures import chain, grid27, hex_matrix, ldlt_system


def dense(pat, v):
    return sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"])).toarray()


def unpack_L(S, Lx):
    """Dense lower factor (new ordering) from the supernodal blocks."""
    L = np.zeros((S.n, S.n))
    for s in range(S.nsuper):
        f, nc = S.superStart[s], S.superStart[s + 1] - S.superStart[s]
        rows = S.Si[S.Sp[s]:S.Sp[s + 1]]
        m = len(rows)
        blk = Lx[S.Lxp[s]:S.Lxp[s + 1]].reshape(nc, m).T
        for c in range(nc):
            L[rows[c:], f + c] = blk[c:, c]
    return L


def test_tridiagonal_closed_form():
    n = 200
    pat, v = chain(n)
    S = O.Sym(pat, perm=np.arange(n), relax=False)
    ok, Lx = O.chol_factor(S, v)
    L = unpack_L(S, Lx)
    k = np.arange(n, dtype=float)
    assert ok
    assert np.allclose(np.diag(L), np.sqrt((k + 2) / (k + 1)), rtol=0, atol=1e-15)
    assert np.allclose(np.diag(L, -1), -np.sqrt((k[:-1] + 1) / (k[:-1] + 2)), rtol=0, atol=1e-15)


@pytest.mark.parametrize("case", ["poisson8", "grid27_6", "hex"])
def test_cholesky_equals_lapack(case):
    pat, v = {"poisson8": lambda: W.poisson2d(8)[:2], "grid27_6": lambda: grid27(6, 6, 5, symmetric=True)[:2],
              "hex": lambda: hex_matrix(4, 3, 3)}[case]()
    S = O.Sym(pat)
    ok, Lx = O.chol_factor(S, v)
    assert ok
    A = dense(pat, v)
    PAP = A[np.ix_(S.perm, S.perm)]
    Lref = np.linalg.cholesky(PAP)
    L = unpack_L(S, Lx)
    assert np.abs(L - Lref).max() / np.abs(Lref).max() < 1e-13


def brute_force_colcounts(A_pattern, perm):
    """Symbolic Cholesky by boolean elimination of the permuted dense pattern."""
    B = A_pattern[np.ix_(perm, perm)].astype(bool)
    n = len(perm)
    cc = np.zeros(n, dtype=int)
    for k in range(n):
        rows = np.nonzero(B[k:, k])[0] + k
        cc[k] = len(rows)
        below = rows[rows > k]
        B[np.ix_(below, below)] = True
    return cc


@pytest.mark.parametrize("seed", range(6))
def test_colcount_brute_force_tiny(seed):
    rng = np.random.default_rng(seed)
    n = 12
    M = rng.random((n, n)) < 0.2
    M = M | M.T | np.eye(n, dtype=bool)
    A = sp.csr_matrix(M.astype(float))
    A.sort_indices()
    pat = W.csr_dict(A)
    for relax in (False, True):
        S = O.Sym(pat, relax=relax)
        assert np.array_equal(S.colCount, brute_force_colcounts(M, S.perm))
        assert sorted(S.perm) == list(range(n))


def test_structure_properties_of_the_reference_tests():
    """la-sparse.test.mjs:65-103: colCount = up-looking structure, own rows first and sorted, nnzLStored identity."""
    pat, v = hex_matrix(5, 4, 3)
    for relax in (False, True):
        S = O.Sym(pat, relax=relax)
        U = O.UpLooking(pat, S.perm)
        assert S.nnzL == U.nnzL
        assert np.array_equal(S.colCount, np.diff(U.Lp))
        assert sorted(S.perm) == list(range(S.n))
        for s in range(S.nsuper):
            f, nc = S.superStart[s], S.superStart[s + 1] - S.superStart[s]
            rows = S.Si[S.Sp[s]:S.Sp[s + 1]]
            assert np.array_equal(rows[:nc], np.arange(f, f + nc))
            assert np.all(np.diff(rows) > 0)
        if not relax:
            nc = np.diff(S.superStart)
            assert S.nnzLStored - int(np.sum(nc * (nc - 1) // 2)) == S.nnzL


def test_supernodal_matches_uplooking_and_scaling():
    """la-sparse.test.mjs:105-127."""
    pat, v = hex_matrix()
    n = pat["n"]
    b = np.sin(0.7 * np.arange(n)) + 0.3
    S = O.Sym(pat)
    ok, Lx = O.chol_factor(S, v)
    assert ok
    xs = O.chol_solve(S, Lx, b)
    U = O.UpLooking(pat, S.perm)
    assert U.factor(v)
    xu = U.solve(b)
    assert np.abs(xs - xu).max() / np.abs(xu).max() < 1e-12
    A = dense(pat, v)
    assert np.linalg.norm(A @ xs - b) / np.linalg.norm(b) < 1e-12
    ok2, Lx2 = O.chol_factor(S, 2 * v)
    x2 = O.chol_solve(S, Lx2, b)
    assert np.abs(2 * x2 - xs).max() <= 1e-12 * np.abs(xu).max()


def test_indefinite_cholesky_fails():
    pat, v = hex_matrix(4, 4, 3)
    w = v.copy()
    A = sp.csr_matrix((w, pat["colidx"], pat["rowptr"]))
    for i in range(pat["n"]):
        for p in range(pat["rowptr"][i], pat["rowptr"][i + 1]):
            if pat["colidx"][p] == i:
                w[p] -= 1e3
    S = O.Sym(pat)
    ok, _ = O.chol_factor(S, w)
    assert ok is False
    ok, _ = O.chol_factor(S, -v)
    assert ok is False
    del A


def test_sylvester_inertia_of_shifted_chain():
    """core-solvers.test.mjs:86-100: sigma between the 10th and 11th eigenvalues gives nNeg = 10."""
    n = 300
    lam = lambda k: 2 - 2 * np.cos(k * np.pi / (n + 1))  # noqa: E731
    sigma = 0.5 * (lam(10) + lam(11))
    pat, v = chain(n, 2 - sigma, -1)
    S = O.Sym(pat)
    ok, Lx, D, nneg, npert = O.ldlt_factor(S, v)
    assert ok and nneg == 10 and npert == 0
    b = np.cos(np.arange(n))
    x = O.ldlt_solve(S, Lx, D, b)
    assert np.linalg.norm(dense(pat, v) @ x - b) / np.linalg.norm(b) < 1e-10


def test_ldlt_perturbed_pivots_and_refinement():
    """la-ldlt.test.mjs:59-68: exact zeros on the diagonal (identity order) force perturbed pivots; 30 refinement
    steps recover the exact solution to 1e-10."""
    pat, v, x, b = ldlt_system(40)
    S = O.Sym(pat, perm=np.arange(pat["n"]))
    ok, Lx, D, nneg, npert = O.ldlt_factor(S, v)
    assert ok and npert > 0
    xr, rel, steps = O.ldlt_solve_refined(S, Lx, D, pat, v, b, max_steps=30)
    assert np.abs(xr - x).max() / np.abs(x).max() < 1e-10


def test_ldlt_equals_cholesky_on_spd():
    pat, v, n = grid27(7, 6, 5, symmetric=True)
    S = O.Sym(pat)
    b = np.cos(np.arange(n))
    ok, Lx = O.chol_factor(S, v)
    okl, Lxl, D, nneg, npert = O.ldlt_factor(S, v)
    assert ok and okl and nneg == 0 and npert == 0
    x1 = O.chol_solve(S, Lx, b)
    x2 = O.ldlt_solve(S, Lxl, D, b)
    X = O.ldlt_solve_many(S, Lxl, D, np.stack([b, 2 * b]))
    assert np.abs(x1 - x2).max() / np.abs(x1).max() < 1e-12
    assert np.abs(X[0] - x2).max() / np.abs(x2).max() < 1e-12
    assert np.abs(X[1] - 2 * x2).max() / np.abs(x2).max() < 1e-12


def test_gmres_nonsymmetric_chain():
    """core-solvers.test.mjs:102-115."""
    n = 500
    pat, _ = chain(n)
    r = np.repeat(np.arange(n), np.diff(pat["rowptr"]))
    c = pat["colidx"]
    val = np.where(c == r, 2.2, np.where(c < r, -1.3, -0.7))
    A = sp.csr_matrix((val, c, pat["rowptr"]), shape=(n, n))
    b = np.sin(0.1 * np.arange(n)) + 1
    x, it, relres, conv = O.gmres(n, lambda v: A @ v, lambda v: v / 2.2, b, tol=1e-12, restart=40, max_iter=2000)
    assert conv
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-11


def fnv1a_reference(pat):
    h = 0x811C9DC5
    words = list(pat["rowptr"][: pat["n"] + 1]) + list(pat["colidx"][: pat["rowptr"][pat["n"]]])
    for wv in words:
        u = int(wv) & 0xFFFFFFFF
        for sh in (0, 8, 16, 24):
            h ^= (u >> sh) & 0xFF
            h = (h * 0x01000193) & 0xFFFFFFFF
    return f"{pat['n']}:{pat['rowptr'][pat['n']]}:{h:x}"


def test_fingerprint_independent():
    for pat in (W.poisson2d(8)[0], grid27(5, 4, 3)[0]):
        assert O.fingerprint(pat) == fnv1a_reference(pat)
