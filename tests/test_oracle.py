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


# ---- symbolic pins: published fill figures and the relaxed-amalgamation thresholds --------------------------------

def tet10_box_pattern(nx, restrict_interior):
    """Node pattern of the tet10 unit-cube mesh of mesh.ts:123-161 (nx^3 hexes, 6 Kuhn tets each, midside nodes on
    the (2 nx + 1)^3 fine grid, nid(i, j, k) = (k (NY + 1) + j) (NX + 1) + i): every pair of nodes of a tet is coupled,
    diagonal included (cellPattern). With restrict_interior the pattern is restricted to the nodes off all six faces,
    keeping the node order (restrict(), scalar.ts:210-233)."""
    N = 2 * nx + 1
    kc = W._kuhn_corners()  # (6, 4, 3) corner offsets, oriented as the mesh orients them
    hk, hj, hi = np.meshgrid(np.arange(nx), np.arange(nx), np.arange(nx), indexing="ij")
    base = 2 * np.stack([hi.ravel(), hj.ravel(), hk.ravel()], 1)  # fine-grid origin of each hex
    edges = [(0, 1), (1, 2), (0, 2), (0, 3), (1, 3), (2, 3)]
    pts = np.concatenate([2 * kc] + [kc[:, [a]] + kc[:, [b]] for a, b in edges], axis=1)  # (6, 10, 3)
    fine = base[:, None, None, :] + pts[None]  # (nh, 6, 10, 3)
    ids = ((fine[..., 2] * N + fine[..., 1]) * N + fine[..., 0]).reshape(-1, 10)
    rows = np.repeat(ids, 10, axis=1).ravel()
    cols = np.tile(ids, (1, 10)).ravel()
    nn = N ** 3
    M = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(nn, nn))
    if restrict_interior:
        g = np.arange(nn)
        i, j, k = g % N, (g // N) % N, g // (N * N)
        inner = (i > 0) & (i < N - 1) & (j > 0) & (j < N - 1) & (k > 0) & (k < N - 1)
        keep = np.nonzero(inner)[0]
        M = M[keep][:, keep]
    return W.csr_dict(M)


# DevLog-045-perf-core-1.md:117-128: AMD symbolic flop per DOF of the 3D unit-cube tet10 heat problem at h 1 / 0.7 / 0.5
# (nx = 8 / 12 / 16): unrestricted node pattern 6.0e4 / 2.7e5 / 7.6e5 per node; restricted to the free DOFs (all six
# faces Dirichlet: 3,375 / 12,167 / 29,791 DOFs) 1.92e5 at h 0.7 and 7.63e5 at h 0.5.
PUBLISHED_FLOPS_PER_DOF = [(8, False, "6.0e+04"), (12, False, "2.7e+05"), (16, False, "7.6e+05"),
                           (12, True, "1.92e+05"), (16, True, "7.63e+05")]


@pytest.mark.parametrize("nx,restricted,published", PUBLISHED_FLOPS_PER_DOF)
def test_amd_fill_matches_published_tet10_figures(nx, restricted, published):
    """The reference publishes the fill its AMD + analysis produce on this mesh family; a different tie-break, hash,
    absorption or amalgamation step moves these figures (the natural order, for instance, gives 2.0e5 instead of
    6.0e4 at nx = 8). Both the oracle and the native symbolic phase must reproduce them to the published digits."""
    pat = tet10_box_pattern(nx, restricted)
    if restricted:
        assert pat["n"] == (2 * nx - 1) ** 3
    digits = len(published.split("e")[0].split(".")[1])
    S = O.Sym(pat)
    assert f"{S.flops / pat['n']:.{digits}e}" == published
    import ampsolve as amp

    sym = amp.analyze(pat)
    assert sym.flops == S.flops and np.array_equal(sym.perm, S.perm) and np.array_equal(sym.Si, S.Si)


def merge_tree_pattern(ncC, ncP, t, g):
    """Natural-order pattern with four fundamental supernodes C, P, H, G (columns in that order): C a clique of ncC
    columns coupled to the first t columns of P; P a clique of ncP columns coupled to all g columns of G; H one column
    coupled to the first column of G (G's second child, which keeps P and G apart); G a clique (the root)."""
    n = ncC + ncP + 1 + g
    C = np.arange(ncC)
    P = ncC + np.arange(ncP)
    H = ncC + ncP
    G = ncC + ncP + 1 + np.arange(g)
    M = np.zeros((n, n), bool)
    for blk in (C, P, G):
        M[np.ix_(blk, blk)] = True
    M[np.ix_(C, P[:t])] = True
    M[np.ix_(P, G)] = True
    M[H, H] = M[H, G[0]] = True
    M = M | M.T
    return W.csr_dict(sp.csr_matrix(M.astype(float))), ncC


# (ncC, ncP, t, g) -> merged? The merged block's explicit zeros are ncC (ncP + g - t) of
# nc m - nc (nc - 1) / 2 stored entries (nc = ncC + ncP, m = ncC + ncP + g); zf = zeros / stored, worked by hand:
MERGE_CASES = [
    ((3, 1, 1, 200), True),     # nc 4: zf 0.741, merged by nc <= 4
    ((15, 1, 1, 10), True),     # nc 16: zf 0.507 < 0.8
    ((16, 1, 1, 10), False),    # nc 17: zf 0.495, above 0.1 and 0.05
    ((15, 1, 1, 49), True),     # nc 16: zf 735 / 920 = 0.79891 < 0.8
    ((15, 1, 1, 50), False),    # nc 16: zf 750 / 936 = 0.80128
    ((47, 1, 1, 2), True),      # nc 48: zf 94 / 1272 = 0.0739 < 0.1
    ((47, 1, 1, 3), False),     # nc 48: zf 141 / 1320 = 0.1068
    ((48, 1, 1, 2), False),     # nc 49: zf 96 / 1323 = 0.0726, not < 0.05
    ((99, 1, 1, 2), True),      # nc 100: zf 198 / 5250 = 0.0377 < 0.05
    ((99, 1, 1, 3), False),     # nc 100: zf 297 / 5350 = 0.0555
]


@pytest.mark.parametrize("params,merged", MERGE_CASES)
def test_relaxed_amalgamation_thresholds(params, merged):
    """sparse.mjs:790-812 (CHOLMOD's defaults): a child supernode joins its parent when nc <= 4, or nc <= 16 and
    zf < 0.8, or nc <= 48 and zf < 0.1, or zf < 0.05. One case just inside and one just outside each threshold."""
    pat, ncC = merge_tree_pattern(*params)
    n = pat["n"]
    for relax in (True, False):
        S = O.Sym(pat, perm=np.arange(n), relax=relax)
        together = S.colSuper[0] == S.colSuper[ncC]
        assert together == (merged and relax)
        assert S.colSuper[0] == S.colSuper[ncC - 1]  # C is one fundamental supernode
        import ampsolve as amp

        sym = amp.analyze(pat, {"perm": np.arange(n, dtype=np.int32), "relax": relax})
        assert np.array_equal(sym.superStart, S.superStart)
    # the merged block stores the explicit zeros: nnzLStored grows by exactly that count over the unrelaxed figure
    ncC_, ncP, t, g = params
    S1 = O.Sym(pat, perm=np.arange(n), relax=True)
    S0 = O.Sym(pat, perm=np.arange(n), relax=False)
    assert S0.nnzL == S1.nnzL
    if merged:
        assert S1.nnzLStored - S0.nnzLStored >= ncC_ * (ncP + g - t)


def test_threaded_factor_is_bitwise_the_single_thread_one():
    """The oracle splits front rows and Schur target columns among threads without changing any entry's operation
    order: factors with 1 and 4 threads are bitwise equal (Cholesky and LDL^T)."""
    pat, v, _ = W.heat3d(20)
    S = O.Sym(pat)
    keep = O.get_threads()
    try:
        O.set_threads(1)
        ok1, L1 = O.chol_factor(S, v)
        _, Ld1, D1, nn1, _ = O.ldlt_factor(S, v)
        O.set_threads(4)
        ok4, L4 = O.chol_factor(S, v)
        _, Ld4, D4, nn4, _ = O.ldlt_factor(S, v)
    finally:
        O.set_threads(keep)
    assert ok1 and ok4 and S.maxFront > 512
    assert np.array_equal(L1, L4) and np.array_equal(Ld1, Ld4) and np.array_equal(D1, D4) and nn1 == nn4 == 0


# ---- P1 assembly (assembleCells / elementSlots / restrict restated in oracle/assembly.c) -----------------------------

def test_assembly_single_reference_tet_closed_form():
    """The unit tet's P1 stiffness is k |T| G G^T with |T| = 1/6, G rows (-1,-1,-1), e1, e2, e3."""
    nodes = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float)
    pat, v = O.assemble_p1(nodes, np.array([[0, 1, 2, 3]]), np.array([2.5]))
    K = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(4, 4)).toarray()
    ref = 2.5 / 6 * np.array([[3, -1, -1, -1], [-1, 1, 0, 0], [-1, 0, 1, 0], [-1, 0, 0, 1]], float)
    assert np.abs(K - ref).max() < 1e-15


@pytest.mark.parametrize("dim", [2, 3])
def test_assembly_row_sums_patch_and_energy(dim):
    """Unrestricted operator: zero row sums (constants), K u = 0 at interior nodes for a linear u with uniform k
    (patch test), and u^T K u = sum_e k_e |T_e| |grad u|^2 for a linear u with random k."""
    if dim == 3:
        nodes, conn = W.box_tet_mesh(6, 5, 4)
        g = np.array([0.3, -1.2, 0.7])
    else:
        nodes, conn = W.rect_tri_mesh(7, 6)
        g = np.array([0.3, -1.2])
    kel = 10.0 ** (2.0 * W.lcg_uniform(5, len(conn)))
    pat, v = O.assemble_p1(nodes, conn, kel)
    n = pat["n"]
    K = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(n, n))
    assert np.abs(K @ np.ones(n)).max() < 1e-13 * np.abs(v).max()
    u = nodes @ g + 2.0
    X = nodes[conn]
    J = np.transpose(X[:, 1:, :] - X[:, :1, :], (0, 2, 1))
    vol = np.abs(np.linalg.det(J)) / (6.0 if dim == 3 else 2.0)
    assert abs(u @ (K @ u) / (g @ g * np.sum(kel * vol)) - 1) < 1e-13
    patU, vU = O.assemble_p1(nodes, conn, np.ones(len(conn)))
    KU = sp.csr_matrix((vU, patU["colidx"], patU["rowptr"]), shape=(n, n))
    lo, hi = nodes.min(0), nodes.max(0)
    interior = np.all((nodes > lo + 1e-12) & (nodes < hi - 1e-12), axis=1)
    assert np.abs((KU @ u)[interior]).max() < 1e-12


def test_assembly_restriction_matches_the_configs():
    """The restricted pattern and values equal the workload generators' (heat3d, poisson2d), which build the
    reference's element pattern with numpy."""
    nodes, tets = W.box_tet_mesh(8, 8, 8)
    kel = 10.0 ** (2.0 * W.lcg_uniform(1, len(tets)))
    pat, v = O.assemble_p1(nodes, tets, kel, nodes[:, 2] > 1e-12)
    pw, vw, _ = W.heat3d(8)
    assert np.array_equal(pat["rowptr"], pw["rowptr"]) and np.array_equal(pat["colidx"], pw["colidx"])
    assert np.abs(v - vw).max() <= 4e-15 * np.abs(vw).max()  # sums in another order: a few ulps of the largest
    nodes, tris = W.rect_tri_mesh(16, 16)
    x, y = nodes[:, 0], nodes[:, 1]
    free = (x > 1e-12) & (x < 1 - 1e-12) & (y > 1e-12) & (y < 1 - 1e-12)
    pat, v = O.assemble_p1(nodes, tris, np.ones(len(tris)), free)
    pw, vw, _ = W.poisson2d(16)
    assert np.array_equal(pat["colidx"], pw["colidx"]) and np.abs(v - vw).max() <= 4e-15 * np.abs(vw).max()


# ---- complex-symmetric LDL^T (ComplexSupernodalLDLT restated in oracle/complex.c) -----------------------------------

def _cdense(pat, re, im):
    A = sp.csr_matrix((re + 1j * im, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))
    return A.toarray()


@pytest.mark.parametrize("N,shift", [(6, 0.0), (6, 13.0), (5, 26.0)])
def test_complex_ldlt_matches_dense_solve(N, shift):
    """The complex factor and solve against LAPACK's complex solve of the dense matrix (the la-complex-solve grids)."""
    from fixtures import complex_grid27

    pat, re, im = complex_grid27(N, shift=shift)
    n = pat["n"]
    S = O.Sym(pat)
    ok, Lr, Li, Dr, Di, npert = O.cldlt_factor(S, re, im)
    assert ok and npert == 0
    br, bi = np.sin(0.3 * np.arange(n)), 0.5 * np.cos(0.7 * np.arange(n))
    xr, xi = O.cldlt_solve(S, Lr, Li, Dr, Di, br, bi)
    x = np.linalg.solve(_cdense(pat, re, im), br + 1j * bi)
    assert np.abs(xr + 1j * xi - x).max() / np.abs(x).max() < 1e-11


def test_complex_ldlt_zero_imaginary_part_is_the_real_factor():
    """With im = 0 the complex factor carries the real LDL^T's values (another summation order: to round-off) and
    zero imaginary parts exactly."""
    from fixtures import complex_grid27

    pat, re, _ = complex_grid27(6, shift=13.0)
    im = np.zeros_like(re)
    S = O.Sym(pat)
    ok, Lr, Li, Dr, Di, npert = O.cldlt_factor(S, re, im)
    okr, Lx, D, nneg, nper = O.ldlt_factor(S, re)
    assert ok and okr and not np.any(Li) and not np.any(Di)
    assert np.abs(Dr - D).max() / np.abs(D).max() < 1e-13
    m = np.zeros(S.nnzLStored, bool)
    for s in range(S.nsuper):
        nc, mm = S.superStart[s + 1] - S.superStart[s], S.Sp[s + 1] - S.Sp[s]
        m[S.Lxp[s]:S.Lxp[s + 1]] = np.tril(np.ones((mm, nc), bool)).T.ravel()
    assert np.abs(Lr[m] - Lx[m]).max() / np.abs(Lx[m]).max() < 1e-12


def test_complex_ldlt_perturbed_pivots_and_refinement():
    """la-ldlt.test.mjs:70-82: exact zero pivots in the identity order are perturbed; the plain solve is visibly wrong
    and 30 complex refinement steps recover the exact solution to 1e-10."""
    from fixtures import complex_ldlt_system

    pat, re, im, xr, xi, br, bi = complex_ldlt_system(40)
    S = O.Sym(pat, perm=np.arange(pat["n"]))
    ok, Lr, Li, Dr, Di, npert = O.cldlt_factor(S, re, im)
    assert ok and npert > 0
    pr, pi = O.cldlt_solve(S, Lr, Li, Dr, Di, br, bi)
    assert np.abs(pr - xr).max() / np.abs(xr).max() > 1e-6
    rr, ri, rel, steps = O.cldlt_solve_refined(S, Lr, Li, Dr, Di, pat, re, im, br, bi, max_steps=30)
    assert np.abs(rr - xr).max() / np.abs(xr).max() < 1e-10 and np.abs(ri - xi).max() / np.abs(xi).max() < 1e-10


# ---- ndOrder (sparse.mjs:230-283), restated in oracle/symbolic.c and csrc/symbolic.cpp -----------------------------

def grid_pattern(nx, ny, nine=False):
    """5-point (or 9-point) grid graph, node i + nx j at (x, y) = (i, j)."""
    Ix, Iy = sp.identity(nx), sp.identity(ny)
    Tx = sp.diags([1, 1, 1], [-1, 0, 1], shape=(nx, nx))
    Ty = sp.diags([1, 1, 1], [-1, 0, 1], shape=(ny, ny))
    M = sp.kron(Ty, Tx) if nine else sp.kron(Iy, Tx) + sp.kron(Ty, Ix)
    M = sp.csr_matrix(M)
    M.sort_indices()
    g = np.arange(nx * ny)
    return W.csr_dict(M), (g % nx).astype(float), (g // nx).astype(float)


def test_nd_order_of_a_path_closed_form():
    """A 100-node path at x = i: split at half the length (50), the first half's node next to the second half is the
    separator, the first half (49 nodes) splits again at 24 with separator 23, the second (50) at 25 with separator
    74; ranges of at most 48 nodes keep their order; the order is first half, second half, separator at every level.
    Laid out vertically (x = 0, y = i) the bisection takes y and gives the same order."""
    n = 100
    M = sp.diags([1, 1, 1], [-1, 0, 1], shape=(n, n), format="csr")
    pat = W.csr_dict(M)
    want = list(range(0, 23)) + list(range(24, 49)) + [23] + list(range(50, 74)) + list(range(75, 100)) + [74, 49]
    i = np.arange(n, dtype=float)
    assert O.nd_order(pat, i, np.zeros(n)).tolist() == want
    assert O.nd_order(pat, np.zeros(n), i).tolist() == want
    # ranges of at most `leaf` nodes keep their order
    assert O.nd_order(pat, i, np.zeros(n), leaf=100).tolist() == list(range(n))
    p48 = W.csr_dict(sp.diags([1, 1, 1], [-1, 0, 1], shape=(48, 48), format="csr"))
    assert O.nd_order(p48, np.arange(48.0), np.zeros(48)).tolist() == list(range(48))


def test_nd_order_axis_rule_and_stable_ties():
    """A 10 x 10 grid has equal extents, so the bisection takes x (>=); the sort is stable, so the first separator --
    the last ten nodes of the order -- is the column x = 4 in node order (4, 14, ..., 94). Taking y, or an unstable
    sort, would change it."""
    pat, x, y = grid_pattern(10, 10)
    perm = O.nd_order(pat, x, y)
    assert perm[-10:].tolist() == [4 + 10 * j for j in range(10)]
    assert sorted(perm.tolist()) == list(range(100))


@pytest.mark.parametrize("nine", [False, True])
def test_nd_order_reduces_fill_and_native_matches_oracle(nine):
    """On grids the dissection order has far less fill than the natural order; the native ordering (amp_nd_order)
    equals the oracle's on grids and on jittered coordinates with many ties broken by node order."""
    import ampsolve as amp

    pat, x, y = grid_pattern(40, 40, nine)
    perm = O.nd_order(pat, x, y)
    S_nd = O.Sym(pat, perm=perm)
    S_nat = O.Sym(pat, perm=np.arange(pat["n"]))
    assert S_nd.nnzL < (0.75 if nine else 0.6) * S_nat.nnzL
    assert np.array_equal(amp.ndOrder(pat, x, y), perm)
    rng = np.random.default_rng(3)
    for leaf in (48, 5, 1):
        xj = np.round(x + rng.uniform(-0.4, 0.4, pat["n"]), 1)
        yj = np.round(y * 1.3 + rng.uniform(-0.4, 0.4, pat["n"]), 1)
        assert np.array_equal(amp.ndOrder(pat, xj, yj, leaf), O.nd_order(pat, xj, yj, leaf))
