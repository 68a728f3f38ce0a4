"""Test matrices restated from the reference's own test fixtures (jwt625/applied-multiphysics packages/core/test)."""
import numpy as np
import scipy.sparse as sp

from ampsolve import workloads as W


def element_pattern(n, elems):
    """buildElementPattern (sparse.mjs:86-110): full symmetric CSR with the diagonal, rows sorted."""
    elems = np.asarray(elems, dtype=np.int64)
    k = elems.shape[1]
    rows = np.repeat(elems, k, axis=1).ravel()
    cols = np.tile(elems, (1, k)).ravel()
    keys = np.unique(np.concatenate([rows * n + cols, np.arange(n, dtype=np.int64) * (n + 1)]))
    r, c = keys // n, keys % n
    rowptr = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(np.bincount(r, minlength=n), out=rowptr[1:])
    return {"n": n, "rowptr": rowptr.astype(np.int32), "colidx": c.astype(np.int32), "nnz": len(c)}


def grid27(Nx, Ny, Nz, zeroEvery=0, shift=0.0, symmetric=False):
    """27-point hex-element pattern with synthetic values: grid() of la-real-front.test.mjs:15-31 (symmetric=False)
    and of la-real-solve.test.mjs:15-31 (symmetric=True, min / max in the off-diagonal formula)."""
    n = Nx * Ny * Nz
    idx = lambda i, j, k: i + Nx * (j + Ny * k)  # noqa: E731
    el = []
    for k in range(Nz - 1):
        for j in range(Ny - 1):
            for i in range(Nx - 1):
                el.append([idx(i + a, j + b, k + c) for a, b, c in
                           [(0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)]])
    pat = element_pattern(n, el)
    r = np.repeat(np.arange(n), np.diff(pat["rowptr"]))
    c = pat["colidx"].astype(np.int64)
    if symmetric:
        off = -1 + 0.001 * ((np.minimum(r, c) * 31 + np.maximum(r, c) * 17) % 11)
    else:
        off = -1 + 0.001 * ((r * 31 + c * 17) % 11)
    if zeroEvery:
        off = np.where((r + c) % zeroEvery == 0, 0.0, off)
    diag = 26.5 + 0.01 * (r % 7) - shift
    v = np.where(r == c, diag, off).astype(np.float64)
    return pat, v, n


def chain(n, d=2.0, o=-1.0):
    """Tridiagonal chain tridiag(o, d, o) (the 1D Laplacian of core-solvers.test.mjs)."""
    A = sp.diags([np.full(n - 1, o), np.full(n, d), np.full(n - 1, o)], [-1, 0, 1], format="csr")
    A.sort_indices()
    return W.csr_dict(A), A.data.astype(np.float64)


def lcg_rnd(seed):
    """rnd() of the reference tests: (Math.imul(seed, 1664525) + 1013904223) >>> 0, / 2^32 - 0.5."""
    state = [seed]

    def rnd():
        state[0] = (state[0] * 1664525 + 1013904223) % (1 << 32)
        return state[0] / 4294967296 - 0.5

    return rnd


def ldlt_system(nb, seed=4242):
    """system(nb, false) of la-ldlt.test.mjs:12-45: 2x2 diagonal blocks [[0, a], [a, 0]] chained; exact x."""
    n = 2 * nb
    rnd = lcg_rnd(seed)
    ent = {}

    def add(i, j, v):
        for a, b in ([(i, j)] if i == j else [(i, j), (j, i)]):
            ent[(a, b)] = ent.get((a, b), 0.0) + v

    for b in range(nb):
        add(2 * b, 2 * b + 1, 2 + rnd())  # the imaginary draw only happens for the complex system
        add(2 * b, 2 * b, 0.0)
        add(2 * b + 1, 2 * b + 1, 0.0)
        if b + 1 < nb:
            add(2 * b + 1, 2 * b + 2, 0.3 * rnd())
    keys = sorted(ent.keys())
    rowptr = np.zeros(n + 1, dtype=np.int32)
    for (i, _j) in keys:
        rowptr[i + 1] += 1
    rowptr = np.cumsum(rowptr).astype(np.int32)
    colidx = np.array([j for (_i, j) in keys], dtype=np.int32)
    vals = np.array([ent[k] for k in keys], dtype=np.float64)
    pat = {"n": n, "rowptr": rowptr, "colidx": colidx, "nnz": len(keys)}
    x = np.array([rnd() for _ in range(n)])
    A = sp.csr_matrix((vals, colidx, rowptr), shape=(n, n))
    return pat, vals, x, A @ x


def hex_matrix(nx=6, ny=5, nz=4):
    """hexMatrix of la-sparse.test.mjs:13-63: Q1 hexes on a graded box, diagonal conductivity, Dirichlet on x = 0."""
    xs = np.array([(i / nx) ** 1.3 for i in range(nx + 1)])
    ys = np.array([j / ny for j in range(ny + 1)])
    zs = np.array([0.2 * (k / nz) ** 0.8 for k in range(nz + 1)])
    N = lambda i, j, k: i + (nx + 1) * (j + (ny + 1) * k)  # noqa: E731
    n = (nx + 1) * (ny + 1) * (nz + 1)
    S = np.array([[1, -1], [-1, 1]], dtype=float)
    M = np.array([[2, 1], [1, 2]], dtype=float)
    elems, kes = [], []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                hx, hy, hz = xs[i + 1] - xs[i], ys[j + 1] - ys[j], zs[k + 1] - zs[k]
                kx, ky, kz = 1 + (i % 2), 0.5, 3 + (k % 3)
                K = np.zeros((8, 8))
                for a in range(8):
                    for b in range(8):
                        ia, ja, ka = a & 1, (a >> 1) & 1, a >> 2
                        ib, jb, kb = b & 1, (b >> 1) & 1, b >> 2
                        K[a, b] = (kx * (S[ia][ib] / hx) * (M[ja][jb] * hy / 6) * (M[ka][kb] * hz / 6)
                                   + ky * (M[ia][ib] * hx / 6) * (S[ja][jb] / hy) * (M[ka][kb] * hz / 6)
                                   + kz * (M[ia][ib] * hx / 6) * (M[ja][jb] * hy / 6) * (S[ka][kb] / hz))
                kes.append(K)
                elems.append([N(i + (a & 1), j + ((a >> 1) & 1), k + (a >> 2)) for a in range(8)])
    elems = np.array(elems)
    A = W.assemble(n, elems, np.array(kes))
    free = np.ones(n, dtype=bool)
    for k in range(nz + 1):
        for j in range(ny + 1):
            free[N(0, j, k)] = False
    R, _ = W.restrict(A, free)
    return W.csr_dict(R), R.data.astype(np.float64)


def complex_grid27(N, zeroEvery=0, shift=0.0):
    """grid() of la-complex-solve.test.mjs:16-36: 27-point pattern on an N^3 grid (hex-element pattern), complex-
    symmetric values; `shift` makes the real part indefinite. Returns (pat, re, im)."""
    n = N * N * N
    idf = lambda i, j, k: i + N * (j + N * k)  # noqa: E731
    el = []
    for k in range(N - 1):
        for j in range(N - 1):
            for i in range(N - 1):
                for a, b, c in ((0, 0, 0), (1, 0, 0), (0, 1, 0), (1, 1, 0), (0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1)):
                    el.append(idf(i + a, j + b, k + c))
    pat = element_pattern(n, np.array(el, dtype=np.int64).reshape(-1, 8))
    rows = np.repeat(np.arange(n), np.diff(pat["rowptr"]))
    cols = pat["colidx"].astype(np.int64)
    lo, hi = np.minimum(rows, cols), np.maximum(rows, cols)
    diag = rows == cols
    re = np.where(diag, 26.5 + 0.01 * (rows % 7) - shift, -1 + 0.001 * ((lo * 31 + hi * 17) % 11))
    if zeroEvery:
        re = np.where(~diag & ((rows + cols) % zeroEvery == 0), 0.0, re)
    im = np.where(diag, 0.2 + 0.001 * (rows % 5), np.where((lo + hi) % 3 == 0, 0.0, 0.01 * ((lo * 7 + hi) % 5)))
    return pat, re.astype(np.float64), im.astype(np.float64)


def complex_ldlt_system(nb, seed=4242):
    """system(nb, true) of la-ldlt.test.mjs:12-45: 2x2 diagonal blocks [[0, a], [a, 0]] with complex a chained by
    complex couplings; the exact solution x = xr + i xi and b = A x. Returns (pat, re, im, xr, xi, br, bi)."""
    n = 2 * nb
    rnd = lcg_rnd(seed)
    ent = {}

    def add(i, j, vr, vi):
        for a, b in ([(i, j)] if i == j else [(i, j), (j, i)]):
            v = ent.get((a, b), (0.0, 0.0))
            ent[(a, b)] = (v[0] + vr, v[1] + vi)

    for b in range(nb):
        vr = 2 + rnd()
        vi = 0.3 * rnd()
        add(2 * b, 2 * b + 1, vr, vi)
        add(2 * b, 2 * b, 0.0, 0.0)
        add(2 * b + 1, 2 * b + 1, 0.0, 0.0)
        if b + 1 < nb:
            vr = 0.3 * rnd()
            vi = 0.1 * rnd()
            add(2 * b + 1, 2 * b + 2, vr, vi)
    keys = sorted(ent.keys())
    rowptr = np.zeros(n + 1, dtype=np.int32)
    for (i, _j) in keys:
        rowptr[i + 1] += 1
    rowptr = np.cumsum(rowptr).astype(np.int32)
    colidx = np.array([j for (_i, j) in keys], dtype=np.int32)
    re = np.array([ent[k][0] for k in keys])
    im = np.array([ent[k][1] for k in keys])
    pat = {"n": n, "rowptr": rowptr, "colidx": colidx, "nnz": len(keys)}
    xr = np.array([rnd() for _ in range(n)])
    xi = np.array([rnd() for _ in range(n)])
    br, bi = np.zeros(n), np.zeros(n)
    for i in range(n):
        for p in range(rowptr[i], rowptr[i + 1]):
            j = colidx[p]
            br[i] += re[p] * xr[j] - im[p] * xi[j]
            bi[i] += re[p] * xi[j] + im[p] * xr[j]
    return pat, re, im, xr, xi, br, bi
