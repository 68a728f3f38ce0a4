"""Shift-invert eigen driver over the device factor: block Arnoldi with thick restart, driven by the device operator
Y = (K - sigma M)^-1 M X (amp_shift_invert_apply), one multi-right-hand-side solve per block of vectors.

Restates packages/core/src/la/arnoldi.mjs (blockArnoldi :39-157, ritzBasis :160-213, ritz :216-278; DGKS
reorthogonalisation, thick restart from the kept Ritz vectors) and the core of shiftInvertEigen in
packages/physics/src/common/eigen.ts:115-216 (factor choice, `op`, one inverse-iteration refinement step, Rayleigh
quotients, M-orthonormalisation inside clusters). The Arnoldi basis and the small Hessenberg problem stay on the host
(numpy; the reference's vector-dense.mjs eigen routines are replaced by LAPACK's); every operator application runs on
the device.
"""
import time

import numpy as np

from .sparse import NumericFactor, analyze


def seeded_random(seed):
    """The reference's LCG stream u - 0.5 (arnoldi.mjs:13-19)."""
    s = seed & 0xFFFFFFFF

    def rnd(count):
        nonlocal s
        out = np.empty(count)
        for i in range(count):
            s = (s * 1664525 + 1013904223) & 0xFFFFFFFF
            out[i] = s / 4294967296.0 - 0.5
        return out

    return rnd


def _eig_sorted(Hm, real_only):
    lam, Y = np.linalg.eig(Hm)
    mag = np.abs(lam)
    real = np.abs(lam.imag) <= 1e-9 * mag
    idx = [i for i in range(len(lam)) if (real[i] if real_only else (real[i] or lam[i].imag > 0))]
    idx.sort(key=lambda i: -mag[i])
    return lam, Y, real, idx


def ritz(H, m, V, nWant, tol, real_only, form_vectors):
    """Ritz pairs of the leading m x m block of H, residual norms |H[m:, :m] y| / |mu| (arnoldi.mjs:216-278)."""
    Hm = H[:m, :m]
    lam, Y, real, idx = _eig_sorted(Hm, real_only)
    want = idx[:nWant]
    rows = V.shape[1] - m
    mus, vecs, res = [], [], []
    converged = len(want) == nWant
    for i in want:
        mu = complex(lam[i].real, 0.0 if real[i] else lam[i].imag)
        y = Y[:, i] if not real[i] else Y[:, i].real / np.linalg.norm(Y[:, i].real)
        r = np.linalg.norm(H[m:m + rows, :m] @ y) / abs(mu)
        res.append(float(r))
        if not r < tol:
            converged = False
        mus.append(mu)
        if form_vectors:
            vecs.append(V[:, :m] @ y)
    return {"mu": mus, "vecs": vecs, "ritzRes": res, "converged": converged}


def ritz_basis(H, m, nKeep, real_only):
    """Orthonormal basis (m x k) of the nKeep dominant Ritz vectors of H_m (arnoldi.mjs:160-213)."""
    Hm = H[:m, :m].copy()
    lam, Y, real, idx = _eig_sorted(Hm, real_only)
    cols, used = [], 0
    for i in idx:
        need = 1 if real[i] else 2
        if used + need > nKeep:
            break
        used += need
        if real[i]:
            cols.append(Y[:, i].real)
        else:
            cols += [Y[:, i].real, Y[:, i].imag]
    kept = []
    for c0 in cols:  # modified Gram-Schmidt twice, dependent columns dropped
        c = np.array(c0, dtype=np.float64)
        n0 = np.linalg.norm(c)
        for _ in range(2):
            for q in kept:
                c -= (q @ c) * q
        nc = np.linalg.norm(c)
        if nc > 1e-8 * n0:
            kept.append(c / nc)
    Q = np.stack(kept, axis=1) if kept else np.zeros((m, 0))
    return Q, Hm


def block_arnoldi(N, op, nWant, blockSize, maxBasis, tol, X0=None, maxRestarts=60, seed=20261014, realOnly=True,
                  nKeep=None):
    """Block Arnoldi with DGKS reorthogonalisation and thick restart for the largest-modulus eigenvalues of `op`
    (arnoldi.mjs:39-157). op maps an (N, b) block to an (N, b) block. Returns mu, vecs, ritzRes, basis, opVectors,
    restarts, converged."""
    rnd = seeded_random(seed)
    b = blockSize
    mMax = max(maxBasis, nWant + 3 * b)
    keepN = min(nKeep if nKeep is not None else nWant + max(2, nWant), mMax - 2 * b)
    ld = mMax
    H = np.zeros((ld + b, ld))
    V = np.zeros((N, ld + b))
    nv = 0
    op_vectors = restarts = 0

    def add_vec(w, col):
        nonlocal nv
        n0 = np.linalg.norm(w)
        for _ in range(2):
            h = V[:, :nv].T @ w
            if col >= 0:
                H[:nv, col] += h
            w = w - V[:, :nv] @ h
            if np.linalg.norm(w) > 0.7 * n0:
                break
        nw = np.linalg.norm(w)
        if not nw > 1e-12 * (n0 or 1):  # breakdown: continue with a random direction, H entry 0
            w = rnd(N)
            for _ in range(2):
                w = w - V[:, :nv] @ (V[:, :nv].T @ w)
            w = w / np.linalg.norm(w)
            nw = 0.0
        else:
            w = w / nw
        if col >= 0:
            H[nv, col] = nw
        V[:, nv] = w
        nv += 1

    for k in range(b):  # initial block: warm vectors folded into b columns, random otherwise
        if X0 is not None and k < len(X0):
            w = np.zeros(N)
            for j in range(k, len(X0), b):
                w += X0[j]
        else:
            w = rnd(N)
        add_vec(w, -1)
    col = 0
    while True:
        converged = False
        while col < mMax:
            nb = min(b, mMax - col)
            W = op(V[:, col:col + nb])
            op_vectors += nb
            for c in range(nb):
                add_vec(np.array(W[:, c]), col + c)
            col += nb
            if col >= nWant + b and ritz(H, col, V[:, :nv], nWant, tol, realOnly, False)["converged"]:
                converged = True
                break
        if converged or restarts >= maxRestarts:
            last = ritz(H, col, V[:, :nv], nWant, tol, realOnly, True)
            last.update(basis=col, opVectors=op_vectors, restarts=restarts)
            return last
        # thick restart: C U = U (Q^T H_m Q) + W (H_low Q)
        restarts += 1
        m = col
        Q, Hm = ritz_basis(H, m, keepN, realOnly)
        k = Q.shape[1]
        U = V[:, :m] @ Q
        Wv = V[:, m:m + b].copy()
        Hn = np.zeros_like(H)
        Hn[:k, :k] = Q.T @ (Hm @ Q)
        Hn[k:k + Wv.shape[1], :k] = H[m:m + Wv.shape[1], :m] @ Q
        H = Hn
        V = np.zeros_like(V)
        V[:, :k] = U
        V[:, k:k + Wv.shape[1]] = Wv
        nv = k + Wv.shape[1]
        col = k


def _csr_mv(pat, val, X):
    import scipy.sparse as sp

    A = sp.csr_matrix((val, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))
    return A @ X


def shiftInvertEigen(pat, K, M, sigma, nWant, blockSize=1, maxBasis=None, tol=1e-10, maxRestarts=60, seed=20261106,
                     factor=None, kind="hip", sym=None, kPsd=False):
    """Eigenpairs of K x = lambda M x nearest sigma (eigen.ts:115-216): factor A = K - sigma M on the device (Cholesky
    when sigma < 0 for a PSD K, else LDL^T), Arnoldi on X -> A^-1 M X with the device operator, one inverse-iteration
    step, Rayleigh quotients, M-orthonormalisation of clusters. Returns {lambda, vecs, residual, stats}."""
    n = int(pat["n"])
    K = np.asarray(K, np.float64)
    M = np.asarray(M, np.float64)
    A = K - sigma * M
    chol = factor == "cholesky" or (factor != "ldlt" and sigma < 0 and kPsd)
    t0 = time.perf_counter()
    sym = sym or analyze(pat)
    num = NumericFactor(sym, 0 if chol else 1, kind)
    ok = num.factor(A)
    if not ok:
        raise RuntimeError("eigen: K - sigma M is not positive definite (sigma not below the spectrum?)" if chol else
                           "eigen: LDL^T of K - sigma M failed (sigma on an eigenvalue?)")
    factor_ms = 1e3 * (time.perf_counter() - t0)
    from .device import DeviceArray

    dM = DeviceArray.from_numpy(M)
    apps = [0]

    def op(X):  # (n, b) host block -> (n, b): A^-1 M X on the device (vectors stored one after another)
        b = X.shape[1]
        dX = DeviceArray.from_numpy(np.ascontiguousarray(X.T))
        Y = num.shift_invert(dM, dX, b).numpy()
        apps[0] += b
        return Y.reshape(b, n).T

    t1 = time.perf_counter()
    ar = block_arnoldi(n, op, nWant, blockSize, maxBasis or max(30, 2 * nWant + 12), tol, None, maxRestarts, seed, True)
    arnoldi_ms = 1e3 * (time.perf_counter() - t1)
    t2 = time.perf_counter()
    X = op(np.stack([np.real(v) for v in ar["vecs"]], axis=1))  # one inverse-iteration step
    KX, MX = _csr_mv(pat, K, X), _csr_mv(pat, M, X)
    lam = np.einsum("ij,ij->j", X, KX) / np.einsum("ij,ij->j", X, MX)
    order = np.argsort(lam, kind="stable")
    X, lam = X[:, order], lam[order]
    done = []
    for i in range(X.shape[1]):  # M-Gram-Schmidt inside clusters (relative gap 1e-6), then M-normalise
        x = X[:, i].copy()
        for j in range(i):
            if abs(lam[j] - lam[i]) > 1e-6 * max(abs(lam[i]), 1e-300):
                continue
            y = done[j]
            x -= (x @ _csr_mv(pat, M, y)) * y
        x /= np.sqrt(x @ _csr_mv(pat, M, x))
        done.append(x)
    V = np.stack(done, axis=1)
    KV, MV = _csr_mv(pat, K, V), _csr_mv(pat, M, V)
    lam = np.einsum("ij,ij->j", V, KV) / np.einsum("ij,ij->j", V, MV)
    res = np.linalg.norm(KV - MV * lam, axis=0) / (np.linalg.norm(KV, axis=0) + np.abs(lam) * np.linalg.norm(MV, axis=0))
    info = num.info
    return {"lambda": lam, "vecs": V, "residual": res,
            "stats": {"n": n, "nnz": int(pat["nnz"]), "factor": "cholesky" if chol else "ldlt", "nnzL": sym.nnzL,
                      "nNegative": 0 if chol else info["nNeg"], "nPerturbed": info["nPerturbed"],
                      "factorMs": factor_ms, "arnoldiMs": arnoldi_ms, "refineMs": 1e3 * (time.perf_counter() - t2),
                      "opApplications": apps[0], "restarts": ar["restarts"], "converged": ar["converged"],
                      "ritzRes": ar["ritzRes"]}}
