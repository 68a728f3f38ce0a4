"""Shift-invert operator applications on the device (amp_shift_invert_apply: Y = (K - sigma M)^-1 M X, one
multi-right-hand-side solve per block) driven by the block Arnoldi of ampsolve/eigen.py (eigen.ts:115-216,
arnoldi.mjs:39-278). Pins: the operator against a sparse LU of the same matrices; the 1D Laplacian's closed-form
eigenvalues 2 - 2 cos(k pi / (n + 1)) (core-solvers.test.mjs:70-80); the eigenvalues nearest an indefinite shift with
Sylvester's count nNeg = 10 (core-solvers.test.mjs:83-100); the PEC unit-cube cavity's lowest k^2 = 2 pi^2 (three
modes, em3d.test.ts:30-32, within 2e-2 as there) against ARPACK on the same discrete pencil, with the inertia of the
gradient null space ((ncell - 1)^3 interior nodes below the shift)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sl

import ampsolve as amp
from ampsolve import eigen as E
from ampsolve import workloads as W
from fixtures import chain

pytestmark = pytest.mark.gpu


def lap1d(n, k):
    return 2 - 2 * np.cos(k * np.pi / (n + 1))


def csr(pat, v):
    return sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))


@pytest.mark.parametrize("kind", ["hip", "hip-f32"])
def test_operator_matches_sparse_lu(device, kind):
    pat, K, M = W.maxwell_km(6)
    n = pat["n"]
    sigma = 15.0
    A = K - sigma * M
    sym = amp.analyze(pat)
    num = amp.NumericFactor(sym, 1, kind)
    assert num.factor(A)
    X = np.stack([W.lcg_normal(60 + k, n) for k in range(5)])  # 5 vectors, one after another
    Yh = num.shift_invert(M, X.ravel(), 5).reshape(5, n)
    Yd = num.shift_invert(device.DeviceArray.from_numpy(M), device.DeviceArray.from_numpy(X.ravel()), 5)
    lu = sl.splu(csr(pat, A).tocsc())
    Yo = lu.solve(csr(pat, M) @ X.T).T
    err = np.abs(Yh - Yo).max() / np.abs(Yo).max()
    print(f"shift-invert operator {kind}: max rel diff vs LU {err:.2e}", flush=True)
    assert err < (1e-11 if kind == "hip" else 1e-3)
    assert np.array_equal(Yh, Yd.numpy().reshape(5, n))


@pytest.mark.parametrize("block", [1, 2])
def test_laplacian_closed_form(block):
    n = 400
    pat, K = chain(n)
    I = np.where(pat["colidx"] == np.repeat(np.arange(n), np.diff(pat["rowptr"])), 1.0, 0.0)
    r = E.shiftInvertEigen(pat, K, I, 0.0, 4, blockSize=block, maxBasis=20, factor="cholesky")
    assert r["stats"]["converged"] and r["stats"]["factor"] == "cholesky"
    assert np.allclose(r["lambda"], [lap1d(n, k) for k in range(1, 5)], rtol=1e-9, atol=0)
    assert r["residual"].max() < 1e-10


def test_indefinite_shift_nearest_pair_and_inertia():
    n = 300
    pat, K = chain(n)
    I = np.where(pat["colidx"] == np.repeat(np.arange(n), np.diff(pat["rowptr"])), 1.0, 0.0)
    sigma = 0.5 * (lap1d(n, 10) + lap1d(n, 11))
    r = E.shiftInvertEigen(pat, K, I, sigma, 2, maxBasis=20, seed=7)
    st = r["stats"]
    assert st["factor"] == "ldlt" and st["nNegative"] == 10 and st["nPerturbed"] == 0 and st["converged"]
    assert np.allclose(r["lambda"], [lap1d(n, 10), lap1d(n, 11)], rtol=1e-9, atol=0)


def test_pec_cavity_lowest_modes():
    ncell, sigma = 10, 15.0
    pat, K, M = W.maxwell_km(ncell)
    r = E.shiftInvertEigen(pat, K, M, sigma, 3, blockSize=3, maxBasis=30, tol=1e-10)
    st = r["stats"]
    k2 = r["lambda"] / np.pi ** 2
    ref = np.sort(sl.eigsh(csr(pat, K), k=3, M=csr(pat, M), sigma=sigma, return_eigenvectors=False))
    print(f"cavity ncell={ncell} n={pat['n']}: k^2/pi^2 {k2}, ARPACK {ref / np.pi ** 2}, op applications "
          f"{st['opApplications']}, Arnoldi {st['arnoldiMs']:.1f} ms", flush=True)
    assert st["converged"] and st["nNegative"] == (ncell - 1) ** 3 and st["nPerturbed"] == 0
    assert np.allclose(r["lambda"], ref, rtol=1e-9, atol=0)
    assert np.all(np.abs(k2 - 2) < 2e-2)
    assert r["residual"].max() < 1e-9
