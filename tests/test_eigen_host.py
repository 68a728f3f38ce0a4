"""The block Arnoldi restatement (ampsolve/eigen.py after arnoldi.mjs:39-278) on the host, with an exact LU operator:
the closed-form eigenvalues of the 1D Laplacian (core-solvers.test.mjs:70-80) for block sizes 1 and 2, and the
eigenvalues nearest an indefinite shift (core-solvers.test.mjs:83-100)."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as sl

from ampsolve import eigen as E
from fixtures import chain


def lap1d(n, k):
    return 2 - 2 * np.cos(k * np.pi / (n + 1))


@pytest.mark.parametrize("block", [1, 2])
def test_arnoldi_lowest_laplacian_eigenvalues(block):
    n = 400
    pat, v = chain(n)
    A = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(n, n)).tocsc()
    lu = sl.splu(A)
    ar = E.block_arnoldi(n, lambda X: lu.solve(X), nWant=4, blockSize=block, maxBasis=20, tol=1e-10, maxRestarts=30,
                         seed=20261014)
    assert ar["converged"]
    lam = np.sort([1 / m.real for m in ar["mu"]])
    for k in range(4):
        assert abs(lam[k] / lap1d(n, k + 1) - 1) < 1e-9
    for v_ in ar["vecs"]:  # Ritz vectors of a positive operator: positive Rayleigh quotient
        y = lu.solve(v_)
        assert v_ @ y > 0


def test_arnoldi_indefinite_shift():
    n = 300
    pat, _ = chain(n)
    sigma = 0.5 * (lap1d(n, 10) + lap1d(n, 11))
    _, v = chain(n, 2 - sigma, -1)
    lu = sl.splu(sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(n, n)).tocsc())
    ar = E.block_arnoldi(n, lambda X: lu.solve(X), nWant=2, blockSize=1, maxBasis=20, tol=1e-10, maxRestarts=30,
                         seed=7)
    assert ar["converged"]
    lam = np.sort([sigma + 1 / m.real for m in ar["mu"]])
    assert abs(lam[0] / lap1d(n, 10) - 1) < 1e-9 and abs(lam[1] / lap1d(n, 11) - 1) < 1e-9


def test_thick_restart_keeps_converging():
    """A basis too small for the wanted pairs forces restarts; the kept Ritz space must still converge."""
    n = 400
    pat, v = chain(n)
    lu = sl.splu(sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(n, n)).tocsc())
    ar = E.block_arnoldi(n, lambda X: lu.solve(X), nWant=6, blockSize=1, maxBasis=14, tol=1e-10, maxRestarts=60,
                         seed=3)
    assert ar["converged"] and ar["restarts"] > 0
    lam = np.sort([1 / m.real for m in ar["mu"]])
    assert np.allclose(lam, [lap1d(n, k) for k in range(1, 7)], rtol=1e-9, atol=0)
