"""The 3D heat configuration at full size (61^3 nodes: 60^3 hexes x 6 Kuhn tets, log-uniform conductivity in [1, 100],
Dirichlet on z = 0, 223,260 free DOFs) against the oracle's complete factor + solve of the same system.

Bars: the device FP64 factor + solve within 1e-12 (max-norm relative) of the oracle's solution, with a relative
residual within 10x of the oracle's; the FP32 factor + FP64 refinement to |b - A x| / |b| <= 1e-10, its error against
the oracle printed and bounded. The oracle runs on every host core (its factor is bitwise the same for any thread
count)."""
import os
import time

import numpy as np
import pytest
import scipy.sparse as sp

import ampsolve as amp
import oracle as O
from ampsolve import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]


@pytest.fixture(scope="module")
def heat61():
    t0 = time.perf_counter()
    pat, v, b = W.heat3d(60)
    print(f"heat61: n={pat['n']} nnz={pat['nnz']} ({time.perf_counter() - t0:.1f} s)", flush=True)
    So = O.Sym(pat)
    keep = O.get_threads()
    O.set_threads(len(os.sched_getaffinity(0)))
    t0 = time.perf_counter()
    ok, Lo = O.chol_factor(So, v)
    tf = time.perf_counter() - t0
    O.set_threads(keep)
    assert ok
    xo = O.chol_solve(So, Lo, b)
    del Lo
    A = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))
    ro = np.linalg.norm(A @ xo - b) / np.linalg.norm(b)
    print(f"heat61: oracle factor {tf:.1f} s ({So.flops / tf / 1e9:.1f} GFLOP/s), relres {ro:.2e}", flush=True)
    return pat, v, b, A, So, xo, ro


def relmax(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def test_fp64_matches_oracle_at_full_size(heat61):
    pat, v, b, A, So, xo, ro = heat61
    sym = amp.analyze(pat)
    assert np.array_equal(sym.perm, So.perm) and np.array_equal(sym.Si, So.Si) and sym.flops == So.flops
    num = amp.factorNumeric(sym, v)
    assert num.ok
    x = amp.solveNumeric(num, b)
    rg = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
    err = relmax(x, xo)
    print(f"heat61 fp64: factor {num.info['ms']:.1f} ms, error vs oracle {err:.2e}, relres {rg:.2e} "
          f"(oracle {ro:.2e})", flush=True)
    assert err <= 1e-12
    assert rg <= 10 * ro


def test_mixed_reaches_1e10_at_full_size(heat61):
    pat, v, b, A, So, xo, ro = heat61
    sym = amp.analyze(pat)
    num = amp.NumericFactor(sym, 0, "hip-f32")
    assert num.factor(v)
    x, info = num.refine(v, b, "ir", maxSteps=10, tol=1e-10)
    rg = np.linalg.norm(A @ x - b) / np.linalg.norm(b)
    err = relmax(x, xo)
    print(f"heat61 mixed: FP32 factor {num.info['ms']:.1f} ms, {info['steps']} refinement steps, relres {rg:.2e}, "
          f"error vs oracle {err:.2e}", flush=True)
    assert info["converged"] and rg <= 1e-10
    assert err <= 1e-8
