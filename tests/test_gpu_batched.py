"""The batched configuration on the device: 512 independent 2D P1 heat systems (grid side U{20..120} cells,
log-uniform conductivity, Dirichlet edges, 2.8M DOFs in all) with 8 normal right-hand sides each, assembled into one
block-diagonal pattern, so one analysis, one plan and one multi-right-hand-side solve cover the whole batch.

Pins: the FP64 factor and solve of the union against the oracle's (the same forest); the captured FP32 factor + FP64
refinement sequence (amp_graph_*, on plans autotuned for the batch) reaches |b - A x| / |b| <= 1e-10 for every system
and right-hand side, agrees with the FP64 direct solution, and gives the exact scaling x / 2 when replayed on 2 A."""
import time

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
import ampsolve as amp
from ampsolve import workloads as W

pytestmark = pytest.mark.gpu

NRHS = 8


@pytest.fixture(scope="module")
def batch():
    print("batched: generating 512 systems", flush=True)
    pat, v, B, sides, offs = W.batch_heat2d()
    A = sp.csr_matrix((v, pat["colidx"], pat["rowptr"]), shape=(pat["n"], pat["n"]))
    sym = amp.analyze(pat)
    So = O.Sym(pat)
    ok, Lo = O.chol_factor(So, v)
    assert ok
    Xo = np.stack([O.chol_solve(So, Lo, B[k]) for k in range(NRHS)])
    return pat, v, B, offs, A, sym, Xo


def per_system(R, offs):
    """Per-system 2-norms of the rows of R (q x n)."""
    return np.sqrt(np.add.reduceat(R ** 2, offs[:-1], axis=1))


def test_union_fp64_matches_oracle(batch):
    print("batched fp64", flush=True)
    pat, v, B, offs, A, sym, Xo = batch
    So = O.Sym(pat)
    assert np.array_equal(sym.perm, So.perm)
    num = amp.factorNumeric(sym, v)
    assert num.ok
    X = num.solve(B.ravel(), nrhs=NRHS).reshape(NRHS, -1)
    err = per_system(X - Xo, offs) / per_system(Xo, offs)
    assert err.max() < 1e-12


def test_captured_mixed_sequence(batch, device):
    print("batched graph", flush=True)
    pat, v, B, offs, A, sym, Xo = batch
    dv = device.DeviceArray.from_numpy(v)
    tune = sym.autotune(dv, kind="hip-f32")  # shape thresholds for this batch, timed on the device
    print(f"batched autotune: {tune}", flush=True)
    num = amp.NumericFactor(sym, 0, "hip-f32")
    dB = device.DeviceArray.from_numpy(B.ravel())
    dX = device.DeviceArray.zeros(dB.shape)
    g = num.graph(dv, dB, dX, NRHS, 3)
    times = []
    for _ in range(3):
        t0 = time.perf_counter()
        g.launch()
        st = g.sync()
        times.append(time.perf_counter() - t0)
        assert st["ok"]
    print(f"batched graph replay: device {st['ms']:.2f} ms, wall {1e3 * min(times):.2f} ms", flush=True)
    X = dX.numpy().reshape(NRHS, -1)
    rel = per_system((A @ X.T).T - B, offs) / per_system(B, offs)
    assert rel.max() <= 1e-10, rel.max()
    err = per_system(X - Xo, offs) / per_system(Xo, offs)
    assert err.max() < 1e-8
    # new values in place: 2 A gives x / 2 from the same graph
    dv.upload(2.0 * v)
    g.launch()
    assert g.sync()["ok"]
    X2 = dX.numpy().reshape(NRHS, -1)
    assert np.abs(2 * X2 - X).max() / np.abs(X).max() < 1e-9
