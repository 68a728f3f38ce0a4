"""The learned parametric sweep on the device: the warm-start network's hand-written kernels (amp_mlp_*) against the
float64 reference (tests/mlp_ref.py, itself pinned by finite differences), the PCG driver, and a reduced sweep
(12^3-cell cube) in which every held-out point is solved to |b - A x| / |b| <= 1e-10, judged against the FP64
direct solve, with the network's warm start compared to the same method without it (zero initial guess)."""
import numpy as np
import pytest
import scipy.sparse as sp

import mlp_ref as R
from ampsolve import sweep as SW
from ampsolve import workloads as W

pytestmark = pytest.mark.gpu


def relnorm(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.mark.parametrize("dims,B", [((8, 32, 300), 37), ((8, 64, 3000), 64)])
def test_network_kernels_match_reference(device, dims, B):
    """Forward, loss, gradient (the 3000-output case takes the depth-split product) and one Adam step."""
    nIn, H, nOut = dims
    net = SW.WarmStartNet(nIn, H, nOut, 64, seed=5)
    th = net.get_params().astype(np.float64)
    # nonzero biases too
    th[nIn * H:nIn * H + H] = 0.1 * (2 * W.lcg_uniform(21, H) - 1)
    th[-nOut:] = 0.1 * (2 * W.lcg_uniform(22, nOut) - 1)
    net.set_params(th.astype(np.float32))
    th = net.get_params().astype(np.float64)
    X = (2 * W.lcg_uniform(23, B * nIn) - 1).reshape(B, nIn).astype(np.float32)
    T = W.lcg_normal(24, B * nOut).reshape(B, nOut).astype(np.float32)
    dX, dT = device.DeviceArray.from_numpy(X), device.DeviceArray.from_numpy(T)
    Y = device.DeviceArray((B, nOut), np.float32)
    net.forward(dX, Y)
    _, _, yr = R.forward(th, X, nIn, H, nOut)
    assert relnorm(Y.numpy(), yr) < 1e-5
    loss = net.gradient(dX, dT)
    lr_, gr = R.loss_and_grad(th, X, T, nIn, H, nOut)
    assert abs(loss - lr_) / lr_ < 1e-5
    g = net.get_grads().astype(np.float64)
    o = 0
    for name, shp in R.sizes(nIn, H, nOut):
        k = int(np.prod(shp))
        assert relnorm(g[o:o + k], gr[o:o + k]) < 1e-4, name
        o += k
    net.adam(1e-3)
    th1, _, _ = R.adam(th, g, np.zeros_like(th), np.zeros_like(th), 1, 1e-3)
    assert np.abs(net.get_params() - th1).max() < 1e-6


def test_pcg_with_exact_preconditioner_converges_in_one_step(device):
    """With the FP64 factor of the same matrix as preconditioner CG stops after one iteration from a zero guess."""
    prob = SW.ParamHeat3D(8)
    sym = SW.analyze(prob.pat)
    th = 2 * W.lcg_uniform(31, 8) - 1
    dv = device.DeviceArray.from_numpy(prob.values(th))
    num = SW.NumericFactor(sym, 0, "hip")
    assert num.factor(dv)
    db = device.DeviceArray.from_numpy(prob.b)
    x = device.DeviceArray.zeros(db.shape)
    info = SW.pcg(num, dv, db, x, tol=1e-10, maxIter=10)
    assert info["converged"] and info["iters"] == 1 and info["initialResidual"] == 1.0


def test_reduced_learned_sweep(device):
    print("sweep: reduced config", flush=True)
    sw = SW.LearnedSweep(ncell=12, hidden=128)
    train = (2 * W.lcg_uniform(41, 300 * 8) - 1).reshape(300, 8)
    held = (2 * W.lcg_uniform(42, 24 * 8) - 1).reshape(24, 8)
    st = sw.fit(train, epochs=150, batch=32, lr=3e-3)
    print(f"sweep: training loss {st['loss'][0]:.3f} -> {st['loss'][-1]:.4f} in {st['train_s']:.2f} s", flush=True)
    assert st["loss"][-1] < 0.1 * st["loss"][0]  # the network learned
    X0 = sw.predict(held)
    it_net, it_zero, err = [], [], []
    for i, th in enumerate(held):
        x, info = sw.solve(th, X0[i])
        v = sw.prob.values(th)
        A = sp.csr_matrix((v, sw.prob.pat["colidx"], sw.prob.pat["rowptr"]), shape=(sw.prob.pat["n"],) * 2)
        xh = x.numpy()
        assert np.linalg.norm(A @ xh - sw.prob.b) / np.linalg.norm(sw.prob.b) <= 1e-10
        xd, _ = sw.solve_direct(th)
        err.append(relnorm(xh, xd.numpy()))  # judged against the FP64 direct solve
        if not info["fallback"]:
            it_net.append(info["iters"])
            z = device.DeviceArray.zeros(X0[i].shape)
            zi = SW.pcg(sw.cache[info["nearest"]], device.DeviceArray.from_numpy(v), sw.db, z, maxIter=100)
            it_zero.append(zi["iters"] if zi["converged"] else 100)
    print(f"sweep: PCG iterations with the network's warm start {np.mean(it_net):.2f}, "
          f"from zero {np.mean(it_zero):.2f} over {len(it_net)} points", flush=True)
    print(f"sweep: largest relative difference from the FP64 direct solution {max(err):.2e}", flush=True)
    assert max(err) < 1e-6  # both at |b - A x| / |b| <= 1e-10; the difference is bounded by the conditioning
    assert len(it_net) >= len(held) // 2
    assert np.mean(it_net) < np.mean(it_zero)
