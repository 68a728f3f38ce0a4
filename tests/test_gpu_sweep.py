"""The learned parametric sweep on the device: the warm-start network's hand-written kernels (amp_mlp_*) against the
float64 reference (tests/mlp_ref.py, itself pinned by finite differences), the PCG driver, the device assembly against the oracle, the batched PCG against the single-system one, and the sweep at the
configuration's size in which every held-out point is solved to |b - A x| / |b| <= 1e-10, judged against the FP64
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


def test_device_assembly_matches_oracle(device):
    """amp_asm: the restricted pattern equals the sweep's; device assembly of given coefficients equals the oracle's
    restatement of assembleCells + restrict bitwise; the device coefficients exp(Phi theta) equal numpy's to 1e-15."""
    import oracle as O

    prob = SW.ParamHeat3D(8)
    prob.device_assembly()
    assert np.array_equal(prob.asm_rowptr, prob.pat["rowptr"]) and np.array_equal(prob.asm_colidx, prob.pat["colidx"])
    thetas = (2 * W.lcg_uniform(71, 3 * 8) - 1).reshape(3, 8)
    K = np.exp(thetas @ prob.Phi.T)  # (3, ne) host coefficients
    dk = device.DeviceArray.from_numpy(K)
    dv = device.DeviceArray((3, prob.pat["nnz"]))
    from ampsolve import _capi as A
    A.check(A.lib().amp_assemble_device(prob.device_assembly(), A.ptr(dk), A.ptr(dv), 3))
    V = dv.numpy()
    for j in range(3):
        pat, vo = O.assemble_p1(prob.nodes, prob.tets, K[j], prob.free)
        assert np.array_equal(pat["colidx"], prob.pat["colidx"])
        assert np.array_equal(V[j], vo), j  # same products and sums in the same order
        assert np.abs(V[j] - prob.values(thetas[j])).max() <= 4e-15 * np.abs(vo).max()
    Vd = prob.device_values(thetas).numpy()  # coefficients computed on the device too
    assert np.abs(Vd - V).max() <= 1e-15 * np.abs(V).max() * 4


def test_batched_pcg_matches_single_system_pcg(device):
    """amp_pcg_batch_device runs each system's CG as amp_pcg_device does (same preconditioner, same iterate): equal
    iteration counts and solutions to round-off."""
    prob = SW.ParamHeat3D(10)
    sym = SW.analyze(prob.pat)
    num = SW.NumericFactor(sym, 0, "hip-f32")
    assert num.factor(prob.device_values(np.zeros(8))[0])
    thetas = 0.5 * (2 * W.lcg_uniform(72, 4 * 8) - 1).reshape(4, 8)
    dv = prob.device_values(thetas)
    db = device.DeviceArray.from_numpy(prob.b)
    X = device.DeviceArray.zeros((4, prob.pat["n"]))
    infos = SW.pcg_batch(num, dv, db, X, 4, tol=1e-10, maxIter=100)
    for j in range(4):
        x1 = device.DeviceArray.zeros((prob.pat["n"],))
        one = SW.pcg(num, dv[j], db, x1, tol=1e-10, maxIter=100)
        assert infos[j]["converged"] and one["converged"]
        assert abs(infos[j]["iters"] - one["iters"]) <= 1
        assert np.abs(X[j].numpy() - x1.numpy()).max() <= 1e-9 * np.abs(x1.numpy()).max()


@pytest.mark.timeout(1200)
def test_learned_sweep_full_config(device):
    """The learned sweep at the configuration's size: 30^3 hexes x 6 tets (29,791 nodes), 2,000 training solves,
    500 held-out points, each preconditioned by the cached FP32 factor nearest in conductivity spread (PCG from the
    network's warm start, device assembly); groups of 8 sharing one factor and the Euclidean-nearest factor are
    measured on a subset for comparison. Every held-out point reaches |b - A x| / |b| <= 1e-10 and is judged against the FP64 direct solve; the
    network's gain over a zero initial guess is measured on the same groups and factors; the time per point is compared
    with the fastest refactoring path (captured FP32 factor + FP64 refinement)."""
    import time

    print("sweep: full config", flush=True)
    sw = SW.LearnedSweep(ncell=30, hidden=256)
    train = (2 * W.lcg_uniform(41, 2000 * 8) - 1).reshape(2000, 8)
    held = (2 * W.lcg_uniform(42, 500 * 8) - 1).reshape(500, 8)
    st = sw.fit(train, epochs=150, batch=64, lr=3e-3)
    print(f"sweep: {len(train)} training solves {st['data_s']:.1f} s, network training {st['train_s']:.1f} s, loss "
          f"{st['loss'][0]:.3f} -> {st['loss'][-1]:.4f}", flush=True)
    assert st["loss"][-1] < 0.1 * st["loss"][0]
    order = sw.order_points(held)
    groups = [order[i:i + sw.group] for i in range(0, len(order), sw.group)]
    device.synchronize()
    t0 = time.perf_counter()
    results = []
    for gi in groups:
        X, infos, a = sw.solve_group(held[gi])
        results.append((gi, X, infos, a))
    device.synchronize()
    t_learned = (time.perf_counter() - t0) / len(held)
    iters = [i["iters"] for _, _, infos, _ in results for i in infos]
    fallbacks = sum(i["fallback"] for _, _, infos, _ in results for i in infos)
    # judge every point: residual and difference from the FP64 direct solve
    worst_res = worst_err = 0.0
    for gi, X, infos, a in results:
        dv = sw.prob.device_values(held[gi])
        for k, p in enumerate(gi):
            v = dv[k].numpy()
            Am = sp.csr_matrix((v, sw.prob.pat["colidx"], sw.prob.pat["rowptr"]), shape=(sw.prob.pat["n"],) * 2)
            x = X[k].numpy()
            worst_res = max(worst_res, np.linalg.norm(Am @ x - sw.prob.b) / np.linalg.norm(sw.prob.b))
            xd, _ = sw.solve_direct(held[p], dv[k])
            worst_err = max(worst_err, relnorm(x, xd.numpy()))
    # the same points and factors from a zero guess (first 64)
    zero_iters, net_iters = [], []
    for gi, X, infos, a in results[:64]:
        Z = device.DeviceArray.zeros((len(gi), sw.prob.pat["n"]))
        _, zi, _ = sw.solve_group(held[gi], x0=Z, factor=a)
        zero_iters += [i["iters"] for i in zi]
        net_iters += [i["iters"] for i in infos]
    # groups of 8 consecutive points sharing one factor (batched PCG), on the first 64 points of the same order
    t0 = time.perf_counter()
    g8 = []
    for i in range(0, 64, 8):
        _, infos, _ = sw.solve_group(held[order[i:i + 8]])
        g8 += infos
    device.synchronize()
    t_g8 = (time.perf_counter() - t0) / 64
    print(f"sweep: groups of 8 sharing a factor: {1e3 * t_g8:.2f} ms per point, "
          f"{np.mean([i['iters'] for i in g8]):.1f} iterations, {sum(i['fallback'] for i in g8)} fallbacks of 64",
          flush=True)
    # the Euclidean-nearest factor instead of the spread-nearest one, same points (first 64)
    eu = []
    for p in order[:64]:
        _, infos, _ = sw.solve_group(held[[p]], factor=sw.nearest(held[p]))
        eu += infos
    print(f"sweep: Euclidean-nearest factor: {np.mean([i['iters'] for i in eu]):.1f} iterations vs spread-nearest "
          f"{np.mean(iters[:64]):.1f} on the same points", flush=True)
    # the fastest refactoring path, per point
    run, keep = sw.refactor_graph()
    dvals = sw.prob.device_values(held[:64])
    device.synchronize()
    t0 = time.perf_counter()
    for k in range(64):
        run(dvals[k])
    device.synchronize()
    t_refactor = (time.perf_counter() - t0) / 64
    print(f"sweep: learned {1e3 * t_learned:.2f} ms per point ({np.mean(iters):.1f} PCG iterations, {fallbacks} "
          f"fallbacks), refactor (captured FP32 factor + IR) {1e3 * t_refactor:.2f} ms per point: "
          f"{t_refactor / t_learned:.2f}x; network warm start {np.mean(net_iters):.1f} iterations vs zero guess "
          f"{np.mean(zero_iters):.1f}; worst relres {worst_res:.2e}, worst difference from FP64 direct "
          f"{worst_err:.2e}", flush=True)
    assert worst_res <= 1e-10
    assert worst_err < 1e-8
    assert np.mean(net_iters) < np.mean(zero_iters)
