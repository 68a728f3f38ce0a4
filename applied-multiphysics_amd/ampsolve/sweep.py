"""Learned warm starts for parametric sweeps (no counterpart in the reference, whose sweeps refactor at every point).

The parametric system is 3D P1 heat on the unit cube (ncell^3 hexes x 6 Kuhn tets, Dirichlet z = 0) with conductivity
k(c) = exp(sum_i theta_i phi_i(c)) at the tet centroid c, phi_i separable cosine modes. The sweep:
  1. assembles the values of every parameter point on the device (amp_param_coefficients_device, then
     amp_assemble_device: assembleCells / elementSlots / restrict of the P1 operator);
  2. solves the training points with the direct solver's mixed path (FP32 factor + FP64 iterative refinement to 1e-12)
     and keeps every training factor resident in HBM (workspace released, solve storage only);
  3. trains the warm-start network theta -> solution on all nodes (WarmStartNet: hand-written device kernels,
     amp_mlp_*);
  4. solves new points one by one (group = 1) or in groups sharing a factor: the network's predictions are the
     initial guesses, a cached factor preconditions CG (amp_pcg_batch_device: one multi-right-hand-side solve per
     iteration for all the group's systems) to |b - A x| / |b| <= tol within a fixed iteration budget, and a point
     that misses it gets a fresh factor (the fallback).
The cached factor of a point (or group) is the one whose conductivity field is closest in the sense that bounds the
preconditioned condition number: for a factor of A(theta_c) preconditioning A(theta), the eigenvalues lie in
[exp(min_e d_e), exp(max_e d_e)] with d_e = Phi_e (theta - theta_c), so the factor minimising the spread
max_e d_e - min_e d_e (the largest over a group's points) is taken, among the Euclidean-nearest candidates of each
point, on a sample of the modes' values. In eight parameters the training points are far apart, so a factor shared by
a group preconditions each member much worse than its own best one: one point per solve is the default.
"""
import ctypes as C
import time

import numpy as np
import scipy.sparse as sp

from . import _capi as A
from . import workloads as W
from .device import DeviceArray, synchronize
from .sparse import NumericFactor, analyze

MODES = np.array([(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1), (2, 0, 0)])


def mode_values(points):
    """phi_i at the given points (..., 3) -> (..., 8)."""
    return np.prod(np.cos(np.pi * MODES * points[..., None, :]), axis=-1)


class ParamHeat3D:
    """Pattern, right-hand side and value map of the parametric heat system: values(theta) = Cmat @ exp(Phi theta)."""

    def __init__(self, ncell=30, rhs_seed=7):
        nodes, tets = W.box_tet_mesh(ncell, ncell, ncell)
        self.ncell, self.nnodes = ncell, len(nodes)
        kc = W._kuhn_corners().astype(np.float64) / ncell
        Kref = W.p1_element_stiffness(kc)  # (6, 4, 4): every hex has the same six shapes
        cen = nodes[tets].mean(axis=1)
        self.Phi = np.prod(np.cos(np.pi * MODES[None, :, :] * cen[:, None, :]), axis=2)  # (ne, 8)
        free = nodes[:, 2] > 1e-12
        self.free = free
        ridx = np.full(len(nodes), -1, np.int64)
        ridx[free] = np.arange(free.sum())
        n = int(free.sum())
        ne = len(tets)
        a = ridx[np.repeat(tets, 4, axis=1)].ravel()  # (ne * 16): row of (e, i, j)
        b = ridx[np.tile(tets, (1, 4))].ravel()       # column
        coef = Kref[np.arange(ne) % 6].reshape(ne * 16)
        el = np.repeat(np.arange(ne), 16)
        keep = (a >= 0) & (b >= 0)
        a, b, coef, el = a[keep], b[keep], coef[keep], el[keep]
        keys = a * n + b
        ukeys, slot = np.unique(keys, return_inverse=True)
        rows, cols = ukeys // n, ukeys % n
        rowptr = np.zeros(n + 1, np.int64)
        np.cumsum(np.bincount(rows, minlength=n), out=rowptr[1:])
        self.pat = {"n": n, "rowptr": rowptr.astype(np.int32), "colidx": cols.astype(np.int32), "nnz": len(ukeys)}
        self.Cmat = sp.csr_matrix((coef, (slot, el)), shape=(len(ukeys), ne))
        self.b = 2.0 * W.lcg_uniform(rhs_seed, n) - 1.0
        self.nodes, self.tets = nodes, tets
        self._asm = None

    def device_assembly(self):
        """The device assembly of this system (amp_asm, created on first use): restricted pattern, modes set."""
        if self._asm is None:
            h = C.c_void_p()
            nodes = np.ascontiguousarray(self.nodes, np.float64)
            tets = np.ascontiguousarray(self.tets, np.int32)
            free = np.ascontiguousarray(self.free, np.uint8)
            A.check(A.lib().amp_asm_create(3, len(nodes), nodes.ctypes.data, len(tets), tets.ctypes.data,
                                           free.ctypes.data, C.byref(h)))
            self._asm = h
            phi = np.ascontiguousarray(self.Phi, np.float64)
            A.check(A.lib().amp_asm_set_modes(h, phi.shape[1], phi.ctypes.data))
            n, nnz = C.c_int32(), C.c_int64()
            rp, ci = C.POINTER(C.c_int32)(), C.POINTER(C.c_int32)()
            A.check(A.lib().amp_asm_get_pattern(h, C.byref(n), C.byref(nnz), C.byref(rp), C.byref(ci)))
            self.asm_rowptr = np.ctypeslib.as_array(rp, shape=(n.value + 1,)).copy()
            self.asm_colidx = np.ctypeslib.as_array(ci, shape=(nnz.value,)).copy()
        return self._asm

    def device_values(self, thetas):
        """Values of the points' matrices on the device, (len(thetas), nnz): k_e = exp(Phi_e theta) and the assembly,
        both on the device."""
        h = self.device_assembly()
        th = np.ascontiguousarray(np.atleast_2d(thetas), np.float64)
        nb = th.shape[0]
        dth = DeviceArray.from_numpy(th)
        dk = DeviceArray((nb, len(self.tets)))
        A.check(A.lib().amp_param_coefficients_device(h, A.ptr(dth), A.ptr(dk), nb))
        dv = DeviceArray((nb, self.pat["nnz"]))
        A.check(A.lib().amp_assemble_device(h, A.ptr(dk), A.ptr(dv), nb))
        return dv

    def __del__(self):
        h = getattr(self, "_asm", None)
        if h is not None and h.value:
            try:
                A.lib().amp_asm_free(h)
            except Exception:
                pass

    def values(self, theta):
        return self.Cmat @ np.exp(self.Phi @ np.asarray(theta, np.float64))

    def to_nodes(self, x):
        """Free-DOF vectors (..., n) -> all nodes (..., nnodes), zero on the Dirichlet face."""
        out = np.zeros(x.shape[:-1] + (self.nnodes,), dtype=x.dtype)
        out[..., self.free] = x
        return out


class WarmStartNet:
    """FP32 network nIn -> H (tanh) -> H (tanh) -> nOut trained and evaluated on the device by amp_mlp_*."""

    def __init__(self, nIn, H, nOut, maxBatch, seed=1):
        self.dims = (nIn, H, nOut)
        h = C.c_void_p()
        A.check(A.lib().amp_mlp_create(nIn, H, nOut, maxBatch, C.byref(h)))
        self._h = h
        self.nparams = int(A.lib().amp_mlp_param_count(h))
        # Glorot-uniform weights from the LCG, zero biases
        parts, s = [], seed
        for rows, cols in ((H, nIn), (H, H), (nOut, H)):
            lim = np.sqrt(6.0 / (rows + cols))
            parts += [lim * (2 * W.lcg_uniform(s, rows * cols) - 1), np.zeros(rows)]
            s += 1
        self.set_params(np.concatenate(parts).astype(np.float32))

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        A.check(A.lib().amp_mlp_set_params(self._h, p, len(p)))

    def get_params(self):
        p = np.empty(self.nparams, np.float32)
        A.check(A.lib().amp_mlp_get_params(self._h, p, len(p)))
        return p

    def get_grads(self):
        g = np.empty(self.nparams, np.float32)
        A.check(A.lib().amp_mlp_get_grads(self._h, g, len(g)))
        return g

    def forward(self, X, Y):
        """X (B x nIn), Y (B x nOut): float32 device tensors."""
        A.check(A.lib().amp_mlp_forward(self._h, A.ptr(X), int(X.shape[0]), A.ptr(Y)))
        return Y

    def gradient(self, X, T):
        loss = C.c_double()
        A.check(A.lib().amp_mlp_gradient(self._h, A.ptr(X), A.ptr(T), int(X.shape[0]), C.byref(loss)))
        return loss.value

    def adam(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        A.check(A.lib().amp_mlp_adam(self._h, float(lr), float(b1), float(b2), float(eps)))

    def train(self, X, T, epochs, batch, lr, seed=3, decay=0.3):
        """Minibatch Adam over device arrays X (N x nIn), T (N x nOut) whose rows the caller has put in random order:
        each epoch visits the contiguous minibatches in an order of its own; the learning rate steps down by `decay`
        at each third of the epochs. Returns the mean loss of each epoch."""
        N = X.shape[0]
        starts = np.arange(0, N, batch)
        hist = []
        for ep in range(epochs):
            rate = lr * decay ** (3 * ep // epochs)
            tot = 0.0
            for s in starts[np.argsort(W.lcg_uniform(seed + ep, len(starts)))]:
                e = min(s + batch, N)
                tot += self.gradient(X[s:e], T[s:e]) * (e - s)
                self.adam(rate)
            hist.append(tot / N)
        return hist

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                A.lib().amp_mlp_free(h)
            except Exception:
                pass


def pcg(num, values, b, x, tol=1e-10, maxIter=50):
    """amp_pcg_device: CG on the matrix `values` preconditioned by the factor `num`, from the guess in x (device)."""
    o = A.PcgOpts()
    o.maxIter, o.tol = int(maxIter), float(tol)
    info = A.PcgInfo()
    A.check(A.lib().amp_pcg_device(num._h, A.ptr(values), A.ptr(b), A.ptr(x), C.byref(o), C.byref(info)))
    return {"iters": info.iters, "converged": bool(info.converged), "relResidual": info.relResidual,
            "initialResidual": info.initialResidual, "ms": info.ms}


def pcg_batch(num, values, b, x, nb, tol=1e-10, maxIter=50, shared_b=True):
    """amp_pcg_batch_device: nb systems (values nb x nnz, x nb x n, device) preconditioned together by `num`."""
    o = A.PcgOpts()
    o.maxIter, o.tol = int(maxIter), float(tol)
    infos = (A.PcgInfo * nb)()
    stride = 0 if shared_b else num.sym.n
    A.check(A.lib().amp_pcg_batch_device(num._h, A.ptr(values), A.ptr(b), stride, A.ptr(x), int(nb), C.byref(o),
                                         infos))
    return [{"iters": i.iters, "converged": bool(i.converged), "relResidual": i.relResidual,
             "initialResidual": i.initialResidual, "ms": i.ms} for i in infos]


class LearnedSweep:
    """The sweep of the module docstring on one device."""

    def __init__(self, ncell=30, hidden=256, maxBatch=512, group=1, candidates=16):
        self.prob = ParamHeat3D(ncell)
        self.sym = analyze(self.prob.pat)
        self.db = DeviceArray.from_numpy(self.prob.b)
        self.net = WarmStartNet(MODES.shape[0], hidden, self.prob.nnodes, maxBatch)
        self.direct = NumericFactor(self.sym, 0, "hip")
        self.group, self.candidates = group, candidates
        g = np.linspace(0.0, 1.0, 9)  # sample of the modes' values for the spread of d_e (corners and edges included)
        self.phi_sample = mode_values(np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3))

    def device_values(self, theta):
        return self.prob.device_values(theta)[0]

    def solve_direct(self, theta, dv=None):
        """Fresh FP64 factor + solve (judging the learned solutions): (x, device ms of factor + solve)."""
        dv = self.device_values(theta) if dv is None else dv
        assert self.direct.factor(dv)
        fms = self.direct.info["ms"]
        x = self.direct.solve(self.db)
        return x, fms + self.direct.kernel_ms("solve_total")[0]

    def fit(self, thetas, epochs=150, batch=64, lr=3e-3, seed=5, chunk=64):
        """Training solves (FP32 factor + FP64 refinement, factors cached), network training. Returns statistics."""
        t0 = time.perf_counter()
        thetas = np.asarray(thetas, np.float64)
        self.cache_theta = thetas
        self.cache, xs = [], []
        for c0 in range(0, len(thetas), chunk):
            dvals = self.prob.device_values(thetas[c0:c0 + chunk])
            for i in range(dvals.shape[0]):
                dv = dvals[i]
                num = NumericFactor(self.sym, 0, "hip-f32")
                assert num.factor(dv)
                x, info = num.refine(dv, self.db, "ir", maxSteps=20, tol=1e-12)
                assert info["converged"], info
                num.release_workspace()
                self.cache.append(num)
                xs.append(x.numpy())
        synchronize()
        t_data = time.perf_counter() - t0
        full = self.prob.to_nodes(np.stack(xs))  # (N, nnodes) FP64
        # targets scaled per node (mean and spread over the training points; Dirichlet nodes stay at zero)
        self.mu = full.mean(0)
        sd = (full - self.mu).std(0)
        self.sigma = np.maximum(sd, 1e-3 * sd.max()) if sd.max() > 0 else np.ones_like(sd)
        order = np.argsort(W.lcg_uniform(seed, len(thetas)))  # rows in random order once; minibatches are slices
        T = DeviceArray.from_numpy(((full[order] - self.mu) / self.sigma).astype(np.float32))
        Xin = DeviceArray.from_numpy(np.asarray(thetas, np.float32)[order])
        t0 = time.perf_counter()
        hist = self.net.train(Xin, T, epochs, batch, lr)
        synchronize()
        t_train = time.perf_counter() - t0
        return {"data_s": t_data, "train_s": t_train, "loss": hist}

    def predict(self, thetas):
        """Network warm starts for a batch of points: (B, n) FP64 device array on the free DOFs."""
        Xin = DeviceArray.from_numpy(np.asarray(thetas, np.float32))
        Y = DeviceArray((len(thetas), self.prob.nnodes), np.float32)
        self.net.forward(Xin, Y)
        full = Y.numpy().astype(np.float64) * self.sigma + self.mu
        return DeviceArray.from_numpy(full[:, self.prob.free])

    def spread(self, thetas, idx):
        """Largest spread max_e d_e - min_e d_e over the points `thetas` (g x 8) for each cached factor in idx."""
        d = (thetas[:, None, :] - self.cache_theta[idx][None, :, :]) @ self.phi_sample.T  # (g, c, samples)
        return (d.max(axis=2) - d.min(axis=2)).max(axis=0)

    def nearest(self, theta):
        return int(np.argmin(np.linalg.norm(self.cache_theta - np.asarray(theta), axis=1)))

    def choose_factor(self, thetas):
        """The cached factor for a group of points: the smallest largest spread among each point's Euclidean-nearest
        candidates."""
        cand = set()
        for th in thetas:
            cand.update(np.argsort(np.linalg.norm(self.cache_theta - th, axis=1))[:self.candidates].tolist())
        cand = np.array(sorted(cand))
        return int(cand[np.argmin(self.spread(np.atleast_2d(thetas), cand))])

    def order_points(self, thetas):
        """A sweep order in which consecutive points are close (greedy nearest neighbour from the first point), so
        that the groups of consecutive points share a good factor."""
        thetas = np.asarray(thetas)
        left = list(range(1, len(thetas)))
        order = [0]
        while left:
            last = thetas[order[-1]]
            j = int(np.argmin(np.linalg.norm(thetas[left] - last, axis=1)))
            order.append(left.pop(j))
        return np.array(order)

    def solve_group(self, thetas, x0=None, tol=1e-10, maxIter=100, factor=None):
        """Batched PCG for a group of points (their values assembled on the device, one cached factor), from the
        network's warm starts (or from x0: a (g, n) device array, zeros for a zero guess); a point that misses tol
        within maxIter gets a fresh factor. Returns (x (g, n) device array, infos, chosen factor)."""
        thetas = np.atleast_2d(np.asarray(thetas, np.float64))
        g = len(thetas)
        a = self.choose_factor(thetas) if factor is None else factor
        dv = self.prob.device_values(thetas)
        X = self.predict(thetas) if x0 is None else x0.clone()
        infos = pcg_batch(self.cache[a], dv, self.db, X, g, tol, maxIter)
        for i, info in enumerate(infos):
            info["factor"] = a
            info["fallback"] = False
            if not info["converged"]:
                x, ms = self.solve_direct(thetas[i], dv[i])
                X[i].copy_(x)
                info["fallback"] = True
                info["ms"] += ms
        return X, infos, a

    def refactor_graph(self, ir_steps=2):
        """The fastest refactoring path for comparison: a captured FP32 factor + solve + FP64 refinement sequence on one
        values buffer (amp_graph); returns (run(dv) -> x device array, the handles to keep)."""
        num = NumericFactor(self.sym, 0, "hip-f32")
        buf = DeviceArray((self.prob.pat["nnz"],))
        x = DeviceArray.zeros((self.prob.pat["n"],))
        buf.copy_(self.device_values(np.zeros(MODES.shape[0])))
        g = num.graph(buf, self.db, x, 1, ir_steps)

        def run(dv):
            buf.copy_(dv)
            g.launch()
            st = g.sync()
            assert st["ok"]
            return x

        return run, (num, g, buf, x)
