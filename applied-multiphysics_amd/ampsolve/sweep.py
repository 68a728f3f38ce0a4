"""Learned warm starts for parametric sweeps (no counterpart in the reference, whose sweeps refactor at every point).

The parametric system is 3D P1 heat on the unit cube (ncell^3 hexes x 6 Kuhn tets, Dirichlet z = 0) with conductivity
k(c) = exp(sum_i theta_i phi_i(c)) at the tet centroid c, phi_i separable cosine modes. The sweep:
  1. solves training points theta with the direct FP64 factor (the data);
  2. trains the warm-start network theta -> solution on all nodes (WarmStartNet: hand-written device kernels,
     amp_mlp_*);
  3. keeps the FP32 factor of every solved training point resident in HBM (workspace released, solve storage only);
  4. for a new theta: x0 = the network's prediction, then preconditioned CG (amp_pcg_device) with the cached factor
     of the nearest solved point to |b - A x| / |b| <= tol within a fixed iteration budget, else a fresh factor (the
     fallback).
The training solves are the direct solver's mixed path (FP32 factor + FP64 iterative refinement to 1e-12), whose
factors are the ones cached.
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


class LearnedSweep:
    """The sweep of the module docstring on one device."""

    def __init__(self, ncell=30, hidden=256, maxBatch=512):
        self.prob = ParamHeat3D(ncell)
        self.sym = analyze(self.prob.pat)
        self.db = DeviceArray.from_numpy(self.prob.b)
        self.net = WarmStartNet(MODES.shape[0], hidden, self.prob.nnodes, maxBatch)
        self.direct = NumericFactor(self.sym, 0, "hip")

    def device_values(self, theta):
        return DeviceArray.from_numpy(self.prob.values(theta))

    def solve_direct(self, theta):
        """Fresh FP64 factor + solve (the refactoring method): (x, device ms of factor + solve)."""
        dv = self.device_values(theta)
        assert self.direct.factor(dv)
        fms = self.direct.info["ms"]
        x = self.direct.solve(self.db)
        return x, fms + self.direct.kernel_ms("solve_total")[0]

    def fit(self, thetas, epochs=150, batch=64, lr=3e-3, seed=5):
        """Training solves (FP32 factor + FP64 refinement, factors cached), network training. Returns statistics."""
        t0 = time.perf_counter()
        self.cache_theta = np.asarray(thetas, np.float64)
        self.cache, xs = [], []
        for th in thetas:
            dv = self.device_values(th)
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

    def nearest(self, theta):
        return int(np.argmin(np.linalg.norm(self.cache_theta - np.asarray(theta), axis=1)))

    def solve(self, theta, x0, tol=1e-10, maxIter=100):
        """PCG from x0 preconditioned by the cached factor of the nearest solved point; a fresh FP64 factor when it
        misses tol within maxIter. Returns (x, info)."""
        dv = self.device_values(theta)
        a = self.nearest(theta)
        x = x0.clone()
        info = pcg(self.cache[a], dv, self.db, x, tol, maxIter)
        info["nearest"] = a
        info["fallback"] = False
        if not info["converged"]:
            x, ms = self.solve_direct(theta)
            info["fallback"] = True
            info["ms"] += ms
        return x, info
