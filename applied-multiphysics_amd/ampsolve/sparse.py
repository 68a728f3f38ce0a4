"""Host API of the device solver, mirroring the reference's sparse direct solver interface.

Names, argument meaning and error behaviour follow packages/core/src/la/sparse.mjs and ldlt.mjs of
jwt625/applied-multiphysics (declarations in sparse.d.mts:28-57 and ldlt.d.mts:40-59):
  analyze(pat, opts) -> Symbolic                        sparse.mjs:666-902
  factorNumeric(sym, values, num, pat, factor=...)      sparse.mjs:941-1125
  solveNumeric(num, b) / solveNumericMulti(num, bs)     sparse.mjs:1261-1310
  SparseCholesky(pat, perm, sym=..., factor=...)        sparse.mjs:1312-1372
  SupernodalLDLT(sym, factor=...)                       ldlt.mjs:417-645
  setDefaultRealFactor(kind)                            front-real.mjs:196-200
The factor kinds are 'hip' (FP64 on the matrix cores), 'hip-f32' (FP32 factor; use with refinement) and 'hip-bf16x3'
(FP32 factor whose bulk trailing updates run as three bf16 products; use with refinement). There is no host fallback:
without libampsolve.so or a GPU every numeric call raises.
A CSR pattern is a dict {'n', 'rowptr', 'colidx', 'nnz'} as the reference's CsrPattern (types.d.mts:5-10).
"""
import ctypes as C
import time

import numpy as np

from . import _capi as A

REAL_FACTOR_KINDS = ("hip", "hip-f32", "hip-bf16x3")
PRECISION_OF_KIND = {"hip": A.AMP_F64, "hip-f32": A.AMP_F32, "hip-bf16x3": A.AMP_F32_BF16X3}
DEFAULT_REAL_FACTOR = "hip"


def setDefaultRealFactor(kind):
    """Validate and set the default factor kind (front-real.mjs:196-200)."""
    global DEFAULT_REAL_FACTOR
    if kind not in REAL_FACTOR_KINDS:
        raise ValueError(f"unknown real factor {kind} ({' | '.join(REAL_FACTOR_KINDS)})")
    DEFAULT_REAL_FACTOR = kind
    return kind


def _resolve_kind(kind):
    k = kind if kind is not None else DEFAULT_REAL_FACTOR
    if k not in REAL_FACTOR_KINDS:
        raise ValueError(f"unknown real factor {k} ({' | '.join(REAL_FACTOR_KINDS)})")
    return k


def _pat_arrays(pat):
    rp = np.ascontiguousarray(pat["rowptr"], dtype=np.int32)
    ci = np.ascontiguousarray(pat["colidx"], dtype=np.int32)
    return rp, ci


def patternFingerprint(pat):
    """'n:nnz:hash' with 32-bit FNV-1a over rowptr and colidx (sparse.mjs:908-925)."""
    rp, ci = _pat_arrays(pat)
    buf = C.create_string_buffer(64)
    A.check(A.lib().amp_pattern_fingerprint(int(pat["n"]), rp, ci if len(ci) else np.zeros(1, np.int32), buf))
    return buf.value.decode()


def amdOrder(pat, denseFactor=10.0):
    """Approximate minimum degree ordering, perm new -> old (sparse.mjs:297-592)."""
    n = int(pat["n"])
    rp, ci = _pat_arrays(pat)
    perm = np.zeros(max(n, 1), dtype=np.int32)
    A.check(A.lib().amp_amd_order(n, rp, ci if len(ci) else np.zeros(1, np.int32), float(denseFactor), perm))
    return perm[:n]


def ndOrder(pat, x, y, leaf=48):
    """Nested dissection by recursive coordinate bisection, perm new -> old (sparse.mjs:230-283)."""
    n = int(pat["n"])
    rp, ci = _pat_arrays(pat)
    xs = np.ascontiguousarray(x, dtype=np.float64)
    ys = np.ascontiguousarray(y, dtype=np.float64)
    if len(xs) != n or len(ys) != n:
        raise ValueError("ndOrder: node coordinates x and y must have length n")
    perm = np.zeros(max(n, 1), dtype=np.int32)
    A.check(A.lib().amp_nd_order(n, rp, ci if len(ci) else np.zeros(1, np.int32), xs if n else np.zeros(1),
                                 ys if n else np.zeros(1), int(leaf), perm))
    return perm[:n]


class Symbolic:
    """Symbolic analysis (the object analyze() returns, sparse.mjs:873-901) owning the native handle and plan."""

    _ARRAYS = ("perm", "pinv", "parent", "colCount", "lowerP", "lowerI", "lowerSrc", "superStart", "colSuper",
               "superParent", "childHead", "childNext", "Sp", "Si")

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        info = A.SymInfo()
        A.check(A.lib().amp_sym_get_info(self._h, C.byref(info)))
        self.n, self.nsuper, self.maxFront, self.nlevels = info.n, info.nsuper, info.maxFront, info.nlevels
        self.nnzL, self.nnzA, self.nnzLStored = int(info.nnzL), int(info.nnzA), int(info.nnzLStored)
        self.nnzPattern = int(info.nnzPattern)
        self.flops, self.fill, self.gemmFlops = info.flops, info.fill, info.gemmFlops
        self.bulkBytes, self.bulkLaunches = info.bulkBytes, int(info.bulkLaunches)
        self.patternKey = info.patternKey.decode()
        arr = A.SymArrays()
        A.check(A.lib().amp_sym_get_arrays(self._h, C.byref(arr)))
        n, ns = self.n, self.nsuper
        lengths = {"perm": n, "pinv": n, "parent": n, "colCount": n, "lowerP": n + 1, "lowerI": self.nnzA,
                   "lowerSrc": self.nnzA, "superStart": ns + 1, "colSuper": n, "superParent": ns, "childHead": ns,
                   "childNext": ns, "Sp": ns + 1}
        for name in self._ARRAYS:
            if name == "Si":
                continue
            setattr(self, name, _copy(getattr(arr, name), lengths[name], np.int32))
        self.Si = _copy(arr.Si, int(self.Sp[-1]) if ns else 0, np.int32)
        self.Lxp = _copy(arr.Lxp, ns + 1, np.int64)
        self.timing = {}

    def autotune(self, values, mode=0, kind="hip"):
        """Time the factor and one solve of `values` (a device tensor on the pattern) on the device for candidate
        shape thresholds and keep the fastest plan (amp_sym_autotune). Call before creating numeric factors."""
        t = A.TuneInfo()
        prec = PRECISION_OF_KIND[kind]
        A.check(A.lib().amp_sym_autotune(self._h, A.ptr(values), int(mode), prec, C.byref(t)))
        self._refresh()
        return {f: getattr(t, f) for f, _ in A.TuneInfo._fields_}

    def set_params(self, params):
        """Apply shape thresholds (a dict as autotune returns; flowWideMin defaults to 160, packMin to 512, fusedPanel to 1,
        fusedSolve to 3) and rebuild the plans (amp_sym_set_params)."""
        t = A.TuneInfo()
        for f in ("smallMax", "smallNarrow", "wideLaunch", "superPanel"):
            setattr(t, f, int(params[f]))
        t.flowWideMin = int(params.get("flowWideMin", 160))
        t.packMin = int(params.get("packMin", 512))
        t.fusedPanel = int(params.get("fusedPanel", 1))
        t.fusedSolve = int(params.get("fusedSolve", 3))
        A.check(A.lib().amp_sym_set_params(self._h, C.byref(t)))
        self._refresh()

    def _refresh(self):
        info = A.SymInfo()
        A.check(A.lib().amp_sym_get_info(self._h, C.byref(info)))
        self.gemmFlops = info.gemmFlops
        self.bulkBytes, self.bulkLaunches = info.bulkBytes, int(info.bulkLaunches)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                A.lib().amp_sym_free(h)
            except Exception:
                pass
            self._h = None


def _copy(p, count, dtype):
    if count == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(p, shape=(count,)).astype(dtype, copy=True)


def analyze(pat, opts=None):
    """Symbolic analysis of a full symmetric CSR pattern (sparse.mjs:666-902). opts: {'perm', 'ordering', 'x', 'y',
    'relax'}: perm (new -> old) to start from; else `ordering` 'amd' (default) or 'nd' (node coordinates x and y of
    length n). timing: {ordering_ms, symbolic_ms} (ordering_ms 0 when perm is given)."""
    opts = opts or {}
    n = int(pat["n"])
    perm = opts.get("perm")
    if perm is not None and len(perm) != n:
        raise ValueError(f"analyze: perm has length {len(perm)}, pattern has {n} rows")
    if perm is None and n > 0 and opts.get("ordering") == "nd":
        x, y = opts.get("x"), opts.get("y")
        if x is None or y is None or len(x) != n or len(y) != n:
            raise ValueError("analyze: ordering nd needs node coordinates x and y of length n")
    rp, ci = _pat_arrays(pat)
    t0 = time.perf_counter()
    ordering_ms = 0.0
    if perm is None and n > 0:
        perm = ndOrder(pat, opts["x"], opts["y"]) if opts.get("ordering") == "nd" else amdOrder(pat)
        ordering_ms = 1e3 * (time.perf_counter() - t0)
    t1 = time.perf_counter()
    h = C.c_void_p()
    permArr = None if perm is None else np.ascontiguousarray(perm, dtype=np.int32)
    A.check(A.lib().amp_analyze(n, rp, ci if len(ci) else np.zeros(1, np.int32),
                                None if permArr is None else permArr.ctypes.data_as(C.c_void_p),
                                1 if opts.get("relax", True) else 0, C.byref(h)))
    sym = Symbolic(h.value)
    sym.timing = {"ordering_ms": ordering_ms, "symbolic_ms": 1e3 * (time.perf_counter() - t1)}
    sym.pattern = {"n": n, "rowptr": rp.copy(), "colidx": ci.copy(), "nnz": int(rp[n]) if n else 0}
    return sym


def checkSymbolic(sym, pat):
    """Throw unless `sym` was analysed from `pat` (sparse.mjs:927-930)."""
    if sym is None or sym.n != int(pat["n"]) or sym.patternKey != patternFingerprint(pat):
        raise ValueError("symbolic analysis is for a different sparsity pattern")


class FactorGraph:
    """A captured factor + solve + refinement sequence (amp_graph_*): launch() replays it on the factor's stream,
    sync() waits and returns the factor status of the replay with its device time."""

    def __init__(self, num, values, B, X, nrhs, irSteps):
        self.num = num
        self._keep = (values, B, X)  # the graph reads these device buffers at every replay
        h = C.c_void_p()
        A.check(A.lib().amp_graph_create(num._h, A.ptr(values), A.ptr(B), A.ptr(X), int(nrhs), int(irSteps),
                                         C.byref(h)))
        self._h = h

    def launch(self):
        A.check(A.lib().amp_graph_launch(self._h))

    def sync(self):
        info = A.FactorInfo()
        A.check(A.lib().amp_graph_sync(self._h, C.byref(info)))
        self.num.ok = bool(info.ok)
        return {"ok": bool(info.ok), "failColumn": info.failColumn, "ms": info.ms}

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                A.lib().amp_graph_free(h)
            except Exception:
                pass
            self._h = None


class NumericFactor:
    """A factor resident in HBM (the NumericFactor of types.d.mts:21-26): sym, ok, kernel, info."""

    def __init__(self, sym, mode, kind):
        self.sym = sym
        self.mode = mode
        self.kernel = kind
        self.precision = PRECISION_OF_KIND[kind]
        h = C.c_void_p()
        A.check(A.lib().amp_num_create(sym._h, mode, self.precision, C.byref(h)))
        self._h = h
        self.ok = False
        self.info = None

    def factor(self, values, pivotTol=1e-13):
        nvalues = int(values.numel()) if hasattr(values, "numel") else len(values)
        info = A.FactorInfo()
        if hasattr(values, "data_ptr") and getattr(values, "is_cuda", False):
            A.check(A.lib().amp_factor_device(self._h, A.ptr(values), nvalues, float(pivotTol), C.byref(info)))
        else:
            v = np.ascontiguousarray(values, dtype=np.float64)
            A.check(A.lib().amp_factor(self._h, A.ptr(v) if nvalues else None, nvalues, float(pivotTol),
                                       C.byref(info)))
        self.ok = bool(info.ok)
        self.info = {"ok": self.ok, "nNeg": info.nNeg, "nPerturbed": info.nPerturbed,
                     "failColumn": info.failColumn, "ms": info.ms}
        return self.ok

    def solve(self, B, out=None, nrhs=None):
        """B: q vectors one after another (numpy, host) or a device tensor; returns X of the same layout."""
        n = self.sym.n
        if hasattr(B, "data_ptr") and getattr(B, "is_cuda", False):
            q = nrhs if nrhs is not None else B.numel() // max(n, 1)
            if out is None:
                out = B.new_empty(B.shape)
            A.check(A.lib().amp_solve_device(self._h, A.ptr(B), A.ptr(out), int(q)))
            return out
        b = np.ascontiguousarray(B, dtype=np.float64)
        q = nrhs if nrhs is not None else (b.size // n if n else 1)
        x = np.zeros_like(b) if out is None else out
        A.check(A.lib().amp_solve(self._h, A.ptr(b) if b.size else None, A.ptr(x) if x.size else None, int(q)))
        return x

    def graph(self, values, B, X, nrhs, irSteps):
        """Capture factor(values) + solve of the nrhs right-hand sides in B into X + irSteps FP64 refinement steps as
        one replayable device graph (device tensors, read again at every replay). Cholesky handles only."""
        return FactorGraph(self, values, B, X, nrhs, irSteps)

    def refine(self, values, b, method="ir", maxSteps=3, restart=40, tol=1e-14, out=None):
        """Refine the factor's solution of one right-hand side on the device (FP64 residual and update):
        method 'ir' (solveRefined, ldlt.mjs:253-275) or 'gmres' (gmres-real.mjs:31-104 preconditioned by the factor).
        values: the unperturbed matrix values on the analysed pattern. Returns (x, info)."""
        o = A.RefineOpts()
        o.method = A.AMP_REFINE_IR if method == "ir" else A.AMP_REFINE_GMRES
        o.maxSteps, o.restart, o.tol = int(maxSteps), int(restart), float(tol)
        info = A.RefineInfo()
        if hasattr(b, "data_ptr") and getattr(b, "is_cuda", False):
            x = b.new_empty(b.shape) if out is None else out
            A.check(A.lib().amp_refine_device(self._h, A.ptr(values), A.ptr(b), A.ptr(x), C.byref(o), C.byref(info)))
        else:
            v = np.ascontiguousarray(values, dtype=np.float64)
            bb = np.ascontiguousarray(b, dtype=np.float64)
            x = np.zeros_like(bb) if out is None else out
            A.check(A.lib().amp_refine(self._h, A.ptr(v) if v.size else None, A.ptr(bb) if bb.size else None,
                                       A.ptr(x) if x.size else None, C.byref(o), C.byref(info)))
        return x, {"steps": info.steps, "converged": bool(info.converged), "relResidual": info.relResidual,
                   "ms": info.ms}

    def shift_invert(self, M, X, nvec):
        """Y = A^-1 (M X) for nvec vectors (eigen.ts:156-166): M on the analysed pattern, A the factored matrix.
        Host arrays (numpy) or device arrays; returns Y of X's kind and layout."""
        if hasattr(X, "data_ptr") and getattr(X, "is_cuda", False):
            Y = X.new_empty(X.shape)
            A.check(A.lib().amp_shift_invert_apply_device(self._h, A.ptr(M), A.ptr(X), A.ptr(Y), int(nvec)))
            return Y
        Mv = np.ascontiguousarray(M, dtype=np.float64)
        Xv = np.ascontiguousarray(X, dtype=np.float64)
        Y = np.empty_like(Xv)
        A.check(A.lib().amp_shift_invert_apply(self._h, A.ptr(Mv), A.ptr(Xv), A.ptr(Y), int(nvec)))
        return Y

    def download(self):
        """(Lx, D) in the reference's layout (D is None for Cholesky)."""
        Lx = np.zeros(max(self.sym.nnzLStored, 1))
        D = np.zeros(max(self.sym.n, 1)) if self.mode == A.AMP_LDLT else None
        A.check(A.lib().amp_download_factor(self._h, A.ptr(Lx), None if D is None else A.ptr(D)))
        return Lx[: self.sym.nnzLStored], (None if D is None else D[: self.sym.n])

    def set_profiling(self, enable=True):
        A.check(A.lib().amp_num_set_profiling(self._h, 1 if enable else 0))

    def set_growth_limit(self, limit):
        """FP32 factors: diagonal blocks whose largest elimination multiplier exceeds `limit` are redone in FP64
        (0: never). The count of the last factor is kernel_ms('diag_fp64_redo')[1]."""
        A.check(A.lib().amp_num_set_growth_limit(self._h, float(limit)))

    def release_workspace(self):
        """Free the update-matrix workspace (solves keep working; the next factor allocates it again)."""
        A.check(A.lib().amp_num_release_workspace(self._h))

    def kernel_ms(self, name):
        ms, cnt = C.c_double(), C.c_int32()
        A.check(A.lib().amp_num_get_kernel_ms(self._h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                A.lib().amp_num_free(h)
            except Exception:
                pass
            self._h = None


def _check_count(sym, values):
    nvalues = int(values.numel()) if hasattr(values, "numel") else len(values)
    if nvalues != sym.nnzPattern:
        raise ValueError(f"factorNumeric: {nvalues} values for a pattern with {sym.nnzPattern} entries")


def factorNumeric(sym, values, num=None, pat=None, factor=None):
    """Numeric LL^T on the device (sparse.mjs:941-1125). Reuses `num` storage when given for the same sym."""
    if pat is not None:
        checkSymbolic(sym, pat)
    _check_count(sym, values)
    kind = _resolve_kind(factor)
    if num is None or num.sym is not sym or num.kernel != kind or num.mode != A.AMP_CHOLESKY:
        num = NumericFactor(sym, A.AMP_CHOLESKY, kind)
    num.factor(values)
    return num


def solveNumeric(num, b, out=None):
    """Solve A x = b (sparse.mjs:1261-1305); b and x in the original ordering."""
    if num is None or not num.ok:
        raise RuntimeError("solveNumeric: no successful numeric factorization")
    return num.solve(b, out, nrhs=1)


def solveNumericMulti(num, bs, outs=None):
    """Several right-hand sides in one pass over the factor (sparse.mjs:1308-1310)."""
    if num is None or not num.ok:
        raise RuntimeError("solveNumericMulti: no successful numeric factorization")
    B = np.stack([np.asarray(b, dtype=np.float64) for b in bs]) if len(bs) else np.zeros((0, num.sym.n))
    X = num.solve(B.reshape(-1), nrhs=len(bs)).reshape(B.shape)
    if outs is not None:
        for o, x in zip(outs, X):
            o[:] = x
        return outs
    return [x for x in X]


class SparseCholesky:
    """Supernodal multifrontal Cholesky for a fixed pattern (sparse.mjs:1312-1372)."""

    def __init__(self, pat, perm=None, sym=None, ordering="amd", x=None, y=None, factor=None):
        t0 = time.perf_counter()
        self.factorKind = factor
        if sym is not None:
            checkSymbolic(sym, pat)
            self.sym = sym
        else:
            self.sym = analyze(pat, {"perm": perm} if perm is not None else {"ordering": ordering, "x": x, "y": y})
        self.n = int(pat["n"])
        self.perm = self.sym.perm
        self.nnzL = self.sym.nnzL
        self.num = None
        self.ok = False
        reused = sym is not None
        self.timing = {"n": self.n, "nnzA": self.sym.nnzA, "nnzL": self.nnzL, "fill": self.sym.fill,
                       "flops": self.sym.flops, "nsuper": self.sym.nsuper, "symbolicReused": reused,
                       "ordering_ms": 0.0 if reused else self.sym.timing["ordering_ms"],
                       "symbolic_ms": 0.0 if reused else
                       1e3 * (time.perf_counter() - t0) - self.sym.timing["ordering_ms"], "factor_ms": 0.0,
                       "nFactor": 0, "solve_ms": 0.0, "nSolve": 0}

    def factor(self, values):
        t0 = time.perf_counter()
        self.num = factorNumeric(self.sym, values, self.num, None, self.factorKind)
        self.ok = self.num.ok
        self.timing["kernel"] = self.num.kernel
        self.timing["factor_ms"] += 1e3 * (time.perf_counter() - t0)
        self.timing["nFactor"] += 1
        return self.ok

    def solve(self, b, out=None):
        t0 = time.perf_counter()
        if not self.ok:
            raise RuntimeError("SparseCholesky: solve without a successful factorization")
        r = solveNumeric(self.num, b, out)
        self.timing["solve_ms"] += 1e3 * (time.perf_counter() - t0)
        self.timing["nSolve"] += 1
        return r

    def solveMany(self, bs, outs=None):
        t0 = time.perf_counter()
        if not self.ok:
            raise RuntimeError("SparseCholesky: solve without a successful factorization")
        r = solveNumericMulti(self.num, bs, outs)
        self.timing["solve_ms"] += 1e3 * (time.perf_counter() - t0)
        self.timing["nSolve"] += len(bs)
        return r


class SupernodalLDLT:
    """Real symmetric LDL^T with static pivot perturbation and inertia (ldlt.mjs:417-645)."""

    def __init__(self, sym, factor=None):
        self.sym = sym
        self.factorKind = factor
        self.kernel = _resolve_kind(factor)
        self.num = None
        self.nnzL = sym.nnzL
        self.nNeg = 0
        self.nPerturbed = 0
        self.ok = False
        self.D = None

    def factor(self, values, pivotTol=1e-13):
        _check_count(self.sym, values)
        kind = _resolve_kind(self.factorKind)
        if self.num is None or self.num.kernel != kind:
            self.num = NumericFactor(self.sym, A.AMP_LDLT, kind)
        self.kernel = kind
        self.num.factor(values, pivotTol)
        info = self.num.info
        self.ok = info["ok"]
        self.nNeg = info["nNeg"]
        self.nPerturbed = info["nPerturbed"]
        self.factorMs = info["ms"]
        return {"ok": self.ok, "nNeg": self.nNeg, "nPerturbed": self.nPerturbed, "ms": info["ms"]}

    def _need(self):
        if not self.ok:
            raise RuntimeError("SupernodalLDLT: solve without a successful factorization")

    def solve(self, b, out=None):
        self._need()
        return self.num.solve(b, out, nrhs=1)

    def solveMany(self, b, q, out=None):
        """q right-hand sides stored one after another in b (length n q)."""
        self._need()
        return self.num.solve(b, out, nrhs=q)

    def solveRefined(self, pat, values, b, maxSteps=3, tol=1e-14):
        """Iterative refinement against the unperturbed values (ldlt.mjs:253-275), on the device.
        Returns {'x', 'relResidual', 'steps'}."""
        self._need()
        checkSymbolic(self.sym, pat)
        x, info = self.num.refine(values, b, "ir", maxSteps=maxSteps, tol=tol)
        return {"x": x, "relResidual": info["relResidual"], "steps": info["steps"]}
