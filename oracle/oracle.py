"""ctypes binding of the CPU oracle (oracle/liboracle.so). TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check in __graft_entry__.py and the CPU-baseline leg of bench.py may import this module; the
shipped library (applied-multiphysics_amd/) never does. Function-by-function citations of the reference are in
oracle.h.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")


class OrcSym(C.Structure):
    _fields_ = [
        ("n", C.c_int32), ("nsuper", C.c_int32), ("maxFront", C.c_int32),
        ("nnzL", C.c_int64), ("nnzA", C.c_int64), ("nnzLStored", C.c_int64),
        ("flops", C.c_double), ("fill", C.c_double),
        ("perm", C.POINTER(C.c_int32)), ("pinv", C.POINTER(C.c_int32)), ("parent", C.POINTER(C.c_int32)),
        ("colCount", C.POINTER(C.c_int32)),
        ("lowerP", C.POINTER(C.c_int32)), ("lowerI", C.POINTER(C.c_int32)), ("lowerSrc", C.POINTER(C.c_int32)),
        ("superStart", C.POINTER(C.c_int32)), ("colSuper", C.POINTER(C.c_int32)),
        ("superParent", C.POINTER(C.c_int32)), ("childHead", C.POINTER(C.c_int32)),
        ("childNext", C.POINTER(C.c_int32)),
        ("Sp", C.POINTER(C.c_int32)), ("Si", C.POINTER(C.c_int32)),
        ("Lxp", C.POINTER(C.c_int64)),
        ("patternKey", C.c_char * 64),
    ]


class OrcUplook(C.Structure):
    _fields_ = [
        ("n", C.c_int32),
        ("perm", C.POINTER(C.c_int32)), ("pinv", C.POINTER(C.c_int32)), ("Cp", C.POINTER(C.c_int32)),
        ("Ci", C.POINTER(C.c_int32)), ("srcIndex", C.POINTER(C.c_int32)), ("parent", C.POINTER(C.c_int32)),
        ("Lp", C.POINTER(C.c_int32)), ("Li", C.POINTER(C.c_int32)),
        ("Lx", C.POINTER(C.c_double)), ("nnzL", C.c_int64),
    ]


APPLY_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


class GmresInfo(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("converged", C.c_int32), ("relres", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_amd.argtypes = [C.c_int32, i32p, i32p, C.c_double, i32p]
        L.orc_amd.restype = C.c_int
        L.orc_nd_order.argtypes = [C.c_int32, i32p, i32p, f64p, f64p, C.c_int32, i32p]
        L.orc_nd_order.restype = C.c_int
        L.orc_analyze.argtypes = [C.c_int32, i32p, i32p, C.c_void_p, C.c_int]
        L.orc_analyze.restype = C.POINTER(OrcSym)
        L.orc_sym_free.argtypes = [C.POINTER(OrcSym)]
        L.orc_fingerprint.argtypes = [C.c_int32, i32p, i32p, C.c_char_p]
        L.orc_chol_factor.argtypes = [C.POINTER(OrcSym), f64p, f64p]
        L.orc_chol_factor.restype = C.c_int
        L.orc_chol_factor_subtree.argtypes = [C.POINTER(OrcSym), f64p, C.c_int32, f64p, C.POINTER(C.c_double)]
        L.orc_chol_factor_subtree.restype = C.c_int
        L.orc_chol_solve.argtypes = [C.POINTER(OrcSym), f64p, f64p, f64p]
        L.orc_chol_solve_multi.argtypes = [C.POINTER(OrcSym), f64p, f64p, C.c_int32, f64p]
        L.orc_ldlt_factor.argtypes = [C.POINTER(OrcSym), f64p, C.c_double, f64p, f64p, C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32)]
        L.orc_ldlt_factor.restype = C.c_int
        L.orc_ldlt_solve.argtypes = [C.POINTER(OrcSym), f64p, f64p, f64p, f64p]
        L.orc_ldlt_solve_many.argtypes = [C.POINTER(OrcSym), f64p, f64p, f64p, C.c_int32, f64p]
        L.orc_ldlt_solve_refined.argtypes = [C.POINTER(OrcSym), f64p, f64p, i32p, i32p, f64p, f64p, C.c_int32,
                                             C.c_double, f64p, C.POINTER(C.c_double)]
        L.orc_ldlt_solve_refined.restype = C.c_int32
        L.orc_uplook_create.argtypes = [C.c_int32, i32p, i32p, i32p]
        L.orc_uplook_create.restype = C.POINTER(OrcUplook)
        L.orc_uplook_factor.argtypes = [C.POINTER(OrcUplook), f64p]
        L.orc_uplook_factor.restype = C.c_int
        L.orc_uplook_solve.argtypes = [C.POINTER(OrcUplook), f64p, f64p]
        L.orc_uplook_free.argtypes = [C.POINTER(OrcUplook)]
        L.orc_assemble_p1.argtypes = [C.c_int32, C.c_int32, f64p, C.c_int32, i32p, f64p, C.c_void_p, i32p, i32p, f64p]
        L.orc_assemble_p1.restype = C.c_int64
        L.orc_cldlt_factor.argtypes = [C.POINTER(OrcSym), f64p, f64p, C.c_double, f64p, f64p, f64p, f64p,
                                       C.POINTER(C.c_int32)]
        L.orc_cldlt_factor.restype = C.c_int
        L.orc_cldlt_solve.argtypes = [C.POINTER(OrcSym), f64p, f64p, f64p, f64p, f64p, f64p, f64p, f64p]
        L.orc_cldlt_solve_refined.argtypes = [C.POINTER(OrcSym), f64p, f64p, f64p, f64p, i32p, i32p, f64p, f64p, f64p,
                                              f64p, C.c_int32, C.c_double, f64p, f64p, C.POINTER(C.c_double)]
        L.orc_cldlt_solve_refined.restype = C.c_int32
        L.orc_set_threads.argtypes = [C.c_int32]
        L.orc_get_threads.restype = C.c_int32
        L.orc_gmres.argtypes = [C.c_int32, APPLY_FN, C.c_void_p, APPLY_FN, C.c_void_p, f64p, C.c_double, C.c_int32,
                                C.c_int32, C.c_void_p, f64p, C.POINTER(GmresInfo)]
        L.orc_csr_matvec.argtypes = [C.c_int32, i32p, i32p, f64p, f64p, f64p]
        _lib = L
    return _lib


def _arr(ptr, n, dtype=np.int32):
    if n == 0:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)


def _pat(pat):
    return (np.ascontiguousarray(pat["rowptr"], dtype=np.int32), np.ascontiguousarray(pat["colidx"], dtype=np.int32))


def amd(pat, dense_factor=10.0):
    n = pat["n"]
    rp, ci = _pat(pat)
    perm = np.zeros(n, dtype=np.int32)
    if lib().orc_amd(n, rp, ci, dense_factor, perm) != 0:
        raise RuntimeError("amdOrder: internal error")
    return perm


def nd_order(pat, x, y, leaf=48):
    """ndOrder (sparse.mjs:230-283): perm new -> old from node coordinates x, y."""
    n = pat["n"]
    rp, ci = _pat(pat)
    perm = np.zeros(n, dtype=np.int32)
    if lib().orc_nd_order(n, rp, ci, np.ascontiguousarray(x, np.float64), np.ascontiguousarray(y, np.float64), leaf,
                          perm) != 0:
        raise RuntimeError("ndOrder: bad arguments or allocation failure")
    return perm


def fingerprint(pat):
    rp, ci = _pat(pat)
    buf = C.create_string_buffer(64)
    lib().orc_fingerprint(pat["n"], rp, ci, buf)
    return buf.value.decode()


class Sym:
    """Symbolic analysis (same fields as analyze() in sparse.mjs:873-901) held by the oracle."""

    def __init__(self, pat, perm=None, relax=True):
        rp, ci = _pat(pat)
        self._pat = (rp, ci)
        pp = None
        if perm is not None:
            self._perm_in = np.ascontiguousarray(perm, dtype=np.int32)
            pp = self._perm_in.ctypes.data_as(C.c_void_p)
        self._p = lib().orc_analyze(pat["n"], rp, ci, pp, 1 if relax else 0)
        if not self._p:
            raise RuntimeError("analyze failed")
        s = self._p.contents
        n, ns = s.n, s.nsuper
        self.n, self.nsuper, self.maxFront = n, ns, s.maxFront
        self.nnzL, self.nnzA, self.nnzLStored = s.nnzL, s.nnzA, s.nnzLStored
        self.flops, self.fill = s.flops, s.fill
        self.patternKey = s.patternKey.decode()
        self.perm = _arr(s.perm, n)
        self.pinv = _arr(s.pinv, n)
        self.parent = _arr(s.parent, n)
        self.colCount = _arr(s.colCount, n)
        self.lowerP = _arr(s.lowerP, n + 1)
        self.lowerI = _arr(s.lowerI, int(s.nnzA))
        self.lowerSrc = _arr(s.lowerSrc, int(s.nnzA))
        self.superStart = _arr(s.superStart, ns + 1)
        self.colSuper = _arr(s.colSuper, n)
        self.superParent = _arr(s.superParent, ns)
        self.childHead = _arr(s.childHead, ns)
        self.childNext = _arr(s.childNext, ns)
        self.Sp = _arr(s.Sp, ns + 1)
        self.Si = _arr(s.Si, int(self.Sp[-1]))
        self.Lxp = _arr(s.Lxp, ns + 1, np.int64)

    def __del__(self):
        try:
            if self._p:
                lib().orc_sym_free(self._p)
                self._p = None
        except Exception:
            pass


def assemble_p1(nodes, conn, kel, free=None):
    """P1 diffusion assembly restricted to the free nodes (assembleCells + elementSlots + restrict): returns the CSR
    pattern dict and the values."""
    nodes = np.ascontiguousarray(nodes, np.float64)
    conn = np.ascontiguousarray(conn, np.int32)
    nn, dim = nodes.shape
    ne, k = conn.shape
    kel = np.ascontiguousarray(kel, np.float64)
    fm = None if free is None else np.ascontiguousarray(free, np.uint8)
    nf = nn if fm is None else int(fm.sum())
    cap = ne * k * k + nn
    rp = np.zeros(nf + 1, np.int32)
    ci = np.zeros(cap, np.int32)
    v = np.zeros(cap)
    nnz = lib().orc_assemble_p1(dim, nn, nodes, ne, conn, kel, None if fm is None else fm.ctypes.data, rp, ci, v)
    return {"n": nf, "rowptr": rp, "colidx": ci[:nnz].copy(), "nnz": int(nnz)}, v[:nnz].copy()


def cldlt_factor(sym, re, im, pivot_tol=1e-13):
    """ComplexSupernodalLDLT.factor: (ok, Lr, Li, Dr, Di, nPerturbed)."""
    L = lib()
    nl = max(sym.nnzLStored, 1)
    Lr, Li = np.zeros(nl), np.zeros(nl)
    Dr, Di = np.zeros(max(sym.n, 1)), np.zeros(max(sym.n, 1))
    npert = C.c_int32()
    ok = L.orc_cldlt_factor(sym._p, np.ascontiguousarray(re, np.float64), np.ascontiguousarray(im, np.float64),
                            float(pivot_tol), Lr, Li, Dr, Di, C.byref(npert))
    return bool(ok), Lr, Li, Dr, Di, npert.value


def cldlt_solve(sym, Lr, Li, Dr, Di, br, bi):
    n = sym.n
    xr, xi = np.zeros(max(n, 1)), np.zeros(max(n, 1))
    lib().orc_cldlt_solve(sym._p, Lr, Li, Dr, Di, np.ascontiguousarray(br, np.float64),
                          np.ascontiguousarray(bi, np.float64), xr, xi)
    return xr[:n], xi[:n]


def cldlt_solve_refined(sym, Lr, Li, Dr, Di, pat, re, im, br, bi, max_steps=3, tol=1e-14):
    n = sym.n
    xr, xi = np.zeros(max(n, 1)), np.zeros(max(n, 1))
    rel = C.c_double()
    rp, ci = _pat(pat)
    steps = lib().orc_cldlt_solve_refined(sym._p, Lr, Li, Dr, Di, rp, ci, np.ascontiguousarray(re, np.float64),
                                          np.ascontiguousarray(im, np.float64), np.ascontiguousarray(br, np.float64),
                                          np.ascontiguousarray(bi, np.float64), int(max_steps), float(tol), xr, xi,
                                          C.byref(rel))
    return xr[:n], xi[:n], rel.value, steps


def set_threads(k):
    """Threads of the numeric factorization (the factor is bitwise the same for every count)."""
    lib().orc_set_threads(int(k))


def get_threads():
    return int(lib().orc_get_threads())


def chol_factor(sym, values):
    """factorNumeric: returns (ok, Lx)."""
    Lx = np.zeros(max(int(sym.nnzLStored), 1), dtype=np.float64)
    ok = lib().orc_chol_factor(sym._p, np.ascontiguousarray(values, dtype=np.float64), Lx)
    return bool(ok), Lx[: int(sym.nnzLStored)]


def chol_factor_subtree(sym, values, root, Lx=None):
    """Oracle factor of the complete subtree of supernode `root`: (ok, flops of the subtree, Lx)."""
    if Lx is None:
        Lx = np.zeros(max(int(sym.nnzLStored), 1), dtype=np.float64)
    fl = C.c_double(0)
    ok = lib().orc_chol_factor_subtree(sym._p, np.ascontiguousarray(values, dtype=np.float64), int(root), Lx,
                                       C.byref(fl))
    return bool(ok), fl.value, Lx


def chol_solve(sym, Lx, b):
    x = np.zeros(sym.n, dtype=np.float64)
    lib().orc_chol_solve(sym._p, np.ascontiguousarray(Lx, dtype=np.float64) if len(Lx) else np.zeros(1),
                         np.ascontiguousarray(b, dtype=np.float64), x)
    return x


def chol_solve_multi(sym, Lx, B):
    B = np.ascontiguousarray(B, dtype=np.float64)
    q = B.shape[0]
    X = np.zeros_like(B)
    lib().orc_chol_solve_multi(sym._p, np.ascontiguousarray(Lx), B.reshape(-1), q, X.reshape(-1))
    return X


def ldlt_factor(sym, values, pivot_tol=1e-13):
    """SupernodalLDLT.factor: returns (ok, Lx, D, nNeg, nPerturbed)."""
    Lx = np.zeros(max(int(sym.nnzLStored), 1), dtype=np.float64)
    D = np.zeros(max(sym.n, 1), dtype=np.float64)
    nn, npt = C.c_int32(0), C.c_int32(0)
    ok = lib().orc_ldlt_factor(sym._p, np.ascontiguousarray(values, dtype=np.float64), pivot_tol, Lx, D,
                               C.byref(nn), C.byref(npt))
    return bool(ok), Lx[: int(sym.nnzLStored)], D[: sym.n], nn.value, npt.value


def ldlt_solve(sym, Lx, D, b):
    x = np.zeros(sym.n, dtype=np.float64)
    lib().orc_ldlt_solve(sym._p, np.ascontiguousarray(Lx), np.ascontiguousarray(D),
                         np.ascontiguousarray(b, dtype=np.float64), x)
    return x


def ldlt_solve_many(sym, Lx, D, B):
    B = np.ascontiguousarray(B, dtype=np.float64)
    q = B.shape[0]
    X = np.zeros_like(B)
    lib().orc_ldlt_solve_many(sym._p, np.ascontiguousarray(Lx), np.ascontiguousarray(D), B.reshape(-1), q,
                              X.reshape(-1))
    return X


def ldlt_solve_refined(sym, Lx, D, pat, values, b, max_steps=3, tol=1e-14):
    rp, ci = _pat(pat)
    x = np.zeros(sym.n, dtype=np.float64)
    rel = C.c_double(0)
    steps = lib().orc_ldlt_solve_refined(sym._p, np.ascontiguousarray(Lx), np.ascontiguousarray(D), rp, ci,
                                         np.ascontiguousarray(values, dtype=np.float64),
                                         np.ascontiguousarray(b, dtype=np.float64), max_steps, tol, x, C.byref(rel))
    return x, rel.value, steps


class UpLooking:
    """UpLookingCholesky (sparse.mjs:1379-1544)."""

    def __init__(self, pat, perm):
        rp, ci = _pat(pat)
        self._p = lib().orc_uplook_create(pat["n"], rp, ci, np.ascontiguousarray(perm, dtype=np.int32))
        u = self._p.contents
        self.n = u.n
        self.nnzL = u.nnzL
        self.Lp = _arr(u.Lp, u.n + 1)

    def factor(self, values):
        return bool(lib().orc_uplook_factor(self._p, np.ascontiguousarray(values, dtype=np.float64)))

    def solve(self, b):
        x = np.zeros(self.n, dtype=np.float64)
        lib().orc_uplook_solve(self._p, np.ascontiguousarray(b, dtype=np.float64), x)
        return x

    def __del__(self):
        try:
            if self._p:
                lib().orc_uplook_free(self._p)
                self._p = None
        except Exception:
            pass


def csr_matvec(pat, values, x):
    rp, ci = _pat(pat)
    y = np.zeros(pat["n"], dtype=np.float64)
    lib().orc_csr_matvec(pat["n"], rp, ci, np.ascontiguousarray(values, dtype=np.float64),
                         np.ascontiguousarray(x, dtype=np.float64), y)
    return y


def gmres(n, matvec, precond, b, tol=1e-10, restart=40, max_iter=400, x0=None):
    """gmres (gmres-real.mjs:31-104) with Python callbacks: matvec(v) -> A v, precond(v) -> M v."""

    def wrap(fn):
        def cb(_ctx, vp, outp):
            v = np.ctypeslib.as_array(vp, shape=(n,))
            out = np.ctypeslib.as_array(outp, shape=(n,))
            out[:] = fn(v.copy())
        return APPLY_FN(cb)

    mv, pc = wrap(matvec), wrap(precond)
    x = np.zeros(n, dtype=np.float64)
    info = GmresInfo()
    x0a = None if x0 is None else np.ascontiguousarray(x0, dtype=np.float64)
    lib().orc_gmres(n, mv, None, pc, None, np.ascontiguousarray(b, dtype=np.float64), tol, restart, max_iter,
                    None if x0a is None else x0a.ctypes.data_as(C.c_void_p), x, C.byref(info))
    return x, info.iterations, info.relres, bool(info.converged)
