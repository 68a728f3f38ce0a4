"""Complex-symmetric supernodal LDL^T on the device, with the interface of ComplexSupernodalLDLT
(packages/core/src/la/ldlt.mjs:646-866; declarations ldlt.d.mts:49-56): factor(re, im) -> {ok, nPerturbed, ms},
solve(br, bi) -> {re, im}, solveRefined(pat, re, im, br, bi) -> {re, im, relResidual, steps}.

A complex-symmetric system (B + iC)(x + iy) = f + ig with B, C real symmetric is the real symmetric system
[[B, C], [C, -B]] [x; -y] = [f; g]. Each node's two unknowns are kept together (interleaved: 2 j, 2 j + 1), so every
entry of the complex pattern becomes the 2 x 2 real block [[b, c], [c, -b]], and the node ordering of the caller's
analysis extends to the real system node by node (the tree of the complex factor with every node doubled). The real
system is factored by the device LDL^T with static pivoting (matrix cores, FP64 or FP32), solved, and refined against
the complex residual, which is the real system's residual. When the imaginary parts dominate the diagonal, the system
is first multiplied by -i (real part C, imaginary part -B), so that the real system's 1 x 1 pivots are the larger
diagonal parts. This is not the reference's complex factor (L and D differ, and nPerturbed counts real pivots); the
solutions agree to the conditioning of the system.
"""
import numpy as np

from .sparse import A as _A
from .sparse import NumericFactor, _resolve_kind, analyze, checkSymbolic


class ComplexSupernodalLDLT:
    """Complex-symmetric LDL^T (no conjugation) on the device over a node analysis from analyze()."""

    def __init__(self, sym, factor=None, pat=None):
        self.sym = sym
        pat = pat if pat is not None else getattr(sym, "pattern", None)
        if pat is None:
            raise ValueError("ComplexSupernodalLDLT: the analysis carries no pattern (pass pat)")
        self.kind = _resolve_kind(factor)
        n = int(pat["n"])
        rp = np.asarray(pat["rowptr"], np.int64)
        ci = np.asarray(pat["colidx"], np.int64)
        self.n, self.nnz = n, int(rp[n]) if n else 0
        cnt = np.diff(rp)
        rp2 = np.zeros(2 * n + 1, np.int64)
        rp2[1::2] = 2 * cnt
        rp2[2::2] = 2 * cnt
        rp2 = np.cumsum(rp2)
        # slots of the 2 x 2 block of node entry p = (i, j): rows 2 i + r, columns 2 j + d
        rows = np.repeat(np.arange(n), cnt)
        off = 2 * (np.arange(self.nnz) - rp[rows])
        self.s00 = rp2[2 * rows] + off
        self.s01 = self.s00 + 1
        self.s10 = rp2[2 * rows + 1] + off
        self.s11 = self.s10 + 1
        ci2 = np.empty(2 * 2 * self.nnz, np.int32)
        ci2[self.s00] = 2 * ci
        ci2[self.s01] = 2 * ci + 1
        ci2[self.s10] = 2 * ci
        ci2[self.s11] = 2 * ci + 1
        self.pat2 = {"n": 2 * n, "rowptr": rp2.astype(np.int32), "colidx": ci2, "nnz": 4 * self.nnz}
        perm2 = np.empty(2 * n, np.int32)
        perm2[0::2] = 2 * np.asarray(sym.perm, np.int64)
        perm2[1::2] = 2 * np.asarray(sym.perm, np.int64) + 1
        self.sym2 = analyze(self.pat2, {"perm": perm2})
        self._diag = np.nonzero(ci == rows)[0]
        self.num = None
        self.nPerturbed = 0
        self.ok = False
        self.rotated = False

    def _real_values(self, re, im):
        b, c = (im, -re) if self.rotated else (re, im)
        v = np.empty(4 * self.nnz)
        v[self.s00] = b
        v[self.s01] = c
        v[self.s10] = c
        v[self.s11] = -b
        return v

    def _real_rhs(self, br, bi):
        f, g = (bi, -br) if self.rotated else (br, bi)
        w = np.empty(2 * self.n)
        w[0::2] = f
        w[1::2] = g
        return w

    def factor(self, re, im, pivotTol=1e-13):
        re = np.ascontiguousarray(re, np.float64)
        im = np.ascontiguousarray(im, np.float64)
        if len(re) != self.nnz or len(im) != self.nnz:
            raise ValueError(f"factorNumeric: {len(re)} values for a pattern with {self.nnz} entries")
        dre, dim_ = np.abs(re[self._diag]), np.abs(im[self._diag])
        self.rotated = bool(dim_.max(initial=0.0) > dre.max(initial=0.0))
        # the reference's tau = pivotTol max_j |a_jj| with |a_jj| = hypot(re, im) (ldlt.mjs:675); the real system scales
        # by max_j |b_jj| of its own diagonal
        hyp = np.hypot(re[self._diag], im[self._diag]).max(initial=0.0)
        bmax = (dim_ if self.rotated else dre).max(initial=0.0)
        tol = pivotTol * (hyp / bmax if bmax > 0 and hyp > 0 else 1.0)
        if self.num is None or self.num.kernel != self.kind:
            self.num = NumericFactor(self.sym2, _A.AMP_LDLT, self.kind)
        self._values = self._real_values(re, im)
        self.num.factor(self._values, tol)
        info = self.num.info
        self.ok, self.nPerturbed = info["ok"], info["nPerturbed"]
        return {"ok": self.ok, "nPerturbed": self.nPerturbed, "ms": info["ms"]}

    def _need(self):
        if not self.ok:
            raise RuntimeError("ComplexSupernodalLDLT: solve without a successful factorization")

    def solve(self, br, bi):
        self._need()
        w = self.num.solve(self._real_rhs(np.asarray(br, np.float64), np.asarray(bi, np.float64)), nrhs=1)
        return {"re": w[0::2].copy(), "im": -w[1::2]}

    def solveRefined(self, pat, re, im, br, bi, maxSteps=3, tol=1e-14):
        """complexSolveRefined (ldlt.mjs:838-862): refinement against the unperturbed complex matrix; the real
        system's residual has the complex residual's norm."""
        self._need()
        checkSymbolic(self.sym, pat)
        v = self._real_values(np.asarray(re, np.float64), np.asarray(im, np.float64))
        w, info = self.num.refine(v, self._real_rhs(np.asarray(br, np.float64), np.asarray(bi, np.float64)), "ir",
                                  maxSteps=maxSteps, tol=tol)
        return {"re": w[0::2].copy(), "im": -w[1::2], "relResidual": info["relResidual"], "steps": info["steps"]}
