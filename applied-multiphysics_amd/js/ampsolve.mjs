// Host layer of the device sparse direct solver, for the reference's JavaScript callers.
//
// Same exports, argument meaning and error messages as jwt625/applied-multiphysics packages/core/src/la/sparse.mjs and
// ldlt.mjs (declarations sparse.d.mts:28-57, ldlt.d.mts:40-59):
//   analyze (sparse.mjs:666-902), patternFingerprint / checkSymbolic (:908-930), factorNumeric (:941-1125),
//   solveNumeric / solveNumericMulti (:1261-1310), SparseCholesky (:1312-1372), SupernodalLDLT (ldlt.mjs:417-645),
//   setDefaultRealFactor (wasm/front-real.mjs:196-200).
// The numeric work runs on the GPU through amp_native.node (a thin Node-API addon over include/ampsolve.h). An
// analysis produced by the reference's own analyze() is accepted unchanged: it is handed over once per object
// (amp_sym_import) and cached. Written in ES2019 so that it runs on the Node of this image.
import { createRequire } from 'module';

const require = createRequire(import.meta.url);
const native = require('./amp_native.node');

export const REAL_FACTOR_KINDS = ['hip', 'hip-f32', 'hip-bf16x3'];
export let DEFAULT_REAL_FACTOR = 'hip';

/**
 * Select the default device factor: 'hip' (FP64), 'hip-f32' (FP32 factor) or 'hip-bf16x3' (FP32 factor whose bulk
 * updates run as three bf16 products); refine the reduced-precision ones with solveRefined.
 */
export const setDefaultRealFactor = (k) => {
  if (!REAL_FACTOR_KINDS.includes(k)) throw new Error(`unknown real factor ${k} (hip | hip-f32 | hip-bf16x3)`);
  return (DEFAULT_REAL_FACTOR = k);
};

function resolveKind(k) {
  const kind = k === null || k === undefined ? DEFAULT_REAL_FACTOR : k;
  if (!REAL_FACTOR_KINDS.includes(kind)) throw new Error(`unknown real factor ${kind} (hip | hip-f32 | hip-bf16x3)`);
  return kind;
}

const precisionOf = (kind) => (kind === 'hip-f32' ? 1 : kind === 'hip-bf16x3' ? 2 : 0);

/** Approximate minimum degree ordering, perm new -> old (sparse.mjs:297-592). */
export function amdOrder(pat, denseFactor = 10) {
  return native.amdOrder(pat.n, pat.rowptr, pat.colidx, denseFactor);
}

/** Nested dissection ordering by recursive coordinate bisection, perm new -> old (sparse.mjs:230-283). */
export function ndOrder(pat, x, y, leaf = 48) {
  return native.ndOrder(pat.n, pat.rowptr, pat.colidx, Float64Array.from(x), Float64Array.from(y), leaf);
}

/** "n:nnz:hash", 32-bit FNV-1a over rowptr and colidx. */
export function patternFingerprint(pat) {
  return native.fingerprint(pat.n, pat.rowptr, pat.colidx);
}

/** Throw unless `sym` was analysed from `pat`. */
export function checkSymbolic(sym, pat) {
  if (!sym || sym.n !== pat.n || sym.patternKey !== patternFingerprint(pat)) throw new Error('symbolic analysis is for a different sparsity pattern');
}

/**
 * Symbolic analysis. opts.perm: fill-reducing order (new -> old) to start from; else opts.ordering 'amd' (default) or
 * 'nd' (needs node coordinates opts.x, opts.y of length n); opts.relax default true. Returns the fields of the
 * reference's analysis object plus a native handle; timing {ordering_ms, symbolic_ms}.
 */
export function analyze(pat, opts = {}) {
  const n = pat.n;
  if (opts.perm && opts.perm.length !== n) throw new Error(`analyze: perm has length ${opts.perm.length}, pattern has ${n} rows`);
  const nd = !opts.perm && n > 0 && opts.ordering === 'nd';
  if (nd && !(opts.x && opts.x.length === n && opts.y && opts.y.length === n)) throw new Error('analyze: ordering nd needs node coordinates x and y of length n');
  const t0 = Date.now();
  let perm = opts.perm ? Int32Array.from(opts.perm) : null;
  if (!perm && n > 0) perm = nd ? ndOrder(pat, opts.x, opts.y) : amdOrder(pat);
  const t1 = Date.now();
  const sym = native.analyze(n, pat.rowptr, pat.colidx, perm, opts.relax !== false);
  sym.nnzPattern = pat.rowptr[n];
  sym.timing = { ordering_ms: opts.perm ? 0 : t1 - t0, symbolic_ms: Date.now() - t1 };
  return sym;
}

const imported = new WeakMap();

/** Native handle of an analysis: its own, or the one made when a reference analysis is first handed over. */
function symHandle(sym, pat) {
  if (sym.handle) return sym.handle;
  let h = imported.get(sym);
  if (!h) {
    const nnz = Number(String(sym.patternKey).split(':')[1]);
    h = native.importSymbolic(sym, pat ? pat.rowptr : null, pat ? pat.colidx : null, Number.isFinite(nnz) ? nnz : 0);
    imported.set(sym, h);
  }
  return h;
}

function patternEntries(sym) {
  if (sym.nnzPattern !== undefined) return sym.nnzPattern;
  return Number(String(sym.patternKey).split(':')[1]);
}

function makeNumeric(sym, pat, mode, kind) {
  const handle = native.numCreate(symHandle(sym, pat), mode, precisionOf(kind));
  const num = { sym, handle, ok: false, kernel: kind, mode, info: null, work: null, ws: null };
  // the factor lives in HBM; Lx (and D) are copied to the host in the reference layout only when read
  Object.defineProperty(num, 'Lx', {
    get() {
      const Lx = new Float64Array(sym.Lxp[sym.nsuper]);
      native.downloadFactor(handle, Lx, null);
      return Lx;
    },
  });
  return num;
}

/**
 * Numeric LL^T of the SPD matrix `values` (aligned with the analysed CSR pattern) on the GPU. Reuses `num` when it
 * belongs to the same analysis and kind. Returns {sym, ok, kernel, ...}; ok false when not positive definite.
 */
export function factorNumeric(sym, values, num = null, pat = null, { factor = null } = {}) {
  if (pat) checkSymbolic(sym, pat);
  const nnzPat = patternEntries(sym);
  if (Number.isFinite(nnzPat) && values.length !== nnzPat) throw new Error(`factorNumeric: ${values.length} values for a pattern with ${nnzPat} entries`);
  const kind = resolveKind(factor);
  if (!num || num.sym !== sym || num.kernel !== kind || num.mode !== 0) num = makeNumeric(sym, pat, 0, kind);
  const info = native.factor(num.handle, values instanceof Float64Array ? values : Float64Array.from(values), 0);
  num.info = info;
  num.ok = info.ok;
  return num;
}

/** Solve A x = b with a numeric factor; b and the result in the original ordering. */
export function solveNumeric(num, b, out = null) {
  if (!num || !num.ok) throw new Error('solveNumeric: no successful numeric factorization');
  const o = out || new Float64Array(num.sym.n);
  native.solve(num.handle, b instanceof Float64Array ? b : Float64Array.from(b), o, 1);
  return o;
}

/** Solve A X = B for several right-hand sides in one pass over the factor. */
export function solveNumericMulti(num, bs, outs = null) {
  if (!num || !num.ok) throw new Error('solveNumericMulti: no successful numeric factorization');
  const n = num.sym.n, q = bs.length;
  const B = new Float64Array(n * q), X = new Float64Array(n * q);
  bs.forEach((b, k) => B.set(b, k * n));
  if (q > 0) native.solve(num.handle, B, X, q);
  return bs.map((_, k) => {
    const o = outs && outs[k] ? outs[k] : new Float64Array(n);
    o.set(X.subarray(k * n, (k + 1) * n));
    return o;
  });
}

/** Supernodal multifrontal Cholesky for a fixed pattern, factored and solved on the GPU. */
export class SparseCholesky {
  constructor(pat, perm = null, { sym = null, ordering = 'amd', x = null, y = null, factor = null } = {}) {
    const tStart = Date.now();
    this.factorKind = factor;
    this.sym = sym || analyze(pat, perm ? { perm } : { ordering, x, y });
    if (sym) checkSymbolic(sym, pat);
    this.pat = pat;
    this.n = pat.n;
    this.perm = this.sym.perm;
    this.nnzL = this.sym.nnzL;
    this.num = null;
    this.ok = false;
    const reused = !!sym;
    const orderingMs = reused || !this.sym.timing ? 0 : this.sym.timing.ordering_ms;
    this.timing = {
      n: this.n, nnzA: this.sym.nnzA, nnzL: this.nnzL, fill: this.sym.fill, flops: this.sym.flops, nsuper: this.sym.nsuper,
      symbolicReused: reused, ordering_ms: orderingMs, symbolic_ms: reused ? 0 : Date.now() - tStart - orderingMs,
      factor_ms: 0, nFactor: 0, solve_ms: 0, nSolve: 0,
    };
  }

  factor(values) {
    const tStart = Date.now();
    this.num = factorNumeric(this.sym, values, this.num, null, { factor: this.factorKind });
    this.ok = this.num.ok;
    this.timing.kernel = this.num.kernel;
    this.timing.factor_ms += Date.now() - tStart;
    this.timing.nFactor++;
    return this.ok;
  }

  solve(b, out) {
    const tStart = Date.now();
    if (!this.ok) throw new Error('SparseCholesky: solve without a successful factorization');
    const r = solveNumeric(this.num, b, out);
    this.timing.solve_ms += Date.now() - tStart;
    this.timing.nSolve++;
    return r;
  }

  solveMany(bs, outs = null) {
    const tStart = Date.now();
    if (!this.ok) throw new Error('SparseCholesky: solve without a successful factorization');
    const r = solveNumericMulti(this.num, bs, outs);
    this.timing.solve_ms += Date.now() - tStart;
    this.timing.nSolve += bs.length;
    return r;
  }
}

/** Real symmetric LDL^T (static pivot perturbation, inertia) on the GPU over an analysis of the pattern. */
export class SupernodalLDLT {
  constructor(sym, { factor = null } = {}) {
    this.sym = sym;
    this.factorKind = factor;
    this.kernel = resolveKind(factor);
    this.num = null;
    this.nnzL = sym.nnzL;
    this.nNeg = 0;
    this.nPerturbed = 0;
    this.ok = false;
    this.ws = null;
  }

  factor(values, { pivotTol = 1e-13 } = {}) {
    const nnzPat = patternEntries(this.sym);
    if (Number.isFinite(nnzPat) && values.length !== nnzPat) throw new Error(`factorNumeric: ${values.length} values for a pattern with ${nnzPat} entries`);
    const kind = resolveKind(this.factorKind);
    if (!this.num || this.num.kernel !== kind) this.num = makeNumeric(this.sym, null, 1, kind);
    this.kernel = kind;
    const info = native.factor(this.num.handle, values instanceof Float64Array ? values : Float64Array.from(values), pivotTol);
    this.ok = this.num.ok = info.ok;
    this.nNeg = info.nNeg;
    this.nPerturbed = info.nPerturbed;
    this.factorMs = info.ms;
    return { ok: info.ok, nNeg: info.nNeg, nPerturbed: info.nPerturbed, ms: info.ms };
  }

  get D() {
    const D = new Float64Array(this.sym.n);
    native.downloadFactor(this.num.handle, new Float64Array(this.sym.Lxp[this.sym.nsuper]), D);
    return D;
  }

  get Lx() {
    return this.num ? this.num.Lx : null;
  }

  solve(b, out) {
    if (!this.ok) throw new Error('SupernodalLDLT: solve without a successful factorization');
    const o = out || new Float64Array(this.sym.n);
    native.solve(this.num.handle, b instanceof Float64Array ? b : Float64Array.from(b), o, 1);
    return o;
  }

  /** q right-hand sides stored one after another in b (length n q). */
  solveMany(b, q, out) {
    if (!this.ok) throw new Error('SupernodalLDLT: solve without a successful factorization');
    const o = out || new Float64Array(this.sym.n * q);
    native.solve(this.num.handle, b instanceof Float64Array ? b : Float64Array.from(b), o, q);
    return o;
  }

  /** Iterative refinement against the unperturbed values; residual and update in FP64 on the GPU. */
  solveRefined(pat, values, b, { maxSteps = 3, tol = 1e-14 } = {}) {
    if (!this.ok) throw new Error('SupernodalLDLT: solve without a successful factorization');
    checkSymbolic(this.sym, pat);
    if (!this.sym.handle && !this.patternSet) {
      native.setPattern(symHandle(this.sym, pat), pat.rowptr, pat.colidx);
      this.patternSet = true;
    }
    const x = new Float64Array(this.sym.n);
    const info = native.refine(this.num.handle, values instanceof Float64Array ? values : Float64Array.from(values), b instanceof Float64Array ? b : Float64Array.from(b), x, 0, maxSteps, 40, tol);
    return { x, relResidual: info.relResidual, steps: info.steps };
  }
}
