// Checks of the JavaScript host layer, run by tests/test_js_host.py: `--cpu` (no device: analysis, fingerprint,
// argument errors) or `--gpu` (factor / solve / LDL^T / refinement through the addon). Prints one JSON line.
import { analyze, amdOrder, ndOrder, patternFingerprint, checkSymbolic, factorNumeric, solveNumeric, solveNumericMulti, SparseCholesky, SupernodalLDLT, setDefaultRealFactor } from '../ampsolve.mjs';

/** 2D 5-point Laplacian on a k x k grid (full symmetric CSR, sorted rows). */
function laplace(k, shift = 0) {
  const n = k * k, rowptr = new Int32Array(n + 1), cols = [], vals = [];
  for (let j = 0; j < k; j++)
    for (let i = 0; i < k; i++) {
      const r = j * k + i;
      const nb = [];
      if (j > 0) nb.push([r - k, -1]);
      if (i > 0) nb.push([r - 1, -1]);
      nb.push([r, 4 - shift]);
      if (i + 1 < k) nb.push([r + 1, -1]);
      if (j + 1 < k) nb.push([r + k, -1]);
      for (const [c, v] of nb) { cols.push(c); vals.push(v); }
      rowptr[r + 1] = cols.length;
    }
  return { pat: { n, rowptr, colidx: Int32Array.from(cols), nnz: cols.length }, val: Float64Array.from(vals) };
}

function residual(pat, val, x, b) {
  let rr = 0, bb = 0;
  for (let i = 0; i < pat.n; i++) {
    let s = 0;
    for (let p = pat.rowptr[i]; p < pat.rowptr[i + 1]; p++) s += val[p] * x[pat.colidx[p]];
    rr += (s - b[i]) ** 2;
    bb += b[i] ** 2;
  }
  return Math.sqrt(rr / bb);
}

const mode = process.argv[2];
const out = {};
const { pat, val } = laplace(30);
const sym = analyze(pat);
out.n = sym.n;
out.nsuper = sym.nsuper;
out.perm = Array.from(sym.perm);
out.Sp = Array.from(sym.Sp);
out.Lxp_last = sym.Lxp[sym.nsuper];
out.patternKey = sym.patternKey;
out.fingerprint = patternFingerprint(pat);
checkSymbolic(sym, pat);
const errors = [];
try { setDefaultRealFactor('fast'); } catch (e) { errors.push(e.message); }
try { factorNumeric(sym, val.subarray(1)); } catch (e) { errors.push(e.message); }
const other = { ...pat, colidx: pat.colidx.slice() };
other.colidx[1] = (other.colidx[1] + 1) % pat.n;
try { new SparseCholesky(other, null, { sym }); } catch (e) { errors.push(e.message); }
out.errors = errors;
// the addon checks element types and lengths against the sizes the library will read
const addonErrors = [];
try { patternFingerprint({ n: pat.n, rowptr: Float64Array.from(pat.rowptr), colidx: pat.colidx }); } catch (e) { addonErrors.push(e.message); }
try { patternFingerprint({ n: pat.n, rowptr: pat.rowptr, colidx: pat.colidx.subarray(0, 10) }); } catch (e) { addonErrors.push(e.message); }
try { analyze({ n: pat.n + 5, rowptr: pat.rowptr, colidx: pat.colidx }); } catch (e) { addonErrors.push(e.message); }
out.addon_errors = addonErrors;
// orderings: ndOrder from the grid coordinates, analyze with ordering 'nd' (and without coordinates), timing split
const gx = Float64Array.from({ length: pat.n }, (_, r) => r % 30), gy = Float64Array.from({ length: pat.n }, (_, r) => Math.floor(r / 30));
out.nd = Array.from(ndOrder(pat, gx, gy));
out.amd = Array.from(amdOrder(pat));
const symNd = analyze(pat, { ordering: 'nd', x: gx, y: gy });
out.nd_perm = Array.from(symNd.perm);
out.nd_timing = Object.keys(symNd.timing).sort();
try { analyze(pat, { ordering: 'nd' }); } catch (e) { out.nd_error = e.message; }
out.kinds = ['hip', 'hip-f32', 'hip-bf16x3'].map((k) => setDefaultRealFactor(k));
setDefaultRealFactor('hip');
if (mode === '--gpu') {
  const b = Float64Array.from({ length: pat.n }, (_, i) => Math.sin(0.3 * i) + 0.1);
  const ch = new SparseCholesky(pat);
  out.ok = ch.factor(val);
  const x = ch.solve(b);
  out.relres = residual(pat, val, x, b);
  const xs = solveNumericMulti(ch.num, [b, b.map((v) => 2 * v)]);
  out.multi = Math.max(...xs[1].map((v, i) => Math.abs(v - 2 * x[i])));
  try { ch.solve(b.subarray(0, pat.n - 1)); } catch (e) { out.short_b = e.message; }
  const bad = new SparseCholesky(pat);
  out.indefinite_ok = bad.factor(laplace(30, 6).val);
  try { bad.solve(b); } catch (e) { out.solve_error = e.message; }
  const ind = laplace(30, 1.3);
  const f = new SupernodalLDLT(sym);
  const r = f.factor(ind.val);
  out.ldlt = { ok: r.ok, nNeg: r.nNeg, nPerturbed: r.nPerturbed };
  const ref = f.solveRefined(pat, ind.val, b, { maxSteps: 5 });
  out.ldlt_relres = ref.relResidual;
  // a reference-shaped analysis object (no native handle) is handed over unchanged
  const plain = Object.assign({}, sym);
  delete plain.handle;
  delete plain.nnzPattern;
  const num = factorNumeric(plain, val, null, pat);
  out.imported_relres = residual(pat, val, solveNumeric(num, b), b);
  const f32 = new SupernodalLDLT(sym, { factor: 'hip-f32' });
  f32.factor(val);
  out.f32_refined = f32.solveRefined(pat, val, b, { maxSteps: 10, tol: 1e-13 }).relResidual;
}
console.log(JSON.stringify(out));
