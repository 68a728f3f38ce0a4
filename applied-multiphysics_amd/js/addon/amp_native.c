/*
 * Thin Node-API (version 8) addon over the C ABI of libampsolve (include/ampsolve.h). It only converts JavaScript
 * typed arrays to pointers and handles to externals; all logic is in the library. Errors from the library become
 * JavaScript exceptions carrying amp_last_error(); a failed pivot is a normal return ({ok: false}).
 */
#include <node_api.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../../include/ampsolve.h"

#define CHECK_NAPI(call)                                    \
  do {                                                      \
    if ((call) != napi_ok) {                                \
      napi_throw_error(env, NULL, "amp_native: " #call);    \
      return NULL;                                          \
    }                                                       \
  } while (0)

static napi_value throw_amp(napi_env env, int rc) {
  char code[16];
  snprintf(code, sizeof code, "%d", rc);
  napi_throw_error(env, code, amp_last_error());
  return NULL;
}

#define CHECK_AMP(call)                 \
  do {                                  \
    int rc_ = (call);                   \
    if (rc_ != AMP_OK) return throw_amp(env, rc_); \
  } while (0)

/* typed array of element type `want` -> data pointer and element count (NULL / 0 for null or undefined); 0 when the
 * value is neither, or a typed array of another element type */
static int typed(napi_env env, napi_value v, napi_typedarray_type want, void **data, size_t *len) {
  napi_valuetype t;
  *data = NULL;
  *len = 0;
  if (napi_typeof(env, v, &t) != napi_ok) return 0;
  if (t == napi_null || t == napi_undefined) return 1;
  bool is;
  if (napi_is_typedarray(env, v, &is) != napi_ok || !is) return 0;
  napi_typedarray_type tt;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &tt, len, data, &ab, &off) != napi_ok) return 0;
  return tt == want;
}
#define I32 napi_int32_array
#define F64 napi_float64_array

static napi_value throw_len(napi_env env, const char *what) {
  napi_throw_range_error(env, NULL, what);
  return NULL;
}

static void sym_finalize(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  amp_sym_free((amp_sym *)data);
}

/* a numeric handle with the sizes its arrays are checked against */
typedef struct {
  amp_num *num;
  int64_t n, nnzPattern, nnzLStored;
} num_ref;

static void num_finalize(napi_env env, void *data, void *hint) {
  (void)env;
  (void)hint;
  num_ref *r = (num_ref *)data;
  amp_num_free(r->num);
  free(r);
}

static napi_value copy_i32(napi_env env, const int32_t *src, size_t n) {
  napi_value ab, arr;
  void *dst;
  if (napi_create_arraybuffer(env, n * 4, &dst, &ab) != napi_ok) return NULL;
  if (n) memcpy(dst, src, n * 4);
  if (napi_create_typedarray(env, napi_int32_array, n, ab, 0, &arr) != napi_ok) return NULL;
  return arr;
}

static napi_value copy_f64_from_i64(napi_env env, const int64_t *src, size_t n) {
  napi_value ab, arr;
  void *dst;
  if (napi_create_arraybuffer(env, n * 8, &dst, &ab) != napi_ok) return NULL;
  for (size_t i = 0; i < n; i++) ((double *)dst)[i] = (double)src[i];
  if (napi_create_typedarray(env, napi_float64_array, n, ab, 0, &arr) != napi_ok) return NULL;
  return arr;
}

static void set_num(napi_env env, napi_value o, const char *k, double v) {
  napi_value x;
  napi_create_double(env, v, &x);
  napi_set_named_property(env, o, k, x);
}

static void set_val(napi_env env, napi_value o, const char *k, napi_value v) { napi_set_named_property(env, o, k, v); }

/* analysis object with the fields of analyze() (sparse.mjs:873-901) plus `handle` */
static napi_value sym_object(napi_env env, amp_sym *h) {
  amp_sym_info info;
  amp_sym_arrays a;
  CHECK_AMP(amp_sym_get_info(h, &info));
  CHECK_AMP(amp_sym_get_arrays(h, &a));
  napi_value o, ext, key;
  CHECK_NAPI(napi_create_object(env, &o));
  CHECK_NAPI(napi_create_external(env, h, sym_finalize, NULL, &ext));
  set_val(env, o, "handle", ext);
  const size_t n = (size_t)info.n, ns = (size_t)info.nsuper;
  set_num(env, o, "n", info.n);
  set_num(env, o, "nsuper", info.nsuper);
  set_num(env, o, "maxFront", info.maxFront);
  set_num(env, o, "nlevels", info.nlevels);
  set_num(env, o, "nnzL", (double)info.nnzL);
  set_num(env, o, "nnzA", (double)info.nnzA);
  set_num(env, o, "nnzLStored", (double)info.nnzLStored);
  set_num(env, o, "flops", info.flops);
  set_num(env, o, "fill", info.fill);
  set_num(env, o, "gemmFlops", info.gemmFlops);
  CHECK_NAPI(napi_create_string_utf8(env, info.patternKey, NAPI_AUTO_LENGTH, &key));
  set_val(env, o, "patternKey", key);
  const size_t nnzA = (size_t)info.nnzA, siLen = ns ? (size_t)a.Sp[ns] : 0;
  set_val(env, o, "perm", copy_i32(env, a.perm, n));
  set_val(env, o, "pinv", copy_i32(env, a.pinv, n));
  set_val(env, o, "parent", copy_i32(env, a.parent, n));
  set_val(env, o, "colCount", copy_i32(env, a.colCount, n));
  set_val(env, o, "lowerP", copy_i32(env, a.lowerP, n + 1));
  set_val(env, o, "lowerI", copy_i32(env, a.lowerI, nnzA));
  set_val(env, o, "lowerSrc", copy_i32(env, a.lowerSrc, nnzA));
  set_val(env, o, "superStart", copy_i32(env, a.superStart, ns + 1));
  set_val(env, o, "colSuper", copy_i32(env, a.colSuper, n));
  set_val(env, o, "superParent", copy_i32(env, a.superParent, ns));
  set_val(env, o, "childHead", copy_i32(env, a.childHead, ns));
  set_val(env, o, "childNext", copy_i32(env, a.childNext, ns));
  set_val(env, o, "Sp", copy_i32(env, a.Sp, ns + 1));
  set_val(env, o, "Si", copy_i32(env, a.Si, siLen));
  set_val(env, o, "Lxp", copy_f64_from_i64(env, a.Lxp, ns + 1));
  return o;
}

/* analyze(n, rowptr, colidx, perm | null, relax) */
static napi_value js_analyze(napi_env env, napi_callback_info cb) {
  size_t argc = 5;
  napi_value argv[5];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  int32_t n;
  bool relax = true;
  CHECK_NAPI(napi_get_value_int32(env, argv[0], &n));
  void *rp, *ci, *perm;
  size_t lrp, lci, lperm;
  if (!typed(env, argv[1], I32, &rp, &lrp) || !typed(env, argv[2], I32, &ci, &lci) ||
      !typed(env, argv[3], I32, &perm, &lperm))
    return throw_amp(env, AMP_ERR_ARG);
  if (n < 0 || !rp || lrp < (size_t)n + 1 || ((const int32_t *)rp)[n] < 0 || lci < (size_t)((const int32_t *)rp)[n])
    return throw_len(env, "analyze: rowptr / colidx shorter than the pattern");
  if (perm && lperm != (size_t)n) return throw_len(env, "analyze: perm length differs from n");
  napi_get_value_bool(env, argv[4], &relax);
  amp_sym *h = NULL;
  int32_t dummy = 0;
  CHECK_AMP(amp_analyze(n, (const int32_t *)rp, ci ? (const int32_t *)ci : &dummy, (const int32_t *)perm, relax, &h));
  return sym_object(env, h);
}

/* importSymbolic(sym, rowptr, colidx, nnzPattern): an analysis made by the reference's own analyze() */
static napi_value js_import(napi_env env, napi_callback_info cb) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  napi_value s = argv[0], v;
  amp_sym_desc d;
  memset(&d, 0, sizeof d);
  CHECK_NAPI(napi_get_named_property(env, s, "n", &v));
  CHECK_NAPI(napi_get_value_int32(env, v, &d.n));
  CHECK_NAPI(napi_get_named_property(env, s, "nsuper", &v));
  CHECK_NAPI(napi_get_value_int32(env, v, &d.nsuper));
  const char *names[] = {"perm", "parent", "colCount", "superStart", "Sp", "Si", "lowerP", "lowerI", "lowerSrc"};
  const int32_t **slots[] = {&d.perm, &d.parent, &d.colCount, &d.superStart, &d.Sp, &d.Si, &d.lowerP, &d.lowerI,
                             &d.lowerSrc};
  size_t lens[9];
  for (int i = 0; i < 9; i++) {
    void *p;
    CHECK_NAPI(napi_get_named_property(env, s, names[i], &v));
    if (!typed(env, v, I32, &p, &lens[i])) return throw_amp(env, AMP_ERR_ARG);
    *slots[i] = (const int32_t *)p;
  }
  /* lengths the library reads: n for perm / parent / colCount, nsuper + 1 for superStart / Sp, Sp[nsuper] for Si,
   * n + 1 for lowerP, lowerP[n] for lowerI / lowerSrc */
  const size_t n = d.n < 0 ? 0 : (size_t)d.n, ns = d.nsuper < 0 ? 0 : (size_t)d.nsuper;
  if (d.n < 0 || d.nsuper < 0 || lens[0] < n || lens[1] < n || (d.colCount && lens[2] < n) || lens[3] < ns + 1 ||
      lens[4] < ns + 1 || lens[6] < n + 1)
    return throw_len(env, "importSymbolic: analysis arrays shorter than n / nsuper");
  if ((ns && (d.Sp[ns] < 0 || lens[5] < (size_t)d.Sp[ns])) || d.lowerP[n] < 0 || lens[7] < (size_t)d.lowerP[n] ||
      lens[8] < (size_t)d.lowerP[n])
    return throw_len(env, "importSymbolic: Si / lowerI / lowerSrc shorter than Sp / lowerP say");
  void *rp, *ci;
  size_t lrp, lci;
  if (!typed(env, argv[1], I32, &rp, &lrp) || !typed(env, argv[2], I32, &ci, &lci)) return throw_amp(env, AMP_ERR_ARG);
  if (rp && (lrp < n + 1 || ((const int32_t *)rp)[n] < 0 || lci < (size_t)((const int32_t *)rp)[n]))
    return throw_len(env, "importSymbolic: rowptr / colidx shorter than the pattern");
  double nnz;
  CHECK_NAPI(napi_get_value_double(env, argv[3], &nnz));
  d.nnzPattern = (int64_t)nnz;
  char key[64] = {0};
  size_t klen;
  if (napi_get_named_property(env, s, "patternKey", &v) == napi_ok &&
      napi_get_value_string_utf8(env, v, key, sizeof key, &klen) == napi_ok)
    d.patternKey = key;
  amp_sym *h = NULL;
  CHECK_AMP(amp_sym_import(&d, &h));
  if (rp) { /* the CSR pattern is optional: only the refinement residual needs it */
    int rc = amp_sym_set_pattern(h, (const int32_t *)rp, (const int32_t *)ci);
    if (rc != AMP_OK) {
      amp_sym_free(h);
      return throw_amp(env, rc);
    }
  }
  napi_value ext;
  CHECK_NAPI(napi_create_external(env, h, sym_finalize, NULL, &ext));
  return ext;
}

/* numCreate(symHandle, mode, precision) */
static napi_value js_num_create(napi_env env, napi_callback_info cb) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  void *h;
  int32_t mode, prec;
  CHECK_NAPI(napi_get_value_external(env, argv[0], &h));
  CHECK_NAPI(napi_get_value_int32(env, argv[1], &mode));
  CHECK_NAPI(napi_get_value_int32(env, argv[2], &prec));
  amp_sym_info info;
  CHECK_AMP(amp_sym_get_info((amp_sym *)h, &info));
  num_ref *r = (num_ref *)calloc(1, sizeof *r);
  if (!r) return throw_len(env, "numCreate: out of host memory");
  int rc = amp_num_create((amp_sym *)h, mode, prec, &r->num);
  if (rc != AMP_OK) {
    free(r);
    return throw_amp(env, rc);
  }
  r->n = info.n;
  r->nnzPattern = info.nnzPattern;
  r->nnzLStored = info.nnzLStored;
  napi_value ext;
  if (napi_create_external(env, r, num_finalize, NULL, &ext) != napi_ok) {
    amp_num_free(r->num);
    free(r);
    napi_throw_error(env, NULL, "amp_native: napi_create_external");
    return NULL;
  }
  return ext;
}

/* factor(numHandle, values, pivotTol) -> {ok, nNeg, nPerturbed, failColumn, ms} */
static napi_value js_factor(napi_env env, napi_callback_info cb) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  void *h, *vals;
  size_t nv;
  double tol;
  CHECK_NAPI(napi_get_value_external(env, argv[0], &h));
  if (!typed(env, argv[1], F64, &vals, &nv)) return throw_amp(env, AMP_ERR_ARG);
  CHECK_NAPI(napi_get_value_double(env, argv[2], &tol));
  amp_factor_info info;
  /* the library compares the count with the pattern's (AMP_ERR_PATTERN) and reads exactly nv values */
  CHECK_AMP(amp_factor(((num_ref *)h)->num, (const double *)vals, (int64_t)nv, tol, &info));
  napi_value o, ok;
  CHECK_NAPI(napi_create_object(env, &o));
  CHECK_NAPI(napi_get_boolean(env, info.ok != 0, &ok));
  set_val(env, o, "ok", ok);
  set_num(env, o, "nNeg", info.nNeg);
  set_num(env, o, "nPerturbed", info.nPerturbed);
  set_num(env, o, "failColumn", info.failColumn);
  set_num(env, o, "ms", info.ms);
  return o;
}

/* solve(numHandle, b, out, nrhs) */
static napi_value js_solve(napi_env env, napi_callback_info cb) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  void *h, *b, *x;
  size_t lb, lx;
  int32_t q;
  CHECK_NAPI(napi_get_value_external(env, argv[0], &h));
  if (!typed(env, argv[1], F64, &b, &lb) || !typed(env, argv[2], F64, &x, &lx)) return throw_amp(env, AMP_ERR_ARG);
  CHECK_NAPI(napi_get_value_int32(env, argv[3], &q));
  const num_ref *r = (const num_ref *)h;
  if (q < 0 || lb < (size_t)(r->n * q) || lx < (size_t)(r->n * q))
    return throw_len(env, "solve: b / out shorter than n x nrhs");
  CHECK_AMP(amp_solve(r->num, (const double *)b, (double *)x, q));
  return argv[2];
}

/* refine(numHandle, values, b, out, method, maxSteps, restart, tol) -> {steps, converged, relResidual, ms} */
static napi_value js_refine(napi_env env, napi_callback_info cb) {
  size_t argc = 8;
  napi_value argv[8];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  void *h, *vals, *b, *x;
  size_t lv, lb, lx;
  amp_refine_opts o;
  CHECK_NAPI(napi_get_value_external(env, argv[0], &h));
  if (!typed(env, argv[1], F64, &vals, &lv) || !typed(env, argv[2], F64, &b, &lb) || !typed(env, argv[3], F64, &x, &lx))
    return throw_amp(env, AMP_ERR_ARG);
  const num_ref *r = (const num_ref *)h;
  if (lv != (size_t)r->nnzPattern) return throw_len(env, "refine: values length differs from the pattern's");
  if (lb < (size_t)r->n || lx < (size_t)r->n) return throw_len(env, "refine: b / out shorter than n");
  CHECK_NAPI(napi_get_value_int32(env, argv[4], &o.method));
  CHECK_NAPI(napi_get_value_int32(env, argv[5], &o.maxSteps));
  CHECK_NAPI(napi_get_value_int32(env, argv[6], &o.restart));
  CHECK_NAPI(napi_get_value_double(env, argv[7], &o.tol));
  amp_refine_info info;
  CHECK_AMP(amp_refine(r->num, (const double *)vals, (const double *)b, (double *)x, &o, &info));
  napi_value res, conv;
  CHECK_NAPI(napi_create_object(env, &res));
  set_num(env, res, "steps", info.steps);
  CHECK_NAPI(napi_get_boolean(env, info.converged != 0, &conv));
  set_val(env, res, "converged", conv);
  set_num(env, res, "relResidual", info.relResidual);
  set_num(env, res, "ms", info.ms);
  return res;
}

/* downloadFactor(numHandle, Lx, D | null) */
static napi_value js_download(napi_env env, napi_callback_info cb) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  void *h, *Lx, *D;
  size_t l1, l2;
  CHECK_NAPI(napi_get_value_external(env, argv[0], &h));
  if (!typed(env, argv[1], F64, &Lx, &l1) || !typed(env, argv[2], F64, &D, &l2)) return throw_amp(env, AMP_ERR_ARG);
  const num_ref *r = (const num_ref *)h;
  if (!Lx || l1 < (size_t)r->nnzLStored || (D && l2 < (size_t)r->n))
    return throw_len(env, "downloadFactor: Lx / D shorter than the factor");
  CHECK_AMP(amp_download_factor(r->num, (double *)Lx, (double *)D));
  return argv[1];
}

/* setPattern(symHandle, rowptr, colidx): CSR pattern for the refinement residual of an imported analysis */
static napi_value js_set_pattern(napi_env env, napi_callback_info cb) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  void *h, *rp, *ci;
  size_t l1, l2;
  CHECK_NAPI(napi_get_value_external(env, argv[0], &h));
  if (!typed(env, argv[1], I32, &rp, &l1) || !typed(env, argv[2], I32, &ci, &l2) || !rp)
    return throw_amp(env, AMP_ERR_ARG);
  amp_sym_info info;
  CHECK_AMP(amp_sym_get_info((amp_sym *)h, &info));
  const size_t n = (size_t)info.n;
  if (l1 < n + 1 || ((const int32_t *)rp)[n] < 0 || l2 < (size_t)((const int32_t *)rp)[n])
    return throw_len(env, "setPattern: rowptr / colidx shorter than the pattern");
  CHECK_AMP(amp_sym_set_pattern((amp_sym *)h, (const int32_t *)rp, (const int32_t *)ci));
  return argv[0];
}

/* fingerprint(n, rowptr, colidx) */
static napi_value js_fingerprint(napi_env env, napi_callback_info cb) {
  size_t argc = 3;
  napi_value argv[3];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  int32_t n, dummy = 0;
  void *rp, *ci;
  size_t l1, l2;
  CHECK_NAPI(napi_get_value_int32(env, argv[0], &n));
  if (!typed(env, argv[1], I32, &rp, &l1) || !typed(env, argv[2], I32, &ci, &l2)) return throw_amp(env, AMP_ERR_ARG);
  if (n < 0 || !rp || l1 < (size_t)n + 1 || ((const int32_t *)rp)[n] < 0 || l2 < (size_t)((const int32_t *)rp)[n])
    return throw_len(env, "fingerprint: rowptr / colidx shorter than the pattern");
  char out[64];
  CHECK_AMP(amp_pattern_fingerprint(n, (const int32_t *)rp, ci ? (const int32_t *)ci : &dummy, out));
  napi_value s;
  CHECK_NAPI(napi_create_string_utf8(env, out, NAPI_AUTO_LENGTH, &s));
  return s;
}

/* the pattern arguments of an ordering call (argv[0..2]); NULL after throwing */
static int order_pattern(napi_env env, napi_value *argv, const char *who, int32_t *n, const int32_t **rp,
                         const int32_t **ci) {
  void *r, *c;
  size_t l1, l2;
  if (napi_get_value_int32(env, argv[0], n) != napi_ok || !typed(env, argv[1], I32, &r, &l1) ||
      !typed(env, argv[2], I32, &c, &l2)) {
    throw_amp(env, AMP_ERR_ARG);
    return 0;
  }
  if (*n < 0 || !r || l1 < (size_t)*n + 1 || ((const int32_t *)r)[*n] < 0 || l2 < (size_t)((const int32_t *)r)[*n]) {
    char msg[96];
    snprintf(msg, sizeof msg, "%s: rowptr / colidx shorter than the pattern", who);
    throw_len(env, msg);
    return 0;
  }
  static const int32_t none = 0;
  *rp = (const int32_t *)r;
  *ci = c ? (const int32_t *)c : &none;
  return 1;
}

/* a new Int32Array of n entries and its storage */
static napi_value new_i32(napi_env env, size_t n, int32_t **data) {
  napi_value ab, arr;
  void *dst;
  if (napi_create_arraybuffer(env, n * 4, &dst, &ab) != napi_ok) return NULL;
  if (napi_create_typedarray(env, napi_int32_array, n, ab, 0, &arr) != napi_ok) return NULL;
  *data = (int32_t *)dst;
  return arr;
}

/* amdOrder(n, rowptr, colidx, denseFactor) -> Int32Array perm (new -> old) */
static napi_value js_amd_order(napi_env env, napi_callback_info cb) {
  size_t argc = 4;
  napi_value argv[4];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  int32_t n;
  const int32_t *rp, *ci;
  if (!order_pattern(env, argv, "amdOrder", &n, &rp, &ci)) return NULL;
  double dense = 10.0;
  if (argc > 3) napi_get_value_double(env, argv[3], &dense);
  int32_t *perm;
  napi_value out = new_i32(env, (size_t)n, &perm);
  if (!out) return throw_amp(env, AMP_ERR_ARG);
  if (n > 0) CHECK_AMP(amp_amd_order(n, rp, ci, dense, perm));
  return out;
}

/* ndOrder(n, rowptr, colidx, x, y, leaf) -> Int32Array perm (new -> old) */
static napi_value js_nd_order(napi_env env, napi_callback_info cb) {
  size_t argc = 6;
  napi_value argv[6];
  CHECK_NAPI(napi_get_cb_info(env, cb, &argc, argv, NULL, NULL));
  int32_t n, leaf = 48;
  const int32_t *rp, *ci;
  if (!order_pattern(env, argv, "ndOrder", &n, &rp, &ci)) return NULL;
  void *x, *y;
  size_t lx, ly;
  if (!typed(env, argv[3], F64, &x, &lx) || !typed(env, argv[4], F64, &y, &ly)) return throw_amp(env, AMP_ERR_ARG);
  if (lx != (size_t)n || ly != (size_t)n) return throw_len(env, "ndOrder: node coordinates x and y must have length n");
  if (argc > 5) napi_get_value_int32(env, argv[5], &leaf);
  int32_t *perm;
  napi_value out = new_i32(env, (size_t)n, &perm);
  if (!out) return throw_amp(env, AMP_ERR_ARG);
  if (n > 0) CHECK_AMP(amp_nd_order(n, rp, ci, (const double *)x, (const double *)y, leaf, perm));
  return out;
}

NAPI_MODULE_INIT() {
  napi_property_descriptor props[] = {
      {"analyze", NULL, js_analyze, NULL, NULL, NULL, napi_default, NULL},
      {"amdOrder", NULL, js_amd_order, NULL, NULL, NULL, napi_default, NULL},
      {"ndOrder", NULL, js_nd_order, NULL, NULL, NULL, napi_default, NULL},
      {"importSymbolic", NULL, js_import, NULL, NULL, NULL, napi_default, NULL},
      {"numCreate", NULL, js_num_create, NULL, NULL, NULL, napi_default, NULL},
      {"factor", NULL, js_factor, NULL, NULL, NULL, napi_default, NULL},
      {"solve", NULL, js_solve, NULL, NULL, NULL, napi_default, NULL},
      {"refine", NULL, js_refine, NULL, NULL, NULL, napi_default, NULL},
      {"downloadFactor", NULL, js_download, NULL, NULL, NULL, napi_default, NULL},
      {"fingerprint", NULL, js_fingerprint, NULL, NULL, NULL, napi_default, NULL},
      {"setPattern", NULL, js_set_pattern, NULL, NULL, NULL, napi_default, NULL},
  };
  napi_define_properties(env, exports, sizeof props / sizeof props[0], props);
  return exports;
}
