/*
 * Oracle (test infrastructure only): numeric supernodal multifrontal Cholesky / LDL^T, triangular solves, iterative
 * refinement and GMRES, restated from packages/core/src/la/{sparse,ldlt,gmres-real}.mjs and
 * wasm/{front-real,solve-real}.mjs. Compiled with -ffp-contract=off: no fused multiply-add anywhere.
 */
#include "oracle.h"

#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/*
 * Dense partial factorization of the m x m column-major front F with nc pivot columns, in place.
 * Per entry (row i, target t) the operation sequence of front-real.mjs:9-16 / sparse.mjs:992-1098 / ldlt.mjs:492-535:
 *   for c = 0, 4, ... with c + 3 < cmax:  F[i,t] = F[i,t] - (((F[i,c] w_c + F[i,c+1] w_c+1) + F[i,c+2] w_c+2)
 *                                                             + F[i,c+3] w_c+3)
 *   then for the remaining c < cmax, skipped when w_c == 0:  F[i,t] = F[i,t] - F[i,c] w_c
 * cmax = t for panel columns and nc for the Schur block; w_c = F[t,c] (Cholesky) or F[t,c] * D[f+c] (LDL^T).
 * Target columns are processed four at a time sharing the source loads (as front-real.mjs:259-321 does), which
 * leaves every entry's sequence unchanged.
 * Pivots: Cholesky (front-real.mjs:367-375) F[k,k] = sqrt(d), F[i,k] *= 1/sqrt(d), fail unless d > 0;
 * LDL^T (front-real.mjs:376-388, ldlt.mjs:505-519) static perturbation to +-tau, nNeg / nPert counts, D[f+k] = d,
 * F[k,k] = 1, F[i,k] *= 1/d, fail on a non-finite small pivot.
 */
typedef struct {
  int ldlt;
  const double *D; /* D + f, LDL^T only */
  double tau;
  int32_t nNeg, nPert;
} front_ctx;

static inline double weight(const double *F, int64_t m, int32_t t, int32_t c, const front_ctx *x) {
  const double v = F[t + c * m];
  return x->ldlt ? v * x->D[c] : v;
}

/* Rows per parallel chunk of a left-looking group update, and the work (rows x groups) below which it stays on one
 * thread. Splitting the rows of a target column among threads leaves every entry's operation sequence unchanged (the
 * groups are applied to each row in increasing order), so the result is bitwise the same for any thread count. */
#define ROW_CHUNK 512
#define PAR_WORK 65536

/* Groups g < ng into target columns t0 .. t0 + 3, rows [r0, r1) with r0 >= t0 + 3 (rows below every target's
 * diagonal); w[jj][4 g + q] are the weights of target t0 + jj. */
static void groups4_rows(double *F, int64_t m, int32_t t0, int32_t ng, const double *w, int64_t r0, int64_t r1) {
  double *T0 = F + (int64_t)t0 * m, *T1 = T0 + m, *T2 = T1 + m, *T3 = T2 + m;
  const int64_t W = 4 * (int64_t)ng;
  for (int32_t g = 0; g < ng; g++) {
    const int32_t c = 4 * g;
    const double *S0 = F + c * m, *S1 = S0 + m, *S2 = S1 + m, *S3 = S2 + m;
    const double a0 = w[c], a1 = w[c + 1], a2 = w[c + 2], a3 = w[c + 3];
    const double b0 = w[W + c], b1 = w[W + c + 1], b2 = w[W + c + 2], b3 = w[W + c + 3];
    const double c0 = w[2 * W + c], c1 = w[2 * W + c + 1], c2 = w[2 * W + c + 2], c3 = w[2 * W + c + 3];
    const double d0 = w[3 * W + c], d1 = w[3 * W + c + 1], d2 = w[3 * W + c + 2], d3 = w[3 * W + c + 3];
    for (int64_t i = r0; i < r1; i++) {
      const double s0 = S0[i], s1 = S1[i], s2 = S2[i], s3 = S3[i];
      T0[i] = T0[i] - (((s0 * a0 + s1 * a1) + s2 * a2) + s3 * a3);
      T1[i] = T1[i] - (((s0 * b0 + s1 * b1) + s2 * b2) + s3 * b3);
      T2[i] = T2[i] - (((s0 * c0 + s1 * c1) + s2 * c2) + s3 * c3);
      T3[i] = T3[i] - (((s0 * d0 + s1 * d1) + s2 * d2) + s3 * d3);
    }
  }
}

/* Groups g < ng into one target column t, rows [r0, r1) (r0 >= t). */
static void groups1_rows(double *F, int64_t m, int32_t t, int32_t ng, const double *w, int64_t r0, int64_t r1) {
  double *T = F + (int64_t)t * m;
  for (int32_t g = 0; g < ng; g++) {
    const int32_t c = 4 * g;
    const double *S0 = F + c * m, *S1 = S0 + m, *S2 = S1 + m, *S3 = S2 + m;
    const double w0 = w[c], w1 = w[c + 1], w2 = w[c + 2], w3 = w[c + 3];
    for (int64_t i = r0; i < r1; i++) T[i] = T[i] - (((S0[i] * w0 + S1[i] * w1) + S2[i] * w2) + S3[i] * w3);
  }
}

/* Apply groups g < ng to target columns t0..t0+nt-1 (nt <= 4); target t updates rows [t, m). The weights are read
 * from the source columns' rows t0 .. t0 + nt - 1, which no target update writes. par: rows may be split among the
 * threads (not inside an enclosing parallel loop). */
static void apply_groups(double *F, int64_t m, int32_t t0, int32_t nt, int32_t ng, const front_ctx *x, int par) {
  if (ng == 0) return;
  double wbuf[4 * 4 * 64];
  double *w = 4 * ng <= 4 * 64 ? wbuf : malloc(sizeof(double) * 16 * (size_t)ng);
  const int64_t W = 4 * (int64_t)ng;
  for (int32_t jj = 0; jj < nt; jj++)
    for (int32_t c = 0; c < 4 * ng; c++) w[jj * W + c] = weight(F, m, t0 + jj, c, x);
  if (nt == 4) {
    /* head rows t0 .. t0+2 (targets whose diagonal is at or above the row) */
    for (int32_t jj = 0; jj < 3; jj++) groups1_rows(F, m, t0 + jj, ng, w + jj * W, t0 + jj, t0 + 3);
    const int64_t r0 = t0 + 3, rows = m - r0;
    if (par && rows * ng >= PAR_WORK && rows > ROW_CHUNK) {
      const int64_t nchunk = (rows + ROW_CHUNK - 1) / ROW_CHUNK;
#pragma omp parallel for schedule(static)
      for (int64_t k = 0; k < nchunk; k++) {
        const int64_t a = r0 + k * ROW_CHUNK, b = a + ROW_CHUNK < m ? a + ROW_CHUNK : m;
        groups4_rows(F, m, t0, ng, w, a, b);
      }
    } else if (rows > 0)
      groups4_rows(F, m, t0, ng, w, r0, m);
  } else {
    for (int32_t jj = 0; jj < nt; jj++) {
      const int64_t r0 = t0 + jj, rows = m - r0;
      if (par && rows * ng >= PAR_WORK && rows > ROW_CHUNK) {
        const int64_t nchunk = (rows + ROW_CHUNK - 1) / ROW_CHUNK;
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < nchunk; k++) {
          const int64_t a = r0 + k * ROW_CHUNK, b = a + ROW_CHUNK < m ? a + ROW_CHUNK : m;
          groups1_rows(F, m, t0 + jj, ng, w + jj * W, a, b);
        }
      } else
        groups1_rows(F, m, t0 + jj, ng, w + jj * W, r0, m);
    }
  }
  if (w != wbuf) free(w);
}

/* Single source column c into target t, rows [t, m), skipped when the weight is zero. */
static void apply_single(double *F, int64_t m, int32_t t, int32_t c, const front_ctx *x) {
  const double w = weight(F, m, t, c, x);
  if (w == 0) return;
  double *T = F + (int64_t)t * m;
  const double *S = F + (int64_t)c * m;
  for (int64_t i = t; i < m; i++) T[i] = T[i] - S[i] * w;
}

static int pivot(double *F, int64_t m, int32_t k, front_ctx *x, double *Dout) {
  double *K = F + (int64_t)k * m;
  double d = K[k];
  if (!x->ldlt) {
    if (!(d > 0)) return 0;
    const double dk = sqrt(d);
    K[k] = dk;
    const double inv = 1 / dk;
    for (int64_t i = k + 1; i < m; i++) K[i] *= inv;
    return 1;
  }
  if (!(fabs(d) >= x->tau)) {
    if (!isfinite(d)) return 0;
    d = d < 0 ? -x->tau : x->tau;
    x->nPert++;
  }
  if (d < 0) x->nNeg++;
  Dout[k] = d;
  K[k] = 1;
  const double inv = 1 / d;
  for (int64_t i = k + 1; i < m; i++) K[i] *= inv;
  return 1;
}

static int front_factor(double *F, int32_t m, int32_t nc, front_ctx *x, double *Dout) {
  const int64_t M = m;
  int32_t k = 0;
  for (; k + 3 < nc; k += 4) {
    apply_groups(F, M, k, 4, k >> 2, x, 1);
    for (int32_t jj = 0; jj < 4; jj++) {
      const int32_t t = k + jj;
      for (int32_t c = k; c < t; c++) apply_single(F, M, t, c, x);
      if (!pivot(F, M, t, x, Dout)) return 0;
    }
  }
  for (; k < nc; k++) {
    const int32_t ng = k >> 2;
    apply_groups(F, M, k, 1, ng, x, 1);
    for (int32_t c = 4 * ng; c < k; c++) apply_single(F, M, k, c, x);
    if (!pivot(F, M, k, x, Dout)) return 0;
  }
  /* Schur block: target column groups are independent (each reads only the factored columns) */
  const int32_t ng = nc >> 2;
  const int32_t nj = (m - nc + 3) / 4;
  const int par = (int64_t)(m - nc) * (m - nc) / 2 * (ng + 1) >= 4 * PAR_WORK;
#pragma omp parallel for schedule(dynamic, 4) if (par)
  for (int32_t q = 0; q < nj; q++) {
    const int32_t j = nc + 4 * q;
    const int32_t nt = m - j < 4 ? m - j : 4;
    apply_groups(F, M, j, nt, ng, x, 0);
    for (int32_t c = 4 * ng; c < nc; c++)
      for (int32_t jj = 0; jj < nt; jj++) apply_single(F, M, j + jj, c, x);
  }
  return 1;
}

/* The multifrontal driver shared by factorNumeric (sparse.mjs:941-1125) and SupernodalLDLT.factor (ldlt.mjs:450-562). */
static int multifrontal(const orc_sym *S, const double *values, double *Lx, front_ctx *x, double *D, int32_t sFirst,
                        int32_t sLast) {
  const int32_t n = S->n, nsuper = S->nsuper;
  if (n == 0) return 1;
  const int64_t mf = S->maxFront;
  double *F = xcalloc((size_t)(mf * mf), 8);
  int32_t *rel = xcalloc(n, 4), *stackDim = xcalloc(nsuper, 4);
  int64_t *stackPos = xcalloc(nsuper + 1, 8);
  int32_t nPos = 0;
  int64_t cap = 1024, top = 0;
  double *stack = xcalloc(cap, 8);
  int ok = 1;
  for (int32_t s = sFirst; s <= sLast && ok; s++) {
    const int32_t f = S->superStart[s], nc = S->superStart[s + 1] - f;
    const int32_t r0 = S->Sp[s], m = S->Sp[s + 1] - r0;
    const int64_t M = m;
    for (int32_t r = 0; r < m; r++) rel[S->Si[r0 + r]] = r;
#pragma omp parallel for schedule(dynamic, 16) if (M * M >= 8 * PAR_WORK)
    for (int64_t b = 0; b < M; b++) memset(F + b * M + b, 0, (size_t)(M - b) * 8); /* zero the lower part */
    for (int32_t c = 0; c < nc; c++) { /* assemble A's own columns (+= onto +0) */
      const int32_t j = f + c;
      double *Fc = F + c * M;
      for (int32_t p = S->lowerP[j]; p < S->lowerP[j + 1]; p++) Fc[rel[S->lowerI[p]]] += values[S->lowerSrc[p]];
    }
    for (int32_t ch = S->childHead[s]; ch != -1; ch = S->childNext[ch]) { /* extend-add, children popped in order */
      const int64_t off = stackPos[--nPos];
      const int64_t mu = stackDim[ch];
      const int32_t cr0 = S->Sp[ch] + (S->superStart[ch + 1] - S->superStart[ch]);
      /* the relative map is injective: each front entry receives at most one add from this child, so columns of
       * the child may go to different threads */
#pragma omp parallel for schedule(dynamic, 16) if (mu * mu >= 8 * PAR_WORK)
      for (int64_t b = 0; b < mu; b++) {
        double *Fb = F + (int64_t)rel[S->Si[cr0 + b]] * M;
        const double *ub = stack + off + b * mu;
        for (int64_t a = b; a < mu; a++) Fb[rel[S->Si[cr0 + a]]] += ub[a];
      }
      top = off;
    }
    x->D = D ? D + f : NULL;
    if (!front_factor(F, m, nc, x, D ? D + f : NULL)) {
      ok = 0;
      break;
    }
    memcpy(Lx + S->Lxp[s], F, (size_t)(M * nc) * 8); /* store the m x nc block */
    const int64_t mu = m - nc;
    stackDim[s] = (int32_t)mu;
    if (mu > 0) {
      const int64_t need = top + mu * mu;
      if (need > cap) {
        int64_t bigger = need > 2 * cap ? need : 2 * cap;
        stack = realloc(stack, (size_t)bigger * 8);
        cap = bigger;
      }
#pragma omp parallel for schedule(dynamic, 16) if (mu * mu >= 8 * PAR_WORK)
      for (int64_t b = 0; b < mu; b++) {
        const double *src = F + (nc + b) * M + nc;
        double *dst = stack + top + b * mu;
        for (int64_t a = b; a < mu; a++) dst[a] = src[a];
      }
      stackPos[nPos++] = top;
      top = need;
    } else
      stackPos[nPos++] = top;
  }
  free(F);
  free(rel);
  free(stackDim);
  free(stackPos);
  free(stack);
  return ok;
}

void orc_set_threads(int32_t k) { omp_set_num_threads(k > 0 ? k : 1); }
int32_t orc_get_threads(void) { return omp_get_max_threads(); }

int orc_chol_factor(const orc_sym *S, const double *values, double *Lx) {
  front_ctx x = {0, NULL, 0, 0, 0};
  return multifrontal(S, values, Lx, &x, NULL, 0, S->nsuper - 1);
}

int orc_chol_factor_subtree(const orc_sym *S, const double *values, int32_t root, double *Lx, double *flops) {
  /* the subtree of `root` is the contiguous postorder range [first, root] */
  int32_t first = root;
  while (S->childHead[first] != -1) {
    int32_t c = S->childHead[first], lowest = c;
    for (; c != -1; c = S->childNext[c])
      if (c < lowest) lowest = c;
    first = lowest;
  }
  double f = 0;
  for (int32_t j = S->superStart[first]; j < S->superStart[root + 1]; j++) f += (double)S->colCount[j] * S->colCount[j];
  *flops = f;
  front_ctx x = {0, NULL, 0, 0, 0};
  return multifrontal(S, values, Lx, &x, NULL, first, root);
}

/* diagScale (ldlt.mjs:419-424): max |a_jj| over the diagonal, 1 when zero */
static double diag_scale(const orc_sym *S, const double *values) {
  double amax = 0;
  for (int32_t j = 0; j < S->n; j++)
    for (int32_t p = S->lowerP[j]; p < S->lowerP[j + 1]; p++)
      if (S->lowerI[p] == j) amax = fmax(amax, fabs(values[S->lowerSrc[p]]));
  return amax > 0 ? amax : 1;
}

int orc_ldlt_factor(const orc_sym *S, const double *values, double pivotTol, double *Lx, double *D, int32_t *nNeg,
                    int32_t *nPert) {
  front_ctx x = {1, NULL, pivotTol * diag_scale(S, values), 0, 0};
  const int ok = multifrontal(S, values, Lx, &x, D, 0, S->nsuper - 1);
  *nNeg = x.nNeg;
  *nPert = x.nPert;
  return ok;
}

/* ---- Cholesky solves, sparse.mjs:1161-1305 ---------------------------------------------------------------------- */
static double dot_below(const double *L, const double *t, int32_t mu) { /* dotBelow, sparse.mjs:1179-1193 */
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int32_t r = 0;
  for (; r + 3 < mu; r += 4) {
    s0 += L[r] * t[r];
    s1 += L[r + 1] * t[r + 1];
    s2 += L[r + 2] * t[r + 2];
    s3 += L[r + 3] * t[r + 3];
  }
  for (; r < mu; r++) s0 += L[r] * t[r];
  return (s0 + s1) + (s2 + s3);
}

static void chol_solve_perm(const orc_sym *S, const double *Lx, double *y, double *t) {
  for (int32_t s = 0; s < S->nsuper; s++) {
    const int32_t f = S->superStart[s], nc = S->superStart[s + 1] - f, r0 = S->Sp[s], m = S->Sp[s + 1] - r0;
    const int32_t mu = m - nc;
    const double *L = Lx + S->Lxp[s];
    for (int32_t c = 0; c < nc; c++) {
      const double *Lc = L + (int64_t)c * m;
      const double yj = (y[f + c] /= Lc[c]);
      for (int32_t r = c + 1; r < nc; r++) y[f + r] -= Lc[r] * yj;
    }
    if (mu == 0) continue;
    for (int32_t r = 0; r < mu; r++) t[r] = 0; /* belowAccumulate, sparse.mjs:1161-1176 */
    int32_t c = 0;
    for (; c + 1 < nc; c += 2) {
      const double *La = L + (int64_t)c * m + nc, *Lb = La + m;
      const double ya = y[f + c], yb = y[f + c + 1];
      for (int32_t r = 0; r < mu; r++) t[r] += La[r] * ya + Lb[r] * yb;
    }
    if (c < nc) {
      const double *La = L + (int64_t)c * m + nc;
      const double ya = y[f + c];
      for (int32_t r = 0; r < mu; r++) t[r] += La[r] * ya;
    }
    const int32_t rb = r0 + nc;
    for (int32_t r = 0; r < mu; r++) y[S->Si[rb + r]] -= t[r];
  }
  for (int32_t s = S->nsuper - 1; s >= 0; s--) {
    const int32_t f = S->superStart[s], nc = S->superStart[s + 1] - f, r0 = S->Sp[s], m = S->Sp[s + 1] - r0;
    const int32_t mu = m - nc, rb = r0 + nc;
    const double *L = Lx + S->Lxp[s];
    for (int32_t r = 0; r < mu; r++) t[r] = y[S->Si[rb + r]];
    for (int32_t c = nc - 1; c >= 0; c--) {
      const double *Lc = L + (int64_t)c * m;
      double w = y[f + c];
      for (int32_t r = c + 1; r < nc; r++) w -= Lc[r] * y[f + r];
      y[f + c] = (w - dot_below(Lc + nc, t, mu)) / Lc[c];
    }
  }
}

void orc_chol_solve(const orc_sym *S, const double *Lx, const double *b, double *x) {
  const int32_t n = S->n;
  double *y = xcalloc(n, 8), *t = xcalloc(S->maxFront, 8);
  for (int32_t i = 0; i < n; i++) y[i] = b[S->perm[i]];
  chol_solve_perm(S, Lx, y, t);
  for (int32_t i = 0; i < n; i++) x[S->perm[i]] = y[i];
  free(y);
  free(t);
}

void orc_chol_solve_multi(const orc_sym *S, const double *Lx, const double *b, int32_t q, double *x) {
  for (int32_t k = 0; k < q; k++) orc_chol_solve(S, Lx, b + (int64_t)k * S->n, x + (int64_t)k * S->n);
}

/* ---- LDL^T solves, ldlt.mjs:565-643 ---------------------------------------------------------------------------- */
void orc_ldlt_solve(const orc_sym *S, const double *Lx, const double *D, const double *b, double *x) {
  const int32_t n = S->n;
  double *y = xcalloc(n, 8);
  for (int32_t i = 0; i < n; i++) y[i] = b[S->perm[i]];
  for (int32_t s = 0; s < S->nsuper; s++) {
    const int32_t f = S->superStart[s], nc = S->superStart[s + 1] - f, r0 = S->Sp[s], m = S->Sp[s + 1] - r0;
    const double *L = Lx + S->Lxp[s];
    for (int32_t c = 0; c < nc; c++) {
      const double *Lc = L + (int64_t)c * m;
      const double yj = y[f + c];
      if (yj == 0) continue;
      for (int32_t r = c + 1; r < nc; r++) y[f + r] -= Lc[r] * yj;
      for (int32_t r = nc; r < m; r++) y[S->Si[r0 + r]] -= Lc[r] * yj;
    }
  }
  for (int32_t j = 0; j < n; j++) y[j] /= D[j];
  for (int32_t s = S->nsuper - 1; s >= 0; s--) {
    const int32_t f = S->superStart[s], nc = S->superStart[s + 1] - f, r0 = S->Sp[s], m = S->Sp[s + 1] - r0;
    const double *L = Lx + S->Lxp[s];
    for (int32_t c = nc - 1; c >= 0; c--) {
      const double *Lc = L + (int64_t)c * m;
      double v = y[f + c];
      for (int32_t r = c + 1; r < nc; r++) v -= Lc[r] * y[f + r];
      for (int32_t r = nc; r < m; r++) v -= Lc[r] * y[S->Si[r0 + r]];
      y[f + c] = v;
    }
  }
  for (int32_t i = 0; i < n; i++) x[S->perm[i]] = y[i];
  free(y);
}

void orc_ldlt_solve_many(const orc_sym *S, const double *Lx, const double *D, const double *b, int32_t q, double *x) {
  const int32_t n = S->n;
  double *y = xcalloc((size_t)n * q, 8), *v = xcalloc(q, 8);
  for (int32_t i = 0; i < n; i++)
    for (int32_t k = 0; k < q; k++) y[(int64_t)i * q + k] = b[(int64_t)k * n + S->perm[i]];
  for (int32_t s = 0; s < S->nsuper; s++) {
    const int32_t f = S->superStart[s], nc = S->superStart[s + 1] - f, r0 = S->Sp[s], m = S->Sp[s + 1] - r0;
    const double *L = Lx + S->Lxp[s];
    for (int32_t c = 0; c < nc; c++) {
      const double *Lc = L + (int64_t)c * m;
      const int64_t jq = (int64_t)(f + c) * q;
      for (int32_t r = c + 1; r < m; r++) {
        const double l = Lc[r];
        if (l == 0) continue;
        const int64_t iq = (int64_t)S->Si[r0 + r] * q;
        for (int32_t k = 0; k < q; k++) y[iq + k] -= l * y[jq + k];
      }
    }
  }
  for (int32_t j = 0; j < n; j++) {
    const double dj = 1 / D[j];
    for (int32_t k = 0; k < q; k++) y[(int64_t)j * q + k] *= dj;
  }
  for (int32_t s = S->nsuper - 1; s >= 0; s--) {
    const int32_t f = S->superStart[s], nc = S->superStart[s + 1] - f, r0 = S->Sp[s], m = S->Sp[s + 1] - r0;
    const double *L = Lx + S->Lxp[s];
    for (int32_t c = nc - 1; c >= 0; c--) {
      const double *Lc = L + (int64_t)c * m;
      const int64_t jq = (int64_t)(f + c) * q;
      for (int32_t k = 0; k < q; k++) v[k] = y[jq + k];
      for (int32_t r = c + 1; r < m; r++) {
        const double l = Lc[r];
        if (l == 0) continue;
        const int64_t iq = (int64_t)S->Si[r0 + r] * q;
        for (int32_t k = 0; k < q; k++) v[k] -= l * y[iq + k];
      }
      for (int32_t k = 0; k < q; k++) y[jq + k] = v[k];
    }
  }
  for (int32_t i = 0; i < n; i++)
    for (int32_t k = 0; k < q; k++) x[(int64_t)k * n + S->perm[i]] = y[(int64_t)i * q + k];
  free(y);
  free(v);
}

int32_t orc_ldlt_solve_refined(const orc_sym *S, const double *Lx, const double *D, const int32_t *rowptr,
                               const int32_t *colidx, const double *values, const double *b, int32_t maxSteps,
                               double tol, double *x, double *relResidual) {
  const int32_t n = S->n; /* ldlt.mjs:253-275 */
  double *r = xcalloc(n, 8), *dx = xcalloc(n, 8);
  orc_ldlt_solve(S, Lx, D, b, x);
  double bn = 0;
  for (int32_t i = 0; i < n; i++) bn += b[i] * b[i];
  bn = sqrt(bn);
  if (bn == 0 || bn != bn) bn = 1; /* `Math.sqrt(bn) || 1` */
  double rel = INFINITY;
  int32_t steps = 0;
  for (; steps <= maxSteps; steps++) {
    double rn = 0;
    for (int32_t i = 0; i < n; i++) {
      double s = b[i];
      for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) s -= values[p] * x[colidx[p]];
      r[i] = s;
      rn += s * s;
    }
    rel = sqrt(rn) / bn;
    if (rel < tol || steps == maxSteps) break;
    orc_ldlt_solve(S, Lx, D, r, dx);
    for (int32_t i = 0; i < n; i++) x[i] += dx[i];
  }
  *relResidual = rel;
  free(r);
  free(dx);
  return steps;
}

void orc_csr_matvec(int32_t n, const int32_t *rowptr, const int32_t *colidx, const double *val, const double *x,
                    double *y) {
  for (int32_t i = 0; i < n; i++) {
    double s = 0;
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) s += val[p] * x[colidx[p]];
    y[i] = s;
  }
}

/* ---- gmres, gmres-real.mjs:31-104 ------------------------------------------------------------------------------- */
static double norm2(const double *v, int32_t n) {
  double s = 0;
  for (int32_t i = 0; i < n; i++) s += v[i] * v[i];
  return sqrt(s);
}

/* r = b - A x, returns |r| (the `residual` closure of gmres-real.mjs:45-49) */
static double gm_residual(int32_t n, orc_apply_fn matvec, void *mctx, const double *b, const double *x, double *r) {
  matvec(mctx, x, r);
  for (int32_t i = 0; i < n; i++) r[i] = b[i] - r[i];
  return norm2(r, n);
}

void orc_gmres(int32_t n, orc_apply_fn matvec, void *mctx, orc_apply_fn precond, void *pctx, const double *b,
               double tol, int32_t restart, int32_t maxIter, const double *x0, double *x, orc_gmres_info *info) {
  const int32_t m = restart;
  if (x0)
    memcpy(x, x0, (size_t)n * 8);
  else
    memset(x, 0, (size_t)n * 8);
  double *V = xcalloc((size_t)(m + 1) * n, 8), *Z = xcalloc((size_t)m * n, 8);
  double *r = xcalloc(n, 8), *w = xcalloc(n, 8), *z = xcalloc(n, 8);
  double *H = xcalloc((size_t)(m + 1) * m, 8), *cs = xcalloc(m, 8), *sn = xcalloc(m, 8), *g = xcalloc(m + 1, 8);
  double *yv = xcalloc(m, 8);
  double bnorm = norm2(b, n);
  if (bnorm == 0 || bnorm != bnorm) bnorm = 1;
  int32_t iters = 0;
  double relres, beta;

  if (x0)
    beta = gm_residual(n, matvec, mctx, b, x, r);
  else {
    memcpy(r, b, (size_t)n * 8);
    beta = bnorm;
  }
  relres = beta / bnorm;
  while (relres > tol && iters < maxIter) {
    for (int32_t i = 0; i < n; i++) V[i] = r[i] / beta;
    memset(g, 0, (size_t)(m + 1) * 8);
    g[0] = beta;
    int32_t k = 0;
    for (; k < m && iters < maxIter; k++) {
      iters++;
      double *Vk = V + (int64_t)k * n, *Zk = Z + (int64_t)k * n;
      precond(pctx, Vk, Zk);
      matvec(mctx, Zk, w);
      for (int32_t j = 0; j <= k; j++) {
        const double *vj = V + (int64_t)j * n;
        double h = 0;
        for (int32_t i = 0; i < n; i++) h += w[i] * vj[i];
        H[j * m + k] = h;
        for (int32_t i = 0; i < n; i++) w[i] -= h * vj[i];
      }
      const double hn = norm2(w, n);
      H[(k + 1) * m + k] = hn;
      if (hn > 0) {
        double *vk = V + (int64_t)(k + 1) * n;
        for (int32_t i = 0; i < n; i++) vk[i] = w[i] / hn;
      }
      for (int32_t j = 0; j < k; j++) {
        const double a = H[j * m + k], c = H[(j + 1) * m + k];
        H[j * m + k] = cs[j] * a + sn[j] * c;
        H[(j + 1) * m + k] = -sn[j] * a + cs[j] * c;
      }
      const double a = H[k * m + k], c = H[(k + 1) * m + k];
      double d = hypot(a, c);
      if (d == 0 || d != d) d = 1e-300;
      cs[k] = a / d;
      sn[k] = c / d;
      H[k * m + k] = d;
      H[(k + 1) * m + k] = 0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      relres = fabs(g[k + 1]) / bnorm;
      if (relres <= tol || hn == 0) {
        k++;
        break;
      }
    }
    for (int32_t i = k - 1; i >= 0; i--) {
      double s = g[i];
      for (int32_t j = i + 1; j < k; j++) s -= H[i * m + j] * yv[j];
      yv[i] = s / H[i * m + i];
    }
    memset(z, 0, (size_t)n * 8);
    for (int32_t j = 0; j < k; j++) {
      const double *zj = Z + (int64_t)j * n;
      const double yj = yv[j];
      for (int32_t i = 0; i < n; i++) z[i] += yj * zj[i];
    }
    for (int32_t i = 0; i < n; i++) x[i] += z[i];
    if (relres <= tol || iters >= maxIter) break;
    beta = gm_residual(n, matvec, mctx, b, x, r);
    relres = beta / bnorm;
  }
  const double trueRes = gm_residual(n, matvec, mctx, b, x, r) / bnorm;
  info->iterations = iters;
  info->relres = trueRes;
  info->converged = trueRes <= fmax(tol * 10, 1e-14);
  free(V);
  free(Z);
  free(r);
  free(w);
  free(z);
  free(H);
  free(cs);
  free(sn);
  free(g);
  free(yv);
}
