/*
 * Oracle (test infrastructure only): complex-symmetric supernodal LDL^T (A = A^T complex, no conjugation) restated from
 * ComplexSupernodalLDLT of packages/core/src/la/ldlt.mjs:648-866: the same multifrontal driver as the real factor
 * (fronts assembled from A's columns and the children's update matrices, children in childHead order), a
 * left-looking update per column in the order of ldlt.mjs:705-720 (skipped when the complex weight is zero), complex
 * pivots perturbed to magnitude tau with their phase kept (:724-734), the solve of :780-829 and the complex iterative
 * refinement of :838-862. Real and imaginary parts in separate arrays, every complex product and quotient in the
 * component order the reference writes.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static void *zc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

int orc_cldlt_factor(const orc_sym *S, const double *re, const double *im, double pivotTol, double *Lr, double *Li,
                     double *Dr, double *Di, int32_t *nPertOut) {
  const int32_t n = S->n, nsuper = S->nsuper;
  *nPertOut = 0;
  if (n == 0) return 1;
  /* diagScale with |a_jj| = hypot(re, im) (ldlt.mjs:419-424, :675) */
  double amax = 0;
  for (int32_t j = 0; j < n; j++)
    for (int32_t p = S->lowerP[j]; p < S->lowerP[j + 1]; p++)
      if (S->lowerI[p] == j) amax = fmax(amax, hypot(re[S->lowerSrc[p]], im[S->lowerSrc[p]]));
  const double tau = pivotTol * (amax > 0 ? amax : 1);
  const int64_t mf = S->maxFront;
  double *Fr = zc((size_t)(mf * mf), 8), *Fi = zc((size_t)(mf * mf), 8);
  int32_t *rel = zc(n, 4), *stackDim = zc(nsuper, 4);
  int64_t *stackPos = zc(nsuper + 1, 8);
  int32_t nPos = 0, nPert = 0;
  int64_t cap = 1024, top = 0;
  double *sr = zc(cap, 8), *si = zc(cap, 8);
  int ok = 1;
  for (int32_t s = 0; s < nsuper && ok; s++) {
    const int32_t f = S->superStart[s], nc = S->superStart[s + 1] - f;
    const int32_t r0 = S->Sp[s], m = S->Sp[s + 1] - r0;
    const int64_t M = m;
    for (int32_t r = 0; r < m; r++) rel[S->Si[r0 + r]] = r;
    for (int64_t b = 0; b < M; b++) {
      memset(Fr + b * M + b, 0, (size_t)(M - b) * 8);
      memset(Fi + b * M + b, 0, (size_t)(M - b) * 8);
    }
    for (int32_t c = 0; c < nc; c++) {
      const int32_t j = f + c;
      for (int32_t p = S->lowerP[j]; p < S->lowerP[j + 1]; p++) {
        const int64_t at = rel[S->lowerI[p]] + c * M;
        Fr[at] += re[S->lowerSrc[p]];
        Fi[at] += im[S->lowerSrc[p]];
      }
    }
    for (int32_t ch = S->childHead[s]; ch != -1; ch = S->childNext[ch]) {
      const int64_t off = stackPos[--nPos];
      const int64_t mu = stackDim[ch];
      const int32_t cr0 = S->Sp[ch] + (S->superStart[ch + 1] - S->superStart[ch]);
      for (int64_t b = 0; b < mu; b++) {
        const int64_t cb = (int64_t)rel[S->Si[cr0 + b]] * M, ub = off + b * mu;
        for (int64_t a = b; a < mu; a++) {
          const int64_t at = rel[S->Si[cr0 + a]] + cb;
          Fr[at] += sr[ub + a];
          Fi[at] += si[ub + a];
        }
      }
      top = off;
    }
    /* column k (rows i >= lim) -= sum_{c < min(k, nc)} L[:, c] D[c] L[k, c] */
#define UPDATE(k, lim)                                                         \
  do {                                                                         \
    const int64_t kb = (int64_t)(k) * M;                                       \
    const int32_t cmax = (k) < nc ? (k) : nc;                                  \
    for (int32_t c = 0; c < cmax; c++) {                                       \
      const int64_t b0 = (int64_t)c * M;                                       \
      const double lr = Fr[(k) + b0], li = Fi[(k) + b0];                       \
      const double dr = Dr[f + c], di = Di[f + c];                             \
      const double wr = lr * dr - li * di, wi = lr * di + li * dr;             \
      if (wr == 0 && wi == 0) continue;                                        \
      for (int64_t i = (lim); i < M; i++) {                                    \
        const double ar = Fr[i + b0], ai = Fi[i + b0];                         \
        Fr[i + kb] -= ar * wr - ai * wi;                                       \
        Fi[i + kb] -= ar * wi + ai * wr;                                       \
      }                                                                        \
    }                                                                          \
  } while (0)
    for (int32_t k = 0; k < nc; k++) {
      UPDATE(k, k);
      const int64_t kb = (int64_t)k * M;
      double dr = Fr[k + kb], di = Fi[k + kb];
      const double mag = hypot(dr, di);
      if (!(mag >= tau)) {
        if (!isfinite(mag)) {
          ok = 0;
          break;
        }
        if (mag == 0)
          dr = tau;
        else {
          dr *= tau / mag;
          di *= tau / mag;
        }
        nPert++;
      }
      Dr[f + k] = dr;
      Di[f + k] = di;
      Fr[k + kb] = 1;
      Fi[k + kb] = 0;
      const double den = dr * dr + di * di, ir = dr / den, ii = -di / den;
      for (int64_t i = k + 1; i < M; i++) {
        const double ar = Fr[i + kb], ai = Fi[i + kb];
        Fr[i + kb] = ar * ir - ai * ii;
        Fi[i + kb] = ar * ii + ai * ir;
      }
    }
    if (!ok) break;
    for (int32_t j = nc; j < m; j++) UPDATE(j, j);
#undef UPDATE
    memcpy(Lr + S->Lxp[s], Fr, (size_t)(M * nc) * 8);
    memcpy(Li + S->Lxp[s], Fi, (size_t)(M * nc) * 8);
    const int64_t mu = m - nc;
    stackDim[s] = (int32_t)mu;
    if (mu > 0) {
      const int64_t need = top + mu * mu;
      if (need > cap) {
        const int64_t bigger = need > 2 * cap ? need : 2 * cap;
        sr = realloc(sr, (size_t)bigger * 8);
        si = realloc(si, (size_t)bigger * 8);
        cap = bigger;
      }
      for (int64_t b = 0; b < mu; b++) {
        const int64_t src = (nc + b) * M + nc, dst = top + b * mu;
        for (int64_t a = b; a < mu; a++) {
          sr[dst + a] = Fr[src + a];
          si[dst + a] = Fi[src + a];
        }
      }
      stackPos[nPos++] = top;
      top = need;
    } else
      stackPos[nPos++] = top;
  }
  *nPertOut = nPert;
  free(Fr);
  free(Fi);
  free(rel);
  free(stackDim);
  free(stackPos);
  free(sr);
  free(si);
  return ok;
}

void orc_cldlt_solve(const orc_sym *S, const double *Lr, const double *Li, const double *Dr, const double *Di,
                     const double *br, const double *bi, double *xr, double *xi) {
  const int32_t n = S->n;
  double *yr = zc(n, 8), *yi = zc(n, 8);
  for (int32_t i = 0; i < n; i++) {
    yr[i] = br[S->perm[i]];
    yi[i] = bi[S->perm[i]];
  }
  for (int32_t s = 0; s < S->nsuper; s++) {
    const int32_t f = S->superStart[s], nc = S->superStart[s + 1] - f, r0 = S->Sp[s], m = S->Sp[s + 1] - r0;
    const int64_t lo = S->Lxp[s];
    for (int32_t c = 0; c < nc; c++) {
      const int64_t cb = lo + (int64_t)c * m;
      const double a = yr[f + c], b = yi[f + c];
      if (a == 0 && b == 0) continue;
      for (int32_t r = c + 1; r < m; r++) {
        const double lr = Lr[cb + r], li = Li[cb + r];
        const int32_t t = S->Si[r0 + r];
        yr[t] -= lr * a - li * b;
        yi[t] -= lr * b + li * a;
      }
    }
  }
  for (int32_t j = 0; j < n; j++) {
    const double ar = Dr[j], ai = Di[j], den = ar * ar + ai * ai, a = yr[j], b = yi[j];
    yr[j] = (a * ar + b * ai) / den;
    yi[j] = (b * ar - a * ai) / den;
  }
  for (int32_t s = S->nsuper - 1; s >= 0; s--) {
    const int32_t f = S->superStart[s], nc = S->superStart[s + 1] - f, r0 = S->Sp[s], m = S->Sp[s + 1] - r0;
    const int64_t lo = S->Lxp[s];
    for (int32_t c = nc - 1; c >= 0; c--) {
      const int64_t cb = lo + (int64_t)c * m;
      double a = yr[f + c], b = yi[f + c];
      for (int32_t r = c + 1; r < m; r++) {
        const double lr = Lr[cb + r], li = Li[cb + r];
        const int32_t t = S->Si[r0 + r];
        a -= lr * yr[t] - li * yi[t];
        b -= lr * yi[t] + li * yr[t];
      }
      yr[f + c] = a;
      yi[f + c] = b;
    }
  }
  for (int32_t i = 0; i < n; i++) {
    xr[S->perm[i]] = yr[i];
    xi[S->perm[i]] = yi[i];
  }
  free(yr);
  free(yi);
}

/* complexSolveRefined (ldlt.mjs:838-862) on the CSR pattern; returns the steps taken, *rel the relative residual */
int32_t orc_cldlt_solve_refined(const orc_sym *S, const double *Lr, const double *Li, const double *Dr,
                                const double *Di, const int32_t *rowptr, const int32_t *colidx, const double *re,
                                const double *im, const double *br, const double *bi, int32_t maxSteps, double tol,
                                double *xr, double *xi, double *relOut) {
  const int32_t n = S->n;
  orc_cldlt_solve(S, Lr, Li, Dr, Di, br, bi, xr, xi);
  double *rr = zc(n, 8), *ri = zc(n, 8), *dr = zc(n, 8), *di = zc(n, 8);
  double bn = 0;
  for (int32_t i = 0; i < n; i++) bn += br[i] * br[i] + bi[i] * bi[i];
  bn = sqrt(bn);
  if (!(bn != 0)) bn = 1; /* `Math.sqrt(bn) || 1` */
  double rel = INFINITY;
  int32_t steps = 0;
  for (; steps <= maxSteps; steps++) {
    double rn = 0;
    for (int32_t i = 0; i < n; i++) {
      double s_r = br[i], s_i = bi[i];
      for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
        const int32_t j = colidx[p];
        s_r -= re[p] * xr[j] - im[p] * xi[j];
        s_i -= re[p] * xi[j] + im[p] * xr[j];
      }
      rr[i] = s_r;
      ri[i] = s_i;
      rn += s_r * s_r + s_i * s_i;
    }
    rel = sqrt(rn) / bn;
    if (rel < tol || steps == maxSteps) break;
    orc_cldlt_solve(S, Lr, Li, Dr, Di, rr, ri, dr, di);
    for (int32_t i = 0; i < n; i++) {
      xr[i] += dr[i];
      xi[i] += di[i];
    }
  }
  *relOut = rel;
  free(rr);
  free(ri);
  free(dr);
  free(di);
  return steps;
}
