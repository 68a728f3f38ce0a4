/*
 * Oracle (test infrastructure only): symbolic analysis restated from packages/core/src/la/sparse.mjs.
 * See oracle.h for what pins each function.
 */
#include "oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define FLIP(i) (-(i)-2)

static void *xcalloc(size_t n, size_t sz) { return calloc(n ? n : 1, sz); }

/* ---- ndOrder, sparse.mjs:230-283 ----------------------------------------------------------------------------- */

/* stable sort of v[0..m) by key (the reference sorts with Array.prototype.sort, which is stable), by merging */
static void nd_sort(int32_t *v, int32_t *buf, int32_t m, const double *key) {
  if (m < 2) return;
  const int32_t h = m / 2;
  nd_sort(v, buf, h, key);
  nd_sort(v + h, buf, m - h, key);
  int32_t i = 0, j = h, k = 0;
  while (i < h && j < m) buf[k++] = key[v[j]] < key[v[i]] ? v[j++] : v[i++];
  while (i < h) buf[k++] = v[i++];
  while (j < m) buf[k++] = v[j++];
  memcpy(v, buf, sizeof(int32_t) * (size_t)m);
}

typedef struct {
  const int32_t *rowptr, *colidx;
  const double *x, *y;
  int32_t leaf;
  int8_t *side; /* 0 unassigned, 1 first half, 2 second half */
  int32_t *order, norder;
} nd_ctx;

/* rec(nodes) of sparse.mjs:236-275 */
static int nd_rec(nd_ctx *c, const int32_t *nodes, int32_t m) {
  if (m <= c->leaf) {
    for (int32_t i = 0; i < m; i++) c->order[c->norder++] = nodes[i];
    return 0;
  }
  double x0 = INFINITY, x1 = -INFINITY, y0 = INFINITY, y1 = -INFINITY;
  for (int32_t i = 0; i < m; i++) {
    const int32_t v = nodes[i];
    if (c->x[v] < x0) x0 = c->x[v];
    if (c->x[v] > x1) x1 = c->x[v];
    if (c->y[v] < y0) y0 = c->y[v];
    if (c->y[v] > y1) y1 = c->y[v];
  }
  const double *key = x1 - x0 >= y1 - y0 ? c->x : c->y;
  int32_t *sorted = malloc(sizeof(int32_t) * (size_t)m), *buf = malloc(sizeof(int32_t) * (size_t)m);
  int32_t *A = malloc(sizeof(int32_t) * (size_t)m), *S = malloc(sizeof(int32_t) * (size_t)m);
  int32_t *B = malloc(sizeof(int32_t) * (size_t)m);
  if (!sorted || !buf || !A || !S || !B) {
    free(sorted), free(buf), free(A), free(S), free(B);
    return -1;
  }
  memcpy(sorted, nodes, sizeof(int32_t) * (size_t)m);
  nd_sort(sorted, buf, m, key);
  const int32_t half = m >> 1;
  for (int32_t i = 0; i < m; i++) c->side[sorted[i]] = i < half ? 1 : 2;
  int32_t na = 0, ns = 0, nb = 0;
  for (int32_t i = 0; i < m; i++) {
    const int32_t v = sorted[i];
    if (i < half) {
      int sep = 0;
      for (int32_t p = c->rowptr[v]; p < c->rowptr[v + 1]; p++) {
        const int32_t w = c->colidx[p];
        if (w != v && c->side[w] == 2) {
          sep = 1;
          break;
        }
      }
      if (sep)
        S[ns++] = v;
      else
        A[na++] = v;
    } else
      B[nb++] = v;
  }
  for (int32_t i = 0; i < m; i++) c->side[sorted[i]] = 0;
  int rc = nd_rec(c, A, na);
  if (!rc) rc = nd_rec(c, B, nb);
  if (!rc)
    for (int32_t i = 0; i < ns; i++) c->order[c->norder++] = S[i];
  free(sorted), free(buf), free(A), free(S), free(B);
  return rc;
}

int orc_nd_order(int32_t n, const int32_t *rowptr, const int32_t *colidx, const double *x, const double *y,
                 int32_t leaf, int32_t *perm) {
  if (leaf < 1) return -1;
  nd_ctx c = {rowptr, colidx, x, y, leaf, xcalloc((size_t)n, 1), perm, 0};
  int32_t *all = xcalloc((size_t)n, sizeof(int32_t));
  if (!c.side || !all) {
    free(c.side), free(all);
    return -1;
  }
  for (int32_t i = 0; i < n; i++) all[i] = i;
  const int rc = nd_rec(&c, all, n);
  free(c.side), free(all);
  return rc;
}

/* ---- amdOrder, sparse.mjs:297-592 ---------------------------------------------------------------------------- */
int orc_amd(int32_t n, const int32_t *rowptr, const int32_t *colidx, double denseFactor, int32_t *out) {
  if (n == 0) return 0;
  int64_t cnz = 0; /* sparse.mjs:300-301: off-diagonal entries */
  for (int32_t i = 0; i < n; i++)
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++)
      if (colidx[p] != i) cnz++;
  const int64_t nzmax = cnz + cnz / 5 + 2 * (int64_t)n; /* :302 */
  int32_t *Cp = xcalloc(n + 1, 4), *Ci = xcalloc(nzmax, 4);
  int32_t *len = xcalloc(n + 1, 4), *nv = xcalloc(n + 1, 4), *next = xcalloc(n + 1, 4), *head = xcalloc(n + 1, 4),
          *elen = xcalloc(n + 1, 4), *degree = xcalloc(n + 1, 4), *w = xcalloc(n + 1, 4), *hhead = xcalloc(n + 1, 4),
          *last = xcalloc(n + 1, 4);
  int64_t q0 = 0;
  for (int32_t i = 0; i < n; i++) { /* :305-310 */
    Cp[i] = (int32_t)q0;
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++)
      if (colidx[p] != i) Ci[q0++] = colidx[p];
  }
  Cp[n] = (int32_t)q0;
  for (int32_t k = 0; k < n; k++) len[k] = Cp[k + 1] - Cp[k];
  double dense = fmax(16.0, denseFactor * sqrt((double)n)); /* :321 */
  if ((double)(n - 2) < dense) dense = (double)(n - 2);
  for (int32_t i = 0; i <= n; i++) {
    head[i] = -1;
    last[i] = -1;
    next[i] = -1;
    hhead[i] = -1;
    nv[i] = 1;
    w[i] = 1;
    elen[i] = 0;
    degree[i] = len[i];
  }
  int32_t mark, lemax = 0;
  /* wclear (:333-339) inlined as a macro-like helper */
#define WCLEAR(mk, lm)                                                                                              \
  do {                                                                                                              \
    if ((mk) < 2 || (int64_t)(mk) + (int64_t)(lm) > 2000000000LL) {                                                 \
      for (int32_t kk = 0; kk < n; kk++)                                                                            \
        if (w[kk] != 0) w[kk] = 1;                                                                                  \
      (mk) = 2;                                                                                                     \
    }                                                                                                               \
  } while (0)
  mark = 0;
  WCLEAR(mark, 0);
  elen[n] = -2;
  Cp[n] = -1;
  w[n] = 0;
  nv[n] = 0;
  int32_t nel = 0;
  for (int32_t i = 0; i < n; i++) { /* :345-361 */
    const int32_t d = degree[i];
    if (d == 0) {
      elen[i] = -2;
      nel++;
      Cp[i] = -1;
      w[i] = 0;
    } else if ((double)d > dense) {
      nv[i] = 0;
      elen[i] = -1;
      nel++;
      Cp[i] = FLIP(n);
      nv[n]++;
    } else {
      if (head[d] != -1) last[head[d]] = i;
      next[i] = head[d];
      head[d] = i;
    }
  }
  int32_t mindeg = 0;
  while (nel < n) {
    int32_t k = -1; /* :366-368 pivot of minimum approximate degree */
    for (; mindeg < n && (k = head[mindeg]) == -1; mindeg++)
      ;
    if (next[k] != -1) last[next[k]] = -1;
    head[mindeg] = next[k];
    const int32_t elenk = elen[k];
    int32_t nvk = nv[k];
    nel += nvk;
    if (elenk > 0 && cnz + mindeg >= nzmax) { /* :374-392 garbage collection */
      for (int32_t j = 0; j < n; j++) {
        const int32_t p = Cp[j];
        if (p >= 0) {
          Cp[j] = Ci[p];
          Ci[p] = FLIP(j);
        }
      }
      int64_t q = 0;
      for (int64_t p = 0; p < cnz;) {
        const int32_t j = FLIP(Ci[p++]);
        if (j >= 0) {
          Ci[q] = Cp[j];
          Cp[j] = (int32_t)q++;
          for (int32_t k3 = 0; k3 < len[j] - 1; k3++) Ci[q++] = Ci[p++];
        }
      }
      cnz = q;
    }
    /* new element Lk, :394-426 */
    int32_t dk = 0;
    nv[k] = -nvk;
    int64_t p = Cp[k];
    const int64_t pk1 = elenk == 0 ? p : cnz;
    int64_t pk2 = pk1;
    for (int32_t k1 = 1; k1 <= elenk + 1; k1++) {
      int32_t e, ln;
      int64_t pj;
      if (k1 > elenk) {
        e = k;
        pj = p;
        ln = len[k] - elenk;
      } else {
        e = Ci[p++];
        pj = Cp[e];
        ln = len[e];
      }
      for (int32_t k2 = 1; k2 <= ln; k2++) {
        const int32_t i = Ci[pj++];
        const int32_t nvi = nv[i];
        if (nvi <= 0) continue;
        dk += nvi;
        nv[i] = -nvi;
        Ci[pk2++] = i;
        if (next[i] != -1) last[next[i]] = last[i];
        if (last[i] != -1)
          next[last[i]] = next[i];
        else
          head[degree[i]] = next[i];
      }
      if (e != k) {
        Cp[e] = FLIP(k);
        w[e] = 0;
      }
    }
    if (elenk != 0) cnz = pk2;
    degree[k] = dk;
    Cp[k] = (int32_t)pk1;
    len[k] = (int32_t)(pk2 - pk1);
    elen[k] = -2;
    /* set differences |Le \ Lk|, :433-445 */
    WCLEAR(mark, lemax);
    for (int64_t pk = pk1; pk < pk2; pk++) {
      const int32_t i = Ci[pk];
      const int32_t eln = elen[i];
      if (eln <= 0) continue;
      const int32_t nvi = -nv[i];
      const int32_t wnvi = mark - nvi;
      for (int64_t q = Cp[i]; q <= (int64_t)Cp[i] + eln - 1; q++) {
        const int32_t e = Ci[q];
        if (w[e] >= mark)
          w[e] -= nvi;
        else if (w[e] != 0)
          w[e] = degree[e] + wnvi;
      }
    }
    /* approximate degree update, element absorption, hashing, :447-497 */
    for (int64_t pk = pk1; pk < pk2; pk++) {
      const int32_t i = Ci[pk];
      const int64_t p1 = Cp[i], p2 = p1 + elen[i] - 1;
      int64_t pn = p1;
      int64_t h = 0;
      int32_t d = 0;
      for (int64_t q = p1; q <= p2; q++) {
        const int32_t e = Ci[q];
        if (w[e] != 0) {
          const int32_t dext = w[e] - mark;
          if (dext > 0) {
            d += dext;
            Ci[pn++] = e;
            h += e;
          } else {
            Cp[e] = FLIP(k);
            w[e] = 0;
          }
        }
      }
      elen[i] = (int32_t)(pn - p1 + 1);
      const int64_t p3 = pn, p4 = p1 + len[i];
      for (int64_t q = p2 + 1; q < p4; q++) {
        const int32_t j = Ci[q];
        const int32_t nvj = nv[j];
        if (nvj <= 0) continue;
        d += nvj;
        Ci[pn++] = j;
        h += j;
      }
      if (d == 0) { /* mass elimination */
        Cp[i] = FLIP(k);
        const int32_t nvi = -nv[i];
        dk -= nvi;
        nvk += nvi;
        nel += nvi;
        nv[i] = 0;
        elen[i] = -1;
      } else {
        if (d < degree[i]) degree[i] = d;
        Ci[pn] = Ci[p3];
        Ci[p3] = Ci[p1];
        Ci[p1] = k;
        len[i] = (int32_t)(pn - p1 + 1);
        h = (h < 0 ? -h : h) % n;
        next[i] = hhead[h];
        hhead[h] = i;
        last[i] = (int32_t)h;
      }
    }
    degree[k] = dk;
    if (dk > lemax) lemax = dk;
    mark = mark + lemax;
    WCLEAR(mark, lemax);
    /* supervariable detection, :502-530 */
    for (int64_t pk = pk1; pk < pk2; pk++) {
      int32_t i = Ci[pk];
      if (nv[i] >= 0) continue;
      const int32_t h = last[i];
      i = hhead[h];
      hhead[h] = -1;
      for (; i != -1 && next[i] != -1; i = next[i], mark++) {
        const int32_t ln = len[i], eln = elen[i];
        for (int64_t q = (int64_t)Cp[i] + 1; q <= (int64_t)Cp[i] + ln - 1; q++) w[Ci[q]] = mark;
        int32_t jlast = i;
        for (int32_t j = next[i]; j != -1;) {
          int ok = len[j] == ln && elen[j] == eln;
          for (int64_t q = (int64_t)Cp[j] + 1; ok && q <= (int64_t)Cp[j] + ln - 1; q++)
            if (w[Ci[q]] != mark) ok = 0;
          if (ok) {
            Cp[j] = FLIP(i);
            nv[i] += nv[j];
            nv[j] = 0;
            elen[j] = -1;
            j = next[j];
            next[jlast] = j;
          } else {
            jlast = j;
            j = next[j];
          }
        }
      }
    }
    /* finalize Lk, :532-553 */
    int64_t q = pk1;
    for (int64_t pk = pk1; pk < pk2; pk++) {
      const int32_t i = Ci[pk];
      const int32_t nvi = -nv[i];
      if (nvi <= 0) continue;
      nv[i] = nvi;
      int32_t d = degree[i] + dk - nvi;
      if (n - nel - nvi < d) d = n - nel - nvi;
      if (head[d] != -1) last[head[d]] = i;
      next[i] = head[d];
      last[i] = -1;
      head[d] = i;
      if (d < mindeg) mindeg = d;
      degree[i] = d;
      Ci[q++] = i;
    }
    nv[k] = nvk;
    if ((len[k] = (int32_t)(q - pk1)) == 0) {
      Cp[k] = -1;
      w[k] = 0;
    }
    if (elenk != 0) cnz = q;
  }
#undef WCLEAR
  /* postorder of the assembly tree, :555-590 */
  for (int32_t i = 0; i < n; i++) Cp[i] = FLIP(Cp[i]);
  for (int32_t j = 0; j <= n; j++) head[j] = -1;
  for (int32_t j = n; j >= 0; j--) {
    if (nv[j] > 0) continue;
    if (Cp[j] == -1) { /* JS reads head[-1] as undefined (stored as 0) and drops the write to head[-1] */
      next[j] = 0;
      continue;
    }
    next[j] = head[Cp[j]];
    head[Cp[j]] = j;
  }
  for (int32_t e = n; e >= 0; e--) {
    if (nv[e] <= 0) continue;
    if (Cp[e] != -1) {
      next[e] = head[Cp[e]];
      head[Cp[e]] = e;
    }
  }
  int32_t o = 0;
  for (int32_t i = 0; i <= n; i++) {
    if (Cp[i] != -1) continue;
    int32_t top = 0;
    w[0] = i;
    while (top >= 0) {
      const int32_t pp = w[top];
      const int32_t ii = head[pp];
      if (ii == -1) {
        top--;
        if (pp < n) out[o++] = pp;
      } else {
        head[pp] = next[ii];
        w[++top] = ii;
      }
    }
  }
  free(Cp);
  free(Ci);
  free(len);
  free(nv);
  free(next);
  free(head);
  free(elen);
  free(degree);
  free(w);
  free(hhead);
  free(last);
  return o == n ? 0 : -1;
}

/* ---- postorder, sparse.mjs:597-625 ----------------------------------------------------------------------------- */
static void postorder(const int32_t *parent, int32_t n, int32_t *post) {
  int32_t *head = xcalloc(n, 4), *next = xcalloc(n, 4), *stack = xcalloc(n, 4);
  for (int32_t j = 0; j < n; j++) head[j] = -1;
  for (int32_t j = n - 1; j >= 0; j--) {
    if (parent[j] == -1) continue;
    next[j] = head[parent[j]];
    head[parent[j]] = j;
  }
  int32_t k = 0;
  for (int32_t j = 0; j < n; j++) {
    if (parent[j] != -1) continue;
    int32_t top = 0;
    stack[0] = j;
    while (top >= 0) {
      const int32_t p = stack[top];
      const int32_t i = head[p];
      if (i == -1) {
        top--;
        post[k++] = p;
      } else {
        head[p] = next[i];
        stack[++top] = i;
      }
    }
  }
  free(head);
  free(next);
  free(stack);
}

/* ---- etreeOf, sparse.mjs:628-645 ------------------------------------------------------------------------------- */
static void etree(int32_t n, const int32_t *rowptr, const int32_t *colidx, const int32_t *perm, const int32_t *pinv,
                  int32_t *parent) {
  int32_t *anc = xcalloc(n, 4);
  for (int32_t i = 0; i < n; i++) parent[i] = anc[i] = -1;
  for (int32_t k = 0; k < n; k++) {
    const int32_t v = perm[k];
    for (int32_t p = rowptr[v]; p < rowptr[v + 1]; p++) {
      int32_t i = pinv[colidx[p]];
      while (i != -1 && i < k) {
        const int32_t inext = anc[i];
        anc[i] = k;
        if (inext == -1) parent[i] = k;
        i = inext;
      }
    }
  }
  free(anc);
}

void orc_fingerprint(int32_t n, const int32_t *rowptr, const int32_t *colidx, char *out) {
  uint32_t h = 0x811c9dc5u; /* FNV-1a over the bytes of rowptr then colidx, sparse.mjs:908-925 */
#define MIX(v)                                                                                                      \
  do {                                                                                                              \
    uint32_t vv = (uint32_t)(v);                                                                                    \
    h ^= vv & 0xffu;                                                                                                \
    h *= 0x01000193u;                                                                                               \
    h ^= (vv >> 8) & 0xffu;                                                                                         \
    h *= 0x01000193u;                                                                                               \
    h ^= (vv >> 16) & 0xffu;                                                                                        \
    h *= 0x01000193u;                                                                                               \
    h ^= vv >> 24;                                                                                                  \
    h *= 0x01000193u;                                                                                               \
  } while (0)
  for (int32_t i = 0; i <= n; i++) MIX(rowptr[i]);
  const int32_t nnz = rowptr[n];
  for (int32_t p = 0; p < nnz; p++) MIX(colidx[p]);
#undef MIX
  snprintf(out, 48, "%d:%d:%x", n, nnz, h);
}

void orc_sym_free(orc_sym *s) {
  if (!s) return;
  free(s->perm);
  free(s->pinv);
  free(s->parent);
  free(s->colCount);
  free(s->lowerP);
  free(s->lowerI);
  free(s->lowerSrc);
  free(s->superStart);
  free(s->colSuper);
  free(s->superParent);
  free(s->childHead);
  free(s->childNext);
  free(s->Sp);
  free(s->Si);
  free(s->Lxp);
  free(s);
}

static int cmp_i32(const void *a, const void *b) {
  const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* ---- analyze, sparse.mjs:666-902 ------------------------------------------------------------------------------- */
orc_sym *orc_analyze(int32_t n, const int32_t *rowptr, const int32_t *colidx, const int32_t *perm0in, int relax) {
  orc_sym *S = xcalloc(1, sizeof(orc_sym));
  orc_fingerprint(n, rowptr, colidx, S->patternKey);
  S->n = n;
  if (n == 0) { /* :671-675 */
    S->lowerP = xcalloc(1, 4);
    S->superStart = xcalloc(1, 4);
    S->Sp = xcalloc(1, 4);
    S->Lxp = xcalloc(1, 8);
    S->fill = 1;
    return S;
  }
  int32_t *perm0 = xcalloc(n, 4);
  if (perm0in)
    memcpy(perm0, perm0in, (size_t)n * 4);
  else if (orc_amd(n, rowptr, colidx, 10.0, perm0) != 0) {
    free(perm0);
    orc_sym_free(S);
    return NULL;
  }
  /* etree and postorder, :682-696 */
  int32_t *pinv = xcalloc(n, 4), *parent0 = xcalloc(n, 4), *post = xcalloc(n, 4);
  for (int32_t i = 0; i < n; i++) pinv[perm0[i]] = i;
  etree(n, rowptr, colidx, perm0, pinv, parent0);
  postorder(parent0, n, post);
  int32_t *perm = xcalloc(n, 4), *postInv = xcalloc(n, 4), *parent = xcalloc(n, 4);
  for (int32_t k = 0; k < n; k++) perm[k] = perm0[post[k]];
  for (int32_t k = 0; k < n; k++) postInv[post[k]] = k;
  for (int32_t k = 0; k < n; k++) {
    const int32_t p = parent0[post[k]];
    parent[k] = p == -1 ? -1 : postInv[p];
  }
  for (int32_t i = 0; i < n; i++) pinv[perm[i]] = i;
  free(parent0);
  free(post);
  free(postInv);
  free(perm0);
  /* lower CSC of P A P^T with source slots, :697-718 */
  int32_t *lowerP = xcalloc(n + 1, 4);
  for (int32_t v = 0; v < n; v++) {
    const int32_t pv = pinv[v];
    for (int32_t p = rowptr[v]; p < rowptr[v + 1]; p++)
      if (pinv[colidx[p]] >= pv) lowerP[pv + 1]++;
  }
  for (int32_t j = 0; j < n; j++) lowerP[j + 1] += lowerP[j];
  const int32_t nnzA = lowerP[n];
  int32_t *lowerI = xcalloc(nnzA, 4), *lowerSrc = xcalloc(nnzA, 4), *nxt = xcalloc(n, 4);
  memcpy(nxt, lowerP, (size_t)n * 4);
  for (int32_t v = 0; v < n; v++) {
    const int32_t pv = pinv[v];
    for (int32_t p = rowptr[v]; p < rowptr[v + 1]; p++) {
      const int32_t pi = pinv[colidx[p]];
      if (pi >= pv) {
        const int32_t q = nxt[pv]++;
        lowerI[q] = pi;
        lowerSrc[q] = p;
      }
    }
  }
  free(nxt);
  /* column counts, Gilbert-Ng-Peyton, :719-753 */
  int32_t *colCount = xcalloc(n, 4);
  {
    int32_t *first = xcalloc(n, 4), *maxfirst = xcalloc(n, 4), *prevleaf = xcalloc(n, 4), *ancestor = xcalloc(n, 4);
    for (int32_t i = 0; i < n; i++) first[i] = maxfirst[i] = prevleaf[i] = -1;
    for (int32_t k = 0; k < n; k++) {
      colCount[k] = first[k] == -1 ? 1 : 0;
      for (int32_t j = k; j != -1 && first[j] == -1; j = parent[j]) first[j] = k;
    }
    for (int32_t i = 0; i < n; i++) ancestor[i] = i;
    for (int32_t j = 0; j < n; j++) {
      if (parent[j] != -1) colCount[parent[j]]--;
      for (int32_t p = lowerP[j]; p < lowerP[j + 1]; p++) {
        const int32_t i = lowerI[p];
        if (i <= j || first[j] <= maxfirst[i]) continue;
        maxfirst[i] = first[j];
        const int32_t jprev = prevleaf[i];
        prevleaf[i] = j;
        colCount[j]++;
        if (jprev == -1) continue;
        int32_t q = jprev;
        while (q != ancestor[q]) q = ancestor[q];
        for (int32_t s = jprev, sp; s != q; s = sp) {
          sp = ancestor[s];
          ancestor[s] = q;
        }
        colCount[q]--;
      }
      if (parent[j] != -1) ancestor[j] = parent[j];
    }
    for (int32_t j = 0; j < n; j++)
      if (parent[j] != -1) colCount[parent[j]] += colCount[j];
    free(first);
    free(maxfirst);
    free(prevleaf);
    free(ancestor);
  }
  int64_t nnzL = 0;
  double flops = 0;
  for (int32_t j = 0; j < n; j++) {
    nnzL += colCount[j];
    flops += (double)colCount[j] * (double)colCount[j];
  }
  /* fundamental supernodes, :760-766 */
  int32_t *nchild = xcalloc(n, 4), *fStart = xcalloc(n + 1, 4);
  for (int32_t j = 0; j < n; j++)
    if (parent[j] != -1) nchild[parent[j]]++;
  int32_t nsf = 0;
  fStart[nsf++] = 0;
  for (int32_t j = 1; j < n; j++)
    if (!(parent[j - 1] == j && colCount[j - 1] == colCount[j] + 1 && nchild[j] == 1)) fStart[nsf++] = j;
  fStart[nsf] = n; /* nsf = number of fundamental supernodes */
  free(nchild);
  int32_t *colSuperF = xcalloc(n, 4);
  for (int32_t s = 0; s < nsf; s++)
    for (int32_t j = fStart[s]; j < fStart[s + 1]; j++) colSuperF[j] = s;
  /* relaxed amalgamation, :770-814 */
  int32_t *superStart = xcalloc(nsf + 1, 4);
  int32_t nsuper = 0;
  if (relax && nsf > 1) {
    int32_t *sf = xcalloc(nsf, 4), *sl = xcalloc(nsf, 4), *sparent = xcalloc(nsf, 4);
    double *sm = xcalloc(nsf, 8), *sz = xcalloc(nsf, 8), *snnz = xcalloc(nsf, 8);
    unsigned char *alive = xcalloc(nsf, 1);
    for (int32_t s = 0; s < nsf; s++) {
      alive[s] = 1;
      sf[s] = fStart[s];
      sl[s] = fStart[s + 1] - 1;
      sm[s] = colCount[fStart[s]];
      const double nc = fStart[s + 1] - fStart[s];
      snnz[s] = nc * sm[s] - (nc * (nc - 1)) / 2;
      const int32_t pl = parent[sl[s]];
      sparent[s] = pl == -1 ? -1 : colSuperF[pl];
    }
    for (int32_t s = 0; s < nsf; s++) {
      const int32_t P = sparent[s];
      if (P == -1 || sf[P] != sl[s] + 1) continue;
      const double ncC = sl[s] - sf[s] + 1, ncP = sl[P] - sf[P] + 1;
      const double nc = ncC + ncP;
      const double m = ncC + sm[P];
      const double stored = nc * m - (nc * (nc - 1)) / 2;
      const double zeros = stored - snnz[s] - snnz[P] + sz[s] + sz[P];
      const double zf = zeros / stored;
      const int merge = nc <= 4 || (nc <= 16 && zf < 0.8) || (nc <= 48 && zf < 0.1) || zf < 0.05;
      if (!merge) continue;
      sf[P] = sf[s];
      sm[P] = m;
      sz[P] = zeros;
      snnz[P] += snnz[s];
      alive[s] = 0;
    }
    for (int32_t s = 0; s < nsf; s++)
      if (alive[s]) superStart[nsuper++] = sf[s];
    superStart[nsuper] = n;
    free(sf);
    free(sl);
    free(sparent);
    free(sm);
    free(sz);
    free(snnz);
    free(alive);
  } else {
    for (int32_t s = 0; s <= nsf; s++) superStart[s] = fStart[s];
    nsuper = nsf;
  }
  free(fStart);
  free(colSuperF);
  int32_t *colSuper = xcalloc(n, 4), *superParent = xcalloc(nsuper, 4);
  for (int32_t s = 0; s < nsuper; s++)
    for (int32_t j = superStart[s]; j < superStart[s + 1]; j++) colSuper[j] = s;
  for (int32_t s = 0; s < nsuper; s++) {
    const int32_t pl = parent[superStart[s + 1] - 1];
    superParent[s] = pl == -1 ? -1 : colSuper[pl];
  }
  /* row structures, :820-867; children linked in decreasing order */
  int32_t *childHead = xcalloc(nsuper, 4), *childNext = xcalloc(nsuper, 4);
  for (int32_t s = 0; s < nsuper; s++) childHead[s] = childNext[s] = -1;
  for (int32_t s = 0; s < nsuper; s++) {
    const int32_t P = superParent[s];
    if (P == -1) continue;
    childNext[s] = childHead[P];
    childHead[P] = s;
  }
  int32_t *Sp = xcalloc(nsuper + 1, 4);
  int32_t **rowsOf = xcalloc(nsuper, sizeof(int32_t *));
  int32_t *rowsLen = xcalloc(nsuper, 4);
  int32_t *marker = xcalloc(n, 4), *below = xcalloc(n, 4);
  for (int32_t i = 0; i < n; i++) marker[i] = -1;
  int32_t maxFront = 0;
  for (int32_t s = 0; s < nsuper; s++) {
    const int32_t f = superStart[s], l = superStart[s + 1] - 1;
    int32_t nb = 0;
    for (int32_t j = f; j <= l; j++)
      for (int32_t p = lowerP[j]; p < lowerP[j + 1]; p++) {
        const int32_t i = lowerI[p];
        if (i > l && marker[i] != s) {
          marker[i] = s;
          below[nb++] = i;
        }
      }
    for (int32_t c = childHead[s]; c != -1; c = childNext[c]) {
      const int32_t *rc = rowsOf[c];
      const int32_t ncc = superStart[c + 1] - superStart[c];
      for (int32_t r = ncc; r < rowsLen[c]; r++) {
        const int32_t i = rc[r];
        if (i > l && marker[i] != s) {
          marker[i] = s;
          below[nb++] = i;
        }
      }
    }
    const int32_t len = l - f + 1 + nb;
    int32_t *rows = xcalloc(len, 4);
    for (int32_t j = f; j <= l; j++) rows[j - f] = j;
    memcpy(rows + (l - f + 1), below, (size_t)nb * 4);
    qsort(rows + (l - f + 1), nb, 4, cmp_i32);
    rowsOf[s] = rows;
    rowsLen[s] = len;
    Sp[s + 1] = Sp[s] + len;
    if (len > maxFront) maxFront = len;
  }
  free(marker);
  free(below);
  int32_t *Si = xcalloc(Sp[nsuper], 4);
  int64_t *Lxp = xcalloc(nsuper + 1, 8);
  for (int32_t s = 0; s < nsuper; s++) {
    memcpy(Si + Sp[s], rowsOf[s], (size_t)rowsLen[s] * 4);
    Lxp[s + 1] = Lxp[s] + (int64_t)rowsLen[s] * (superStart[s + 1] - superStart[s]);
    free(rowsOf[s]);
  }
  free(rowsOf);
  free(rowsLen);
  S->nsuper = nsuper;
  S->maxFront = maxFront;
  S->nnzL = nnzL;
  S->nnzA = nnzA;
  S->nnzLStored = Lxp[nsuper];
  S->flops = flops;
  S->fill = (double)nnzL / (double)nnzA;
  S->perm = perm;
  S->pinv = pinv;
  S->parent = parent;
  S->colCount = colCount;
  S->lowerP = lowerP;
  S->lowerI = lowerI;
  S->lowerSrc = lowerSrc;
  S->superStart = superStart;
  S->colSuper = colSuper;
  S->superParent = superParent;
  S->childHead = childHead;
  S->childNext = childNext;
  S->Sp = Sp;
  S->Si = Si;
  S->Lxp = Lxp;
  return S;
}

/* ---- UpLookingCholesky, sparse.mjs:1379-1544 ------------------------------------------------------------------- */
static int32_t uplook_ereach(const orc_uplook *u, int32_t k, int32_t *mark, int32_t *s, int32_t stamp) {
  int32_t top = u->n;
  mark[k] = stamp;
  for (int32_t p = u->Cp[k]; p < u->Cp[k + 1]; p++) {
    int32_t i = u->Ci[p];
    if (i > k) continue;
    int32_t len = 0;
    for (; mark[i] != stamp; i = u->parent[i]) {
      s[len++] = i;
      mark[i] = stamp;
    }
    while (len > 0) s[--top] = s[--len];
  }
  return top;
}

orc_uplook *orc_uplook_create(int32_t n, const int32_t *rowptr, const int32_t *colidx, const int32_t *perm) {
  orc_uplook *u = xcalloc(1, sizeof(orc_uplook));
  u->n = n;
  u->perm = xcalloc(n, 4);
  memcpy(u->perm, perm, (size_t)n * 4);
  u->pinv = xcalloc(n, 4);
  for (int32_t i = 0; i < n; i++) u->pinv[perm[i]] = i;
  u->Cp = xcalloc(n + 1, 4);
  for (int32_t i = 0; i < n; i++)
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
      const int32_t pi = u->pinv[i], pj = u->pinv[colidx[p]];
      if (pi <= pj) u->Cp[pj + 1]++;
    }
  for (int32_t i = 0; i < n; i++) u->Cp[i + 1] += u->Cp[i];
  const int32_t nnzC = u->Cp[n];
  u->Ci = xcalloc(nnzC, 4);
  u->srcIndex = xcalloc(nnzC, 4);
  int32_t *next = xcalloc(n, 4);
  memcpy(next, u->Cp, (size_t)n * 4);
  for (int32_t i = 0; i < n; i++)
    for (int32_t p = rowptr[i]; p < rowptr[i + 1]; p++) {
      const int32_t pi = u->pinv[i], pj = u->pinv[colidx[p]];
      if (pi <= pj) {
        const int32_t q = next[pj]++;
        u->Ci[q] = pi;
        u->srcIndex[q] = p;
      }
    }
  free(next);
  u->parent = xcalloc(n, 4);
  int32_t *anc = xcalloc(n, 4);
  for (int32_t i = 0; i < n; i++) u->parent[i] = anc[i] = -1;
  for (int32_t k = 0; k < n; k++)
    for (int32_t p = u->Cp[k]; p < u->Cp[k + 1]; p++) {
      int32_t i = u->Ci[p];
      while (i != -1 && i < k) {
        const int32_t inext = anc[i];
        anc[i] = k;
        if (inext == -1) u->parent[i] = k;
        i = inext;
      }
    }
  free(anc);
  int32_t *mark = xcalloc(n, 4), *s = xcalloc(n, 4), *cc = xcalloc(n, 4);
  for (int32_t k = 0; k < n; k++) {
    const int32_t top = uplook_ereach(u, k, mark, s, k + 1);
    for (int32_t q = top; q < n; q++) cc[s[q]]++;
    cc[k]++;
  }
  u->Lp = xcalloc(n + 1, 4);
  for (int32_t k = 0; k < n; k++) u->Lp[k + 1] = u->Lp[k] + cc[k];
  u->nnzL = u->Lp[n];
  u->Li = xcalloc(u->nnzL, 4);
  u->Lx = xcalloc(u->nnzL, 8);
  free(mark);
  free(s);
  free(cc);
  return u;
}

int orc_uplook_factor(orc_uplook *u, const double *values) {
  const int32_t n = u->n;
  int32_t *mark = xcalloc(n, 4), *s = xcalloc(n, 4), *c = xcalloc(n, 4);
  double *x = xcalloc(n, 8);
  int32_t stamp = 1, ok = 1;
  for (int32_t k = 0; k < n; k++) c[k] = u->Lp[k];
  for (int32_t k = 0; k < n && ok; k++) {
    stamp++;
    const int32_t top = uplook_ereach(u, k, mark, s, stamp);
    for (int32_t p = u->Cp[k]; p < u->Cp[k + 1]; p++) x[u->Ci[p]] = values[u->srcIndex[p]];
    double d = x[k];
    x[k] = 0;
    for (int32_t q = top; q < n; q++) {
      const int32_t i = s[q];
      const double lki = x[i] / u->Lx[u->Lp[i]];
      x[i] = 0;
      const int32_t ci = c[i];
      for (int32_t p = u->Lp[i] + 1; p < ci; p++) x[u->Li[p]] -= u->Lx[p] * lki;
      d -= lki * lki;
      u->Li[ci] = k;
      u->Lx[ci] = lki;
      c[i] = ci + 1;
    }
    if (!(d > 0)) {
      ok = 0;
      break;
    }
    u->Li[c[k]] = k;
    u->Lx[c[k]] = sqrt(d);
    c[k]++;
  }
  free(mark);
  free(s);
  free(c);
  free(x);
  return ok;
}

void orc_uplook_solve(const orc_uplook *u, const double *b, double *out) {
  const int32_t n = u->n;
  double *y = xcalloc(n, 8);
  for (int32_t i = 0; i < n; i++) y[i] = b[u->perm[i]];
  for (int32_t j = 0; j < n; j++) {
    y[j] /= u->Lx[u->Lp[j]];
    const double yj = y[j];
    for (int32_t p = u->Lp[j] + 1; p < u->Lp[j + 1]; p++) y[u->Li[p]] -= u->Lx[p] * yj;
  }
  for (int32_t j = n - 1; j >= 0; j--) {
    double v = y[j];
    for (int32_t p = u->Lp[j] + 1; p < u->Lp[j + 1]; p++) v -= u->Lx[p] * y[u->Li[p]];
    y[j] = v / u->Lx[u->Lp[j]];
  }
  for (int32_t i = 0; i < n; i++) out[u->perm[i]] = y[i];
  free(y);
}

void orc_uplook_free(orc_uplook *u) {
  if (!u) return;
  free(u->perm);
  free(u->pinv);
  free(u->Cp);
  free(u->Ci);
  free(u->srcIndex);
  free(u->parent);
  free(u->Lp);
  free(u->Li);
  free(u->Lx);
  free(u);
}
