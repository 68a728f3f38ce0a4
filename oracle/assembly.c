/*
 * Oracle (test infrastructure only): P1 diffusion assembly restated from the reference's host code:
 *   buildElementPattern (packages/core/src/la/sparse.mjs:86-110), elementSlots (sparse.mjs:113-125),
 *   assembleCells (packages/physics/src/common/scalar.ts:96-144) with physGrad (lagrange.ts:148-187) for P1 simplices,
 *   restrict (scalar.ts:210-233).
 * The element matrix of a P1 simplex is k_e (grad N_a . grad N_b) |T| (constant gradients; the reference's quadrature
 * sum of the point weights is the measure det J / d! up to rounding, taken exactly here). The scatter adds each
 * element's entries in element order onto a zero-filled K.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

static int cmp_i32(const void *a, const void *b) {
  const int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

/* physGrad for P1 (dN constant): gradients G (k x dim) and det J */
static double p1_grad(int dim, const double *X, double *G) {
  if (dim == 2) {
    static const double dN[6] = {-1, -1, 1, 0, 0, 1};
    double a = 0, b = 0, c = 0, d = 0;
    for (int i = 0; i < 3; i++) {
      const double gx = dN[2 * i], gy = dN[2 * i + 1];
      a += X[2 * i] * gx;
      b += X[2 * i] * gy;
      c += X[2 * i + 1] * gx;
      d += X[2 * i + 1] * gy;
    }
    const double det = a * d - b * c;
    for (int i = 0; i < 3; i++) {
      const double gx = dN[2 * i], gy = dN[2 * i + 1];
      G[2 * i] = (d * gx - c * gy) / det;
      G[2 * i + 1] = (-b * gx + a * gy) / det;
    }
    return det;
  }
  static const double dN[12] = {-1, -1, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1};
  double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < 4; i++)
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) J[3 * r + c] += X[3 * i + r] * dN[3 * i + c];
  const double a = J[0], b = J[1], c = J[2], d = J[3], e = J[4], f = J[5], g = J[6], h = J[7], k = J[8];
  const double A = e * k - f * h, B = -(d * k - f * g), C = d * h - e * g;
  const double det = a * A + b * B + c * C;
  const double inv[9] = {A / det,  -(b * k - c * h) / det, (b * f - c * e) / det,  B / det, (a * k - c * g) / det,
                         -(a * f - c * d) / det, C / det, -(a * h - b * g) / det, (a * e - b * d) / det};
  for (int i = 0; i < 4; i++) {
    const double g0 = dN[3 * i], g1 = dN[3 * i + 1], g2 = dN[3 * i + 2];
    G[3 * i] = inv[0] * g0 + inv[3] * g1 + inv[6] * g2;
    G[3 * i + 1] = inv[1] * g0 + inv[4] * g1 + inv[7] * g2;
    G[3 * i + 2] = inv[2] * g0 + inv[5] * g1 + inv[8] * g2;
  }
  return det;
}

int64_t orc_assemble_p1(int32_t dim, int32_t nn, const double *nodes, int32_t ne, const int32_t *conn,
                        const double *kel, const uint8_t *freeMask, int32_t *rowptr, int32_t *colidx,
                        double *values) {
  const int k = dim + 1;
  /* buildElementPattern */
  int32_t *cap = calloc((size_t)nn + 1, 4);
  for (int32_t i = 0; i < nn; i++) cap[i + 1] = 1;
  for (int64_t q = 0; q < (int64_t)ne * k; q++) cap[conn[q] + 1] += k - 1;
  for (int32_t i = 0; i < nn; i++) cap[i + 1] += cap[i];
  int32_t *buf = malloc(((size_t)cap[nn] + 1) * 4), *fill = malloc(((size_t)nn + 1) * 4);
  memcpy(fill, cap, (size_t)nn * 4);
  for (int32_t i = 0; i < nn; i++) buf[fill[i]++] = i;
  for (int32_t e = 0; e < ne; e++)
    for (int a = 0; a < k; a++) {
      const int32_t r = conn[(int64_t)k * e + a];
      for (int b = 0; b < k; b++)
        if (b != a) buf[fill[r]++] = conn[(int64_t)k * e + b];
    }
  int32_t *frp = calloc((size_t)nn + 1, 4);
  int32_t nnz = 0;
  for (int32_t i = 0; i < nn; i++) {
    qsort(buf + cap[i], (size_t)(cap[i + 1] - cap[i]), 4, cmp_i32);
    for (int32_t p = cap[i]; p < cap[i + 1]; p++)
      if (p == cap[i] || buf[p] != buf[p - 1]) buf[nnz++] = buf[p];
    frp[i + 1] = nnz;
  }
  /* assembleCells: zero-filled K, element matrices scattered through elementSlots in element order */
  double *K = calloc((size_t)nnz + 1, 8);
  const double fact = dim == 2 ? 2.0 : 6.0;
  for (int32_t e = 0; e < ne; e++) {
    double X[12], G[12], Ke[16];
    for (int a = 0; a < k; a++)
      for (int d = 0; d < dim; d++) X[dim * a + d] = nodes[(int64_t)dim * conn[(int64_t)k * e + a] + d];
    const double w = p1_grad(dim, X, G) / fact;
    for (int a = 0; a < k; a++)
      for (int b = a; b < k; b++) {
        double s = 0;
        for (int d = 0; d < dim; d++) s += G[dim * a + d] * G[dim * b + d];
        Ke[k * a + b] = Ke[k * b + a] = s * w;
      }
    for (int a = 0; a < k; a++) {
      const int32_t r = conn[(int64_t)k * e + a];
      for (int b = 0; b < k; b++) {
        const int32_t c = conn[(int64_t)k * e + b];
        int32_t lo = frp[r], hi = frp[r + 1] - 1, slot = -1; /* elementSlots: binary search */
        while (lo <= hi) {
          const int32_t mid = (lo + hi) >> 1;
          if (buf[mid] == c) {
            slot = mid;
            break;
          }
          if (buf[mid] < c) lo = mid + 1;
          else hi = mid - 1;
        }
        K[slot] += kel[e] * Ke[k * a + b];
      }
    }
  }
  /* restrict */
  int32_t *map = malloc(((size_t)nn + 1) * 4);
  int32_t nf = 0;
  for (int32_t i = 0; i < nn; i++) map[i] = (!freeMask || freeMask[i]) ? nf++ : -1;
  int64_t z = 0;
  rowptr[0] = 0;
  for (int32_t i = 0; i < nn; i++) {
    if (map[i] < 0) continue;
    for (int32_t q = frp[i]; q < frp[i + 1]; q++)
      if (map[buf[q]] >= 0) {
        colidx[z] = map[buf[q]];
        values[z] = K[q];
        z++;
      }
    rowptr[map[i] + 1] = (int32_t)z;
  }
  free(cap);
  free(buf);
  free(fill);
  free(frp);
  free(K);
  free(map);
  return z;
}
