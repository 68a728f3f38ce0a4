// Device assembly of P1 diffusion matrices (amp_asm_*): the element pattern, the element slots and the restriction are
// built once on the host; every assembly then runs on the device as a gather over each restricted entry's element
// contributions, in element order.
//
// Restates, for P1 simplices (tri3 / tet4) and a per-element coefficient k_e:
//   buildElementPattern (packages/core/src/la/sparse.mjs:86-110): full symmetric CSR, rows sorted, diagonal present;
//   elementSlots (sparse.mjs:113-125): the CSR slot of every local entry (e, a, b);
//   assembleCells (packages/physics/src/common/scalar.ts:96-144): Ke[a][b] = k_e s_ab w with s_ab = grad N_a . grad N_b
//     from physGrad (lagrange.ts:148-187) and w the element measure, scattered K[slot(e, a, b)] += Ke in element order;
//   restrict (scalar.ts:210-233): rows and columns of the free nodes, order kept.
// P1 gradients are constant over the element, so the quadrature sum of the reference (sum over points of w_q, which is
// the element measure up to rounding) is replaced by the measure det J / d! itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "handles.h"

using amp::DevBuf;
using amp::fail;

namespace amp {
void launchAssembleGather(const int32_t* zptr, const int32_t* contrib, const double* Kgeo, const double* k, int32_t ne,
                          double* values, int64_t nnz, int32_t nb, int32_t kk, hipStream_t st);
void launchParamCoefficients(const double* Phi, int32_t nm, int32_t ne, const double* theta, double* k, int32_t nb,
                             hipStream_t st);
}  // namespace amp

struct amp_asm {
  int32_t dim = 0, nn = 0, ne = 0, kpe = 0, nm = 0;
  int32_t n = 0;  // free DOFs
  std::vector<int32_t> rowptr, colidx;  // restricted pattern
  std::vector<int32_t> fullRowptr, fullColidx, slots, src, map;
  hipStream_t stream = nullptr;
  DevBuf zptr, contrib, Kgeo, Phi;
  ~amp_asm() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// physGrad (lagrange.ts:148-187) for P1: gradients of the nodal shape functions and det J (dN constant)
double p1Grad(int32_t dim, const double* X, double* G) {
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

}  // namespace

extern "C" {

int amp_asm_create(int32_t dim, int32_t nn, const double* nodes, int32_t ne, const int32_t* conn,
                   const uint8_t* freeMask, amp_asm** out) {
  if (!out || (dim != 2 && dim != 3) || nn < 0 || ne < 0 || (nn > 0 && !nodes) || (ne > 0 && !conn))
    return fail(AMP_ERR_ARG, "amp_asm_create: bad arguments");
  *out = nullptr;
  const int32_t k = dim + 1;
  for (int64_t q = 0; q < int64_t(ne) * k; q++)
    if (conn[q] < 0 || conn[q] >= nn) return fail(AMP_ERR_ARG, "amp_asm_create: connectivity out of range");
  std::unique_ptr<amp_asm> h(new amp_asm());
  h->dim = dim;
  h->nn = nn;
  h->ne = ne;
  h->kpe = k;
  // buildElementPattern (sparse.mjs:86-110): counting sort of each row's neighbours, sorted, duplicates dropped
  std::vector<int32_t> cap(size_t(nn) + 1, 0);
  for (int32_t i = 0; i < nn; i++) cap[i + 1] = 1;
  for (int64_t q = 0; q < int64_t(ne) * k; q++) cap[conn[q] + 1] += k - 1;
  for (int32_t i = 0; i < nn; i++) cap[i + 1] += cap[i];
  std::vector<int32_t> buf(cap[nn]);
  std::vector<int32_t> fillp(cap.begin(), cap.end() - 1);
  for (int32_t i = 0; i < nn; i++) buf[fillp[i]++] = i;
  for (int32_t e = 0; e < ne; e++)
    for (int a = 0; a < k; a++) {
      const int32_t r = conn[int64_t(k) * e + a];
      for (int b = 0; b < k; b++)
        if (b != a) buf[fillp[r]++] = conn[int64_t(k) * e + b];
    }
  h->fullRowptr.assign(size_t(nn) + 1, 0);
  int32_t nnz = 0;
  for (int32_t i = 0; i < nn; i++) {
    std::sort(buf.begin() + cap[i], buf.begin() + cap[i + 1]);
    for (int32_t p = cap[i]; p < cap[i + 1]; p++)
      if (p == cap[i] || buf[p] != buf[p - 1]) buf[nnz++] = buf[p];
    h->fullRowptr[i + 1] = nnz;
  }
  h->fullColidx.assign(buf.begin(), buf.begin() + nnz);
  // elementSlots (sparse.mjs:113-125): slot of (e, a, b) by binary search in row conn[e][a]
  h->slots.resize(size_t(ne) * k * k);
  for (int32_t e = 0; e < ne; e++)
    for (int a = 0; a < k; a++) {
      const int32_t r = conn[int64_t(k) * e + a];
      for (int b = 0; b < k; b++) {
        const int32_t c = conn[int64_t(k) * e + b];
        const int32_t* lo = h->fullColidx.data() + h->fullRowptr[r];
        const int32_t* hi = h->fullColidx.data() + h->fullRowptr[r + 1];
        h->slots[(size_t(e) * k + a) * k + b] = int32_t(std::lower_bound(lo, hi, c) - h->fullColidx.data());
      }
    }
  // restrict (scalar.ts:210-233): free nodes keep their order
  h->map.assign(nn, -1);
  int32_t nf = 0;
  for (int32_t i = 0; i < nn; i++)
    if (!freeMask || freeMask[i]) h->map[i] = nf++;
  h->n = nf;
  h->rowptr.assign(size_t(nf) + 1, 0);
  for (int32_t i = 0; i < nn; i++) {
    if (h->map[i] < 0) continue;
    for (int32_t q = h->fullRowptr[i]; q < h->fullRowptr[i + 1]; q++)
      if (h->map[h->fullColidx[q]] >= 0) {
        h->colidx.push_back(h->map[h->fullColidx[q]]);
        h->src.push_back(q);
      }
    h->rowptr[h->map[i] + 1] = int32_t(h->colidx.size());
  }
  // contributions of every restricted entry, in element order: (element, local entry a k + b of its matrix)
  std::vector<int32_t> zOfSlot(nnz, -1);
  for (size_t z = 0; z < h->src.size(); z++) zOfSlot[h->src[z]] = int32_t(z);
  const size_t nz = h->src.size();
  std::vector<int32_t> zptr(nz + 1, 0);
  for (size_t q = 0; q < h->slots.size(); q++)
    if (zOfSlot[h->slots[q]] >= 0) zptr[zOfSlot[h->slots[q]] + 1]++;
  for (size_t z = 0; z < nz; z++) zptr[z + 1] += zptr[z];
  std::vector<int32_t> contrib(2 * size_t(zptr[nz]));
  std::vector<int32_t> at(zptr.begin(), zptr.end() - 1);
  for (int32_t e = 0; e < ne; e++)
    for (int ab = 0; ab < k * k; ab++) {
      const int32_t z = zOfSlot[h->slots[size_t(e) * k * k + ab]];
      if (z < 0) continue;
      const int32_t p = at[z]++;
      contrib[2 * size_t(p)] = e;
      contrib[2 * size_t(p) + 1] = ab;
    }
  // geometric element matrices |T| grad N_a . grad N_b (the a <= b entries, mirrored as the reference mirrors them)
  std::vector<double> Kgeo(size_t(ne) * k * k);
  const double fact = dim == 2 ? 2.0 : 6.0;
  for (int32_t e = 0; e < ne; e++) {
    double X[12], G[12];
    for (int a = 0; a < k; a++)
      for (int d = 0; d < dim; d++) X[dim * a + d] = nodes[int64_t(dim) * conn[int64_t(k) * e + a] + d];
    const double det = p1Grad(dim, X, G);
    if (!(det > 0)) return fail(AMP_ERR_ARG, "amp_asm_create: non-positive Jacobian in element " + std::to_string(e));
    const double w = det / fact;
    double* Ke = Kgeo.data() + size_t(e) * k * k;
    for (int a = 0; a < k; a++)
      for (int b = a; b < k; b++) {
        double s = 0;
        for (int d = 0; d < dim; d++) s += G[dim * a + d] * G[dim * b + d];
        Ke[k * a + b] = Ke[k * b + a] = s * w;
      }
  }
  HIP_TRY(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  HIP_TRY(amp::upload(h->zptr, zptr));
  HIP_TRY(amp::upload(h->contrib, contrib));
  HIP_TRY(amp::upload(h->Kgeo, Kgeo));
  *out = h.release();
  return AMP_OK;
}

int amp_asm_get_pattern(const amp_asm* h, int32_t* n, int64_t* nnz, const int32_t** rowptr, const int32_t** colidx) {
  if (!h || !n || !nnz || !rowptr || !colidx) return fail(AMP_ERR_ARG, "amp_asm_get_pattern: bad arguments");
  *n = h->n;
  *nnz = int64_t(h->colidx.size());
  *rowptr = h->rowptr.data();
  *colidx = h->colidx.data();
  return AMP_OK;
}

int amp_assemble_device(amp_asm* h, const double* d_k, double* d_values, int32_t nb) {
  if (!h || nb < 0 || (nb > 0 && (!d_k || !d_values))) return fail(AMP_ERR_ARG, "amp_assemble_device: bad arguments");
  const int64_t nz = int64_t(h->colidx.size());
  if (nb == 0 || nz == 0) return AMP_OK;
  HIP_TRY(hipDeviceSynchronize());  // coefficients produced on any stream are complete
  amp::launchAssembleGather(h->zptr.as<int32_t>(), h->contrib.as<int32_t>(), h->Kgeo.as<double>(), d_k, h->ne,
                            d_values, nz, nb, h->kpe * h->kpe, h->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return AMP_OK;
}

int amp_asm_set_modes(amp_asm* h, int32_t nm, const double* Phi) {
  if (!h || nm < 1 || nm > 64 || !Phi) return fail(AMP_ERR_ARG, "amp_asm_set_modes: bad arguments");
  std::vector<double> p(Phi, Phi + size_t(h->ne) * nm);
  HIP_TRY(amp::upload(h->Phi, p));
  h->nm = nm;
  return AMP_OK;
}

int amp_param_coefficients_device(amp_asm* h, const double* d_theta, double* d_k, int32_t nb) {
  if (!h || nb < 0 || (nb > 0 && (!d_theta || !d_k))) return fail(AMP_ERR_ARG, "amp_param_coefficients: bad arguments");
  if (!h->nm) return fail(AMP_ERR_STATE, "amp_param_coefficients: no modes (amp_asm_set_modes)");
  if (nb == 0 || h->ne == 0) return AMP_OK;
  HIP_TRY(hipDeviceSynchronize());
  amp::launchParamCoefficients(h->Phi.as<double>(), h->nm, h->ne, d_theta, d_k, nb, h->stream);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(h->stream));
  return AMP_OK;
}

void amp_asm_free(amp_asm* h) { delete h; }

}  // extern "C"
