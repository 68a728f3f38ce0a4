// Device refinement drivers behind amp_refine / amp_refine_device (include/ampsolve.h).
//
// Iterative refinement restates solveRefined (packages/core/src/la/ldlt.mjs:253-275): x = solve(b), then up to
// maxSteps times r = b - A x (CSR row order, FP64), stop when |r| / |b| < tol, x += solve(r).
// GMRES-IR restates the restarted right-preconditioned GMRES of gmres-real.mjs:31-104 (Givens rotations on the
// Hessenberg column, true residual at exit) with the factor's solve as the preconditioner and the factor's solution
// as the initial guess; the Arnoldi step orthogonalizes with two passes of classical Gram-Schmidt (one reduction per
// pass on the GPU instead of one per basis vector) in place of modified Gram-Schmidt.
// Every vector lives in HBM in FP64; only the Hessenberg matrix and the scalars of the rotations are on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "factor_kernels.h"
#include "handles.h"
#include "refine_kernels.h"

using amp::DevBuf;
using amp::fail;

namespace {

int ensure(DevBuf& b, size_t bytes) {
  if (b.bytes < bytes) HIP_TRY(b.alloc(bytes));
  return AMP_OK;
}

int patternToDevice(amp_sym* h) {
  if (h->rowptr.empty()) return fail(AMP_ERR_STATE, "refine: the CSR pattern is unknown (amp_sym_set_pattern)");
  if (!h->dRowptr.p) {
    HIP_TRY(amp::upload(h->dRowptr, h->rowptr));
    HIP_TRY(amp::upload(h->dColidx, h->colidx));
  }
  return AMP_OK;
}

struct Ctx {
  amp_num* num;
  int32_t n;
  hipStream_t st;
  const int32_t *rp, *ci;
  const double* val;
  double* small;  // device scratch for reduction results (kMaxRestart + 2)
  double* partial;

  // host copy of out[0..nv) = V_j . w
  int dots(const double* V, int64_t ldv, int32_t nv, const double* w, double* host) {
    amp::launchMultiDot(V, ldv, nv, w, n, partial, small, st);
    HIP_TRY(hipMemcpyAsync(host, small, size_t(nv) * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMP_OK;
  }
  int norm(const double* w, double* out) {
    int rc = dots(w, 0, 1, w, out);
    *out = std::sqrt(*out);
    return rc;
  }
  // r = b - A x, returns |r|
  int residual(const double* b, const double* x, double* r, double* rn) {
    amp::launchResidual(rp, ci, val, x, b, r, n, st);
    return norm(r, rn);
  }
};

int runIR(Ctx& c, const double* d_b, double* d_x, double bn, const amp_refine_opts* o, amp_refine_info* info) {
  amp_num* num = c.num;
  int rc = amp::solveOnDevice(num, d_b, d_x, 1);
  if (rc) return rc;
  double rel = INFINITY;
  int32_t steps = 0;
  for (; steps <= o->maxSteps; steps++) {
    double rn = 0;
    rc = c.residual(d_b, d_x, num->r.as<double>(), &rn);
    if (rc) return rc;
    rel = rn / bn;
    if (rel < o->tol || steps == o->maxSteps) break;
    rc = amp::solveOnDevice(num, num->r.as<double>(), num->dvec.as<double>(), 1);
    if (rc) return rc;
    amp::launchAxpby(1.0, num->dvec.as<double>(), 1.0, d_x, c.n, c.st);
  }
  info->steps = steps;
  info->relResidual = rel;
  info->converged = rel < o->tol;
  return AMP_OK;
}

int runGmres(Ctx& c, const double* d_b, double* d_x, double bn, const amp_refine_opts* o, amp_refine_info* info) {
  amp_num* num = c.num;
  const int32_t n = c.n, m = o->restart;
  const int64_t ld = n;
  double* V = num->V.as<double>();
  double* Z = num->Z.as<double>();
  double* r = num->r.as<double>();
  double* w = num->w.as<double>();
  int rc = amp::solveOnDevice(num, d_b, d_x, 1);  // x0 = the factor's solution
  if (rc) return rc;
  std::vector<double> H(size_t(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1), y(m), h(m + 1), h2(m + 1);
  double beta = 0;
  rc = c.residual(d_b, d_x, r, &beta);
  if (rc) return rc;
  double relres = beta / bn;
  int32_t iters = 0;
  while (relres > o->tol && iters < o->maxSteps) {
    amp::launchAxpby(1.0 / beta, r, 0.0, V, n, c.st);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int32_t k = 0;
    for (; k < m && iters < o->maxSteps; k++) {
      iters++;
      double* Vk = V + k * ld;
      double* Zk = Z + k * ld;
      rc = amp::solveOnDevice(num, Vk, Zk, 1);
      if (rc) return rc;
      amp::launchResidual(c.rp, c.ci, c.val, Zk, nullptr, w, n, c.st);  // w = A z_k
      // two passes of classical Gram-Schmidt against V_0 .. V_k
      rc = c.dots(V, ld, k + 1, w, h.data());
      if (rc) return rc;
      HIP_TRY(hipMemcpyAsync(c.small, h.data(), size_t(k + 1) * 8, hipMemcpyHostToDevice, c.st));
      amp::launchMultiAxpy(V, ld, k + 1, c.small, w, n, c.st);
      rc = c.dots(V, ld, k + 1, w, h2.data());
      if (rc) return rc;
      HIP_TRY(hipMemcpyAsync(c.small, h2.data(), size_t(k + 1) * 8, hipMemcpyHostToDevice, c.st));
      amp::launchMultiAxpy(V, ld, k + 1, c.small, w, n, c.st);
      for (int32_t j = 0; j <= k; j++) H[size_t(j) * m + k] = h[j] + h2[j];
      double hn = 0;
      rc = c.norm(w, &hn);
      if (rc) return rc;
      H[size_t(k + 1) * m + k] = hn;
      if (hn > 0) amp::launchAxpby(1.0 / hn, w, 0.0, V + (k + 1) * ld, n, c.st);
      for (int32_t j = 0; j < k; j++) {
        const double a = H[size_t(j) * m + k], b = H[size_t(j + 1) * m + k];
        H[size_t(j) * m + k] = cs[j] * a + sn[j] * b;
        H[size_t(j + 1) * m + k] = -sn[j] * a + cs[j] * b;
      }
      const double a = H[size_t(k) * m + k], b = H[size_t(k + 1) * m + k];
      double d = std::hypot(a, b);
      if (!(d != 0) || std::isnan(d)) d = 1e-300;  // `Math.hypot(a, c) || 1e-300`
      cs[k] = a / d;
      sn[k] = b / d;
      H[size_t(k) * m + k] = d;
      H[size_t(k + 1) * m + k] = 0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      relres = std::fabs(g[k + 1]) / bn;
      if (relres <= o->tol || hn == 0) {
        k++;
        break;
      }
    }
    for (int32_t i = k - 1; i >= 0; i--) {
      double s = g[i];
      for (int32_t j = i + 1; j < k; j++) s -= H[size_t(i) * m + j] * y[j];
      y[i] = s / H[size_t(i) * m + i];
    }
    HIP_TRY(hipMemcpyAsync(c.small, y.data(), size_t(k) * 8, hipMemcpyHostToDevice, c.st));
    amp::launchCombine(Z, ld, k, c.small, d_x, n, c.st);
    if (relres <= o->tol || iters >= o->maxSteps) break;
    rc = c.residual(d_b, d_x, r, &beta);
    if (rc) return rc;
    relres = beta / bn;
  }
  double tr = 0;
  rc = c.residual(d_b, d_x, r, &tr);
  if (rc) return rc;
  info->steps = iters;
  info->relResidual = tr / bn;
  info->converged = info->relResidual <= std::max(o->tol * 10, 1e-14);
  return AMP_OK;
}

int refineDevice(amp_num* num, const double* d_values, const double* d_b, double* d_x, const amp_refine_opts* o,
                 amp_refine_info* info) {
  amp_sym* h = num->sym;
  const int32_t n = h->S.n;
  if (!num->ok) return fail(AMP_ERR_STATE, "refine: no successful numeric factorization");
  if (o->method != AMP_REFINE_IR && o->method != AMP_REFINE_GMRES) return fail(AMP_ERR_ARG, "refine: unknown method");
  if (o->method == AMP_REFINE_GMRES && (o->restart < 1 || o->restart > amp::kMaxRestart))
    return fail(AMP_ERR_ARG, "refine: restart must be in [1, 64]");
  if (o->maxSteps < 0) return fail(AMP_ERR_ARG, "refine: negative step count");
  int rc = patternToDevice(h);
  if (rc) return rc;
  const size_t vb = size_t(std::max(n, 1)) * 8;
  if ((rc = ensure(num->r, vb)) || (rc = ensure(num->dvec, vb)) || (rc = ensure(num->w, vb))) return rc;
  if ((rc = ensure(num->partial, size_t(amp::partialEntries(amp::kMaxRestart + 2)) * 8))) return rc;
  if ((rc = ensure(num->small, size_t(amp::kMaxRestart + 2) * 8))) return rc;
  if (o->method == AMP_REFINE_GMRES) {
    if ((rc = ensure(num->V, vb * (o->restart + 1))) || (rc = ensure(num->Z, vb * o->restart))) return rc;
  }
  if (!num->evR0) {
    HIP_TRY(hipEventCreate(&num->evR0));
    HIP_TRY(hipEventCreate(&num->evR1));
  }
  hipEvent_t e0 = num->evR0, e1 = num->evR1;  // the inner solves use ev0 / ev1 themselves
  HIP_TRY(hipEventRecord(e0, num->stream));
  Ctx c{num, n, num->stream, h->dRowptr.as<int32_t>(), h->dColidx.as<int32_t>(), d_values, num->small.as<double>(),
        num->partial.as<double>()};
  double bn = 0;
  rc = c.norm(d_b, &bn);
  if (rc) return rc;
  if (!(bn != 0) || std::isnan(bn)) bn = 1;  // `Math.sqrt(bn) || 1`
  rc = o->method == AMP_REFINE_IR ? runIR(c, d_b, d_x, bn, o, info) : runGmres(c, d_b, d_x, bn, o, info);
  if (rc) return rc;
  HIP_TRY(hipEventRecord(e1, num->stream));
  HIP_TRY(hipStreamSynchronize(num->stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
  info->ms = ms;
  return AMP_OK;
}

// Preconditioned conjugate gradients from the guess in d_x (Hestenes-Stiefel recurrences; two dot products per
// iteration read back for the step lengths and the stopping test on the recursively updated residual, with the true
// residual recomputed at exit).
int runPcg(amp_num* num, const double* d_values, const double* d_b, double* d_x, const amp_pcg_opts* o,
           amp_pcg_info* info) {
  amp_sym* h = num->sym;
  const int32_t n = h->S.n;
  if (!num->ok) return fail(AMP_ERR_STATE, "pcg: no successful numeric factorization");
  if (o->maxIter < 0) return fail(AMP_ERR_ARG, "pcg: negative iteration count");
  int rc = patternToDevice(h);
  if (rc) return rc;
  const size_t vb = size_t(std::max(n, 1)) * 8;
  if ((rc = ensure(num->r, vb)) || (rc = ensure(num->dvec, vb)) || (rc = ensure(num->w, vb)) ||
      (rc = ensure(num->V, 2 * vb)))
    return rc;
  if ((rc = ensure(num->partial, size_t(amp::partialEntries(amp::kMaxRestart + 2)) * 8))) return rc;
  if ((rc = ensure(num->small, size_t(amp::kMaxRestart + 2) * 8))) return rc;
  if (!num->evR0) {
    HIP_TRY(hipEventCreate(&num->evR0));
    HIP_TRY(hipEventCreate(&num->evR1));
  }
  HIP_TRY(hipEventRecord(num->evR0, num->stream));
  Ctx c{num, n, num->stream, h->dRowptr.as<int32_t>(), h->dColidx.as<int32_t>(), d_values, num->small.as<double>(),
        num->partial.as<double>()};
  double* r = num->r.as<double>();
  double* z = num->dvec.as<double>();
  double* q = num->w.as<double>();
  double* p = num->V.as<double>();
  double bn = 0, rn = 0;
  if ((rc = c.norm(d_b, &bn))) return rc;
  if (!(bn != 0) || std::isnan(bn)) bn = 1;
  if ((rc = c.residual(d_b, d_x, r, &rn))) return rc;
  info->initialResidual = rn / bn;
  double rel = rn / bn, rz = 0, pq = 0;
  int32_t it = 0;
  if (rel > o->tol && o->maxIter > 0) {
    amp::launchClearFlag(num->fc.err, num->stream);  // the solves' wait-failure flag, read once at the end
    if ((rc = amp::solveOnDeviceNoSync(num, r, z, 1))) return rc;
    HIP_TRY(hipMemcpyAsync(p, z, vb, hipMemcpyDeviceToDevice, num->stream));
    if ((rc = c.dots(r, 0, 1, z, &rz))) return rc;
    for (it = 1; it <= o->maxIter; it++) {
      amp::launchResidual(c.rp, c.ci, d_values, p, nullptr, q, n, num->stream);  // q = A p
      if ((rc = c.dots(p, 0, 1, q, &pq))) return rc;
      if (!(pq > 0)) break;  // the preconditioned operator is not positive on p: stop with the current iterate
      const double alpha = rz / pq;
      amp::launchAxpby(alpha, p, 1.0, d_x, n, num->stream);
      amp::launchAxpby(-alpha, q, 1.0, r, n, num->stream);
      if ((rc = c.norm(r, &rn))) return rc;
      rel = rn / bn;
      if (rel <= o->tol) break;
      if ((rc = amp::solveOnDeviceNoSync(num, r, z, 1))) return rc;
      double rzNew = 0;
      if ((rc = c.dots(r, 0, 1, z, &rzNew))) return rc;
      amp::launchAxpby(1.0, z, rzNew / rz, p, n, num->stream);  // p = z + beta p
      rz = rzNew;
    }
    it = std::min(it, o->maxIter);
    if ((rc = c.residual(d_b, d_x, r, &rn))) return rc;  // true residual at exit
    rel = rn / bn;
    if ((rc = amp::solveWaitFailed(num))) return rc;
  }
  HIP_TRY(hipEventRecord(num->evR1, num->stream));
  HIP_TRY(hipStreamSynchronize(num->stream));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, num->evR0, num->evR1));
  info->iters = it;
  info->relResidual = rel;
  info->converged = rel <= o->tol;
  info->ms = ms;
  return AMP_OK;
}

// Preconditioned CG on nb independent systems A_j x_j = b_j in lockstep, every A_j on the analysed pattern (values of
// system j at j nnz), one shared preconditioner (the factor held by num): each iteration's preconditioner step is one
// multi-right-hand-side solve of all systems. A system stops updating (zero step) once it meets tol; the loop ends
// when all have, or after maxIter iterations. The dot products of all systems are read back together (three per
// iteration); the true residual of every system is recomputed at exit.
int runPcgBatch(amp_num* num, const double* d_values, const double* d_b, int64_t bStride, double* d_x, int32_t nb,
                const amp_pcg_opts* o, amp_pcg_info* info) {
  amp_sym* h = num->sym;
  const int32_t n = h->S.n;
  const int64_t nnz = h->S.nnzPattern;
  if (!num->ok) return fail(AMP_ERR_STATE, "pcg: no successful numeric factorization");
  if (o->maxIter < 0) return fail(AMP_ERR_ARG, "pcg: negative iteration count");
  int rc = patternToDevice(h);
  if (rc) return rc;
  const size_t vb = size_t(std::max(n, 1)) * nb * 8;
  // r, z, q, p (nb vectors each), per-system coefficients (2 nb), reduction scratch
  if ((rc = ensure(num->r, vb)) || (rc = ensure(num->dvec, vb)) || (rc = ensure(num->w, vb)) ||
      (rc = ensure(num->V, vb)))
    return rc;
  if ((rc = ensure(num->partial, size_t(amp::partialEntries(std::max(nb, amp::kMaxRestart + 2))) * 8))) return rc;
  if ((rc = ensure(num->small, size_t(std::max(4 * nb, amp::kMaxRestart + 2)) * 8))) return rc;
  if (!num->evR0) {
    HIP_TRY(hipEventCreate(&num->evR0));
    HIP_TRY(hipEventCreate(&num->evR1));
  }
  hipStream_t st = num->stream;
  HIP_TRY(hipEventRecord(num->evR0, st));
  const int32_t* rp = h->dRowptr.as<int32_t>();
  const int32_t* ci = h->dColidx.as<int32_t>();
  double* r = num->r.as<double>();
  double* z = num->dvec.as<double>();
  double* q = num->w.as<double>();
  double* p = num->V.as<double>();
  double* part = num->partial.as<double>();
  double* dd = num->small.as<double>();  // [0, nb): reduction results; [nb, 2 nb): a_j; [2 nb, 3 nb): c_j
  std::vector<double> hv(nb), bn(nb), rn(nb), rz(nb), coef(2 * size_t(nb));
  std::vector<char> done(nb, 0);
  auto dots = [&](const double* u, const double* w, std::vector<double>& out) -> int {
    amp::launchBatchDot(u, w, n, nb, part, dd, st);
    HIP_TRY(hipMemcpyAsync(out.data(), dd, size_t(nb) * 8, hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    return AMP_OK;
  };
  auto setCoef = [&](const std::vector<double>& a, const std::vector<double>* c) -> int {
    for (int32_t j = 0; j < nb; j++) {
      coef[j] = a[j];
      coef[nb + j] = c ? (*c)[j] : 1.0;
    }
    HIP_TRY(hipMemcpyAsync(dd + nb, coef.data(), 2 * size_t(nb) * 8, hipMemcpyHostToDevice, st));
    return AMP_OK;
  };
  // |b_j|: a shared right-hand side is read through the residual of x = 0 (r = b)
  amp::launchBatchResidual(rp, ci, d_values, nnz, d_x, d_b, bStride, r, n, nb, st);
  {
    // b_j . b_j from r_j + A_j x_j would round differently; copy b_j into q and take its norm
    if (bStride == 0)
      for (int32_t j = 0; j < nb; j++)
        HIP_TRY(hipMemcpyAsync(q + int64_t(j) * n, d_b, size_t(n) * 8, hipMemcpyDeviceToDevice, st));
    else
      for (int32_t j = 0; j < nb; j++)
        HIP_TRY(hipMemcpyAsync(q + int64_t(j) * n, d_b + int64_t(j) * bStride, size_t(n) * 8,
                               hipMemcpyDeviceToDevice, st));
    if ((rc = dots(q, q, bn))) return rc;
  }
  if ((rc = dots(r, r, rn))) return rc;
  for (int32_t j = 0; j < nb; j++) {
    bn[j] = std::sqrt(bn[j]);
    if (!(bn[j] != 0) || std::isnan(bn[j])) bn[j] = 1;
    rn[j] = std::sqrt(rn[j]);
    info[j] = amp_pcg_info{};
    info[j].initialResidual = rn[j] / bn[j];
    done[j] = !(rn[j] / bn[j] > o->tol) || o->maxIter == 0;
  }
  int32_t it = 0;
  auto allDone = [&] {
    for (char d : done)
      if (!d) return false;
    return true;
  };
  if (!allDone()) {
    amp::launchClearFlag(num->fc.err, st);  // the solves' wait-failure flag, read once after the iterations
    if ((rc = amp::solveOnDeviceNoSync(num, r, z, nb))) return rc;
    HIP_TRY(hipMemcpyAsync(p, z, vb, hipMemcpyDeviceToDevice, st));
    if ((rc = dots(r, z, rz))) return rc;
    std::vector<double> alpha(nb), negAlpha(nb), beta(nb), pq(nb), rzNew(nb);
    for (it = 1; it <= o->maxIter && !allDone(); it++) {
      amp::launchBatchResidual(rp, ci, d_values, nnz, p, nullptr, 0, q, n, nb, st);  // q_j = A_j p_j
      if ((rc = dots(p, q, pq))) return rc;
      for (int32_t j = 0; j < nb; j++) {
        if (!done[j] && !(pq[j] > 0)) done[j] = 1;  // not positive on p_j: stop that system with its iterate
        alpha[j] = done[j] ? 0.0 : rz[j] / pq[j];
        negAlpha[j] = -alpha[j];
        if (!done[j]) info[j].iters = it;
      }
      if ((rc = setCoef(alpha, nullptr))) return rc;
      amp::launchBatchAxpby(dd + nb, p, dd + 2 * nb, d_x, n, nb, st);  // x_j += a_j p_j
      if ((rc = setCoef(negAlpha, nullptr))) return rc;
      amp::launchBatchAxpby(dd + nb, q, dd + 2 * nb, r, n, nb, st);  // r_j -= a_j q_j
      if ((rc = dots(r, r, rn))) return rc;
      for (int32_t j = 0; j < nb; j++)
        if (!done[j] && !(std::sqrt(rn[j]) / bn[j] > o->tol)) done[j] = 1;
      if (allDone()) break;
      if ((rc = amp::solveOnDeviceNoSync(num, r, z, nb))) return rc;
      if ((rc = dots(r, z, rzNew))) return rc;
      std::vector<double> one(nb, 1.0);
      for (int32_t j = 0; j < nb; j++) {
        beta[j] = done[j] ? 1.0 : rzNew[j] / rz[j];
        if (!done[j]) rz[j] = rzNew[j];
      }
      // p_j = z_j + beta_j p_j for the running systems (a converged one keeps its p: it takes zero steps)
      for (int32_t j = 0; j < nb; j++) one[j] = done[j] ? 0.0 : 1.0;
      if ((rc = setCoef(one, &beta))) return rc;
      amp::launchBatchAxpby(dd + nb, z, dd + 2 * nb, p, n, nb, st);
    }
  }
  if (it > 0 && (rc = amp::solveWaitFailed(num))) return rc;
  // true residuals at exit
  amp::launchBatchResidual(rp, ci, d_values, nnz, d_x, d_b, bStride, r, n, nb, st);
  if ((rc = dots(r, r, rn))) return rc;
  HIP_TRY(hipEventRecord(num->evR1, st));
  HIP_TRY(hipStreamSynchronize(st));
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, num->evR0, num->evR1));
  for (int32_t j = 0; j < nb; j++) {
    info[j].relResidual = std::sqrt(rn[j]) / bn[j];
    info[j].converged = info[j].relResidual <= o->tol;
    info[j].ms = ms;
  }
  return AMP_OK;
}

// Y = A^-1 (M X) for nvec vectors (eigen.ts:156-166, `op`): FP64 products with M on the analysed CSR pattern, then
// one multi-right-hand-side solve with the factor of A = K - sigma M held by num.
int runShiftInvert(amp_num* num, const double* d_M, const double* d_X, double* d_Y, int32_t nvec) {
  amp_sym* h = num->sym;
  const int32_t n = h->S.n;
  if (!num->ok) return fail(AMP_ERR_STATE, "shift-invert: no successful numeric factorization");
  int rc = patternToDevice(h);
  if (rc) return rc;
  if ((rc = ensure(num->opw, size_t(n) * nvec * 8))) return rc;
  double* w = num->opw.as<double>();
  amp::launchResidualMulti(h->dRowptr.as<int32_t>(), h->dColidx.as<int32_t>(), d_M, d_X, nullptr, w, n, nvec,
                           num->stream);
  HIP_TRY(hipGetLastError());
  return amp::solveOnDevice(num, w, d_Y, nvec);
}

}  // namespace

extern "C" {

int amp_pcg_batch_device(amp_num* num, const double* d_values, const double* d_b, int64_t bStride, double* d_x,
                         int32_t nb, const amp_pcg_opts* o, amp_pcg_info* info) {
  if (!num || !o || !info || nb < 1 || nb > 64 || bStride < 0 || (num->sym->S.n > 0 && (!d_values || !d_b || !d_x)))
    return fail(AMP_ERR_ARG, "amp_pcg_batch_device: bad arguments");
  if (num->sym->S.n == 0) {
    for (int32_t j = 0; j < nb; j++) info[j] = amp_pcg_info{0, 1, 0, 0, 0};
    return AMP_OK;
  }
  if (bStride != 0 && bStride < num->sym->S.n) return fail(AMP_ERR_ARG, "amp_pcg_batch_device: bStride < n");
  HIP_TRY(amp::afterCaller(num->stream, num->inEv));
  return runPcgBatch(num, d_values, d_b, bStride, d_x, nb, o, info);
}

int amp_shift_invert_apply_device(amp_num* num, const double* d_M, const double* d_X, double* d_Y, int32_t nvec) {
  if (!num || nvec < 0 || (nvec > 0 && num->sym->S.n > 0 && (!d_M || !d_X || !d_Y)))
    return fail(AMP_ERR_ARG, "amp_shift_invert_apply: bad arguments");
  if (num->sym->S.n == 0 || nvec == 0) return AMP_OK;
  if (int64_t(num->sym->S.n) * nvec > INT32_MAX) return fail(AMP_ERR_ARG, "amp_shift_invert_apply: n x nvec too large");
  HIP_TRY(amp::afterCaller(num->stream, num->inEv));
  return runShiftInvert(num, d_M, d_X, d_Y, nvec);
}

int amp_shift_invert_apply(amp_num* num, const double* M, const double* X, double* Y, int32_t nvec) {
  if (!num || nvec < 0 || (nvec > 0 && num->sym->S.n > 0 && (!M || !X || !Y)))
    return fail(AMP_ERR_ARG, "amp_shift_invert_apply: bad arguments");
  const int32_t n = num->sym->S.n;
  if (n == 0 || nvec == 0) return AMP_OK;
  if (int64_t(n) * nvec > INT32_MAX) return fail(AMP_ERR_ARG, "amp_shift_invert_apply: n x nvec too large");
  if (!num->ok) return fail(AMP_ERR_STATE, "shift-invert: no successful numeric factorization");
  const size_t nv = size_t(num->sym->S.nnzPattern), bytes = size_t(n) * nvec * 8;
  int rc = ensure(num->dvec, bytes);
  if (rc || (rc = ensure(num->w, bytes))) return rc;
  HIP_TRY(hipMemcpyAsync(num->values.p, M, nv * 8, hipMemcpyHostToDevice, num->stream));
  HIP_TRY(hipMemcpyAsync(num->dvec.p, X, bytes, hipMemcpyHostToDevice, num->stream));
  rc = runShiftInvert(num, num->values.as<double>(), num->dvec.as<double>(), num->w.as<double>(), nvec);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(Y, num->w.p, bytes, hipMemcpyDeviceToHost, num->stream));
  HIP_TRY(hipStreamSynchronize(num->stream));
  return AMP_OK;
}

int amp_pcg_device(amp_num* num, const double* d_values, const double* d_b, double* d_x, const amp_pcg_opts* o,
                   amp_pcg_info* info) {
  if (!num || !o || !info || (num->sym->S.n > 0 && (!d_values || !d_b || !d_x)))
    return fail(AMP_ERR_ARG, "amp_pcg_device: bad arguments");
  *info = amp_pcg_info{};
  if (num->sym->S.n == 0) {
    info->converged = 1;
    return AMP_OK;
  }
  HIP_TRY(amp::afterCaller(num->stream, num->inEv));
  return runPcg(num, d_values, d_b, d_x, o, info);
}


int amp_sym_set_pattern(amp_sym* h, const int32_t* rowptr, const int32_t* colidx) {
  if (!h || !rowptr || (h->S.n > 0 && !colidx)) return fail(AMP_ERR_ARG, "amp_sym_set_pattern: bad arguments");
  const int32_t n = h->S.n;
  if (int64_t(rowptr[n]) != h->S.nnzPattern) return fail(AMP_ERR_PATTERN, "amp_sym_set_pattern: entry count differs");
  h->rowptr.assign(rowptr, rowptr + n + 1);
  h->colidx.assign(colidx, colidx + rowptr[n]);
  h->dRowptr.reset();
  h->dColidx.reset();
  return AMP_OK;
}

int amp_refine_device(amp_num* num, const double* d_values, const double* d_b, double* d_x, const amp_refine_opts* o,
                      amp_refine_info* info) {
  if (!num || !o || !info || (num->sym->S.n > 0 && (!d_values || !d_b || !d_x)))
    return fail(AMP_ERR_ARG, "amp_refine: bad arguments");
  *info = amp_refine_info{};
  if (num->sym->S.n == 0) {
    info->converged = 1;
    return AMP_OK;
  }
  HIP_TRY(amp::afterCaller(num->stream, num->inEv));
  return refineDevice(num, d_values, d_b, d_x, o, info);
}

int amp_refine(amp_num* num, const double* values, const double* b, double* x, const amp_refine_opts* o,
               amp_refine_info* info) {
  if (!num || !o || !info || (num->sym->S.n > 0 && (!values || !b || !x)))
    return fail(AMP_ERR_ARG, "amp_refine: bad arguments");
  const int32_t n = num->sym->S.n;
  *info = amp_refine_info{};
  if (n == 0) {
    info->converged = 1;
    return AMP_OK;
  }
  if (!num->ok) return fail(AMP_ERR_STATE, "refine: no successful numeric factorization");
  const size_t nv = size_t(num->sym->S.nnzPattern);
  HIP_TRY(hipMemcpyAsync(num->values.p, values, nv * 8, hipMemcpyHostToDevice, num->stream));
  HIP_TRY(hipMemcpyAsync(num->b.p, b, size_t(n) * 8, hipMemcpyHostToDevice, num->stream));
  int rc = refineDevice(num, num->values.as<double>(), num->b.as<double>(), num->x.as<double>(), o, info);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(x, num->x.p, size_t(n) * 8, hipMemcpyDeviceToHost, num->stream));
  HIP_TRY(hipStreamSynchronize(num->stream));
  return AMP_OK;
}

}  // extern "C"
