// C ABI of the library (include/ampsolve.h): handles, HBM storage, plan upload and the factor / solve drivers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "handles.h"
#include "refine_kernels.h"

using amp::DevBuf;
using amp::fail;
using amp::upload;

namespace amp {
namespace {
thread_local std::string g_error;
}
int fail(int code, const std::string& msg) {
  g_error = msg;
  return code;
}
}  // namespace amp

namespace {

// Largest elimination multiplier of a diagonal block an FP32 factor keeps in FP32 arithmetic
constexpr double kDefaultGrowthLimit = 1e4;

// Factor and solve plans of the analysis; AMP_ERR_ARG when the system exceeds the plans' index ranges (the handle's
// plans are then unchanged).
int buildPlans(amp_sym* h, const amp::PlanParams& p) {
  try {
    amp::FactorPlan F = amp::buildFactorPlan(h->S, p);
    amp::PlanParams q = p;
    q.fusedPanel = 0;
    if (F.panelMaxTasks > amp::panelResidentTasks())  // a front's panel step must fit the device at once
      F = amp::buildFactorPlan(h->S, q);
    amp::FactorPlan F3 = F.panelCounters > 0 ? amp::buildFactorPlan(h->S, q) : amp::FactorPlan();
    if (!F3.launches.empty()) {  // the task arrays the two forms share on the device must be the same
      auto same = [](const auto& x, const auto& y) {
        return x.size() == y.size() && (x.empty() || std::memcmp(x.data(), y.data(), x.size() * sizeof(x[0])) == 0);
      };
      if (!same(F.zeros, F3.zeros) || !same(F.adst, F3.adst) || !same(F.smallFronts, F3.smallFronts) ||
          !same(F.ea, F3.ea) || !same(F.eaRanges, F3.eaRanges))
        return fail(AMP_ERR_INTERNAL, "analyze: the two forms of the factor plan disagree outside the panel chain");
    }
    amp::SolvePlan SP = amp::buildSolvePlan(h->S, F, p);
    h->F = std::move(F);
    h->F3 = std::move(F3);
    h->SP = std::move(SP);
  } catch (const std::length_error& e) {
    return fail(AMP_ERR_ARG, std::string("analyze: ") + e.what());
  }
  return AMP_OK;
}

int finishSym(amp_sym* h) {
  const amp::Symbolic& S = h->S;
  if (int rc = buildPlans(h, h->params)) return rc;
  h->diagSrc.assign(S.n, -1);
  for (int32_t j = 0; j < S.n; j++)
    for (int32_t q = S.lowerP[j]; q < S.lowerP[j + 1]; q++)
      if (S.lowerI[q] == j) h->diagSrc[j] = S.lowerSrc[q];
  return AMP_OK;
}

int symToDevice(amp_sym* h) {
  if (h->onDevice) return AMP_OK;
  const amp::Symbolic& S = h->S;
  const amp::FactorPlan& F = h->F;
  HIP_TRY(upload(h->superStart, S.superStart));
  HIP_TRY(upload(h->Sp, S.Sp));
  HIP_TRY(upload(h->Si, S.Si));
  HIP_TRY(upload(h->Lxp, S.Lxp));
  HIP_TRY(upload(h->Uoff, F.Uoff));
  HIP_TRY(upload(h->Woff, F.Woff));
  HIP_TRY(upload(h->childHead, S.childHead));
  HIP_TRY(upload(h->childNext, S.childNext));
  HIP_TRY(upload(h->rel, F.rel));
  HIP_TRY(upload(h->lowerP, S.lowerP));
  HIP_TRY(upload(h->lowerSrc, S.lowerSrc));
  HIP_TRY(upload(h->amap, F.amap));
  HIP_TRY(upload(h->adst, F.adst));
  HIP_TRY(upload(h->perm, S.perm));
  HIP_TRY(upload(h->dDiagSrc, h->diagSrc));
  HIP_TRY(upload(h->zeros, F.zeros));
  HIP_TRY(upload(h->smallFronts, F.smallFronts));
  HIP_TRY(upload(h->ea, F.ea));
  HIP_TRY(upload(h->eaRanges, F.eaRanges));
  HIP_TRY(upload(h->diag, F.diag));
  HIP_TRY(upload(h->panels, F.panels));
  HIP_TRY(upload(h->gemm, F.gemm));
  HIP_TRY(upload(h->splits, F.splits));
  HIP_TRY(upload(h->srefs, F.srefs));
  HIP_TRY(upload(h->sFronts, h->SP.fronts));
  HIP_TRY(upload(h->sBlocks, h->SP.blocks));
  h->dfp.zeros = h->zeros.as<amp::ZeroTask>();
  h->dfp.adst = h->adst.as<int64_t>();
  h->dfp.smallFronts = h->smallFronts.as<int32_t>();
  h->dfp.ea = h->ea.as<amp::EaTask>();
  h->dfp.eaRanges = h->eaRanges.as<amp::EaRange>();
  h->dfp.diag = h->diag.as<amp::DiagTask>();
  h->dfp.panels = h->panels.as<amp::PanelTask>();
  h->dfp.panelCounters = F.panelCounters;
  h->dfp.gemm = h->gemm.p;
  h->dfp.splits = h->splits.as<amp::SplitTask>();
  h->dfp.srefs = h->srefs.as<amp::SplitRef>();
  h->dfp.launches = F.launches;
  h->dfp3 = h->dfp;
  if (!h->F3.launches.empty()) {
    // the three-launch form shares every task array but those of the panel chain and the bulk launches
    HIP_TRY(upload(h->diag3, h->F3.diag));
    HIP_TRY(upload(h->gemm3, h->F3.gemm));
    HIP_TRY(upload(h->splits3, h->F3.splits));
    HIP_TRY(upload(h->srefs3, h->F3.srefs));
    h->dfp3.diag = h->diag3.as<amp::DiagTask>();
    h->dfp3.panels = nullptr;
    h->dfp3.panelCounters = 0;
    h->dfp3.gemm = h->gemm3.p;
    h->dfp3.splits = h->splits3.as<amp::SplitTask>();
    h->dfp3.srefs = h->srefs3.as<amp::SplitRef>();
    h->dfp3.launches = h->F3.launches;
  }
  h->dsp.fronts = h->sFronts.as<int32_t>();
  h->dsp.blocks = h->sBlocks.as<amp::BlockTask>();
  h->dsp.launches = h->SP.launches;
  h->dsp.fvEntries = h->SP.fvEntries;
  h->onDevice = true;
  return AMP_OK;
}

// entries of each of the four split arrays (a multiple of 64, so every array starts 128-byte aligned)
int64_t splitPitch(const amp::FactorPlan& F) { return (F.splitEntries + 63) / 64 * 64; }

template <typename T>
amp::FactorArgs<T> factorArgs(amp_num* num) {
  amp_sym* h = num->sym;
  amp::FactorArgs<T> a{};
  a.Lx = num->Lx.as<T>();
  a.U = num->U.as<T>();
  a.W = num->W.as<T>();
  a.D = num->D.as<T>();
  a.pert = num->pert.as<uint8_t>();
  a.failCol = num->failCol.as<int32_t>();
  a.superStart = h->superStart.as<int32_t>();
  a.Sp = h->Sp.as<int32_t>();
  a.Lxp = h->Lxp.as<int64_t>();
  a.Uoff = h->Uoff.as<int64_t>();
  a.Woff = h->Woff.as<int64_t>();
  a.childHead = h->childHead.as<int32_t>();
  a.childNext = h->childNext.as<int32_t>();
  a.rel = h->rel.as<int32_t>();
  a.Si = h->Si.as<int32_t>();
  a.lowerP = h->lowerP.as<int32_t>();
  a.lowerSrc = h->lowerSrc.as<int32_t>();
  a.amap = h->amap.as<int32_t>();
  a.ldlt = num->mode == AMP_LDLT;
  a.tau = 0;
  a.bulkBf16 = num->prec == AMP_F32_BF16X3;
  a.growthLimit = std::is_same<T, float>::value ? num->growthLimit : 0.0;
  a.fallbacks = num->counts.as<int32_t>() + 2;
  a.split = num->split.as<__bf16>();
  a.splitPitch = splitPitch(factorPlan(num));
  a.panelCtr = num->panelCtr.as<int32_t>();
  a.panelErr = num->counts.as<int32_t>() + 3;
  return a;
}

template <typename T>
amp::SolveArgs<T> solveArgs(amp_num* num) {
  amp_sym* h = num->sym;
  amp::SolveArgs<T> a{};
  a.Lx = num->Lx.as<T>();
  a.W = num->W.as<T>();
  a.D = num->mode == AMP_LDLT ? num->D.as<T>() : nullptr;
  a.superStart = h->superStart.as<int32_t>();
  a.Sp = h->Sp.as<int32_t>();
  a.Si = h->Si.as<int32_t>();
  a.Lxp = h->Lxp.as<int64_t>();
  a.Woff = h->Woff.as<int64_t>();
  a.childHead = h->childHead.as<int32_t>();
  a.childNext = h->childNext.as<int32_t>();
  a.rel = h->rel.as<int32_t>();
  a.perm = h->perm.as<int32_t>();
  a.n = h->S.n;
  a.ldlt = num->mode == AMP_LDLT;
  return a;
}

// max |a_jj| over the diagonal (diagScale, ldlt.mjs:419-424), computed on the host from a copy of the diagonal
__global__ void k_gather_diag(const double* __restrict__ values, const int32_t* __restrict__ diagSrc,
                              double* __restrict__ out, int32_t n) {
  for (int32_t j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x)
    out[j] = diagSrc[j] >= 0 ? values[diagSrc[j]] : 0.0;
}

template <typename T>
int runFactor(amp_num* num, const double* d_values, double pivotTol, amp_factor_info* info) {
  amp_sym* h = num->sym;
  const int32_t n = h->S.n;
  if (!num->U.p && factorPlan(num).uEntries > 0) HIP_TRY(num->U.alloc(size_t(factorPlan(num).uEntries) * sizeof(T)));  // released before
  amp::FactorArgs<T> a = factorArgs<T>(num);
  if (num->mode == AMP_LDLT && n > 0) {
    // tau = pivotTol * max |a_jj| needs the diagonal on the host before the launch sequence
    std::vector<double> diag(n);
    hipLaunchKernelGGL(k_gather_diag, dim3(std::min(1024, (n + 255) / 256)), dim3(256), 0, num->stream, d_values,
                       h->dDiagSrc.as<int32_t>(), num->y.as<double>(), n);
    HIP_TRY(hipMemcpyAsync(diag.data(), num->y.p, size_t(n) * 8, hipMemcpyDeviceToHost, num->stream));
    HIP_TRY(hipStreamSynchronize(num->stream));
    double amax = 0;
    for (double d : diag) amax = std::max(amax, std::fabs(d));
    a.tau = pivotTol * (amax > 0 ? amax : 1);
  }
  HIP_TRY(hipEventRecord(num->ev0, num->stream));
  if (n > 0)
    amp::launchFactorPlan<T>(deviceFactorPlan(num), a, d_values, num->stream, num->side, num->syncEv,
                             num->profiling ? &num->prof : nullptr);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(num->ev1, num->stream));
  int32_t counts[4] = {0, 0, 0, 0}, failCol = INT32_MAX;
  if (n > 0) {
    if (num->mode == AMP_LDLT) amp::launchLdltCounts<T>(a, n, num->counts.as<int32_t>(), num->stream);
    HIP_TRY(hipMemcpyAsync(&failCol, num->failCol.p, 4, hipMemcpyDeviceToHost, num->stream));
    HIP_TRY(hipMemcpyAsync(counts, num->counts.p, 16, hipMemcpyDeviceToHost, num->stream));
  }
  HIP_TRY(hipStreamSynchronize(num->stream));
  if (counts[3]) {
    amp::launchClearFlag(num->counts.as<int32_t>() + 3, num->stream);  // reports this factor only
    return fail(AMP_ERR_INTERNAL, "factor: a panel-step wait for the launch's workgroups timed out");
  }
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, num->ev0, num->ev1));
  for (auto it = num->kernelMs.begin(); it != num->kernelMs.end();)
    it = it->first.rfind("solve_", 0) == 0 ? std::next(it) : num->kernelMs.erase(it);
  if (num->profiling) {
    static const char* names[] = {"zero",       "assemble",  "small_front", "extend_add", "diag_block",
                                  "gemm_tile", "split_panel", "panel_step", "gemm_bf16x3"};
    for (size_t i = 0; i < num->prof.used; i++) {
      float k = 0;
      HIP_TRY(hipEventElapsedTime(&k, num->prof.events[2 * i], num->prof.events[2 * i + 1]));
      auto& slot = num->kernelMs[names[num->prof.kinds[i]]];
      slot.first += k;
      slot.second += 1;
      auto& lv = num->kernelMs[std::string(names[num->prof.kinds[i]]) + "@" + std::to_string(num->prof.levels[i])];
      lv.first += k;
      lv.second += 1;
    }
  }
  auto& fb = num->kernelMs["diag_fp64_redo"];  // FP32 factors: diagonal blocks redone in FP64 (count)
  fb.first = 0;
  fb.second = counts[2];
  num->ok = failCol == INT32_MAX;
  if (info) {
    info->ok = num->ok ? 1 : 0;
    info->failColumn = num->ok ? -1 : failCol;
    info->nNeg = num->mode == AMP_LDLT ? counts[0] : 0;
    info->nPerturbed = num->mode == AMP_LDLT ? counts[1] : 0;
    info->ms = ms;
  }
  return AMP_OK;
}

int ensureVecs(amp_num* num, int32_t nrhs) {
  const amp_sym* h = num->sym;
  const size_t n = size_t(std::max(h->S.n, 1)) * nrhs;
  const size_t es = num->prec == AMP_F64 ? 8 : 4;
  if (num->liveGraphs > 0) {  // captured graphs hold the current addresses: no buffer may move
    const size_t fvn = size_t(std::max<int64_t>(h->SP.fvEntries, 1)) * nrhs * es;
    const size_t pzn =
        size_t(std::max<int64_t>(h->SP.initSlots, 1)) * amp::kTileEdge * std::min(nrhs, amp::kFlowRhs) * es;
    if (num->b.bytes < n * 8 || num->y.bytes < std::max(n * es, size_t(h->S.n) * 8) || num->fv.bytes < fvn ||
        num->pz.bytes < pzn)
      return fail(AMP_ERR_STATE, "solve: " + std::to_string(nrhs) +
                                     " right-hand sides need larger solve buffers, which captured graphs of this "
                                     "factor still use (free them first)");
    return AMP_OK;
  }
  if (num->b.bytes < n * 8) {
    HIP_TRY(num->b.alloc(n * 8));
    HIP_TRY(num->x.alloc(n * 8));
  }
  if (num->y.bytes < std::max(n * es, size_t(h->S.n) * 8)) HIP_TRY(num->y.alloc(std::max(n * es, size_t(h->S.n) * 8)));
  const size_t fvn = size_t(std::max<int64_t>(h->SP.fvEntries, 1)) * nrhs * es;
  if (num->fv.bytes < fvn) HIP_TRY(num->fv.alloc(fvn));
  const size_t pzn = size_t(std::max<int64_t>(h->SP.initSlots, 1)) * amp::kTileEdge * std::min(nrhs, amp::kFlowRhs) * es;
  if (num->pz.bytes < pzn) HIP_TRY(num->pz.alloc(pzn));
  return AMP_OK;
}

// The solve with the given scratch (y, fv, pz sized as ensureVecs sizes them). record: keep the per-class times of
// the handle (host bookkeeping); the allocation-free entry point passes its own scratch and records the total only.
int runSolveWith(amp_num* num, const double* d_b, double* d_x, int32_t nrhs, void* y, void* fv, void* pz,
                 bool record) {
  amp::ProfileSink* prof = record && num->profiling ? &num->solveProf : nullptr;
  if (prof) prof->used = 0;
  // the wait-failure flag reports this solve only (cleared by a kernel: no memset node if the caller captures)
  amp::launchClearFlag(num->fc.err, num->stream);
  HIP_TRY(hipEventRecord(num->ev0, num->stream));
  if (num->prec == AMP_F64)
    HIP_TRY((amp::launchSolvePlan<double, double>(num->sym->dsp, solveArgs<double>(num), d_b, d_x, nrhs,
                                                  static_cast<double*>(y), static_cast<double*>(fv),
                                                  static_cast<double*>(pz), num->fc, num->stream, prof)));
  else
    HIP_TRY((amp::launchSolvePlan<float, float>(num->sym->dsp, solveArgs<float>(num), d_b, d_x, nrhs,
                                                static_cast<float*>(y), static_cast<float*>(fv),
                                                static_cast<float*>(pz), num->fc, num->stream, prof)));
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipEventRecord(num->ev1, num->stream));
  int32_t waitErr = 0;
  HIP_TRY(hipMemcpyAsync(&waitErr, num->fc.err, 4, hipMemcpyDeviceToHost, num->stream));
  HIP_TRY(hipStreamSynchronize(num->stream));
  if (waitErr) return fail(AMP_ERR_INTERNAL, "solve: a block dependency wait timed out");
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, num->ev0, num->ev1));
  num->solveMs = ms;
  return AMP_OK;
}

int runSolve(amp_num* num, const double* d_b, double* d_x, int32_t nrhs) {
  if (!num->ok) return fail(AMP_ERR_STATE, "solve: no successful numeric factorization");
  int rc = ensureVecs(num, nrhs);
  if (rc) return rc;
  rc = runSolveWith(num, d_b, d_x, nrhs, num->y.p, num->fv.p, num->pz.p, true);
  if (rc) return rc;
  const float ms = num->solveMs;
  amp::ProfileSink* prof = num->profiling ? &num->solveProf : nullptr;
  for (auto it = num->kernelMs.begin(); it != num->kernelMs.end();)
    it = it->first.rfind("solve_", 0) == 0 ? num->kernelMs.erase(it) : std::next(it);
  auto& total = num->kernelMs["solve_total"];  // recorded with or without profiling
  total.first = ms;
  total.second = 1;
  if (prof) {
    static const char* names[] = {"solve_small_fwd", "solve_gather", "solve_fwd_flow", "solve_small_bwd",
                                  "solve_bwd_flow", "solve_bwd_gather", "solve_bwd_init", "solve_fwd_fused",
                                  "solve_bwd_fused"};
    for (size_t i = 0; i < prof->used; i++) {
      float k = 0;
      HIP_TRY(hipEventElapsedTime(&k, prof->events[2 * i], prof->events[2 * i + 1]));
      auto& slot = num->kernelMs[names[prof->kinds[i] - 16]];
      slot.first += k;
      slot.second += 1;
      auto& lv = num->kernelMs[std::string(names[prof->kinds[i] - 16]) + "@" + std::to_string(prof->levels[i])];
      lv.first += k;
      lv.second += 1;
    }
  }
  return AMP_OK;
}

}  // namespace

namespace amp {
int solveOnDevice(amp_num* num, const double* d_b, double* d_x, int32_t nrhs) { return runSolve(num, d_b, d_x, nrhs); }

int solveOnDeviceNoSync(amp_num* num, const double* d_b, double* d_x, int32_t nrhs) {
  if (!num->ok) return fail(AMP_ERR_STATE, "solve: no successful numeric factorization");
  int rc = ensureVecs(num, nrhs);
  if (rc) return rc;
  if (num->prec == AMP_F64)
    HIP_TRY((amp::launchSolvePlan<double, double>(num->sym->dsp, solveArgs<double>(num), d_b, d_x, nrhs,
                                                  num->y.as<double>(), num->fv.as<double>(), num->pz.as<double>(),
                                                  num->fc, num->stream, nullptr)));
  else
    HIP_TRY((amp::launchSolvePlan<float, float>(num->sym->dsp, solveArgs<float>(num), d_b, d_x, nrhs,
                                                num->y.as<float>(), num->fv.as<float>(), num->pz.as<float>(), num->fc,
                                                num->stream, nullptr)));
  HIP_TRY(hipGetLastError());
  return AMP_OK;
}

int solveWaitFailed(amp_num* num) {
  int32_t waitErr = 0;
  HIP_TRY(hipMemcpyAsync(&waitErr, num->fc.err, 4, hipMemcpyDeviceToHost, num->stream));
  HIP_TRY(hipStreamSynchronize(num->stream));
  return waitErr ? fail(AMP_ERR_INTERNAL, "solve: a block dependency wait timed out") : AMP_OK;
}

SolveScratch solveScratchNeed(const amp_num* num, int32_t nrhs) {
  const amp_sym* h = num->sym;
  const size_t es = num->prec == AMP_F64 ? 8 : 4;
  const size_t n = size_t(std::max(h->S.n, 1)) * nrhs;
  auto up = [](size_t b) { return (b + 255) / 256 * 256; };
  SolveScratch s;
  s.yBytes = up(std::max(n * es, size_t(h->S.n) * 8));
  s.fvBytes = up(size_t(std::max<int64_t>(h->SP.fvEntries, 1)) * nrhs * es);
  s.pzBytes = up(size_t(std::max<int64_t>(h->SP.initSlots, 1)) * kTileEdge * std::min(nrhs, kFlowRhs) * es);
  return s;
}

int solveWithScratch(amp_num* num, const double* d_b, double* d_x, int32_t nrhs, void* y, void* fv, void* pz) {
  return runSolveWith(num, d_b, d_x, nrhs, y, fv, pz, false);
}
}  // namespace amp

extern "C" {

const char* amp_last_error(void) { return amp::g_error.c_str(); }

int amp_amd_order(int32_t n, const int32_t* rowptr, const int32_t* colidx, double denseFactor, int32_t* perm) {
  if (n < 0 || (n > 0 && (!rowptr || !colidx || !perm))) return fail(AMP_ERR_ARG, "amp_amd_order: bad arguments");
  if (!amp::amdOrder(n, rowptr, colidx, denseFactor, perm)) return fail(AMP_ERR_INTERNAL, "amdOrder: internal error");
  return AMP_OK;
}

int amp_nd_order(int32_t n, const int32_t* rowptr, const int32_t* colidx, const double* x, const double* y, int32_t leaf,
                 int32_t* perm) {
  if (n < 0 || leaf < 1 || (n > 0 && (!rowptr || !colidx || !x || !y || !perm)))
    return fail(AMP_ERR_ARG, "amp_nd_order: bad arguments");
  for (int32_t i = 0; i < n; i++)
    if (rowptr[i + 1] < rowptr[i]) return fail(AMP_ERR_ARG, "amp_nd_order: rowptr decreases");
  for (int32_t p = 0; p < (n > 0 ? rowptr[n] : 0); p++)
    if (colidx[p] < 0 || colidx[p] >= n) return fail(AMP_ERR_ARG, "amp_nd_order: column index out of range");
  amp::ndOrder(n, rowptr, colidx, x, y, leaf, perm);
  return AMP_OK;
}

int amp_pattern_fingerprint(int32_t n, const int32_t* rowptr, const int32_t* colidx, char* out) {
  if (n < 0 || !rowptr || (n > 0 && !colidx) || !out) return fail(AMP_ERR_ARG, "amp_pattern_fingerprint: bad arguments");
  const std::string k = amp::patternFingerprint(n, rowptr, colidx);
  std::snprintf(out, 48, "%s", k.c_str());
  return AMP_OK;
}

int amp_analyze(int32_t n, const int32_t* rowptr, const int32_t* colidx, const int32_t* perm, int32_t relax,
                amp_sym** out) {
  if (!out || n < 0 || !rowptr || (n > 0 && !colidx)) return fail(AMP_ERR_ARG, "amp_analyze: bad arguments");
  *out = nullptr;
  std::unique_ptr<amp_sym> h(new amp_sym());
  if (!amp::analyzePattern(n, rowptr, colidx, perm, relax != 0, h->S))
    return fail(AMP_ERR_INTERNAL, "analyze: amdOrder internal error");
  h->rowptr.assign(rowptr, rowptr + n + 1);
  h->colidx.assign(colidx, colidx + rowptr[n]);
  if (int rc = finishSym(h.get())) return rc;
  *out = h.release();
  return AMP_OK;
}

int amp_sym_import(const amp_sym_desc* d, amp_sym** out) {
  if (!d || !out || d->n < 0 || d->nsuper < 0) return fail(AMP_ERR_ARG, "amp_sym_import: bad arguments");
  *out = nullptr;
  const int32_t n = d->n, ns = d->nsuper;
  if (n > 0 && (!d->perm || !d->parent || !d->superStart || !d->Sp || !d->Si || !d->lowerP || !d->lowerI || !d->lowerSrc))
    return fail(AMP_ERR_ARG, "amp_sym_import: missing array");
  std::unique_ptr<amp_sym> h(new amp_sym());
  amp::Symbolic& S = h->S;
  S.n = n;
  S.nsuper = ns;
  S.nnzPattern = d->nnzPattern;
  S.patternKey = d->patternKey ? d->patternKey : "";
  S.perm.assign(d->perm, d->perm + n);
  S.pinv.resize(n);
  for (int32_t i = 0; i < n; i++) {
    if (S.perm[i] < 0 || S.perm[i] >= n) return fail(AMP_ERR_ARG, "amp_sym_import: perm out of range");
    S.pinv[S.perm[i]] = i;
  }
  // every index the plans and kernels will follow is checked here, so that a corrupted descriptor is an argument
  // error rather than an out-of-range access later
  auto bad = [](const char* what) { return fail(AMP_ERR_ARG, std::string("amp_sym_import: ") + what); };
  if ((n > 0) != (ns > 0) || ns > n) return bad("nsuper does not fit n");
  if (d->nnzPattern < 0) return bad("nnzPattern");
  {
    std::vector<char> seen(n, 0);
    for (int32_t i = 0; i < n; i++) {
      if (seen[S.perm[i]]) return bad("perm is not a permutation");
      seen[S.perm[i]] = 1;
    }
  }
  S.parent.assign(d->parent, d->parent + n);
  for (int32_t j = 0; j < n; j++)
    if (S.parent[j] != -1 && (S.parent[j] <= j || S.parent[j] >= n)) return bad("parent is not a postordered tree");
  S.superStart.assign(d->superStart, d->superStart + ns + 1);
  if (ns > 0 && (S.superStart[ns] != n || S.superStart[0] != 0)) return bad("superStart");
  for (int32_t s = 0; s < ns; s++)
    if (S.superStart[s + 1] <= S.superStart[s]) return bad("superStart does not increase");
  S.Sp.assign(d->Sp, d->Sp + ns + 1);
  if (ns > 0 && S.Sp[0] != 0) return bad("Sp[0]");
  for (int32_t s = 0; s < ns; s++)
    if (int64_t(S.Sp[s + 1]) - S.Sp[s] < int64_t(S.superStart[s + 1]) - S.superStart[s])
      return bad("Sp shorter than the supernode");
  S.Si.assign(d->Si, d->Si + (ns ? d->Sp[ns] : 0));
  for (int32_t s = 0; s < ns; s++) {  // own columns first, then strictly increasing rows below them
    const int32_t f = S.superStart[s], nc = S.superStart[s + 1] - f;
    for (int32_t r = 0; r < S.Sp[s + 1] - S.Sp[s]; r++) {
      const int32_t i = S.Si[S.Sp[s] + r];
      if (r < nc ? i != f + r : (i >= n || i <= S.Si[S.Sp[s] + r - 1])) return bad("Si row structure");
    }
  }
  S.lowerP.assign(d->lowerP, d->lowerP + n + 1);
  if (n > 0 && S.lowerP[0] != 0) return bad("lowerP[0]");
  for (int32_t j = 0; j < n; j++)
    if (S.lowerP[j + 1] < S.lowerP[j]) return bad("lowerP does not increase");
  S.nnzA = n ? d->lowerP[n] : 0;
  S.lowerI.assign(d->lowerI, d->lowerI + S.nnzA);
  S.lowerSrc.assign(d->lowerSrc, d->lowerSrc + S.nnzA);
  for (int64_t q = 0; q < S.nnzA; q++)
    if (S.lowerSrc[q] < 0 || S.lowerSrc[q] >= d->nnzPattern) return bad("lowerSrc out of the pattern's range");
  {  // every lower entry of a column lies in its supernode's row structure
    std::vector<int32_t> mark(n, -1);
    for (int32_t s = 0; s < ns; s++) {
      for (int32_t r = S.Sp[s]; r < S.Sp[s + 1]; r++) mark[S.Si[r]] = s;
      for (int32_t j = S.superStart[s]; j < S.superStart[s + 1]; j++)
        for (int32_t q = S.lowerP[j]; q < S.lowerP[j + 1]; q++)
          if (S.lowerI[q] < j || S.lowerI[q] >= n || mark[S.lowerI[q]] != s) return bad("lowerI outside the row structure");
    }
  }
  S.colSuper.resize(n);
  for (int32_t s = 0; s < ns; s++)
    for (int32_t j = S.superStart[s]; j < S.superStart[s + 1]; j++) S.colSuper[j] = s;
  S.superParent.resize(ns);
  for (int32_t s = 0; s < ns; s++) {
    const int32_t pl = S.parent[S.superStart[s + 1] - 1];
    S.superParent[s] = pl < 0 ? -1 : S.colSuper[pl];
    if (S.superParent[s] != -1 && S.superParent[s] <= s) return fail(AMP_ERR_ARG, "amp_sym_import: tree not postordered");
  }
  S.childHead.assign(ns, -1);
  S.childNext.assign(ns, -1);
  for (int32_t s = 0; s < ns; s++) {
    const int32_t P = S.superParent[s];
    if (P == -1) continue;
    S.childNext[s] = S.childHead[P];
    S.childHead[P] = s;
  }
  {  // a child's rows below its columns all appear in its parent's structure (the relative maps need them)
    std::vector<int32_t> mark(n, -1);
    for (int32_t P = 0; P < ns; P++) {
      for (int32_t r = S.Sp[P]; r < S.Sp[P + 1]; r++) mark[S.Si[r]] = P;
      for (int32_t c = S.childHead[P]; c != -1; c = S.childNext[c])
        for (int32_t r = S.Sp[c] + (S.superStart[c + 1] - S.superStart[c]); r < S.Sp[c + 1]; r++)
          if (mark[S.Si[r]] != P) return bad("a child's rows are missing from its parent's structure");
    }
  }
  S.Lxp.assign(ns + 1, 0);
  S.maxFront = 0;
  for (int32_t s = 0; s < ns; s++) {
    const int32_t m = S.Sp[s + 1] - S.Sp[s], nc = S.superStart[s + 1] - S.superStart[s];
    if (m < nc) return fail(AMP_ERR_ARG, "amp_sym_import: Sp shorter than the supernode");
    S.Lxp[s + 1] = S.Lxp[s] + int64_t(m) * nc;
    S.maxFront = std::max(S.maxFront, m);
  }
  S.nnzLStored = S.Lxp[ns];
  S.colCount.assign(n, 0);
  S.nnzL = 0;
  S.flops = 0;
  for (int32_t s = 0; s < ns; s++) {
    const int32_t m = S.Sp[s + 1] - S.Sp[s];
    for (int32_t j = S.superStart[s]; j < S.superStart[s + 1]; j++) {
      const int32_t cc = d->colCount ? d->colCount[j] : m - (j - S.superStart[s]);
      if (cc < 1 || cc > n - j) return bad("colCount out of range");
      S.colCount[j] = cc;
      S.nnzL += cc;
      S.flops += double(cc) * cc;
    }
  }
  S.fill = S.nnzA ? double(S.nnzL) / double(S.nnzA) : 1;
  if (int rc = finishSym(h.get())) return rc;
  *out = h.release();
  return AMP_OK;
}

int amp_sym_get_info(const amp_sym* h, amp_sym_info* info) {
  if (!h || !info) return fail(AMP_ERR_ARG, "amp_sym_get_info: bad arguments");
  const amp::Symbolic& S = h->S;
  info->n = S.n;
  info->nsuper = S.nsuper;
  info->maxFront = S.maxFront;
  info->nlevels = h->F.nlevels;
  info->nnzL = S.nnzL;
  info->nnzA = S.nnzA;
  info->nnzLStored = S.nnzLStored;
  info->nnzPattern = S.nnzPattern;
  info->flops = S.flops;
  info->fill = S.fill;
  info->gemmFlops = h->F.gemmFlops;
  info->bulkBytes = h->F.bulkBytes;
  info->bulkLaunches = h->F.bulkLaunches;
  std::snprintf(info->patternKey, sizeof info->patternKey, "%s", S.patternKey.c_str());
  return AMP_OK;
}

int amp_sym_get_arrays(const amp_sym* h, amp_sym_arrays* a) {
  if (!h || !a) return fail(AMP_ERR_ARG, "amp_sym_get_arrays: bad arguments");
  const amp::Symbolic& S = h->S;
  a->perm = S.perm.data();
  a->pinv = S.pinv.data();
  a->parent = S.parent.data();
  a->colCount = S.colCount.data();
  a->lowerP = S.lowerP.data();
  a->lowerI = S.lowerI.data();
  a->lowerSrc = S.lowerSrc.data();
  a->superStart = S.superStart.data();
  a->colSuper = S.colSuper.data();
  a->superParent = S.superParent.data();
  a->childHead = S.childHead.data();
  a->childNext = S.childNext.data();
  a->Sp = S.Sp.data();
  a->Si = S.Si.data();
  a->Lxp = S.Lxp.data();
  return AMP_OK;
}

void amp_sym_free(amp_sym* h) { delete h; }

int amp_num_create(amp_sym* h, int32_t mode, int32_t precision, amp_num** out) {
  if (!h || !out || (mode != AMP_CHOLESKY && mode != AMP_LDLT) ||
      (precision != AMP_F64 && precision != AMP_F32 && precision != AMP_F32_BF16X3))
    return fail(AMP_ERR_ARG, "amp_num_create: bad arguments");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return fail(AMP_ERR_HIP, "amp_num_create: no HIP device");
  int rc = symToDevice(h);
  if (rc) return rc;
  std::unique_ptr<amp_num> num(new amp_num());
  num->sym = h;
  h->liveNums++;
  num->mode = mode;
  num->prec = precision;
  const size_t es = precision == AMP_F64 ? 8 : 4;
  const amp::FactorPlan& F = factorPlan(num.get());
  HIP_TRY(hipStreamCreateWithFlags(&num->stream, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&num->side, hipStreamNonBlocking));
  num->syncEv.resize(size_t(F.signals) + 2, nullptr);  // signals, the final join, the initial fork
  for (hipEvent_t& e : num->syncEv) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  HIP_TRY(hipEventCreateWithFlags(&num->inEv, hipEventDisableTiming));
  HIP_TRY(hipEventCreate(&num->ev0));
  HIP_TRY(hipEventCreate(&num->ev1));
  HIP_TRY(num->Lx.alloc(size_t(F.nnzLStored) * es));
  HIP_TRY(num->U.alloc(size_t(F.uEntries) * es));
  HIP_TRY(num->W.alloc(size_t(F.wEntries) * es));
  HIP_TRY(num->D.alloc(size_t(std::max(h->S.n, 1)) * es));
  HIP_TRY(num->pert.alloc(size_t(std::max(h->S.n, 1))));
  HIP_TRY(num->failCol.alloc(4));
  HIP_TRY(num->counts.alloc(16));  // nNeg, nPerturbed, FP64 diagonal-block redos, panel-step wait failure
  HIP_TRY(hipMemset(num->counts.p, 0, 16));
  if (F.panelCounters > 0) HIP_TRY(num->panelCtr.alloc(size_t(F.panelCounters) * 4));
  if (precision == AMP_F32_BF16X3 && F.splitEntries > 0)
    HIP_TRY(num->split.alloc(size_t(4 * splitPitch(F)) * sizeof(uint16_t)));
  // static pivoting (LDL^T) can grow the multipliers; a Cholesky factor of an SPD matrix cannot (|L|^2 <= |A|)
  num->growthLimit = precision == AMP_F64 || mode == AMP_CHOLESKY ? 0.0 : kDefaultGrowthLimit;
  HIP_TRY(num->values.alloc(size_t(std::max<int64_t>(h->S.nnzPattern, 1)) * 8));
  {
    const amp::SolvePlan& SP = h->SP;
    int32_t nflow = 0;
    for (const amp::SolveLaunch& L : SP.launches) nflow += amp::solveKindIsFlow(L.kind);
    // flags of both directions, error flag, pass number, task counters of the flow and fused launches, counters of
    // the fused launches
    const size_t bytes = size_t(2 * SP.flowFlags + 2 + nflow + SP.doneSlots) * 4;
    HIP_TRY(num->flow.alloc(bytes));
    HIP_TRY(hipMemset(num->flow.p, 0, bytes));
    int32_t* w = num->flow.as<int32_t>();
    num->fc.fwdFlag = w;
    num->fc.bwdFlag = w + SP.flowFlags;
    num->fc.err = w + 2 * SP.flowFlags;
    num->fc.epoch = w + 2 * SP.flowFlags + 1;
    num->fc.ctr = w + 2 * SP.flowFlags + 2;
    num->fc.nctr = nflow;
    num->fc.done = num->fc.ctr + nflow;
    num->fc.ndone = int32_t(SP.doneSlots);
    // solved-value mailboxes of both directions (pass number 0 in every word: never current)
    const size_t boxBytes = size_t(std::max<int64_t>(SP.flowFlags, 1)) * amp::kBoxWords * 8;
    HIP_TRY(num->boxes.alloc(2 * boxBytes));
    HIP_TRY(hipMemset(num->boxes.p, 0, 2 * boxBytes));
    num->fc.fwdBox = num->boxes.as<uint64_t>();
    num->fc.bwdBox = num->fc.fwdBox + boxBytes / 8;
  }
  rc = ensureVecs(num.get(), 1);
  if (rc) return rc;
  *out = num.release();
  return AMP_OK;
}

int amp_factor_device(amp_num* num, const double* d_values, int64_t nvalues, double pivotTol, amp_factor_info* info) {
  if (!num || (nvalues > 0 && !d_values)) return fail(AMP_ERR_ARG, "amp_factor: bad arguments");
  if (nvalues != num->sym->S.nnzPattern)
    return fail(AMP_ERR_PATTERN, "factorNumeric: " + std::to_string(nvalues) + " values for a pattern with " +
                                     std::to_string(num->sym->S.nnzPattern) + " entries");
  num->ok = false;
  HIP_TRY(amp::afterCaller(num->stream, num->inEv));
  return num->prec == AMP_F64 ? runFactor<double>(num, d_values, pivotTol, info)
                              : runFactor<float>(num, d_values, pivotTol, info);
}

int amp_factor(amp_num* num, const double* values, int64_t nvalues, double pivotTol, amp_factor_info* info) {
  if (!num || (nvalues > 0 && !values)) return fail(AMP_ERR_ARG, "amp_factor: bad arguments");
  if (nvalues != num->sym->S.nnzPattern)
    return fail(AMP_ERR_PATTERN, "factorNumeric: " + std::to_string(nvalues) + " values for a pattern with " +
                                     std::to_string(num->sym->S.nnzPattern) + " entries");
  if (nvalues > 0)
    HIP_TRY(hipMemcpyAsync(num->values.p, values, size_t(nvalues) * 8, hipMemcpyHostToDevice, num->stream));
  return amp_factor_device(num, num->values.as<double>(), nvalues, pivotTol, info);
}

int amp_solve_device(amp_num* num, const double* d_b, double* d_x, int32_t nrhs) {
  if (!num || nrhs < 0 || (nrhs > 0 && num->sym->S.n > 0 && (!d_b || !d_x))) return fail(AMP_ERR_ARG, "amp_solve: bad arguments");
  HIP_TRY(amp::afterCaller(num->stream, num->inEv));
  return runSolve(num, d_b, d_x, nrhs);
}

int amp_solve(amp_num* num, const double* b, double* x, int32_t nrhs) {
  if (!num || nrhs < 0 || (nrhs > 0 && num->sym->S.n > 0 && (!b || !x))) return fail(AMP_ERR_ARG, "amp_solve: bad arguments");
  if (!num->ok) return fail(AMP_ERR_STATE, "solve: no successful numeric factorization");
  const int32_t n = num->sym->S.n;
  if (n == 0 || nrhs == 0) return AMP_OK;
  int rc = ensureVecs(num, nrhs);
  if (rc) return rc;
  const size_t bytes = size_t(n) * nrhs * 8;
  HIP_TRY(hipMemcpyAsync(num->b.p, b, bytes, hipMemcpyHostToDevice, num->stream));
  rc = runSolve(num, num->b.as<double>(), num->x.as<double>(), nrhs);
  if (rc) return rc;
  HIP_TRY(hipMemcpyAsync(x, num->x.p, bytes, hipMemcpyDeviceToHost, num->stream));
  HIP_TRY(hipStreamSynchronize(num->stream));
  return AMP_OK;
}

int amp_download_factor(amp_num* num, double* Lx, double* D) {
  if (!num || !Lx) return fail(AMP_ERR_ARG, "amp_download_factor: bad arguments");
  const amp::Symbolic& S = num->sym->S;
  if (num->prec == AMP_F64) {
    HIP_TRY(hipMemcpy(Lx, num->Lx.p, size_t(S.nnzLStored) * 8, hipMemcpyDeviceToHost));
    if (D && num->mode == AMP_LDLT) HIP_TRY(hipMemcpy(D, num->D.p, size_t(S.n) * 8, hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(size_t(std::max<int64_t>(S.nnzLStored, int64_t(S.n))));
    HIP_TRY(hipMemcpy(tmp.data(), num->Lx.p, size_t(S.nnzLStored) * 4, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < S.nnzLStored; i++) Lx[i] = tmp[i];
    if (D && num->mode == AMP_LDLT) {
      HIP_TRY(hipMemcpy(tmp.data(), num->D.p, size_t(S.n) * 4, hipMemcpyDeviceToHost));
      for (int32_t i = 0; i < S.n; i++) D[i] = tmp[i];
    }
  }
  return AMP_OK;
}

int amp_num_set_profiling(amp_num* num, int32_t enable) {
  if (!num) return fail(AMP_ERR_ARG, "amp_num_set_profiling: bad arguments");
  num->profiling = enable != 0;
  if (num->profiling && num->prof.events.empty()) {
    const size_t nl = deviceFactorPlan(num).launches.size();
    num->prof.events.resize(2 * nl + 2, nullptr);
    num->prof.kinds.resize(nl + 1, 0);
    num->prof.levels.resize(nl + 1, 0);
    for (hipEvent_t& e : num->prof.events) HIP_TRY(hipEventCreate(&e));
    const size_t ns = num->sym->dsp.launches.size();
    num->solveProf.events.resize(2 * ns + 2, nullptr);
    num->solveProf.kinds.resize(ns + 1, 0);
    num->solveProf.levels.resize(ns + 1, 0);
    for (hipEvent_t& e : num->solveProf.events) HIP_TRY(hipEventCreate(&e));
  }
  return AMP_OK;
}

int amp_num_get_kernel_ms(amp_num* num, const char* kernel, double* total_ms, int32_t* launches) {
  if (!num || !kernel || !total_ms || !launches) return fail(AMP_ERR_ARG, "amp_num_get_kernel_ms: bad arguments");
  auto it = num->kernelMs.find(kernel);
  *total_ms = it == num->kernelMs.end() ? 0.0 : it->second.first;
  *launches = it == num->kernelMs.end() ? 0 : it->second.second;
  return AMP_OK;
}

int amp_num_set_growth_limit(amp_num* num, double limit) {
  if (!num || !(limit >= 0)) return fail(AMP_ERR_ARG, "amp_num_set_growth_limit: bad arguments");
  num->growthLimit = num->prec == AMP_F64 ? 0.0 : limit;
  return AMP_OK;
}

int amp_num_release_workspace(amp_num* num) {
  if (!num) return fail(AMP_ERR_ARG, "amp_num_release_workspace: bad arguments");
  if (num->liveGraphs > 0)
    return fail(AMP_ERR_STATE, "amp_num_release_workspace: captured graphs of this factor still use the workspace");
  HIP_TRY(hipStreamSynchronize(num->stream));
  num->U.reset();
  return AMP_OK;
}

void amp_num_free(amp_num* num) { delete num; }

}  // extern "C"

struct amp_graph {
  amp_num* num = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  DevBuf r, d;  // residual and correction, n x nrhs FP64
  hipEvent_t e0 = nullptr, e1 = nullptr;
  bool counted = false;  // counted in num->liveGraphs
  ~amp_graph() {
    if (counted) num->liveGraphs--;
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

namespace {

// Enqueue (no host synchronisation) the factor, the solve and irSteps refinement steps on the handle's stream.
template <typename T>
int enqueueSequence(amp_graph* g, const double* d_values, const double* d_b, double* d_x, int32_t q, int32_t irSteps) {
  amp_num* num = g->num;
  amp_sym* h = num->sym;
  const int32_t n = h->S.n;
  hipStream_t st = num->stream;
  amp::launchFactorPlan<T>(deviceFactorPlan(num), factorArgs<T>(num), d_values, st, num->side, num->syncEv, nullptr);
  const amp::SolveArgs<T> sa = solveArgs<T>(num);
  auto solve = [&](const double* b, double* x) {
    return amp::launchSolvePlan<T, T>(h->dsp, sa, b, x, q, num->y.as<T>(), num->fv.as<T>(), num->pz.as<T>(), num->fc,
                                      st, nullptr);
  };
  HIP_TRY(solve(d_b, d_x));
  double* r = g->r.as<double>();
  double* d = g->d.as<double>();
  for (int32_t it = 0; it < irSteps; it++) {
    amp::launchResidualMulti(h->dRowptr.as<int32_t>(), h->dColidx.as<int32_t>(), d_values, d_x, d_b, r, n, q, st);
    HIP_TRY(solve(r, d));
    amp::launchAxpby(1.0, d, 1.0, d_x, n * q, st);
  }
  HIP_TRY(hipGetLastError());
  return AMP_OK;
}

}  // namespace

extern "C" {

int amp_graph_create(amp_num* num, const double* d_values, const double* d_b, double* d_x, int32_t nrhs,
                     int32_t irSteps, amp_graph** out) {
  if (!num || !out || nrhs < 1 || irSteps < 0 || !d_values || !d_b || !d_x)
    return fail(AMP_ERR_ARG, "amp_graph_create: bad arguments");
  *out = nullptr;
  if (num->mode != AMP_CHOLESKY) return fail(AMP_ERR_ARG, "amp_graph_create: Cholesky handles only");
  amp_sym* h = num->sym;
  const int32_t n = h->S.n;
  if (n == 0) return fail(AMP_ERR_ARG, "amp_graph_create: empty system");
  if (int64_t(n) * nrhs > INT32_MAX) return fail(AMP_ERR_ARG, "amp_graph_create: n x nrhs too large");
  if (irSteps > 0) {
    if (h->rowptr.empty()) return fail(AMP_ERR_STATE, "amp_graph_create: the CSR pattern is unknown (amp_sym_set_pattern)");
    if (!h->dRowptr.p) {
      HIP_TRY(upload(h->dRowptr, h->rowptr));
      HIP_TRY(upload(h->dColidx, h->colidx));
    }
  }
  int rc = ensureVecs(num, nrhs);
  if (rc) return rc;
  std::unique_ptr<amp_graph> g(new amp_graph());
  g->num = num;
  HIP_TRY(g->r.alloc(size_t(n) * nrhs * 8));
  HIP_TRY(g->d.alloc(size_t(n) * nrhs * 8));
  HIP_TRY(hipEventCreate(&g->e0));
  HIP_TRY(hipEventCreate(&g->e1));
  HIP_TRY(amp::afterCaller(num->stream, num->inEv));
  // one uncaptured pass first: kernel occupancy queries and first-use runtime work stay out of the capture
  rc = num->prec == AMP_F64 ? enqueueSequence<double>(g.get(), d_values, d_b, d_x, nrhs, 0)
                            : enqueueSequence<float>(g.get(), d_values, d_b, d_x, nrhs, 0);
  if (rc) return rc;
  HIP_TRY(hipStreamSynchronize(num->stream));
  HIP_TRY(hipStreamBeginCapture(num->stream, hipStreamCaptureModeRelaxed));
  rc = num->prec == AMP_F64 ? enqueueSequence<double>(g.get(), d_values, d_b, d_x, nrhs, irSteps)
                            : enqueueSequence<float>(g.get(), d_values, d_b, d_x, nrhs, irSteps);
  const hipError_t ce = hipStreamEndCapture(num->stream, &g->graph);
  if (rc) return rc;
  if (ce != hipSuccess) return fail(AMP_ERR_HIP, std::string("amp_graph_create: capture: ") + hipGetErrorString(ce));
  HIP_TRY(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
  num->liveGraphs++;
  g->counted = true;
  *out = g.release();
  return AMP_OK;
}

int amp_graph_launch(amp_graph* g) {
  if (!g) return fail(AMP_ERR_ARG, "amp_graph_launch: bad arguments");
  HIP_TRY(amp::afterCaller(g->num->stream, g->num->inEv));
  amp::launchClearFlag(g->num->fc.err, g->num->stream);  // the wait-failure flag reports this replay only
  amp::launchClearFlag(g->num->counts.as<int32_t>() + 3, g->num->stream);  // and the panel-step one
  HIP_TRY(hipEventRecord(g->e0, g->num->stream));
  HIP_TRY(hipGraphLaunch(g->exec, g->num->stream));
  HIP_TRY(hipEventRecord(g->e1, g->num->stream));
  return AMP_OK;
}

int amp_graph_sync(amp_graph* g, amp_factor_info* info) {
  if (!g) return fail(AMP_ERR_ARG, "amp_graph_sync: bad arguments");
  amp_num* num = g->num;
  int32_t failCol = INT32_MAX, waitErr = 0, panelErr = 0;
  HIP_TRY(hipMemcpyAsync(&failCol, num->failCol.p, 4, hipMemcpyDeviceToHost, num->stream));
  HIP_TRY(hipMemcpyAsync(&waitErr, num->fc.err, 4, hipMemcpyDeviceToHost, num->stream));
  HIP_TRY(hipMemcpyAsync(&panelErr, num->counts.as<int32_t>() + 3, 4, hipMemcpyDeviceToHost, num->stream));
  HIP_TRY(hipStreamSynchronize(num->stream));
  if (waitErr) return fail(AMP_ERR_INTERNAL, "graph replay: a block dependency wait timed out");
  if (panelErr) return fail(AMP_ERR_INTERNAL, "graph replay: a panel-step wait for the launch's workgroups timed out");
  float ms = 0;
  HIP_TRY(hipEventElapsedTime(&ms, g->e0, g->e1));
  num->ok = failCol == INT32_MAX;
  if (info) {
    info->ok = num->ok ? 1 : 0;
    info->failColumn = num->ok ? -1 : failCol;
    info->nNeg = 0;
    info->nPerturbed = 0;
    info->ms = ms;
  }
  return AMP_OK;
}

void amp_graph_free(amp_graph* g) { delete g; }

}  // extern "C"

namespace {

int rebuildPlans(amp_sym* h, const amp::PlanParams& p) {
  if (int rc = buildPlans(h, p)) return rc;
  h->params = p;
  h->onDevice = false;
  return symToDevice(h);
}

bool validParams(const amp::PlanParams& p) {
  return p.smallMax >= 1 && p.smallMax <= amp::kSmallMax && p.smallNarrow >= 1 && p.smallNarrow <= p.smallMax &&
         p.wideLaunch >= 1 && p.superPanel >= amp::kTileEdge && p.superPanel % amp::kTileEdge == 0 &&
         p.flowWideMin >= 0 && p.packMin >= 0 && (p.fusedPanel == 0 || p.fusedPanel == 1) &&
         p.fusedSolve >= 0 && p.fusedSolve <= 3;
}

// factor + one solve (zero right-hand side), device milliseconds of the second of two runs
int timePlan(amp_sym* h, const double* d_values, int32_t mode, int32_t precision, double* ms) {
  amp_num* raw = nullptr;
  int rc = amp_num_create(h, mode, precision, &raw);
  if (rc) return rc;
  std::unique_ptr<amp_num> num(raw);
  HIP_TRY(hipMemset(num->b.p, 0, size_t(std::max(h->S.n, 1)) * 8));
  for (int rep = 0; rep < 2; rep++) {
    amp_factor_info fi{};
    rc = num->prec == AMP_F64 ? runFactor<double>(num.get(), d_values, 1e-13, &fi)
                              : runFactor<float>(num.get(), d_values, 1e-13, &fi);
    if (rc) return rc;
    if (!fi.ok) return fail(AMP_ERR_ARG, "amp_sym_autotune: the factorization fails on these values");
    rc = runSolve(num.get(), num->b.as<double>(), num->x.as<double>(), 1);
    if (rc) return rc;
    *ms = fi.ms + num->solveMs;
  }
  return AMP_OK;
}

}  // namespace

extern "C" {

int amp_sym_set_params(amp_sym* h, const amp_tune_info* t) {
  if (!h || !t) return fail(AMP_ERR_ARG, "amp_sym_set_params: bad arguments");
  if (h->liveNums) return fail(AMP_ERR_STATE, "amp_sym_set_params: numeric handles of this analysis exist");
  amp::PlanParams p;
  p.smallMax = t->smallMax;
  p.smallNarrow = t->smallNarrow;
  p.wideLaunch = t->wideLaunch;
  p.superPanel = t->superPanel;
  p.flowWideMin = t->flowWideMin;
  p.packMin = t->packMin;
  p.fusedPanel = t->fusedPanel;
  p.fusedSolve = t->fusedSolve;
  if (!validParams(p)) return fail(AMP_ERR_ARG, "amp_sym_set_params: thresholds out of range");
  return rebuildPlans(h, p);
}

int amp_sym_autotune(amp_sym* h, const double* d_values, int32_t mode, int32_t precision, amp_tune_info* info) {
  if (!h || !d_values || (mode != AMP_CHOLESKY && mode != AMP_LDLT) ||
      (precision != AMP_F64 && precision != AMP_F32 && precision != AMP_F32_BF16X3))
    return fail(AMP_ERR_ARG, "amp_sym_autotune: bad arguments");
  if (h->liveNums) return fail(AMP_ERR_STATE, "amp_sym_autotune: numeric handles of this analysis exist");
  if (h->S.n == 0) return fail(AMP_ERR_ARG, "amp_sym_autotune: empty system");
  HIP_TRY(hipDeviceSynchronize());  // d_values complete, whichever stream produced it
  amp::PlanParams best;  // the defaults
  int rc = rebuildPlans(h, best);
  if (rc) return rc;
  double bestMs = 0;
  rc = timePlan(h, d_values, mode, precision, &bestMs);
  if (rc) return rc;
  const double defaultMs = bestMs;
  int32_t tried = 1;
  // one threshold at a time, from the defaults
  const int32_t superPanels[] = {128, 512}, wides[] = {256, 1024, 1 << 30}, smalls[] = {64}, narrows[] = {16, 64};
  const int32_t flowWides[] = {64, 1 << 30};  // forward solve launches: wide tasks from 64 tasks on, or never
  const int32_t packs[] = {128, 1 << 30};     // whole fronts of at most 8 / 16 rows packed from 128 on, or never
  auto attempt = [&](amp::PlanParams p) -> int {
    if (!validParams(p)) return AMP_OK;
    int r = rebuildPlans(h, p);
    if (r) return r;
    double ms = 0;
    r = timePlan(h, d_values, mode, precision, &ms);
    if (r) return r;
    tried++;
    if (ms < bestMs) {
      bestMs = ms;
      best = p;
    }
    return AMP_OK;
  };
  for (int32_t v : superPanels) {
    amp::PlanParams p = best;
    p.superPanel = v;
    if ((rc = attempt(p))) return rc;
  }
  for (int32_t v : wides) {
    amp::PlanParams p = best;
    p.wideLaunch = v;
    if ((rc = attempt(p))) return rc;
  }
  for (int32_t v : smalls) {
    amp::PlanParams p = best;
    p.smallMax = v;
    p.smallNarrow = std::min(p.smallNarrow, v);
    if ((rc = attempt(p))) return rc;
  }
  for (int32_t v : narrows) {
    amp::PlanParams p = best;
    p.smallNarrow = std::min(v, p.smallMax);
    if ((rc = attempt(p))) return rc;
  }
  for (int32_t v : flowWides) {
    amp::PlanParams p = best;
    p.flowWideMin = v;
    if ((rc = attempt(p))) return rc;
  }
  for (int32_t v : packs) {
    amp::PlanParams p = best;
    p.packMin = v;
    if ((rc = attempt(p))) return rc;
  }
  {  // the panel chain as three launches per step instead of one
    amp::PlanParams p = best;
    p.fusedPanel = 0;
    if ((rc = attempt(p))) return rc;
  }
  {  // the solve's runs of tiled levels as two launches per level and direction instead of one per run
    amp::PlanParams p = best;
    for (int32_t v = 0; v < 4; v++) {
      p.fusedSolve = v;
      if (v != best.fusedSolve && (rc = attempt(p))) return rc;
    }
  }
  rc = rebuildPlans(h, best);
  if (rc) return rc;
  if (info) {
    info->smallMax = best.smallMax;
    info->smallNarrow = best.smallNarrow;
    info->wideLaunch = best.wideLaunch;
    info->superPanel = best.superPanel;
    info->flowWideMin = best.flowWideMin;
    info->packMin = best.packMin;
    info->fusedPanel = best.fusedPanel;
    info->fusedSolve = best.fusedSolve;
    info->candidates = tried;
    info->defaultMs = defaultMs;
    info->bestMs = bestMs;
  }
  return AMP_OK;
}

}  // extern "C"
