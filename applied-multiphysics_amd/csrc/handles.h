// Internal definitions of the C ABI handles (include/ampsolve.h) shared by the library's host translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/ampsolve.h"
#include "factor_kernels.h"
#include "plan.h"
#include "symbolic.h"

namespace amp {

int fail(int code, const std::string& msg);

#define HIP_TRY(call)                                                                                    \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess) return ::amp::fail(AMP_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

// The *_device entry points read device inputs that the caller may have produced with work still queued on the null
// (legacy default) stream, the default stream of most frameworks: before its first launch, the handle's stream waits
// on an event recorded there. Work on other streams of the caller must be complete before the call.
inline hipError_t afterCaller(hipStream_t st, hipEvent_t ev) {
  hipError_t e = hipEventRecord(ev, nullptr);
  if (e == hipSuccess) e = hipStreamWaitEvent(st, ev, 0);
  return e;
}

// Owning device allocation.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void reset() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  hipError_t alloc(size_t b) {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = b;
    return hipMalloc(&p, std::max<size_t>(b, 256));
  }
  template <typename T>
  T* as() const {
    return static_cast<T*>(p);
  }
};

template <typename T>
hipError_t upload(DevBuf& d, const std::vector<T>& v) {
  hipError_t e = d.alloc(v.size() * sizeof(T));
  if (e != hipSuccess || v.empty()) return e;
  return hipMemcpy(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
}

int solveOnDevice(amp_num* num, const double* d_b, double* d_x, int32_t nrhs);
// The solve's launches on the handle's stream, with no synchronisation: the caller clears the wait-failure flag
// (launchClearFlag) before its first such solve and checks it (solveWaitFailed) after its last, for drivers that solve
// once per iteration and synchronise for their own scalars anyway.
int solveOnDeviceNoSync(amp_num* num, const double* d_b, double* d_x, int32_t nrhs);
int solveWaitFailed(amp_num* num);  // reads the flag (synchronises the handle's stream): AMP_OK or AMP_ERR_INTERNAL

// Bytes of the solve scratch (working-precision y, front vectors fv, backward partial products pz, each rounded up to
// 256 bytes) that nrhs right-hand sides need, and the solve with scratch supplied by the caller (no allocation).
struct SolveScratch {
  size_t yBytes = 0, fvBytes = 0, pzBytes = 0;
  size_t total() const { return yBytes + fvBytes + pzBytes; }
};
SolveScratch solveScratchNeed(const amp_num* num, int32_t nrhs);
int solveWithScratch(amp_num* num, const double* d_b, double* d_x, int32_t nrhs, void* y, void* fv, void* pz);

}  // namespace amp

struct amp_sym {
  amp::Symbolic S;
  amp::FactorPlan F;
  // F with each fused panel step as three launches, built when F fuses them: FP64 factors run this form (the fused
  // kernel's FP64 instance cannot share a CU with the FP64 bulk updates); empty otherwise
  amp::FactorPlan F3;
  amp::SolvePlan SP;
  std::vector<int32_t> diagSrc;  // CSR slot of each diagonal entry (new column order), -1 when absent
  std::vector<int32_t> rowptr, colidx;  // the analysed CSR pattern (for the refinement residual), empty if unknown
  amp::DevBuf dRowptr, dColidx;
  bool onDevice = false;
  // device copies of the structure and of the task arrays
  amp::DevBuf superStart, Sp, Si, Lxp, Uoff, Woff, childHead, childNext, rel, lowerP, lowerSrc, amap, adst, perm, dDiagSrc;
  amp::DevBuf zeros, smallFronts, ea, eaRanges, diag, panels, gemm, splits, srefs, sFronts, sBlocks;
  amp::DeviceFactorPlan dfp;
  amp::DevBuf diag3, gemm3, splits3, srefs3;  // task arrays of F3 (the others are F's)
  amp::DeviceFactorPlan dfp3;
  amp::DeviceSolvePlan dsp;
  amp::PlanParams params;  // the thresholds the plans were built with
  int32_t liveNums = 0;    // numeric handles on this analysis (their storage is laid out by the plans)
};

struct amp_num {
  amp_sym* sym = nullptr;  // counted in sym->liveNums while the handle lives
  int32_t mode = AMP_CHOLESKY, prec = AMP_F64;
  hipStream_t stream = nullptr, side = nullptr;  // side: bulk trailing updates overlapping the panel chain
  std::vector<hipEvent_t> syncEv;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, evR0 = nullptr, evR1 = nullptr;
  hipEvent_t inEv = nullptr;  // afterCaller
  amp::DevBuf Lx, U, W, D, pert, failCol, counts, tau, values, b, x, y, fv, pz;
  amp::DevBuf panelCtr;  // counters of the kPanel launches
  amp::DevBuf split;  // bf16 bulk products (AMP_F32_BF16X3): the split panel of the current bulk launch
  // refinement work space (FP64): r, d, w, Krylov bases V (restart + 1 vectors) and Z (restart), reductions
  amp::DevBuf r, dvec, w, V, Z, partial, small;
  amp::DevBuf opw;  // shift-invert operator: the products M X (n x nvec, FP64)
  // tiled-supernode solve synchronisation: forward and backward block flags, the wait-failure flag
  amp::DevBuf flow, boxes;
  amp::FlowCtl fc{};
  bool ok = false;
  bool profiling = false;
  int32_t liveGraphs = 0;  // captured graphs holding this handle's buffer addresses (amp_graph_create / _free)
  double growthLimit = 0;  // FP32 factors: FP64 redo of diagonal blocks with a larger elimination multiplier
  amp::ProfileSink prof, solveProf;
  float solveMs = 0;
  std::map<std::string, std::pair<double, int32_t>> kernelMs;
  ~amp_num() {
    if (sym) sym->liveNums--;
    for (hipEvent_t e : prof.events) (void)hipEventDestroy(e);
    for (hipEvent_t e : solveProf.events) (void)hipEventDestroy(e);
    if (ev0) (void)hipEventDestroy(ev0);
    if (evR0) (void)hipEventDestroy(evR0);
    if (evR1) (void)hipEventDestroy(evR1);
    if (ev1) (void)hipEventDestroy(ev1);
    if (inEv) (void)hipEventDestroy(inEv);
    for (hipEvent_t e : syncEv) (void)hipEventDestroy(e);
    if (side) (void)hipStreamDestroy(side);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

// The factor plan a numeric handle runs (host and device forms): F3 for FP64 factors when the analysis has one.
inline const amp::FactorPlan& factorPlan(const amp_num* num) {
  return num->prec == AMP_F64 && !num->sym->F3.launches.empty() ? num->sym->F3 : num->sym->F;
}
inline const amp::DeviceFactorPlan& deviceFactorPlan(const amp_num* num) {
  return num->prec == AMP_F64 && !num->sym->F3.launches.empty() ? num->sym->dfp3 : num->sym->dfp;
}

