// Launch interface of the factor and solve kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "kernels_common.h"
#include "plan.h"

namespace amp {

// Device copies of a FactorPlan's task arrays (launch records stay on the host).
struct DeviceFactorPlan {
  const ZeroTask* zeros = nullptr;
  const int64_t* adst = nullptr;
  const int32_t* smallFronts = nullptr;
  const EaTask* ea = nullptr;
  const EaRange* eaRanges = nullptr;
  const DiagTask* diag = nullptr;
  const PanelTask* panels = nullptr;
  int32_t panelCounters = 0;
  const void* gemm = nullptr;
  const SplitTask* splits = nullptr;
  const SplitRef* srefs = nullptr;
  std::vector<Launch> launches;
};

// Optional per-launch timing: events[2 i], events[2 i + 1] bracket launch i; kinds[i] is its launch kind (or
// kProfGemmBf16x3 for a bulk launch run as bf16 products) and levels[i] the elimination-tree level it serves.
constexpr int32_t kProfGemmBf16x3 = 8;
struct ProfileSink {
  std::vector<hipEvent_t> events;
  std::vector<int32_t> kinds, levels;
  size_t used = 0;
};

// Enqueue the factor plan on `main`, with the launches marked kSyncSide on `side`; `sync` holds
// FactorPlan::signals + 2 events (the signals, the final join, and the initial fork: the side stream first waits for
// everything enqueued on `main` before the call). The side stream is joined into `main` before returning.
template <typename T>
void launchFactorPlan(const DeviceFactorPlan& dp, const FactorArgs<T>& a, const double* values, hipStream_t main,
                      hipStream_t side, const std::vector<hipEvent_t>& sync, ProfileSink* prof);
template <typename T>
void launchLdltCounts(const FactorArgs<T>& a, int32_t n, int32_t* out, hipStream_t st);

// ---- solves ----
struct DeviceSolvePlan {
  const int32_t* fronts = nullptr;
  const BlockTask* blocks = nullptr;
  std::vector<SolveLaunch> launches;
  int64_t fvEntries = 0;
};

template <typename T>
struct SolveArgs {
  const T* Lx;
  const T* W;
  const T* D;  // LDL^T pivots, nullptr for Cholesky
  const int32_t* superStart;
  const int32_t* Sp;
  const int32_t* Si;
  const int64_t* Lxp;
  const int64_t* Woff;
  const int32_t* childHead;
  const int32_t* childNext;
  const int32_t* rel;
  const int32_t* perm;
  int32_t n;
  int32_t ldlt;
};

// Synchronisation state of the tiled-supernode solve launches, owned by one numeric handle (launches of one handle
// run in stream order). Flags hold the pass number of the last pass that completed the block, so they never need
// clearing; the pass number lives in device memory (*epoch) and is raised by a kernel at the start of every pass
// of at most kFlowRhs right-hand sides, so a captured graph raises it on every replay. err is set when a wait gives
// up.
// The solved values of a block also travel in mailboxes (fwdBox / bwdBox): per flag, kFlowRhs x 64 values, each in
// 64-bit words that carry the pass number in their high half (one word per FP32 value, two per FP64 value: low and
// high half of the bits). A consumer accepts a word only when it carries the current pass number, so the values need
// no fence to be read safely; the flags only tell a consumer how many blocks are worth reading.
struct FlowCtl {
  int32_t* fwdFlag;
  int32_t* bwdFlag;
  int32_t* err;
  int32_t* epoch;
  uint64_t* fwdBox;
  uint64_t* bwdBox;
  int32_t* ctr;  // task counters of the flow launches of a pass (cleared by the kernel that raises the pass number)
  int32_t nctr;
  int32_t* done;  // counters of the fused launches (SolvePlan::doneSlots, stored after the task counters, cleared
                  // with them)
  int32_t ndone;
};
constexpr int kBoxWords = 2 * kFlowRhs * 64;  // mailbox words per flag (FP64 values)

// Bulk trailing-update launch of an FP32 factor on the bf16 matrix cores (three bf16 products per FP32 product; modes 0
// and 1 of GemmTask, tiles of kGemmEdge).
void launchGemmBf16x3(const GemmTask* tasks, const SplitRef* refs, unsigned grid, const FactorArgs<float>& a,
                      hipStream_t st);
// Split launch in front of such a bulk launch: the panels of `splits` into the split arrays of `a`.
void launchSplitPanels(const SplitTask* splits, int64_t count, int32_t maxRows, int32_t maxPitch,
                       const FactorArgs<float>& a, hipStream_t st);

// Dynamic LDS for a side-stream bulk launch of `kernel` (`threads` per workgroup, `staticLds` bytes of static LDS):
// padding that lets one workgroup fewer fit on a CU than without it (0 when the kernel fits at most two), so that the
// panel chain's kernels find room beside the bulk updates on every CU. Cached per kernel and device.
size_t bulkLds(const void* kernel, int threads, size_t staticLds);

// Fused panel-step launch (kPanel) and the kernel that clears the launches' counters (n ints) before a factor.
template <typename T>
void launchPanelSteps(const PanelTask* tasks, unsigned count, const FactorArgs<T>& a, hipStream_t st);
void launchClearInts(int32_t* p, int32_t n, hipStream_t st);
// Workgroups of the fused panel-step kernels resident at once on the current device (the fewest over their instances;
// INT32_MAX when the device cannot be queried).
int32_t panelResidentTasks();

// Clear one int32 device flag with a kernel (a kernel node, not a memset node, when the caller captures the stream).
void launchClearFlag(int32_t* flag, hipStream_t st);

// Solve q systems in place on the device: b (q vectors of n, original order) -> x. y (n q), fv (fvEntries q) and pz
// (initSlots x 64 x min(q, kFlowRhs)) are scratch of the working precision S; Lx / W / D are the factor's precision
// T. Raises *fc.epoch once per pass. Enqueues kernels only (no host synchronisation), so it can be captured.
template <typename T, typename S>
hipError_t launchSolvePlan(const DeviceSolvePlan& dp, const SolveArgs<T>& a, const double* b, double* x, int32_t q,
                           S* y, S* fv, S* pz, const FlowCtl& fc, hipStream_t st, ProfileSink* prof);

}  // namespace amp
