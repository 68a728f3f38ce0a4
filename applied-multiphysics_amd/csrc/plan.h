// Execution plan of the device multifrontal factorization and of the supernodal solves.
//
// Built once per symbolic structure on the host. It fixes the HBM layout (L blocks at the reference's Lxp offsets,
// one update matrix per supernode, the inverted diagonal blocks of tiled supernodes) and a flat sequence of launches,
// each a list of independent tasks. Levels of the supernodal elimination tree (level = height above the leaves) run
// in order, so every child is complete before its parent starts. Task offsets are relative to buffer bases, so the
// same plan is replayed by any executor.
#pragma once
#include <cstdint>
#include <vector>

#include "symbolic.h"

namespace amp {

constexpr int kTileEdge = 64;   // panel width of tiled fronts (diagonal blocks); block size of their solves
constexpr int kGemmEdge = 128;  // output-tile edge of update products in launches with enough tiles
constexpr int kWideLaunch = 512;  // tiles of edge kGemmEdge a launch needs to use them (else kTileEdge)
constexpr int kEaWidth = 16;    // parent columns per extend-add work item into an update matrix
constexpr int kEaWidthF = 4;    // parent columns per extend-add work item into the factor columns (on the chain)
constexpr int kSuperPanel = 256;  // columns brought up to date together (left-looking) in a tiled front
constexpr int kSmallMax = 96;   // fronts with m <= kSmallMax are factored (and solved) whole by one workgroup
constexpr int kWaveFrontMax = 32;  // fronts with m <= this are factored by one wave each (four per workgroup)
constexpr int kZeroChunk = 1 << 18;

enum Buf : int8_t { kBufLx = 0, kBufU = 1, kBufW = 2 };

// One output tile (rows x cols <= 64 x 64) of
//   mode 0: C = C - A diag(d) B^T      mode 1: C = -A diag(d) B^T      mode 2: C = A B^T diag(d)^-1
// column-major operands; d indexes the depth (modes 0, 1) or the output column (mode 2). Only LDL^T uses d.
struct GemmTask {
  int64_t a_off, b_off, c_off;
  int32_t lda, ldb, ldc;
  int32_t rows, cols, k;
  int32_t d_off;
  int8_t a_buf, b_buf, c_buf, mode;
};

// Split of a panel for the bf16 bulk products of FP32 factors (k_gemm_bf16x3): rows [r0, m) of columns
// [c0, c0 + w) of a tiled front (FP32 at Lx[src], src = Lxp[s] + r0 + c0 m, leading dimension m) into bf16 heads and
// tails (x = head + tail + O(2^-16 x)), in chunks of kSplitDepth depths: depth d of row r0 + i at
// dst + (d / kSplitDepth) rows kSplitDepth + i kSplitDepth + d % kSplitDepth, so that a chunk of consecutive rows is
// contiguous; depths [w, dp) are zero (dp: w rounded up to the chunk). LDL^T also splits the pivot-scaled copy (the B
// operand, pivots at D[d_off ..]).
struct SplitTask {
  int64_t src, dst;
  int32_t m, rows, w, dp, d_off, pad;
};

// Split-buffer operands of a bulk product task (aligned with FactorPlan::gemm; unused for the tasks of other
// launches): element offsets of the first rows of the A and B tiles in their panel's first chunk, and the chunk pitch
// of that panel (rows kSplitDepth).
struct SplitRef {
  int64_t a, b;
  int32_t ld, pad;
};
constexpr int kSplitDepth = 32;  // depth chunk of k_gemm_bf16x3: split rows are padded to a multiple of it

// Factor the nb x nb diagonal block at Lx[f_off] (leading dimension ld) in place and write the inverse of its lower
// factor (unit lower for LDL^T) to W[w_off] (leading dimension kTileEdge). col0: global index of its first column.
struct DiagTask {
  int64_t f_off, w_off;
  int32_t ld, nb, col0, pad;
};

// Add every child's update matrix into columns [j0, j1) (front coordinates) of front s. The children's columns that
// land there are listed in EaRange entries [r0, r1), children in the reference's order (childHead chain).
struct EaTask {
  int32_t s, j0, j1, r0, r1, pad;
};

// Columns [b0, b1) of child c's update matrix (they map to increasing parent columns through the relative map).
struct EaRange {
  int32_t c, b0, b1, pad;
};

struct ZeroTask {
  int64_t off, len;  // Lx entries
};

// One workgroup of a fused panel step (kPanel launch): the 64-column step whose diagonal block is at Lx[f_off]
// (leading dimension ld, nb columns, first global column col0) of a tiled front, for one row tile of its panel. Every
// task factors the diagonal block itself; the owner also stores its factor, the inverse of its lower factor (W[w_off])
// and the pivot records. Own rows [ro, ro + rows) below the block (relative to its first row; rows == 0: none) become
// L = A L11^-T (LDL^T: times D^-1); then the same rows of the ucols columns that follow the block inside its
// super-panel are updated, C -= L D Lc^T, with Lc (the panel rows [nb, nb + ucols)) formed by the task from A. Tasks
// that overwrite inputs of the front's step (the owner, and the row tiles that hold Lc) wait until all ntask tasks of
// the front's step have read them: counter ctr (one per front and launch) counts the tasks that have. A front's tasks
// are consecutive in the launch, so with workgroups dispatched in order the writers of every front whose tasks are
// all resident complete, and no wait can hold the launch (ntask must not exceed the resident workgroups).
// flags: kPanelPre: the block was factored by the launch before (the next block of its super-panel's previous step),
// so its tasks load W and D instead, and none is the owner; kPanelNext (the first row tile, when the super-panel has
// a next step): after the update of its first column tile, which is the next step's diagonal block, the task
// factors that block and stores it as that step's owner would.
struct PanelTask {
  int64_t f_off, w_off;
  int32_t ld, nb, col0, ro, rows, ucols, ctr, ntask, owner, flags;
};
constexpr int32_t kPanelPre = 1, kPanelNext = 2;
constexpr int kPanelUpdateTiles = kSuperPanel / kTileEdge - 1;  // column tiles a panel task updates, at most

enum LaunchKind : int32_t {
  kZero = 0,
  kAssemble = 1,
  kSmall = 2,
  kExtendAdd = 3,
  kDiag = 4,
  kGemm = 5,
  kSplit = 6,
  kPanel = 7,
};

struct Launch {
  LaunchKind kind;
  int32_t level;
  int64_t first, count;  // range in the task array of this kind (kSmall: in smallFronts; kAssemble: lower entries)
  int32_t maxM;          // kSmall: largest front of the launch (LDS sizing); kSplit: most rows of a task
  int32_t threads;       // kSmall: workgroup size, or 0 for one wave per front (maxM <= kWaveFrontMax)
  int32_t edge;          // kGemm: output-tile edge of the launch's tasks (kTileEdge or kGemmEdge); kSplit: largest
                         // row pitch of the launch's tasks; kSmall with threads 0: lanes per front when a wave takes
                         // several (8 or 16, at least maxM), else 0
  int32_t sync;          // kSyncSide | kSyncWait | kSyncSignal
};

// Two streams: the main stream and a side stream for the bulk trailing updates of large fronts, which overlap the
// next super-panel's factorization on the main stream. kSyncSide: run on the side stream; kSyncWait: first wait for
// the other stream's most recent signal (enqueue order); kSyncSignal: signal after the launch.
constexpr int32_t kSyncSide = 1, kSyncWait = 2, kSyncSignal = 4;

struct FactorPlan {
  int32_t n = 0, nsuper = 0, nlevels = 0;
  int64_t nnzLStored = 0, uEntries = 0, wEntries = 0, nnzA = 0, nnzPattern = 0;
  double flops = 0;
  double gemmFlops = 0;  // algorithmic flops of the gemm tasks: lower-triangle outputs, triangular solves as such
  // algorithmic HBM bytes of the wide bulk product launches (those with a panel split in front): every lower entry of
  // their outputs read and written (mode 0) or written (mode 1), 4 bytes each way, and every entry of the split
  // panels read once (4 bytes: an FP32 value, or its bf16 head and tail)
  double bulkBytes = 0;
  int32_t bulkLaunches = 0;
  // per supernode
  std::vector<int32_t> level, isBig;
  std::vector<int64_t> Uoff;  // offset of the mu x mu update matrix (leading dimension mu), -1 when mu == 0
  std::vector<int64_t> Woff;  // tiled supernodes: offset of their ceil(nc/64) inverted 64 x 64 diagonal blocks
  std::vector<int32_t> rel;   // aligned with Si: position of each below row of a child inside its parent's front
  // per lower entry of A (aligned with lowerP / lowerI / lowerSrc)
  std::vector<int32_t> amap;  // offset inside its supernode's m x nc block (row + c m)
  std::vector<int64_t> adst;  // tiled supernodes: absolute Lx index; -1 for supernodes factored in LDS
  // task arrays
  std::vector<ZeroTask> zeros;
  std::vector<int32_t> smallFronts;
  std::vector<EaTask> ea;
  std::vector<EaRange> eaRanges;
  std::vector<DiagTask> diag;
  std::vector<PanelTask> panels;  // kPanel launches (fused panel steps)
  int32_t panelCounters = 0;      // counters of the kPanel launches (one per front and launch, PanelTask::ctr)
  int32_t panelMaxTasks = 0;      // most tasks of one front in a kPanel launch
  std::vector<GemmTask> gemm;
  std::vector<SplitTask> splits;  // kSplit launches: the panels of bulk launches with kGemmEdge tiles
  std::vector<SplitRef> srefs;    // aligned with gemm
  int64_t splitEntries = 0;       // largest split buffer a kSplit launch writes (entries of one array)
  std::vector<Launch> launches;
  int32_t signals = 0;  // launches with kSyncSignal
};

constexpr int kFlowWideMinTasks = 160;  // default of PlanParams::flowWideMin (see the solve plan below)
// Whole fronts of at most 8 (or 16) rows of a level are packed 8 (or 4) to a wave, in launches of their own, when the
// level has at least PlanParams::packMin of them (levels of many tiny fronts, where one wave per front is bound by
// the latency of its handful of loads); fewer go with the fronts of the next size class.
constexpr int kPackMinFronts = 512;

// Thresholds that pick kernels by supernode shape; the defaults are the constants above, and a device autotuner may
// choose others (each at most its constant, which sizes the kernels' static storage).
struct PlanParams {
  int32_t smallMax = kSmallMax;       // height m at or below which a front is factored and solved whole in LDS
  int32_t smallNarrow = 32;           // height at or below which a whole-front workgroup has 64 threads (else 256);
                                      // fronts of at most kWaveFrontMax rows take one wave each regardless
  int32_t wideLaunch = kWideLaunch;   // tiles a product launch needs to use kGemmEdge tiles
  int32_t superPanel = kSuperPanel;   // columns of a super-panel (a multiple of kTileEdge)
  int32_t flowWideMin = kFlowWideMinTasks;  // wide tasks a forward solve launch needs to use them (else one block per
                                            // task)
  int32_t packMin = kPackMinFronts;         // fronts of a size class a level needs to pack them several to a wave
  int32_t fusedPanel = 1;  // a panel step of tiled fronts is one kPanel launch (else diagonal blocks, panel products
                           // and the super-panel's update as three launches); super-panels of at most kSuperPanel
                           // columns only
  int32_t fusedSolve = 3;  // runs of consecutive levels of tiled fronts are solved by one launch per direction (bit 0:
                           // forward, bit 1: backward; else a gather or initialisation launch and a flow launch per
                           // level)
};

FactorPlan buildFactorPlan(const Symbolic& S, const PlanParams& prm = PlanParams());

// ---- solves ------------------------------------------------------------------------------------------------------
// Every supernode s owns a front vector v_s of length m (at Sp[s] in a buffer of Sp[nsuper] entries per right-hand
// side). Forward: v_s = [y_own; 0] + sum of the children's below parts (through rel); x_own = L11^-1 v_own is written
// to y; v_below -= L21 x_own and stays in v_s for the parent. Backward: z = y_own - L21^T y[below];
// x_own = L11^-T z. Supernodes with m <= kSmallMax are solved whole by one workgroup. Tiled supernodes are solved by
// one dependency-driven launch per level and direction: a task per 64-row tile (forward) or 64-column block
// (backward) waits on per-block completion flags of the blocks it consumes, so the block chain of a supernode runs
// inside one launch.
constexpr int kFlowRhs = 8;       // right-hand sides per pass of the tiled-supernode solve kernels
constexpr int kSolveWaveMax = 64;  // whole-front supernodes with m <= this are solved by one wave each
constexpr int kGatherRows = 512;
// Largest front (64-column blocks) of a level that a fused solve launch takes: a fused launch lets a level's gather or
// initialisation tasks stream beside the chains of the level above, which slows the longest chains; with fronts of
// up to 16 blocks (3D heat, 30^3 nodes) one launch per direction still wins (0.99 against 1.03 ms per pass), with
// the 59-block root of 61^3 nodes it loses (3.67 against 3.03 ms).
constexpr int kFusedSolveMaxBlocks = 16;  // front-vector rows per gather workgroup or task of the forward solve
constexpr int kInitRows = 256;    // below rows per task of the backward z = -L21^T y[below] partial products

// Blocks per task of the tiled-supernode solve launches: a task has one 256-thread quarter per 64-row (forward) or
// 64-column (backward) block. Wide tasks pass their blocks from one quarter to the next without a round trip through
// device memory, but stream a wider panel per workgroup: they are used for forward launches with many tasks (levels of
// many small fronts), where the round trips dominate; launches with few tasks (the large fronts near the root) keep one
// block per task, where the streaming per workgroup does. A wide task has four quarters in the kernels of FP32 factors
// and two in those of FP64 factors (twice the registers per thread); the plan holds a task list for each precision.
constexpr int kFlowQuadsF32 = 4, kFlowQuadsF64 = 2;

// One task of a tiled-supernode solve launch; `flag0` is the supernode's first flag.
// Forward: rows [r0, r1) of supernode s, 64 rows per quarter. A chain task (writer = w >= 1) owns the blocks
// blk .. blk + w - 1 (r0 = 64 blk, r1 = min(nc, 64 (blk + w))): it consumes the solutions of blocks 0 .. blk - 1, then
// solves its blocks in order, raising each one's flag. A below task (writer = 0, blk = nb) takes below rows and
// consumes every block.
// Backward: a chain task owns the column blocks blk .. blk + w - 1 (columns [r0, r1)): it consumes blocks
// blk + w .. nb - 1 and, for its block blk + i, the np partial products at slots p0 + i np .. p0 + i np + np - 1,
// then solves its blocks from the last one down. Backward initialisation: partial product of column block blk over
// below rows [r0, r1) into slot p0.
// Fused launches (kSolveFwdFused, kSolveBwdFused) hold the tasks of several levels, one block per chain task, and
// also the level's gather tasks (forward, writer = -1: rows [r0, r1) of the front vector of s) or initialisation tasks
// (backward, writer = 0). A task that reads what other tasks of the launch stored first waits until counter `wslot`
// of the pass reaches `wcount`, and a task whose stores others read raises counter `dslot` when it is done (-1: none):
// the gather tasks of s wait for the below tasks of its children, the chain and below tasks of s for the gather tasks
// of s; backward, the initialisation tasks of s wait for the chain tasks of its parent, the chain tasks of s for the
// initialisation tasks of s.
struct BlockTask {
  int32_t s, blk;
  int32_t r0, r1;
  int32_t writer, flag0;
  int32_t p0, np;
  int32_t wslot = -1, wcount = 0, dslot = -1;
};

enum SolveKind : int32_t {
  kSolveSmallFwd = 0,
  kSolveGather = 1,
  kSolveFwdFlow = 2,
  kSolveSmallBwd = 3,
  kSolveBwdFlow = 4,
  kSolveBwdGather = 5,  // not emitted (the backward partial products read y[below] through the row structure); the
                        // number stays reserved so that the kinds' numbers in profiles are unchanged
  kSolveBwdInit = 6,
  kSolveFwdFused = 7,  // gather and forward flow tasks of a run of levels (task lists of BlockTask, one block each)
  kSolveBwdFused = 8,  // initialisation and backward flow tasks of a run of levels
};

inline bool solveKindIsForward(SolveKind k) {
  return k == kSolveSmallFwd || k == kSolveGather || k == kSolveFwdFlow || k == kSolveFwdFused;
}
inline bool solveKindIsFlow(SolveKind k) {
  return k == kSolveFwdFlow || k == kSolveBwdFlow || k == kSolveFwdFused || k == kSolveBwdFused;
}

struct SolveLaunch {
  SolveKind kind;
  int32_t level;
  int64_t first, count;  // small / gather: range in `fronts`; flow kinds: range in `blocks` for FP32 factors; init:
                         // range in `blocks` of the initialisation tasks
  int32_t maxM;          // small: largest front height; gather: largest m, which sets the number of row slices per
                         // front
  int64_t first2 = 0, count2 = 0;  // flow kinds: range in `blocks` of the same launch for FP64 factors
  int32_t nq = 1, nq2 = 1;          // flow kinds: blocks per task of the two lists (1, or kFlowQuadsF32 / kFlowQuadsF64)
  int32_t seg = 0;                  // small kinds: lanes per front when a wave takes several (8 or 16), else 0
};

struct SolvePlan {
  int64_t fvEntries = 0;  // Sp[nsuper]
  int64_t flowFlags = 0;  // 64-column blocks of all tiled supernodes (one flag each per direction)
  int64_t initSlots = 0;  // 64-entry partial products of the backward initialisation
  int64_t doneSlots = 0;  // counters of the fused launches (BlockTask::wslot / dslot)
  std::vector<int32_t> fronts;
  std::vector<BlockTask> blocks;
  std::vector<SolveLaunch> launches;
};

SolvePlan buildSolvePlan(const Symbolic& S, const FactorPlan& F, const PlanParams& prm = PlanParams());

}  // namespace amp
