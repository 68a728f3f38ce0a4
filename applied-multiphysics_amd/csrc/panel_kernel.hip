// Fused panel step of tiled fronts (gfx950): the factorization of a 64-column step's diagonal block, the triangular
// solve of its panel and the update of the later columns of its super-panel, in one launch per step.
//
// The three-launch form (k_diag_block, then k_gemm_tile in modes 2 and 0) spends most of a step in launch latency:
// each step's critical path is three dependent kernels. Here every workgroup owns one 64-row tile of the panel and
// eliminates the diagonal block itself (the elimination is deterministic, so every workgroup holds the same factor and
// inverse), forms its rows of L on the matrix cores, forms the rows of L that head the super-panel's later columns
// (the column tiles of its update) the same way, and updates its rows of those columns. The workgroup of the first
// row tile updates the next step's diagonal block and eliminates it at once, so the later steps of a super-panel
// start from its stored inverse instead of eliminating the block again. The reference's operations per
// entry are those of sparse.mjs:992-1098 (Cholesky) and ldlt.mjs:490-560 (LDL^T); the products are the ones
// k_gemm_tile computes (same operands, same depth order), so the factor is the three-launch one, bit for bit.
#include <algorithm>
#include <type_traits>

#include "diag_core.h"
#include "factor_kernels.h"

namespace amp {

namespace {

constexpr int kPanelPitch = kTileEdge + 4;     // LDS row pitch of the depth-major operand tiles
constexpr uint32_t kPanelSpinLimit = 1u << 22;  // polls of the launch counter before a wait gives up

// C (64 x 64) = A B^T over 64 depths, A at As[k P + i] and B at Bs[k P + j] (LDS, depth-major, zeros past the
// operands' extents). Eight waves in 4 x 2, each 16 rows x 32 columns (two 16 x 16 blocks); the matrix cores get the
// transposed product, so component r of acc[bj] is C[16 (w % 4) + lane % 16][32 (w / 4) + 16 bj + mfmaRow(lane, r)].
template <typename T>
__device__ inline void tileProduct(const T* As, const T* Bs, int w, int lane, typename Vec4<T>::type (&acc)[2]) {
  const int wr = 16 * (w & 3), wc = 32 * (w >> 2);
  acc[0] = typename Vec4<T>::type{0, 0, 0, 0};
  acc[1] = typename Vec4<T>::type{0, 0, 0, 0};
#pragma unroll
  for (int k4 = 0; k4 < kTileEdge; k4 += 4) {
    const int kk = (k4 + (lane >> 4)) * kPanelPitch + (lane & 15);
    const T af = As[kk + wr];
    const T b0 = Bs[kk + wc], b1 = Bs[kk + wc + 16];
    acc[0] = mfma16x16x4(b0, af, acc[0]);
    acc[1] = mfma16x16x4(b1, af, acc[1]);
  }
}

// Wait (thread 0 polls, the workgroup follows) until `target` workgroups have counted themselves on *ctr. A wait that
// gives up sets *err (the host reports the factorization as failed) and the workgroup goes on, so the launch ends.
__device__ inline void waitLaunchCount(int32_t* ctr, int32_t target, int32_t* err) {
  if (threadIdx.x == 0) {
    uint32_t spins = 0;
    while (__hip_atomic_load(ctr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++spins >= kPanelSpinLimit) {
        atomicOr(err, 1);
        break;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  __syncthreads();
}

}  // namespace

// One workgroup per PanelTask (eight waves). Steps:
//  1. issue the loads of the own row tile and of the column tiles that precede it (their rows of A, depths < nb);
//  2. eliminate the diagonal block (diagCore: factored columns and the rows of W = L11^-1 in registers; GROWTH as
//     k_diag_block: an FP64 redo when the largest multiplier exceeds the limit); kPanelPre: load its W and pivots;
//  3. count the workgroup on the front's counter (the block and its A rows are read); the owner waits for every task
//     of the front, then stores L11, W and the pivot records;
//  4. with W in LDS (the B operand), P = A W^T for the own tile and each column tile (L = P, LDL^T: L = P D^-1);
//  5. with the own L in LDS (the A operand) and each column tile's L (times D for LDL^T) as the B operand, the update
//     C -= L_own D L_c^T of the own rows in that tile's columns (read, subtract, write);
//  6. kPanelNext: eliminate the next block (updated in step 5, from registers through LDS) and store its L11, W and
//     pivot records;
//  7. store the own L rows (in place; a tile that is also a column tile first waits for every task of the front).
// The column tiles of the update are the panel's first row tiles after the block (tile c: panel rows
// [nb + 64 c, nb + 64 c + min(64, ucols - 64 c))), c <= the own tile's index: every entry written is at or below the
// front's diagonal in the columns of the super-panel (the diagonal tile also writes its upper part, which no kernel
// reads, as the product kernel does).
template <typename T, bool LDLT, bool GROWTH>
__global__ __launch_bounds__(64 * kDiagWaves, sizeof(T) == 8 ? 4 : 1) void k_panel_step(const PanelTask* __restrict__ tasks, FactorArgs<T> a) {
  using V4 = typename Vec4<T>::type;
  constexpr int E = kTileEdge, P = kPanelPitch, J = kDiagSlots, NCT = kPanelUpdateTiles, NW = kDiagWaves;
  constexpr int B = kDiagStep;
  // the elimination's scratch and, after it, the two operand tiles
  constexpr size_t kOpBytes = size_t(2) * E * P * sizeof(T), kDiagBytes = size_t(3) * 2 * B * E * sizeof(double);
  __shared__ __attribute__((aligned(32))) double ldsRaw[(kOpBytes > kDiagBytes ? kOpBytes : kDiagBytes) / sizeof(double)];
  __shared__ double gred[NW];
  __shared__ T dsh[E];  // pivots d_k (LDL^T; 1 past nb)
  chainPriority();
  const PanelTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  T* F = a.Lx + t.f_off;  // the diagonal block; front row ro + i, column c of the block's columns at F[ro + i + c ld]
  const int64_t ld = t.ld;
  const int nb = t.nb;
  const bool hasRows = t.rows > 0;
  const int own = hasRows ? (t.ro - nb) / E : -1;                 // the own tile's index among the panel's row tiles
  const int nct = (t.ucols + E - 1) / E;                          // column tiles of the super-panel's update
  const int nup = hasRows ? min(own + 1, nct) : 0;                // column tiles the own rows update
  const int nred = hasRows ? min(own, nct) : 0;                   // update tiles headed by other tasks' rows
  const int li = tid & 63, lk = tid >> 6;                         // loader: row li, depths lk + 8 q
  // 1. A rows of the own tile (pre[0]) and of column tiles 0 .. nred - 1 (pre[1 + c]); addresses clamped into the
  // operands, zeros chosen when staged
  T pre[1 + NCT][8] = {};
  auto tileRows = [&](int c) { return c < 0 ? t.rows : min(E, t.ucols - E * c); };
  auto tileRow0 = [&](int c) { return c < 0 ? t.ro : nb + E * c; };
  auto loadTile = [&](T (&dst)[8], int c) {
    const T* base = F + tileRow0(c) + min(li, tileRows(c) - 1);
#pragma unroll
    for (int q = 0; q < 8; q++) dst[q] = base[int64_t(min(lk + 8 * q, nb - 1)) * ld];
  };
  // FP64: the A rows are loaded after the elimination, before the workgroup counts itself (their registers would
  // otherwise stay live through it, and the kernel would not fit beside a bulk-update workgroup); the growth check is
  // an FP32 path only
  constexpr bool kEarly = sizeof(T) == 4;
  if constexpr (kEarly) {
    if (hasRows) loadTile(pre[0], -1);
#pragma unroll
    for (int c = 0; c < NCT; c++)
      if (c < nred) loadTile(pre[1 + c], c);
  }
  const bool pivotLane = lane < nb && w == (lane / B) % NW;  // the lane that records column `lane`'s pivot
  T wT[J];  // rows of W (stored precision): slot s holds row B (NW (s / B) + w) + s % B, column lane
  T dT = T(1);
  // the A rows and the block are read (every load of this workgroup has returned): count the workgroup
  auto arrive = [&]() {
#pragma unroll
    for (int c = 0; c < 1 + NCT; c++)
#pragma unroll
      for (int q = 0; q < 8; q++) asm volatile("" ::"v"(pre[c][q]));
    __syncthreads();
    if (tid == 0) __hip_atomic_fetch_add(a.panelCtr + t.ctr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  };
  if (t.flags & kPanelPre) {
    // 2'. the block was factored by the launch before: its W and pivots
    const T* Wt = a.W + t.w_off;
#pragma unroll
    for (int s = 0; s < J; s++) {
      const int j = B * (NW * (s / B) + w) + s % B;
      wT[s] = Wt[j + E * lane];
    }
    if (pivotLane && LDLT) dT = a.D[t.col0 + lane];
    if constexpr (!kEarly) {
      if (hasRows) loadTile(pre[0], -1);
#pragma unroll
      for (int c = 0; c < NCT; c++)
        if (c < nred) loadTile(pre[1 + c], c);
    }
    arrive();
  } else {
  // 2. the diagonal block. Without the growth check (whose FP64 redo reads the block again) the workgroup counts
  // itself as soon as the block and its A rows are loaded, so that the owner's store does not wait for the other
  // workgroups' eliminations
  const DiagTask dt{t.f_off, t.w_off, t.ld, t.nb, t.col0, 0};
  T recD = T(0), gmax = T(0);
  uint8_t recP = 0;
  int32_t recFail = INT32_MAX;
  T rA[J], rW[J];
  constexpr bool kEarlyArrive = kEarly && !GROWTH;
  diagLoad(dt, F, w, lane, rA, rW);
  if constexpr (kEarlyArrive) {
#pragma unroll
    for (int s = 0; s < J; s++) asm volatile("" ::"v"(rA[s]));
    arrive();
  }
  diagEliminate<T, LDLT>(dt, T(a.tau), w, lane, rA, rW, reinterpret_cast<T*>(ldsRaw), recD, recP, recFail, gmax);
  bool redo = false;
  if constexpr (GROWTH && std::is_same<T, float>::value) {
    float g = gmax;
    for (int o = 32; o > 0; o >>= 1) g = fmaxf(g, __shfl_xor(g, o, 64));
    if (lane == 0) gred[w] = double(g);
    __syncthreads();
    double gb = 0;
#pragma unroll
    for (int i = 0; i < NW; i++) gb = fmax(gb, gred[i]);
    // NaN multipliers (a failed pivot) keep the FP32 result, whose failure is recorded the same way
    if (gb > a.growthLimit) {
      redo = true;
      double dA[J], dW[J], dD = 0, dg = 0;
      uint8_t dP = 0;
      int32_t dFail = INT32_MAX;
      diagCore<double, LDLT>(dt, F, a.tau, w, lane, dA, dW, ldsRaw, dD, dP, dFail, dg);
#pragma unroll
      for (int s = 0; s < J; s++) wT[s] = T(dW[s]);
      if (pivotLane) dT = T(dD);
      // 3. (the redone block) count, then the owner stores
      arrive();
      if (t.owner) {
        waitLaunchCount(a.panelCtr + t.ctr, t.ntask, a.panelErr);
        if (tid == 0) atomicAdd(a.fallbacks, 1);
        diagStore<double, LDLT>(dt, a, w, lane, dA, dW, dD, dP, dFail);
      }
    }
  }
  if (!redo) {
#pragma unroll
    for (int s = 0; s < J; s++) wT[s] = rW[s];
    if (pivotLane) dT = recD;
    if constexpr (!kEarly) {
      if (hasRows) loadTile(pre[0], -1);
#pragma unroll
      for (int c = 0; c < NCT; c++)
        if (c < nred) loadTile(pre[1 + c], c);
    }
    // 3. count (unless counted already), then the owner stores
    if constexpr (!kEarlyArrive) arrive();
    if (t.owner) {
      waitLaunchCount(a.panelCtr + t.ctr, t.ntask, a.panelErr);
      diagStore<T, LDLT>(dt, a, w, lane, rA, rW, recD, recP, recFail);
    }
  }
  }
  if (!hasRows) return;  // a block with no rows below it
  T* opA = reinterpret_cast<T*>(ldsRaw);
  T* opB = opA + E * P;
  const int wr = 16 * (w & 3), wc = 32 * (w >> 2);
  const int row = wr + (lane & 15);  // the output row of this lane's accumulator components
  // column of component r of acc[bj]
  auto colOf = [&](int bj, int r) { return wc + 16 * bj + mfmaRow<T>(lane, r); };
  // staging of prefetched A rows as the A (opA[k P + i] = A[i][k]) or B operand (opB[k P + j] = A[j][k])
  auto stage = [&](T* op, const T (&src)[8], int rows) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int k = lk + 8 * q;
      op[k * P + li] = (li < rows && k < nb) ? src[q] : T(0);
    }
  };
  // an accumulator pair as an operand tile with the product's rows as the operand's rows: op[col P + row] = v
  auto stageAcc = [&](T* op, const T (&v)[2][4]) {
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) op[colOf(bj, r) * P + row] = v[bj][r];
  };
  // 4. P = A_own W^T: W as the B operand (opB[k P + j] = W[j][k], lane k), the pivots
  __syncthreads();  // the elimination's scratch is free
#pragma unroll
  for (int s = 0; s < J; s++) {
    const int j = B * (NW * (s / B) + w) + s % B;
    opB[lane * P + j] = (j < nb && lane < nb) ? wT[s] : T(0);
  }
  if (pivotLane) dsh[lane] = LDLT ? dT : T(1);
  if (tid >= nb && tid < E) dsh[tid] = T(1);
  stage(opA, pre[0], t.rows);
  __syncthreads();
  V4 acc[2];
  tileProduct<T>(opA, opB, w, lane, acc);
  // L = P (LDL^T: P d^-1, one reciprocal per column as the product kernel forms it)
  T lOwn[2][4];
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int r = 0; r < 4; r++) lOwn[bj][r] = LDLT ? acc[bj][r] * (T(1) / dsh[colOf(bj, r)]) : acc[bj][r];
  const bool next = t.flags & kPanelNext;
  const int nbn = min(E, t.ucols);  // (next) the next block's columns
  T nxt[2][4];                      // (next) this lane's entries of the updated next block
  if (nup > 0) {
    // 5. Q = L_own W (L_own as the A operand, W^T as the B operand: opB[k P + j] = W[k][j], W's row k = lane's slot)
    __syncthreads();  // the product has read opA and opB
    stageAcc(opA, lOwn);
#pragma unroll
    for (int s = 0; s < J; s++) {
      const int k = B * (NW * (s / B) + w) + s % B;
      opB[k * P + lane] = (k < nb && lane < nb) ? wT[s] : T(0);
    }
    __syncthreads();
    tileProduct<T>(opA, opB, w, lane, acc);
    T qv[2][4];
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) qv[bj][r] = acc[bj][r];
    __syncthreads();
    stageAcc(opA, qv);  // Q as the A operand of every update
    // per column tile c: C -= Q A_c^T (A_c: the panel rows heading the tile's columns, as the B operand; for LDL^T
    // L D L_c^T = L D D^-1 W A_c^T, so the pivots cancel)
#pragma unroll
    for (int c = 0; c < NCT; c++) {
      if (c >= nup) break;
      const int wcols = tileRows(c);  // columns of this update tile
      T* C = F + t.ro + int64_t(nb + E * c) * ld;
      T cv[2][4];  // this lane's C entries, read before the product so that they arrive meanwhile
#pragma unroll
      for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          cv[bj][r] = C[min(row, t.rows - 1) + int64_t(min(colOf(bj, r), wcols - 1)) * ld];
      if (c > 0) __syncthreads();  // the previous update has read opB
      stage(opB, c == own ? pre[0] : pre[1 + (c < NCT ? c : 0)], wcols);
      __syncthreads();
      tileProduct<T>(opA, opB, w, lane, acc);
      // (next: tile 0 is the next block; its lower part is stored once factored)
      const bool blockTile = next && c == 0;
#pragma unroll
      for (int bj = 0; bj < 2; bj++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int col = colOf(bj, r);
          const T v = cv[bj][r] - acc[bj][r];
          if (blockTile) nxt[bj][r] = v;
          if (row < t.rows && col < wcols && !(blockTile && row < nbn && row >= col)) C[row + int64_t(col) * ld] = v;
        }
    }
  }
  // 6. (kPanelNext: the first row tile, whose first update tile is the next step's diagonal block) the updated block,
  // from the registers into LDS (column-major, pitch P, its lower part), factored and stored as that step's owner
  // would store it
  if (next) {
    T* blk = opB;
    __syncthreads();  // the last product has read opB
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int col = colOf(bj, r);
        blk[row + col * P] = (row < nbn && row >= col) ? nxt[bj][r] : T(0);
      }
    __syncthreads();
    const DiagTask rd{0, 0, P, nbn, t.col0 + E, 0};                                 // read from LDS
    const DiagTask st{t.f_off + E + int64_t(E) * ld, t.w_off + E * E, t.ld, nbn, t.col0 + E, 0};  // store
    T nD = T(0), ng = T(0);
    uint8_t nP = 0;
    int32_t nFail = INT32_MAX;
    T nA[J], nW[J];
    diagCore<T, LDLT>(rd, blk, T(a.tau), w, lane, nA, nW, opA, nD, nP, nFail, ng);
    bool nredo = false;
    if constexpr (GROWTH && std::is_same<T, float>::value) {
      float g = ng;
      for (int o = 32; o > 0; o >>= 1) g = fmaxf(g, __shfl_xor(g, o, 64));
      if (lane == 0) gred[w] = double(g);
      __syncthreads();
      double gb = 0;
#pragma unroll
      for (int i = 0; i < NW; i++) gb = fmax(gb, gred[i]);
      if (gb > a.growthLimit) {
        nredo = true;
        double dA[J], dW[J], dD = 0, dg = 0;
        uint8_t dP = 0;
        int32_t dFail = INT32_MAX;
        diagCore<double, LDLT>(rd, blk, a.tau, w, lane, dA, dW, reinterpret_cast<double*>(opA), dD, dP, dFail, dg);
        if (tid == 0) atomicAdd(a.fallbacks, 1);
        diagStore<double, LDLT>(st, a, w, lane, dA, dW, dD, dP, dFail);
      }
    }
    if (!nredo) diagStore<T, LDLT>(st, a, w, lane, nA, nW, nD, nP, nFail);
  }
  // 7. the own rows of L, in place (a column tile: once every task of the front has read it)
  if (own < nct) waitLaunchCount(a.panelCtr + t.ctr, t.ntask, a.panelErr);
  T* Lr = F + t.ro;
#pragma unroll
  for (int bj = 0; bj < 2; bj++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int col = colOf(bj, r);
      if (row < t.rows && col < nb) Lr[row + int64_t(col) * ld] = lOwn[bj][r];
    }
}

template <typename T>
void launchPanelSteps(const PanelTask* tasks, unsigned count, const FactorArgs<T>& a, hipStream_t st) {
  // the instance with the growth check only when an FP32 factor asks for it (as launchDiagBlocks)
  const bool growth = std::is_same<T, float>::value && a.growthLimit > 0;
  constexpr bool F32 = std::is_same<T, float>::value;  // FP64 factors have no growth instance
  auto go = [&](auto kernel) { hipLaunchKernelGGL(kernel, dim3(count), dim3(64 * kDiagWaves), 0, st, tasks, a); };
  if (a.ldlt)
    growth ? go(k_panel_step<T, true, F32>) : go(k_panel_step<T, true, false>);
  else
    growth ? go(k_panel_step<T, false, F32>) : go(k_panel_step<T, false, false>);
}

int32_t panelResidentTasks() {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return INT32_MAX;
  const void* kernels[] = {(const void*)k_panel_step<double, false, false>, (const void*)k_panel_step<double, true, false>,
                           (const void*)k_panel_step<float, false, false>,  (const void*)k_panel_step<float, true, false>,
                           (const void*)k_panel_step<float, false, true>,   (const void*)k_panel_step<float, true, true>};
  int32_t fewest = INT32_MAX;
  for (const void* k : kernels) {
    int perCu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, k, 64 * kDiagWaves, 0) != hipSuccess) return INT32_MAX;
    fewest = std::min(fewest, perCu * cus);
  }
  return fewest;
}

// Counters of the kPanel launches (cleared by a kernel, so that a captured factor clears them on every replay).
__global__ void k_clear_ints(int32_t* p, int32_t n) {
  for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

void launchClearInts(int32_t* p, int32_t n, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_clear_ints, dim3(std::min(256, (n + 255) / 256)), dim3(256), 0, st, p, n);
}

template void launchPanelSteps<double>(const PanelTask*, unsigned, const FactorArgs<double>&, hipStream_t);
template void launchPanelSteps<float>(const PanelTask*, unsigned, const FactorArgs<float>&, hipStream_t);

}  // namespace amp
