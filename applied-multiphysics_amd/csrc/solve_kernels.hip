// Device supernodal triangular solves (forward L, optional D^-1, backward L^T) over the level schedule of the plan.
//
// Same mathematics as the reference solves (packages/core/src/la/sparse.mjs:1161-1305 for Cholesky, ldlt.mjs:565-643
// for LDL^T): per supernode an in-block triangular solve and the product with the below block; the scatter into
// ancestor rows is carried by each supernode's front vector and gathered by the parent through the relative map, so
// no two workgroups ever write the same entry (deterministic, no atomics).
#include <map>
#include <mutex>

#include "factor_kernels.h"

namespace amp {

__global__ void k_clear_flag(int32_t* flag) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *flag = 0;
}

// Raise the pass number of the tiled-supernode solve flags (wrapping to 1; flags start at 0) and clear the task
// counters of the pass's flow launches and the counters of its fused launches (stored after them).
__global__ void k_epoch_bump(int32_t* epoch, int32_t* ctr, int32_t nctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *epoch = *epoch == INT32_MAX ? 1 : *epoch + 1;
  for (int32_t i = threadIdx.x; i < nctr; i += blockDim.x) ctr[i] = 0;
}

// Next task of a flow launch for the calling workgroup: tasks are handed out in list order as workgroups ask for them,
// so every task a running task waits on was taken earlier by a workgroup that is running or done, and the launch
// drains whatever the number of co-resident workgroups (no cooperative launch needed).
__device__ inline int32_t nextTask(int32_t* ctr) {
  __shared__ int32_t s_i;
  __syncthreads();  // the previous task's readers of s_i are done
  if (threadIdx.x == 0) s_i = atomicAdd(ctr, 1);
  __syncthreads();
  return s_i;
}

template <typename S>
__global__ void k_perm_in(const double* __restrict__ b, S* __restrict__ y, const int32_t* __restrict__ perm, int32_t n,
                          int32_t q) {
  const int64_t total = int64_t(n) * q;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t k = t / n, i = t % n;
    y[t] = S(b[k * n + perm[i]]);
  }
}

template <typename S>
__global__ void k_perm_out(const S* __restrict__ y, double* __restrict__ x, const int32_t* __restrict__ perm, int32_t n,
                           int32_t q) {
  const int64_t total = int64_t(n) * q;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t k = t / n, i = t % n;
    x[k * n + perm[i]] = double(y[t]);
  }
}

template <typename T, typename S>
__global__ void k_diag_scale(S* __restrict__ y, const T* __restrict__ D, int32_t n, int32_t q) {
  const int64_t total = int64_t(n) * q;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x)
    y[t] = y[t] / S(D[t % n]);
}

// Front vector of supernode s for right-hand side k: own rows from y, the rest zero, plus every child's below part.
template <typename T, typename S>
__device__ inline void gatherFront(const SolveArgs<T>& a, int32_t s, const S* yk, const S* fk, S* v, int32_t nc,
                                   int32_t m) {
  const int32_t f = a.superStart[s];
  for (int i = threadIdx.x; i < m; i += blockDim.x) v[i] = i < nc ? yk[f + i] : S(0);
  __syncthreads();
  for (int32_t c = a.childHead[s]; c != -1; c = a.childNext[c]) {
    const int32_t ncc = a.superStart[c + 1] - a.superStart[c], muc = a.Sp[c + 1] - a.Sp[c] - ncc;
    const int32_t o = a.Sp[c] + ncc;
    for (int r = threadIdx.x; r < muc; r += blockDim.x) v[a.rel[o + r]] += fk[o + r];
    __syncthreads();
  }
}

// threads of the whole-front solve workgroups: one per row of the largest whole front
constexpr int kSmallSolveThreads = 128;
static_assert(kSmallMax <= kSmallSolveThreads, "a whole-front solve keeps one row per thread");

// Supernodes with m <= kSmallMax, one workgroup each (at least m threads), all right-hand sides of the pass together.
// Thread r keeps row r of the front vector in registers and loads its entry of the next column of L while the current
// one is applied; per column the pivot's thread forms the q solution values and broadcasts them through LDS
// (double-buffered, so one barrier per column).
template <typename T, typename S, int Q>
__global__ __launch_bounds__(256) void k_solve_small_fwd(const int32_t* __restrict__ fronts, SolveArgs<T> a, S* y,
                                                         S* fv, int32_t q, int64_t fvN) {
  __shared__ S v[Q][kSmallMax];
  __shared__ S xc[2][Q];
  const int32_t s = fronts[blockIdx.x];
  const int32_t f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s];
  const T* L = a.Lx + a.Lxp[s];
  const int r = threadIdx.x;
  const bool own = r < m;
  T lcur = (own && nc > 0) ? L[r] : T(0);  // L[r][col] for col = 0
  for (int32_t k = 0; k < q; k++) gatherFront<T, S>(a, s, y + int64_t(k) * a.n, fv + k * fvN, v[k], nc, m);
  S vr[Q];
#pragma unroll
  for (int k = 0; k < Q; k++) vr[k] = (own && k < q) ? v[k][r] : S(0);
  for (int col = 0; col < nc; col++) {
    const T lnext = (own && col + 1 < nc) ? L[r + int64_t(col + 1) * m] : T(0);
    const int b = col & 1;
    if (r == col) {
      const S dinv = a.ldlt ? S(1) : S(1) / S(lcur);  // one division per column for every right-hand side
#pragma unroll
      for (int k = 0; k < Q; k++) {
        if (k >= q) continue;
        const S x = a.ldlt ? vr[k] : vr[k] * dinv;
        vr[k] = x;
        xc[b][k] = x;
      }
    }
    __syncthreads();
    if (own && r > col) {
#pragma unroll
      for (int k = 0; k < Q; k++)
        if (k < q) vr[k] -= S(lcur) * xc[b][k];
    }
    lcur = lnext;
  }
  if (own) {
#pragma unroll
    for (int k = 0; k < Q; k++) {
      if (k >= q) continue;
      if (r < nc) y[int64_t(k) * a.n + f + r] = vr[k];
      fv[k * fvN + a.Sp[s] + r] = vr[k];
    }
  }
}

// Supernodes with m <= kSolveWaveMax, one wave each and four per workgroup: lane r keeps row r of the front vector
// of every right-hand side in registers, each column's solution value is broadcast with readlane, so the column loop
// needs no barrier and no LDS; the columns of L are loaded eight at a time, one group ahead.
constexpr int kWaveGroup = 8;

// out[j] = p[(c0 + dir j) stride] for the indices in [0, n) when `valid`, else 0: one group of independent loads.
template <typename T>
__device__ inline void loadGroup(const T* p, int64_t stride, int32_t c0, int dir, int32_t n, bool valid,
                                 T (&out)[kWaveGroup]) {
#pragma unroll
  for (int j = 0; j < kWaveGroup; j++) {
    const int32_t c = c0 + dir * j;
    out[j] = (valid && c >= 0 && c < n) ? p[int64_t(c) * stride] : T(0);
  }
}

template <typename T, typename S, int Q>
__global__ __launch_bounds__(256) void k_solve_wave_fwd(const int32_t* __restrict__ fronts, int32_t count,
                                                        SolveArgs<T> a, S* y, S* fv, int32_t q, int64_t fvN) {
  __shared__ S v[4][Q][kSolveWaveMax];
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int64_t idx = int64_t(blockIdx.x) * 4 + w;
  if (idx >= count) return;  // whole waves only: no workgroup barrier below
  const int32_t s = fronts[idx];
  const int32_t f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s];
  const T* L = a.Lx + a.Lxp[s];
#pragma unroll
  for (int k = 0; k < Q; k++)
    if (k < q) v[w][k][lane] = lane < nc ? y[int64_t(k) * a.n + f + lane] : S(0);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  for (int32_t c = a.childHead[s]; c != -1; c = a.childNext[c]) {
    const int32_t ncc = a.superStart[c + 1] - a.superStart[c], muc = a.Sp[c + 1] - a.Sp[c] - ncc;
    const int32_t o = a.Sp[c] + ncc;
    if (lane < muc) {
      const int32_t r = a.rel[o + lane];
#pragma unroll
      for (int k = 0; k < Q; k++)
        if (k < q) v[w][k][r] += fv[k * fvN + o + lane];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  S xr[Q];
#pragma unroll
  for (int k = 0; k < Q; k++) xr[k] = k < q ? v[w][k][lane] : S(0);
  T cur[kWaveGroup], nxt[kWaveGroup];
  const bool rowValid = lane < m;
  loadGroup(L + lane, m, 0, 1, nc, rowValid, cur);
  for (int32_t c0 = 0; c0 < nc; c0 += kWaveGroup) {
    if (c0 + kWaveGroup < nc) loadGroup(L + lane, m, c0 + kWaveGroup, 1, nc, rowValid, nxt);
#pragma unroll
    for (int j = 0; j < kWaveGroup; j++) {
      const int32_t col = c0 + j;
      if (col >= nc) break;
      const S l = lane > col ? S(cur[j]) : S(0);
      // the pivot's reciprocal, once per column for every right-hand side
      const S dinv = a.ldlt ? S(1) : S(1) / S(laneValue(cur[j], col));
#pragma unroll
      for (int k = 0; k < Q; k++) {
        if (k >= q) continue;
        const S x = a.ldlt ? laneValue(xr[k], col) : laneValue(xr[k], col) * dinv;
        xr[k] = lane == col ? x : xr[k] - l * x;
      }
    }
#pragma unroll
    for (int j = 0; j < kWaveGroup; j++) cur[j] = nxt[j];
  }
#pragma unroll
  for (int k = 0; k < Q; k++) {
    if (k >= q) continue;
    if (lane < nc) y[int64_t(k) * a.n + f + lane] = xr[k];
    if (lane < m) fv[k * fvN + a.Sp[s] + lane] = xr[k];
  }
}

// Whole fronts of at most G (8 or 16) rows, 64 / G per wave: segment g of the wave (lanes G g .. G g + G - 1) solves
// one front by the steps of k_solve_wave_fwd, lane r of the segment keeping row r, each column's solution values read
// from the segment's pivot lane with ds_bpermute. All G columns of a lane's row of L are loaded at once. The column
// loop runs in step across the wave (it ends when no segment has a column left) and a segment's own sizes enter
// through selects, so all lanes of a segment are active at every cross-lane read.
template <typename T, typename S, int Q, int G>
__global__ __launch_bounds__(256) void k_solve_seg_fwd(const int32_t* __restrict__ fronts, int32_t count,
                                                       SolveArgs<T> a, S* y, S* fv, int32_t q, int64_t fvN) {
  constexpr int FW = 64 / G;  // fronts per wave
  __shared__ S v[4][Q][64];   // the wave's front vectors: segment g at G g
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int sl = lane % G, base = lane - sl;
  const int64_t idx0 = (int64_t(blockIdx.x) * 4 + w) * FW;
  if (idx0 >= count) return;  // whole waves only: no workgroup barrier below
  const int64_t idx = idx0 + lane / G;
  int32_t s = -1, f = 0, nc = 0, m = 0;
  if (idx < count) {  // segments past the end keep an empty front
    s = fronts[idx];
    f = a.superStart[s];
    nc = a.superStart[s + 1] - f;
    m = a.Sp[s + 1] - a.Sp[s];
  }
  const T* L = a.Lx + (s >= 0 ? a.Lxp[s] : 0);
  T lr[G];  // this lane's row of L: lr[c] = L[sl][c]
#pragma unroll
  for (int c = 0; c < G; c++) lr[c] = (sl < m && c < nc) ? L[sl + int64_t(c) * m] : T(0);
#pragma unroll
  for (int k = 0; k < Q; k++)
    if (k < q) v[w][k][lane] = sl < nc ? y[int64_t(k) * a.n + f + sl] : S(0);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  for (int32_t c = s >= 0 ? a.childHead[s] : -1; c != -1; c = a.childNext[c]) {
    const int32_t ncc = a.superStart[c + 1] - a.superStart[c], muc = a.Sp[c + 1] - a.Sp[c] - ncc;
    const int32_t o = a.Sp[c] + ncc;
    if (sl < muc) {
      const int32_t r = a.rel[o + sl];
#pragma unroll
      for (int k = 0; k < Q; k++)
        if (k < q) v[w][k][base + r] += fv[k * fvN + o + sl];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  S xr[Q];
#pragma unroll
  for (int k = 0; k < Q; k++) xr[k] = k < q ? v[w][k][lane] : S(0);
#pragma unroll
  for (int col = 0; col < G; col++) {
    if (!waveAny(col < nc)) break;
    const bool act = col < nc;
    const S l = sl > col ? S(lr[col]) : S(0);
    // the pivot's reciprocal, once per column for every right-hand side
    const S dinv = a.ldlt ? S(1) : S(1) / S(segValue<G>(lr[col], col));
#pragma unroll
    for (int k = 0; k < Q; k++) {
      if (k >= q) continue;
      const S xc = segValue<G>(xr[k], col);
      const S x = a.ldlt ? xc : xc * dinv;
      xr[k] = !act ? xr[k] : sl == col ? x : xr[k] - l * x;
    }
  }
  if (s < 0) return;
#pragma unroll
  for (int k = 0; k < Q; k++) {
    if (k >= q) continue;
    if (sl < nc) y[int64_t(k) * a.n + f + sl] = xr[k];
    if (sl < m) fv[k * fvN + a.Sp[s] + sl] = xr[k];
  }
}

// Values that tasks of one fused launch hand to each other (front-vector rows, partial products, backward solutions)
// are stored and loaded with device-coherent accesses (written through to the point of coherence, never served from
// a non-coherent cache), so a hand-over needs no cache-wide write-back or invalidation, only the completion of the
// stores before the producer's count (CO = true); CO = false: ordinary accesses, for values handed over between
// launches.
template <bool CO, typename S>
__device__ inline S ldv(const S* p) {
  if constexpr (!CO) {
    return *p;
  } else if constexpr (sizeof(S) == 4) {
    return __uint_as_float(__hip_atomic_load(reinterpret_cast<uint32_t*>(const_cast<S*>(p)), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT));
  } else {
    return __longlong_as_double(int64_t(__hip_atomic_load(reinterpret_cast<uint64_t*>(const_cast<S*>(p)),
                                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)));
  }
}
template <bool CO, typename S>
__device__ inline void stv(S* p, S v) {
  if constexpr (!CO) {
    *p = v;
  } else if constexpr (sizeof(S) == 4) {
    __hip_atomic_store(reinterpret_cast<uint32_t*>(p), __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(reinterpret_cast<uint64_t*>(p), uint64_t(__double_as_longlong(v)), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Front vectors of the tiled supernodes of a level, rows [kGatherRows y, kGatherRows (y + 1)) of front fronts[x] per
// workgroup (x, y), so that the large fronts near the root are gathered by many workgroups. Every workgroup walks each
// child's whole relative map once for all right-hand sides and keeps the rows that land in its slice (the map
// increases, so a lane stops at the first row past the slice); per entry the order is the single-workgroup one: own
// value or zero, then the children in order. gatherSlice: rows [i0, i1) of supernode s (256 threads).
template <typename T, typename S, int Q, bool CO = false>
__device__ __forceinline__ void gatherSlice(int32_t s, int32_t i0, int32_t i1, const SolveArgs<T>& a, const S* y,
                                            S* fv, int32_t q, int64_t fvN, S (&vs)[Q][kGatherRows]) {
  const int32_t f = a.superStart[s], nc = a.superStart[s + 1] - f;
  // the slice of v is accumulated in LDS (children in order, one barrier each) and stored once
  for (int32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
#pragma unroll
    for (int k = 0; k < Q; k++)
      if (k < q) vs[k][i - i0] = i < nc ? y[int64_t(k) * a.n + f + i] : S(0);
  __syncthreads();
  for (int32_t c = a.childHead[s]; c != -1; c = a.childNext[c]) {
    const int32_t ncc = a.superStart[c + 1] - a.superStart[c], muc = a.Sp[c + 1] - a.Sp[c] - ncc;
    const int32_t o = a.Sp[c] + ncc;
    // U entries per lane are read before any is used, so walking a long map to the slice costs a round trip per
    // U x 256 entries, not per 256
    constexpr int U = Q == 1 ? 8 : 2;
    for (int32_t r0 = threadIdx.x; r0 < muc; r0 += U * int32_t(blockDim.x)) {
      int32_t iv[U];
      S vv[U][Q];
#pragma unroll
      for (int u = 0; u < U; u++) {
        const int32_t r = r0 + u * int32_t(blockDim.x);
        iv[u] = r < muc ? a.rel[o + r] : INT32_MAX;
#pragma unroll
        for (int k = 0; k < Q; k++) vv[u][k] = (r < muc && k < q) ? ldv<CO>(fv + k * fvN + o + r) : S(0);
      }
      bool past = false;
#pragma unroll
      for (int u = 0; u < U; u++) {
        past = past || iv[u] >= i1;
        if (!past && iv[u] >= i0)
#pragma unroll
          for (int k = 0; k < Q; k++)
            if (k < q) vs[k][iv[u] - i0] += vv[u][k];
      }
      if (past) break;
    }
    __syncthreads();
  }
  for (int32_t i = i0 + threadIdx.x; i < i1; i += blockDim.x)
#pragma unroll
    for (int k = 0; k < Q; k++)
      if (k < q) stv<CO>(fv + k * fvN + a.Sp[s] + i, vs[k][i - i0]);
}

template <typename T, typename S, int Q>
__global__ __launch_bounds__(256) void k_solve_gather(const int32_t* __restrict__ fronts, SolveArgs<T> a, S* y, S* fv,
                                                      int32_t q, int64_t fvN) {
  const int32_t s = fronts[blockIdx.x];
  const int32_t m = a.Sp[s + 1] - a.Sp[s];
  const int32_t i0 = int32_t(blockIdx.y) * kGatherRows, i1 = min(m, i0 + kGatherRows);
  if (i0 >= m) return;  // whole workgroup
  __shared__ S vs[Q][kGatherRows];
  gatherSlice<T, S, Q>(s, i0, i1, a, y, fv, q, fvN, vs);
}

namespace {

constexpr uint32_t kFlowSpinLimit = 1u << 22;  // about half a second of polling before a wait gives up
constexpr int kFlowRun = 8;                    // blocks consumed per synchronisation of a flow task

// The run of blocks a consumer stages at once: from block `k` up to `lim` (exclusive; lim < k walks downwards) when
// the flag of block lim - 1 (lim + 1 downwards) is raised (a block is published only after every block it consumes,
// so then the whole run is complete), else block k alone. Never waits: the consumer's mailbox loads wait for the
// values themselves (one device round trip after they are stored), and the flags only let a consumer that starts late
// take a run of finished blocks in one pass.
__device__ inline int32_t flowRun(const int32_t* flags, int32_t k, int32_t lim, int32_t want) {
  __shared__ int32_t s_r;
  const int32_t step = lim > k ? 1 : -1, last = lim - step;
  if (last == k) return lim;  // a run of one block: nothing to look up (uniform, no barrier)
  if (threadIdx.x == 0) {
    const int32_t fl = __hip_atomic_load(flags + last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_r = fl == want ? lim : k + step;
  }
  __syncthreads();
  return s_r;
}

// Mailbox word index of value r (0..63) of right-hand side jj of the block with flag `flag` (see FlowCtl).
__device__ inline int64_t boxSlot(int32_t flag, int32_t jj, int32_t r) {
  return (int64_t(flag) * kFlowRhs + jj) * 64 + r;
}

// Store / load one solved value with the pass number in every word (device-coherent 64-bit atomic accesses, each
// single-copy atomic). A load spins until its words carry `want`; a wait that exceeds the spin limit sets *err and
// returns what it read (the host reports the error).
__device__ inline void boxStore(uint64_t* box, int64_t slot, int32_t want, float v) {
  const uint64_t tag = uint64_t(uint32_t(want)) << 32;
  __hip_atomic_store(box + 2 * slot, tag | __float_as_uint(v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline void boxStore(uint64_t* box, int64_t slot, int32_t want, double v) {
  const uint64_t tag = uint64_t(uint32_t(want)) << 32;
  const uint64_t bits = uint64_t(__double_as_longlong(v));
  __hip_atomic_store(box + 2 * slot, tag | (bits & 0xffffffffull), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  __hip_atomic_store(box + 2 * slot + 1, tag | (bits >> 32), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ inline uint64_t boxWord(const uint64_t* p, int32_t want, int32_t* err) {
  uint64_t w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (uint32_t spins = 0; uint32_t(w >> 32) != uint32_t(want);) {
    // give up at the spin limit, or early once any other wait of the launch has given up
    if (++spins >= kFlowSpinLimit ||
        ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
      atomicOr(err, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(1);
    w = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  return w;
}
// Counters of the fused launches (cleared by the kernel that raises the pass number). The values a counter guards are
// stored and loaded device-coherently (ldv / stv with CO), so the counters need no fences: raiseCount adds one to
// counter `slot` once every wave has waited for its stores to complete and the workgroup has met at a barrier;
// waitCount: thread 0 polls counter `slot` until it reaches `count` (bounded like boxWord), and the workgroup barrier
// holds every wave's loads until then. A slot below 0: nothing.
__device__ inline void waitCount(const FlowCtl& c, int32_t slot, int32_t count) {
  if (slot < 0) return;
  if (threadIdx.x == 0) {
    // the waits of a launch's later levels last long and are many: their polls back off, so that they do not crowd
    // the device-coherent traffic of the running chains
    for (uint32_t spins = 0; __hip_atomic_load(c.done + slot, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < count;) {
      if (++spins >= kFlowSpinLimit ||
          ((spins & 255) == 0 && __hip_atomic_load(c.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)) {
        atomicOr(c.err, 1);
        break;
      }
      if (spins < 16)
        __builtin_amdgcn_s_sleep(1);
      else if (spins < 64)
        __builtin_amdgcn_s_sleep(4);
      else
        __builtin_amdgcn_s_sleep(16);
    }
  }
  __syncthreads();
}

__device__ inline void raiseCount(const FlowCtl& c, int32_t slot) {
  if (slot < 0) return;
  __builtin_amdgcn_s_waitcnt(0);  // this wave's stores have completed (the barrier alone waits for LDS only)
  __syncthreads();
  if (threadIdx.x == 0) __hip_atomic_fetch_add(c.done + slot, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <typename S>
__device__ inline S boxLoad(const uint64_t* box, int64_t slot, int32_t want, int32_t* err);
template <>
__device__ inline float boxLoad<float>(const uint64_t* box, int64_t slot, int32_t want, int32_t* err) {
  return __uint_as_float(uint32_t(boxWord(box + 2 * slot, want, err)));
}
template <>
__device__ inline double boxLoad<double>(const uint64_t* box, int64_t slot, int32_t want, int32_t* err) {
  const uint64_t lo = boxWord(box + 2 * slot, want, err) & 0xffffffffull;
  const uint64_t hi = boxWord(box + 2 * slot + 1, want, err) & 0xffffffffull;
  return __longlong_as_double(int64_t((hi << 32) | lo));
}

}  // namespace

// Forward solve of the tiled supernodes of one level. A task has NQ quarters of 256 threads, each owning 64 rows
// (r0 + 64 qd ..); thread (wave w, lane) of a quarter holds row 16 w + (lane & 15) and the columns (lane >> 4) + 4 i of
// every consumed 64-column block, so a row sum is completed by two lane shuffles. The x blocks are staged in LDS once
// for all quarters, a run of complete blocks at a time, and a thread's L values are loaded one block ahead. Below
// tasks (writer == 0) store v_below - sum into the front vector. A chain task owns the `writer` blocks blk, blk + 1,
// ...: after the blocks before blk, quarter i forms v = v_own - sum for block blk + i, applies the block's inverted
// diagonal block (x = W v), publishes x (y, mailbox, flag) and hands it to the later quarters through LDS, which apply
// it to their rows before the next quarter's turn: inside a task, a block passes to the next without a round trip
// through device memory. Q is the number of right-hand sides the kernel is built for (q <= Q).
template <typename T, typename S, int Q, int NQ, bool CO = false>
__device__ __forceinline__ void fwdTask(const BlockTask t, const SolveArgs<T>& a, S* y, S* fv, int32_t q,
                                        int64_t fvN, const FlowCtl& c, int32_t want, S (&xs)[Q][kFlowRun * kTileEdge],
                                        S (&vh)[Q][kTileEdge]) {
  constexpr int E = kTileEdge;
  const int tid = threadIdx.x, qd = tid >> 8, lane = tid & 63;
  const int rq = 16 * ((tid >> 6) & 3) + (lane & 15), cp = lane >> 4;  // row in the quarter, column phase
  const int32_t s = t.s, f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s];
  const int32_t r = t.r0 + E * qd + rq;
  const bool live = r < t.r1;
  const T* Lr = a.Lx + a.Lxp[s] + (live ? r : t.r0) + int64_t(cp) * m;
  const int32_t nb = (nc + E - 1) / E;
  const bool own = qd < t.writer;            // the quarter solves block blk + qd
  const bool local0 = !t.writer && nb == 1;  // a below task of a single-block supernode forms x_0 = W_0 v_own itself
  T wv[16];  // row rq of the quarter's inverted diagonal block (quarter 0's W_0 for local0), columns cp + 4 i
  if (own || (local0 && qd == 0)) {
    const T* Wt = a.W + a.Woff[s] + int64_t(own ? t.blk + qd : 0) * E * E;
#pragma unroll
    for (int i = 0; i < 16; i++) wv[i] = Wt[rq + (cp + 4 * i) * E];
  }
  S acc[Q], vown[Q];  // acc: the thread's column phase of the row sums; vown: the row's front-vector value (read
                      // before any wait: the gather wrote it earlier)
#pragma unroll
  for (int jj = 0; jj < Q; jj++) {
    acc[jj] = S(0);
    vown[jj] = (cp == 0 && live && jj < q) ? ldv<CO>(fv + jj * fvN + a.Sp[s] + r) : S(0);
  }
  // blocks the row consumes: those before blk (all of them for a below task), then in a chain task the task's own
  // blocks before the quarter's
  const int32_t steps = t.writer ? t.blk : nb;
  const int32_t total = steps + (own ? qd : 0);
  T cur[16], nxt[16];
  auto loadL = [&](T* dst, int32_t k) {
    const int32_t cb = k * E + cp;
#pragma unroll
    for (int i = 0; i < 16; i++) dst[i] = (live && cb + 4 * i < nc) ? Lr[int64_t(k * E + 4 * i) * m] : T(0);
  };
  // acc += L[r, block] x_block for x in LDS at xb (column cp + 4 i of the block at xb[cp + 4 i])
  auto apply = [&](const T* lv, int xo) {
#pragma unroll
    for (int jj = 0; jj < Q; jj++) {
      if (jj < q) {
        S sum = S(0);
#pragma unroll
        for (int i = 0; i < 16; i++) sum += S(lv[i]) * xs[jj][xo + 4 * i];
        acc[jj] += sum;
      }
    }
  };
  // sum over the four column phases of a row (lanes 16 apart): every lane gets the row's total
  auto rowTotal = [](S v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
  };
  if (total > 0) loadL(cur, 0);
  if (local0) {
    if (tid < E)
      for (int jj = 0; jj < q; jj++) vh[jj][tid] = tid < nc ? ldv<CO>(fv + jj * fvN + a.Sp[s] + tid) : S(0);
    __syncthreads();
    if (qd == 0) {
#pragma unroll
      for (int jj = 0; jj < Q; jj++) {
        S sum = S(0);
        if (jj < q) {
#pragma unroll
          for (int i = 0; i < 16; i++) sum += S(wv[i]) * vh[jj][cp + 4 * i];
        }
        sum = rowTotal(sum);
        if (cp == 0) xs[jj][rq] = (jj < q && rq < nc) ? sum : S(0);
      }
    }
    __syncthreads();
    apply(cur, cp);
  }
  for (int32_t k0 = 0; k0 < (local0 ? 0 : steps);) {
    const int32_t k1 = flowRun(c.fwdFlag + t.flag0, k0, min(steps, k0 + kFlowRun), want);
    for (int idx = tid; idx < (k1 - k0) * E; idx += 256 * NQ) {
      const int32_t col = k0 * E + idx;
#pragma unroll
      for (int jj = 0; jj < Q; jj++)
        if (jj < q)
          xs[jj][idx] = col < nc ? boxLoad<S>(c.fwdBox, boxSlot(t.flag0 + k0 + idx / E, jj, idx % E), want, c.err)
                                 : S(0);
    }
    __syncthreads();
    for (int32_t k = k0; k < k1; k++) {
      if (k + 1 < total) loadL(nxt, k + 1);
      apply(cur, (k - k0) * E + cp);
#pragma unroll
      for (int i = 0; i < 16; i++) cur[i] = nxt[i];
    }
    k0 = k1;
    if (k0 < steps) __syncthreads();  // the next run is staged into xs
  }
  if (!t.writer) {
#pragma unroll
    for (int jj = 0; jj < Q; jj++) {
      const S v = rowTotal(acc[jj]);
      if (jj < q && cp == 0 && live) stv<CO>(fv + jj * fvN + a.Sp[s] + r, vown[jj] - v);
    }
    return;
  }
  // the task's blocks in order: quarter i solves block blk + i, the later quarters apply it
  for (int32_t i = 0; i < t.writer; i++) {
    const int32_t kb = t.blk + i;
    if (qd == i) {
#pragma unroll
      for (int jj = 0; jj < Q; jj++) {
        const S v = rowTotal(acc[jj]);
        if (cp == 0) vh[jj][rq] = (jj < q && live) ? vown[jj] - v : S(0);
      }
    }
    __syncthreads();  // v of block kb in vh (the previous hand-over in xs has been applied)
    if (qd == i) {
#pragma unroll
      for (int jj = 0; jj < Q; jj++) {
        S sum = S(0);
        if (jj < q) {
#pragma unroll
          for (int i2 = 0; i2 < 16; i2++) sum += S(wv[i2]) * vh[jj][cp + 4 * i2];
        }
        sum = rowTotal(sum);
        if (cp == 0) {
          xs[jj][rq] = (jj < q && live) ? sum : S(0);  // hand-over to the later quarters
          if (jj < q && live) {
            y[int64_t(jj) * a.n + f + r] = sum;
            if (nb > 1) boxStore(c.fwdBox, boxSlot(t.flag0 + kb, jj, rq), want, sum);
          }
        }
      }
    }
    __syncthreads();  // x of block kb in xs[.][0, 64); its mailbox words are stored
    // raise the block's flag (relaxed: consumers validate the mailbox words themselves); single blocks have no waiters
    if (nb > 1 && tid == 256 * i)
      __hip_atomic_store(c.fwdFlag + t.flag0 + kb, want, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (qd > i && qd < t.writer) {
      if (steps + i + 1 < total) loadL(nxt, steps + i + 1);
      apply(cur, cp);
#pragma unroll
      for (int i2 = 0; i2 < 16; i2++) cur[i2] = nxt[i2];
    }
  }
}

// Persistent form: workgroups take tasks in list order (nextTask). Every task waits only on tasks listed before it, so
// the lowest unfinished task never waits on an unfinished one and is being run: the launch always drains.
template <typename T, typename S, int Q, int NQ>
__global__ __launch_bounds__(256 * NQ) void k_solve_fwd_flow(const BlockTask* __restrict__ tasks, int32_t ntasks,
                                                             int32_t* ctr, SolveArgs<T> a, S* y, S* fv, int32_t q,
                                                             int64_t fvN, FlowCtl c) {
  __shared__ S xs[Q][kFlowRun * kTileEdge];
  __shared__ S vh[Q][kTileEdge];
  const int32_t want = *c.epoch;  // this pass's number (raised by the kernel that starts the pass)
  for (int32_t i = nextTask(ctr); i < ntasks; i = nextTask(ctr))
    fwdTask<T, S, Q, NQ>(tasks[i], a, y, fv, q, fvN, c, want, xs, vh);
}

template <typename T, typename S, int Q>
__global__ __launch_bounds__(256) void k_solve_small_bwd(const int32_t* __restrict__ fronts, SolveArgs<T> a, S* y,
                                                         int32_t q) {
  // thread i keeps solution row i (i < nc, at most one row per thread) in registers: z_i = y_i - sum_r L[nc + r][i]
  // y[below r], then the backward substitution loads its entry of the next row of L while the current one is applied,
  // one barrier per column (the broadcast of the pivot's values is double-buffered)
  __shared__ S xb[Q][kSmallMax];
  __shared__ S xc[2][Q];
  const int32_t s = fronts[blockIdx.x];
  const int32_t f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s], mu = m - nc;
  const T* L = a.Lx + a.Lxp[s];
  const int32_t* rows = a.Si + a.Sp[s] + nc;
  for (int32_t k = 0; k < q; k++)
    for (int r = threadIdx.x; r < mu; r += blockDim.x) xb[k][r] = y[int64_t(k) * a.n + rows[r]];
  __syncthreads();
  const int i = threadIdx.x;
  const bool own = i < nc;
  const T* Li = L + int64_t(own ? i : 0) * m;
  S zr[Q];
#pragma unroll
  for (int k = 0; k < Q; k++) zr[k] = (own && k < q) ? y[int64_t(k) * a.n + f + i] : S(0);
  if (own) {
    // eight entries of the column's below part loaded before any is applied (one round trip per eight rows)
    for (int r0 = 0; r0 < mu; r0 += 8) {
      S lv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) lv[u] = r0 + u < mu ? S(Li[nc + r0 + u]) : S(0);
#pragma unroll
      for (int u = 0; u < 8; u++)
#pragma unroll
        for (int k = 0; k < Q; k++)
          if (k < q && r0 + u < mu) zr[k] -= lv[u] * xb[k][r0 + u];
    }
  }
  T lcur = (own && nc > 0) ? Li[nc - 1] : T(0);  // L[col][i] for col = nc - 1
  for (int col = nc - 1; col >= 0; col--) {
    const T lnext = (own && col > 0) ? Li[col - 1] : T(0);
    const int b = col & 1;
    if (i == col) {
      const S dinv = a.ldlt ? S(1) : S(1) / S(lcur);  // one division per column for every right-hand side
#pragma unroll
      for (int k = 0; k < Q; k++) {
        if (k >= q) continue;
        const S x = a.ldlt ? zr[k] : zr[k] * dinv;
        zr[k] = x;
        xc[b][k] = x;
      }
    }
    __syncthreads();
    if (own && i < col) {
#pragma unroll
      for (int k = 0; k < Q; k++)
        if (k < q) zr[k] -= S(lcur) * xc[b][k];
    }
    lcur = lnext;
  }
  if (own) {
#pragma unroll
    for (int k = 0; k < Q; k++)
      if (k < q) y[int64_t(k) * a.n + f + i] = zr[k];
  }
}

// Backward counterpart of k_solve_wave_fwd: lane i keeps solution row i of the supernode; the below rows' values are
// held by lanes 0..mu-1 and broadcast with readlane, lane i walks column i of L.
template <typename T, typename S, int Q>
__global__ __launch_bounds__(256) void k_solve_wave_bwd(const int32_t* __restrict__ fronts, int32_t count,
                                                        SolveArgs<T> a, S* y, int32_t q) {
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int64_t idx = int64_t(blockIdx.x) * 4 + w;
  if (idx >= count) return;
  const int32_t s = fronts[idx];
  const int32_t f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s], mu = m - nc;
  const T* L = a.Lx + a.Lxp[s];
  const int32_t* rows = a.Si + a.Sp[s] + nc;
  S xb[Q], z[Q];
  const int32_t br = lane < mu ? rows[lane] : 0;
#pragma unroll
  for (int k = 0; k < Q; k++) {
    xb[k] = (k < q && lane < mu) ? y[int64_t(k) * a.n + br] : S(0);
    z[k] = (k < q && lane < nc) ? y[int64_t(k) * a.n + f + lane] : S(0);
  }
  // z_i -= sum_r L[nc + r][i] xb_r: lane i walks the below rows of its column i
  const bool colValid = lane < nc;
  const T* Li = L + int64_t(colValid ? lane : 0) * m;
  T cur[kWaveGroup], nxt[kWaveGroup];
  loadGroup(Li + nc, 1, 0, 1, mu, colValid, cur);
  for (int32_t r0 = 0; r0 < mu; r0 += kWaveGroup) {
    if (r0 + kWaveGroup < mu) loadGroup(Li + nc, 1, r0 + kWaveGroup, 1, mu, colValid, nxt);
#pragma unroll
    for (int j = 0; j < kWaveGroup; j++) {
      const int32_t r = r0 + j;
      if (r >= mu) break;
      const S l = S(cur[j]);
#pragma unroll
      for (int k = 0; k < Q; k++)
        if (k < q) z[k] -= l * laneValue(xb[k], r);
    }
#pragma unroll
    for (int j = 0; j < kWaveGroup; j++) cur[j] = nxt[j];
  }
  // in-block backward substitution, last column first: lane i < col reads L[col][i] from its column i
  loadGroup(Li, 1, nc - 1, -1, nc, colValid, cur);
  for (int32_t c0 = nc - 1; c0 >= 0; c0 -= kWaveGroup) {
    if (c0 - kWaveGroup >= 0) loadGroup(Li, 1, c0 - kWaveGroup, -1, nc, colValid, nxt);
#pragma unroll
    for (int j = 0; j < kWaveGroup; j++) {
      const int32_t col = c0 - j;
      if (col < 0) break;
      const S l = lane < col ? S(cur[j]) : S(0);
      const S dinv = a.ldlt ? S(1) : S(1) / S(laneValue(cur[j], col));
#pragma unroll
      for (int k = 0; k < Q; k++) {
        if (k >= q) continue;
        const S x = a.ldlt ? laneValue(z[k], col) : laneValue(z[k], col) * dinv;
        z[k] = lane == col ? x : z[k] - l * x;
      }
    }
#pragma unroll
    for (int j = 0; j < kWaveGroup; j++) cur[j] = nxt[j];
  }
#pragma unroll
  for (int k = 0; k < Q; k++)
    if (k < q && lane < nc) y[int64_t(k) * a.n + f + lane] = z[k];
}

// Backward counterpart of k_solve_seg_fwd: lane i of segment g keeps solution row i of the segment's supernode and
// loads column i of L whole; the below rows' values are held by the segment's lanes 0..mu-1 and read with ds_bpermute.
template <typename T, typename S, int Q, int G>
__global__ __launch_bounds__(256) void k_solve_seg_bwd(const int32_t* __restrict__ fronts, int32_t count,
                                                       SolveArgs<T> a, S* y, int32_t q) {
  constexpr int FW = 64 / G;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int sl = lane % G;
  const int64_t idx0 = (int64_t(blockIdx.x) * 4 + w) * FW;
  if (idx0 >= count) return;
  const int64_t idx = idx0 + lane / G;
  int32_t s = -1, f = 0, nc = 0, m = 0;
  if (idx < count) {
    s = fronts[idx];
    f = a.superStart[s];
    nc = a.superStart[s + 1] - f;
    m = a.Sp[s + 1] - a.Sp[s];
  }
  const int32_t mu = m - nc;
  const bool colValid = sl < nc;
  const T* Li = a.Lx + (s >= 0 ? a.Lxp[s] : 0) + int64_t(colValid ? sl : 0) * m;  // column sl of L
  T lc[G], lb[G];  // lc[c] = L[c][sl] (in-block rows), lb[r] = L[nc + r][sl] (below rows)
#pragma unroll
  for (int r = 0; r < G; r++) {
    lc[r] = (colValid && r < nc) ? Li[r] : T(0);
    lb[r] = (colValid && r < mu) ? Li[nc + r] : T(0);
  }
  const int32_t br = (s >= 0 && sl < mu) ? a.Si[a.Sp[s] + nc + sl] : 0;
  S xb[Q], z[Q];
#pragma unroll
  for (int k = 0; k < Q; k++) {
    xb[k] = (k < q && sl < mu) ? y[int64_t(k) * a.n + br] : S(0);
    z[k] = (k < q && colValid) ? y[int64_t(k) * a.n + f + sl] : S(0);
  }
  // z_i -= sum_r L[nc + r][i] xb_r
#pragma unroll
  for (int r = 0; r < G; r++) {
    if (!waveAny(r < mu)) break;
    const S l = S(lb[r]);
#pragma unroll
    for (int k = 0; k < Q; k++) {
      if (k >= q) continue;
      const S xv = segValue<G>(xb[k], r);
      z[k] = r < mu ? z[k] - l * xv : z[k];
    }
  }
  // in-block backward substitution, last column first: lane i < col applies L[col][i]
#pragma unroll
  for (int col = G - 1; col >= 0; col--) {
    if (!waveAny(col < nc)) continue;
    const bool act = col < nc;
    const S l = sl < col ? S(lc[col]) : S(0);
    const S dinv = a.ldlt ? S(1) : S(1) / S(segValue<G>(lc[col], col));
#pragma unroll
    for (int k = 0; k < Q; k++) {
      if (k >= q) continue;
      const S zc = segValue<G>(z[k], col);
      const S x = a.ldlt ? zc : zc * dinv;
      z[k] = !act ? z[k] : sl == col ? x : z[k] - l * x;
    }
  }
#pragma unroll
  for (int k = 0; k < Q; k++)
    if (k < q && colValid) y[int64_t(k) * a.n + f + sl] = z[k];
}

// Sums over the 64 lanes of a wave of V values per lane (V a power of two, at most 64) with V - 1 + log2(64 / V)
// exchanges instead of 6 V. The lanes hold their values permuted: lane l keeps value v in slot v ^ spreadMask<V>(l)
// (the low log2 V lane bits in reverse order). Step j pairs lane l with l ^ 2^j, which differs in the slot bit of
// weight h = V / 2^(j+1): each lane adds the partner's slot i + h to its own slot i (i < h), both the same value, and
// keeps the lower half. Plain butterflies over the remaining lane bits finish; lane l then holds in a[0] the whole-wave
// sum of value spreadMask<V>(l). Every register index is a constant, so no step selects between slots at run time.
template <int V>
__host__ __device__ constexpr int spreadMask(int l) {
  int r = 0;
  for (int w = V >> 1, b = 1; w > 0; w >>= 1, b <<= 1)
    if (l & b) r |= w;
  return r;
}
template <int V, typename S>
__device__ inline S waveSumSpread(S (&a)[V]) {
  static_assert(V > 0 && V <= 64 && (V & (V - 1)) == 0, "a power of two of values, at most one per lane");
  constexpr int LV = __builtin_ctz(V);
#pragma unroll
  for (int j = 0; j < LV; j++)
#pragma unroll
    for (int i = 0; i < (V >> (j + 1)); i++) a[i] += __shfl_xor(a[i + (V >> (j + 1))], 1 << j, 64);
  S v = a[0];
#pragma unroll
  for (int o = V; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Partial products of the backward initialisation: for the task's column block and below rows [r0, r1) (at most
// kInitRows), pz[slot][jj][c] = sum_r L[nc + r, c] y[below r] (y[below] gathered into the front vector). Wave w handles
// columns 16 w .. 16 w + 15, CG at a time; lane i takes rows r0 + i, r0 + 64 + i, ..., so every load instruction reads
// 64 consecutive rows of one column. Each lane gathers its rows' y values once. The CG Q column sums of a group are
// reduced across the wave together (waveSumSpread): lane l accumulates value v = c Q + jj (column c, right-hand side
// jj) in slot v ^ spreadMask(l), which it arranges by loading column c ^ (mask / Q) into slot c and right-hand side
// jj ^ (mask % Q) into slot jj.
template <typename T, typename S, int Q, bool CO = false>
__device__ __forceinline__ void initTask(const BlockTask t, const SolveArgs<T>& a, const S* y, S* pz, int32_t q) {
  constexpr int E = kTileEdge, NR = kInitRows / E;
  constexpr int CG = (Q > 1 && sizeof(S) == 8) ? 2 : 4;  // columns per group (registers: CG Q sums per lane)
  constexpr int V = CG * Q;
  static_assert((Q & (Q - 1)) == 0, "right-hand sides per pass: a power of two");
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int mask = spreadMask<V>(lane & (V - 1)), mc = mask / Q, mj = mask % Q;
  const int32_t s = t.s, nc = a.superStart[s + 1] - a.superStart[s], m = a.Sp[s + 1] - a.Sp[s];
  const T* Lb = a.Lx + a.Lxp[s] + nc;
  const int32_t* rows = a.Si + a.Sp[s] + nc;  // y[below] read through the row structure
  S yv[NR][Q];  // slot js: right-hand side js ^ mj
#pragma unroll
  for (int u = 0; u < NR; u++) {
    const int32_t i = t.r0 + E * u + lane;
    const int32_t yi = i < t.r1 ? rows[i] : 0;
#pragma unroll
    for (int js = 0; js < Q; js++) {
      const int jj = js ^ mj;
      yv[u][js] = (i < t.r1 && jj < q) ? ldv<CO>(y + int64_t(jj) * a.n + yi) : S(0);
    }
  }
#pragma unroll 1
  for (int g = 0; g < 16 / CG; g++) {
    const int cj0 = 16 * w + CG * g;  // first column (in the block) of this group
    S acc[V];                         // slot cs Q + js: column cs ^ mc, right-hand side js ^ mj
#pragma unroll
    for (int v = 0; v < V; v++) acc[v] = S(0);
#pragma unroll
    for (int u = 0; u < NR; u++) {
      const int32_t i = t.r0 + E * u + lane;
#pragma unroll
      for (int cs = 0; cs < CG; cs++) {
        const int32_t col = t.blk * E + cj0 + (cs ^ mc);
        const S lv = (i < t.r1 && col < nc) ? S(Lb[int64_t(col) * m + i]) : S(0);
#pragma unroll
        for (int js = 0; js < Q; js++) acc[cs * Q + js] += lv * yv[u][js];
      }
    }
    const S sum = waveSumSpread<V>(acc);  // lane l < V: value `mask` (column mask / Q, right-hand side mask % Q)
    if (lane < V && mj < q) stv<CO>(pz + (int64_t(t.p0) * q + mj) * E + cj0 + mc, sum);
  }
}

template <typename T, typename S, int Q>
__global__ __launch_bounds__(256) void k_solve_bwd_init(const BlockTask* __restrict__ tasks, SolveArgs<T> a,
                                                        const S* y, S* pz, int32_t q) {
  initTask<T, S, Q>(tasks[blockIdx.x], a, y, pz, q);
}

// Backward solve of the tiled supernodes of one level. A task has NQ quarters of 256 threads, quarter qd owning column
// block blk + qd; wave w of a quarter handles the columns 16 w .. 16 w + 15 (lane & 15) and lane >> 4 selects every
// fourth row, so each load instruction reads four consecutive rows of 16 columns. z starts from y_own minus the
// block's partial products of L21^T y[below] (summed in slot order before any wait); then z -= L[block k, column
// block]^T x_k for the blocks k after the task's (x staged in LDS once for all quarters, a run of complete blocks at a
// time, L loaded one block ahead). The task's blocks are then solved from the last one down: quarter i forms z,
// applies its block's transposed inverted diagonal block (x = W^T z), publishes x and hands it to the earlier quarters
// through LDS, which apply it to their columns before the next quarter's turn.
template <typename T, typename S, int Q, int NQ, bool CO = false>
__device__ __forceinline__ void bwdTask(const BlockTask t, const SolveArgs<T>& a, S* y, const S* pz, int32_t q,
                                        const FlowCtl& c, int32_t want, S (&xs)[Q][kFlowRun * kTileEdge],
                                        S (&zs)[Q][kTileEdge]) {
  constexpr int E = kTileEdge;
  const int tid = threadIdx.x, qd = tid >> 8, lane = tid & 63;
  const int cj = 16 * ((tid >> 6) & 3) + (lane & 15), rs = lane >> 4;
  const int32_t s = t.s, f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s];
  const int32_t nb = (nc + E - 1) / E, cb = t.blk + qd, col = cb * E + cj;
  const bool own = qd < t.writer;
  const bool live = own && col < nc;
  const T* Lc = a.Lx + a.Lxp[s] + int64_t(live ? col : t.r0) * m;
  T wv[16];
  if (own) {
    const T* Wt = a.W + a.Woff[s] + int64_t(cb) * E * E;
#pragma unroll
    for (int i = 0; i < 16; i++) wv[i] = Wt[(rs + 4 * i) + cj * E];
  }
  // the partial products of L21^T y[below] (slot order), summed before any wait, and this column's y value
  S acc[Q], yown[Q];
#pragma unroll
  for (int jj = 0; jj < Q; jj++) {
    S below = S(0);
    if (own && jj < q && rs == 0)
      for (int32_t p = 0; p < t.np; p++) below += ldv<CO>(pz + (int64_t(t.p0 + qd * t.np + p) * q + jj) * E + cj);
    acc[jj] = below;
    yown[jj] = (jj < q && rs == 0 && live) ? y[int64_t(jj) * a.n + f + col] : S(0);
  }
  T cur[16], nxt[16];
  auto loadL = [&](T* dst, int32_t k) {
    const int32_t rb = k * E + rs;
#pragma unroll
    for (int i = 0; i < 16; i++) dst[i] = (live && rb + 4 * i < nc) ? Lc[rb + 4 * i] : T(0);
  };
  // acc += L[block, col]^T x_block for x in LDS at xs[.][xo + 4 i] (row rs + 4 i of the block)
  auto apply = [&](const T* lv, int xo) {
#pragma unroll
    for (int jj = 0; jj < Q; jj++) {
      if (jj < q) {
        S sum = S(0);
#pragma unroll
        for (int i = 0; i < 16; i++) sum += S(lv[i]) * xs[jj][xo + 4 * i];
        acc[jj] += sum;
      }
    }
  };
  // sum over the four row phases of a column (lanes 16 apart): every lane gets the column's total
  auto colTotal = [](S v) {
    v += __shfl_xor(v, 16, 64);
    v += __shfl_xor(v, 32, 64);
    return v;
  };
  const int32_t kExt = t.blk + t.writer - 1;  // blocks after kExt are solved by earlier tasks
  const int32_t kEnd = own ? cb : nb - 1;     // the quarter's column block consumes the blocks after kEnd
  if (nb - 1 > kEnd) loadL(cur, nb - 1);
  for (int32_t k0 = nb - 1; k0 > kExt;) {
    const int32_t k1 = flowRun(c.bwdFlag + t.flag0, k0, max(kExt, k0 - kFlowRun), want);
    // blocks k0 down to k1 + 1 are complete; stage their x (block k at slot k0 - k)
    for (int idx = tid; idx < (k0 - k1) * E; idx += 256 * NQ) {
      const int32_t row = (k0 - idx / E) * E + (idx % E);
#pragma unroll
      for (int jj = 0; jj < Q; jj++)
        if (jj < q)
          xs[jj][idx] = row < nc ? boxLoad<S>(c.bwdBox, boxSlot(t.flag0 + k0 - idx / E, jj, idx % E), want, c.err)
                                 : S(0);
    }
    __syncthreads();
    for (int32_t k = k0; k > k1; k--) {
      if (k - 1 > kEnd) loadL(nxt, k - 1);
      apply(cur, (k0 - k) * E + rs);
#pragma unroll
      for (int i = 0; i < 16; i++) cur[i] = nxt[i];
    }
    k0 = k1;
    if (k0 > kExt) __syncthreads();  // the next run is staged into xs
  }
  // the task's blocks from the last one down: quarter i solves block blk + i, the earlier quarters apply it
  for (int32_t i = t.writer - 1; i >= 0; i--) {
    const int32_t kb = t.blk + i;
    if (qd == i) {
#pragma unroll
      for (int jj = 0; jj < Q; jj++) {
        const S v = colTotal(acc[jj]);
        if (rs == 0) zs[jj][cj] = (jj < q && live) ? yown[jj] - v : S(0);
      }
    }
    __syncthreads();  // z of block kb in zs (the previous hand-over in xs has been applied)
    if (qd == i) {
#pragma unroll
      for (int jj = 0; jj < Q; jj++) {
        S sum = S(0);
        if (jj < q) {
#pragma unroll
          for (int i2 = 0; i2 < 16; i2++) sum += S(wv[i2]) * zs[jj][rs + 4 * i2];
        }
        sum = colTotal(sum);
        if (rs == 0) {
          xs[jj][cj] = (jj < q && live) ? sum : S(0);  // hand-over to the earlier quarters
          if (jj < q && live) {
            stv<CO>(y + int64_t(jj) * a.n + f + col, sum);
            if (kb > 0) boxStore(c.bwdBox, boxSlot(t.flag0 + kb, jj, cj), want, sum);
          }
        }
      }
    }
    __syncthreads();  // x of block kb in xs[.][0, 64); its mailbox words are stored
    // raise the block's flag (relaxed, as forward); block 0 has no waiters
    if (kb > 0 && tid == 256 * i)
      __hip_atomic_store(c.bwdFlag + t.flag0 + kb, want, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (qd < i) {
      if (kb - 1 > kEnd) loadL(nxt, kb - 1);
      apply(cur, rs);
#pragma unroll
      for (int i2 = 0; i2 < 16; i2++) cur[i2] = nxt[i2];
    }
  }
}

// Persistent form, as k_solve_fwd_flow.
template <typename T, typename S, int Q, int NQ>
__global__ __launch_bounds__(256 * NQ) void k_solve_bwd_flow(const BlockTask* __restrict__ tasks, int32_t ntasks,
                                                             int32_t* ctr, SolveArgs<T> a, S* y, const S* pz,
                                                             int32_t q, FlowCtl c) {
  __shared__ S xs[Q][kFlowRun * kTileEdge];
  __shared__ S zs[Q][kTileEdge];
  const int32_t want = *c.epoch;
  for (int32_t i = nextTask(ctr); i < ntasks; i = nextTask(ctr))
    bwdTask<T, S, Q, NQ>(tasks[i], a, y, pz, q, c, want, xs, zs);
}

// Fused launches: the tasks of a run of levels in one launch (one block per chain task), with the gather tasks
// (forward, writer < 0) or initialisation tasks (backward, writer = 0) of each level among them. A task first waits for
// the counter of the tasks whose stores it reads (BlockTask::wslot) and raises its own when done (dslot); every task
// waits only on tasks listed before it, so the launch drains as the flow launches do.
template <typename T, typename S, int Q>
__global__ __launch_bounds__(256) void k_solve_fwd_fused(const BlockTask* __restrict__ tasks, int32_t ntasks,
                                                         int32_t* ctr, SolveArgs<T> a, S* y, S* fv, int32_t q,
                                                         int64_t fvN, FlowCtl c) {
  __shared__ S xs[Q][kFlowRun * kTileEdge];
  __shared__ S vh[Q][kTileEdge];
  __shared__ S vs[Q][kGatherRows];
  const int32_t want = *c.epoch;
  for (int32_t i = nextTask(ctr); i < ntasks; i = nextTask(ctr)) {
    const BlockTask t = tasks[i];
    waitCount(c, t.wslot, t.wcount);
    if (t.writer < 0)
      gatherSlice<T, S, Q, true>(t.s, t.r0, t.r1, a, y, fv, q, fvN, vs);
    else
      fwdTask<T, S, Q, 1, true>(t, a, y, fv, q, fvN, c, want, xs, vh);
    raiseCount(c, t.dslot);
  }
}

template <typename T, typename S, int Q>
__global__ __launch_bounds__(256) void k_solve_bwd_fused(const BlockTask* __restrict__ tasks, int32_t ntasks,
                                                         int32_t* ctr, SolveArgs<T> a, S* y, S* pz, int32_t q,
                                                         FlowCtl c) {
  __shared__ S xs[Q][kFlowRun * kTileEdge];
  __shared__ S zs[Q][kTileEdge];
  const int32_t want = *c.epoch;
  for (int32_t i = nextTask(ctr); i < ntasks; i = nextTask(ctr)) {
    const BlockTask t = tasks[i];
    waitCount(c, t.wslot, t.wcount);
    if (t.writer == 0)
      initTask<T, S, Q, true>(t, a, y, pz, q);
    else
      bwdTask<T, S, Q, 1, true>(t, a, y, pz, q, c, want, xs, zs);
    raiseCount(c, t.dslot);
  }
}

namespace {

// Co-resident workgroups of `threads` threads for `kernel` on device `dev` (cached per kernel and device).
int residentGroups(const void* kernel, int threads, int dev) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, int> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find({kernel, dev});
  if (it != cache.end()) return it->second;
  int cus = 0, perCu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, threads, 0) != hipSuccess)
    return 0;
  return cache[{kernel, dev}] = cus * perCu;
}

// Launch of a persistent flow kernel (NQ quarters of 256 threads per workgroup) over `ntasks` tasks with at most the
// co-resident number of workgroups (more would only wait for a slot); an ordinary launch on the pass's stream, so it
// shares the hardware queue of the pass's other kernels.
template <int NQ, typename K, typename... Args>
hipError_t launchFlow(K kernel, int64_t ntasks, hipStream_t st, const BlockTask* tasks, int32_t* ctr, Args... args) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  const int cap = residentGroups((const void*)kernel, 256 * NQ, dev);
  if (cap <= 0) return hipErrorInvalidConfiguration;
  const unsigned grid = unsigned(std::min<int64_t>(ntasks, cap));
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(256 * NQ), 0, st, tasks, int32_t(ntasks), ctr, args...);
  return hipGetLastError();
}

template <typename T, typename S>
hipError_t solvePass(const DeviceSolvePlan& dp, const SolveArgs<T>& a, const double* b, double* x, int32_t q, S* y, S* fv,
               S* pz, const FlowCtl& fc, hipStream_t st, ProfileSink* prof) {
  const int32_t n = a.n;
  const unsigned g = unsigned(std::min<int64_t>((int64_t(n) * q + 255) / 256, 65536));
  hipLaunchKernelGGL(k_perm_in<S>, dim3(g), dim3(256), 0, st, b, y, a.perm, n, q);
  bool scaled = false;
  hipError_t err = hipSuccess;
  int32_t nflow = 0;  // flow launches so far: index of the next one's task counter
  for (const SolveLaunch& L : dp.launches) {
    if (a.ldlt && !scaled && !solveKindIsForward(L.kind)) {
      hipLaunchKernelGGL((k_diag_scale<T, S>), dim3(g), dim3(256), 0, st, y, a.D, n, q);
      scaled = true;
    }
    const unsigned grid = unsigned(L.count);
    const bool rec = prof && 2 * prof->used + 1 < prof->events.size();
    if (rec) {
      (void)hipEventRecord(prof->events[2 * prof->used], st);
      prof->kinds[prof->used] = 16 + L.kind;
      prof->levels[prof->used] = L.level;
    }
    switch (L.kind) {
      case kSolveSmallFwd:
        if (L.seg > 0) {  // 64 / seg fronts per wave
          const unsigned wg = unsigned((L.count + 4 * (64 / L.seg) - 1) / (4 * (64 / L.seg)));
          auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(wg), dim3(256), 0, st, dp.fronts + L.first, int32_t(L.count), a, y, fv, q,
                               dp.fvEntries);
          };
          if (L.seg == 8)
            q == 1 ? go(k_solve_seg_fwd<T, S, 1, 8>) : go(k_solve_seg_fwd<T, S, kFlowRhs, 8>);
          else
            q == 1 ? go(k_solve_seg_fwd<T, S, 1, 16>) : go(k_solve_seg_fwd<T, S, kFlowRhs, 16>);
        } else if (L.maxM <= kSolveWaveMax) {
          const unsigned wg = unsigned((L.count + 3) / 4);
          if (q == 1)
            hipLaunchKernelGGL((k_solve_wave_fwd<T, S, 1>), dim3(wg), dim3(256), 0, st, dp.fronts + L.first,
                               int32_t(L.count), a, y, fv, q, dp.fvEntries);
          else
            hipLaunchKernelGGL((k_solve_wave_fwd<T, S, kFlowRhs>), dim3(wg), dim3(256), 0, st, dp.fronts + L.first,
                               int32_t(L.count), a, y, fv, q, dp.fvEntries);
        } else if (q == 1)
          hipLaunchKernelGGL((k_solve_small_fwd<T, S, 1>), dim3(grid), dim3(kSmallSolveThreads), 0, st,
                             dp.fronts + L.first, a, y, fv, q, dp.fvEntries);
        else
          hipLaunchKernelGGL((k_solve_small_fwd<T, S, kFlowRhs>), dim3(grid), dim3(kSmallSolveThreads), 0, st,
                             dp.fronts + L.first, a, y, fv, q, dp.fvEntries);
        break;
      case kSolveGather: {
        const dim3 gg(grid, unsigned((L.maxM + kGatherRows - 1) / kGatherRows));
        if (q == 1)
          hipLaunchKernelGGL((k_solve_gather<T, S, 1>), gg, dim3(256), 0, st, dp.fronts + L.first, a, y, fv, q,
                             dp.fvEntries);
        else
          hipLaunchKernelGGL((k_solve_gather<T, S, kFlowRhs>), gg, dim3(256), 0, st, dp.fronts + L.first, a, y, fv, q,
                             dp.fvEntries);
        break;
      }
      case kSolveFwdFlow: {
        // the launch's task list for this factor precision, with its blocks per task
        constexpr int NQ = sizeof(T) == 4 ? kFlowQuadsF32 : kFlowQuadsF64;
        const BlockTask* tk = dp.blocks + (NQ == kFlowQuadsF32 ? L.first : L.first2);
        const int64_t nt = NQ == kFlowQuadsF32 ? L.count : L.count2;
        const bool wide = (NQ == kFlowQuadsF32 ? L.nq : L.nq2) == NQ;
        int32_t* ctr = fc.ctr + nflow++;
        if (wide)
          err = q == 1 ? launchFlow<NQ>(k_solve_fwd_flow<T, S, 1, NQ>, nt, st, tk, ctr, a, y, fv, q, dp.fvEntries, fc)
                       : launchFlow<NQ>(k_solve_fwd_flow<T, S, kFlowRhs, NQ>, nt, st, tk, ctr, a, y, fv, q,
                                        dp.fvEntries, fc);
        else
          err = q == 1 ? launchFlow<1>(k_solve_fwd_flow<T, S, 1, 1>, nt, st, tk, ctr, a, y, fv, q, dp.fvEntries, fc)
                       : launchFlow<1>(k_solve_fwd_flow<T, S, kFlowRhs, 1>, nt, st, tk, ctr, a, y, fv, q, dp.fvEntries,
                                       fc);
        break;
      }
      case kSolveSmallBwd:
        if (L.seg > 0) {
          const unsigned wg = unsigned((L.count + 4 * (64 / L.seg) - 1) / (4 * (64 / L.seg)));
          auto go = [&](auto kernel) {
            hipLaunchKernelGGL(kernel, dim3(wg), dim3(256), 0, st, dp.fronts + L.first, int32_t(L.count), a, y, q);
          };
          if (L.seg == 8)
            q == 1 ? go(k_solve_seg_bwd<T, S, 1, 8>) : go(k_solve_seg_bwd<T, S, kFlowRhs, 8>);
          else
            q == 1 ? go(k_solve_seg_bwd<T, S, 1, 16>) : go(k_solve_seg_bwd<T, S, kFlowRhs, 16>);
        } else if (L.maxM <= kSolveWaveMax) {
          const unsigned wg = unsigned((L.count + 3) / 4);
          if (q == 1)
            hipLaunchKernelGGL((k_solve_wave_bwd<T, S, 1>), dim3(wg), dim3(256), 0, st, dp.fronts + L.first,
                               int32_t(L.count), a, y, q);
          else
            hipLaunchKernelGGL((k_solve_wave_bwd<T, S, kFlowRhs>), dim3(wg), dim3(256), 0, st, dp.fronts + L.first,
                               int32_t(L.count), a, y, q);
        } else if (q == 1)
          hipLaunchKernelGGL((k_solve_small_bwd<T, S, 1>), dim3(grid), dim3(kSmallSolveThreads), 0, st,
                             dp.fronts + L.first, a, y, q);
        else
          hipLaunchKernelGGL((k_solve_small_bwd<T, S, kFlowRhs>), dim3(grid), dim3(kSmallSolveThreads), 0, st,
                             dp.fronts + L.first, a, y, q);
        break;
      case kSolveBwdGather:  // never emitted by the plan
        return hipErrorInvalidValue;
      case kSolveBwdInit:
        if (q == 1)
          hipLaunchKernelGGL((k_solve_bwd_init<T, S, 1>), dim3(grid), dim3(256), 0, st, dp.blocks + L.first, a, y,
                             pz, q);
        else
          hipLaunchKernelGGL((k_solve_bwd_init<T, S, kFlowRhs>), dim3(grid), dim3(256), 0, st, dp.blocks + L.first, a,
                             y, pz, q);
        break;
      case kSolveFwdFused: {
        int32_t* ctr = fc.ctr + nflow++;
        const BlockTask* tk = dp.blocks + L.first;
        err = q == 1 ? launchFlow<1>(k_solve_fwd_fused<T, S, 1>, L.count, st, tk, ctr, a, y, fv, q, dp.fvEntries, fc)
                     : launchFlow<1>(k_solve_fwd_fused<T, S, kFlowRhs>, L.count, st, tk, ctr, a, y, fv, q,
                                     dp.fvEntries, fc);
        break;
      }
      case kSolveBwdFused: {
        int32_t* ctr = fc.ctr + nflow++;
        const BlockTask* tk = dp.blocks + L.first;
        err = q == 1 ? launchFlow<1>(k_solve_bwd_fused<T, S, 1>, L.count, st, tk, ctr, a, y, pz, q, fc)
                     : launchFlow<1>(k_solve_bwd_fused<T, S, kFlowRhs>, L.count, st, tk, ctr, a, y, pz, q, fc);
        break;
      }
      case kSolveBwdFlow: {
        // backward launches keep one block per task (their column loads make a wide task's panel slow to stream)
        const BlockTask* tk = dp.blocks + (sizeof(T) == 4 ? L.first : L.first2);
        const int64_t nt = sizeof(T) == 4 ? L.count : L.count2;
        int32_t* ctr = fc.ctr + nflow++;
        err = q == 1 ? launchFlow<1>(k_solve_bwd_flow<T, S, 1, 1>, nt, st, tk, ctr, a, y, pz, q, fc)
                     : launchFlow<1>(k_solve_bwd_flow<T, S, kFlowRhs, 1>, nt, st, tk, ctr, a, y, pz, q, fc);
        break;
      }
    }
    if (rec) {
      (void)hipEventRecord(prof->events[2 * prof->used + 1], st);
      prof->used++;
    }
    if (err != hipSuccess) return err;
  }
  if (a.ldlt && !scaled) hipLaunchKernelGGL((k_diag_scale<T, S>), dim3(g), dim3(256), 0, st, y, a.D, n, q);
  hipLaunchKernelGGL(k_perm_out<S>, dim3(g), dim3(256), 0, st, y, x, a.perm, n, q);
  return hipSuccess;
}

}  // namespace

template <typename T, typename S>
hipError_t launchSolvePlan(const DeviceSolvePlan& dp, const SolveArgs<T>& a, const double* b, double* x, int32_t q,
                           S* y, S* fv, S* pz, const FlowCtl& fc, hipStream_t st, ProfileSink* prof) {
  const int32_t n = a.n;
  if (n == 0 || q == 0) return hipSuccess;
  for (int32_t k0 = 0; k0 < q; k0 += kFlowRhs) {
    const int32_t qc = std::min(kFlowRhs, q - k0);
    hipLaunchKernelGGL(k_epoch_bump, dim3(1), dim3(256), 0, st, fc.epoch, fc.ctr, fc.nctr + fc.ndone);
    const hipError_t e = solvePass<T, S>(dp, a, b + int64_t(k0) * n, x + int64_t(k0) * n, qc, y, fv, pz, fc, st, prof);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

template hipError_t launchSolvePlan<double, double>(const DeviceSolvePlan&, const SolveArgs<double>&, const double*, double*,
                                              int32_t, double*, double*, double*, const FlowCtl&, hipStream_t, ProfileSink*);
template hipError_t launchSolvePlan<float, float>(const DeviceSolvePlan&, const SolveArgs<float>&, const double*, double*,
                                            int32_t, float*, float*, float*, const FlowCtl&, hipStream_t, ProfileSink*);

void launchClearFlag(int32_t* flag, hipStream_t st) { hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(64), 0, st, flag); }

}  // namespace amp
