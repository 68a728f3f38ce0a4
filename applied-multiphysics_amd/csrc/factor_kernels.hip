// Device kernels of the multifrontal factorization (Cholesky and LDL^T with static pivoting), gfx950.
//
// Every kernel executes one launch of the host plan (plan.h): a list of independent tasks, one workgroup per task.
// The arithmetic per front is the reference's dense partial factorization (packages/core/src/la/sparse.mjs:992-1098,
// ldlt.mjs:492-535, pivots as wasm/front-real.mjs:367-389), reorganised as a blocked right-looking factorization
// whose trailing products run on the matrix cores.
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include "factor_kernels.h"
#include "pivot.h"

namespace amp {


template <typename T>
__global__ void k_zero(const ZeroTask* __restrict__ tasks, T* __restrict__ Lx) {
  const ZeroTask t = tasks[blockIdx.x];
  T* p = Lx + t.off;
  for (int64_t i = threadIdx.x; i < t.len; i += blockDim.x) p[i] = T(0);
}

template <typename T>
__global__ void k_assemble(const int64_t* __restrict__ adst, const int32_t* __restrict__ lowerSrc,
                           const double* __restrict__ values, T* __restrict__ Lx, int64_t count) {
  for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < count; q += int64_t(gridDim.x) * blockDim.x) {
    const int64_t d = adst[q];
    if (d >= 0) Lx[d] = T(0.0 + values[lowerSrc[q]]);
  }
}

// Whole front in LDS: assemble A's columns, extend-add the children (in the reference's child order), factor the nc
// pivot columns right-looking, store the m x nc block and the mu x mu update matrix. The triangular loops give each
// wave whole columns and its lanes the rows of a column (no index division per entry).
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_small_front(const int32_t* __restrict__ fronts, FactorArgs<T> a,
                                                    const double* __restrict__ values, int32_t maxM) {
  constexpr int NW = NT / 64;
  chainPriority();
  extern __shared__ unsigned char smem[];
  T* F = reinterpret_cast<T*>(smem);
  T* wcol = F + maxM * maxM;
  __shared__ T s_inv, s_mult;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int32_t s = fronts[blockIdx.x];
  const int32_t f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s], mu = m - nc;
  for (int idx = tid; idx < m * m; idx += NT) F[idx] = T(0);
  __syncthreads();
  for (int32_t q = a.lowerP[f] + tid; q < a.lowerP[f + nc]; q += NT) F[a.amap[q]] = T(0.0 + values[a.lowerSrc[q]]);
  __syncthreads();
  for (int32_t c = a.childHead[s]; c != -1; c = a.childNext[c]) {
    const int32_t ncc = a.superStart[c + 1] - a.superStart[c], muc = a.Sp[c + 1] - a.Sp[c] - ncc;
    const int32_t* rel = a.rel + a.Sp[c] + ncc;
    const T* Uc = a.U + a.Uoff[c];
    for (int b = wv; b < muc; b += NW) {  // child column b, its rows r >= b
      T* Fb = F + rel[b] * m;
      const T* Ub = Uc + b * muc;
      for (int r = b + lane; r < muc; r += 64) Fb[rel[r]] += Ub[r];
    }
    __syncthreads();
  }
  for (int k = 0; k < nc; k++) {
    if (tid == 0) {
      T inv, mult;
      F[k + k * m] = pivotValue<T>(F[k + k * m], f + k, a, &inv, &mult);
      s_inv = inv;
      s_mult = mult;
    }
    __syncthreads();
    const T inv = s_inv, mult = s_mult;
    for (int i = k + 1 + tid; i < m; i += NT) {
      const T l = F[i + k * m] * inv;
      F[i + k * m] = l;
      wcol[i] = l * mult;
    }
    __syncthreads();
    const T* Fk = F + k * m;
    for (int j = k + 1 + wv; j < m; j += NW) {  // column j, rows i >= j
      T* Fj = F + j * m;
      const T wj = wcol[j];
      for (int i = j + lane; i < m; i += 64) Fj[i] -= Fk[i] * wj;
    }
    __syncthreads();
  }
  T* L = a.Lx + a.Lxp[s];
  for (int idx = tid; idx < m * nc; idx += NT) L[idx] = F[idx];
  if (mu > 0) {
    T* Us = a.U + a.Uoff[s];
    for (int b = wv; b < mu; b += NW)
      for (int r = lane; r < mu; r += 64) Us[r + b * mu] = F[(nc + r) + (nc + b) * m];
  }
}

// Fronts with m <= MB (16 or 32), one wave each and four per workgroup. The wave assembles its front in its own LDS
// block as k_small_front does; then lane j takes column j of the symmetric front into registers (rows i < j from row
// j of the stored lower triangle) and the nc elimination steps run there: the pivot column's entries reach every lane
// through readlane, and lane j finds its own multiplier in its row k, so no step needs LDS or a barrier. Columns
// return through LDS (lower triangle, upper zero) for coalesced stores of L and U.
template <typename T, int MB>
__global__ __launch_bounds__(256) void k_small_front_wave(const int32_t* __restrict__ fronts, int32_t count,
                                                          FactorArgs<T> a, const double* __restrict__ values) {
  __shared__ T Fs[4][MB * MB];
  chainPriority();
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int64_t idx = int64_t(blockIdx.x) * 4 + w;
  if (idx >= count) return;  // whole waves only: no workgroup barrier below
  T* F = Fs[w];
  const int32_t s = fronts[idx];
  const int32_t f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s], mu = m - nc;
  for (int i = lane; i < m * m; i += 64) F[i] = T(0);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  for (int32_t q = a.lowerP[f] + lane; q < a.lowerP[f + nc]; q += 64) F[a.amap[q]] = T(0.0 + values[a.lowerSrc[q]]);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  for (int32_t c = a.childHead[s]; c != -1; c = a.childNext[c]) {
    const int32_t ncc = a.superStart[c + 1] - a.superStart[c], muc = a.Sp[c + 1] - a.Sp[c] - ncc;
    const int32_t* rel = a.rel + a.Sp[c] + ncc;
    const T* Uc = a.U + a.Uoff[c];
    for (int i = lane; i < muc * muc; i += 64) {
      const int r = i % muc, b = i / muc;
      if (r >= b) F[rel[r] + rel[b] * m] += Uc[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  const bool own = lane < m;
  T cv[MB];
#pragma unroll
  for (int i = 0; i < MB; i++) cv[i] = (own && i < m) ? (i >= lane ? F[i + lane * m] : F[lane + i * m]) : T(0);
#pragma unroll
  for (int k = 0; k < MB; k++) {
    if (k >= nc) continue;
    T inv, mult;
    const T piv = pivotRule<T>(laneValue(cv[k], k), f + k, a, &inv, &mult, lane == 0);
    const T ljm = lane > k ? cv[k] * inv * mult : T(0);  // l_j d_k (row k of column j is F[j][k] by symmetry)
    if (lane == k) cv[k] = piv;
#pragma unroll
    for (int i = k + 1; i < MB; i++) {
      const T li = laneValue(cv[i], k) * inv;
      cv[i] = lane == k ? li : cv[i] - li * ljm;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
  for (int i = 0; i < MB; i++)
    if (own && i < m) F[i + lane * m] = i >= lane ? cv[i] : T(0);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  T* L = a.Lx + a.Lxp[s];
  for (int i = lane; i < m * nc; i += 64) L[i] = F[i];
  if (mu > 0) {
    T* Us = a.U + a.Uoff[s];
    for (int i = lane; i < mu * mu; i += 64) {
      const int r = i % mu, b = i / mu;
      Us[i] = F[(nc + r) + (nc + b) * m];
    }
  }
}

// Fronts with m <= G (8 or 16), 64 / G per wave: segment g (lanes G g .. G g + G - 1) takes one front through the
// steps of k_small_front_wave, lane j of the segment holding column j, the pivot column's entries read from the
// segment's lane k with ds_bpermute. The elimination loop runs in step across the wave (its exit is taken when no
// segment has a column left) and a segment's own sizes enter through selects, so all lanes of a segment are active at
// every cross-lane read; the assembly and store loops may diverge between segments.
template <typename T, int G>
__global__ __launch_bounds__(256) void k_small_front_seg(const int32_t* __restrict__ fronts, int32_t count,
                                                         FactorArgs<T> a, const double* __restrict__ values) {
  constexpr int FW = 64 / G;  // fronts per wave
  __shared__ T Fs[4][FW][G * G];
  chainPriority();
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
  const int sl = lane % G;  // lane within the segment
  const int64_t idx0 = (int64_t(blockIdx.x) * 4 + w) * FW;
  if (idx0 >= count) return;  // whole waves only: no workgroup barrier below
  const int64_t idx = idx0 + lane / G;
  T* F = Fs[w][lane / G];
  int32_t s = -1, f = 0, nc = 0, m = 0;
  if (idx < count) {  // segments past the end keep an empty front
    s = fronts[idx];
    f = a.superStart[s];
    nc = a.superStart[s + 1] - f;
    m = a.Sp[s + 1] - a.Sp[s];
  }
  const int32_t mu = m - nc;
  for (int i = sl; i < m * m; i += G) F[i] = T(0);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (s >= 0)
    for (int32_t q = a.lowerP[f] + sl; q < a.lowerP[f + nc]; q += G) F[a.amap[q]] = T(0.0 + values[a.lowerSrc[q]]);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  for (int32_t c = s >= 0 ? a.childHead[s] : -1; c != -1; c = a.childNext[c]) {
    const int32_t ncc = a.superStart[c + 1] - a.superStart[c], muc = a.Sp[c + 1] - a.Sp[c] - ncc;
    const int32_t* rel = a.rel + a.Sp[c] + ncc;
    const T* Uc = a.U + a.Uoff[c];
    for (int i = sl; i < muc * muc; i += G) {
      const int r = i % muc, b = i / muc;
      if (r >= b) F[rel[r] + rel[b] * m] += Uc[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  }
  const bool own = sl < m;
  T cv[G];
#pragma unroll
  for (int i = 0; i < G; i++) cv[i] = (own && i < m) ? (i >= sl ? F[i + sl * m] : F[sl + i * m]) : T(0);
#pragma unroll
  for (int k = 0; k < G; k++) {
    if (!waveAny(k < nc)) break;
    const bool act = k < nc;
    T inv, mult;
    const T piv = pivotRule<T>(segValue<G>(cv[k], k), f + k, a, &inv, &mult, act && sl == 0);
    const T ljm = sl > k ? cv[k] * inv * mult : T(0);  // l_j d_k (row k of column j is F[j][k] by symmetry)
    if (act && sl == k) cv[k] = piv;
#pragma unroll
    for (int i = k + 1; i < G; i++) {
      const T li = segValue<G>(cv[i], k) * inv;
      cv[i] = !act ? cv[i] : sl == k ? li : cv[i] - li * ljm;
    }
  }
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
#pragma unroll
  for (int i = 0; i < G; i++)
    if (own && i < m) F[i + sl * m] = i >= sl ? cv[i] : T(0);
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (s < 0) return;
  T* L = a.Lx + a.Lxp[s];
  for (int i = sl; i < m * nc; i += G) L[i] = F[i];
  if (mu > 0) {
    T* Us = a.U + a.Uoff[s];
    for (int i = sl; i < mu * mu; i += G) {
      const int r = i % mu, b = i / mu;
      Us[i] = F[(nc + r) + (nc + b) * m];
    }
  }
}

// Extend-add into parent columns [j0, j1) of a tiled front. The plan lists, per work item, the column range of
// every child that lands there (each child column maps to one parent column through the increasing relative map).
// Children are added one after another (reference order); a child's entries land on distinct parent entries, so the
// waves split a child's columns and their row chunks, each lane keeping eight rows in flight.
template <typename T>
__global__ __launch_bounds__(256) void k_extend_add(const EaTask* __restrict__ tasks,
                                                    const EaRange* __restrict__ ranges, FactorArgs<T> a) {
  constexpr int R = 8, CH = 64 * R, NWV = 4;  // rows in flight per lane, rows per chunk, waves
  chainPriority();
  const EaTask t = tasks[blockIdx.x];
  const int32_t wave = int32_t(threadIdx.x >> 6);
  const bool factorCols = t.j0 < a.superStart[t.s + 1] - a.superStart[t.s];
  const int32_t s = t.s;
  const int32_t nc = a.superStart[s + 1] - a.superStart[s], m = a.Sp[s + 1] - a.Sp[s], mu = m - nc;
  T* L = a.Lx + a.Lxp[s];
  T* Us = mu > 0 ? a.U + a.Uoff[s] : nullptr;
  const int lane = threadIdx.x & 63;
  for (int32_t q = t.r0; q < t.r1; q++) {
    const EaRange g = ranges[q];
    const int32_t c = g.c;
    const int32_t ncc = a.superStart[c + 1] - a.superStart[c], muc = a.Sp[c + 1] - a.Sp[c] - ncc;
    const int32_t* rel = a.rel + a.Sp[c] + ncc;
    const T* Uc = a.U + a.Uoff[c];
    // child column b lands in parent column rel[b], rows rel[r] for r >= b (lane r - b mod 64, R rows in flight)
    auto addRows = [&](int32_t b, int32_t r, int32_t rEnd) {
      const int32_t j = rel[b];
      const T* src = Uc + int64_t(b) * muc;
      T* dst = j < nc ? L + int64_t(j) * m : Us + int64_t(j - nc) * mu;
      const int32_t sh = j < nc ? 0 : nc;  // parent row -> row of the column's storage
      for (; r + 64 * (R - 1) < rEnd; r += 64 * R) {
        int32_t d[R];
        T v[R], w[R];
#pragma unroll
        for (int u = 0; u < R; u++) {
          d[u] = rel[r + 64 * u] - sh;
          v[u] = src[r + 64 * u];
        }
#pragma unroll
        for (int u = 0; u < R; u++) w[u] = dst[d[u]];
#pragma unroll
        for (int u = 0; u < R; u++) dst[d[u]] = w[u] + v[u];
      }
      for (; r < rEnd; r += 64) dst[rel[r] - sh] += src[r];
    };
    if (factorCols) {
      // into the factor columns (on the panel chain, few work items): the rows of each child column are cut into
      // chunks of 64 R and the (column, chunk) pairs are dealt to the waves in turn, so that a few long columns still
      // keep every wave busy
      int32_t slot = 0;
      for (int32_t b = g.b0; b < g.b1; b++) {
        const int32_t nch = (muc - b + CH - 1) / CH;
        const int32_t first = (wave - slot % NWV + NWV) % NWV;  // this wave's first chunk of the column
        slot += nch;
        for (int32_t ch = first; ch < nch; ch += NWV) addRows(b, b + ch * CH + lane, min(muc, b + (ch + 1) * CH));
      }
    } else {
      // into an update matrix (many work items): the waves split the child's columns
      for (int32_t b = g.b0 + wave; b < g.b1; b += NWV) addRows(b, b + lane, muc);
    }
    __syncthreads();
  }
}

// Output tiles of up to G x G (G = kGemmEdge or kTileEdge). G = 128: eight waves in 4 x 2, each owning 32 rows x 64
// columns; G = 64: four waves in 2 x 2, each 32 x 32 (MFMA blocks of 16 x 16), so every SIMD runs two waves of a
// 128-tile. The matrix cores are fed the transposed product (operand roles of A and B swapped), so the accumulator
// components of a lane hold 16 consecutive rows of one column of C and the epilogue reads and writes C straight from
// registers along columns. The depth is staged through LDS in chunks of KC (A and B depth-major, one LDS read per
// operand fragment); the next chunk is loaded into registers while the current one feeds the matrix cores. 128-tiles
// double-buffer the chunks in LDS (one barrier per chunk).
// Workgroups are dispatched round-robin over the eight XCDs; each XCD is given a contiguous run of the task list,
// whose neighbouring tiles share operand panels in that XCD's L2.
// chain != 0 (launches of the main stream): the waves raise their issue priority over the side stream's bulk updates,
// whose launches (chain == 0) take padding LDS so that each CU keeps room for the chain's workgroups (bulkLds).
// Loads are issued on every path (addresses clamped into the operands, chunks past the end re-read its last depth),
// and the entries outside an operand are replaced by zeros when a chunk is written to LDS: the wait before a chunk is
// written then counts only the loads issued before it, so the chunk loaded two steps ahead stays in flight.
// LDLT: the kernel built for LDL^T products scales the B operand by the pivots of its depths (B D) when a chunk is
// written to LDS (update tasks; the triangular-solve tasks are not scaled).
// static LDS of a k_gemm_tile instance (its two operand buffers)
template <typename T, int G>
constexpr size_t gemmTileLds() {
  return size_t(2) * (G == 128 ? 2 : 1) * ((G == 128 || sizeof(T) == 8) ? 16 : 32) * (G + 4) * sizeof(T);
}

template <typename T, int G, bool LDLT>
__global__ __launch_bounds__(G == 128 ? 512 : 256, G == 128 && !LDLT ? 4 : 1) void k_gemm_tile(const GemmTask* __restrict__ tasks, T* Lx, T* U,
                                                                   T* W, const T* __restrict__ D, int32_t chain) {
  using V4 = typename Vec4<T>::type;
  constexpr int WRW = G == 128 ? 4 : 2, WCW = 2, NT = 64 * WRW * WCW;  // wave grid, threads
  constexpr int WTR = G / WRW, WTC = G / WCW, NBI = WTR / 16, NBJ = WTC / 16;
  // depth chunk, LDS row pitch (16 for 128-tiles and FP64: two FP64 128-workgroups fit per CU, fewer loads in flight
  // per thread keep more waves resident; FP32 64-tiles take 32)
  constexpr int KC = (G == 128 || sizeof(T) == 8) ? 16 : 32, PA = G + 4;
  constexpr int LS = NT / G, NL = KC / LS;            // loader: depth stride, depths per thread
  constexpr int NBUF = G == 128 ? 2 : 1;
  __shared__ T As[NBUF][KC * PA];
  __shared__ T Bs[NBUF][KC * PA];
  static_assert(sizeof(As) + sizeof(Bs) == gemmTileLds<T, G>(), "LDS of the operand buffers");
  if (chain) chainPriority();
  const int nblk = gridDim.x, q8 = nblk / 8, r8 = nblk % 8, x8 = blockIdx.x % 8;
  const GemmTask t = tasks[x8 * q8 + min(x8, r8) + blockIdx.x / 8];
  // buffers chosen by selects of the kernel arguments (not an indexed array), so that the compiler keeps them in the
  // global address space (global loads and stores rather than flat ones, which LDS waits would also cover)
  auto base = [&](int32_t b) { return b == 0 ? Lx : b == 1 ? U : W; };
  const T* A = base(t.a_buf) + t.a_off;
  const T* B = base(t.b_buf) + t.b_off;
  T* C = base(t.c_buf) + t.c_off;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave % WRW) * WTR, wc = (wave / WRW) * WTC;
  const bool active = wr < t.rows && wc < t.cols;  // uniform per wave
  const bool scaleDepth = LDLT && t.mode != 2;
  const int li = tid % G, lk = tid / G;  // loader: row li, depths lk + LS r
  const bool rowA = li < t.rows, rowB = li < t.cols;
  // in-tile offsets are 32-bit (depth x leading dimension < 2^31, checked by the plan)
  const T* Ar = A + min(li, t.rows - 1);
  const T* Br = B + min(li, t.cols - 1);
  V4 acc[NBI][NBJ];
#pragma unroll
  for (int bi = 0; bi < NBI; bi++)
#pragma unroll
    for (int bj = 0; bj < NBJ; bj++) acc[bi][bj] = V4{0, 0, 0, 0};
  // two register sets of loaded depth values: a chunk is loaded two chunks before its turn to feed the matrix
  // cores, so its global-memory latency is covered by two chunks of products
  T pa[2][NL], pb[2][NL];
  T pd[LDLT ? 2 : 1][LDLT ? NL : 1];  // LDLT: pivots of the loaded depths
  auto load = [&](auto set, int k0) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int r = 0; r < NL; r++) {
      const int k = min(k0 + lk + LS * r, t.k - 1);
      pa[S][r] = Ar[k * t.lda];
      pb[S][r] = Br[k * t.ldb];
      if constexpr (LDLT) pd[S][r] = D[t.d_off + k];
    }
  };
  auto stash = [&](auto set, int buf, int k0) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int r = 0; r < NL; r++) {
      const bool kin = k0 + lk + LS * r < t.k;
      T bv = pb[S][r];
      if constexpr (LDLT) bv = scaleDepth ? bv * pd[S][r] : bv;
      As[buf][(lk + LS * r) * PA + li] = (kin && rowA) ? pa[S][r] : T(0);
      Bs[buf][(lk + LS * r) * PA + li] = (kin && rowB) ? bv : T(0);
    }
  };
  auto compute = [&](int buf) {
    if (!active) return;
#pragma unroll
    for (int k4 = 0; k4 < KC; k4 += 4) {
      const int kk = (k4 + (lane >> 4)) * PA + (lane & 15);
      T af[NBI], bf[NBJ];
#pragma unroll
      for (int b = 0; b < NBI; b++) af[b] = As[buf][kk + wr + 16 * b];
#pragma unroll
      for (int b = 0; b < NBJ; b++) bf[b] = Bs[buf][kk + wc + 16 * b];
#pragma unroll
      for (int bi = 0; bi < NBI; bi++)
#pragma unroll
        for (int bj = 0; bj < NBJ; bj++) acc[bi][bj] = mfma16x16x4(bf[bj], af[bi], acc[bi][bj]);
    }
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  // component r of acc[bi][bj]: C row wr + 16 bi + (lane & 15), C column wc + 16 bj + mfmaRow(lane, r)
  const int row0 = wr + (lane & 15);
  // C -= A B^T (mode 0): the last one or two steps load no chunk, so their register sets read the C entries of column
  // blocks 0 and 1 instead (entry r NBI + bi of a block in pa[S] then pb[S]), which arrive while the last chunks feed
  // the matrix cores
  static_assert(4 * NBI <= 2 * NL, "a column block of C fits in one register set");
  auto prefetchC = [&](auto set, int bj) {
    constexpr int S = decltype(set)::value;
    if (t.mode != 0 || !active) return;
    const T* Cc = C;
#pragma unroll
    for (int r = 0; r < 4; r++) {
      Cc = C + int64_t(min(wc + 16 * bj + mfmaRow<T>(lane, r), t.cols - 1)) * t.ldc;
#pragma unroll
      for (int bi = 0; bi < NBI; bi++) {
        const int u = r * NBI + bi;
        const T v = Cc[min(row0 + 16 * bi, t.rows - 1)];
        if (u < NL)
          pa[S][u < NL ? u : 0] = v;
        else
          pb[S][u < NL ? 0 : u - NL] = v;
      }
    }
  };
  auto prefetched = [&](auto set, int u) -> T {
    constexpr int S = decltype(set)::value;
    return u < NL ? pa[S][u < NL ? u : 0] : pb[S][u < NL ? 0 : u - NL];
  };
  load(S0{}, 0);
  load(S1{}, KC);
  int k0 = 0;
  if constexpr (NBUF == 2) {
    // double-buffered LDS: chunk c is read from buffer c & 1 while chunk c + 1 (already in registers) is written to
    // the other, and chunk c + 2 is loaded into the register set chunk c left; one barrier per chunk. Chunk c's
    // register set is c & 1 (the loop is unrolled by two so that the sets are compile-time).
    stash(S0{}, 0, 0);
    __syncthreads();
    // `set`: register set of chunk c + 1; buffer of chunk c is set ^ 1; `ld`: whether chunk c + 2 is loaded
    auto step = [&](auto set, int k0, auto ld) {
      constexpr int S = decltype(set)::value;
      // the buffer index is made opaque, so that the LDS addresses of both buffers are not precomputed and kept live
      // across the loop
      int buf = S ^ 1;
      asm volatile("" : "+s"(buf));
      if constexpr (decltype(ld)::value)
        load(std::integral_constant<int, S ^ 1>{}, k0 + 2 * KC);
      else
        prefetchC(std::integral_constant<int, S ^ 1>{}, S ^ 1);  // blocks 0, 1 into sets 0, 1
      compute(buf);
      if (k0 + KC < t.k) stash(set, buf ^ 1, k0 + KC);
      __syncthreads();
    };
    // pairs of chunks while a chunk two ahead exists (the pair's second load may re-read the last depth), then the
    // last one or two chunks without loads: at most one chunk is loaded in vain per tile
    for (; k0 + 2 * KC < t.k; k0 += 2 * KC) {
      step(S1{}, k0, std::true_type{});
      step(S0{}, k0 + KC, std::true_type{});
    }
    step(S1{}, k0, std::false_type{});
    if (k0 + KC < t.k) step(S0{}, k0 + KC, std::false_type{});
  } else {
    // one LDS buffer: chunk c (register set c & 1) is stashed, then chunk c + 2 is loaded into the same set while
    // chunk c feeds the matrix cores
    auto step = [&](auto set, int k0, auto ld) {
      int buf = 0;  // opaque, as above
      asm volatile("" : "+s"(buf));
      stash(set, buf, k0);
      __syncthreads();
      if constexpr (decltype(ld)::value)
        load(set, k0 + 2 * KC);
      else
        prefetchC(set, decltype(set)::value);  // blocks 0, 1 into sets 0, 1
      compute(buf);
      __syncthreads();
    };
    // pairs while a chunk two ahead exists, then the rest without loads, as above
    for (; k0 + 2 * KC < t.k; k0 += 2 * KC) {
      step(S0{}, k0, std::true_type{});
      step(S1{}, k0 + KC, std::true_type{});
    }
    step(S0{}, k0, std::false_type{});
    if (k0 + KC < t.k) step(S1{}, k0 + KC, std::false_type{});
  }
  if (!active) return;
  if (t.mode == 0) {
    const bool pre1 = k0 + KC < t.k;  // the second drain step ran: block 1 is in register set 1
    // C -= A B^T, one 16-column block at a time: its C entries are all read (addresses clamped into the tile) before
    // any is written, so a block costs one memory round trip (the compiler keeps a store before every later load of
    // C; reading and writing entry by entry cost a round trip per entry)
    constexpr int BJB = 1;  // column blocks per batch (more cost registers and were not faster)
#pragma unroll
    for (int bj0 = 0; bj0 < NBJ; bj0 += BJB) {
      T cv[BJB][4][NBI];
#pragma unroll
      for (int b = 0; b < BJB; b++) {
        const int bj = bj0 + b;
        if (bj == 0 || (bj == 1 && pre1)) {
#pragma unroll
          for (int r = 0; r < 4; r++)
#pragma unroll
            for (int bi = 0; bi < NBI; bi++)
              cv[b][r][bi] = bj == 0 ? prefetched(S0{}, r * NBI + bi) : prefetched(S1{}, r * NBI + bi);
        } else {
#pragma unroll
          for (int r = 0; r < 4; r++) {
            const T* Cc = C + int64_t(min(wc + 16 * bj + mfmaRow<T>(lane, r), t.cols - 1)) * t.ldc;
#pragma unroll
            for (int bi = 0; bi < NBI; bi++) cv[b][r][bi] = Cc[min(row0 + 16 * bi, t.rows - 1)];
          }
        }
      }
#pragma unroll
      for (int b = 0; b < BJB; b++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int col = wc + 16 * (bj0 + b) + mfmaRow<T>(lane, r);
          T* Cc = C + int64_t(col) * t.ldc;
#pragma unroll
          for (int bi = 0; bi < NBI; bi++) {
            const int row = row0 + 16 * bi;
            if (col < t.cols && row < t.rows) Cc[row] = cv[b][r][bi] - acc[bi][bj0 + b][r];
          }
        }
    }
    return;
  }
#pragma unroll
  for (int bj = 0; bj < NBJ; bj++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int col = wc + 16 * bj + mfmaRow<T>(lane, r);
      if (col >= t.cols) continue;
      const T dinv = (LDLT && t.mode == 2) ? T(1) / D[t.d_off + col] : T(1);
      T* Cc = C + int64_t(col) * t.ldc;
#pragma unroll
      for (int bi = 0; bi < NBI; bi++) {
        const int row = row0 + 16 * bi;
        if (row >= t.rows) continue;
        const T v = acc[bi][bj][r];
        if (t.mode == 1)
          Cc[row] = -v;
        else
          Cc[row] = LDLT ? v * dinv : v;
      }
    }
}

// nNeg / nPerturbed over the columns before the first failure (the counts the reference returns, ldlt.mjs:505-519).
template <typename T>
__global__ __launch_bounds__(256) void k_ldlt_counts(const T* __restrict__ D, const uint8_t* __restrict__ pert,
                                                     const int32_t* __restrict__ failCol, int32_t n,
                                                     int32_t* __restrict__ out) {
  __shared__ int32_t sn[256], sp[256];
  const int32_t lim = min(n, *failCol);
  int32_t cn = 0, cp = 0;
  for (int32_t j = threadIdx.x; j < lim; j += 256) {
    cn += D[j] < T(0);
    cp += pert[j];
  }
  sn[threadIdx.x] = cn;
  sp[threadIdx.x] = cp;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      sn[threadIdx.x] += sn[threadIdx.x + w];
      sp[threadIdx.x] += sp[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sn[0];
    out[1] = sp[0];
  }
}

__global__ void k_set_int(int32_t* p, int32_t v) { *p = v; }

// ---- launchers ------------------------------------------------------------------------------------------------
size_t bulkLds(const void* kernel, int threads, size_t staticLds) {
  static std::mutex mu;
  static std::map<std::pair<const void*, int>, size_t> pads;  // (kernel, device) -> dynamic LDS bytes
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lock(mu);
  auto it = pads.find({kernel, dev});
  if (it != pads.end()) return it->second;
  size_t pad = 0;
  int perCu = 0, ldsCu = 0, ldsBlock = 0;
  // kernels that fit two workgroups or fewer per CU are left alone: halving their residency cost more bulk throughput
  // than the chain gained (measured on the 128-tile FP64 and bf16 bulk kernels)
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, kernel, threads, 0) == hipSuccess && perCu >= 3 &&
      hipDeviceGetAttribute(&ldsCu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev) == hipSuccess &&
      hipDeviceGetAttribute(&ldsBlock, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess) {
    // smallest workgroup footprint (static + padding, in 1 KiB steps) at which only perCu - 1 workgroups fit
    const size_t want = size_t(ldsCu) / size_t(perCu) + 1;
    size_t p = want > staticLds ? (want - staticLds + 1023) / 1024 * 1024 : 0;
    if (p > 0 && staticLds + p <= size_t(ldsBlock)) {
      int fit = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&fit, kernel, threads, p) == hipSuccess && fit == perCu - 1)
        pad = p;
    }
  }
  pads.emplace(std::make_pair(kernel, dev), pad);
  return pad;
}

template <typename T>
void launchFactorPlan(const DeviceFactorPlan& dp, const FactorArgs<T>& a, const double* values, hipStream_t main,
                      hipStream_t side, const std::vector<hipEvent_t>& sync, ProfileSink* prof) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, main, a.failCol, INT32_MAX);
  if (a.panelCtr) launchClearInts(a.panelCtr, dp.panelCounters, main);
  if (a.fallbacks) hipLaunchKernelGGL(k_set_int, dim3(1), dim3(1), 0, main, a.fallbacks, 0);
  if (prof) prof->used = 0;
  hipEvent_t last[2] = {nullptr, nullptr};  // most recent signal of each stream
  size_t ev = 0;
  // fork: the side stream starts after everything enqueued on `main` so far (and joins a capture of `main`)
  (void)hipEventRecord(sync.back(), main);
  (void)hipStreamWaitEvent(side, sync.back(), 0);
  for (const Launch& L : dp.launches) {
    const unsigned grid = unsigned(L.count);
    const int si = (L.sync & kSyncSide) ? 1 : 0;
    hipStream_t st = si ? side : main;
    if ((L.sync & kSyncWait) && last[1 - si]) (void)hipStreamWaitEvent(st, last[1 - si], 0);
    // panel splits serve the bf16 bulk products only (their wait, above, still orders the side stream)
    if (L.kind == kSplit && !(std::is_same<T, float>::value && a.bulkBf16)) continue;
    if (prof && 2 * prof->used + 1 < prof->events.size()) {
      (void)hipEventRecord(prof->events[2 * prof->used], st);
      prof->kinds[prof->used] = L.kind;
      prof->levels[prof->used] = L.level;
    }
    switch (L.kind) {
      case kZero:
        hipLaunchKernelGGL(k_zero<T>, dim3(grid), dim3(256), 0, st, dp.zeros + L.first, a.Lx);
        break;
      case kAssemble: {
        const unsigned g = unsigned(std::min<int64_t>((L.count + 255) / 256, 65536));
        hipLaunchKernelGGL(k_assemble<T>, dim3(g), dim3(256), 0, st, dp.adst, a.lowerSrc, values, a.Lx, L.count);
        break;
      }
      case kSmall: {
        if (L.threads == 0 && L.edge > 0) {  // 64 / edge fronts per wave
          const unsigned wg = unsigned((L.count + 4 * (64 / L.edge) - 1) / (4 * (64 / L.edge)));
          if (L.edge == 8)
            hipLaunchKernelGGL((k_small_front_seg<T, 8>), dim3(wg), dim3(256), 0, st, dp.smallFronts + L.first,
                               int32_t(L.count), a, values);
          else
            hipLaunchKernelGGL((k_small_front_seg<T, 16>), dim3(wg), dim3(256), 0, st, dp.smallFronts + L.first,
                               int32_t(L.count), a, values);
          break;
        }
        if (L.threads == 0) {  // one wave per front
          const unsigned wg = unsigned((L.count + 3) / 4);
          if (L.maxM <= 16)
            hipLaunchKernelGGL((k_small_front_wave<T, 16>), dim3(wg), dim3(256), 0, st, dp.smallFronts + L.first,
                               int32_t(L.count), a, values);
          else
            hipLaunchKernelGGL((k_small_front_wave<T, kWaveFrontMax>), dim3(wg), dim3(256), 0, st,
                               dp.smallFronts + L.first, int32_t(L.count), a, values);
          break;
        }
        const size_t lds = (size_t(L.maxM) * L.maxM + L.maxM) * sizeof(T);
        if (L.threads == 64)
          hipLaunchKernelGGL((k_small_front<T, 64>), dim3(grid), dim3(64), lds, st, dp.smallFronts + L.first, a, values,
                             L.maxM);
        else
          hipLaunchKernelGGL((k_small_front<T, 256>), dim3(grid), dim3(256), lds, st, dp.smallFronts + L.first, a,
                             values, L.maxM);
        break;
      }
      case kExtendAdd:
        hipLaunchKernelGGL(k_extend_add<T>, dim3(grid), dim3(256), 0, st, dp.ea + L.first, dp.eaRanges, a);
        break;
      case kDiag:
        launchDiagBlocks<T>(dp.diag + L.first, grid, a, st);
        break;
      case kPanel:
        launchPanelSteps<T>(dp.panels + L.first, grid, a, st);
        break;
      case kSplit:
        if constexpr (std::is_same<T, float>::value)
          launchSplitPanels(dp.splits + L.first, L.count, L.maxM, L.edge, a, st);
        break;
      case kGemm: {
        const GemmTask* gt = reinterpret_cast<const GemmTask*>(dp.gemm) + L.first;
        if constexpr (std::is_same<T, float>::value) {
          // bulk launches (side stream: modes 0 / 1 only) of a factor with bf16 bulk updates, on their split panel
          if (a.bulkBf16 && L.edge == kGemmEdge && (L.sync & kSyncSide)) {
            launchGemmBf16x3(gt, dp.srefs + L.first, grid, a, st);
            if (prof && 2 * prof->used + 1 < prof->events.size()) prof->kinds[prof->used] = kProfGemmBf16x3;
            break;
          }
        }
        // main-stream launches: chain priority; side stream: padding LDS that leaves room for the chain
        auto go = [&](auto kernel, int threads, size_t staticLds) {
          const size_t pad = si ? bulkLds(reinterpret_cast<const void*>(kernel), threads, staticLds) : 0;
          hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), pad, st, gt, a.Lx, a.U, a.W, a.D, 1 - si);
        };
        if (L.edge == kGemmEdge) {
          if (a.ldlt)
            go(k_gemm_tile<T, kGemmEdge, true>, 512, gemmTileLds<T, kGemmEdge>());
          else
            go(k_gemm_tile<T, kGemmEdge, false>, 512, gemmTileLds<T, kGemmEdge>());
        } else if (a.ldlt) {
          go(k_gemm_tile<T, kTileEdge, true>, 256, gemmTileLds<T, kTileEdge>());
        } else {
          go(k_gemm_tile<T, kTileEdge, false>, 256, gemmTileLds<T, kTileEdge>());
        }
        break;
      }
    }
    if (prof && 2 * prof->used + 1 < prof->events.size()) {
      (void)hipEventRecord(prof->events[2 * prof->used + 1], st);
      prof->used++;
    }
    if ((L.sync & kSyncSignal) && ev < sync.size()) {
      (void)hipEventRecord(sync[ev], st);
      last[si] = sync[ev++];
    }
  }
  if (last[1] && ev < sync.size()) {  // join the side stream into the main stream
    (void)hipEventRecord(sync[ev], side);
    (void)hipStreamWaitEvent(main, sync[ev], 0);
  }
}

template <typename T>
void launchLdltCounts(const FactorArgs<T>& a, int32_t n, int32_t* out, hipStream_t st) {
  hipLaunchKernelGGL(k_ldlt_counts<T>, dim3(1), dim3(256), 0, st, a.D, a.pert, a.failCol, n, out);
}

template void launchFactorPlan<double>(const DeviceFactorPlan&, const FactorArgs<double>&, const double*, hipStream_t,
                                       hipStream_t, const std::vector<hipEvent_t>&, ProfileSink*);
template void launchFactorPlan<float>(const DeviceFactorPlan&, const FactorArgs<float>&, const double*, hipStream_t,
                                      hipStream_t, const std::vector<hipEvent_t>&, ProfileSink*);
template void launchLdltCounts<double>(const FactorArgs<double>&, int32_t, int32_t*, hipStream_t);
template void launchLdltCounts<float>(const FactorArgs<float>&, int32_t, int32_t*, hipStream_t);

}  // namespace amp
