// Device supernodal triangular solves (forward L, optional D^-1, backward L^T) over the level schedule of the plan.
//
// Same mathematics as the reference solves (packages/core/src/la/sparse.mjs:1161-1305 for Cholesky, ldlt.mjs:565-643
// for LDL^T): per supernode an in-block triangular solve and the product with the below block; the scatter into
// ancestor rows is carried by each supernode's front vector and gathered by the parent through the relative map, so
// no two workgroups ever write the same entry (deterministic, no atomics).
#include "factor_kernels.h"

namespace amp {

namespace {

template <typename S>
__device__ inline S waveSum(S v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

}  // namespace

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

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_solve_small_fwd(const int32_t* __restrict__ fronts, SolveArgs<T> a, S* y,
                                                         S* fv, int32_t q, int64_t fvN) {
  __shared__ S v[kSmallMax];
  __shared__ S s_x;
  const int32_t s = fronts[blockIdx.x];
  const int32_t f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s];
  const T* L = a.Lx + a.Lxp[s];
  for (int32_t k = 0; k < q; k++) {
    S* yk = y + int64_t(k) * a.n;
    S* fk = fv + k * fvN;
    gatherFront<T, S>(a, s, yk, fk, v, nc, m);
    for (int col = 0; col < nc; col++) {
      if (threadIdx.x == 0) {
        const S xc = a.ldlt ? v[col] : v[col] / S(L[col + col * m]);
        v[col] = xc;
        s_x = xc;
      }
      __syncthreads();
      const S xc = s_x;
      for (int r = col + 1 + threadIdx.x; r < m; r += blockDim.x) v[r] -= S(L[r + col * m]) * xc;
      __syncthreads();
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      if (i < nc) yk[f + i] = v[i];
      fk[a.Sp[s] + i] = v[i];
    }
    __syncthreads();
  }
}

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_solve_gather(const int32_t* __restrict__ fronts, SolveArgs<T> a, S* y, S* fv,
                                                      int32_t q, int64_t fvN) {
  const int32_t s = fronts[blockIdx.x];
  const int32_t nc = a.superStart[s + 1] - a.superStart[s], m = a.Sp[s + 1] - a.Sp[s];
  for (int32_t k = 0; k < q; k++) gatherFront<T, S>(a, s, y + int64_t(k) * a.n, fv + k * fvN, fv + k * fvN + a.Sp[s], nc, m);
}

// The inverted 64 x 64 diagonal block of block `blk` of supernode s into LDS (row i, column j at Wsh[i + j * 65]).
template <typename T, typename S>
__device__ inline void loadInverse(const SolveArgs<T>& a, int32_t s, int32_t blk, S* Wsh) {
  const T* W = a.W + a.Woff[s] + int64_t(blk) * kTileEdge * kTileEdge;
  for (int idx = threadIdx.x; idx < kTileEdge * kTileEdge; idx += blockDim.x)
    Wsh[(idx & 63) + (idx >> 6) * (kTileEdge + 1)] = S(W[idx]);
}

// Forward step of one 64-column block of a tiled supernode: x = Winv v_blk (every work item recomputes it from the
// inverse staged in LDS), then v[r] -= L[r, blk] x for the item's rows, four threads per row over 16 columns each;
// the writer item stores x into y.
template <typename T, typename S>
__global__ __launch_bounds__(256) void k_solve_fwd_block(const BlockTask* __restrict__ tasks, SolveArgs<T> a, S* y,
                                                         S* fv, int32_t q, int64_t fvN) {
  constexpr int E = kTileEdge, P = kTileEdge + 1;
  __shared__ S Wsh[E * P];
  __shared__ S vb[E], xb[E];
  __shared__ S part[4][kSolveRows];
  const BlockTask t = tasks[blockIdx.x];
  const int32_t s = t.s, f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s];
  const int32_t c0 = t.blk * E, nb = min(E, nc - c0);
  const T* L = a.Lx + a.Lxp[s];
  const int tid = threadIdx.x, rr = tid & 63, g = tid >> 6;
  loadInverse<T, S>(a, s, t.blk, Wsh);
  for (int32_t k = 0; k < q; k++) {
    S* fvs = fv + k * fvN + a.Sp[s];
    if (tid < E) vb[tid] = tid < nb ? fvs[c0 + tid] : S(0);
    __syncthreads();
    if (tid < E) {
      S sum = 0;
#pragma unroll 16
      for (int kk = 0; kk < E; kk++) sum += Wsh[tid + kk * P] * vb[kk];
      xb[tid] = sum;
    }
    __syncthreads();
    for (int32_t r0 = t.r0; r0 < t.r1; r0 += 64) {
      const int32_t r = r0 + rr;
      S sum = 0;
      if (r < t.r1) {
        const T* Lr = L + r + int64_t(c0 + 16 * g) * m;
#pragma unroll
        for (int j = 0; j < 16; j++) sum += S(Lr[int64_t(j) * m]) * xb[16 * g + j];
      }
      part[g][rr] = sum;
      __syncthreads();
      if (g == 0 && r < t.r1) fvs[r] -= (part[0][rr] + part[1][rr]) + (part[2][rr] + part[3][rr]);
      __syncthreads();
    }
    if (t.writer && tid < nb) y[int64_t(k) * a.n + f + c0 + tid] = xb[tid];
    __syncthreads();
  }
}

template <typename T, typename S>
__global__ __launch_bounds__(256) void k_solve_small_bwd(const int32_t* __restrict__ fronts, SolveArgs<T> a, S* y,
                                                         int32_t q) {
  __shared__ S z[kSmallMax], xb[kSmallMax];
  __shared__ S s_x;
  const int32_t s = fronts[blockIdx.x];
  const int32_t f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s], mu = m - nc;
  const T* L = a.Lx + a.Lxp[s];
  const int32_t* rows = a.Si + a.Sp[s] + nc;
  for (int32_t k = 0; k < q; k++) {
    S* yk = y + int64_t(k) * a.n;
    for (int r = threadIdx.x; r < mu; r += blockDim.x) xb[r] = yk[rows[r]];
    __syncthreads();
    for (int i = threadIdx.x; i < nc; i += blockDim.x) {
      S sum = yk[f + i];
      for (int r = 0; r < mu; r++) sum -= S(L[nc + r + i * m]) * xb[r];
      z[i] = sum;
    }
    __syncthreads();
    for (int col = nc - 1; col >= 0; col--) {
      if (threadIdx.x == 0) {
        const S xc = a.ldlt ? z[col] : z[col] / S(L[col + col * m]);
        z[col] = xc;
        s_x = xc;
      }
      __syncthreads();
      const S xc = s_x;
      for (int i = threadIdx.x; i < col; i += blockDim.x) z[i] -= S(L[col + i * m]) * xc;
      __syncthreads();
    }
    for (int i = threadIdx.x; i < nc; i += blockDim.x) yk[f + i] = z[i];
    __syncthreads();
  }
}

// z_i = y_own[i] - L21[:, i] . y[below] for the item's (at most 16) columns: y[below] is staged through LDS in chunks,
// each wave owns four columns and its lanes run over the rows.
template <typename T, typename S>
__global__ __launch_bounds__(256) void k_solve_bwd_init(const BlockTask* __restrict__ tasks, SolveArgs<T> a, S* y,
                                                        S* fv, int32_t q, int64_t fvN) {
  constexpr int CH = 2048;
  __shared__ S yb[CH];
  const BlockTask t = tasks[blockIdx.x];
  const int32_t s = t.s, f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s], mu = m - nc;
  const T* L = a.Lx + a.Lxp[s];
  const int32_t* rows = a.Si + a.Sp[s] + nc;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int32_t i0 = t.r0 + 4 * wave;
  for (int32_t k = 0; k < q; k++) {
    const S* yk = y + int64_t(k) * a.n;
    S acc[4] = {0, 0, 0, 0};
    for (int32_t base = 0; base < mu; base += CH) {
      const int32_t len = min(CH, mu - base);
      __syncthreads();
      for (int r = threadIdx.x; r < len; r += blockDim.x) yb[r] = yk[rows[base + r]];
      __syncthreads();
#pragma unroll
      for (int c = 0; c < 4; c++) {
        if (i0 + c >= t.r1) continue;
        const T* Li = L + nc + base + int64_t(i0 + c) * m;
        S sum = 0;
        for (int r = lane; r < len; r += 64) sum += S(Li[r]) * yb[r];
        acc[c] += sum;
      }
    }
    S* fvs = fv + k * fvN + a.Sp[s];
#pragma unroll
    for (int c = 0; c < 4; c++) {
      const S v = waveSum(acc[c]);
      if (lane == 0 && i0 + c < t.r1) fvs[i0 + c] = yk[f + i0 + c] - v;
    }
  }
}

// Backward step of one block: x = Winv^T z_blk (every item recomputes it from the inverse staged in LDS), then
// z_i -= L[blk, i] . x for the item's columns i (one wave per column); the writer item stores x into y.
template <typename T, typename S>
__global__ __launch_bounds__(256) void k_solve_bwd_block(const BlockTask* __restrict__ tasks, SolveArgs<T> a, S* y,
                                                         S* fv, int32_t q, int64_t fvN) {
  constexpr int E = kTileEdge, P = kTileEdge + 1;
  __shared__ S Wsh[E * P];
  __shared__ S zb[E], xb[E];
  const BlockTask t = tasks[blockIdx.x];
  const int32_t s = t.s, f = a.superStart[s], nc = a.superStart[s + 1] - f, m = a.Sp[s + 1] - a.Sp[s];
  const int32_t c0 = t.blk * E, nb = min(E, nc - c0);
  const T* L = a.Lx + a.Lxp[s];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  loadInverse<T, S>(a, s, t.blk, Wsh);
  for (int32_t k = 0; k < q; k++) {
    S* fvs = fv + k * fvN + a.Sp[s];
    if (tid < E) zb[tid] = tid < nb ? fvs[c0 + tid] : S(0);
    __syncthreads();
    if (tid < E) {
      S sum = 0;
#pragma unroll 16
      for (int kk = 0; kk < E; kk++) sum += Wsh[kk + tid * P] * zb[kk];
      xb[tid] = sum;
    }
    __syncthreads();
    for (int32_t i = t.r0 + wave; i < t.r1; i += 4) {
      S v = lane < nb ? S(L[(c0 + lane) + int64_t(i) * m]) * xb[lane] : S(0);
      v = waveSum(v);
      if (lane == 0) fvs[i] -= v;
    }
    if (t.writer && tid < nb) y[int64_t(k) * a.n + f + c0 + tid] = xb[tid];
    __syncthreads();
  }
}

template <typename T, typename S>
void launchSolvePlan(const DeviceSolvePlan& dp, const SolveArgs<T>& a, const double* b, double* x, int32_t q, S* y,
                     S* fv, hipStream_t st, ProfileSink* prof) {
  const int32_t n = a.n;
  if (n == 0 || q == 0) return;
  const unsigned g = unsigned(std::min<int64_t>((int64_t(n) * q + 255) / 256, 65536));
  hipLaunchKernelGGL(k_perm_in<S>, dim3(g), dim3(256), 0, st, b, y, a.perm, n, q);
  bool scaled = false;
  for (const SolveLaunch& L : dp.launches) {
    if (a.ldlt && !scaled && (L.kind == kSolveSmallBwd || L.kind == kSolveBwdInit)) {
      hipLaunchKernelGGL((k_diag_scale<T, S>), dim3(g), dim3(256), 0, st, y, a.D, n, q);
      scaled = true;
    }
    const unsigned grid = unsigned(L.count);
    const bool rec = prof && 2 * prof->used + 1 < prof->events.size();
    if (rec) {
      (void)hipEventRecord(prof->events[2 * prof->used], st);
      prof->kinds[prof->used] = 16 + L.kind;
    }
    switch (L.kind) {
      case kSolveSmallFwd:
        hipLaunchKernelGGL((k_solve_small_fwd<T, S>), dim3(grid), dim3(L.maxM <= 64 ? 64 : 128), 0, st,
                           dp.fronts + L.first, a, y, fv, q, dp.fvEntries);
        break;
      case kSolveGather:
        hipLaunchKernelGGL((k_solve_gather<T, S>), dim3(grid), dim3(256), 0, st, dp.fronts + L.first, a, y, fv, q,
                           dp.fvEntries);
        break;
      case kSolveFwdBlock:
        hipLaunchKernelGGL((k_solve_fwd_block<T, S>), dim3(grid), dim3(256), 0, st, dp.blocks + L.first, a, y, fv, q,
                           dp.fvEntries);
        break;
      case kSolveSmallBwd:
        hipLaunchKernelGGL((k_solve_small_bwd<T, S>), dim3(grid), dim3(L.maxM <= 64 ? 64 : 128), 0, st,
                           dp.fronts + L.first, a, y, q);
        break;
      case kSolveBwdInit:
        hipLaunchKernelGGL((k_solve_bwd_init<T, S>), dim3(grid), dim3(256), 0, st, dp.blocks + L.first, a, y, fv, q,
                           dp.fvEntries);
        break;
      case kSolveBwdBlock:
        hipLaunchKernelGGL((k_solve_bwd_block<T, S>), dim3(grid), dim3(256), 0, st, dp.blocks + L.first, a, y, fv, q,
                           dp.fvEntries);
        break;
    }
    if (rec) {
      (void)hipEventRecord(prof->events[2 * prof->used + 1], st);
      prof->used++;
    }
  }
  if (a.ldlt && !scaled) hipLaunchKernelGGL((k_diag_scale<T, S>), dim3(g), dim3(256), 0, st, y, a.D, n, q);
  hipLaunchKernelGGL(k_perm_out<S>, dim3(g), dim3(256), 0, st, y, x, a.perm, n, q);
}

template void launchSolvePlan<double, double>(const DeviceSolvePlan&, const SolveArgs<double>&, const double*, double*,
                                              int32_t, double*, double*, hipStream_t, ProfileSink*);
template void launchSolvePlan<float, float>(const DeviceSolvePlan&, const SolveArgs<float>&, const double*, double*,
                                            int32_t, float*, float*, hipStream_t, ProfileSink*);
template void launchSolvePlan<float, double>(const DeviceSolvePlan&, const SolveArgs<float>&, const double*, double*,
                                             int32_t, double*, double*, hipStream_t, ProfileSink*);

}  // namespace amp
