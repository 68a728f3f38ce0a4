// Bulk trailing-update products of the FP32 factor on the bf16 matrix cores (gfx950, v_mfma_f32_16x16x32_bf16).
//
// Each FP32 operand entry x is split into a bf16 head h = bf16(x) and tail l = bf16(x - h) (x - h is exact in FP32),
// and the product a b is accumulated in FP32 as h_a h_b + h_a l_b + l_a h_b: three bf16 products per entry pair, the
// dropped l_a l_b and the tails' own rounding leave a relative error of about 2^-16 per product (FP32: 2^-24), at
// 16 / 3 of the FP32 matrix-core rate. The factor's panel chain (diagonal blocks, panel solves, updates inside a
// super-panel) stays FP32; only the launches of the side stream (every later column and the Schur complement from a
// factored super-panel) take this kernel, and FP64 iterative refinement restores the solution's accuracy.
//
// Launched with padding LDS (bulkLds) so that each CU keeps room for the panel chain's workgroups.
// Tiles of 128 x 128 (modes 0: C -= A B^T and 1: C = -A B^T of plan.h, LDL^T: B scaled by the pivots of its depths
// before the split); eight waves in 4 x 2, each 32 rows x 64 columns of 16 x 16 blocks. Depth chunks of 32: a thread
// loads 8 consecutive depths of one row of A and of B (FP32, coalesced along rows), splits them and writes the heads
// and tails to LDS row-major (depth contiguous, 16-byte groups XOR-swizzled by row so that the 16 rows a quarter-wave
// reads fall in distinct banks). Two LDS buffers, the chunk after next in registers, one barrier per chunk.
#include <type_traits>

#include "factor_kernels.h"

namespace amp {

namespace {

typedef __bf16 bf16x8 __attribute__((__vector_size__(16)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kG = 128, kKC = 32;
constexpr int kRowBytes = kKC * 2;  // bf16 depth row of a tile row in LDS (64 bytes, four 16-byte groups)

// byte offset of 16-byte group g (depths 8 g .. 8 g + 7) of tile row r
__device__ inline int ldsOff(int r, int g) { return r * kRowBytes + 16 * (g ^ ((r >> 2) & 3)); }

__device__ inline void split8(const float* x, bf16x8& h, bf16x8& l) {
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const __bf16 hj = (__bf16)x[j];
    h[j] = hj;
    l[j] = (__bf16)(x[j] - (float)hj);
  }
}

}  // namespace

template <bool LDLT>
__global__ __launch_bounds__(512, 2) void k_gemm_bf16x3(const GemmTask* __restrict__ tasks, float* Lx, float* U,
                                                        float* W, const float* __restrict__ D) {
  constexpr int WTR = 32, WTC = 64, NBI = WTR / 16, NBJ = WTC / 16;
  __shared__ __attribute__((aligned(16))) unsigned char lds[2][4][kG * kRowBytes];  // [buf][A head, A tail, B h, B t]
  const int nblk = gridDim.x, q8 = nblk / 8, r8 = nblk % 8, x8 = blockIdx.x % 8;
  const GemmTask t = tasks[x8 * q8 + min(x8, r8) + blockIdx.x / 8];
  auto base = [&](int32_t b) { return b == 0 ? Lx : b == 1 ? U : W; };
  const float* A = base(t.a_buf) + t.a_off;
  const float* B = base(t.b_buf) + t.b_off;
  float* C = base(t.c_buf) + t.c_off;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave & 3) * WTR, wc = (wave >> 2) * WTC;
  const bool active = wr < t.rows && wc < t.cols;  // uniform per wave
  // loader: tile row li, depths k0 + 8 lg .. + 7
  const int li = tid & (kG - 1), lg = tid >> 7;
  const bool rowA = li < t.rows, rowB = li < t.cols;
  const float* Ar = A + min(li, t.rows - 1);
  const float* Br = B + min(li, t.cols - 1);
  f4 acc[NBI][NBJ];
#pragma unroll
  for (int bi = 0; bi < NBI; bi++)
#pragma unroll
    for (int bj = 0; bj < NBJ; bj++) acc[bi][bj] = f4{0, 0, 0, 0};
  float pa[2][8], pb[2][8];
  float pd[LDLT ? 2 : 1][LDLT ? 8 : 1];
  auto load = [&](auto set, int k0) {
    constexpr int S = decltype(set)::value;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int k = min(k0 + 8 * lg + j, t.k - 1);  // clamped: every load is issued, out-of-range depths zeroed later
      pa[S][j] = Ar[int64_t(k) * t.lda];
      pb[S][j] = Br[int64_t(k) * t.ldb];
      if constexpr (LDLT) pd[S][j] = D[t.d_off + k];
    }
  };
  auto stash = [&](auto set, int buf, int k0) {
    constexpr int S = decltype(set)::value;
    float xa[8], xb[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const bool kin = k0 + 8 * lg + j < t.k;
      float bv = pb[S][j];
      if constexpr (LDLT) bv *= pd[S][j];
      xa[j] = (kin && rowA) ? pa[S][j] : 0.0f;
      xb[j] = (kin && rowB) ? bv : 0.0f;
    }
    bf16x8 ah, al, bh, bl;
    split8(xa, ah, al);
    split8(xb, bh, bl);
    const int off = ldsOff(li, lg);
    *reinterpret_cast<bf16x8*>(&lds[buf][0][off]) = ah;
    *reinterpret_cast<bf16x8*>(&lds[buf][1][off]) = al;
    *reinterpret_cast<bf16x8*>(&lds[buf][2][off]) = bh;
    *reinterpret_cast<bf16x8*>(&lds[buf][3][off]) = bl;
  };
  auto compute = [&](int buf) {
    if (!active) return;
    const int g = lane >> 4, r = lane & 15;
    bf16x8 ah[NBI], al[NBI];
#pragma unroll
    for (int bi = 0; bi < NBI; bi++) {
      const int off = ldsOff(wr + 16 * bi + r, g);
      ah[bi] = *reinterpret_cast<const bf16x8*>(&lds[buf][0][off]);
      al[bi] = *reinterpret_cast<const bf16x8*>(&lds[buf][1][off]);
    }
#pragma unroll
    for (int bj = 0; bj < NBJ; bj++) {
      const int off = ldsOff(wc + 16 * bj + r, g);
      const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&lds[buf][2][off]);
      const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&lds[buf][3][off]);
#pragma unroll
      for (int bi = 0; bi < NBI; bi++) {
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], bl, acc[bi][bj], 0, 0, 0);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[bi], bh, acc[bi][bj], 0, 0, 0);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[bi], bh, acc[bi][bj], 0, 0, 0);
      }
    }
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  load(S0{}, 0);
  load(S1{}, kKC);
  stash(S0{}, 0, 0);
  __syncthreads();
  // chunk c is read from buffer c & 1 while chunk c + 1 (in registers) is written to the other buffer and chunk c + 2
  // is loaded into the register set chunk c left; `set` is the register set of chunk c + 1
  auto step = [&](auto set, int k0, auto ld) {
    constexpr int S = decltype(set)::value;
    int buf = S ^ 1;
    asm volatile("" : "+s"(buf));
    if constexpr (decltype(ld)::value) load(std::integral_constant<int, S ^ 1>{}, k0 + 2 * kKC);
    compute(buf);
    if (k0 + kKC < t.k) stash(set, buf ^ 1, k0 + kKC);
    __syncthreads();
  };
  int k0 = 0;
  for (; k0 + 2 * kKC < t.k; k0 += 2 * kKC) {
    step(S1{}, k0, std::true_type{});
    step(S0{}, k0 + kKC, std::true_type{});
  }
  step(S1{}, k0, std::false_type{});
  if (k0 + kKC < t.k) step(S0{}, k0 + kKC, std::false_type{});
  if (!active) return;
  // component q of acc[bi][bj]: C row wr + 16 bi + 4 (lane / 16) + q, column wc + 16 bj + lane % 16
  const int rowq = wr + 4 * (lane >> 4), col0 = wc + (lane & 15);
#pragma unroll
  for (int bj = 0; bj < NBJ; bj++) {
    const int col = col0 + 16 * bj;
    if (col >= t.cols) continue;
    float* Cc = C + int64_t(col) * t.ldc;
    if (t.mode == 0) {
      float cv[NBI][4];
#pragma unroll
      for (int bi = 0; bi < NBI; bi++)
#pragma unroll
        for (int q = 0; q < 4; q++) cv[bi][q] = Cc[min(rowq + 16 * bi + q, t.rows - 1)];
#pragma unroll
      for (int bi = 0; bi < NBI; bi++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int row = rowq + 16 * bi + q;
          if (row < t.rows) Cc[row] = cv[bi][q] - acc[bi][bj][q];
        }
    } else {
#pragma unroll
      for (int bi = 0; bi < NBI; bi++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
          const int row = rowq + 16 * bi + q;
          if (row < t.rows) Cc[row] = -acc[bi][bj][q];
        }
    }
  }
}

void launchGemmBf16x3(const GemmTask* tasks, unsigned grid, float* Lx, float* U, float* W, const float* D, bool ldlt,
                      hipStream_t st) {
  constexpr size_t lds = sizeof(unsigned char[2][4][kG * kRowBytes]);
  if (ldlt)
    hipLaunchKernelGGL(k_gemm_bf16x3<true>, dim3(grid), dim3(512),
                       bulkLds(reinterpret_cast<const void*>(&k_gemm_bf16x3<true>), 512, lds), st, tasks, Lx, U, W, D);
  else
    hipLaunchKernelGGL(k_gemm_bf16x3<false>, dim3(grid), dim3(512),
                       bulkLds(reinterpret_cast<const void*>(&k_gemm_bf16x3<false>), 512, lds), st, tasks, Lx, U, W,
                       D);
}

}  // namespace amp
