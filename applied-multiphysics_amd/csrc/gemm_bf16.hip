// Bulk trailing-update products of the FP32 factor on the bf16 matrix cores (gfx950, v_mfma_f32_16x16x32_bf16).
//
// Each FP32 operand entry x is split into a bf16 head h = bf16(x) and tail l = bf16(x - h) (x - h is exact in FP32),
// and the product a b is accumulated in FP32 as h_a h_b + h_a l_b + l_a h_b: three bf16 products per entry pair, the
// dropped l_a l_b and the tails' own rounding leave a relative error of about 2^-16 per product (FP32: 2^-24), at
// 16 / 3 of the FP32 matrix-core rate. The factor's panel chain (diagonal blocks, panel solves, updates inside a
// super-panel) stays FP32; only the launches of the side stream (every later column and the Schur complement from a
// factored super-panel) take this kernel, and FP64 iterative refinement restores the solution's accuracy.
//
// Every operand row of a bulk launch is a row of its super-panel, which tens of tiles read. The panel is therefore
// split once, by k_split_panel in front of the bulk launch: heads and tails in chunks of 32 depths, each chunk holding
// its rows one after another, 32 depths (64 bytes) per row: the layout the matrix cores take an operand fragment in
// (8 consecutive depths of one row per lane), zero-padded to the chunk; for LDL^T the B operand is split from the
// pivot-scaled panel. A tile's chunk is then 8 KB of contiguous heads (and tails), which the product kernel loads
// 16 bytes per lane, 1 KB per instruction, with no conversion.
//
// Tiles of 128 x 128 (modes 0: C -= A B^T and 1: C = -A B^T of plan.h); eight waves in 4 x 2, each 32 rows x 64
// columns of 16 x 16 blocks, the matrix cores fed the transposed product so that C is read and written along
// columns. Depth chunks of 32: a thread loads the heads and tails of 8 consecutive depths of one row
// of A and of B and writes them to LDS row-major (16-byte groups XOR-swizzled by row so that the 16 rows a
// quarter-wave reads fall in distinct banks). Two LDS buffers, the chunk after next in registers, one barrier per
// chunk.
#include <algorithm>
#include <type_traits>
#include <utility>

#include "factor_kernels.h"

namespace amp {

namespace {

typedef __bf16 bf16x8 __attribute__((__vector_size__(16)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kG = 128, kKC = kSplitDepth;
constexpr int kFullDepth = kSuperPanel;  // depth of the bulk products of a whole super-panel
constexpr int kRowBytes = kKC * 2;  // bf16 depth row of a tile row in LDS (64 bytes, four 16-byte groups)
static_assert(kKC == 32, "four 16-byte groups of depths per tile row");

// byte offset of 16-byte group g (depths 8 g .. 8 g + 7) of tile row r
__device__ inline int ldsOff(int r, int g) { return r * kRowBytes + 16 * (g ^ ((r >> 2) & 3)); }

// f(integral_constant<int, i>) for i = 0, 1, ... in order
template <typename F, int... I>
__device__ inline void forEachIndex(std::integer_sequence<int, I...>, F& f) {
  (f(std::integral_constant<int, I>{}), ...);
}

__device__ inline void split8(const float* x, bf16x8& h, bf16x8& l) {
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const __bf16 hj = (__bf16)x[j];
    h[j] = hj;
    l[j] = (__bf16)(x[j] - (float)hj);
  }
}

}  // namespace

// Panel split (SplitTask of plan.h): work items of 16 rows x one chunk (4 groups of 8 depths) per wave, lanes 0-15 on
// consecutive rows, so that every FP32 load instruction reads four 64-byte row runs (rows are contiguous in the
// column-major panel) and the wave's 16-byte stores cover 1 KB of the chunk contiguously. Depths at or past the panel
// width are written as zeros.
template <bool LDLT>
__global__ __launch_bounds__(256) void k_split_panel(const SplitTask* __restrict__ tasks, const float* __restrict__ Lx,
                                                     const float* __restrict__ D, __bf16* Ah, __bf16* Al, __bf16* Bh,
                                                     __bf16* Bl) {
  const SplitTask t = tasks[blockIdx.y];
  const int lane = threadIdx.x & 63;
  const int32_t ng4 = t.dp / 32;                  // quads of depth groups per row
  const int64_t items = int64_t((t.rows + 15) / 16) * ng4;  // (16-row block, group quad) per wave
  const int64_t waves = int64_t(gridDim.x) * 4;
  for (int64_t it = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6); it < items; it += waves) {
    const int32_t r = int32_t(it / ng4) * 16 + (lane & 15);
    const int32_t g = int32_t(it % ng4) * 4 + (lane >> 4);
    if (r >= t.rows) continue;
    float x[8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int32_t d = 8 * g + j;
      x[j] = d < t.w ? Lx[t.src + r + int64_t(d) * t.m] : 0.0f;
    }
    const int64_t o = t.dst + int64_t(g >> 2) * t.rows * kKC + int64_t(r) * kKC + 8 * (g & 3);
    bf16x8 h, l;
    split8(x, h, l);
    *reinterpret_cast<bf16x8*>(Ah + o) = h;
    *reinterpret_cast<bf16x8*>(Al + o) = l;
    if constexpr (LDLT) {
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int32_t d = 8 * g + j;
        x[j] = d < t.w ? x[j] * D[t.d_off + d] : 0.0f;
      }
      split8(x, h, l);
      *reinterpret_cast<bf16x8*>(Bh + o) = h;
      *reinterpret_cast<bf16x8*>(Bl + o) = l;
    }
  }
}

__global__ __launch_bounds__(512, 4) void k_gemm_bf16x3(const GemmTask* __restrict__ tasks,
                                                        const SplitRef* __restrict__ refs, float* Lx, float* U,
                                                        const __bf16* __restrict__ Ah, const __bf16* __restrict__ Al,
                                                        const __bf16* __restrict__ Bh, const __bf16* __restrict__ Bl) {
  constexpr int WTR = 32, WTC = 64, NBI = WTR / 16, NBJ = WTC / 16;
  __shared__ __attribute__((aligned(16))) unsigned char lds[2][4][kG * kRowBytes];  // [buf][A head, A tail, B h, B t]
  const int nblk = gridDim.x, q8 = nblk / 8, r8 = nblk % 8, x8 = blockIdx.x % 8;
  const int ti = x8 * q8 + min(x8, r8) + blockIdx.x / 8;
  const GemmTask t = tasks[ti];
  const SplitRef sr = refs[ti];
  float* C = (t.c_buf == 0 ? Lx : U) + t.c_off;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = (wave & 3) * WTR, wc = (wave >> 2) * WTC;
  const bool active = wr < t.rows && wc < t.cols;  // uniform per wave
  // loader: tile row li, depths k0 + 8 lg .. + 7 (four lanes per row: each instruction reads 16 rows' 64 contiguous
  // bytes); rows past the tile's edge read its last row and are zeroed
  const int li = tid >> 2, lg = tid & 3;
  const bool rowA = li < t.rows, rowB = li < t.cols;
  const int64_t oa = sr.a + int64_t(min(li, t.rows - 1)) * kKC + 8 * lg;
  const int64_t ob = sr.b + int64_t(min(li, t.cols - 1)) * kKC + 8 * lg;
  const int64_t cp = sr.ld;  // chunk pitch
  f4 acc[NBI][NBJ];
#pragma unroll
  for (int bi = 0; bi < NBI; bi++)
#pragma unroll
    for (int bj = 0; bj < NBJ; bj++) acc[bi][bj] = f4{0, 0, 0, 0};
  bf16x8 ra[3][2], rb[3][2];  // [register set][head, tail]
  auto load = [&](auto set, int k0) {
    constexpr int S = decltype(set)::value;
    const int64_t c = int64_t(k0 / kKC) * cp;
    ra[S][0] = *reinterpret_cast<const bf16x8*>(Ah + oa + c);
    ra[S][1] = *reinterpret_cast<const bf16x8*>(Al + oa + c);
    rb[S][0] = *reinterpret_cast<const bf16x8*>(Bh + ob + c);
    rb[S][1] = *reinterpret_cast<const bf16x8*>(Bl + ob + c);
  };
  auto stash = [&](auto set, int buf) {
    constexpr int S = decltype(set)::value;
    const bf16x8 z = {};
    const int off = ldsOff(li, lg);
    *reinterpret_cast<bf16x8*>(&lds[buf][0][off]) = rowA ? ra[S][0] : z;
    *reinterpret_cast<bf16x8*>(&lds[buf][1][off]) = rowA ? ra[S][1] : z;
    *reinterpret_cast<bf16x8*>(&lds[buf][2][off]) = rowB ? rb[S][0] : z;
    *reinterpret_cast<bf16x8*>(&lds[buf][3][off]) = rowB ? rb[S][1] : z;
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
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bl, ah[bi], acc[bi][bj], 0, 0, 0);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, al[bi], acc[bi][bj], 0, 0, 0);
        acc[bi][bj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(bh, ah[bi], acc[bi][bj], 0, 0, 0);
      }
    }
  };
  using S0 = std::integral_constant<int, 0>;
  using S1 = std::integral_constant<int, 1>;
  // the matrix cores are fed the transposed product (B's fragment first), so component q of acc[bi][bj] is C row
  // wr + 16 bi + lane % 16, column wc + 16 bj + 4 (lane / 16) + q: the 16 lanes of a row group cover 64 consecutive
  // bytes of a C column, and one load or store instruction four such runs
  const int row0 = wr + (lane & 15), colq = wc + 4 * (lane >> 4);
  if (t.k == kFullDepth) {
    // a full super-panel (the common case): its chunks at compile time, each loaded three chunks ahead (three register
    // sets), and the C entries of mode 0 read while the last chunk runs (unconditional, at clamped addresses, so that
    // the chunk loads before them are waited for alone; a step earlier the kernel would spill a register)
    constexpr int NC = kFullDepth / kKC;
    static_assert(NBI == 2, "two row blocks per wave");
    float cv[NBJ][4][NBI];
    load(S0{}, 0);
    load(S1{}, kKC);
    load(std::integral_constant<int, 2>{}, 2 * kKC);
    stash(S0{}, 0);
    __syncthreads();
    auto fixedStep = [&](auto cc) {
      constexpr int c = decltype(cc)::value;
      if constexpr (c + 3 < NC) load(std::integral_constant<int, c % 3>{}, (c + 3) * kKC);
      if constexpr (c == NC - 1) {
        const float* Cr = C + min(row0, t.rows - 1);
        const int d16 = min(row0 + 16, t.rows - 1) - min(row0, t.rows - 1);
#pragma unroll
        for (int bj = 0; bj < NBJ; bj++)
#pragma unroll
          for (int q = 0; q < 4; q++) {
            const float* Cc = Cr + int64_t(min(colq + 16 * bj + q, t.cols - 1)) * t.ldc;
            cv[bj][q][0] = Cc[0];
            cv[bj][q][1] = Cc[d16];
          }
      }
      compute(c & 1);
      if constexpr (c + 1 < NC) {
        stash(std::integral_constant<int, (c + 1) % 3>{}, (c + 1) & 1);
        __syncthreads();
      }
    };
    forEachIndex(std::make_integer_sequence<int, NC>{}, fixedStep);
    if (!active) return;
#pragma unroll
    for (int bj = 0; bj < NBJ; bj++)
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int col = colq + 16 * bj + q;
        if (col >= t.cols) continue;
        float* Cc = C + int64_t(col) * t.ldc;
#pragma unroll
        for (int bi = 0; bi < NBI; bi++) {
          const int row = row0 + 16 * bi;
          if (row < t.rows) Cc[row] = (t.mode == 0 ? cv[bj][q][bi] : 0.0f) - acc[bi][bj][q];
        }
      }
    return;
  }
  // the split rows hold whole depth chunks (zero past the panel width): every chunk below t.k is in bounds
  load(S0{}, 0);
  if (kKC < t.k) load(S1{}, kKC);
  stash(S0{}, 0);
  __syncthreads();
  // chunk c is read from buffer c & 1 while chunk c + 1 (in registers) is written to the other buffer and chunk c + 2
  // is loaded into the register set chunk c left; `set` is the register set of chunk c + 1
  auto step = [&](auto set, int k0, auto ld) {
    constexpr int S = decltype(set)::value;
    int buf = S ^ 1;
    asm volatile("" : "+s"(buf));
    if constexpr (decltype(ld)::value) load(std::integral_constant<int, S ^ 1>{}, k0 + 2 * kKC);
    compute(buf);
    if (k0 + kKC < t.k) stash(set, buf ^ 1);
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
  // mode 0 reads a 16-column block's C entries (addresses clamped into the tile), then writes them
#pragma unroll
  for (int bj = 0; bj < NBJ; bj++) {
    float cv[4][NBI];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const float* Cc = C + int64_t(min(colq + 16 * bj + q, t.cols - 1)) * t.ldc;
#pragma unroll
      for (int bi = 0; bi < NBI; bi++) cv[q][bi] = t.mode == 0 ? Cc[min(row0 + 16 * bi, t.rows - 1)] : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int col = colq + 16 * bj + q;
      if (col >= t.cols) continue;
      float* Cc = C + int64_t(col) * t.ldc;
#pragma unroll
      for (int bi = 0; bi < NBI; bi++) {
        const int row = row0 + 16 * bi;
        if (row < t.rows) Cc[row] = cv[q][bi] - acc[bi][bj][q];
      }
    }
  }
}

void launchSplitPanels(const SplitTask* splits, int64_t count, int32_t maxRows, int32_t maxPitch,
                       const FactorArgs<float>& a, hipStream_t st) {
  if (count <= 0) return;
  const int64_t p = a.splitPitch;
  __bf16 *Ah = a.split, *Al = a.split + p, *Bh = a.split + 2 * p, *Bl = a.split + 3 * p;
  // workgroups of four waves per task, enough for every (16-row block, group quad) of the largest task
  const int64_t items = int64_t((maxRows + 15) / 16) * std::max(1, maxPitch / 32);
  const unsigned gx = unsigned(std::min<int64_t>((items + 3) / 4, 1024));
  if (a.ldlt)
    hipLaunchKernelGGL(k_split_panel<true>, dim3(gx, unsigned(count)), dim3(256), 0, st, splits, a.Lx, a.D, Ah, Al,
                       Bh, Bl);
  else
    hipLaunchKernelGGL(k_split_panel<false>, dim3(gx, unsigned(count)), dim3(256), 0, st, splits, a.Lx, a.D, Ah, Al,
                       Bh, Bl);
}

void launchGemmBf16x3(const GemmTask* tasks, const SplitRef* refs, unsigned grid, const FactorArgs<float>& a,
                      hipStream_t st) {
  const int64_t p = a.splitPitch;
  const __bf16 *Ah = a.split, *Al = a.split + p;
  const __bf16* Bh = a.ldlt ? a.split + 2 * p : Ah;  // Cholesky: B is split from the same (unscaled) panel
  const __bf16* Bl = a.ldlt ? a.split + 3 * p : Al;
  hipLaunchKernelGGL(k_gemm_bf16x3, dim3(grid), dim3(512), 0, st, tasks, refs, a.Lx, a.U, Ah, Al, Bh, Bl);
}

}  // namespace amp
