// Warm-start network of the parametric sweep: a multilayer perceptron nIn -> nHidden -> nHidden -> nOut (tanh on the
// hidden layers, linear output), FP32, batch-major activations, trained on the device by these kernels (forward,
// backward, Adam). The products run on the matrix cores (FP32 16 x 16 x 4 MFMA, 64 x 64 output tiles); reductions
// (loss, bias gradients) run in a fixed order, so a training step is deterministic.
//
// Parameter layout (flat FP32): W1 (nHidden x nIn, row-major), b1, W2 (nHidden x nHidden), b2, W3 (nOut x nHidden), b3.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <memory>
#include <string>
#include <vector>

#include "handles.h"
#include "kernels_common.h"

using amp::DevBuf;
using amp::fail;

namespace amp {
namespace {

// Strided operand: element (r, c) at p[r * sr + c * sc].
struct Mat {
  const float* p;
  int64_t sr, sc;
};

enum Act : int32_t { kActNone = 0, kActTanh = 1 };

// C[M x N] (row-major, ldc) = A[M x K] B[K x N] (+ bias[j]) (then tanh), or C += when accumulate. Four waves of 32 x 32
// (2 x 2 MFMA blocks), depth staged through LDS in chunks of 16. With split > 1, slice z = blockIdx.z covers depths
// [z kPer, (z + 1) kPer) and stores its raw product to C + z M N (row-major, ldc = N): k_mlp_reduce finishes.
__global__ __launch_bounds__(256) void k_mlp_gemm(Mat A, Mat B, float* C, int64_t ldc, int32_t M, int32_t N, int32_t K,
                                                  int32_t kPer, const float* __restrict__ bias, int32_t act,
                                                  int32_t accumulate) {
  constexpr int T = 64, KC = 16, P = T + 4;
  __shared__ float As[KC * P], Bs[KC * P];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int i0 = blockIdx.y * T, j0 = blockIdx.x * T;
  const int wr = (wave & 1) * 32, wc = (wave >> 1) * 32;
  const bool split = gridDim.z > 1;
  const int32_t kb = blockIdx.z * kPer, ke = min(K, kb + kPer);
  v4f acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++) acc[a][b] = v4f{0, 0, 0, 0};
  for (int k0 = kb; k0 < ke; k0 += KC) {
    for (int idx = tid; idx < KC * T; idx += 256) {
      const int r = idx % T, k = idx / T;  // A: row i0 + r, depth k0 + k; B: column j0 + r
      const int kk = k0 + k;
      As[k * P + r] = (i0 + r < M && kk < ke) ? A.p[int64_t(i0 + r) * A.sr + int64_t(kk) * A.sc] : 0.f;
      Bs[k * P + r] = (j0 + r < N && kk < ke) ? B.p[int64_t(kk) * B.sr + int64_t(j0 + r) * B.sc] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int k4 = 0; k4 < KC; k4 += 4) {
      const int kk = (k4 + (lane >> 4)) * P + (lane & 15);
      float af[2], bf[2];
      for (int b = 0; b < 2; b++) {
        af[b] = As[kk + wr + 16 * b];
        bf[b] = Bs[kk + wc + 16 * b];
      }
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = mfma16x16x4(af[a], bf[b], acc[a][b]);
    }
    __syncthreads();
  }
  float* Cz = split ? C + int64_t(blockIdx.z) * M * N : C;
  const int64_t ld = split ? N : ldc;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int r = 0; r < 4; r++) {
        const int i = i0 + wr + 16 * a + mfmaRow<float>(lane, r), j = j0 + wc + 16 * b + (lane & 15);
        if (i >= M || j >= N) continue;
        float v = acc[a][b][r];
        float* dst = Cz + int64_t(i) * ld + j;
        if (!split) {
          if (accumulate) v += *dst;
          if (bias) v += bias[j];
          if (act == kActTanh) v = tanhf(v);
        }
        *dst = v;
      }
}

// C = sum over the slices (slice order) of S[z] (+ C when accumulate, + bias, then tanh)
__global__ void k_mlp_reduce(const float* __restrict__ S, int32_t slices, float* C, int64_t ldc, int32_t M, int32_t N,
                             const float* __restrict__ bias, int32_t act, int32_t accumulate) {
  const int64_t total = int64_t(M) * N;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total; t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = t / N, j = t % N;
    float v = 0.f;
    for (int z = 0; z < slices; z++) v += S[int64_t(z) * total + t];
    float* dst = C + i * ldc + j;
    if (accumulate) v += *dst;
    if (bias) v += bias[j];
    if (act == kActTanh) v = tanhf(v);
    *dst = v;
  }
}

// dZ = dA * (1 - A^2) (tanh backward, in place on dA)
__global__ void k_tanh_back(const float* __restrict__ A, float* __restrict__ dA, int64_t n) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x)
    dA[t] *= 1.f - A[t] * A[t];
}

// out[j] = sum_b X[b][j] over the batch (bias gradients), one thread per column, batch order
__global__ void k_colsum(const float* __restrict__ X, int32_t B, int32_t N, float* __restrict__ out) {
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int b = 0; b < B; b++) s += X[int64_t(b) * N + j];
    out[j] = s;
  }
}

// dY = 2 (Y - T) / (B N) in place on Y; partial[block] = sum of (Y - T)^2 over the block's strided entries
__global__ __launch_bounds__(256) void k_mse(float* __restrict__ Y, const float* __restrict__ Tg, int64_t n,
                                             float scale, double* __restrict__ partial) {
  __shared__ double red[256];
  double s = 0;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
    const float d = Y[t] - Tg[t];
    s += double(d) * d;
    Y[t] = scale * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Adam (bias-corrected) on every parameter
__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       int64_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n; t += int64_t(gridDim.x) * blockDim.x) {
    const float gt = g[t];
    const float mt = b1 * m[t] + (1.f - b1) * gt;
    const float vt = b2 * v[t] + (1.f - b2) * gt * gt;
    m[t] = mt;
    v[t] = vt;
    p[t] -= lr * (mt / c1) / (sqrtf(vt / c2) + eps);
  }
}

unsigned blocksFor(int64_t n) { return unsigned(std::min<int64_t>((n + 255) / 256, 4096)); }

constexpr int kMaxSplit = 64;

// Products with few output tiles and a deep K are split over depth slices (into `scratch`, kMaxSplit M N floats at
// most) and reduced in slice order, so the result does not depend on the schedule.
void gemm(Mat A, Mat B, float* C, int64_t ldc, int32_t M, int32_t N, int32_t K, const float* bias, int32_t act,
          int32_t accumulate, float* scratch, int64_t scratchFloats, hipStream_t st) {
  const int64_t tiles = int64_t((N + 63) / 64) * ((M + 63) / 64);
  int32_t split = 1;
  if (tiles < 128 && K >= 2048 && scratch) {
    split = int32_t(std::min<int64_t>({kMaxSplit, (256 + tiles - 1) / tiles, K / 512}));
    while (split > 1 && int64_t(split) * M * N > scratchFloats) split--;
  }
  const int32_t kPer = (K + split - 1) / split;
  const dim3 grid((N + 63) / 64, (M + 63) / 64, split);
  if (split == 1) {
    hipLaunchKernelGGL(k_mlp_gemm, grid, dim3(256), 0, st, A, B, C, ldc, M, N, K, K, bias, act, accumulate);
    return;
  }
  hipLaunchKernelGGL(k_mlp_gemm, grid, dim3(256), 0, st, A, B, scratch, int64_t(N), M, N, K, kPer, nullptr,
                     int32_t(kActNone), 0);
  hipLaunchKernelGGL(k_mlp_reduce, dim3(blocksFor(int64_t(M) * N)), dim3(256), 0, st, scratch, split, C, ldc, M, N,
                     bias, act, accumulate);
}

}  // namespace
}  // namespace amp

struct amp_mlp {
  int32_t nIn = 0, nH = 0, nOut = 0, maxBatch = 0;
  int64_t nParams = 0, step = 0;
  hipStream_t stream = nullptr;
  hipEvent_t inEv = nullptr;  // afterCaller
  DevBuf params, grads, m, v, h1, h2, y, partial;
  DevBuf dh;       // dz2, dz1 (2 x maxBatch x nHidden)
  DevBuf scratch;  // depth slices of split products
  ~amp_mlp() {
    if (inEv) (void)hipEventDestroy(inEv);
    if (stream) (void)hipStreamDestroy(stream);
  }
  // parameter offsets
  int64_t oW1() const { return 0; }
  int64_t ob1() const { return oW1() + int64_t(nH) * nIn; }
  int64_t oW2() const { return ob1() + nH; }
  int64_t ob2() const { return oW2() + int64_t(nH) * nH; }
  int64_t oW3() const { return ob2() + nH; }
  int64_t ob3() const { return oW3() + int64_t(nOut) * nH; }
};

namespace {

constexpr int kMseBlocks = 1024;

// activations of the batch: h1, h2 (B x nH), y (B x nOut)
void forward(amp_mlp* n, const float* x, int32_t B, float* y) {
  using amp::Mat;
  const float* P = n->params.as<float>();
  float* h1 = n->h1.as<float>();
  float* h2 = n->h2.as<float>();
  float* sc = n->scratch.as<float>();
  const int64_t scn = int64_t(n->scratch.bytes / 4);
  // h1 = tanh(x W1^T + b1): A = x (B x nIn), B(k, j) = W1[j][k]
  amp::gemm(Mat{x, n->nIn, 1}, Mat{P + n->oW1(), 1, n->nIn}, h1, n->nH, B, n->nH, n->nIn, P + n->ob1(), amp::kActTanh,
            0, sc, scn, n->stream);
  amp::gemm(Mat{h1, n->nH, 1}, Mat{P + n->oW2(), 1, n->nH}, h2, n->nH, B, n->nH, n->nH, P + n->ob2(), amp::kActTanh, 0,
            sc, scn, n->stream);
  amp::gemm(Mat{h2, n->nH, 1}, Mat{P + n->oW3(), 1, n->nH}, y, n->nOut, B, n->nOut, n->nH, P + n->ob3(), amp::kActNone,
            0, sc, scn, n->stream);
}

}  // namespace

extern "C" {

int amp_mlp_create(int32_t nIn, int32_t nHidden, int32_t nOut, int32_t maxBatch, amp_mlp** out) {
  if (!out || nIn < 1 || nHidden < 1 || nOut < 1 || maxBatch < 1) return fail(AMP_ERR_ARG, "amp_mlp_create: bad arguments");
  *out = nullptr;
  std::unique_ptr<amp_mlp> n(new amp_mlp());
  n->nIn = nIn;
  n->nH = nHidden;
  n->nOut = nOut;
  n->maxBatch = maxBatch;
  n->nParams = n->ob3() + nOut;
  HIP_TRY(hipStreamCreateWithFlags(&n->stream, hipStreamNonBlocking));
  HIP_TRY(hipEventCreateWithFlags(&n->inEv, hipEventDisableTiming));
  const size_t pb = size_t(n->nParams) * 4;
  HIP_TRY(n->params.alloc(pb));
  HIP_TRY(n->grads.alloc(pb));
  HIP_TRY(n->m.alloc(pb));
  HIP_TRY(n->v.alloc(pb));
  HIP_TRY(hipMemset(n->params.p, 0, pb));
  HIP_TRY(hipMemset(n->m.p, 0, pb));
  HIP_TRY(hipMemset(n->v.p, 0, pb));
  HIP_TRY(n->h1.alloc(size_t(maxBatch) * nHidden * 4));
  HIP_TRY(n->h2.alloc(size_t(maxBatch) * nHidden * 4));
  HIP_TRY(n->y.alloc(size_t(maxBatch) * nOut * 4));
  HIP_TRY(n->partial.alloc(kMseBlocks * 8));
  HIP_TRY(n->dh.alloc(size_t(2) * maxBatch * nHidden * 4));
  HIP_TRY(n->scratch.alloc(size_t(amp::kMaxSplit) * maxBatch * nHidden * 4));
  *out = n.release();
  return AMP_OK;
}

int64_t amp_mlp_param_count(const amp_mlp* n) { return n ? n->nParams : -1; }

int amp_mlp_set_params(amp_mlp* n, const float* host, int64_t count) {
  if (!n || !host || count != n->nParams) return fail(AMP_ERR_ARG, "amp_mlp_set_params: bad arguments");
  HIP_TRY(hipMemcpy(n->params.p, host, size_t(count) * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemset(n->m.p, 0, size_t(count) * 4));
  HIP_TRY(hipMemset(n->v.p, 0, size_t(count) * 4));
  n->step = 0;
  return AMP_OK;
}

int amp_mlp_get_params(amp_mlp* n, float* host, int64_t count) {
  if (!n || !host || count != n->nParams) return fail(AMP_ERR_ARG, "amp_mlp_get_params: bad arguments");
  HIP_TRY(hipStreamSynchronize(n->stream));
  HIP_TRY(hipMemcpy(host, n->params.p, size_t(count) * 4, hipMemcpyDeviceToHost));
  return AMP_OK;
}

int amp_mlp_get_grads(amp_mlp* n, float* host, int64_t count) {
  if (!n || !host || count != n->nParams) return fail(AMP_ERR_ARG, "amp_mlp_get_grads: bad arguments");
  HIP_TRY(hipStreamSynchronize(n->stream));
  HIP_TRY(hipMemcpy(host, n->grads.p, size_t(count) * 4, hipMemcpyDeviceToHost));
  return AMP_OK;
}

int amp_mlp_forward(amp_mlp* n, const float* d_x, int32_t batch, float* d_y) {
  if (!n || !d_x || !d_y || batch < 1 || batch > n->maxBatch) return fail(AMP_ERR_ARG, "amp_mlp_forward: bad arguments");
  HIP_TRY(amp::afterCaller(n->stream, n->inEv));
  forward(n, d_x, batch, d_y);
  HIP_TRY(hipGetLastError());
  HIP_TRY(hipStreamSynchronize(n->stream));
  return AMP_OK;
}

// Gradients of the mean squared error (1 / (B nOut)) sum (y - t)^2 at the batch (d_x: B x nIn, d_t: B x nOut, device)
// into the handle's gradient buffer; *loss = the mean squared error.
int amp_mlp_gradient(amp_mlp* n, const float* d_x, const float* d_t, int32_t B, double* loss) {
  using amp::Mat;
  if (!n || !d_x || !d_t || B < 1 || B > n->maxBatch) return fail(AMP_ERR_ARG, "amp_mlp_gradient: bad arguments");
  hipStream_t st = n->stream;
  HIP_TRY(amp::afterCaller(st, n->inEv));
  float* y = n->y.as<float>();
  forward(n, d_x, B, y);
  const int64_t ny = int64_t(B) * n->nOut;
  hipLaunchKernelGGL(amp::k_mse, dim3(kMseBlocks), dim3(256), 0, st, y, d_t, ny, float(2.0 / double(ny)),
                     n->partial.as<double>());
  const float* P = n->params.as<float>();
  float* G = n->grads.as<float>();
  float* h1 = n->h1.as<float>();
  float* h2 = n->h2.as<float>();
  float* dz2 = n->dh.as<float>();
  float* dz1 = dz2 + int64_t(n->maxBatch) * n->nH;
  float* sc = n->scratch.as<float>();
  const int64_t scn = int64_t(n->scratch.bytes / 4);
  const int32_t H = n->nH;
  // y now holds dY. Layer 3: dW3 = dY^T h2, db3 = colsum dY, dz2 = (dY W3) * (1 - h2^2)
  amp::gemm(Mat{y, 1, n->nOut}, Mat{h2, H, 1}, G + n->oW3(), H, n->nOut, H, B, nullptr, amp::kActNone, 0, sc, scn, st);
  hipLaunchKernelGGL(amp::k_colsum, dim3(amp::blocksFor(n->nOut)), dim3(256), 0, st, y, B, n->nOut, G + n->ob3());
  amp::gemm(Mat{y, n->nOut, 1}, Mat{P + n->oW3(), H, 1}, dz2, H, B, H, n->nOut, nullptr, amp::kActNone, 0, sc, scn, st);
  hipLaunchKernelGGL(amp::k_tanh_back, dim3(amp::blocksFor(int64_t(B) * H)), dim3(256), 0, st, h2, dz2, int64_t(B) * H);
  // layer 2: dW2 = dz2^T h1, db2 = colsum dz2, dz1 = (dz2 W2) * (1 - h1^2)
  amp::gemm(Mat{dz2, 1, H}, Mat{h1, H, 1}, G + n->oW2(), H, H, H, B, nullptr, amp::kActNone, 0, sc, scn, st);
  hipLaunchKernelGGL(amp::k_colsum, dim3(amp::blocksFor(H)), dim3(256), 0, st, dz2, B, H, G + n->ob2());
  amp::gemm(Mat{dz2, H, 1}, Mat{P + n->oW2(), H, 1}, dz1, H, B, H, H, nullptr, amp::kActNone, 0, sc, scn, st);
  hipLaunchKernelGGL(amp::k_tanh_back, dim3(amp::blocksFor(int64_t(B) * H)), dim3(256), 0, st, h1, dz1, int64_t(B) * H);
  // layer 1: dW1 = dz1^T x, db1 = colsum dz1
  amp::gemm(Mat{dz1, 1, H}, Mat{d_x, n->nIn, 1}, G + n->oW1(), n->nIn, H, n->nIn, B, nullptr, amp::kActNone, 0, sc, scn,
            st);
  hipLaunchKernelGGL(amp::k_colsum, dim3(amp::blocksFor(H)), dim3(256), 0, st, dz1, B, H, G + n->ob1());
  HIP_TRY(hipGetLastError());
  std::vector<double> part(kMseBlocks);
  HIP_TRY(hipMemcpyAsync(part.data(), n->partial.p, kMseBlocks * 8, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  double s = 0;
  for (double p : part) s += p;
  if (loss) *loss = s / double(ny);
  return AMP_OK;
}

int amp_mlp_adam(amp_mlp* n, float lr, float beta1, float beta2, float eps) {
  if (!n || !(lr > 0) || !(beta1 >= 0 && beta1 < 1) || !(beta2 >= 0 && beta2 < 1) || !(eps > 0))
    return fail(AMP_ERR_ARG, "amp_mlp_adam: bad arguments");
  n->step++;
  const float c1 = float(1.0 - std::pow(double(beta1), double(n->step)));
  const float c2 = float(1.0 - std::pow(double(beta2), double(n->step)));
  hipLaunchKernelGGL(amp::k_adam, dim3(amp::blocksFor(n->nParams)), dim3(256), 0, n->stream, n->params.as<float>(),
                     n->grads.as<float>(), n->m.as<float>(), n->v.as<float>(), n->nParams, lr, beta1, beta2, eps, c1, c2);
  HIP_TRY(hipGetLastError());
  return AMP_OK;
}

void amp_mlp_free(amp_mlp* n) { delete n; }

}  // extern "C"
