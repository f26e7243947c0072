// Hand-written kernels of the transport-surrogate MLP: 64 -> kH -> kH -> kH -> 64 (kH = 256)
// (55 inputs / 55 outputs zero-padded to 64), GELU (tanh form) on the hidden layers.
//
// Training (FP32 master weights): tiled FP32 GEMM with either operand transposed, bias +
// GELU forward, GELU backward, masked-MSE gradient, bias-gradient reduction, Adam.
// Inference: BF16 weights packed into MFMA fragment order (kx_mlp_wave.h) and the
// wave-level forward pass of kx_mlp_wave.h, 64 states per wavefront.
//
// Parameter block (FP32, contiguous): W1[kH][64] b1[kH] W2[kH][kH] b2[kH]
// W3[kH][kH] b3[kH] W4[64][kH] b4[64].
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/kinetix_mi.h"
#include "kx_mlp_wave.h"

namespace {

using kxmlp::kIn;
using kxmlp::kH;
using kxmlp::kOut;
using kxmlp::kValidOut;
using kxmlp::kW1;
using kxmlp::kB1;
using kxmlp::kW2;
using kxmlp::kB2;
using kxmlp::kW3;
using kxmlp::kB3;
using kxmlp::kW4;
using kxmlp::kB4;
using kxmlp::kParams;
constexpr int kWeightsBf16 = kxmlp::kFragTotal;

__device__ __forceinline__ float gelu(float x)
{
  const float u = 0.7978845608028654f * (x + 0.044715f * x * x * x);
  return 0.5f * x * (1.0f + tanhf(u));
}

__device__ __forceinline__ float gelu_grad(float x)
{
  const float x2 = x * x;
  const float u = 0.7978845608028654f * (x + 0.044715f * x2 * x);
  const float th = tanhf(u);
  return 0.5f * (1.0f + th) + 0.5f * x * (1.0f - th * th) * 0.7978845608028654f * (1.0f + 3.0f * 0.044715f * x2);
}

__device__ __forceinline__ uint16_t to_bf16(float f)
{
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
  u += 0x7fffu + ((u >> 16) & 1u);      // round to nearest even
  return (uint16_t)(u >> 16);
}

// C[MxN] = alpha * op(A) op(B) + beta * C, row-major; op(A) is MxK (A or A^T stored KxM),
// op(B) is KxN (B or B^T stored NxK). M, N multiples of 64, K multiple of 16.
template <bool TA, bool TB>
__global__ void __launch_bounds__(256) gemm_f32(int M, int N, int K, float alpha,
                                                const float* __restrict__ A, int lda,
                                                const float* __restrict__ B, int ldb, float beta,
                                                float* __restrict__ C, int ldc)
{
  __shared__ float As[16][68];
  __shared__ float Bs[16][68];
  const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
  const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
  float acc[4][4] = {};
  for (int k0 = 0; k0 < K; k0 += 16) {
    for (int e = threadIdx.x; e < 1024; e += 256) {
      if (!TA) { const int mm = e >> 4, kk = e & 15; As[kk][mm] = A[(size_t)(m0 + mm) * lda + k0 + kk]; }
      else     { const int kk = e >> 6, mm = e & 63; As[kk][mm] = A[(size_t)(k0 + kk) * lda + m0 + mm]; }
      if (!TB) { const int kk = e >> 6, nn = e & 63; Bs[kk][nn] = B[(size_t)(k0 + kk) * ldb + n0 + nn]; }
      else     { const int nn = e >> 4, kk = e & 15; Bs[kk][nn] = B[(size_t)(n0 + nn) * ldb + k0 + kk]; }
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      float a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) { a[i] = As[kk][ty * 4 + i]; b[i] = Bs[kk][tx * 4 + i]; }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fmaf(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      float* c = &C[(size_t)(m0 + ty * 4 + i) * ldc + n0 + tx * 4 + j];
      *c = alpha * acc[i][j] + (beta != 0.0f ? beta * *c : 0.0f);
    }
}

// Z[M][N] += b[N]; H = gelu(Z) if H != nullptr
__global__ void __launch_bounds__(256) bias_act(int M, int N, float* __restrict__ Z,
                                                const float* __restrict__ b, float* __restrict__ H)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)M * N) return;
  const float z = Z[i] + b[i % N];
  Z[i] = z;
  if (H) H[i] = gelu(z);
}

// dZ = dH * gelu'(Z)
__global__ void __launch_bounds__(256) gelu_back(int M, int N, const float* __restrict__ dH,
                                                 const float* __restrict__ Z, float* __restrict__ dZ)
{
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= (size_t)M * N) return;
  dZ[i] = dH[i] * gelu_grad(Z[i]);
}

// dY = 2 (Y - T) / (M * valid) on the first `valid` columns, 0 on the padding; per-block
// partial sums of the squared error go to part[blockIdx.x].
__global__ void __launch_bounds__(256) mse_grad(int M, int N, int valid, const float* __restrict__ Y,
                                                const float* __restrict__ T, float* __restrict__ dY,
                                                float* __restrict__ part)
{
  __shared__ float red[256];
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  float se = 0.0f;
  if (i < (size_t)M * N) {
    const int col = (int)(i % N);
    const float d = (col < valid) ? Y[i] - T[i] : 0.0f;
    dY[i] = 2.0f * d / ((float)M * (float)valid);
    se = d * d;
  }
  red[threadIdx.x] = se;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// db[j] = sum_i dZ[i][j]  (one workgroup per column, fixed summation order)
__global__ void __launch_bounds__(256) bias_grad(int M, int N, const float* __restrict__ dZ,
                                                 float* __restrict__ db)
{
  __shared__ float red[256];
  const int j = blockIdx.x;
  float s = 0.0f;
  for (int i = threadIdx.x; i < M; i += 256) s += dZ[(size_t)i * N + j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) db[j] = red[0];
}

__global__ void __launch_bounds__(256) adam(int n, float* __restrict__ w, const float* __restrict__ g,
                                            float* __restrict__ m, float* __restrict__ v, float lr,
                                            float b1, float b2, float eps, float c1, float c2)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// BF16 weights in MFMA fragment order: fragment (ob, kb) of layer l, lane i + 16 g,
// element e = W_l[16 ob + i][kphi(kb, g, e)] (kx_mlp_wave.h).
__global__ void __launch_bounds__(256) pack_bf16(const float* __restrict__ p, uint16_t* __restrict__ wb)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kWeightsBf16) return;
  int base, K, w;
  if (i < kxmlp::kF2)      { base = kxmlp::kF1; K = kIn; w = kW1; }
  else if (i < kxmlp::kF3) { base = kxmlp::kF2; K = kH;  w = kW2; }
  else if (i < kxmlp::kF4) { base = kxmlp::kF3; K = kH;  w = kW3; }
  else                     { base = kxmlp::kF4; K = kH;  w = kW4; }
  const int local = i - base, e = local & 7, lane = (local >> 3) & 63, frag = local >> 9;
  const int KB = K / 32, kb = frag % KB, ob = frag / KB;
  const int row = 16 * ob + (lane & 15), col = kxmlp::kphi(kb, lane >> 4, e);
  wb[i] = to_bf16(p[w + row * K + col]);
}

// BF16-MFMA inference of n states (row-major FP32 X and Y, n x 64): 4 waves x 64 states
// per workgroup; the inputs go through LDS once to reach the MFMA fragment layout.
constexpr int kXS = 68;       // LDS row stride (bf16) of the staged inputs

__global__ void __launch_bounds__(256, 1) mlp_infer_bf16(const uint16_t* __restrict__ wb,
                                                         const float* __restrict__ p, int n_states,
                                                         const float* __restrict__ X, float* __restrict__ Y)
{
  __shared__ __attribute__((aligned(16))) uint16_t xs[4][64 * kXS];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long s0 = (long long)blockIdx.x * 256 + wv * 64;
  // stage the wave's 64 input rows as BF16 (each pass reads one 256-B row, coalesced)
  for (int e = lane; e < 64 * kIn; e += 64) {
    const int row = e >> 6, col = e & 63;
    const float x = (s0 + row < n_states) ? X[(s0 + row) * kIn + col] : 0.0f;
    xs[wv][row * kXS + col] = to_bf16(x);
  }
  __syncthreads();
  kxmlp::f32x4 y[kOut / 16][4];
  kxmlp::forward<kXS>(wb, p, xs[wv], y, lane);
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int sb = 0; sb < 4; ++sb) {
    const long long s = s0 + 16 * sb + j;
    if (s < n_states) {
#pragma unroll
      for (int ob = 0; ob < kOut / 16; ++ob)
        *(kxmlp::f32x4*)(Y + s * kOut + 16 * ob + 4 * g) = y[ob][sb];
    }
  }
}

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// Self-test of the BF16 MFMA layout: D = A B for closed-form 16x16 A and B.
__global__ void __launch_bounds__(64) bf16_layout_probe(float* __restrict__ d)
{
  const int l = threadIdx.x, i = l & 15, g = l >> 4;
  bf16x4 a, b;
  for (int e = 0; e < 4; ++e) {
    const int k = 4 * g + e;
    a[e] = (short)to_bf16((float)(i + 16 * k + 1) / 16.0f);      // A[i][k]
    b[e] = (short)to_bf16((float)((k + 1) * 3 + i));             // B[k][j = i]
  }
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) d[(4 * g + r) * 16 + i] = acc[r];
}

// Self-test of the v_mfma_f32_16x16x32_bf16 layout the inference path relies on:
// A[i][k] in element k % 8 of lane i + 16 (k / 8), B[k][j] likewise with j, D as above;
// small integer entries so that every product and sum is exact.
__global__ void __launch_bounds__(64) bf16_k32_layout_probe(float* __restrict__ d)
{
  const int l = threadIdx.x, i = l & 15, g = l >> 4;
  kxmlp::u32x4 a, b;
  for (int h = 0; h < 4; ++h) {
    uint32_t wa = 0, wbv = 0;
    for (int q = 0; q < 2; ++q) {
      const int k = 8 * g + 2 * h + q;
      wa |= (uint32_t)to_bf16((float)((i * 7 + k * 3) % 17 - 8)) << (16 * q);     // A[i][k]
      wbv |= (uint32_t)to_bf16((float)((k * 5 + i * 11) % 13 - 6)) << (16 * q);   // B[k][j = i]
    }
    a[h] = wa;
    b[h] = wbv;
  }
  kxmlp::f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = kxmlp::mfma(a, b, acc);
  for (int r = 0; r < 4; ++r) d[(4 * g + r) * 16 + i] = acc[r];
}

inline int grid1(size_t n) { return (int)((n + 255) / 256); }

int hip_status() { return hipGetLastError() == hipSuccess ? KX_OK : KX_ERR_HIP; }

}  // namespace

extern "C" {

int kx_mlp_width(void) { return kH; }
int kx_mlp_n_params(void) { return kParams; }
int kx_mlp_n_weights_bf16(void) { return kWeightsBf16; }

int kx_selftest_mfma_bf16_layout(float* d_out, void* stream)
{
  if (!d_out) return KX_ERR_INVALID_ARG;
  hipLaunchKernelGGL(bf16_layout_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, d_out);
  return hip_status();
}

int kx_selftest_mfma_bf16_k32_layout(float* d_out, void* stream)
{
  if (!d_out) return KX_ERR_INVALID_ARG;
  hipLaunchKernelGGL(bf16_k32_layout_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, d_out);
  return hip_status();
}

// acts (FP32, 6 blocks of B x kH = Z1 H1 Z2 H2 Z3 H3 and Y (B x 64) at the end)
int kx_mlp_forward(const float* p, int B, const float* X, float* acts, void* stream)
{
  if (!p || !X || !acts || B <= 0 || B % 64) return KX_ERR_INVALID_ARG;
  hipStream_t s = (hipStream_t)stream;
  const size_t L = (size_t)B * kH;
  float *Z1 = acts, *H1 = Z1 + L, *Z2 = H1 + L, *H2 = Z2 + L, *Z3 = H2 + L, *H3 = Z3 + L, *Yo = H3 + L;
  const dim3 gH(kH / 64, B / 64), gO(kOut / 64, B / 64);
  hipLaunchKernelGGL((gemm_f32<false, true>), gH, dim3(256), 0, s, B, kH, kIn, 1.f, X, kIn, p + kW1, kIn, 0.f, Z1, kH);
  hipLaunchKernelGGL(bias_act, dim3(grid1(L)), dim3(256), 0, s, B, kH, Z1, p + kB1, H1);
  hipLaunchKernelGGL((gemm_f32<false, true>), gH, dim3(256), 0, s, B, kH, kH, 1.f, H1, kH, p + kW2, kH, 0.f, Z2, kH);
  hipLaunchKernelGGL(bias_act, dim3(grid1(L)), dim3(256), 0, s, B, kH, Z2, p + kB2, H2);
  hipLaunchKernelGGL((gemm_f32<false, true>), gH, dim3(256), 0, s, B, kH, kH, 1.f, H2, kH, p + kW3, kH, 0.f, Z3, kH);
  hipLaunchKernelGGL(bias_act, dim3(grid1(L)), dim3(256), 0, s, B, kH, Z3, p + kB3, H3);
  hipLaunchKernelGGL((gemm_f32<false, true>), gO, dim3(256), 0, s, B, kOut, kH, 1.f, H3, kH, p + kW4, kH, 0.f, Yo, kOut);
  hipLaunchKernelGGL(bias_act, dim3(grid1((size_t)B * kOut)), dim3(256), 0, s, B, kOut, Yo, p + kB4, (float*)nullptr);
  return hip_status();
}

// Gradient of the masked MSE (first 55 outputs) w.r.t. all parameters, written to g
// (same layout as p). work: 2 x B x kH floats. loss_part: ceil(B*64/256) floats.
int kx_mlp_backward(const float* p, int B, const float* X, const float* acts, const float* T,
                    float* g, float* work, float* loss_part, void* stream)
{
  if (!p || !X || !acts || !T || !g || !work || !loss_part || B <= 0 || B % 64)
    return KX_ERR_INVALID_ARG;
  hipStream_t s = (hipStream_t)stream;
  const size_t L = (size_t)B * kH;
  const float *Z1 = acts, *H1 = Z1 + L, *Z2 = H1 + L, *H2 = Z2 + L, *Z3 = H2 + L, *H3 = Z3 + L, *Yo = H3 + L;
  float *dA = work, *dB = work + L;       // dH / dZ scratch (B x kH each)
  float* dY = dB;                          // B x 64 fits in dB
  hipLaunchKernelGGL(mse_grad, dim3(grid1((size_t)B * kOut)), dim3(256), 0, s, B, kOut, kValidOut, Yo, T, dY, loss_part);
  // layer 4: dW4 = dY^T H3, db4, dH3 = dY W4
  hipLaunchKernelGGL((gemm_f32<true, false>), dim3(kH / 64, kOut / 64), dim3(256), 0, s, kOut, kH, B, 1.f, dY, kOut, H3, kH, 0.f, g + kW4, kH);
  hipLaunchKernelGGL(bias_grad, dim3(kOut), dim3(256), 0, s, B, kOut, dY, g + kB4);
  hipLaunchKernelGGL((gemm_f32<false, false>), dim3(kH / 64, B / 64), dim3(256), 0, s, B, kH, kOut, 1.f, dY, kOut, p + kW4, kH, 0.f, dA, kH);
  // layer 3
  hipLaunchKernelGGL(gelu_back, dim3(grid1(L)), dim3(256), 0, s, B, kH, dA, Z3, dA);
  hipLaunchKernelGGL((gemm_f32<true, false>), dim3(kH / 64, kH / 64), dim3(256), 0, s, kH, kH, B, 1.f, dA, kH, H2, kH, 0.f, g + kW3, kH);
  hipLaunchKernelGGL(bias_grad, dim3(kH), dim3(256), 0, s, B, kH, dA, g + kB3);
  hipLaunchKernelGGL((gemm_f32<false, false>), dim3(kH / 64, B / 64), dim3(256), 0, s, B, kH, kH, 1.f, dA, kH, p + kW3, kH, 0.f, dB, kH);
  // layer 2
  hipLaunchKernelGGL(gelu_back, dim3(grid1(L)), dim3(256), 0, s, B, kH, dB, Z2, dB);
  hipLaunchKernelGGL((gemm_f32<true, false>), dim3(kH / 64, kH / 64), dim3(256), 0, s, kH, kH, B, 1.f, dB, kH, H1, kH, 0.f, g + kW2, kH);
  hipLaunchKernelGGL(bias_grad, dim3(kH), dim3(256), 0, s, B, kH, dB, g + kB2);
  hipLaunchKernelGGL((gemm_f32<false, false>), dim3(kH / 64, B / 64), dim3(256), 0, s, B, kH, kH, 1.f, dB, kH, p + kW2, kH, 0.f, dA, kH);
  // layer 1
  hipLaunchKernelGGL(gelu_back, dim3(grid1(L)), dim3(256), 0, s, B, kH, dA, Z1, dA);
  hipLaunchKernelGGL((gemm_f32<true, false>), dim3(kIn / 64, kH / 64), dim3(256), 0, s, kH, kIn, B, 1.f, dA, kH, X, kIn, 0.f, g + kW1, kIn);
  hipLaunchKernelGGL(bias_grad, dim3(kH), dim3(256), 0, s, B, kH, dA, g + kB1);
  return hip_status();
}

int kx_mlp_adam(float* p, const float* g, float* m, float* v, float lr, float b1, float b2,
                float eps, int step, void* stream)
{
  if (!p || !g || !m || !v || step < 1) return KX_ERR_INVALID_ARG;
  const float c1 = 1.0f / (1.0f - powf(b1, (float)step));
  const float c2 = 1.0f / (1.0f - powf(b2, (float)step));
  hipLaunchKernelGGL(adam, dim3(grid1(kParams)), dim3(256), 0, (hipStream_t)stream, kParams, p, g, m, v,
                     lr, b1, b2, eps, c1, c2);
  return hip_status();
}

int kx_mlp_pack_bf16(const float* p, uint16_t* wb, void* stream)
{
  if (!p || !wb) return KX_ERR_INVALID_ARG;
  hipLaunchKernelGGL(pack_bf16, dim3(grid1(kWeightsBf16)), dim3(256), 0, (hipStream_t)stream, p, wb);
  return hip_status();
}

// Y[n][64] = MLP(X[n][64]) with BF16 weights/activations and FP32 accumulation.
int kx_mlp_infer_bf16(const uint16_t* wb, const float* p, int n, const float* X, float* Y, void* stream)
{
  if (!wb || !p || n < 0 || (n > 0 && (!X || !Y))) return KX_ERR_INVALID_ARG;
  if (n == 0) return KX_OK;
  hipLaunchKernelGGL(mlp_infer_bf16, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                     wb, p, n, X, Y);
  return hip_status();
}

}  // extern "C"
