// Hand-written kernels of the transport-surrogate MLP: 64 -> 256 -> 256 -> 256 -> 64
// (55 inputs / 55 outputs zero-padded to 64), GELU (tanh form) on the hidden layers.
//
// Training (FP32 master weights): tiled FP32 GEMM with either operand transposed, bias +
// GELU forward, GELU backward, masked-MSE gradient, bias-gradient reduction, Adam.
// Inference: one fused kernel per 128-state tile on BF16 MFMA (v_mfma_f32_16x16x16_bf16,
// FP32 accumulation); activations ping-pong between two LDS buffers, weights stream
// through LDS in K-chunks of 32.
//
// Parameter block (FP32, contiguous): W1[256][64] b1[256] W2[256][256] b2[256]
// W3[256][256] b3[256] W4[64][256] b4[64]; the BF16 block holds the four W's in the same
// order (biases stay FP32).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/kinetix_mi.h"

namespace {

constexpr int kIn = 64, kH = 256, kOut = 64, kValidOut = 55;
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH, kB3 = kW3 + kH * kH, kW4 = kB3 + kH, kB4 = kW4 + kOut * kH;
constexpr int kParams = kB4 + kOut;
constexpr int kWb1 = 0, kWb2 = kWb1 + kH * kIn, kWb3 = kWb2 + kH * kH, kWb4 = kWb3 + kH * kH;
constexpr int kWeightsBf16 = kWb4 + kOut * kH;

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

__global__ void __launch_bounds__(256) pack_bf16(const float* __restrict__ p, uint16_t* __restrict__ wb)
{
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= kWeightsBf16) return;
  int src;
  if (i < kWb2) src = kW1 + i;
  else if (i < kWb3) src = kW2 + (i - kWb2);
  else if (i < kWb4) src = kW3 + (i - kWb3);
  else src = kW4 + (i - kWb4);
  wb[i] = to_bf16(p[src]);
}

// ---------------------------------------------------------------------------------------
// Fused BF16-MFMA inference. v_mfma_f32_16x16x16_bf16 lane layout (checked on the device
// by kx_selftest_mfma_layout): A[i][k] in lane i + 16 (k / 4), element k % 4;
// B[k][j] in lane j + 16 (k / 4), element k % 4; D[i][j] in element i % 4 of lane j + 16 (i / 4).
typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 128;        // states per workgroup
constexpr int kHS = 260;          // LDS row stride (bf16) of the activation buffers
constexpr int kWS = 36;           // LDS row stride (bf16) of a 32-wide weight chunk

template <int KIN, int NOUT, bool LAST>
__device__ __forceinline__ void mlp_layer(const uint16_t* __restrict__ Wg, const float* __restrict__ bias,
                                          const uint16_t* Hin, uint16_t* Hout, uint16_t* Wc,
                                          float* __restrict__ out, long long s0, int n_states)
{
  constexpr int CPW = NOUT / 4;                  // output columns per wave
  constexpr int CB = CPW / 16;                   // column blocks per wave
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int li = lane & 15, lg = lane >> 4;
  floatx4 acc[8][CB];
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c) acc[r][c] = (floatx4){0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < KIN; k0 += 32) {
    // stage W[:, k0:k0+32] (NOUT x 32 bf16) in LDS: 8 bf16 (16 B) per load
    for (int e = threadIdx.x; e < NOUT * 4; e += 256) {
      const int row = e >> 2, q = e & 3;
      const uint4 v = *(const uint4*)(Wg + (size_t)row * KIN + k0 + 8 * q);
      uint16_t* d = Wc + row * kWS + 8 * q;
      *(uint2*)d = make_uint2(v.x, v.y);
      *(uint2*)(d + 4) = make_uint2(v.z, v.w);
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < 32; ks += 16) {
      bf16x4 b[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c)
        b[c] = *(const bf16x4*)(Wc + (wv * CPW + c * 16 + li) * kWS + ks + 4 * lg);
#pragma unroll
      for (int r = 0; r < 8; ++r) {
        const bf16x4 a = *(const bf16x4*)(Hin + (r * 16 + li) * kHS + k0 + ks + 4 * lg);
#pragma unroll
        for (int c = 0; c < CB; ++c)
          acc[r][c] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b[c], acc[r][c], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int r = 0; r < 8; ++r)
#pragma unroll
    for (int c = 0; c < CB; ++c) {
      const int col = wv * CPW + c * 16 + li;
      const float bc = bias[col];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int row = r * 16 + 4 * lg + i;
        const float z = acc[r][c][i] + bc;
        if (LAST) {
          if (s0 + row < n_states) out[(size_t)(s0 + row) * kOut + col] = z;
        } else {
          Hout[row * kHS + col] = to_bf16(gelu(z));
        }
      }
    }
  __syncthreads();
}

__global__ void __launch_bounds__(256) mlp_infer_bf16(const uint16_t* __restrict__ wb,
                                                      const float* __restrict__ p, int n_states,
                                                      const float* __restrict__ X, float* __restrict__ Y)
{
  __shared__ __attribute__((aligned(16))) uint16_t H0[kTile * kHS];
  __shared__ __attribute__((aligned(16))) uint16_t H1[kTile * kHS];
  __shared__ __attribute__((aligned(16))) uint16_t Wc[kH * kWS];
  const long long s0 = (long long)blockIdx.x * kTile;
  for (int e = threadIdx.x; e < kTile * kIn; e += 256) {
    const int row = e >> 6, col = e & 63;
    const float x = (s0 + row < n_states) ? X[(size_t)(s0 + row) * kIn + col] : 0.0f;
    H0[row * kHS + col] = to_bf16(x);
  }
  __syncthreads();
  mlp_layer<kIn, kH, false>(wb + kWb1, p + kB1, H0, H1, Wc, nullptr, s0, n_states);
  mlp_layer<kH, kH, false>(wb + kWb2, p + kB2, H1, H0, Wc, nullptr, s0, n_states);
  mlp_layer<kH, kH, false>(wb + kWb3, p + kB3, H0, H1, Wc, nullptr, s0, n_states);
  mlp_layer<kH, kOut, true>(wb + kWb4, p + kB4, H1, nullptr, Wc, Y, s0, n_states);
}

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

inline int grid1(size_t n) { return (int)((n + 255) / 256); }

int hip_status() { return hipGetLastError() == hipSuccess ? KX_OK : KX_ERR_HIP; }

}  // namespace

extern "C" {

int kx_mlp_n_params(void) { return kParams; }
int kx_mlp_n_weights_bf16(void) { return kWeightsBf16; }

int kx_selftest_mfma_bf16_layout(float* d_out, void* stream)
{
  if (!d_out) return KX_ERR_INVALID_ARG;
  hipLaunchKernelGGL(bf16_layout_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, d_out);
  return hip_status();
}

// acts (FP32, 7 blocks of B x 256 = Z1 H1 Z2 H2 Z3 H3 and Y (B x 64) at the end)
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
// (same layout as p). work: 2 x B x 256 floats. loss_part: ceil(B*64/256) floats.
int kx_mlp_backward(const float* p, int B, const float* X, const float* acts, const float* T,
                    float* g, float* work, float* loss_part, void* stream)
{
  if (!p || !X || !acts || !T || !g || !work || !loss_part || B <= 0 || B % 64)
    return KX_ERR_INVALID_ARG;
  hipStream_t s = (hipStream_t)stream;
  const size_t L = (size_t)B * kH;
  const float *Z1 = acts, *H1 = Z1 + L, *Z2 = H1 + L, *H2 = Z2 + L, *Z3 = H2 + L, *H3 = Z3 + L, *Yo = H3 + L;
  float *dA = work, *dB = work + L;       // dH / dZ scratch (B x 256 each)
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
  hipLaunchKernelGGL(mlp_infer_bf16, dim3((n + kTile - 1) / kTile), dim3(256), 0, (hipStream_t)stream,
                     wb, p, n, X, Y);
  return hip_status();
}

}  // extern "C"
