// Device self-test of the MFMA operand/result lane layout that the generated mixed-precision
// kernels rely on (v_mfma_f32_16x16x4_f32, wave64):
//   A (16x4):  lane l holds A[l % 16][l / 16]
//   B (4x16):  lane l holds B[l / 16][l % 16]
//   D (16x16): lane l, accumulator element i holds D[4 * (l / 16) + i][l % 16]
// One wave computes D = A B from closed-form A and B and writes D row-major; the caller
// compares with the host product. All indices are within one 64-lane wave.
#include <hip/hip_runtime.h>

#include "../../include/kinetix_mi.h"
#include "kx_rcp.h"

typedef float floatx4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(64) kx_mfma_layout_probe(float* __restrict__ d)
{
  const int l = threadIdx.x;
  const int i = l % 16, k = l / 16;
  const float a = (float)(i + 16 * k + 1);            // A[i][k]
  const float b = (float)((k + 1) * 100 + i);         // B[k][j = i]
  floatx4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) d[(4 * (l / 16) + r) * 16 + (l % 16)] = acc[r];
}

extern "C" int kx_selftest_mfma_layout(float* d_out, void* stream)
{
  if (!d_out) return KX_ERR_INVALID_ARG;
  hipLaunchKernelGGL(kx_mfma_layout_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, d_out);
  return hipGetLastError() == hipSuccess ? KX_OK : KX_ERR_HIP;
}

// v_mfma_f64_16x16x4_f64 (the FP64 Wilke contraction of kx_bk2_fp64_mfma). Layout the kernel
// relies on: A (16x4) lane l holds A[l % 16][l / 16], B (4x16) lane l holds B[l / 16][l % 16],
// and lane l, accumulator element r holds D[4 * r + l / 16][l % 16] (rows interleaved across
// the lane groups, unlike the FP32 16x16x4 result). The probe writes each lane's raw
// accumulators (out[4 * l + r]) so that the host checks this.
// mode 0: A[i][k] = i + 16k + 1, B[k][j] = 100 (k + 1) + j (every operand entry used);
// mode 1: A[i][0] = i, A[i][1] = 1, B[0][j] = 16, B[1][j] = j, else 0, so D[i][j] = 16 i + j
// names its own position.
typedef double doublex4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(64) kx_mfma_f64_layout_probe(int mode, double* __restrict__ d)
{
  const int l = threadIdx.x;
  const int i = l % 16, k = l / 16;
  double a, b;
  if (mode == 0) {
    a = (double)(i + 16 * k + 1);
    b = (double)((k + 1) * 100 + i);
  } else {
    a = k == 0 ? (double)i : (k == 1 ? 1.0 : 0.0);
    b = k == 0 ? 16.0 : (k == 1 ? (double)i : 0.0);
  }
  doublex4 acc = {0.0, 0.0, 0.0, 0.0};
  acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) d[4 * l + r] = acc[r];
}

extern "C" int kx_selftest_mfma_f64_layout(int mode, double* d_out, void* stream)
{
  if (!d_out || (mode != 0 && mode != 1)) return KX_ERR_INVALID_ARG;
  hipLaunchKernelGGL(kx_mfma_f64_layout_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, mode,
                     d_out);
  return hipGetLastError() == hipSuccess ? KX_OK : KX_ERR_HIP;
}

// kx_rcp (the transport kernels' reciprocal) on n inputs: out[i] = kx_rcp(in[i]).
__global__ void __launch_bounds__(256) kx_rcp_probe(int n, const double* __restrict__ in,
                                                    double* __restrict__ out)
{
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = kx_rcp(in[i]);
}

extern "C" int kx_selftest_rcp(int n, const double* d_in, double* d_out, void* stream)
{
  if (n < 0 || (n > 0 && (!d_in || !d_out))) return KX_ERR_INVALID_ARG;
  if (n == 0) return KX_OK;
  hipLaunchKernelGGL(kx_rcp_probe, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n,
                     d_in, d_out);
  return hipGetLastError() == hipSuccess ? KX_OK : KX_ERR_HIP;
}
