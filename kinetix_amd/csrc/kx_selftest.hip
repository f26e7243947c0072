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
