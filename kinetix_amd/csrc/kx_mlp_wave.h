// Wave-level BF16-MFMA forward pass of the transport-surrogate MLP
// (64 -> kH -> kH -> kH -> 64, GELU on the hidden layers) for 64 states per wavefront.
//
// Layout choices (gfx950, v_mfma_f32_16x16x32_bf16, wave64):
//  * every layer is computed transposed, Z^T = W H^T: the MFMA's A operand is a 16 x 32
//    block of weights (rows = output features), B is a 32 x 16 block of activations
//    (columns = states), D is 16 features x 16 states;
//  * the D layout of 16x16 MFMAs (row i of column j in element i % 4 of lane j + 16 (i / 4))
//    is, element for element, the B layout of the next layer (row k of column j in element
//    k % 8 of lane j + 16 (k / 8)) once the next layer's k index is permuted: element e of
//    lane group g in k-block kb holds feature kphi(kb, g, e), i.e. 4 features from output
//    block 2 kb and 4 from block 2 kb + 1. The weights are pre-packed with the same
//    permutation, so activations never leave registers between layers;
//  * weights stream from L2 as pre-packed fragments (one fragment = 64 lanes x 8 bf16 =
//    1 KB, contiguous), each reused for the wave's 4 column blocks of 16 states, requested
//    ahead of their use; the hidden layers are software-pipelined so one output block's
//    MFMAs overlap the previous block's GELU;
//  * bias is the accumulator's initial value; GELU (tanh form, written as x * sigmoid(2u))
//    costs one exp2 and one reciprocal; the results are packed to BF16 with v_cvt_pk_bf16_f32.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace kxmlp {

// Hidden width: 256, the transport surrogate's specified architecture (55 -> 3 x 256 GELU ->
// 55). The code takes any multiple of 64 (the training GEMM tiles); the fused inference
// kernel's time follows the width (2.65 ms at 256, 2.13 at 192, 1.62 at 128 for 4M states).
constexpr int kIn = 64, kH = 256, kOut = 64, kValidOut = 55;
static_assert(kH % 64 == 0, "training GEMM tiles are 64 wide");
// FP32 parameter block: W1[kH][64] b1[kH] W2[kH][kH] b2[kH] W3[kH][kH] b3[kH] W4[64][kH] b4[64]
constexpr int kW1 = 0, kB1 = kW1 + kH * kIn, kW2 = kB1 + kH, kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH, kB3 = kW3 + kH * kH, kW4 = kB3 + kH, kB4 = kW4 + kOut * kH;
constexpr int kParams = kB4 + kOut;
static_assert(kB1 % 4 == 0 && kB2 % 4 == 0 && kB3 % 4 == 0 && kB4 % 4 == 0, "16-B bias loads");

// BF16 fragment block: layer l holds (out/16) x (in/32) fragments of 512 bf16, ordered
// [output block][k block][lane][8]
constexpr int kFrag = 512;
constexpr int kF1 = 0;
constexpr int kF2 = kF1 + (kH / 16) * (kIn / 32) * kFrag;
constexpr int kF3 = kF2 + (kH / 16) * (kH / 32) * kFrag;
constexpr int kF4 = kF3 + (kH / 16) * (kH / 32) * kFrag;
constexpr int kFragTotal = kF4 + (kOut / 16) * (kH / 32) * kFrag;

// input feature held by element e (0..7) of lane group g (0..3) in k-block kb
__host__ __device__ constexpr int kphi(int kb, int g, int e)
{
  return 32 * kb + (e < 4 ? 4 * g + e : 16 + 4 * g + (e - 4));
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
// Weights and biases are read through global-address-space pointers (global_load, counted by
// vmcnt only); through generic pointers they become flat loads, which also count against
// lgkmcnt, so every wait for an LDS read would wait for the weight prefetches too.
typedef const u32x4 __attribute__((address_space(1))) gu32x4;
typedef const float __attribute__((address_space(1))) gfloat;
typedef const f32x4 __attribute__((address_space(1))) gf32x4;

__device__ __forceinline__ unsigned pack2(float a, float b)
{
  return __builtin_bit_cast(unsigned, __builtin_convertvector((f32x2){a, b}, bf16x2));
}

__device__ __forceinline__ f32x4 mfma(u32x4 a, u32x4 b, f32x4 c)
{
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a),
                                                  __builtin_bit_cast(bf16x8, b), c, 0, 0, 0);
}

// GELU, tanh form: 0.5 x (1 + tanh u) = x / (1 + exp(-2u)), u = sqrt(2/pi) (x + 0.044715 x^3);
// exp(-2u) = exp2(-x (c0 + c1 x^2)) with c0 = 2 sqrt(2/pi) log2(e), c1 = 0.044715 c0.
__device__ __forceinline__ float gelu_fast(float x)
{
  constexpr float c0 = 2.302208198144325f, c1 = 0.10294324037962224f;
  const float a = x * __builtin_fmaf(c1, x * x, c0);
  return x * __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(-a));
}

// activations of one pass (SB blocks of 16 states) for one layer: [k block][state block]
// fragments, kH / 32 x SB x 4 registers per layer. A pass of SB = 4 covers the wave's 64
// states (input + output layer: 2 kH / 8 x 4 registers, 256 at kH = 256); SB = 2 covers 32
// states with half that, at twice the weight reads per state.
template <int KB, int SB = 4>
using Act = u32x4[KB][SB];

// B fragments of the first layer from LDS: xs holds the pass's states as rows of XS bf16
// (feature f at column f).
template <int XS, int SB = 4>
__device__ __forceinline__ void load_input(const uint16_t* xs, Act<kIn / 32, SB>& in, int lane)
{
  static_assert(XS % 4 == 0, "8-byte LDS reads");
  const int j = lane & 15, g = lane >> 4;
#pragma unroll
  for (int kb = 0; kb < kIn / 32; ++kb)
#pragma unroll
    for (int sb = 0; sb < SB; ++sb) {
      const uint16_t* r = xs + (16 * sb + j) * XS + 32 * kb + 4 * g;
      const uint2 lo = *(const uint2*)r;           // features kphi(kb, g, 0..3)
      const uint2 hi = *(const uint2*)(r + 16);    // features kphi(kb, g, 4..7)
      in[kb][sb] = (u32x4){lo.x, lo.y, hi.x, hi.y};
    }
}

// Software-pipelined over output blocks with two accumulator sets used alternately (no
// copies between them): while the MFMAs of block ob + 1 run, bias + GELU is applied to block
// ob, so the MFMA pipe and the VALU overlap; the fragments and bias of block ob + 2 are
// requested one block ahead (bias first: vmcnt retires loads in order).
template <int KB, int SB>
__device__ __forceinline__ void mfma_block(const u32x4 (&a)[KB], const f32x4& b,
                                           const Act<KB, SB>& in, f32x4 (&acc)[SB])
{
#pragma unroll
  for (int sb = 0; sb < SB; ++sb) acc[sb] = b;
#pragma unroll
  for (int kb = 0; kb < KB; ++kb)
#pragma unroll
    for (int sb = 0; sb < SB; ++sb) acc[sb] = mfma(a[kb], in[kb][sb], acc[sb]);
}

template <int SB>
__device__ __forceinline__ void gelu_block(const f32x4 (&acc)[SB], Act<kH / 32, SB>& out, int ob)
{
#pragma unroll
  for (int sb = 0; sb < SB; ++sb) {
    const unsigned lo = pack2(gelu_fast(acc[sb][0]), gelu_fast(acc[sb][1]));
    const unsigned hi = pack2(gelu_fast(acc[sb][2]), gelu_fast(acc[sb][3]));
    out[ob >> 1][sb][(ob & 1) * 2] = lo;
    out[ob >> 1][sb][(ob & 1) * 2 + 1] = hi;
  }
}

template <int KB>
__device__ __forceinline__ void load_block(gu32x4* __restrict__ wf, gfloat* __restrict__ bias,
                                           int ob, int lane, u32x4 (&a)[KB], f32x4& b)
{
  b = *(gf32x4*)(bias + 16 * ob + 4 * (lane >> 4));
#pragma unroll
  for (int kb = 0; kb < KB; ++kb) a[kb] = wf[(ob * KB + kb) * 64 + lane];
}

template <int KB, int SB>
__device__ __forceinline__ void hidden_layer(gu32x4* __restrict__ wf, gfloat* __restrict__ bias,
                                             const Act<KB, SB>& in, Act<kH / 32, SB>& out, int lane)
{
  constexpr int OB = kH / 16;
  u32x4 a0[KB], a1[KB];
  f32x4 b0, b1;
  f32x4 accA[SB], accB[SB];
  load_block<KB>(wf, bias, 0, lane, a0, b0);
  load_block<KB>(wf, bias, 1, lane, a1, b1);
  mfma_block<KB>(a0, b0, in, accA);                       // block 0
#pragma unroll
  for (int ob = 0; ob < OB; ob += 2) {
    __builtin_amdgcn_sched_barrier(0);
    if (ob + 2 < OB) load_block<KB>(wf, bias, ob + 2, lane, a0, b0);
    mfma_block<KB>(a1, b1, in, accB);                     // block ob + 1 ...
    gelu_block(accA, out, ob);                            // ... overlaps block ob's GELU
    __builtin_amdgcn_sched_barrier(0);
    if (ob + 3 < OB) load_block<KB>(wf, bias, ob + 3, lane, a1, b1);
    if (ob + 2 < OB) mfma_block<KB>(a0, b0, in, accA);    // block ob + 2 ...
    gelu_block(accB, out, ob + 1);                        // ... overlaps block ob + 1's GELU
  }
}

// y[ob][sb][e]: output feature 16 ob + 4 (lane / 16) + e of state 16 sb + lane % 16
template <int SB>
__device__ __forceinline__ void output_layer(gu32x4* __restrict__ wf, gfloat* __restrict__ bias,
                                             const Act<kH / 32, SB>& in, f32x4 (&y)[kOut / 16][SB], int lane)
{
  const int g = lane >> 4;
  u32x4 a_cur[kH / 32], a_nxt[kH / 32];
  f32x4 b_cur = *(gf32x4*)(bias + 4 * g), b_nxt;
#pragma unroll
  for (int kb = 0; kb < kH / 32; ++kb) a_cur[kb] = wf[kb * 64 + lane];
#pragma unroll
  for (int ob = 0; ob < kOut / 16; ++ob) {
    if (ob + 1 < kOut / 16) {
      b_nxt = *(gf32x4*)(bias + 16 * (ob + 1) + 4 * g);
#pragma unroll
      for (int kb = 0; kb < kH / 32; ++kb) a_nxt[kb] = wf[((ob + 1) * (kH / 32) + kb) * 64 + lane];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int sb = 0; sb < SB; ++sb) y[ob][sb] = b_cur;
#pragma unroll
    for (int kb = 0; kb < kH / 32; ++kb)
#pragma unroll
      for (int sb = 0; sb < SB; ++sb) y[ob][sb] = mfma(a_cur[kb], in[kb][sb], y[ob][sb]);
#pragma unroll
    for (int kb = 0; kb < kH / 32; ++kb) a_cur[kb] = a_nxt[kb];
    b_cur = b_nxt;
  }
}

// Forward pass for 16 SB states (a wave's 64 at SB = 4). wf: fragment block (kFragTotal
// bf16, 16-B aligned); p: FP32 parameter block (biases); both in global memory (device
// allocations: they are read through global-address-space pointers); xs: LDS rows of the
// standardised inputs of the pass's first state onwards.
template <int XS, int SB = 4>
__device__ __forceinline__ void forward(const uint16_t* __restrict__ wfrag, const float* __restrict__ p,
                                        const uint16_t* xs, f32x4 (&y)[kOut / 16][SB], int lane)
{
  // Opaque copies of the (uniform) parameter pointers: when forward() runs more than once per
  // wave (passes over 32 states), the weight loads are otherwise loop-invariant and the
  // compiler hoists or merges all 320 fragments (1,280 VGPRs) across the passes.
  asm volatile("" : "+s"(wfrag), "+s"(p));
  gu32x4* wf = (gu32x4*)wfrag;           // back to global: the asm copies are generic pointers
  gfloat* pg = (gfloat*)p;
  Act<kIn / 32, SB> x0;
  load_input<XS, SB>(xs, x0, lane);
  Act<kH / 32, SB> h1, h2;
  hidden_layer<kIn / 32>(wf + kF1 / 8, pg + kB1, x0, h1, lane);
  hidden_layer<kH / 32>(wf + kF2 / 8, pg + kB2, h1, h2, lane);
  hidden_layer<kH / 32>(wf + kF3 / 8, pg + kB3, h2, h1, lane);
  output_layer(wf + kF4 / 8, pg + kB4, h1, y, lane);
}

}  // namespace kxmlp
