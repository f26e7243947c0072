"""Mechanism-specific kernels of the learned transport surrogate (gfx950 HIP).

The surrogate predicts the mixture-averaged transport outputs y = (ln mu, ln lambda,
ln D_1m .. ln D_Nm) as  y = b(T, p, Y) + mean + std * MLP(x),  where b is a cheap FP32
physics baseline and the MLP (55 -> 3x256 GELU -> 55, BF16 MFMA inference; hand-written
kernels in kinetix_amd/csrc/kx_mlp.hip) learns the standardised residual:

  * mu baseline: the Wilke mixing rule (the formula the reference uses,
    kinetix/core/mix_transport.py:297-357) with its three constant pair matrices C1^2,
    C1 C2, C2^2 replaced by rank-r factorisations (their singular values fall by ~30x per
    rank), so the O(N^2) pair sum costs O(r N). For GRI-3.0, r = 5 (sixth singular values
    <= 2.5e-6 of the first) keeps mu within 2.4e-6 of the full-rank rule over 20,000
    transport states, against the 1e-3 the surrogate answers to; r = 8 would be 1e-11 at
    ~950 more FP32 FMAs per wavefront of 64 states;
  * lambda baseline: the reference's O(N) conductivity mixing rule (mix_transport.py:261-294);
  * D_km baseline: the binary diffusion coefficient of species k in N2 (N2 in O2), at the
    state's T and p, from the same pair fits.

Kernels:
  kx_sur_infer     the inference path in one pass: inputs, baselines and hull flag (one state
                   per lane), the MLP on BF16 MFMA (64 states per wavefront,
                   kinetix_amd/csrc/kx_mlp_wave.h), outputs mu, lambda, rho*D_km (and D_km)
                   in FP64 for states not flagged for the plain-kernel fallback
  kx_sur_features  state -> standardised MLP input (n x 64, FP32), baseline (n x 64) and an
                   out-of-hull flag per state (input outside the training box); training
  kx_sur_targets   plain-kernel labels -> standardised training targets (n x 64)
  kx_sur_outputs   baseline + de-standardised MLP output -> outputs (the composed path
                   features -> kx_mlp_infer_bf16 -> outputs, a cross-check of kx_sur_infer)
"""
from __future__ import annotations

import math

import numpy as np

from . import emit_hip as E

IN = 64            # padded MLP input / output width
RANK = 5
BLOCK = 256


def wilke_factors(t, rank=RANK):
    C1, C2 = t['C1'], t['C2']
    out = []
    for M in (C1 * C1, C1 * C2, C2 * C2):
        U, s, Vt = np.linalg.svd(M)
        out.append((U[:, :rank] * s[:rank], Vt[:rank]))
    return out


def ref_pairs(species):
    """Bath-gas partner of every species for the D baseline: N2 (O2 for N2 itself)."""
    n2 = species.index('N2') if 'N2' in species else 0
    alt = species.index('O2') if 'O2' in species else (1 if n2 == 0 else 0)
    return [n2 if k != n2 else alt for k in range(len(species))]


def baseline_numpy(t, species, T, p, Y):
    """Reference implementation of the baselines (FP64) for tests and training checks."""
    W = t['W']
    lnT = np.log(T)
    P = np.stack([lnT ** i for i in range(5)])
    X = np.maximum(Y, 0) / W[:, None]
    X /= X.sum(0)
    eta = t['visc'] @ P
    fac = wilke_factors(t)
    # S_k = sum_j X_j (C1 + C2 eta_k/eta_j)^2 with each constant matrix in rank-r form
    S = (fac[0][0] @ (fac[0][1] @ X)) + 2 * eta * (fac[1][0] @ (fac[1][1] @ (X / eta))) \
        + eta ** 2 * (fac[2][0] @ (fac[2][1] @ (X / eta ** 2)))
    mu = np.sqrt(T) * (X * eta ** 2 / S).sum(0)
    lam_k = t['cond'] @ P
    lam = np.sqrt(T) * 0.5 * ((X * lam_k).sum(0) + 1.0 / (X / lam_k).sum(0))
    pr = ref_pairs(species)
    d = np.array([t['diff'][k, pr[k]] for k in range(len(species))]) @ P
    D = d * T ** 1.5 / p[None] if not t.get('rcp_diff', 0) else T ** 1.5 / (d * p[None])
    return np.log(mu), np.log(lam), np.log(D)


XS = 68            # LDS row stride (bf16) of the staged MLP inputs in the fused kernel


def emit(t, info, passes=2):
    N = int(t['n_species'])
    if N + 2 > IN:
        return ''                                   # surrogate width is fixed at 64
    if passes not in (1, 2, 4):
        raise ValueError('passes of the wavefront\'s 64 states: 1, 2 or 4')
    sb_n = 4 // passes                              # 16-state blocks per pass
    sp = info['species']
    W, rcpW = t['W'], t['rcpW']
    lit = E.lit
    fac = wilke_factors(t)
    pr = ref_pairs(sp)
    rcp_diff = bool(int(t.get('rcp_diff', 0)))
    S = E._Src()

    def fpoly(c):
        return (f'{float(c[0])!r}f + {float(c[1])!r}f * lT + {float(c[2])!r}f * lT2 + '
                f'{float(c[3])!r}f * lT3 + {float(c[4])!r}f * lT4')

    def state_model(diff=True):
        """Per-state code (one state per lane; T, p, id in scope): raw inputs raw[] and the
        FP32 physics baselines bl[] = (ln mu, ln lambda, ln D_1m .. ln D_Nm); with diff=False
        only bl[0], bl[1] (the fused kernel recomputes the diffusion baselines per output)."""
        S('// FP32 throughout; logarithms and reciprocals on the hardware approximations')
        S('// (v_log_f32, v_rcp_f32: ~1 ulp), which the training features share')
        S(f'float raw[{N + 2}];')
        S('raw[0] = (float)((T - 1400.0) * (1.0 / 1100.0));')
        S(f'const float lnp = kx_lnf((float)p);')
        S(f'raw[1] = (lnp - {float(math.log(101325.0))!r}f) * {1.0 / 1.3!r}f;')
        for k in range(N):
            S(f'raw[{k + 2}] = (float)state[id + offsetT + {k} * offset];')
        S('// baselines (FP32): mole fractions, Wilke viscosity via rank-%d factors, conductivity'
          ' mixing rule, binary diffusion in the bath gas' % RANK)
        S('const float lT = kx_lnf((float)T), lT2 = lT * lT, lT3 = lT2 * lT, lT4 = lT2 * lT2;')
        S('const float sqT = sqrtf((float)T);')
        for k in range(N):
            S(f'const float y{k} = fmaxf(0.0f, raw[{k + 2}]) * {float(rcpW[k])!r}f;')
        S('const float ysum = ' + ' + '.join(f'y{k}' for k in range(N)) + ';')
        S('const float rsum = __builtin_amdgcn_rcpf(ysum);')
        for k in range(N):
            S(f'const float x{k} = y{k} * rsum;')
        for k in range(N):
            S(f'const float e{k} = {fpoly(t["visc"][k])};')
            S(f'const float ie{k} = __builtin_amdgcn_rcpf(e{k});')
        for m in range(RANK):
            for a, (U, V) in enumerate(fac):
                if a == 0:
                    terms = [f'{float(V[m, j])!r}f * x{j}' for j in range(N)]
                elif a == 1:
                    terms = [f'{float(V[m, j])!r}f * (x{j} * ie{j})' for j in range(N)]
                else:
                    terms = [f'{float(V[m, j])!r}f * (x{j} * ie{j} * ie{j})' for j in range(N)]
                S(f'const float v{a}_{m} = ' + ' + '.join(terms) + ';')
        S('float mu = 0.0f;')
        for k in range(N):
            s1 = ' + '.join(f'{float(fac[0][0][k, m])!r}f * v0_{m}' for m in range(RANK))
            s2 = ' + '.join(f'{float(fac[1][0][k, m])!r}f * v1_{m}' for m in range(RANK))
            s3 = ' + '.join(f'{float(fac[2][0][k, m])!r}f * v2_{m}' for m in range(RANK))
            S(f'{{ const float Sk = ({s1}) + 2.0f * e{k} * ({s2}) + e{k} * e{k} * ({s3});')
            S(f'  mu += x{k} * e{k} * e{k} * __builtin_amdgcn_rcpf(Sk); }}')
        S('float l1 = 0.0f, l2 = 0.0f;')
        for k in range(N):
            S(f'{{ const float l = {fpoly(t["cond"][k])}; l1 += x{k} * l; l2 += x{k} * __builtin_amdgcn_rcpf(l); }}')
        S(f'float bl[{N + 2 if diff else 2}];')
        S('bl[0] = kx_lnf(sqT * mu);')
        S('bl[1] = kx_lnf(sqT * 0.5f * (l1 + __builtin_amdgcn_rcpf(l2)));')
        S('const float lT15p = 1.5f * lT - lnp;')
        if diff:
            for k in range(N):
                d = t['diff'][k, pr[k]]
                expr = f'-kx_lnf({fpoly(d)})' if rcp_diff else f'kx_lnf({fpoly(d)})'
                S(f'bl[{k + 2}] = lT15p + {expr};')

    def hull_and_inputs(store):
        """out = 1 if any raw input lies outside the hull box (or is NaN or infinite); store(i,
        expr) receives the standardised input i: (raw - mean) times the hardware reciprocal of
        the standard deviation (~1 ulp FP32, far below the BF16 rounding the MLP input takes;
        a correctly rounded divide costs ~10 instructions per input).

        The box test is min_i min(raw_i - lo_i, hi_i - raw_i) >= 0 in floating point (with
        FP32 denormals kept, x - y >= 0 exactly when x >= y), and NaN/infinity show up in
        sum_i raw_i * 0. With finite bounds (every trained box) this is the same flag as
        per-input compares, without the 3 (N + 2) 64-bit compare masks the compiler otherwise
        hoists and spills to VGPR lanes; with the infinite bounds of the training pass's
        identity normalisation it also flags infinite inputs, a flag that pass ignores."""
        S('float hmin = 1.0f, hnan = 0.0f;')
        S('#pragma unroll')
        S(f'for (int i = 0; i < {N + 2}; ++i) {{')
        S(f'  hmin = fminf(hmin, fminf(raw[i] - norm[{2 * IN} + i], norm[{3 * IN} + i] - raw[i]));')
        S('  hnan += raw[i] * 0.0f;')
        S('  ' + store('i', f'(raw[i] - norm[i]) * __builtin_amdgcn_rcpf(norm[{IN} + i])'))
        S('}')
        S('const int out = !(hmin >= 0.0f) | (hnan != hnan);')

    def rcp_mbar():
        S('double rcpMbar = 0.0;')
        for k in range(N):
            S(f'rcpMbar += fmax(0.0, state[id + offsetT + {k} * offset]) * {lit(rcpW[k])};')

    S('#include "kx_mlp_wave.h"', 0)
    S('__device__ __forceinline__ float kx_lnf(float x) { return __builtin_amdgcn_logf(x) * 0.6931471805599453f; }', 0)
    # ---------------- features (training and the composed inference path)
    S(f'extern "C" __global__ void __launch_bounds__({BLOCK})', 0)
    S('kx_sur_features(const int n_states, const long long offsetT, const long long offset,', 0)
    S('                const double* __restrict__ p_state, const double* __restrict__ state,', 0)
    S('                const float* __restrict__ norm, float* __restrict__ xin,', 0)
    S('                float* __restrict__ base, int* __restrict__ flags, const double Tref)', 0)
    S('{', 0)
    S(E.POOL_MARK)
    S(f'const long long id = (long long)blockIdx.x * {BLOCK} + threadIdx.x;')
    S('if (id >= n_states) return;')
    S('const double T = Tref * state[id];')
    S('const double p = p_state[id];')
    S('// norm: in_mean[64], in_std[64], hull_lo[64], hull_hi[64] (raw input units)')
    state_model()
    S(f'float* xr = xin + id * {IN};')
    hull_and_inputs(lambda i, e: f'xr[{i}] = {e};')
    S(f'for (int i = {N + 2}; i < {IN}; ++i) xr[i] = 0.0f;')
    S('flags[id] = out;')
    S(f'float* br = base + id * {IN};')
    S(f'for (int i = 0; i < {N + 2}; ++i) br[i] = bl[i];')
    S(f'for (int i = {N + 2}; i < {IN}; ++i) br[i] = 0.0f;')
    S('}', 0)

    # ---------------- training targets from plain-kernel labels
    S(f'extern "C" __global__ void __launch_bounds__({BLOCK})', 0)
    S('kx_sur_targets(const int n_states, const long long offset,', 0)
    S('               const double* __restrict__ visc, const double* __restrict__ cond,', 0)
    S('               const double* __restrict__ Dkm, const float* __restrict__ base,', 0)
    S('               const float* __restrict__ ostat, float* __restrict__ tgt)', 0)
    S('{', 0)
    S(f'const long long id = (long long)blockIdx.x * {BLOCK} + threadIdx.x;')
    S('if (id >= n_states) return;')
    S(f'const float* br = base + id * {IN};')
    S(f'float* tr = tgt + id * {IN};')
    S('// ostat: out_mean[64], out_std[64]')
    S(f'tr[0] = (float)(log(visc[id]) - (double)br[0] - (double)ostat[0]) / ostat[{IN}];')
    S(f'tr[1] = (float)(log(cond[id]) - (double)br[1] - (double)ostat[1]) / ostat[{IN + 1}];')
    S(f'for (int k = 0; k < {N}; ++k)')
    S(f'  tr[k + 2] = (float)(log(Dkm[k * offset + id]) - (double)br[k + 2] - (double)ostat[k + 2]) / ostat[{IN} + k + 2];')
    S(f'for (int i = {N + 2}; i < {IN}; ++i) tr[i] = 0.0f;')
    S('}', 0)

    # ---------------- outputs of the composed path
    S(f'extern "C" __global__ void __launch_bounds__({BLOCK})', 0)
    S('kx_sur_outputs(const int n_states, const long long offsetT, const long long offset,', 0)
    S('               const double* __restrict__ p_state, const double* __restrict__ state,', 0)
    S('               const float* __restrict__ base, const float* __restrict__ resid,', 0)
    S('               const float* __restrict__ ostat, const int* __restrict__ flags,', 0)
    S('               double* __restrict__ viscosity, double* __restrict__ conductivity,', 0)
    S('               double* __restrict__ rhoD, double* __restrict__ Dkm, const double Tref)', 0)
    S('{', 0)
    S(f'const long long id = (long long)blockIdx.x * {BLOCK} + threadIdx.x;')
    S('if (id >= n_states || flags[id]) return;')
    S('const double T = Tref * state[id];')
    S('const double p = p_state[id];')
    S(f'const float* br = base + id * {IN};')
    S(f'const float* rr = resid + id * {IN};')
    S(f'viscosity[id] = exp((double)br[0] + (double)ostat[0] + (double)ostat[{IN}] * (double)rr[0]);')
    S(f'conductivity[id] = exp((double)br[1] + (double)ostat[1] + (double)ostat[{IN + 1}] * (double)rr[1]);')
    rcp_mbar()
    S(f'const double rho = p * {lit(1.0 / float(t["R"]))} / (T * rcpMbar);')
    S(f'for (int k = 0; k < {N}; ++k) {{')
    S(f'  const double D = exp((double)br[k + 2] + (double)ostat[k + 2] + (double)ostat[{IN} + k + 2] * (double)rr[k + 2]);')
    S('  rhoD[k * offset + id] = rho * D;')
    S('  if (Dkm) Dkm[k * offset + id] = D;')
    S('}')
    S('}', 0)

    # ---------------- fused inference: features + MLP (BF16 MFMA) + outputs in one pass
    dtab = np.array([t['diff'][k, pr[k]] for k in range(N)], dtype=np.float32)      # (N, 5)
    S(f'__device__ const float kx_sur_dpoly[{5 * N}] = {{' +
      ', '.join(f'{float(v)!r}f' for v in dtab.reshape(-1)) + '};', 0)
    S(f'// One state per lane for the physics part; the MLP (kx_mlp_wave.h) in {passes} pass(es) of', 0)
    S(f'// {64 // passes} states per wavefront (at hidden width 256 two passes of 32: the activations take', 0)
    S('// 128 VGPRs instead of 256 -- 2.81 ms for 4M states against 3.74 in one pass; at width 128', 0)
    S('// one pass of 64 wins, 1.53 against 1.64 ms); two waves share a SIMD (one wave\'s MFMAs overlap the other\'s', 0)
    S('// feature and GELU VALU work). LDS carries the inputs', 0)
    S('// and four floats per state (ln mu and ln lambda baselines + output means, ln T,', 0)
    S('// 1.5 ln T - ln p) between the two layouts; the diffusion baselines are recomputed per', 0)
    S('// output from the FP32 fit table. States flagged out of hull (or beyond n_states) get', 0)
    S('// no output here: the caller routes them to the plain kernel.', 0)
    S('extern "C" __global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))', 0)
    S('kx_sur_infer(const int n_states, const long long offsetT, const long long offset,', 0)
    S('             const double* __restrict__ p_state, const double* __restrict__ state,', 0)
    S('             const float* __restrict__ norm, const float* __restrict__ ostat,', 0)
    S('             const uint16_t* __restrict__ wfrag, const float* __restrict__ params,', 0)
    S('             int* __restrict__ flags, double* __restrict__ viscosity,', 0)
    S('             double* __restrict__ conductivity, double* __restrict__ rhoD,', 0)
    S('             double* __restrict__ Dkm, const double Tref,', 0)
    S('             int* __restrict__ fb_queue, int* __restrict__ fb_count)', 0)
    S('{', 0)
    S(E.POOL_MARK)
    S(f'__shared__ __attribute__((aligned(16))) uint16_t xs[4][64 * {XS}];')
    S('__shared__ __attribute__((aligned(16))) float ss[4][64 * 4];')
    S('__shared__ double rs[4][64];')
    S('__shared__ int fs[4][64];')
    S(f'__shared__ float dps[{5 * N}];')
    S(f'for (int i = threadIdx.x; i < {5 * N}; i += 256) dps[i] = kx_sur_dpoly[i];')
    S('const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;')
    S('const long long id = (long long)blockIdx.x * 256 + threadIdx.x;')
    S(f'uint16_t* xr = xs[wv] + lane * {XS};')
    S('if (id < n_states) {')
    S('const double T = Tref * state[id];')
    S('const double p = p_state[id];')
    state_model(diff=False)
    S(f'float xn[{IN}];')
    hull_and_inputs(lambda i, e: f'xn[{i}] = {e};')
    S(f'for (int i = {N + 2}; i < {IN}; ++i) xn[i] = 0.0f;')
    S(f'for (int i = 0; i < {IN}; i += 2)')
    S('  *(unsigned*)(xr + i) = kxmlp::pack2(xn[i], xn[i + 1]);')
    S('*(kxmlp::f32x4*)(ss[wv] + 4 * lane) = (kxmlp::f32x4){bl[0] + ostat[0], bl[1] + ostat[1], lT, lT15p};')
    S(f'rs[wv][lane] = p * {lit(1.0 / float(t["R"]))} / (T * (double)ysum);')
    S('fs[wv][lane] = out;')
    S('flags[id] = out;')
    S('// flagged states are queued for the plain kernel (vector atomic on global memory)')
    S('if (out && fb_queue) fb_queue[atomicAdd(fb_count, 1)] = (int)id;')
    S('} else {')
    S(f'for (int i = 0; i < {IN}; i += 2) *(unsigned*)(xr + i) = 0u;')
    S('fs[wv][lane] = 1;')
    S('}')
    S('__syncthreads();')
    S('const int j = lane & 15, g = lane >> 4;')
    S('#pragma unroll 1')
    S(f'for (int h = 0; h < {passes}; ++h) {{')
    S(f'kxmlp::f32x4 y[kxmlp::kOut / 16][{sb_n}];')
    S(f'kxmlp::forward<{XS}, {sb_n}>(wfrag, params, xs[wv] + {16 * sb_n} * h * {XS}, y, lane);')
    S('#pragma unroll')
    S(f'for (int sb = 0; sb < {sb_n}; ++sb) {{')
    S(f'  const int s = {16 * sb_n} * h + 16 * sb + j;')
    S('  const long long sid = (long long)blockIdx.x * 256 + wv * 64 + s;')
    S('  if (fs[wv][s]) continue;')
    S('  const double rho = rs[wv][s];')
    S('  const kxmlp::f32x4 q = *(const kxmlp::f32x4*)(ss[wv] + 4 * s);')
    S('  const float lT = q[2], lT15p = q[3], lT2 = lT * lT, lT3 = lT2 * lT, lT4 = lT2 * lT2;')
    S('  #pragma unroll')
    S('  for (int ob = 0; ob < kxmlp::kOut / 16; ++ob)')
    S('  #pragma unroll')
    S('  for (int e = 0; e < 4; ++e) {')
    S('    const int f = 16 * ob + 4 * g + e;')
    S(f'    if (f >= {N + 2}) continue;')
    S('    float b;')
    S('    if (f < 2) {')
    S('      b = f == 0 ? q[0] : q[1];')
    S('    } else {')
    S('      const float* c = dps + 5 * (f - 2);')
    S('      const float d = c[0] + c[1] * lT + c[2] * lT2 + c[3] * lT3 + c[4] * lT4;')
    S(f'      b = lT15p {"-" if rcp_diff else "+"} kx_lnf(d) + ostat[f];')
    S('    }')
    S(f'    const float v = b + ostat[{IN} + f] * y[ob][sb][e];')
    S('    const double D = (double)__builtin_amdgcn_exp2f(v * 1.4426950408889634f);')
    S('    if (f == 0) viscosity[sid] = D;')
    S('    else if (f == 1) conductivity[sid] = D;')
    S('    else {')
    S('      rhoD[(f - 2) * offset + sid] = rho * D;')
    S('      if (Dkm) Dkm[(f - 2) * offset + sid] = D;')
    S('    }')
    S('  }')
    S('}')
    S('}')
    S('}', 0)

    # ---------------- audit / fallback plumbing: a queue of state ids for the plain kernel
    S('// counters: [0] states queued as out of hull, [1] failed audits, [3] the whole-call')
    S('// fallback count (kx_sur_decide). The queue holds the')
    S('// audit samples a0 + i * stride, then the out-of-hull ids kx_sur_infer appends; the plain')
    S('// kernel kx_bk2_fp64_queue runs over it in place.')
    S(f'extern "C" __global__ void __launch_bounds__({BLOCK})', 0)
    S('kx_sur_queue_init(const int n_audit, const long long a0, const long long stride,', 0)
    S('                  int* __restrict__ queue, int* __restrict__ counters)', 0)
    S('{', 0)
    S(f'const long long i = (long long)blockIdx.x * {BLOCK} + threadIdx.x;')
    S('if (i < 4) counters[i] = 0;')
    S('if (i < n_audit) queue[i] = (int)(a0 + i * stride);')
    S('}', 0)
    S('// The whole-call fallback decided on the device: counters[3] = n_states when the failed')
    S('// audits exceed max_fail * n_audit, else 0; the plain kernel then runs with that count.')
    S(f'extern "C" __global__ void __launch_bounds__(64)', 0)
    S('kx_sur_decide(const int n_states, const int n_audit, const double max_fail,', 0)
    S('              int* __restrict__ counters)', 0)
    S('{', 0)
    S('if (threadIdx.x == 0)')
    S('  counters[3] = (n_audit > 0 && (double)counters[1] > max_fail * (double)n_audit) ? n_states : 0;')
    S('}', 0)
    return S.text()
