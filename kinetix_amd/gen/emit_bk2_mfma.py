"""FP64 transport kernel with the Wilke viscosity sum on FP64 MFMA (kx_bk2_fp64_mfma).

Same arguments, outputs and FP64 arithmetic class as kx_bk2_fp64 (the reference's transport
kernel, benchmark/okl/transportProps.okl:11-47); only the O(N^2) Wilke mixing sum of
kinetix_viscosity (kinetix/core/mix_transport.py:297-357, the k/j double loop at :337-345)
changes form. With eta_k = sqrt(mu_k / sqrt(T)) from its fit and the generated Wilke factors
Phi_kj^(1/2) = C1_kj + C2_kj eta_k / eta_j,

    S_k = sum_j x_j Phi_kj
        = sum_j C1_kj^2 x_j + eta_k sum_j (2 C1_kj C2_kj) (x_j / eta_j)
                            + eta_k^2 sum_j C2_kj^2 (x_j / eta_j^2),

exactly: three products of constant N x N matrices with per-state vectors. Over the 64 states
of a wavefront they are three GEMMs (N x N) x (N x 64), run on v_mfma_f64_16x16x4_f64 with
FP64 accumulation. Every term is positive (C1, C2 > 0, x >= 0, eta > 0), so the reordering
changes the sum by rounding only.

Per wavefront:
  * phase 1, one state per lane (VALU): mole fractions, the symmetric diffusion pair sums and
    rho*D_k, then the conductivity rule -- the kx_bk2_fp64 code;
  * phase 2, per block of 16 states: B operands (x_j, x_j/eta_j, x_j/eta_j^2 for 4 species
    j x 16 states) are formed directly in the MFMA lane layout, lane l taking species
    4*ks + l/16 of state l%16: it reads Y_j of that state (an L1/L2 hit), and ln T and Mbar
    of the state from the wavefront's LDS slot. The A operands (C1^2, 2 C1 C2, C2^2 in
    fragment order, ~86 KB for GRI-3.0) sit in LDS, shared by the workgroup's 8 waves;
  * the accumulators hold S_k in the result layout (lane l, element r: species
    16 m + 4 r + l/16 of state l%16). Each lane adds x_k eta_k^2 / S_k over its 16 species,
    two xor-shuffles sum the four lane groups, and the lane that owns the state keeps it.

Lane layouts of v_mfma_f64_16x16x4_f64, measured on the device by kx_selftest_mfma_f64_layout:
A[i][k] in lane i + 16 k, B[k][j] in lane j + 16 k, and D[i][j] in element i / 4 of lane
j + 16 (i % 4) -- not the FP32 16x16x4 result layout (element i % 4 of lane j + 16 (i / 4)).
"""
from __future__ import annotations

import numpy as np

from . import emit_hip as E

WAVES = 8                  # waves per workgroup: 512 threads share the LDS tables


def tiles(N):
    """(m-tiles of 16 species rows, K-steps of 4 species)."""
    return (N + 15) // 16, (N + 3) // 4


def a_fragments(t):
    """A operands in fragment order [ks][m][g][lane]: lane l of fragment (ks, m, g) holds
    M_g[16 m + l % 16][4 ks + l // 16] with M_0 = C1^2, M_1 = 2 C1 C2, M_2 = C2^2 (zero
    outside N x N)."""
    N = int(t['n_species'])
    C1, C2 = np.asarray(t['C1'], np.float64), np.asarray(t['C2'], np.float64)
    mats = (C1 * C1, 2.0 * C1 * C2, C2 * C2)
    M, KS = tiles(N)
    out = np.zeros((KS, M, 3, 64))
    for ks in range(KS):
        for m in range(M):
            for lane in range(64):
                k, j = 16 * m + lane % 16, 4 * ks + lane // 16
                if k < N and j < N:
                    for g in range(3):
                        out[ks, m, g, lane] = mats[g][k, j]
    return out


def species_table(t):
    """Per species: the viscosity fit (5 coefficients) and 1/W_k, [N][6]."""
    N = int(t['n_species'])
    tab = np.zeros((N, 6))
    tab[:, :5] = np.asarray(t['visc'], np.float64)
    tab[:, 5] = np.asarray(t['rcpW'], np.float64)
    return tab


def lds_bytes(t):
    N = int(t['n_species'])
    M, KS = tiles(N)
    return 8 * (KS * M * 3 * 64 + N * 6 + WAVES * 2 * 64)


def _lits(a):
    return ', '.join(E._literal(float(v)) for v in np.asarray(a).reshape(-1))


def emit(t, info):
    N = int(t['n_species'])
    M, KS = tiles(N)
    block = 64 * WAVES
    A = a_fragments(t)
    nA = A.size
    W = t['W']
    S = E._Src()
    S('typedef double kx_doublex4 __attribute__((ext_vector_type(4)));', 0)
    S(f'__device__ const double kx_bk2_wilke_A[{nA}] = {{{_lits(A)}}};', 0)
    S(f'__device__ const double kx_bk2_wilke_sp[{N * 6}] = {{{_lits(species_table(t))}}};', 0)
    S('__device__ double kx_bk2_mfma_sink[64];   // stores of lanes past the last state', 0)
    S(f'extern "C" __global__ void __launch_bounds__({block})', 0)
    S('kx_bk2_fp64_mfma(const int n_states, const long long offsetT, const long long offset,', 0)
    S('            const double pressure,', 0)
    S('            const double* __restrict__ state, double* __restrict__ conductivity,', 0)
    S('            double* __restrict__ viscosity, double* __restrict__ rhoD,', 0)
    S('            const double Tref, const int* __restrict__ n_dev)', 0)
    S('{', 0)
    S('// n_dev (optional): a state count known only on the device')
    S('const long long nn = n_dev ? min((long long)n_states, (long long)*n_dev) : (long long)n_states;')
    S(f'if ((long long)blockIdx.x * {block} >= nn) return;   // whole workgroup past the end')
    S(E.POOL_MARK)
    S(f'__shared__ double wA[{nA}];')
    S(f'__shared__ double wsp[{N * 6}];')
    S(f'__shared__ double wst[{WAVES}][2][64];   // per wave: ln T and Mbar of its states')
    for line in E.lds_stage('wA', 'kx_bk2_wilke_A', nA, block):
        S(line)
    for line in E.lds_stage('wsp', 'kx_bk2_wilke_sp', N * 6, block):
        S(line)
    S('__syncthreads();')
    S('const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;')
    S('const int g16 = lane >> 4, c16 = lane & 15;')
    S(f'const long long id = (long long)blockIdx.x * {block} + threadIdx.x;')
    S('const long long base = id - lane;                 // the wavefront\'s first state')
    S('if (base >= nn) return;                           // whole wavefront past the end')
    S('// lanes past the end recompute the wavefront\'s first state (valid data for the MFMAs)')
    S('// and store nothing')
    S('const bool live = id < nn;')
    S('const long long sid = live ? id : base;')
    S('const double p = pressure;')
    E._emit_bk2_prelude(S, t, 'sid')
    E._emit_diffusion_pairs(S, t)
    S('const double TsqrT = T * sqrT;')
    S(f'const double rho = p * {E.lit(1.0 / float(t["R"]))} * rcpT * Mbar;')
    S('const double pMbar = p * Mbar;')
    # lanes past the end store into a sink slot of their own instead of branching around the
    # stores (a branch there splits the straight-line code and the register allocator then
    # spills ~10^4 VGPRs)
    S('double* const rD = live ? rhoD + id : kx_bk2_mfma_sink + lane;')
    S('const long long ro = live ? offset : 0;')
    for k in range(N):
        S(f'{{ const double D = TsqrT * (Mbar - {E.lit(W[k])} * x{k}) / (pMbar * s{k});')
        S(f'  rD[{k} * ro] = rho * D; }}')
    E._emit_conductivity(S, t)
    S('*(live ? conductivity + id : kx_bk2_mfma_sink + lane) = lambda;')
    S('__builtin_amdgcn_sched_barrier(0);')
    S('// ---- viscosity: the Wilke sum as three GEMMs on FP64 MFMA, 16 states at a time')
    S('wst[wv][0][lane] = lnT;')
    S('wst[wv][1][lane] = Mbar;')
    S('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); __builtin_amdgcn_wave_barrier(); '
      '__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");')
    S('double vis = 0.0;')
    S('#pragma unroll 1')
    S('for (int sb = 0; sb < 4; ++sb) {')
    S('  // an opaque zero per block: keeps the ~170 operand loads below inside the loop instead of')
    S('  // hoisted in front of it (they do not depend on sb), where they would pin ~340 VGPRs')
    S('  int z0 = 0;')
    S('  asm volatile("" : "+v"(z0));')
    S('  const int ln = lane + z0, gg = g16 + z0;')
    S('  const int sl = 16 * sb + c16;                   // this lane\'s state in the B and D layouts')
    S('  const double lt = wst[wv][0][sl], mb = wst[wv][1][sl];')
    S('  const double lt2 = lt * lt, lt3 = lt2 * lt, lt4 = lt3 * lt;')
    S('  const long long cs = base + sl < nn ? base + sl : base;')
    S('  const double* __restrict__ ys = state + cs + offsetT;')
    S(f'  kx_doublex4 acc[{M}][3];')
    S(f'  for (int m = 0; m < {M}; ++m) for (int g = 0; g < 3; ++g) acc[m][g] = (kx_doublex4){{0.0, 0.0, 0.0, 0.0}};')
    for ks in range(KS):
        S(f'  {{ // K-step {ks}: species {4 * ks}..{min(4 * ks + 3, N - 1)}')
        S(f'    const int j = {4 * ks} + gg;')
        guard = 4 * ks + 3 >= N
        if guard:
            S(f'    const bool ok = j < {N};')
            S('    const int jj = ok ? j : 0;')
        else:
            S('    const int jj = j;')
        S('    const double* c = &wsp[6 * jj];')
        S('    const double eta = c[0] + c[1] * lt + c[2] * lt2 + c[3] * lt3 + c[4] * lt4;')
        S('    const double re = kx_rcp(eta);')
        xs = 'fmax(0.0, ys[jj * offset]) * c[5] * mb'
        S(f'    const double u = {"ok ? " + xs + " : 0.0" if guard else xs};')
        S('    const double v = u * re, w = v * re;')
        for m in range(M):
            for g, b in enumerate(('u', 'v', 'w')):
                off = ((ks * M + m) * 3 + g) * 64
                S(f'    acc[{m}][{g}] = __builtin_amdgcn_mfma_f64_16x16x4f64(wA[{off} + ln], {b}, acc[{m}][{g}], 0, 0, 0);')
        S('  }')
    S('  // S_k = D0 + eta_k (D1 + eta_k D2) for species 16 m + 4 r + g16 of state sl')
    S('  double part = 0.0;')
    for m in range(M):
        for r in range(4):
            kmax = 16 * m + 4 * r + 3       # largest k this (m, r) reaches
            S('  {')
            S(f'    const int k = {16 * m + 4 * r} + gg;')
            if kmax >= N:
                S(f'    const bool ok = k < {N};')
                S('    const int kk = ok ? k : 0;')
            else:
                S('    const int kk = k;')
            S('    const double* c = &wsp[6 * kk];')
            S('    const double eta = c[0] + c[1] * lt + c[2] * lt2 + c[3] * lt3 + c[4] * lt4;')
            S(f'    const double Sk = acc[{m}][0][{r}] + eta * (acc[{m}][1][{r}] + eta * acc[{m}][2][{r}]);')
            S('    const double xk = fmax(0.0, ys[kk * offset]) * c[5] * mb;')
            term = 'xk * eta * eta / Sk'
            S(f'    part += {"ok ? " + term + " : 0.0" if kmax >= N else term};')
            S('  }')
    S('  part += __shfl_xor(part, 16);')
    S('  part += __shfl_xor(part, 32);')
    S('  if (g16 == sb) vis = part;                       // lane 16 sb + c16 owns state sl')
    S('}')
    S('*(live ? viscosity + id : kx_bk2_mfma_sink + lane) = sqrT * vis;')
    S('}', 0)
    return S.text()
