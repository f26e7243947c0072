"""The surrogate's plain pass, cooperative: kx_bk2_fp64_queue with several lanes per state.

kx_sur_transport runs the plain FP64 transport kernel over a queue of state ids -- the 1 %
audit samples, then the out-of-hull states -- reading each state in place and writing its
results over the surrogate's. A few 10^4 states cannot fill the GPU one state per lane: that
pass lasts one full BK2 lane program (~47k instructions) at one wavefront per SIMD, ~0.37 ms
with the rho*D audit, whatever the count. Here LPS = 16 lanes share a state, each taking the
species rows k = q, q + 16, ... of every row-wise sum:

  * mole fractions and eta_k (species-parallel, group sums by xor shuffles), x_k and 1/eta_k
    into the state's LDS rows;
  * diffusion row sums s_k = sum_{j != k} x_j / D'_kj over j ascending -- each pair is
    evaluated for both of its rows instead of once (the symmetric form needs the other row's
    sum) -- from the same fits (transportProps.okl:41-47, mix_transport.py:414-471);
  * Wilke S_k = sum_j x_j (C1_kj + C2_kj eta_k / eta_j)^2 per own row and the group sum of
    x_k eta_k^2 / S_k (mix_transport.py:297-357); the conductivity rule's two sums
    (mix_transport.py:261-294) likewise;
  * the audit of the surrogate's mu, lambda and rho*D_k (read before they are overwritten).

Mechanism tables (triangular diffusion fits, C1, C2, the viscosity and conductivity fits,
1/W, W) are staged in LDS once per workgroup (~105 KB for GRI-3.0); with the states' x and
1/eta rows the workgroup of 1024 lanes (64 states) fits the 160 KB of one CU, four
wavefronts per SIMD. The workgroups stride over the queue 64 slots at a time; the host caps
the grid at the workgroups the GPU holds at once (one per CU): a grid over the longest
possible queue (every state out of hull) dispatched ~66k workgroups for 4M states, each
allocating its 156 KB of LDS only to return. The module exports the lanes per state as the
int `kx_bk2q_lanes`: its presence tells the host to launch the capped, strided grid.

Measured on 41,943 queued states (GRI-3.0): 293 us, against ~370 us for the one-lane pass.
Its bound is LDS bandwidth: every lane reads its own pair fit and x_j per ordered pair,
~220 KB per state. A row-split form (one state per lane, rows spread over 8 waves, table
coefficients as wave-uniform scalar loads) moved that traffic off LDS but ran 359 us: the
~21 KB of coefficients each wave streams miss the scalar cache, and every group of pairs
waits out an L2 round trip.
"""
from __future__ import annotations

import numpy as np

from . import emit_hip as E
from . import emit_rolled as RL

LPS = 16
BLOCK = 1024


def tables(t):
    """Triangular pair fits [N(N-1)/2][5] (pair (k, j), k > j, at k(k-1)/2 + j), C1 and C2
    [k][j], and per species [12]: viscosity fit, conductivity fit, 1/W, W."""
    N = int(t['n_species'])
    d = np.asarray(t['diff'])
    tri = np.array([d[k, j] for k in range(1, N) for j in range(k)]).reshape(-1, 5)
    sp = np.zeros((N, 12))
    sp[:, 0:5] = np.asarray(t['visc'])
    sp[:, 5:10] = np.asarray(t['cond'])
    sp[:, 10] = np.asarray(t['rcpW'])
    sp[:, 11] = np.asarray(t['W'])
    return tri, np.asarray(t['C1'], np.float64), np.asarray(t['C2'], np.float64), sp


def lds_bytes(t, lps=LPS, block=BLOCK):
    N = int(t['n_species'])
    return 8 * (5 * N * (N - 1) // 2 + 2 * N * N + 12 * N + 2 * N * (block // lps))


def emit(t, info, lps=LPS, block=BLOCK, unroll=8):
    N = int(t['n_species'])
    if block % lps or 64 % lps:
        raise ValueError('lanes per state must divide the wavefront and the workgroup')
    spb = block // lps                       # states per pass of the workgroup
    tri, C1, C2, sp = tables(t)
    nt, nc = tri.size, N * N
    o_c1, o_c2, o_sp = nt, nt + nc, nt + 2 * nc
    ntab = o_sp + sp.size
    pre = 'kx_bk2q'
    allv = np.concatenate([tri.reshape(-1), C1.reshape(-1), C2.reshape(-1), sp.reshape(-1)])
    S = E._Src()
    S(RL._table(f'{pre}_tab', allv), 0)
    S(f'__device__ __forceinline__ double {pre}_gsum(double v)', 0)
    S('{', 0)
    m = lps // 2
    while m:
        S(f'v += __shfl_xor(v, {m}, {lps});')
        m //= 2
    S('return v;')
    S('}', 0)
    S(f'extern "C" __device__ const int kx_bk2q_lanes = {lps};', 0)
    S(f'extern "C" __global__ void __launch_bounds__({block})', 0)
    S('kx_bk2_fp64_queue(const int n_audit,', 0)
    S('            const int* __restrict__ queue, int* __restrict__ counters,', 0)
    S('            const int* __restrict__ flags, const double tol, const double tol_rhoD,', 0)
    S('            const long long offsetT, const long long offset,', 0)
    S('            const double* __restrict__ p_state,', 0)
    S('            const double* __restrict__ state, double* __restrict__ conductivity,', 0)
    S('            double* __restrict__ viscosity, double* __restrict__ rhoD,', 0)
    S('            double* __restrict__ Dkm, const double Tref)', 0)
    S('{', 0)
    S('// queue entries [0, n_audit) are audit samples, then counters[0] out-of-hull states;')
    S('// counters: [0] out-of-hull states queued, [1] failed audits')
    S('const long long valid = (long long)n_audit + max(counters[0], 0);')
    S(f'if ((long long)blockIdx.x * {spb} >= valid) return;   // whole workgroup past the queue')
    S(f'__shared__ double tab[{ntab}];')
    for line in E.lds_stage('tab', f'{pre}_tab', ntab, block):
        S(line)
    S(f'__shared__ double sx[{N} * {spb}], sre[{N} * {spb}];   // [k][state]: x_k, 1/eta_k')
    S('__syncthreads();')
    S(f'const int q = threadIdx.x & {lps - 1}, st = threadIdx.x / {lps};')
    S('double* const X = sx + st;')
    S('double* const RE = sre + st;')
    S(f'const double* const Cv = tab + {o_sp};                  // [k][12]: visc, cond, 1/W, W')
    S('// grid-stride over the queue: the grid is capped at the workgroups the GPU holds at once')
    S(f'for (long long slot0 = (long long)blockIdx.x * {spb}; slot0 < valid;')
    S(f'     slot0 += (long long)gridDim.x * {spb}) {{')
    S('const long long slot = slot0 + st;')
    S('const bool live = slot < valid;')
    S('const long long id = queue[live ? slot : slot0];     // dead groups redo a valid state, store nothing')
    S('const double p = p_state[id];')
    S('const double T = Tref * state[id];')
    S('const double lnT = log(T);')
    S('const double rcpT = 1.0 / T;')
    S('const double sqrT = sqrt(T);')
    S('const double lnT2 = lnT * lnT;')
    S('const double lnT3 = lnT * lnT * lnT;')
    S('const double lnT4 = lnT * lnT * lnT * lnT;')
    S('double part = 0.0;')
    S(f'for (int k = q; k < {N}; k += {lps}) {{')
    S('  const double xk = fmax(0.0, state[id + offsetT + k * offset]) * Cv[12 * k + 10];')
    S(f'  X[k * {spb}] = xk;')
    S('  part += xk;')
    S('}')
    S(f'const double Mbar = 1.0 / {pre}_gsum(part);')
    S(f'for (int k = q; k < {N}; k += {lps}) {{')
    S('  const double* c = Cv + 12 * k;')
    S(f'  X[k * {spb}] *= Mbar;')
    S(f'  RE[k * {spb}] = kx_rcp(c[0] + c[1] * lnT + c[2] * lnT2 + c[3] * lnT3 + c[4] * lnT4);')
    S('}')
    S('// the state\'s rows are written by the lanes of its group, all in this wavefront')
    S('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); __builtin_amdgcn_wave_barrier(); '
      '__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");')
    S('const double TsqrT = T * sqrT;')
    S(f'const double rho = p * {E._literal(1.0 / float(t["R"]))} * rcpT * Mbar;')
    S('const double pMbar = p * Mbar;')
    S('const bool audit = live && slot < n_audit && flags[id] == 0;')
    S('double bad = 0.0;                                  // a rho*D_k of the audit is off')
    S('double vpart = 0.0, s1 = 0.0, s2 = 0.0;')
    S('#pragma unroll 1')
    S(f'for (int k = q; k < {N}; k += {lps}) {{')
    S(f'  const double xk = X[k * {spb}];')
    S('  // diffusion: s_k over j ascending, the pair fit of (max, min) in the triangular table')
    S('  double s = 0.0;')
    S(f'  #pragma unroll {unroll}')
    S(f'  for (int j = 0; j < {N}; ++j) {{')
    S('    if (j == k) continue;')
    S('    const int pi = k > j ? k * (k - 1) / 2 + j : j * (j - 1) / 2 + k;')
    S('    const double* c = tab + 5 * pi;')
    if bool(int(t.get('rcp_diff', 0))):
        S('    const double rD = c[0] + c[1] * lnT + c[2] * lnT2 + c[3] * lnT3 + c[4] * lnT4;')
    else:
        S('    const double rD = kx_rcp(c[0] + c[1] * lnT + c[2] * lnT2 + c[3] * lnT3 + c[4] * lnT4);')
    S(f'    s += X[j * {spb}] * rD;')
    S('  }')
    S('  const double* cw = Cv + 12 * k;')
    S('  const double D = TsqrT * (Mbar - cw[11] * xk) / (pMbar * s);')
    S('  const double v = rho * D;')
    S('  if (audit && !(fabs(rhoD[k * offset + id] / v - 1.0) <= tol_rhoD)) bad = 1.0;')
    S('  if (live) {')
    S('    rhoD[k * offset + id] = v;')
    S('    if (Dkm) Dkm[k * offset + id] = D;')
    S('  }')
    S('  // Wilke, row k')
    S('  const double e = cw[0] + cw[1] * lnT + cw[2] * lnT2 + cw[3] * lnT3 + cw[4] * lnT4;')
    S('  double sw = 0.0;')
    S(f'  #pragma unroll {unroll}')
    S(f'  for (int j = 0; j < {N}; ++j) {{')
    S(f'    const double f = tab[{o_c1} + {N} * k + j] + tab[{o_c2} + {N} * k + j] * e * RE[j * {spb}];')
    S(f'    sw += X[j * {spb}] * f * f;')
    S('  }')
    S('  vpart += xk * e * e / sw;')
    S('  // conductivity')
    S('  const double l = cw[5] + cw[6] * lnT + cw[7] * lnT2 + cw[8] * lnT3 + cw[9] * lnT4;')
    S('  s1 += xk * l;')
    S('  s2 += xk / l;')
    S('}')
    S(f'const double mu = sqrT * {pre}_gsum(vpart);')
    S(f'const double lambda = sqrT * (0.5 * ({pre}_gsum(s1) + 1.0 / {pre}_gsum(s2)));')
    S(f'bad = {pre}_gsum(bad);')
    S('if (live && q == 0) {')
    S('  if (audit) {   // audit of a state the surrogate answered')
    S('    const double ev = fabs(viscosity[id] / mu - 1.0);')
    S('    const double ec = fabs(conductivity[id] / lambda - 1.0);')
    S('    if (!(ev <= tol) || !(ec <= tol) || bad > 0.0) atomicAdd(counters + 1, 1);')
    S('  }')
    S('  viscosity[id] = mu;')
    S('  conductivity[id] = lambda;')
    S('}')
    S('// the next pass rewrites this state group\'s x and 1/eta rows')
    S('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); __builtin_amdgcn_wave_barrier(); '
      '__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");')
    S('}')
    S('}', 0)
    return S.text()
