"""Transport (BK2) for large mechanisms: kx_bk2_fp64 / kx_bk2_fp64_pstate over state tiles.

The loop form with one state per lane (one lane's x, 1/eta and pair-sum arrays of N doubles
each) keeps ~5 KB per lane at 160 species in scratch; with ~10^4 lanes per XCD that working set
streams through L2 to HBM: 290 GB per 2M-state launch against 5.4 GB of inputs and outputs,
84 ms. Here a workgroup of 1024 lanes takes a tile of 64 states, one per lane, and its 16
wavefronts split the species rows (transportProps.okl:41-47, mix_transport.py:261-471):

  * LDS holds the tile's x_j, and then the diffusion row sums, then 1/eta_j, as [j][state]:
    2 N x 64 doubles (160 KB at 160 species, the whole LDS of a CU); nothing per state lives
    in scratch;
  * diffusion, each pair once: the wavefront of row k takes the pairs j > k, adds
    x_j / D'_kj to its own row sum and x_k / D'_kj to row j's sum in LDS with an atomic add
    (ds_add_f64: the order of those additions varies from run to run, within rounding).
    Rows k and 31 - k of every block of 32 share a wavefront, so the wavefronts have about
    the same number of pairs (mix_transport.py:414-471);
  * Wilke, row by row: S_k = sum_j x_j (C1_kj + C2_kj eta_k / eta_j)^2
    (mix_transport.py:297-357), and the conductivity rule's two sums (:261-294);
  * within a wavefront the row, and so every table address, is uniform: the pair fits (upper
    triangle, row k contiguous), C1/C2 and the species fits are scalar loads shared by the
    tile's 64 states;
  * mu and lambda partial sums of the 16 wavefronts meet in LDS and are added in wavefront
    order.

Measured on 2,097,152 states of the 160-species mechanism: 37.1 ms (loop form 84.3 ms). Tried:
32-state tiles with a row per half-wavefront (row-dependent vector loads of 64 B pair
records, each pair evaluated for both rows): 96-107 ms; 64-state tiles with both rows per
pair: 66.8-69.9 ms. The module exports the lanes per state of these kernels as
`<kernel>_lanes_per_state` (16): the host multiplies the state count by it to size the grid.
"""
from __future__ import annotations

import numpy as np

from . import emit_hip as E

TS = 64                      # states per tile (workgroup), one per lane
HALVES = 16                  # wavefronts per workgroup: the row classes
BLOCK = TS * HALVES          # 1024 lanes


def lds_bytes(t):
    N = int(t['n_species'])
    return 8 * (2 * N * TS + (0 if N >= 3 * HALVES else 3 * HALVES * TS))


def _poly(c, k):
    return (f'{c}[5 * {k}] + {c}[5 * {k} + 1] * lnT + {c}[5 * {k} + 2] * lnT2 + '
            f'{c}[5 * {k} + 3] * lnT3 + {c}[5 * {k} + 4] * lnT4')


def upper_fits(t):
    """The D'_kj fits of the pairs j > k, row k contiguous: pair (k, j) at
    k (N - 1) - k (k - 1) / 2 + (j - k - 1)."""
    N = int(t['n_species'])
    d = np.asarray(t['diff'], dtype=np.float64)
    return np.array([d[j, k] for k in range(N) for j in range(k + 1, N)]).reshape(-1, 5)


def emit(t, info, pstate=False, unroll=4):
    N = int(t['n_species'])
    rcp_diff = bool(int(t.get('rcp_diff', 0)))
    name = 'kx_bk2_fp64_pstate' if pstate else 'kx_bk2_fp64'
    pre = f'{name}_t'
    S = E._Src()

    def table(nm, arr):
        a = np.asarray(arr, dtype=np.float64).reshape(-1)
        S(f'__device__ const double {nm}[{a.size}] = {{' + ', '.join(repr(float(v)) for v in a) + '};', 0)
    table(f'{pre}_fit', upper_fits(t))
    table(f'{pre}_wil', np.stack([np.asarray(t['C1']), np.asarray(t['C2'])], axis=-1))
    table(f'{pre}_visc', t['visc'])
    table(f'{pre}_cond', t['cond'])
    table(f'{pre}_rcpW', t['rcpW'])
    table(f'{pre}_W', t['W'])
    S(f'extern "C" __device__ const int {name}_lanes_per_state = {HALVES};', 0)
    S(f'extern "C" __global__ void __launch_bounds__({BLOCK})', 0)
    S(f'{name}(const int n_states, const long long offsetT, const long long offset,', 0)
    S('            ' + ('const double* __restrict__ p_state,' if pstate else 'const double pressure,'), 0)
    S('            const double* __restrict__ state, double* __restrict__ conductivity,', 0)
    S('            double* __restrict__ viscosity, double* __restrict__ rhoD,', 0)
    S('            ' + ('double* __restrict__ Dkm, ' if pstate else '') + 'const double Tref,', 0)
    S('            const int* __restrict__ n_dev)', 0)
    S('{', 0)
    S('// n_dev: a device-side state count (the surrogate\'s whole-call fallback), or null')
    S('const long long n_eff = n_dev ? min((long long)n_states, (long long)max(*n_dev, 0)) : (long long)n_states;')
    S(f'const long long t0 = (long long)blockIdx.x * {TS};')
    S('if (t0 >= n_eff) return;                           // whole tile past the end')
    S(f'__shared__ double sx[{N} * {TS}], sre[{N} * {TS}];   // [j][state]: x_j, 1/eta_j')
    alias = N >= 3 * HALVES                     # the three partial sums fit over the x rows
    if not alias:
        S(f'__shared__ double red[3 * {HALVES} * {TS}];')
    S('const int s = threadIdx.x & 63;')
    S('const int h = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);   // row class: one per wave')
    S('const long long id0 = t0 + s;')
    S('const bool live = id0 < n_eff;')
    S('const long long id = live ? id0 : t0;              // lanes past the end redo the tile\'s first state, store nothing')
    S('const double p = ' + ('p_state[id];' if pstate else 'pressure;'))
    S('const double T = Tref * state[id];')
    S('const double lnT = log(T);')
    S('const double rcpT = 1.0 / T;')
    S('const double sqrT = sqrt(T);')
    S('const double lnT2 = lnT * lnT;')
    S('const double lnT3 = lnT * lnT * lnT;')
    S('const double lnT4 = lnT * lnT * lnT * lnT;')
    S('double part = 0.0;')
    S('#pragma unroll 1')
    S(f'for (int k = h; k < {N}; k += {HALVES}) {{')
    S(f'  const double y = fmax(0.0, state[id + offsetT + k * offset]) * {pre}_rcpW[k];')
    S(f'  sx[k * {TS} + s] = y;')
    S('  part += y;')
    S('}')
    S(f'sre[h * {TS} + s] = part;                          // row-class partial sums of Y_k / W_k')
    S('__syncthreads();')
    S('double ysum = 0.0;')
    S('#pragma unroll')
    S(f'for (int v = 0; v < {HALVES}; ++v) ysum += sre[v * {TS} + s];')
    S('const double Mbar = 1.0 / ysum;')
    S('__syncthreads();                                   // partials read before 1/eta overwrites them')
    NI = 2 * ((N + 2 * HALVES - 1) // (2 * HALVES))
    S('#pragma unroll 1')
    S(f'for (int k = h; k < {N}; k += {HALVES}) {{')
    S(f'  sx[k * {TS} + s] *= Mbar;')
    S(f'  sre[k * {TS} + s] = 0.0;                     // the diffusion row sums')
    S('}')
    S('__syncthreads();')
    S('const double TsqrT = T * sqrT;')
    S(f'const double rho = p * {E._literal(1.0 / float(t["R"]))} * rcpT * Mbar;')
    S('const double pMbar = p * Mbar;')
    S('// diffusion, each pair once: the wavefront of row k takes the pairs j > k and adds')
    S('// x_k / D_kj to row j with an LDS atomic (rows k and 31 - k of each block of 32 share a')
    S('// wavefront, so every wavefront has about the same number of pairs)')
    S('#pragma unroll 1')
    S(f'for (int i = 0; i < {NI}; ++i) {{')
    S(f'  const int k = {2 * HALVES} * (i >> 1) + ((i & 1) ? {2 * HALVES - 1} - h : h);')
    S(f'  if (k >= {N}) continue;')
    S(f'  const double xk = sx[k * {TS} + s];')
    S(f'  const int base = k * {N - 1} - k * (k - 1) / 2 - k - 1;   // pair (k, j) at base + j')
    S('  double sk = 0.0;')
    S(f'  #pragma unroll {unroll}')
    S(f'  for (int j = k + 1; j < {N}; ++j) {{')
    S(f'    const double* c = {pre}_fit + 5 * (base + j);')
    poly = 'c[0] + c[1] * lnT + c[2] * lnT2 + c[3] * lnT3 + c[4] * lnT4'
    S(f'    const double rD = {poly if rcp_diff else "kx_rcp(" + poly + ")"};')
    S(f'    sk += sx[j * {TS} + s] * rD;')
    S(f'    atomicAdd(sre + j * {TS} + s, xk * rD);')
    S('  }')
    S(f'  atomicAdd(sre + k * {TS} + s, sk);')
    S('}')
    S('__syncthreads();')
    S('#pragma unroll 1')
    S(f'for (int k = h; k < {N}; k += {HALVES}) {{')
    S(f'  const double D = TsqrT * (Mbar - {pre}_W[k] * sx[k * {TS} + s]) / (pMbar * sre[k * {TS} + s]);')
    S('  if (live) {')
    S('    rhoD[k * offset + id] = rho * D;')
    if pstate:
        S('    Dkm[k * offset + id] = D;')
    S('  }')
    S('}')
    S('__syncthreads();                                   // row sums read before 1/eta overwrites them')
    S('#pragma unroll 1')
    S(f'for (int k = h; k < {N}; k += {HALVES}) sre[k * {TS} + s] = kx_rcp({_poly(pre + "_visc", "k")});')
    S('__syncthreads();')
    S('double vpart = 0.0, s1 = 0.0, s2 = 0.0;')
    S('#pragma unroll 1')
    S(f'for (int k = h; k < {N}; k += {HALVES}) {{')
    S(f'  const double xk = sx[k * {TS} + s];')
    S(f'  const double e = {_poly(pre + "_visc", "k")};')
    S(f'  const double* __restrict__ pw = {pre}_wil + k * {2 * N};')
    S('  double sw = 0.0;')
    S(f'  #pragma unroll {2 * unroll}')
    S(f'  for (int j = 0; j < {N}; ++j) {{')
    S(f'    const double f = pw[2 * j] + pw[2 * j + 1] * e * sre[j * {TS} + s];')
    S(f'    sw += sx[j * {TS} + s] * f * f;')
    S('  }')
    S('  vpart += xk * e * e / sw;')
    S(f'  const double l = {_poly(pre + "_cond", "k")};')
    S('  s1 += xk * l;')
    S('  s2 += xk / l;')
    S('}')
    S('__syncthreads();                                   // every row done with x and 1/eta')
    if alias:
        S(f'double* const red = sx;                            // [3][{HALVES}][{TS}] over the x rows')
    S(f'red[(0 * {HALVES} + h) * {TS} + s] = vpart;')
    S(f'red[(1 * {HALVES} + h) * {TS} + s] = s1;')
    S(f'red[(2 * {HALVES} + h) * {TS} + s] = s2;')
    S('__syncthreads();')
    S('if (h == 0 && live) {')
    S('  double m = 0.0, a1 = 0.0, a2 = 0.0;')
    S('  #pragma unroll')
    S(f'  for (int v = 0; v < {HALVES}; ++v) {{')
    S(f'    m += red[(0 * {HALVES} + v) * {TS} + s];')
    S(f'    a1 += red[(1 * {HALVES} + v) * {TS} + s];')
    S(f'    a2 += red[(2 * {HALVES} + v) * {TS} + s];')
    S('  }')
    S('  viscosity[id] = sqrT * m;')
    S('  conductivity[id] = sqrT * (0.5 * (a1 + 1.0 / a2));')
    S('}')
    S('}', 0)
    if alias:
        assert 3 * HALVES <= N, 'the partials alias the x rows'
    return S.text()
