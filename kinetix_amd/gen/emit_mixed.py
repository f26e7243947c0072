"""Mixed-precision BK1 kernel (kx_bk1_mixed) for gfx950.

Precision split (per state):
  * Arrhenius exponentials: argument in FP64, split into integer + fraction, the fraction
    through the FP32 hardware exp2 (v_exp_f32), rescaled exactly in FP64 (so no FP32 range
    limits and ~1.5e-7 relative error instead of |argument| * 6e-8);
  * falloff (Lindemann/Troe/SRI) corrections and third-body factors applied in FP32;
  * Gibbs exponentials / equilibrium constants, the forward-minus-reverse net rate and the
    final per-species sums in FP64;
  * stoichiometric accumulation wdot = q . nu on FP32 MFMA (v_mfma_f32_16x16x4_f32, FP32
    accumulation), per 64-state wave tile, 256-state workgroup tile.

MFMA mapping, per wave: D[species, state] += nu^T[species, reaction] Q^T[reaction, state].
Reactions are processed in chunks of 16 (4 MFMA K-steps of 4); each lane writes the FP32 net
rates of its own state for the chunk into LDS (one 64 B row per state) and reads back the
B operand for state blocks of 16. The nu operand (A) is a per-lane FP32 table, stored only
for (chunk, 16-species block) tiles that contain a non-zero coefficient: species are ordered
by how many reactions they take part in and reactions by the species blocks they touch, so
GRI-3.0 needs 48 of the 84 tiles. Reactions are split into two blocks whose FP32
accumulators are combined in FP64 in the epilogue (through LDS, transposed to one row per
state).

Lane layout of v_mfma_f32_16x16x4_f32 (checked on the device by kx_selftest_mfma_layout):
A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in element i%4 of lane j + 16(i/4).
"""
from __future__ import annotations

import numpy as np

from . import emit_hip as E

CHUNK = 16          # reactions per LDS staging round (4 K-steps of 4)
MB = 16             # species per MFMA row block
NBLK = 2            # reaction blocks folded in FP64
QSTRIDE = 20        # floats per state row of the q staging buffer (16 + pad)
WSTRIDE = 68        # floats per state row of the transposed accumulator tile (64 + pad)
WAVES = 4           # waves per workgroup (256-state tile)


def plan(t):
    net = t['nu_r'] - t['nu_f']
    R, N = net.shape
    active = [k for k in range(N) if np.any(net[:, k] != 0)]
    freq = np.array([np.count_nonzero(net[:, k]) for k in active])
    order = [active[i] for i in np.argsort(-freq, kind='stable')]
    nmb = (len(order) + MB - 1) // MB
    pos = {k: i for i, k in enumerate(order)}

    def key(r):
        blocks = sorted({pos[k] // MB for k in np.nonzero(net[r])[0] if k in pos})
        return (len(blocks), tuple(blocks), r)

    rx = sorted(range(R), key=key)
    chunks = [rx[i:i + CHUNK] for i in range(0, R, CHUNK)]
    tiles = []                       # (chunk index, mblock)
    for c, ch in enumerate(chunks):
        for m in range(nmb):
            sp = order[m * MB:(m + 1) * MB]
            if np.any(net[np.ix_(ch, sp)] != 0):
                tiles.append((c, m))
    table = np.zeros((len(tiles), 64, 4), dtype=np.float32)
    for ti, (c, m) in enumerate(tiles):
        ch = chunks[c]
        for lane in range(64):
            g, i = lane >> 4, lane & 15
            sidx = m * MB + i
            if sidx >= len(order):
                continue
            for s in range(4):
                slot = 4 * g + s
                if slot < len(ch):
                    table[ti, lane, s] = net[ch[slot], order[sidx]]
    split = (len(chunks) + NBLK - 1) // NBLK
    blocks = [list(range(b, min(b + split, len(chunks)))) for b in range(0, len(chunks), split)]
    return {'order': order, 'pos': pos, 'nmb': nmb, 'chunks': chunks, 'tiles': tiles,
            'table': table, 'blocks': blocks}


def mfma_per_state(t):
    p = plan(t)
    # each tile: 4 K-steps x 4 state blocks of 16 per 64-state wave
    return {'mfma_instr_per_wave': len(p['tiles']) * 16,
            'mfma_flops_per_state': len(p['tiles']) * 16 * 2048 // 64,
            'nu_tiles': len(p['tiles']), 'nu_tiles_dense': len(p['chunks']) * p['nmb']}


DEFAULT_OPTS = {'sync': 'wave', 'kstep_outer': True, 'wform': True}


def emit(t, info, opts=None):
    o = dict(DEFAULT_OPTS)
    o.update(opts or {})
    N = int(t['n_species'])
    p = plan(t)
    em = E._RateEmitter(t, info)
    em.sched_window = 0
    if o['sync'] == 'wave':
        # each wave only touches its own LDS rows: a wave-scope fence + wave barrier orders
        # the cross-lane LDS traffic without stalling the other waves of the workgroup
        SYNC = ('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); __builtin_amdgcn_wave_barrier(); '
                '__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");')
    else:
        SYNC = '__syncthreads();'
    S = E._Src()
    tab = p['table'].reshape(-1)
    nt = len(p['tiles'])
    S(f'__device__ const float kx_bk1_mixed_nu[{max(nt, 1) * 256}] = {{'
      + ', '.join(f'{v:.1f}f' for v in tab) + '};', 0)
    S('typedef float kx_floatx4 __attribute__((ext_vector_type(4)));', 0)
    S('__device__ __forceinline__ double kx_pow2(double y)', 0)
    S('{', 0)
    S('// 2^y with the FP32 hardware exp2 on the fractional part and exact FP64 scaling')
    S('const double n = __builtin_rint(y);')
    S('const float m = __builtin_amdgcn_exp2f((float)(y - n));')
    S('return __builtin_ldexp((double)m, (int)n);')
    S('}', 0)
    S(f'extern "C" __global__ void __launch_bounds__({64 * WAVES})', 0)
    S('kx_bk1_mixed(const int n_states, const long long offsetT, const long long offset,', 0)
    S('             const double pressure_R, const double pressure_,', 0)
    S('             const double* __restrict__ state, double* __restrict__ rates, const double Tref)', 0)
    S('{', 0)
    S(E.POOL_MARK)
    S(f'__shared__ __attribute__((aligned(16))) float qs[{WAVES}][64][{QSTRIDE}];')
    S(f'__shared__ __attribute__((aligned(16))) float wt[{WAVES}][64][{WSTRIDE}];')
    # the nu tiles (A operands) are copied once per workgroup into LDS and read from there
    S(f'__shared__ __attribute__((aligned(16))) float nu_s[{max(nt, 1) * 256}];')
    S(f'for (int i = threadIdx.x; i < {max(nt, 1) * 64}; i += {64 * WAVES})')
    S('  *(kx_floatx4*)&nu_s[4 * i] = *(const kx_floatx4*)&kx_bk1_mixed_nu[4 * i];')
    S('__syncthreads();')
    S('const int lane = threadIdx.x & 63;')
    S('const int wv = threadIdx.x >> 6;')
    S(f'const long long id = (long long)blockIdx.x * {64 * WAVES} + threadIdx.x;')
    S('const bool live = id < n_states;')
    S('// lanes past the end still take part in the wave-wide MFMAs with q = 0')
    S('const double T = live ? Tref * state[id] : 1500.0;')
    S('const double rcpT = 1.0 / T;')
    S('const double lnT = log(T);')
    S(f'const double log2T = lnT * {E.lit(1.4426950408889634)};')
    S('const double T2 = T * T;')
    S('const double T3 = T * T * T;')
    S('const double T4 = T * T * T * T;')
    S('const double rcpT2 = rcpT * rcpT;')
    S('const float Tf = (float)T;')
    S('const float rcpTf = (float)rcpT;')
    S('const double P = pressure_, lnP = log(P);   // PLOG reactions (uniform per launch)')
    S('(void)P; (void)lnP; (void)rcpTf; (void)Tf; (void)rcpT2; (void)T2;')
    rcpW = t['rcpW']
    for k in range(N):
        S(f'double c{k} = live ? fmax(0.0, state[id + offsetT + {k} * offset]) * {E.lit(rcpW[k])} : 0.0;')
    S('const double rcpMbar = ' + ' + '.join(f'c{k}' for k in range(N)) + ';')
    S('const double rho = live ? pressure_R * rcpT / rcpMbar : 0.0;')
    for k in range(N):
        S(f'c{k} *= rho;')
    sidx = em.emit_prelude_w(S, 'c') if o['wform'] else em.emit_prelude(S, 'c')
    nmb = p['nmb']
    S(f'kx_floatx4 acc[{nmb}][4];')
    S(f'for (int m = 0; m < {nmb}; ++m) for (int b = 0; b < 4; ++b) acc[m][b] = (kx_floatx4){{0.f, 0.f, 0.f, 0.f}};')
    tiles_by_chunk = {}
    for ti, (c, m) in enumerate(p['tiles']):
        tiles_by_chunk.setdefault(c, []).append((ti, m))
    seen = {}
    pos = p['pos']
    for bi, blk in enumerate(p['blocks']):
        S(f'// ======== reaction block {bi}: FP32 MFMA accumulation')
        for c in blk:
            ch = p['chunks'][c]
            S(f'// ---- chunk {c}: reactions ' + ' '.join(str(r + 1) for r in ch))
            for r in ch:
                if o['wform']:
                    em.emit_reaction_w(S, r, sidx)
                else:
                    em.emit_reaction(S, r, sidx, seen, 'c', mixed=True)
                S(f'const float f{r} = (float)q{r};')
            vals = [f'f{r}' for r in ch] + ['0.f'] * (CHUNK - len(ch))
            S('{')
            S('  float* row = &qs[wv][lane][0];')
            for j in range(0, CHUNK, 4):
                S(f'  *(kx_floatx4*)(row + {j}) = (kx_floatx4){{{vals[j]}, {vals[j + 1]}, {vals[j + 2]}, {vals[j + 3]}}};')
            S('}')
            S(SYNC)
            S('{')
            for (ti, m) in tiles_by_chunk.get(c, []):
                S(f'  const kx_floatx4 a{m} = *(const kx_floatx4*)&nu_s[({ti} * 64 + lane) * 4];')
            if o['kstep_outer']:
                # consecutive MFMAs write different accumulators (no back-to-back dependency)
                for nb in range(4):
                    S(f'  const kx_floatx4 b{nb} = *(const kx_floatx4*)&qs[wv][{nb * 16} + (lane & 15)][4 * (lane >> 4)];')
                for s_ in range(4):
                    for nb in range(4):
                        for (ti, m) in tiles_by_chunk.get(c, []):
                            S(f'  acc[{m}][{nb}] = __builtin_amdgcn_mfma_f32_16x16x4f32(a{m}[{s_}], b{nb}[{s_}], acc[{m}][{nb}], 0, 0, 0);')
            else:
                S('  for (int nb = 0; nb < 4; ++nb) {')
                S('    const kx_floatx4 b = *(const kx_floatx4*)&qs[wv][nb * 16 + (lane & 15)][4 * (lane >> 4)];')
                for (ti, m) in tiles_by_chunk.get(c, []):
                    for s_ in range(4):
                        S(f'    acc[{m}][nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(a{m}[{s_}], b[{s_}], acc[{m}][nb], 0, 0, 0);')
                S('  }')
            S('}')
            S(SYNC)
        # end of block: the first block parks its transposed tile in LDS until the epilogue
        if bi == 0:
            S('// transpose block-0 accumulators to one row per state: D[species, state] -> wt[state][species]')
            S(f'for (int m = 0; m < {nmb}; ++m) for (int nb = 0; nb < 4; ++nb)')
            S('  *(kx_floatx4*)&wt[wv][nb * 16 + (lane & 15)][m * 16 + 4 * (lane >> 4)] = acc[m][nb];')
            if len(p['blocks']) > 1:
                S(f'for (int m = 0; m < {nmb}; ++m) for (int b = 0; b < 4; ++b) acc[m][b] = (kx_floatx4){{0.f, 0.f, 0.f, 0.f}};')
    S('// ======== epilogue: per-species sums of the blocks in FP64, mass rates, heat release')
    S(SYNC)
    act = em.active
    for k in act:
        S(f'double wd{k} = (double)wt[wv][lane][{pos[k]}];')
    if len(p['blocks']) > 1:
        S(SYNC)
        S(f'for (int m = 0; m < {nmb}; ++m) for (int nb = 0; nb < 4; ++nb)')
        S('  *(kx_floatx4*)&wt[wv][nb * 16 + (lane & 15)][m * 16 + 4 * (lane >> 4)] = acc[m][nb];')
        S(SYNC)
        for k in act:
            S(f'wd{k} += (double)wt[wv][lane][{pos[k]}];')
    S('if (!live) return;')
    W = t['W']
    for k in range(N):
        val = f'{E.lit(W[k])} * wd{k}' if k in act else '0.0'
        S(f'rates[id + offsetT + {k} * offset] = {val};')
    S('double sum_h = 0.0;')
    E._emit_h_RT_sum(S, t, act, wvar='wd')
    S(f'rates[id] = {E.lit(-float(t["R"]))} * T * sum_h;')
    S('}', 0)
    return S.text()
