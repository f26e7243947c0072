"""Cooperative (several lanes per state) production-rate kernels for large mechanisms.

The one-state-per-lane loop forms (emit_rolled) keep each state's species arrays -- w, the
Gibbs terms, omega-dot: ~4 KB at 160 species -- in per-lane scratch. With 64 lanes and many
waves per CU that working set spills through L2 to HBM: the 160-species / 1,000-reaction
mechanism moves ~290 GB per 2M-state launch against 5.4 GB of inputs and outputs. Here a
state is shared by a group of LPS lanes and its arrays live in LDS:

  * species phase (species k = g, g + LPS, ...): concentrations C_k, the Gibbs terms
    g_k log2(e) and w_k = C_k exp(g_k) (FP64, as in the other forms) into LDS; group sums
    (shuffles within the group) give 1/Mbar and the third-body total C_m;
  * reaction phase (reactions r = g, g + LPS, ...): the w-form net rate
    q_r = 2^(log2 k_f - sum_reactants g log2 e) (prod_reactants w - C0^-dnu prod_products w)
    with the falloff / third-body factors of the loop forms (emit_rolled.emit_falloff), and
    omega-dot_k += nu_kr q_r as FP64 atomic adds on the state's LDS array (ds_add_f64);
  * output phase: W_k omega-dot_k and the heat release rate (group sum).

The reaction descriptors are packed to 64 bytes per reaction (three FP64 coefficients,
species indices and stoichiometric coefficients as bytes) and sorted by kind, so that the
LPS lanes of an iteration mostly take the same branch; the SPW = 64 / LPS states of a
wavefront read the same descriptors (one broadcast load serves them all).

Arithmetic as the loop forms (kinetix/core/reaction_rates.py:616-1333 rolled, in the w-form;
benchmark/okl/productionRates.okl:10-62 around it): the FP64 variant evaluates the Arrhenius
exponent and falloff in FP64; the mixed variant the Arrhenius exponent and falloff in FP32,
Gibbs terms, net rates and sums in FP64. Only the order of the omega-dot additions differs.
"""
from __future__ import annotations

import numpy as np

from . import emit_hip as E
from . import emit_rolled as RL
from . import mechanism as M

LPS = 32                 # lanes per state (16 and 64 measured slower at 160 species)
WG = 64                  # threads per workgroup (one wavefront), one per state of its block


def _u32_table(name, arr):
    a = np.asarray(arr, dtype=np.uint64).reshape(-1)
    return (f'__device__ const unsigned {name}[{max(a.size, 1)}] __attribute__((aligned(16))) = '
            f'{{{", ".join(f"{int(v)}u" for v in a) if a.size else "0u"}}};')


def packed(t):
    """Reaction descriptors sorted by kind: order (original index of each row), rxd [R][3]
    FP64 (log2 A, beta, T_a log2 e), rxf [R][4] FP32 copies, rxi [R][8] uint32 (reactant and
    product species bytes, packed metadata, third-body/PLOG index, net species bytes x2, net
    coefficients as signed bytes x2), falloff rows rxfo [3 + 10 F] / rxff [3 + 3 F] (three
    leading pad entries, so that row s of the low-pressure and Troe/SRI parameters is read as
    p[3..12] with p = table + 10 s)."""
    N, R = int(t['n_species']), int(t['n_reactions'])
    if N > 255:
        raise ValueError('packed descriptors hold species indices in bytes (N <= 255)')
    d = RL.bk1_data(t)
    meta, dbl, flt = d['meta'], d['dbl'], d['flt']
    order = sorted(range(R), key=lambda r: (int(meta[r, 0]), int(meta[r, 6]), int(meta[r, 2]),
                                            int(meta[r, 3]), r))
    rxd = np.zeros((R, 3))
    rxf = np.zeros((R, 4), dtype=np.float32)
    rxi = np.zeros((R, 8), dtype=np.uint64)
    fo, ff = [0.0] * 3, [0.0] * 3
    nfo = 0

    def bytes4(vals, fill=255):
        v = [(int(x) & 255) if x >= 0 else fill for x in vals] + [fill] * (4 - len(vals))
        return v[0] | (v[1] << 8) | (v[2] << 16) | (v[3] << 24)
    # reactant / product lists padded with species index N: a slot whose w is 1 and g is 0,
    # so every reaction reads the same number of entries with no branch

    for i, r in enumerate(order):
        kind, rev, nf, nr, nnet, dnu, mk, midx = (int(x) for x in meta[r])
        if not -8 <= dnu <= 7 or nnet > 8:
            raise ValueError(f'reaction {r + 1}: stoichiometry outside the packed range')
        fslot = 0
        if kind in (M.LINDEMANN, M.TROE, M.SRI):
            fslot = nfo
            nfo += 1
            fo += [float(v) for v in dbl[r, 3:13]]
            ff += [float(v) for v in flt[r, 3:6]]
        if nfo >= 2048:
            raise ValueError('more than 2047 falloff reactions')
        rxd[i] = dbl[r, :3]
        rxf[i, :3] = flt[r, :3]
        m = kind | (rev << 4) | (nf << 5) | (nr << 8) | (nnet << 11) | ((dnu + 8) << 15) | (mk << 19) | (fslot << 21)
        idx = r if kind == M.PLOG else midx
        nsp = [int(x) for x in d['nsp'][r]]
        nnu = [int(round(float(x))) for x in d['nnu'][r]]
        rxi[i] = (bytes4(d['fsp'][r], N), bytes4(d['rsp'][r], N), m, idx,
                  bytes4(nsp[:4]), bytes4(nsp[4:]),
                  bytes4([v & 255 for v in nnu[:4]], 0), bytes4([v & 255 for v in nnu[4:]], 0))
    mx = int(max(meta[:, 2].max(), meta[:, 3].max())) if R else 1
    return dict(order=order, rxd=rxd, rxf=rxf, rxi=rxi, rxfo=np.array(fo), rxff=np.array(ff, dtype=np.float32),
                data=d, mx=mx, mnet=int(meta[:, 4].max()) if R else 1)


def emit_bk1(t, info, variant='fp64', lps=LPS, prefetch=None, layout='state', pad=0):
    """kx_bk1_<variant> ('fp64' or 'mixed_valu') for one mechanism, LPS lanes per state.
    prefetch (default: the mixed variant only): request the next reaction's descriptors one
    iteration ahead. Measured at 160 species / 1,000 reactions, 2M states: mixed 21.5 -> 18.9
    ms with it, FP64 23.6 -> 25.9 ms (the longer FP64 live ranges)."""
    prefetch = (variant != 'fp64') if prefetch is None else prefetch
    N, R = int(t['n_species']), int(t['n_reactions'])
    if 64 % lps:
        raise ValueError('lanes per state must divide 64')
    spw = 64 // lps
    # LDS layout of the per-state arrays: 'interleaved' [k][state], or 'state' [state][k] with
    # each state's row padded by `pad` doubles
    if layout == 'interleaved':
        SP, ks, base = N + 1, spw, 'st'
    else:
        SP, ks = N + 1 + pad, 1
        base = f'st * {SP}'
    name = f'kx_bk1_{variant}'
    pre = f'{name}_c'
    mixed = variant != 'fp64'
    P = packed(t)
    d = P['data']
    S = E._Src()
    tab = RL._table
    for line in (tab(f'{pre}_rxd', P['rxd']),
                 tab(f'{pre}_rxf', P['rxf'], 'float').replace('] = {', '] __attribute__((aligned(16))) = {', 1),
                 _u32_table(f'{pre}_rxi', P['rxi']), tab(f'{pre}_rxfo', P['rxfo']),
                 tab(f'{pre}_rxff', P['rxff'], 'float'),
                 tab(f'{pre}_setoff', d['setoff'], 'int'), tab(f'{pre}_setsp', d['setsp'], 'int'),
                 tab(f'{pre}_sete', d['sete']),
                 tab(f'{pre}_plo', d['plo'], 'int'), tab(f'{pre}_plP', d['plP']), tab(f'{pre}_pllnP', d['pllnP']),
                 tab(f'{pre}_plrcp', d['plrcp']), tab(f'{pre}_plko', d['plko'], 'int'),
                 tab(f'{pre}_plnk', d['plnk'], 'int'), tab(f'{pre}_plk', d['plk']),
                 tab(f'{pre}_rcpW', t['rcpW']), tab(f'{pre}_W', t['W']), tab(f'{pre}_tmid', t['T_mid']),
                 tab(f'{pre}_glo', t['g_lo']), tab(f'{pre}_ghi', t['g_hi']),
                 tab(f'{pre}_hlo', t['h_lo']), tab(f'{pre}_hhi', t['h_hi'])):
        S(line, 0)
    S(f'__device__ __forceinline__ double {pre}_pow2(double y)', 0)
    S('{', 0)
    if mixed:
        S('const double n = __builtin_rint(y);')
        S('const float m = __builtin_amdgcn_exp2f((float)(y - n));')
        S('return __builtin_ldexp((double)m, (int)n);')
    else:
        S('return exp2(y);')
    S('}', 0)
    S(f'__device__ __forceinline__ double {pre}_gsum(double v)', 0)
    S('{', 0)
    S(f'// sum over the {lps} lanes of a state group (xor butterfly inside the group)')
    m = lps // 2
    while m:
        S(f'v += __shfl_xor(v, {m}, {lps});')
        m //= 2
    S('return v;')
    S('}', 0)
    for line in RL.plog_helpers(pre):
        S(line, 0)
    sync = ('__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront"); __builtin_amdgcn_wave_barrier(); '
            '__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");')
    S(f'extern "C" __global__ void __launch_bounds__({WG})', 0)
    S(f'{name}(const int n_states, const long long offsetT, const long long offset,', 0)
    S('            const double pressure_R, const double pressure_,', 0)
    S('            const double* __restrict__ state, double* __restrict__ rates, const double Tref)', 0)
    S('{', 0)
    S(f'// the workgroup (one wavefront) covers 64 states, {spw} at a time in groups')
    S(f'// of {lps} lanes; per state in LDS: C_k, w_k, g_k log2 e, omega-dot_k, one row per state')
    S('// (' + ('[state][k]: the 32 lanes of a group gather species k of 32 reactions, spread over' if layout != 'interleaved'
               else '[k][state]: the groups read the same species of the same reaction, from') + '')
    S('// ' + ('twice the banks of a [k][state] interleave: 2-4 % faster at 160 species' if layout != 'interleaved'
              else 'neighbouring banks') + '); slot N is the padding entry of the reactant lists')
    S(f'__shared__ double sC[{SP} * {spw}], sW[{SP} * {spw}], sG[{SP} * {spw}], sD[{SP} * {spw}];')
    S(f'const int g = threadIdx.x & {lps - 1}, st = threadIdx.x / {lps};')
    S(f'double* __restrict__ const C = sC + {base};')
    S(f'double* __restrict__ const Wv = sW + {base};')
    S(f'double* __restrict__ const G = sG + {base};')
    S(f'double* __restrict__ const Wd = sD + {base};')
    S(f'if (g == 0) {{ Wv[{N * ks}] = 1.0; G[{N * ks}] = 0.0; }}')
    S('const double P = pressure_, lnP = log(P);   // PLOG reactions (uniform per launch)')
    S('(void)P; (void)lnP;')
    S('// States are interleaved across workgroups (at iteration it the resident workgroups work on')
    S('// neighbouring states) and the workgroup order is made XCD-aware: consecutive workgroups')
    S('// are dispatched round-robin to the 8 XCDs, each with its own L2, so neighbouring states go')
    S('// to workgroups b, b + 8, b + 16, ... of one XCD. A cache line of Y (16 states) is then')
    S('// fetched once into one L2 and a line of rates written whole, instead of once per XCD.')
    S('const unsigned G8 = gridDim.x / 8, b = blockIdx.x;')
    S('const unsigned xb = b < 8 * G8 ? (b % 8) * G8 + b / 8 : b;')
    S('#pragma unroll 1')
    S(f'for (int it = 0; it < {64 // spw}; ++it) {{')
    S(f'const long long id0 = ((long long)it * gridDim.x + xb) * {spw} + st;')
    S('const bool live = id0 < n_states;')
    S('const long long id = live ? id0 : 0;          // groups past the end recompute state 0, store nothing')
    S('const double T = Tref * state[id];')
    S('const double rcpT = 1.0 / T;')
    S('const double lnT = log(T);')
    S('const double T2 = T * T, T3 = T * T * T, T4 = T * T * T * T;')
    S('const double log2T = lnT * 1.4426950408889634;')
    S('const float Tf = (float)T, rcpTf = (float)rcpT, log2Tf = (float)log2T;')
    S('(void)Tf; (void)rcpTf; (void)log2Tf;')
    S('double part = 0.0;')
    S(f'for (int k = g; k < {N}; k += {lps}) {{')
    S(f'  const double c = fmax(0.0, state[id + offsetT + k * offset]) * {pre}_rcpW[k];')
    S(f'  C[k * {ks}] = c;')
    S('  part += c;')
    S('}')
    S(f'const double rho = pressure_R * rcpT / {pre}_gsum(part);')
    S('double cm = 0.0;')
    S(f'for (int k = g; k < {N}; k += {lps}) {{')
    S(f'  const double c = C[k * {ks}] * rho;')
    S(f'  const double* a = (T > {pre}_tmid[k] ? {pre}_ghi : {pre}_glo) + 7 * k;')
    S('  const double gk = a[0] * (1.0 - lnT) + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4 + a[5] * rcpT - a[6];')
    S(f'  C[k * {ks}] = c;')
    S(f'  G[k * {ks}] = gk * 1.4426950408889634;')
    S(f'  Wv[k * {ks}] = c * exp(gk);')
    S(f'  Wd[k * {ks}] = 0.0;')
    S('  cm += c;')
    S('}')
    S(f'const double Cm = {pre}_gsum(cm);')
    S(f'const double C0 = {repr(float(t["one_atm_R"]))} * rcpT;')
    S(f'const double rcpC0 = {repr(float(t["R_one_atm"]))} * T;')
    S(sync)
    if prefetch:
        # software pipeline: the descriptors of the lane's next reaction are requested at the
        # top of each iteration (their L2 latency then overlaps this reaction's arithmetic)
        S(f'const int r0 = g < {R} ? g : 0;')
        S(f'uint4 n0 = *(const uint4*)({pre}_rxi + 8 * r0), n1 = *(const uint4*)({pre}_rxi + 8 * r0 + 4);')
        if mixed:
            S(f'float4 nc = *(const float4*)({pre}_rxf + 4 * r0);')
        else:
            S(f'double nc0 = {pre}_rxd[3 * r0], nc1 = {pre}_rxd[3 * r0 + 1], nc2 = {pre}_rxd[3 * r0 + 2];')
    S(f'for (int r = g; r < {R}; r += {lps}) {{')
    if prefetch:
        S('  const uint4 i0 = n0, i1 = n1;')
        if mixed:
            S('  const float4 fr = nc;')
        else:
            S('  const double dr0 = nc0, dr1 = nc1, dr2 = nc2;')
        S(f'  const int rn = r + {lps} < {R} ? r + {lps} : r;')
        S(f'  n0 = *(const uint4*)({pre}_rxi + 8 * rn);')
        S(f'  n1 = *(const uint4*)({pre}_rxi + 8 * rn + 4);')
        if mixed:
            S(f'  nc = *(const float4*)({pre}_rxf + 4 * rn);')
        else:
            S(f'  nc0 = {pre}_rxd[3 * rn]; nc1 = {pre}_rxd[3 * rn + 1]; nc2 = {pre}_rxd[3 * rn + 2];')
    else:
        S(f'  const uint4 i0 = *(const uint4*)({pre}_rxi + 8 * r);')
        S(f'  const uint4 i1 = *(const uint4*)({pre}_rxi + 8 * r + 4);')
        if mixed:
            S(f'  const float4 fr = *(const float4*)({pre}_rxf + 4 * r);')
        else:
            S(f'  const double dr0 = {pre}_rxd[3 * r], dr1 = {pre}_rxd[3 * r + 1], dr2 = {pre}_rxd[3 * r + 2];')
    S('  const unsigned mt = i0.z;')
    S('  const int kind = mt & 15, nf = (mt >> 5) & 7, nr = (mt >> 8) & 7, nnet = (mt >> 11) & 15;')
    S('  const int dnu = (int)((mt >> 15) & 15) - 8, mk = (mt >> 19) & 3, fs = mt >> 21;')
    if mixed:
        S(f'  double ex = kind == {M.PLOG} ? {pre}_plog((int)i0.w, P, lnP, log2T, rcpT)')
        S('                              : (double)(fr.x + fr.y * log2Tf - fr.z * rcpTf);   // FP32 Arrhenius exponent')
    else:
        S(f'  double ex = kind == {M.PLOG} ? {pre}_plog((int)i0.w, P, lnP, log2T, rcpT) : dr0 + dr1 * log2T - dr2 * rcpT;')
    S('  double fwd = 1.0, rev = 1.0;')
    for i in range(P['mx']):
        S(f'  {{ const int k = ((i0.x >> {8 * i}) & 255) * {ks}; ex -= G[k]; fwd *= Wv[k]; }}')
    for i in range(P['mx']):
        S(f'  rev *= Wv[((i0.y >> {8 * i}) & 255) * {ks}];')
    S('  (void)nf; (void)nr;')
    S(f'  double F = {pre}_pow2(ex);')
    S(f'  if (kind != 0 && kind != {M.PLOG}) {{')
    S('    double M;')
    S('    if (mk == 3) {')
    S('      M = Cm;')
    S(f'      for (int j = {pre}_setoff[i0.w]; j < {pre}_setoff[i0.w + 1]; ++j) M += {pre}_sete[j] * C[{pre}_setsp[j] * {ks}];')
    S('    } else {')
    S(f'      M = mk == 2 ? C[i0.w * {ks}] : Cm;')
    S('    }')
    S(f'    const double* dp = {pre}_rxfo + 10 * fs;      // dp[3..12]: falloff row fs')
    S(f'    const float* fp = {pre}_rxff + 3 * fs;        // fp[3..5]')
    S('    (void)dp; (void)fp;')
    RL.emit_falloff(S, pre, mixed)
    S('  }')
    S('  double q;')
    S('  if (mt & 16) {')
    S('    double c0 = 1.0;')
    S('    for (int i = 0; i < (dnu < 0 ? -dnu : dnu); ++i) c0 *= dnu < 0 ? C0 : rcpC0;')
    S('    q = F * (fwd - c0 * rev);')
    S('  } else {')
    S('    q = F * fwd;')
    S('  }')
    for i in range(P['mnet']):
        w_sp, w_nu, sh = ('i1.x', 'i1.z', 8 * i) if i < 4 else ('i1.y', 'i1.w', 8 * (i - 4))
        S(f'  if ({i} < nnet) atomicAdd(&Wd[(({w_sp} >> {sh}) & 255) * {ks}], (double)(signed char)(({w_nu} >> {sh}) & 255) * q);')
    S('}')
    S(sync)
    S('double sh = 0.0;')
    S(f'for (int k = g; k < {N}; k += {lps}) {{')
    S(f'  const double wd = Wd[k * {ks}];')
    S(f'  if (live) rates[id + offsetT + k * offset] = {pre}_W[k] * wd;')
    S(f'  const double* a = (T > {pre}_tmid[k] ? {pre}_hhi : {pre}_hlo) + 6 * k;')
    S('  sh += wd * (a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4 + a[5] * rcpT);')
    S('}')
    S(f'sh = {pre}_gsum(sh);')
    S(f'if (live && g == 0) rates[id] = {repr(-float(t["R"]))} * T * sh;')
    S(sync)
    S('}')
    S('}', 0)
    return S.text()
