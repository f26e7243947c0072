"""Loop ("rolled") forms of the FP64 kernels for large mechanisms (gfx950 HIP).

The straight-line kernels of emit_hip keep every species quantity in registers and expand
every pair term; past ~100 species that code no longer fits the register file or the
instruction cache (and takes hours to compile). The rolled transport kernels keep the
per-state species arrays in per-lane private memory (scratch, interleaved across lanes so
every array access of a wavefront is one contiguous line), read the mechanism's
coefficients as uniform (scalar-cache) loads from __device__ tables, and loop over species
and pairs. The production-rate kernels of these mechanisms are emit_coop's (several lanes
per state, arrays in LDS); this module keeps their reaction tables (bk1_data), the shared
falloff code (emit_falloff) and the one-state loop core of the kinetix_species_rates device
routine (emit_bk1_core). Same arithmetic and summation order as the straight-line forms:

  BK2: kinetix/core/mix_transport.py:261-471 and benchmark/okl/transportProps.okl:11-47
       (symmetric diffusion pair loop, Wilke viscosity, conductivity mixing rule);
  BK1: kinetix/core/reaction_rates.py:616-1333 and benchmark/okl/productionRates.okl:10-62
       (rolled reaction loop; reaction descriptors in tables).
"""
from __future__ import annotations

import math

import numpy as np

from . import emit_hip as E
from . import mechanism as M

ROLL_N = 96          # mechanisms with more species than this get the loop forms
BK2_BLOCK = 8        # j-species held in registers by the blocked BK2 pair loops (8: 2 waves/SIMD)
BK2_KUNROLL = 2      # k-loop unroll (overlaps the next pair coefficients' loads)


def _table(name, arr, ctype='double'):
    a = np.asarray(arr).reshape(-1)
    if ctype == 'double':
        body = ', '.join(repr(float(v)) for v in a)
    elif ctype == 'float':
        body = ', '.join(f'{float(np.float32(v))!r}f' for v in a)
    else:
        body = ', '.join(str(int(v)) for v in a)
    return f'__device__ const {ctype} {name}[{max(a.size, 1)}] = {{{body if a.size else "0"}}};'


def _poly(tab, row):
    return (f'({tab}[5 * ({row})] + {tab}[5 * ({row}) + 1] * lnT + {tab}[5 * ({row}) + 2] * lnT2 + '
            f'{tab}[5 * ({row}) + 3] * lnT3 + {tab}[5 * ({row}) + 4] * lnT4)')


def emit_bk2(t, info, pstate=False):
    N = int(t['n_species'])
    rcp_diff = bool(int(t.get('rcp_diff', 0)))
    name = 'kx_bk2_fp64_pstate' if pstate else 'kx_bk2_fp64'
    pre = f'{name}_r'
    S = E._Src()
    d = np.asarray(t['diff'])
    pairs = np.array([d[k, j] for k in range(1, N) for j in range(k)])
    S(_table(f'{pre}_rcpW', t['rcpW']), 0)
    S(_table(f'{pre}_W', t['W']), 0)
    S(_table(f'{pre}_visc', t['visc']), 0)
    S(_table(f'{pre}_cond', t['cond']), 0)
    S(_table(f'{pre}_diff', pairs), 0)
    S(_table(f'{pre}_C1', t['C1']), 0)
    S(_table(f'{pre}_C2', t['C2']), 0)
    S(f'extern "C" __global__ void __launch_bounds__({E.BLOCK})', 0)
    if pstate:
        S(f'{name}(const int n_states, const long long offsetT, const long long offset,', 0)
        S('            const double* __restrict__ p_state,', 0)
    else:
        S(f'{name}(const int n_states, const long long offsetT, const long long offset,', 0)
        S('            const double pressure,', 0)
    S('            const double* __restrict__ state, double* __restrict__ conductivity,', 0)
    S('            double* __restrict__ viscosity, double* __restrict__ rhoD,', 0)
    S('            ' + ('double* __restrict__ Dkm, ' if pstate else '') + 'const double Tref,', 0)
    S('            const int* __restrict__ n_dev)', 0)
    S('{', 0)
    S(f'const long long id = (long long)blockIdx.x * {E.BLOCK} + threadIdx.x;')
    S('if (id >= n_states || (n_dev && id >= *n_dev)) return;')
    S('const double p = ' + ('p_state[id];' if pstate else 'pressure;'))
    S('const double T = Tref * state[id];')
    S('const double lnT = log(T);')
    S('const double rcpT = 1.0 / T;')
    S('const double sqrT = sqrt(T);')
    S('const double lnT2 = lnT * lnT;')
    S('const double lnT3 = lnT * lnT * lnT;')
    S('const double lnT4 = lnT * lnT * lnT * lnT;')
    S(f'double x[{N}], s[{N}], re[{N}];       // per-lane private arrays')
    S('double rcpMbar = 0.0;')
    S('#pragma unroll 1')
    S(f'for (int k = 0; k < {N}; ++k) {{')
    S(f'  x[k] = fmax(0.0, state[id + offsetT + k * offset]) * {pre}_rcpW[k];')
    S('  rcpMbar += x[k];')
    S('}')
    S('const double Mbar = 1.0 / rcpMbar;')
    S('#pragma unroll 1')
    S(f'for (int k = 0; k < {N}; ++k) {{ x[k] *= Mbar; s[k] = 0.0; }}')
    S('// ---- mixture-averaged diffusion: symmetric pair sums S_k = sum_{j!=k} X_j / D_kj,')
    S('// register-blocked: a block of 16 j-species (x_j and their S_j sums) stays in registers')
    S('// while k runs over the later species (one private load and one update per k)')
    JB = BK2_BLOCK

    def rd(pidx):
        if rcp_diff:
            return _poly(pre + '_diff', pidx)
        return f'kx_rcp({_poly(pre + "_diff", pidx)})'    # positive pair fits (file header)
    for j0 in range(0, N, JB):
        nb = min(JB, N - j0)
        S('{')
        for i in range(nb):
            S(f'  const double xj{i} = x[{j0 + i}];')
            S(f'  double sj{i} = 0.0;')
        # diagonal block: pairs (k, j) with j0 <= j < k < j0 + nb, unrolled
        for kk in range(1, nb):
            k = j0 + kk
            S(f'  {{ double sk = 0.0;')
            for i in range(kk):
                pidx = k * (k - 1) // 2 + j0 + i
                S(f'    {{ const double r_ = {rd(str(pidx))}; sk += xj{i} * r_; sj{i} += xj{kk} * r_; }}')
            S(f'    sj{kk} += sk; }}')
        if j0 + nb < N:
            S(f'  #pragma unroll {BK2_KUNROLL}')
            S(f'  for (int k = {j0 + nb}; k < {N}; ++k) {{')
            S('    const double xk = x[k];')
            S(f'    const int pb = k * (k - 1) / 2 + {j0};')
            S('    double sk = 0.0;')
            for i in range(nb):
                S(f'    {{ const double r_ = {rd(f"pb + {i}")}; sk += xj{i} * r_; sj{i} += xk * r_; }}')
            S('    s[k] += sk;')
            S('  }')
        for i in range(nb):
            S(f'  s[{j0 + i}] += sj{i};')
        S('}')
    S('const double TsqrT = T * sqrT;')
    S(f'const double rho = p * {repr(1.0 / float(t["R"]))} * rcpT * Mbar;')
    S('const double pMbar = p * Mbar;')
    S('#pragma unroll 1')
    S(f'for (int k = 0; k < {N}; ++k) {{')
    S(f'  const double D = TsqrT * (Mbar - {pre}_W[k] * x[k]) / (pMbar * s[k]);')
    S('  rhoD[k * offset + id] = rho * D;')
    if pstate:
        S('  Dkm[k * offset + id] = D;')
    S('}')
    S('// ---- viscosity (Wilke), register-blocked over 16 j-species: S_k = sum_j X_j Phi_kj')
    S(f'double sv[{N}], ev[{N}];')
    S('#pragma unroll 1')
    S(f'for (int k = 0; k < {N}; ++k) {{ const double e = {_poly(pre + "_visc", "k")}; ev[k] = e; re[k] = 1.0 / e; sv[k] = 0.0; }}')
    for j0 in range(0, N, JB):
        nb = min(JB, N - j0)
        S('{')
        for i in range(nb):
            S(f'  const double xw{i} = x[{j0 + i}], rw{i} = re[{j0 + i}];')
        S(f'  #pragma unroll {BK2_KUNROLL}')
        S(f'  for (int k = 0; k < {N}; ++k) {{')
        S('    const double e = ev[k];')
        S(f'    const double* c1 = {pre}_C1 + k * {N} + {j0};')
        S(f'    const double* c2 = {pre}_C2 + k * {N} + {j0};')
        S('    double acc = 0.0;')
        for i in range(nb):
            S(f'    {{ const double f = c1[{i}] + c2[{i}] * e * rw{i}; acc += xw{i} * f * f; }}')
        S('    sv[k] += acc;')
        S('  }')
        S('}')
    S('double vis = 0.0;')
    S('#pragma unroll 1')
    S(f'for (int k = 0; k < {N}; ++k) vis += x[k] * ev[k] * ev[k] / sv[k];')
    S('viscosity[id] = sqrT * vis;')
    S('// ---- conductivity')
    S('double sum1 = 0.0, sum2 = 0.0;')
    S('#pragma unroll 1')
    S(f'for (int k = 0; k < {N}; ++k) {{')
    S(f'  const double l = {_poly(pre + "_cond", "k")};')
    S('  sum1 += x[k] * l;')
    S('  sum2 += x[k] / l;')
    S('}')
    S('conductivity[id] = sqrT * (0.5 * (sum1 + 1.0 / sum2));')
    S('}', 0)
    return S.text()


def bk1_data(t):
    """Reaction descriptors of the loop-form BK1 as arrays (one row per reaction): the
    reactant/product molecule lists, net stoichiometry, metadata, FP64/FP32 rate
    coefficients, efficiency sets and PLOG points."""
    N, R = int(t['n_species']), int(t['n_reactions'])
    nu_f, nu_r = np.asarray(t['nu_f']).reshape(R, N), np.asarray(t['nu_r']).reshape(R, N)
    kind, rev = np.asarray(t['kind']), np.asarray(t['reversible'])
    collider, has_eff = np.asarray(t['collider']), np.asarray(t['has_eff'])
    eff = np.asarray(t['eff']).reshape(R, N)
    L2E = 1.4426950408889634
    # unique efficiency sets (sparse: species, eff - 1)
    sets, set_of = [], {}
    for r in range(R):
        if has_eff[r]:
            key = tuple(eff[r])
            if key not in set_of:
                set_of[key] = len(sets)
                sets.append(key)
    set_off, set_sp, set_e = [0], [], []
    for key in sets:
        for k, e in enumerate(key):
            if e != 1.0:
                set_sp.append(k)
                set_e.append(e - 1.0)
        set_off.append(len(set_sp))
    # per reaction: expanded reactant/product lists (<= 4 each), net stoichiometry (<= 8)
    MX, MS = 4, 8
    f_sp = np.full((R, MX), -1); r_sp = np.full((R, MX), -1)
    n_sp = np.full((R, MS), -1); n_nu = np.zeros((R, MS))
    meta = np.zeros((R, 8), dtype=np.int64)     # kind, rev, nf, nr, nnet, dnu, mkind, midx
    dbl = np.zeros((R, 13))                     # log2A, beta, TaL2E, l2ratio, dbeta, dTaL2E, troe(4), (sri a b rc d e in 6..10)
    for r in range(R):
        fl = [k for k in range(N) for _ in range(int(nu_f[r, k]))]
        rl = [k for k in range(N) for _ in range(int(nu_r[r, k]))]
        if len(fl) > MX or len(rl) > MX:
            raise ValueError(f'reaction {r}: more than {MX} reactant/product molecules')
        f_sp[r, :len(fl)] = fl
        r_sp[r, :len(rl)] = rl
        net = [(k, int(nu_r[r, k] - nu_f[r, k])) for k in range(N) if nu_r[r, k] != nu_f[r, k]]
        if len(net) > MS:
            raise ValueError(f'reaction {r}: more than {MS} species')
        for i, (k, v) in enumerate(net):
            n_sp[r, i], n_nu[r, i] = k, v
        if has_eff[r]:
            mk, mi = 3, set_of[tuple(eff[r])]
        elif collider[r] >= 0:
            mk, mi = 2, int(collider[r])
        else:
            mk, mi = 1, 0
        meta[r] = (kind[r], rev[r], len(fl), len(rl), len(net), int(sum(v for _, v in net)), mk, mi)
        A, b, Ta = float(t['A'][r]), float(t['beta'][r]), float(t['Ta'][r])
        row = [math.log2(A), b, Ta * L2E] + [0.0] * 10
        if kind[r] in (M.LINDEMANN, M.TROE, M.SRI):
            row[3] = math.log2(float(t['A0'][r]) / A)
            row[4] = float(t['beta0'][r]) - b
            row[5] = (float(t['Ta0'][r]) - Ta) * L2E
        if kind[r] == M.TROE:
            row[6:10] = [float(v) for v in np.asarray(t['troe']).reshape(R, 4)[r]]
        if kind[r] == M.SRI:
            row[6:11] = [float(v) for v in np.asarray(t['sri']).reshape(R, 5)[r]]
        dbl[r] = row
    # FP32 copies of the Arrhenius and low-pressure exponent coefficients (mixed variant)
    flt = dbl[:, :6].astype(np.float32)
    # PLOG points: per reaction a range of points (ascending P), per point a range of
    # Arrhenius rates (log2 A, beta, Ta log2 e) summed at that pressure
    em = E._RateEmitter(t, None)
    plo, plP, pllnP, plrcp, plko, plnk, plk = [0], [], [], [], [], [], []
    for r in range(R):
        if kind[r] == M.PLOG:
            pts = em.plog_points(r)
            for i, (P, rates) in enumerate(pts):
                plP.append(P)
                pllnP.append(math.log(P))
                plrcp.append(1.0 / (math.log(pts[i + 1][0]) - math.log(P)) if i + 1 < len(pts) else 0.0)
                plko.append(len(plk) // 3)
                plnk.append(len(rates))
                for A, b, Ta in rates:
                    plk += [math.log2(A), b, Ta * L2E]
        plo.append(len(plP))
    return dict(flt=flt, plo=np.array(plo), plP=np.array(plP), pllnP=np.array(pllnP),
                plrcp=np.array(plrcp), plko=np.array(plko), plnk=np.array(plnk), plk=np.array(plk),
                fsp=f_sp, rsp=r_sp, nsp=n_sp, nnu=n_nu, meta=meta, dbl=dbl,
                setoff=np.array(set_off), setsp=np.array(set_sp), sete=np.array(set_e),
                nsets=len(sets), MX=MX, MS=MS)


def _bk1_tables(t, pre):
    """Declarations of the loop-form BK1 tables (bk1_data), prefixed pre."""
    d = bk1_data(t)
    flt, plo, plP, pllnP, plrcp = d['flt'], d['plo'], d['plP'], d['pllnP'], d['plrcp']
    plko, plnk, plk, f_sp, r_sp, n_sp, n_nu = d['plko'], d['plnk'], d['plk'], d['fsp'], d['rsp'], d['nsp'], d['nnu']
    meta, dbl, set_off, set_sp, set_e = d['meta'], d['dbl'], d['setoff'], d['setsp'], d['sete']
    out = [_table(f'{pre}_flt', flt, 'float'),
           _table(f'{pre}_plo', np.array(plo), 'int'), _table(f'{pre}_plP', np.array(plP)),
           _table(f'{pre}_pllnP', np.array(pllnP)), _table(f'{pre}_plrcp', np.array(plrcp)),
           _table(f'{pre}_plko', np.array(plko), 'int'), _table(f'{pre}_plnk', np.array(plnk), 'int'),
           _table(f'{pre}_plk', np.array(plk)),_table(f'{pre}_fsp', f_sp, 'int'), _table(f'{pre}_rsp', r_sp, 'int'),
           _table(f'{pre}_nsp', n_sp, 'int'), _table(f'{pre}_nnu', n_nu),
           _table(f'{pre}_meta', meta, 'int'), _table(f'{pre}_dbl', dbl),
           _table(f'{pre}_setoff', np.array(set_off), 'int'), _table(f'{pre}_setsp', np.array(set_sp), 'int'),
           _table(f'{pre}_sete', np.array(set_e)),
           _table(f'{pre}_rcpW', t['rcpW']), _table(f'{pre}_W', t['W']),
           _table(f'{pre}_tmid', t['T_mid']), _table(f'{pre}_glo', t['g_lo']), _table(f'{pre}_ghi', t['g_hi']),
           _table(f'{pre}_hlo', t['h_lo']), _table(f'{pre}_hhi', t['h_hi'])]
    return out, d['nsets'], d['MX'], d['MS']


def plog_helpers(pre):
    """Device helpers of the loop-form PLOG evaluation (tables of prefix pre), as lines."""
    H = E._Src()
    H('// log2 of the summed Arrhenius rates of PLOG point i (FP64)', 0)
    H(f'__device__ double {pre}_plog_point(int i, double log2T, double rcpT)', 0)
    H('{', 0)
    H(f'const double* c = {pre}_plk + 3 * {pre}_plko[i];')
    H(f'const int nk = {pre}_plnk[i];')
    H('if (nk == 1) return c[0] + c[1] * log2T - c[2] * rcpT;')
    H('double s = 0.0;')
    H('for (int j = 0; j < nk; ++j) s += exp2(c[3 * j] + c[3 * j + 1] * log2T - c[3 * j + 2] * rcpT);')
    H('return log2(s);')
    H('}', 0)
    H('// log2 k_f of PLOG reaction r: linear in ln P between the bracketing points, clamped', 0)
    H('// outside them (reaction_rates.py:1064-1175); P is uniform over the launch', 0)
    H(f'__device__ double {pre}_plog(int r, double P, double lnP, double log2T, double rcpT)', 0)
    H('{', 0)
    H(f'const int o0 = {pre}_plo[r], o1 = {pre}_plo[r + 1];')
    H(f'if (P <= {pre}_plP[o0]) return {pre}_plog_point(o0, log2T, rcpT);')
    H(f'if (P > {pre}_plP[o1 - 1]) return {pre}_plog_point(o1 - 1, log2T, rcpT);')
    H('int i = o0;')
    H(f'while (i + 2 < o1 && P > {pre}_plP[i + 1]) ++i;')
    H(f'const double L1 = {pre}_plog_point(i, log2T, rcpT), L2 = {pre}_plog_point(i + 1, log2T, rcpT);')
    H(f'return L1 + (L2 - L1) * ((lnP - {pre}_pllnP[i]) * {pre}_plrcp[i]);')
    H('}', 0)
    return H.text().split('\n')


def emit_falloff(S, pre, mixed):
    """Third-body / falloff factor of one reaction, shared by the loop forms: multiplies F by
    M (three-body) or by Pr/(1+Pr) * F_cent-term (Lindemann, Troe, SRI), as the straight-line
    forms do. In scope: kind, M (the reaction's third-body concentration), dp (a row whose
    [3..12] hold the low-pressure Arrhenius offsets and Troe/SRI parameters), fp (FP32 copies,
    [3..5]), T, rcpT, log2T, Tf, rcpTf, log2Tf and F; kind is neither elementary nor PLOG.
    Helpers of prefix pre (pow2) precede it."""
    S(f'    if (kind == {M.THREE_BODY}) {{')
    S('      F *= M;')
    S('    } else {')
    if mixed:
        S('      // clamped below FLT_MAX: Pr / (1 + Pr) must stay finite (-> 1) in FP32')
        S('      // the FP32 exponent is capped at 100 so exp2f stays finite (no 0 * inf at M = 0)')
        S('      const float Pr = fminf((float)M * __builtin_amdgcn_exp2f(fminf(fp[3] + fp[4] * log2Tf - fp[5] * rcpTf, 100.0f)), 1e30f);')
        S('      const float Pr1 = Pr * __builtin_amdgcn_rcpf(1.0f + Pr);')
        S('      float Fc = 1.0f;')
        S(f'      if (kind == {M.TROE}) {{')
        S('        const float A_ = (float)dp[6], rT3 = (float)dp[7], rT1 = (float)dp[8];')
        S('        const float term3 = dp[9] >= 0.0 ? __builtin_amdgcn_exp2f((float)(-dp[9] * 1.4426950408889634) * rcpTf) : 0.0f;')
        S('        const float logFc = 0.3010299956639812f * __builtin_amdgcn_logf((1.0f - A_) * '
          '__builtin_amdgcn_exp2f(-rT3 * 1.4426950408889634f * Tf) + A_ * __builtin_amdgcn_exp2f(-rT1 * 1.4426950408889634f * Tf) + term3);')
        S('        const float logPr = 0.3010299956639812f * __builtin_amdgcn_logf(Pr + 1e-37f);')
        S('        const float tc = -0.4f - 0.67f * logFc + logPr;')
        S('        const float tf = tc * __builtin_amdgcn_rcpf(0.75f - 1.27f * logFc - 0.14f * tc + 1e-37f);')
        S('        Fc = __builtin_amdgcn_exp2f(3.3219280948873622f * logFc * __builtin_amdgcn_rcpf(1.0f + tf * tf));')
        S(f'      }} else if (kind == {M.SRI}) {{')
        S('        const float logPr = 0.3010299956639812f * __builtin_amdgcn_logf(Pr + 1e-37f);')
        S('        Fc = (float)dp[9] * powf(Tf, (float)dp[10]) * powf((float)dp[6] * expf((float)(-dp[7]) * rcpTf) + '
          'expf((float)(-dp[8]) * Tf), __builtin_amdgcn_rcpf(1.0f + logPr * logPr));')
        S('      }')
        S('      F *= (double)(Pr1 * Fc);')
    else:
        S(f'      const double Pr = M * {pre}_pow2(dp[3] + dp[4] * log2T - dp[5] * rcpT);')
        S('      const double Pr1 = Pr / (1.0 + Pr);')
        S('      double Fc = 1.0;')
        S(f'      if (kind == {M.TROE}) {{')
        S('        const double term3 = dp[9] >= 0.0 ? exp(-dp[9] * rcpT) : 0.0;')
        S('        const double logFc = log10((1.0 - dp[6]) * exp(-dp[7] * T) + dp[6] * exp(-dp[8] * T) + term3);')
        S(f'        const double logPr = log10(Pr + {E.CFLOAT_MIN});')
        S('        const double tc = -0.4 - 0.67 * logFc + logPr;')
        S(f'        const double tf = tc / (0.75 - 1.27 * logFc - 0.14 * tc + {E.CFLOAT_MIN});')
        S('        Fc = exp10(logFc / (1.0 + tf * tf));')
        S(f'      }} else if (kind == {M.SRI}) {{')
        S(f'        const double logPr = log10(Pr + {E.CFLOAT_MIN});')
        S('        Fc = dp[9] * pow(T, dp[10]) * pow(dp[6] * exp(-dp[7] * rcpT) + exp(-dp[8] * T), 1.0 / (1.0 + logPr * logPr));')
        S('      }')
        S('      F *= Pr1 * Fc;')
    S('    }')


def emit_bk1_core(S, t, pre, nsets, MX, MS, mixed):
    """Loop-form production rates of one state, shared by the loop-form kernels and the
    kinetix_species_rates device routine. In scope: T, rcpT, lnT, T2, T3, T4, log2T, Tf,
    rcpTf, log2Tf, P, lnP and the private arrays w[N] (concentrations C_k on entry, w_k =
    C_k exp(g_k) on exit) and wd[N] (omega-dot on exit). Tables and helpers of prefix pre
    (_bk1_tables, the pow2 and PLOG helpers) precede it."""
    N, R = int(t['n_species']), int(t['n_reactions'])
    S(f'double gl[{N}], Ms[{max(nsets, 1)}];')
    S('double Cm = 0.0;')
    S('#pragma unroll 1')
    S(f'for (int k = 0; k < {N}; ++k) {{ Cm += w[k]; wd[k] = 0.0; }}')
    S('// third-body concentrations of the efficiency sets')
    S('#pragma unroll 1')
    S(f'for (int s = 0; s < {nsets}; ++s) {{')
    S('  double m = Cm;')
    S(f'  for (int i = {pre}_setoff[s]; i < {pre}_setoff[s + 1]; ++i) m += {pre}_sete[i] * w[{pre}_setsp[i]];')
    S('  Ms[s] = m;')
    S('}')
    S('// Gibbs terms: gl = g log2(e), w = C exp(g) (FP64); colliders read C before this')
    S(f'double Ccol[{N}];')
    S('#pragma unroll 1')
    S(f'for (int k = 0; k < {N}; ++k) {{')
    S(f'  const double* a = (T > {pre}_tmid[k] ? {pre}_ghi : {pre}_glo) + 7 * k;')
    S('  const double g = a[0] * (1.0 - lnT) + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4 + a[5] * rcpT - a[6];')
    S('  gl[k] = g * 1.4426950408889634;')
    S('  Ccol[k] = w[k];')
    S('  w[k] *= exp(g);')
    S('}')
    S(f'const double C0 = {repr(float(t["one_atm_R"]))} * rcpT;')
    S(f'const double rcpC0 = {repr(float(t["R_one_atm"]))} * T;')
    S('#pragma unroll 1')
    S(f'for (int r = 0; r < {R}; ++r) {{')
    S(f'  const int* mt = {pre}_meta + 8 * r;')
    S(f'  const double* dp = {pre}_dbl + 13 * r;')
    S('  const int kind = mt[0], nf = mt[2], nr = mt[3];')
    if mixed:
        S(f'  const float* fp = {pre}_flt + 6 * r;')
        S(f'  double ex = kind == {M.PLOG} ? {pre}_plog(r, P, lnP, log2T, rcpT)')
        S('                              : (double)(fp[0] + fp[1] * log2Tf - fp[2] * rcpTf);   // FP32 Arrhenius exponent')
    else:
        S(f'  double ex = kind == {M.PLOG} ? {pre}_plog(r, P, lnP, log2T, rcpT) : dp[0] + dp[1] * log2T - dp[2] * rcpT;')
    S('  double fwd = 1.0, rev = 1.0;')
    S(f'  for (int i = 0; i < nf; ++i) {{ const int k = {pre}_fsp[{MX} * r + i]; ex -= gl[k]; fwd *= w[k]; }}')
    S(f'  for (int i = 0; i < nr; ++i) rev *= w[{pre}_rsp[{MX} * r + i]];')
    S(f'  double F = {pre}_pow2(ex);')
    S(f'  if (kind != 0 && kind != {M.PLOG}) {{')
    S('    const double M = mt[6] == 3 ? Ms[mt[7]] : (mt[6] == 2 ? Ccol[mt[7]] : Cm);')
    emit_falloff(S, pre, mixed)
    S('  }')
    S('  double q;')
    S('  if (mt[1]) {')
    S('    double c0 = 1.0;')
    S('    const int dnu = mt[5];')
    S('    for (int i = 0; i < (dnu < 0 ? -dnu : dnu); ++i) c0 *= dnu < 0 ? C0 : rcpC0;')
    S('    q = F * (fwd - c0 * rev);')
    S('  } else {')
    S('    q = F * fwd;')
    S('  }')
    S(f'  for (int i = 0; i < mt[4]; ++i) wd[{pre}_nsp[{MS} * r + i]] += {pre}_nnu[{MS} * r + i] * q;')
    S('}')
