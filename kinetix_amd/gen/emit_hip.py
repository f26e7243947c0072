"""gfx950 HIP emitter: writes mechanism-specialised CDNA4 kernels from the IR tables.

One file per mechanism (``<out>/kernels.hip``), compiled ahead of time to a code object
(``kernels.hsaco``) that the host library loads with hipModuleLoad. Kernels:

  kx_bk1_fp64     species production rates + heat release rate, FP64
                  (the reference's productionRates kernel around kinetix_species_rates and
                  kinetix_enthalpy_RT: benchmark/okl/productionRates.okl:10-62,
                  kinetix/core/reaction_rates.py:616-1333)
  kx_bk2_fp64     mixture-averaged conductivity, viscosity (Wilke) and rho*D_km, FP64
                  (benchmark/okl/transportProps.okl:11-47, kinetix/core/mix_transport.py:261-471)
  kx_thermo_fp64  rho, cp_k, rho*cp (benchmark/okl/thermoCoeffs.okl:10-39)

Design (MI355X-first, not a translation of the reference's rolled C):
  * one state per lane, SoA loads/stores coalesce across the 64-wide wavefront;
  * the mechanism is straight-line code: every table index is a compile-time constant, so
    per-state arrays (concentrations, Gibbs exponentials, wdot) live in registers, not in
    scratch, and no rate constant / q_r array is ever materialised;
  * BK1 (w-form) is emitted in a register-liveness order (liveness.py): each species' Gibbs
    term and w = C exp(g) are computed at its first use, its omega-dot is stored after its
    last reaction, and the reactions are ordered to keep the live set small;
  * k = A * exp(beta lnT - Ta/T) (A kept exact instead of exp(ln A + ...));
  * Troe's 10^x is exp10(x).
"""
from __future__ import annotations

import math
import os

import numpy as np

from . import mechanism as M

BLOCK = 256
CFLOAT_MIN = '1e-300'


# FP64 values that gfx9 VOP3 instructions encode as inline constants (no materialisation).
_INLINE = {0.0, 0.5, -0.5, 1.0, -1.0, 2.0, -2.0, 4.0, -4.0}


class ConstPool:
    """Per-kernel table of FP64 constants, emitted as a __device__ array and read with
    compile-time offsets. Keeping them in memory (scalar loads next to their use) instead
    of as literals stops the compiler from hoisting ~10^3-10^4 64-bit immediates into SGPRs
    for the whole kernel and spilling them."""

    def __init__(self, name, lds=False):
        self.name = name
        self.values = []
        self.index = {}
        self.lds = lds            # stage the table in LDS once per workgroup
        self.ref = f'{name}_s' if lds else name

    def __call__(self, x):
        x = float(x)
        if x in _INLINE or math.isinf(x):
            return _literal(x)
        key = x.hex()
        if key not in self.index:
            self.index[key] = len(self.values)
            self.values.append(x)
        return f'{self.ref}[{self.index[key]}]'

    def prologue(self, block):
        """LDS staging code placed at the top of the kernel body (before any early exit)."""
        if not self.lds or not self.values:
            return ''
        n = len(self.values)
        return (f'  __shared__ double {self.ref}[{n}];\n'
                + ''.join('  ' + line + '\n' for line in lds_stage(self.ref, self.name, n, block))
                + '  __syncthreads();\n')

    def declaration(self):
        if not self.values:
            return ''
        body = ', '.join(_literal(v) for v in self.values)
        return f'__device__ double {self.name}[{len(self.values)}] = {{{body}}};\n'


def lds_stage(dst, src, n, block, batch=16):
    """Lines of C++ copying src[0, n) (global) to dst (LDS) with `block` threads: each thread
    issues `batch` loads (indices clamped in bounds, so they form one block) before its LDS
    writes. A plain strided copy waits out one L2 round trip per element and thread: ~20 us
    for a 10^4-entry table, idle time for every workgroup the table's LDS keeps alone on a CU."""
    step = batch * block
    return [f'for (int b_ = threadIdx.x; b_ < {n}; b_ += {step}) {{',
            f'  double t_[{batch}];',
            '  #pragma unroll',
            f'  for (int r_ = 0; r_ < {batch}; ++r_) t_[r_] = {src}[min(b_ + r_ * {block}, {n - 1})];',
            '  #pragma unroll',
            f'  for (int r_ = 0; r_ < {batch}; ++r_) if (b_ + r_ * {block} < {n}) {dst}[b_ + r_ * {block}] = t_[r_];',
            '}']


_POOL = None


def lit(x):
    """FP64 constant: from the active kernel's pool if one is set, else a literal."""
    return _POOL(x) if _POOL is not None else _literal(x)


def _literal(x):
    """Exact FP64 literal."""
    x = float(x)
    if x == 0.0:
        return '0.0'
    if math.isinf(x):
        return '__builtin_inf()' if x > 0 else '(-__builtin_inf())'
    r = repr(x)
    if 'e' not in r and '.' not in r:
        r += '.0'
    return r


class _Src:
    def __init__(self):
        self.lines = []

    def __call__(self, s='', ind=1):
        self.lines.append('  ' * ind + s if s else '')

    def text(self):
        return '\n'.join(self.lines)


def _prod(var, nu):
    """Product of concentrations with multiplicity, species order; '' if none."""
    terms = []
    for k, c in enumerate(nu):
        terms += [f'{var}{k}'] * int(c)
    return '*'.join(terms)


# --------------------------------------------------------------------------------------
# BK1
# --------------------------------------------------------------------------------------

class _RateEmitter:
    """Emits the kinetics of one mechanism for one lane (FP64)."""

    def __init__(self, t, info):
        self.t = t
        self.N = int(t['n_species'])
        self.R = int(t['n_reactions'])
        self.nu_f = t['nu_f']
        self.nu_r = t['nu_r']
        self.net = self.nu_r - self.nu_f
        self.kind = t['kind']
        self.rev = t['reversible'].astype(bool)
        self.has_eff = t['has_eff'].astype(bool)
        self.active = [k for k in range(self.N) if np.any(self.net[:, k] != 0)]
        self.equations = info.get('equations') if info else None
        self.sched_window = 1      # emit a scheduling barrier after every N-th reaction

    # unique efficiency vectors -> eff variables
    def efficiency_sets(self):
        sets, idx = [], {}
        for r in range(self.R):
            if self.has_eff[r] and self.kind[r] != M.ELEMENTARY:
                key = tuple(self.t['eff'][r])
                if key not in idx:
                    idx[key] = len(sets)
                    sets.append(key)
        return sets, idx

    def arrh_key(self, beta, Ta):
        beta, Ta = float(beta), float(Ta)
        if Ta == 0.0 and beta in (0.0, -2.0, -1.0, 1.0, 2.0):
            return None
        return (beta, Ta)

    def arrh_expr(self, A, beta, Ta, expvar):
        beta, Ta = float(beta), float(Ta)
        if Ta == 0.0:
            if beta == 0.0:
                return lit(A)
            if beta == -2.0:
                return f'{lit(A)} * rcpT2'
            if beta == -1.0:
                return f'{lit(A)} * rcpT'
            if beta == 1.0:
                return f'{lit(A)} * T'
            if beta == 2.0:
                return f'{lit(A)} * T2'
        return f'{lit(A)} * {expvar}'

    def exp_arg(self, beta, Ta):
        beta, Ta = float(beta), float(Ta)
        parts = []
        if beta != 0.0:
            parts.append(f'{lit(beta)} * lnT')
        if Ta != 0.0:
            parts.append(f'{lit(-Ta)} * rcpT')
        return ' + '.join(parts)

    def reaction_order(self):
        """Group reactions by their first Arrhenius exponent key so shared exps stay local."""
        t = self.t
        first = {}
        keys = []
        for r in range(self.R):
            k = self.arrh_key(t['beta'][r], t['Ta'][r])
            keys.append(k)
            if k is not None and k not in first:
                first[k] = r
        return sorted(range(self.R), key=lambda r: (first.get(keys[r], r), r))

    def emit_prelude(self, S, cvar='c'):
        """Third-body sums, Gibbs exponentials (FP64) and C0; returns the eff-set index."""
        t, N = self.t, self.N
        S('// ---- third-body concentrations')
        S(f'const double Cm = ' + ' + '.join(f'{cvar}{k}' for k in range(N)) + ';')
        sets, sidx = self.efficiency_sets()
        for i, eff in enumerate(sets):
            terms = ['Cm']
            for k, e in enumerate(eff):
                if e == 1.0:
                    continue
                if e == 2.0:
                    terms.append(f'+ {cvar}{k}')
                elif e == 0.0:
                    terms.append(f'- {cvar}{k}')
                else:
                    terms.append(f'+ {lit(e - 1.0)} * {cvar}{k}')
            S(f'const double eff{i} = ' + ' '.join(terms) + ';')
        need_g = sorted({k for r in range(self.R) if self.rev[r]
                         for k in range(N) if self.net[r, k] != 0})
        need_rg = sorted({k for r in range(self.R) if self.rev[r]
                          for k in range(N) if self.net[r, k] < 0})
        S('// ---- Gibbs exponentials exp(g_k/RT); NASA-7 piece chosen by T > T_mid without')
        S('// branching (both pieces evaluated, result selected), so the code stays one block')
        tmid = t['T_mid']
        for tm in sorted({float(tmid[k]) for k in need_g}):
            S(f'const bool hi{_tag(tm)} = T > {lit(tm)};')
        for k in need_g:
            S(f'double eg{k} = exp(hi{_tag(tmid[k])} ? ({self._gibbs(t["g_hi"][k])}) : ({self._gibbs(t["g_lo"][k])}));')
        for k in need_rg:
            S(f'const double rg{k} = 1.0 / eg{k};')
        S(f'const double C0 = {lit(float(t["one_atm_R"]))} * rcpT;')
        S(f'const double rcpC0 = {lit(float(t["R_one_atm"]))} * T;')
        return sidx

    def emit_reaction(self, S, r, sidx, seen, cvar='c', mixed=False):
        """Emit `double q<r>` = net rate of progress of reaction r (FP64).

        mixed=False: every rate coefficient in FP64.
        mixed=True:  Arrhenius exponentials through the FP32 hardware exp2 (FP64 argument,
                     exact power-of-two scaling) and falloff corrections in FP32; K_c and the
                     forward-minus-reverse subtraction stay FP64."""
        t, N = self.t, self.N
        pre = []
        kinf = self._rate_const(pre, t['A'][r], t['beta'][r], t['Ta'][r], seen, mixed)
        kind = self.kind[r]
        k0 = None
        if kind in (M.LINDEMANN, M.TROE, M.SRI):
            k0 = self._rate_const(pre, t['A0'][r], t['beta0'][r], t['Ta0'][r], seen, mixed)
        for line in pre:            # shared exps live at function scope
            S(line)
        S(f'double q{r};')
        S(f'{{ // {r + 1}: {self.equations[r] if self.equations else ""}')
        if kind == M.ELEMENTARY:
            kf = kinf
        elif kind == M.THREE_BODY:
            kf = f'({kinf}) * {self._third(r, sidx, cvar)}'
        elif kind == M.PLOG:
            # reached only through the concentration-form options (bk1_wform=False, or
            # emit_mixed's wform=False); the default w-form emits PLOG (emit_reaction_w)
            raise NotImplementedError('PLOG needs the w-form emitter (bk1_wform=True)')
        elif not mixed:
            S(f'  const double kinf = {kinf};')
            S(f'  const double Pr = ({k0}) * {self._third(r, sidx, cvar)} / (kinf + {CFLOAT_MIN});')
            if kind == M.LINDEMANN:
                S(f'  const double kf = kinf * Pr / (1.0 + Pr);')
            elif kind == M.TROE:
                A_, rT3, rT1, T2 = t['troe'][r]
                term3 = f' + exp({lit(-T2)} * rcpT)' if T2 >= 0 else ''
                S(f'  const double logFc = log10({lit(1.0 - A_)} * exp({lit(-rT3)} * T) + '
                  f'{lit(A_)} * exp({lit(-rT1)} * T){term3});')
                S(f'  const double logPr = log10(Pr + {CFLOAT_MIN});')
                S(f'  const double tc = -0.4 - 0.67 * logFc + logPr;')
                S(f'  const double tf = tc / (0.75 - 1.27 * logFc - 0.14 * tc + {CFLOAT_MIN});')
                S(f'  const double F = exp10(logFc / (1.0 + tf * tf));')
                S(f'  const double kf = kinf * (Pr / (1.0 + Pr)) * F;')
            elif kind == M.SRI:
                a, b, rc, d, e = t['sri'][r]
                S(f'  const double logPr = log10(Pr + {CFLOAT_MIN});')
                S(f'  const double F = {lit(d)} * pow(T, {lit(e)}) * pow({lit(a)} * exp({lit(-b)} * rcpT)'
                  f' + exp({lit(-rc)} * T), 1.0 / (1.0 + logPr * logPr));')
                S(f'  const double kf = kinf * (Pr / (1.0 + Pr)) * F;')
            kf = 'kf'
        else:
            # FP32 falloff correction; k_inf and k_0 arrive FP64-scaled
            S(f'  const double kinf = {kinf};')
            S(f'  const float Pr = (float)(({k0}) * {self._third(r, sidx, cvar)}) * '
              f'__builtin_amdgcn_rcpf((float)kinf + 1e-37f);')
            S('  const float Pr1 = Pr * __builtin_amdgcn_rcpf(1.0f + Pr);')
            if kind == M.LINDEMANN:
                S('  const double kf = kinf * (double)Pr1;')
            elif kind == M.TROE:
                A_, rT3, rT1, T2 = t['troe'][r]
                L2E = 1.4426950408889634
                term3 = (f' + __builtin_amdgcn_exp2f({float(-T2 * L2E)!r}f * rcpTf)' if T2 >= 0 else '')
                S(f'  const float logFc = {0.30102999566398120!r}f * __builtin_amdgcn_logf('
                  f'{float(1.0 - A_)!r}f * __builtin_amdgcn_exp2f({float(-rT3 * L2E)!r}f * Tf) + '
                  f'{float(A_)!r}f * __builtin_amdgcn_exp2f({float(-rT1 * L2E)!r}f * Tf){term3});')
                S(f'  const float logPr = {0.30102999566398120!r}f * __builtin_amdgcn_logf(Pr + 1e-37f);')
                S('  const float tc = -0.4f - 0.67f * logFc + logPr;')
                S('  const float tf = tc * __builtin_amdgcn_rcpf(0.75f - 1.27f * logFc - 0.14f * tc + 1e-37f);')
                S(f'  const float F = __builtin_amdgcn_exp2f({3.3219280948873622!r}f * logFc * '
                  f'__builtin_amdgcn_rcpf(1.0f + tf * tf));')
                S('  const double kf = kinf * (double)(Pr1 * F);')
            elif kind == M.SRI:
                a, b, rc, d, e = t['sri'][r]
                S(f'  const float logPr = {0.30102999566398120!r}f * __builtin_amdgcn_logf(Pr + 1e-37f);')
                S(f'  const float F = {float(d)!r}f * powf(Tf, {float(e)!r}f) * powf({float(a)!r}f * '
                  f'expf({float(-b)!r}f * rcpTf) + expf({float(-rc)!r}f * Tf), '
                  f'__builtin_amdgcn_rcpf(1.0f + logPr * logPr));')
                S('  const double kf = kinf * (double)(Pr1 * F);')
            kf = 'kf'
        fwd = _prod(cvar, self.nu_f[r]) or '1.0'
        if self.rev[r]:
            rev = _prod(cvar, self.nu_r[r]) or '1.0'
            S(f'  q{r} = ({kf}) * ({fwd} - {self._kr(r)} * {rev});')
        else:
            S(f'  q{r} = ({kf}) * ({fwd});')
        S('}')

    # ---- "w-form": q_r = [k_f exp(-sum_reactants g)] (prod_r w - C0^-dnu prod_p w), w = C exp(g)
    def emit_prelude_w(self, S, cvar='c'):
        """Third-body sums from C, then gl_k = g_k log2(e) and w_k = C_k exp(g_k) (FP64).
        Concentrations are dead afterwards. Returns the eff-set index."""
        t, N = self.t, self.N
        sidx = self.emit_third_body(S, cvar)
        used = sorted({k for r in range(self.R) for k in range(N)
                       if self.nu_f[r, k] or (self.rev[r] and self.nu_r[r, k])})
        S('// ---- Gibbs energies g_k (FP64): gl = g log2(e) for the rate exponents, w = C exp(g)')
        tmid = t['T_mid']
        for tm in sorted({float(tmid[k]) for k in used}):
            S(f'const bool hi{_tag(tm)} = T > {lit(tm)};')
        L2E = 1.4426950408889634
        for k in used:
            S(f'const double g{k} = hi{_tag(tmid[k])} ? ({self._gibbs(t["g_hi"][k])}) : ({self._gibbs(t["g_lo"][k])});')
            S(f'const double gl{k} = g{k} * {lit(L2E)};')
            S(f'const double w{k} = {cvar}{k} * exp(g{k});')
        S(f'const double C0 = {lit(float(t["one_atm_R"]))} * rcpT;')
        S(f'const double rcpC0 = {lit(float(t["R_one_atm"]))} * T;')
        return sidx

    def emit_third_body(self, S, cvar='c'):
        """Cm, the efficiency-set sums eff<i> and explicit colliders M<k> (FP64) from the
        concentrations. Returns the eff-set index."""
        t, N = self.t, self.N
        S('// ---- third-body concentrations')
        S(f'const double Cm = ' + ' + '.join(f'{cvar}{k}' for k in range(N)) + ';')
        sets, sidx = self.efficiency_sets()
        for i, eff in enumerate(sets):
            terms = ['Cm']
            for k, e in enumerate(eff):
                if e == 1.0:
                    continue
                if e == 2.0:
                    terms.append(f'+ {cvar}{k}')
                elif e == 0.0:
                    terms.append(f'- {cvar}{k}')
                else:
                    terms.append(f'+ {lit(e - 1.0)} * {cvar}{k}')
            S(f'const double eff{i} = ' + ' '.join(terms) + ';')
        colliders = sorted({int(t['collider'][r]) for r in range(self.R)
                            if t['collider'][r] >= 0 and not self.has_eff[r]})
        for k in colliders:
            S(f'const double M{k} = {cvar}{k};')
        return sidx

    def _third_w(self, r, sidx):
        if self.has_eff[r]:
            return f'eff{sidx[tuple(self.t["eff"][r])]}'
        if self.t['collider'][r] >= 0:
            return f'M{int(self.t["collider"][r])}'
        return 'Cm'

    def _log2_rate(self, A, beta, Ta, extra=''):
        """FP64 expression of log2(A T^beta exp(-Ta/T)) (+ extra)."""
        L2E = 1.4426950408889634
        parts = [lit(math.log2(A))] if A != 1.0 else []
        if float(beta) != 0.0:
            parts.append(f'{lit(beta)} * log2T')
        if float(Ta) != 0.0:
            parts.append(f'{lit(-float(Ta) * L2E)} * rcpT')
        if not parts:
            parts = ['0.0']
        return ' + '.join(parts) + extra

    def plog_points(self, r):
        """[(P [Pa], [(A, beta, Ta), ...]), ...] of PLOG reaction r, ascending in P."""
        t = self.t
        o0, o1 = int(t['plog_off'][r]), int(t['plog_off'][r + 1])
        pts = []
        for i in range(o0, o1):
            k0, nk = int(t['plog_koff'][i]), int(t['plog_nk'][i])
            rates = [(float(t['plog_A'][j]), float(t['plog_k'][j][1]), float(t['plog_k'][j][2]))
                     for j in range(k0, k0 + nk)]
            if any(A <= 0.0 for A, _, _ in rates):
                raise NotImplementedError(f'reaction {r + 1}: PLOG rates with A <= 0')
            pts.append((float(t['plog_P'][i]), rates))
        if any(b[0] <= a[0] for a, b in zip(pts, pts[1:])):
            raise ValueError(f'reaction {r + 1}: PLOG pressures not strictly ascending')
        return pts

    def _plog_point_log2k(self, rates):
        """FP64 expression of log2 of the sum of the point's Arrhenius rates."""
        terms = [self._log2_rate(A, b, Ta) for A, b, Ta in rates]
        if len(terms) == 1:
            return f'({terms[0]})'
        return 'log2(' + ' + '.join(f'exp2({x})' for x in terms) + ')'

    def _plog_log2k(self, r):
        """FP64 expression of log2 k_f of PLOG reaction r at the pressure P (lnP = ln P)."""
        pts = self.plog_points(r)
        L = [self._plog_point_log2k(rates) for _, rates in pts]
        if len(pts) == 1:
            return L[0]
        expr = L[-1]                                               # P > last point
        for i in range(len(pts) - 2, -1, -1):
            p1, p2 = pts[i][0], pts[i + 1][0]
            rcp = 1.0 / (math.log(p2) - math.log(p1))
            seg = (f'{L[i]} + ({L[i + 1]} - {L[i]}) * ((lnP - {lit(math.log(p1))}) * {lit(rcp)})')
            expr = f'(P <= {lit(p2)} ? {seg} : {expr})'
        return f'(P <= {lit(pts[0][0])} ? {L[0]} : {expr})'

    def _log2_rate32(self, A, beta, Ta):
        """FP32 expression of log2(A T^beta exp(-Ta/T)) with inline literal constants."""
        L2E = 1.4426950408889634
        parts = [f'{float(math.log2(A))!r}f'] if A != 1.0 else []
        if float(beta) != 0.0:
            parts.append(f'{float(beta)!r}f * log2Tf')
        if float(Ta) != 0.0:
            parts.append(f'{float(-float(Ta) * L2E)!r}f * rcpTf')
        return ' + '.join(parts) if parts else '0.0f'

    def emit_reaction_w(self, S, r, sidx, pow2='kx_pow2', fp64_falloff=False):
        """Emit `double q<r>` in the w-form. The scale factor k_f exp(-sum g_r) comes from
        one pow2() of an FP64 exponent (the Gibbs sum folded into the Arrhenius exponent);
        falloff corrections are FP32; the bracket difference is FP64."""
        t, N = self.t, self.N
        kind = self.kind[r]
        gsum = ''.join(f' - gl{k}' * int(self.nu_f[r, k]) for k in range(N) if self.nu_f[r, k])
        S(f'double q{r};')
        S(f'{{ // {r + 1}: {self.equations[r] if self.equations else ""}')
        if kind == M.PLOG:
            # log2 k_f interpolated linearly in ln P between the bracketing pressure points
            # and clamped outside them (reaction_rates.py:1064-1175); P is the launch's
            # scalar pressure, so every branch is uniform. FP64 in every variant.
            S(f'  const double ylk = {self._plog_log2k(r)};')
            S(f'  double F = {pow2}(ylk{gsum});')
        elif getattr(self, 'arrh32', False) and not fp64_falloff:
            # Arrhenius exponent in FP32 (inline literals); the Gibbs sum stays FP64
            S(f'  const float ya = {self._log2_rate32(t["A"][r], t["beta"][r], t["Ta"][r])};')
            S(f'  double F = {pow2}((double)ya{gsum});')
        else:
            S(f'  double F = {pow2}({self._log2_rate(t["A"][r], t["beta"][r], t["Ta"][r], gsum)});')
        if kind == M.THREE_BODY:
            S(f'  F *= {self._third_w(r, sidx)};')
        elif kind in (M.LINDEMANN, M.TROE, M.SRI) and fp64_falloff:
            ratio = math.log2(t['A0'][r] / t['A'][r])
            db = float(t['beta0'][r]) - float(t['beta'][r])
            dT = float(t['Ta0'][r]) - float(t['Ta'][r])
            S(f'  const double Pr = {self._third_w(r, sidx)} * {pow2}({self._log2_rate(2.0 ** ratio, db, dT)});')
            S('  const double Pr1 = Pr / (1.0 + Pr);')
            if kind == M.LINDEMANN:
                S('  F *= Pr1;')
            elif kind == M.TROE:
                A_, rT3, rT1, T2 = t['troe'][r]
                term3 = f' + exp({lit(-T2)} * rcpT)' if T2 >= 0 else ''
                S(f'  const double logFc = log10({lit(1.0 - A_)} * exp({lit(-rT3)} * T) + '
                  f'{lit(A_)} * exp({lit(-rT1)} * T){term3});')
                S(f'  const double logPr = log10(Pr + {CFLOAT_MIN});')
                S(f'  const double tc = -0.4 - 0.67 * logFc + logPr;')
                S(f'  const double tf = tc / (0.75 - 1.27 * logFc - 0.14 * tc + {CFLOAT_MIN});')
                S(f'  F *= Pr1 * exp10(logFc / (1.0 + tf * tf));')
            else:
                a, b, rc, d, e = t['sri'][r]
                S(f'  const double logPr = log10(Pr + {CFLOAT_MIN});')
                S(f'  F *= Pr1 * {lit(d)} * pow(T, {lit(e)}) * pow({lit(a)} * exp({lit(-b)} * rcpT)'
                  f' + exp({lit(-rc)} * T), 1.0 / (1.0 + logPr * logPr));')
        elif kind in (M.LINDEMANN, M.TROE, M.SRI):
            ratio = math.log2(t['A0'][r] / t['A'][r])
            db = float(t['beta0'][r]) - float(t['beta'][r])
            dT = float(t['Ta0'][r]) - float(t['Ta'][r])
            # clamped below FLT_MAX: Pr / (1 + Pr) must stay finite (-> 1) in FP32; the FP32
            # exponent is capped at 100 so exp2f stays finite (no 0 * inf at M = 0)
            if getattr(self, 'arrh32', False):
                S(f'  const float Pr = fminf((float){self._third_w(r, sidx)} * '
                  f'__builtin_amdgcn_exp2f(fminf({self._log2_rate32(2.0 ** ratio, db, dT)}, 100.0f)), 1e30f);')
            else:
                S(f'  const float Pr = (float)fmin({self._third_w(r, sidx)} * '
                  f'{pow2}({self._log2_rate(2.0 ** ratio, db, dT)}), 1e30);')
            S('  const float Pr1 = Pr * __builtin_amdgcn_rcpf(1.0f + Pr);')
            if kind == M.LINDEMANN:
                S('  F *= (double)Pr1;')
            elif kind == M.TROE:
                A_, rT3, rT1, T2 = t['troe'][r]
                L2E = 1.4426950408889634
                term3 = (f' + __builtin_amdgcn_exp2f({float(-T2 * L2E)!r}f * rcpTf)' if T2 >= 0 else '')
                S(f'  const float logFc = {0.30102999566398120!r}f * __builtin_amdgcn_logf('
                  f'{float(1.0 - A_)!r}f * __builtin_amdgcn_exp2f({float(-rT3 * L2E)!r}f * Tf) + '
                  f'{float(A_)!r}f * __builtin_amdgcn_exp2f({float(-rT1 * L2E)!r}f * Tf){term3});')
                S(f'  const float logPr = {0.30102999566398120!r}f * __builtin_amdgcn_logf(Pr + 1e-37f);')
                S('  const float tc = -0.4f - 0.67f * logFc + logPr;')
                S('  const float tf = tc * __builtin_amdgcn_rcpf(0.75f - 1.27f * logFc - 0.14f * tc + 1e-37f);')
                S(f'  const float Fc = __builtin_amdgcn_exp2f({3.3219280948873622!r}f * logFc * '
                  f'__builtin_amdgcn_rcpf(1.0f + tf * tf));')
                S('  F *= (double)(Pr1 * Fc);')
            else:
                a, b, rc, d, e = t['sri'][r]
                S(f'  const float logPr = {0.30102999566398120!r}f * __builtin_amdgcn_logf(Pr + 1e-37f);')
                S(f'  const float Fs = {float(d)!r}f * powf(Tf, {float(e)!r}f) * powf({float(a)!r}f * '
                  f'expf({float(-b)!r}f * rcpTf) + expf({float(-rc)!r}f * Tf), '
                  f'__builtin_amdgcn_rcpf(1.0f + logPr * logPr));')
                S('  F *= (double)(Pr1 * Fs);')
        fwd = _prod('w', self.nu_f[r]) or '1.0'
        if self.rev[r]:
            dnu = int(self.net[r].sum())
            rev = _prod('w', self.nu_r[r]) or '1.0'
            c0 = '*'.join(['C0' if dnu < 0 else 'rcpC0'] * abs(dnu))
            rev_term = f'({c0}) * {rev}' if c0 else rev
            S(f'  q{r} = F * ({fwd} - {rev_term});')
        else:
            S(f'  q{r} = F * ({fwd});')
        S('}')

    def emit_species_rates(self, S, cvar='c', wvar='w'):
        """FP64 kinetix_species_rates body: registers c{k} -> w{k} (+=)."""
        N = self.N
        sidx = self.emit_prelude(S, cvar)
        S('// ---- reactions')
        seen = {}
        for i, r in enumerate(self.reaction_order()):
            self.emit_reaction(S, r, sidx, seen, cvar)
            for k in range(N):
                nu = int(self.net[r, k])
                if nu == 1:
                    S(f'{wvar}{k} += q{r};')
                elif nu == -1:
                    S(f'{wvar}{k} -= q{r};')
                elif nu != 0:
                    S(f'{wvar}{k} += {lit(nu)} * q{r};')
            if self.sched_window and (i % self.sched_window == 0):
                S('__builtin_amdgcn_sched_barrier(0);')

    def _gibbs(self, a):
        return (f'{lit(a[0])} * (1.0 - lnT) + {lit(a[1])} * T + {lit(a[2])} * T2 + '
                f'{lit(a[3])} * T3 + {lit(a[4])} * T4 + {lit(a[5])} * rcpT - {lit(a[6])}')

    def _rate_const(self, pre, A, beta, Ta, seen, mixed=False):
        """Expression of A T^beta exp(-Ta/T); a shared exp is appended to pre on first use."""
        key = self.arrh_key(beta, Ta)
        if key is None:
            return self.arrh_expr(A, beta, Ta, None)
        if key not in seen:
            name = f'ea{len(seen)}'
            seen[key] = name
            if mixed:
                b, Tk = key
                L2E = 1.4426950408889634
                parts = []
                if b != 0.0:
                    parts.append(f'{lit(b)} * log2T')
                if Tk != 0.0:
                    parts.append(f'{lit(-Tk * L2E)} * rcpT')
                pre.append(f'const double {name} = kx_pow2({" + ".join(parts)});')
            else:
                pre.append(f'const double {name} = exp({self.exp_arg(beta, Ta)});')
        return self.arrh_expr(A, beta, Ta, seen[key])

    def _third(self, r, sidx, cvar):
        if self.has_eff[r]:
            return f'eff{sidx[tuple(self.t["eff"][r])]}'
        if self.t['collider'][r] >= 0:
            return f'{cvar}{int(self.t["collider"][r])}'
        return 'Cm'

    def _kr(self, r):
        terms = []
        dnu = 0
        for k in range(self.N):
            nu = int(self.net[r, k])
            dnu += nu
            if nu < 0:
                terms += [f'rg{k}'] * (-nu)
            elif nu > 0:
                terms += [f'eg{k}'] * nu
        terms += ['C0' if dnu < 0 else 'rcpC0'] * abs(dnu)
        return '(' + '*'.join(terms) + ')' if terms else '1.0'


def _emit_bk1_fp64(t, info, sched_window=1, wform=False, variant='fp64', arrh32=False):
    """variant 'fp64': plain FP64 kernel. variant 'mixed_valu': the mixed-precision rate
    coefficients (FP32 exp2 with FP64 scaling, FP32 falloff; with arrh32 the Arrhenius
    exponent too is formed in FP32, then added to the FP64 Gibbs sum) with FP64 VALU
    accumulation of wdot instead of MFMA (w-form only)."""
    N = int(t['n_species'])
    em = _RateEmitter(t, info)
    em.sched_window = sched_window
    mixed = variant == 'mixed_valu'
    em.arrh32 = mixed and arrh32
    if mixed:
        wform = True
    S = _Src()
    if mixed:
        S('__device__ __forceinline__ double kx_pow2v(double y)', 0)
        S('{', 0)
        S('const double n = __builtin_rint(y);')
        S('const float m = __builtin_amdgcn_exp2f((float)(y - n));')
        S('return __builtin_ldexp((double)m, (int)n);')
        S('}', 0)
    S(f'extern "C" __global__ void __launch_bounds__({BLOCK})', 0)
    S(f'kx_bk1_{variant}(const int n_states, const long long offsetT, const long long offset,', 0)
    S('            const double pressure_R, const double pressure_,', 0)
    S('            const double* __restrict__ state, double* __restrict__ rates, const double Tref)', 0)
    S('{', 0)
    S(POOL_MARK)
    S(f'const long long id = (long long)blockIdx.x * {BLOCK} + threadIdx.x;')
    S('if (id >= n_states) return;')
    S('const double T = Tref * state[id];')
    S('const double rcpT = 1.0 / T;')
    S('const double lnT = log(T);')
    S('const double T2 = T * T;')
    S('const double T3 = T * T * T;')
    S('const double T4 = T * T * T * T;')
    S('const double rcpT2 = rcpT * rcpT;')
    S(f'const double log2T = lnT * {lit(1.4426950408889634)};')
    S('const float Tf = (float)T;')
    S('const float rcpTf = (float)rcpT;')
    S('const float log2Tf = (float)log2T;')
    S('const double P = pressure_, lnP = log(P);   // PLOG reactions (uniform per launch)')
    S('(void)P; (void)lnP; (void)log2T; (void)rcpT2; (void)Tf; (void)rcpTf; (void)log2Tf;')
    S('// ---- Y -> concentrations (negative mass fractions clipped to 0)')
    rcpW = t['rcpW']
    for k in range(N):
        S(f'double c{k} = fmax(0.0, state[id + offsetT + {k} * offset]) * {lit(rcpW[k])};')
    S('const double rcpMbar = ' + ' + '.join(f'c{k}' for k in range(N)) + ';')
    S('const double rho = pressure_R * rcpT / rcpMbar;')
    for k in range(N):
        S(f'c{k} *= rho;')
    for k in em.active:
        S(f'double wd{k} = 0.0;')
    if wform:
        sidx = em.emit_prelude_w(S, 'c')
        for i, r in enumerate(em.reaction_order()):
            if mixed:
                em.emit_reaction_w(S, r, sidx, pow2='kx_pow2v', fp64_falloff=False)
            else:
                em.emit_reaction_w(S, r, sidx, pow2='exp2', fp64_falloff=True)
            for k in range(N):
                nu = int(em.net[r, k])
                if nu == 1:
                    S(f'wd{k} += q{r};')
                elif nu == -1:
                    S(f'wd{k} -= q{r};')
                elif nu != 0:
                    S(f'wd{k} += {lit(nu)} * q{r};')
            if sched_window and (i % sched_window == 0):
                S('__builtin_amdgcn_sched_barrier(0);')
    else:
        em.emit_species_rates(S, wvar='wd')
    S('// ---- mass production rates and heat release rate')
    W = t['W']
    for k in range(N):
        val = f'{lit(W[k])} * wd{k}' if k in em.active else '0.0'
        S(f'rates[id + offsetT + {k} * offset] = {val};')
    S('double sum_h = 0.0;')
    _emit_h_RT_sum(S, t, em.active, wvar='wd')
    S(f'rates[id] = {lit(-float(t["R"]))} * T * sum_h;')
    S('}', 0)
    return S.text()


def _emit_bk1_live(t, info, variant='fp64', arrh32=False, order=None, waves=0):
    """w-form BK1 (FP64 or mixed_valu) in a register-liveness order: reactions follow
    liveness.order_reactions; g_k and w_k are computed at the species' first use; wd_k
    starts at its first net coefficient and is stored (with its enthalpy term added to the
    heat release sum) after its last one. The NASA-7 coefficients of both pieces sit in an
    LDS table read at a per-lane offset (T > T_mid selects the piece without a branch).
    waves > 0 asks the compiler for that many waves per SIMD (register budget 512/waves):
    the mixed kernel runs 8 % faster at 2 (some VGPRs go to scratch, latency hides behind
    the second wave); the FP64 kernel, with more live FP64 state, is faster at 1. Tried for
    the mixed kernel's 128 spilled VGPRs, not kept: reading Y_k again at each species' first
    use (through an opaque pointer) instead of keeping its concentration live, and scheduling
    barriers every 1-16 reactions -- 126-166 spilled VGPRs in every combination."""
    from . import liveness
    N = int(t['n_species'])
    em = _RateEmitter(t, info)
    mixed = variant == 'mixed_valu'
    em.arrh32 = mixed and arrh32
    kname = f'kx_bk1_{variant}'
    G, Wm, D = liveness.use_masks(t)
    if order is None:
        order = liveness.order_reactions(t)
    order = [int(r) for r in order]
    assert sorted(order) == list(range(em.R))
    # NASA-7 table: [piece: 0 = T > T_mid, 1 = T <= T_mid][species][16]: Gibbs in 0..6,
    # enthalpy in 8..13
    PT = np.zeros((2, N, 16))
    for k in range(N):
        PT[0, k, :7], PT[0, k, 8:14] = t['g_hi'][k], t['h_hi'][k]
        PT[1, k, :7], PT[1, k, 8:14] = t['g_lo'][k], t['h_lo'][k]
    S = _Src()
    S(f'__device__ const double {kname}_poly[{PT.size}] = {{' +
      ', '.join(_literal(v) for v in PT.reshape(-1)) + '};', 0)
    if mixed:
        S('__device__ __forceinline__ double kx_pow2v(double y)', 0)
        S('{', 0)
        S('const double n = __builtin_rint(y);')
        S('const float m = __builtin_amdgcn_exp2f((float)(y - n));')
        S('return __builtin_ldexp((double)m, (int)n);')
        S('}', 0)
    attr = f' __attribute__((amdgpu_waves_per_eu({waves}, {waves})))' if waves else ''
    S(f'extern "C" __global__ void __launch_bounds__({BLOCK}){attr}', 0)
    S(f'{kname}(const int n_states, const long long offsetT, const long long offset,', 0)
    S('            const double pressure_R, const double pressure_,', 0)
    S('            const double* __restrict__ state, double* __restrict__ rates, const double Tref)', 0)
    S('{', 0)
    S(POOL_MARK)
    S(f'__shared__ double poly_s[{PT.size}];')
    for line in lds_stage('poly_s', f'{kname}_poly', PT.size, BLOCK):
        S(line)
    S('__syncthreads();')
    S(f'const long long id = (long long)blockIdx.x * {BLOCK} + threadIdx.x;')
    S('if (id >= n_states) return;')
    S('const double T = Tref * state[id];')
    S('const double rcpT = 1.0 / T;')
    S('const double lnT = log(T);')
    S('const double T2 = T * T;')
    S('const double T3 = T * T * T;')
    S('const double T4 = T * T * T * T;')
    S('const double rcpT2 = rcpT * rcpT;')
    S('const double omlnT = 1.0 - lnT;')
    S(f'const double log2T = lnT * {lit(1.4426950408889634)};')
    S('const float Tf = (float)T;')
    S('const float rcpTf = (float)rcpT;')
    S('const float log2Tf = (float)log2T;')
    S('const double P = pressure_, lnP = log(P);   // PLOG reactions (uniform per launch)')
    S('(void)P; (void)lnP; (void)log2T; (void)rcpT2; (void)Tf; (void)rcpTf; (void)log2Tf;')
    S('// ---- Y -> concentrations (negative mass fractions clipped to 0)')
    rcpW = t['rcpW']
    for k in range(N):
        S(f'double c{k} = fmax(0.0, state[id + offsetT + {k} * offset]) * {lit(rcpW[k])};')
    S('const double rcpMbar = ' + ' + '.join(f'c{k}' for k in range(N)) + ';')
    S('const double rho = pressure_R * rcpT / rcpMbar;')
    for k in range(N):
        S(f'c{k} *= rho;')
    sidx = em.emit_third_body(S, 'c')
    tmid = t['T_mid']
    S('// ---- NASA-7 piece per T_mid: a per-lane offset into the LDS table')
    for tm in sorted({float(v) for v in tmid}):
        S(f'const double* P{_tag(tm)} = poly_s + (T > {lit(tm)} ? 0 : {N * 16});')
    S(f'const double C0 = {lit(float(t["one_atm_R"]))} * rcpT;')
    S(f'const double rcpC0 = {lit(float(t["R_one_atm"]))} * T;')
    S('double sum_h = 0.0;')
    L2E = 1.4426950408889634
    W = t['W']
    remaining = D.sum(0).astype(int)
    have_w, have_wd = set(), set()
    for r in order:
        for k in np.flatnonzero(Wm[r]):
            k = int(k)
            if k in have_w:
                continue
            have_w.add(k)
            P, o = f'P{_tag(tmid[k])}', 16 * k
            S(f'const double g{k} = {P}[{o}] * omlnT + {P}[{o + 1}] * T + {P}[{o + 2}] * T2 + '
              f'{P}[{o + 3}] * T3 + {P}[{o + 4}] * T4 + {P}[{o + 5}] * rcpT - {P}[{o + 6}];')
            S(f'const double gl{k} = g{k} * {lit(L2E)};')
            S(f'const double w{k} = c{k} * exp(g{k});')
        em.emit_reaction_w(S, r, sidx, pow2='kx_pow2v' if mixed else 'exp2', fp64_falloff=not mixed)
        for k in np.flatnonzero(D[r]):
            k = int(k)
            nu = int(em.net[r, k])
            term = f'q{r}' if nu == 1 else (f'-q{r}' if nu == -1 else f'{lit(nu)} * q{r}')
            if k in have_wd:
                S(f'wd{k} += {term};')
            else:
                have_wd.add(k)
                S(f'double wd{k} = {term};')
            remaining[k] -= 1
            if remaining[k] == 0:
                P, o = f'P{_tag(tmid[k])}', 16 * k + 8
                S(f'rates[id + offsetT + {k} * offset] = {lit(W[k])} * wd{k};')
                S(f'sum_h += wd{k} * ({P}[{o}] + {P}[{o + 1}] * T + {P}[{o + 2}] * T2 + '
                  f'{P}[{o + 3}] * T3 + {P}[{o + 4}] * T4 + {P}[{o + 5}] * rcpT);')
    for k in range(N):
        if k not in have_wd:
            S(f'rates[id + offsetT + {k} * offset] = 0.0;')
    S(f'rates[id] = {lit(-float(t["R"]))} * T * sum_h;')
    S('}', 0)
    return S.text()


def _emit_h_RT_sum(S, t, active, wvar='w'):
    tmid = t['T_mid']
    for tm in sorted({float(tmid[k]) for k in active}):
        S(f'const bool hh{_tag(tm)} = T > {lit(tm)};')
    for k in active:
        S(f'sum_h += {wvar}{k} * (hh{_tag(tmid[k])} ? ({_hpoly(t["h_hi"][k])}) : ({_hpoly(t["h_lo"][k])}));')


def _tag(tm):
    return str(int(round(float(tm) * 1000)))


def _hpoly(a):
    return (f'{lit(a[0])} + {lit(a[1])} * T + {lit(a[2])} * T2 + {lit(a[3])} * T3 + '
            f'{lit(a[4])} * T4 + {lit(a[5])} * rcpT')


# --------------------------------------------------------------------------------------
# BK2
# --------------------------------------------------------------------------------------

def _poly4(c):
    return (f'{lit(c[0])} + {lit(c[1])} * lnT + {lit(c[2])} * lnT2 + {lit(c[3])} * lnT3 + '
            f'{lit(c[4])} * lnT4')


def _emit_bk2_prelude(S, t, sid):
    """T powers, mole fractions x{k} and Mbar of state `sid` (transportProps.okl:12-35)."""
    N = int(t['n_species'])
    S(f'const double T = Tref * state[{sid}];')
    S('const double lnT = log(T);')
    S('const double rcpT = 1.0 / T;')
    S('const double sqrT = sqrt(T);')
    S('const double lnT2 = lnT * lnT;')
    S('const double lnT3 = lnT * lnT * lnT;')
    S('const double lnT4 = lnT * lnT * lnT * lnT;')
    rcpW = t['rcpW']
    for k in range(N):
        S(f'double x{k} = fmax(0.0, state[{sid} + offsetT + {k} * offset]) * {lit(rcpW[k])};')
    S('const double rcpMbar = ' + ' + '.join(f'x{k}' for k in range(N)) + ';')
    S('const double Mbar = 1.0 / rcpMbar;')
    for k in range(N):
        S(f'x{k} *= Mbar;')


def _emit_diffusion_pairs(S, t):
    """Row sums s{k} = sum_{j != k} x_j / D'_kj over the symmetric pair fits, each pair read
    once and added to both rows (mix_transport.py:414-471)."""
    N = int(t['n_species'])
    rcp_diff = bool(int(t.get('rcp_diff', 0)))
    S('// ---- mixture-averaged diffusion: symmetric pair sums S_k = sum_{j!=k} X_j / D_kj')
    for k in range(N):
        S(f'double s{k} = 0.0;')
    d = t['diff']
    for k in range(1, N):
        S('__builtin_amdgcn_sched_barrier(0);')
        for j in range(k):
            if rcp_diff:
                S(f'{{ const double rD = {_poly4(d[k, j])}; s{k} += x{j} * rD; s{j} += x{k} * rD; }}')
            else:
                # the fit is positive over its window: a Newton-refined reciprocal
                S(f'{{ const double rD = kx_rcp({_poly4(d[k, j])}); s{k} += x{j} * rD; s{j} += x{k} * rD; }}')


def _emit_conductivity(S, t):
    """lambda = sqrt(T) * (sum x_k l_k + 1 / sum x_k / l_k) / 2 (mix_transport.py:261-294)."""
    N = int(t['n_species'])
    S('// ---- conductivity')
    S('double sum1 = 0.0, sum2 = 0.0;')
    for k in range(N):
        S(f'{{ const double l = {_poly4(t["cond"][k])}; sum1 += x{k} * l; sum2 += x{k} / l; }}')
    S('const double lambda = sqrT * (0.5 * (sum1 + 1.0 / sum2));')


def _emit_bk2_fp64(t, info, pstate=False, block=BLOCK, queue=False):
    """pstate=False: kx_bk2_fp64, scalar p (the reference's transport kernel).
    pstate=True: kx_bk2_fp64_pstate, p [Pa] per state, also writes D_km.
    queue=True: kx_bk2_fp64_queue, the plain kernel over a device queue of state ids (the
    surrogate's audit samples, then its out-of-hull states), p per state, reading each state
    and writing its outputs in place; audit entries first compare mu and lambda with the
    surrogate values they replace.
    block: threads per workgroup (the host launches with the kernel's launch bound). The
    kernel's LDS constant pool (~85 KB for GRI-3.0) admits one workgroup per CU, so 512
    threads (8 waves) is what gives 2 waves/SIMD: 8.4 ms instead of 14.5 at 256 threads for
    4M states, despite ~160 VGPRs spilled to scratch; 768 and 1024 spill far more and are
    2-3x slower. The queue form is latency-bound (a few 10^4 states): 256 threads, one wave per
    SIMD on every CU, each wave free to use 372 of the 512 registers without a spill. It is
    emitted only for a mechanism whose tables do not fit the cooperative queue kernel's LDS
    (emit_bk2_queue, 16 lanes per state)."""
    N = int(t['n_species'])
    S = _Src()
    S(f'extern "C" __global__ void __launch_bounds__({block})', 0)
    if queue:
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
        S('// counters: [0] out-of-hull states queued, [1] failed audits. The grid covers the')
        S('// longest possible queue (every state out of hull); workgroups past the actual end')
        S('// return before staging the LDS table, after the working ones were dispatched.')
        S('const long long valid = (long long)n_audit + max(counters[0], 0);')
        S(f'if ((long long)blockIdx.x * {block} >= valid) return;   // whole workgroup past the queue')
        S(POOL_MARK)
        S(f'const long long slot = (long long)blockIdx.x * {block} + threadIdx.x;')
        S('if (slot >= valid) return;')
        S('const long long id = queue[slot];')
        S('const double p = p_state[id];')
    elif pstate:
        S('kx_bk2_fp64_pstate(const int n_states, const long long offsetT, const long long offset,', 0)
        S('            const double* __restrict__ p_state,', 0)
    else:
        S('kx_bk2_fp64(const int n_states, const long long offsetT, const long long offset,', 0)
        S('            const double pressure,', 0)
    if not queue:
        S('            const double* __restrict__ state, double* __restrict__ conductivity,', 0)
        S('            double* __restrict__ viscosity, double* __restrict__ rhoD,', 0)
        if pstate:
            S('            double* __restrict__ Dkm, const double Tref,', 0)
        else:
            S('            const double Tref,', 0)
        S('            const int* __restrict__ n_dev)', 0)
        S('{', 0)
        S('// n_dev (optional): a state count known only on the device; workgroups wholly past it')
        S('// return before staging the LDS table')
        S(f'if (n_dev && (long long)blockIdx.x * {block} >= *n_dev) return;')
        S(POOL_MARK)
        S(f'const long long id = (long long)blockIdx.x * {block} + threadIdx.x;')
        S('if (id >= n_states || (n_dev && id >= *n_dev)) return;')
        S('const double p = ' + ('p_state[id];' if pstate else 'pressure;'))
    _emit_bk2_prelude(S, t, 'id')
    rcpW, W = t['rcpW'], t['W']
    # diffusion first: x and the N row sums are live
    _emit_diffusion_pairs(S, t)
    S('const double TsqrT = T * sqrT;')
    S(f'const double rho = p * {lit(1.0 / float(t["R"]))} * rcpT * Mbar;')
    S('const double pMbar = p * Mbar;')
    def stores(with_dkm):
        for k in range(N):
            S(f'{{ const double D = TsqrT * (Mbar - {lit(W[k])} * x{k}) / (pMbar * s{k});')
            if queue:
                # audit of rho*D: the surrogate's value is read before it is overwritten (an
                # unconditional load; the flag masks non-audit entries, no branch per store)
                S(f'  const double v = rho * D; const double e = fabs(rhoD[{k} * offset + id] / v - 1.0);')
                S(f'  bad_rhoD |= !(e <= tol_rhoD); rhoD[{k} * offset + id] = v;'
                  + (f' Dkm[{k} * offset + id] = D; }}' if with_dkm else ' }'))
            else:
                S(f'  rhoD[{k} * offset + id] = rho * D;' + (f' Dkm[{k} * offset + id] = D; }}' if with_dkm else ' }'))
    if queue:
        S('bool bad_rhoD = false;   // max_k |rhoD_k(surrogate) / rhoD_k - 1| > tol_rhoD, or NaN')
        # D_km optional: one uniform branch around the stores (a branch per store splits the
        # kernel into ~10^2 blocks, and the scheduler then spills ~10^4 VGPRs)
        S('if (Dkm) {')
        stores(True)
        S('} else {')
        stores(False)
        S('}')
    else:
        stores(pstate)
    # viscosity: Wilke, eta_k = poly, Phi^(1/2)_kj = C1 + C2 eta_k / eta_j
    S('__builtin_amdgcn_sched_barrier(0);')
    S('// ---- viscosity (Wilke)')
    for k in range(N):
        S(f'const double re{k} = kx_rcp({_poly4(t["visc"][k])});')
    S('double vis = 0.0;')
    C1, C2 = t['C1'], t['C2']
    for k in range(N):
        S('__builtin_amdgcn_sched_barrier(0);')
        S('{')
        S(f'  const double e = {_poly4(t["visc"][k])};')
        S('  double s = 0.0;')
        for j in range(N):
            S(f'  {{ const double f = {lit(C1[k, j])} + {lit(C2[k, j])} * e * re{j}; s += x{j} * f * f; }}')
        S(f'  vis += x{k} * e * e / s;')
        S('}')
    S('__builtin_amdgcn_sched_barrier(0);')
    S('const double mu = sqrT * vis;')
    if not queue:
        S('viscosity[id] = mu;')
    _emit_conductivity(S, t)
    if queue:
        S('if (slot < n_audit && flags[id] == 0) {   // audit of a state the surrogate answered')
        S('  const double ev = fabs(viscosity[id] / mu - 1.0);')
        S('  const double ec = fabs(conductivity[id] / lambda - 1.0);')
        S('  if (!(ev <= tol) || !(ec <= tol) || bad_rhoD) atomicAdd(counters + 1, 1);')
        S('}')
        S('viscosity[id] = mu;')
    S('conductivity[id] = lambda;')
    S('}', 0)
    return S.text()


# --------------------------------------------------------------------------------------
# thermoCoeffs
# --------------------------------------------------------------------------------------

def _emit_thermo_fp64(t, info):
    """thermoCoeffs (thermoCoeffs.okl:1-41, thermodynamics.py:41-62): rho, cp_k and rho*cp.
    The NASA-7 c_p coefficients of both pieces sit in an LDS table read at a per-lane offset
    (T > T_mid picks the piece without a branch), as in the production-rate kernels: with the
    coefficients as literals behind a per-species select, the compiler materialised the selects
    early -- 260 VGPRs at 53 species, 506 spilled at 160. One pass over the species: Mbar
    factors out of the mean, sum_k c_k y_k Mbar = Mbar sum_k c_k y_k, so no per-species value
    stays live past its species; a scheduling barrier every 8 species keeps the compiler from
    hoisting all N loads of Y to the top."""
    N = int(t['n_species'])
    S = _Src()
    CP = np.zeros((2, N, 8))
    for k in range(N):
        CP[0, k, :5], CP[1, k, :5] = t['cp_hi'][k], t['cp_lo'][k]
    S(f'__device__ const double kx_thermo_fp64_cp[{CP.size}] = {{' +
      ', '.join(_literal(v) for v in CP.reshape(-1)) + '};', 0)
    S(f'extern "C" __global__ void __launch_bounds__({BLOCK})', 0)
    S('kx_thermo_fp64(const int n_states, const long long offsetT, const long long offset,', 0)
    S('               const double pressure_R, const double* __restrict__ state,', 0)
    S('               double* __restrict__ rho_out, double* __restrict__ cp,', 0)
    S('               double* __restrict__ rhoCp, const double Tref)', 0)
    S('{', 0)
    S(POOL_MARK)
    S(f'__shared__ double cp_s[{CP.size}];')
    for line in lds_stage('cp_s', 'kx_thermo_fp64_cp', CP.size, BLOCK):
        S(line)
    S('__syncthreads();')
    S(f'const long long id = (long long)blockIdx.x * {BLOCK} + threadIdx.x;')
    S('if (id >= n_states) return;')
    S('const double T = Tref * state[id];')
    S('const double rcpT = 1.0 / T;')
    S('const double T2 = T * T;')
    S('const double T3 = T * T * T;')
    S('const double T4 = T * T * T * T;')
    rcpW = t['rcpW']
    R = float(t['R'])
    tmid = t['T_mid']
    for tm in sorted({float(x) for x in tmid}):
        S(f'const double* C{_tag(tm)} = cp_s + (T > {lit(tm)} ? 0 : {N * 8});')
    S('double rcpMbar = 0.0, cy = 0.0;')
    for k in range(N):
        if k % 8 == 0 and k:
            S('__builtin_amdgcn_sched_barrier(0);')
        S(f'{{ const double y = fmax(0.0, state[id + offsetT + {k} * offset]) * {lit(rcpW[k])};')
        S('  rcpMbar += y;')
        S(f'  const double* a = C{_tag(tmid[k])} + {8 * k};')
        S('  const double c = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;')
        S(f'  cp[{k} * offset + id] = c * {lit(R * rcpW[k])};')
        S('  cy += c * y; }')
    S('const double Mbar = 1.0 / rcpMbar;')
    S('const double rho = pressure_R * rcpT * Mbar;')
    S('rho_out[id] = rho;')
    S('const double mean = cy * Mbar;')
    S(f'rhoCp[id] = rho * (mean * {lit(R)} * rcpMbar);')
    S('}', 0)
    return S.text()


HEADER = '''// GENERATED by kinetix_amd.gen.emit_hip for mechanism "{name}" -- do not edit.
// {n_species} species ({n_active} active), {n_reactions} reactions. Target: gfx950 (CDNA4).
#include <hip/hip_runtime.h>
#include <math.h>
#include "kx_rcp.h"      // kx_rcp: Newton-refined FP64 reciprocal (kinetix_amd/csrc)
'''


POOL_MARK = '// @@pool-prologue@@'
LDS_POOL_MAX_BYTES = 100 * 1024


def _with_pool(name, fn, *args, lds=False, block=BLOCK, lds_max=LDS_POOL_MAX_BYTES):
    global _POOL
    _POOL = ConstPool(name, lds=lds)
    try:
        body = fn(*args)
        if lds and len(_POOL.values) * 8 > lds_max:
            # too large for LDS: re-emit reading the table from global memory
            _POOL = ConstPool(name, lds=False)
            body = fn(*args)
        return _POOL.declaration() + body.replace('  ' + POOL_MARK, _POOL.prologue(block).rstrip('\n'))
    finally:
        _POOL = None


DEFAULT_OPTS = {'bk1_sched_window': 0, 'bk1_wform': True, 'mixed': None, 'lds_pool': True,
                'mixed_lds_pool': False, 'mixed_arrh32': True, 'bk1_live': True,
                'mixed_waves': 2, 'bk2_block': 512, 'coop_lps': 32}


def generate(tables, info, opts=None):
    o = dict(DEFAULT_OPTS)
    o.update(opts or {})
    parts = [HEADER.format(name=info['name'], n_species=info['n_species'],
                           n_active=info['n_active_species'], n_reactions=info['n_reactions'])]
    from . import emit_rolled
    if o.get('rolled') or int(tables['n_species']) > emit_rolled.ROLL_N:
        # large mechanism: loop forms (private per-lane species arrays, uniform table loads);
        # no MFMA-accumulation variant and no surrogate (its width is fixed at 64)
        # production rates: several lanes per state, species arrays in LDS (emit_coop)
        from . import emit_coop
        parts.append(emit_coop.emit_bk1(tables, info, 'fp64', o['coop_lps']))
        parts.append(emit_coop.emit_bk1(tables, info, 'mixed_valu', o['coop_lps']))
        if 'cond' in tables:
            # transport over 64-state tiles (species rows across wavefronts, x and 1/eta in
            # LDS) up to 160 species; beyond that, or asked for (bk2_loop), the one-lane loop form
            from . import emit_bk2_tile as BT
            tile = BT.lds_bytes(tables) <= 160 * 1024 and not o.get('bk2_loop')
            for ps in (False, True):
                parts.append(BT.emit(tables, info, ps) if tile else emit_rolled.emit_bk2(tables, info, ps))
        parts.append(_with_pool('kx_thermo_fp64_k', _emit_thermo_fp64, tables, info))
        return '\n\n'.join(parts) + '\n'
    order = None
    if o['bk1_live']:
        from . import liveness
        order = liveness.order_reactions(tables)
        parts.append(_with_pool('kx_bk1_fp64_k', _emit_bk1_live, tables, info, 'fp64', False, order,
                                lds=o['lds_pool']))
    else:
        parts.append(_with_pool('kx_bk1_fp64_k', _emit_bk1_fp64, tables, info, o['bk1_sched_window'],
                                o['bk1_wform'], lds=o['lds_pool']))
    if 'cond' in tables:
        parts.append(_with_pool('kx_bk2_fp64_k', _emit_bk2_fp64, tables, info, False, o['bk2_block'],
                                lds=o['lds_pool'], block=o['bk2_block']))
        parts.append(_with_pool('kx_bk2_fp64_pstate_k', _emit_bk2_fp64, tables, info, True, o['bk2_block'],
                                lds=o['lds_pool'], block=o['bk2_block']))
        from . import emit_bk2_queue as BQ
        if BQ.lds_bytes(tables) <= 160 * 1024:
            # several lanes per state: the pass over a few 10^4 queued states is latency-bound
            parts.append(BQ.emit(tables, info))
        else:
            parts.append(_with_pool('kx_bk2_fp64_queue_k', _emit_bk2_fp64, tables, info, True,
                                    256, True, lds=o['lds_pool'], block=256))
        from . import emit_bk2_mfma as BM
        # the constant pool goes to LDS only beside the Wilke tables within 160 KB per CU
        parts.append(_with_pool('kx_bk2_fp64_mfma_k', BM.emit, tables, info, lds=o['lds_pool'],
                                block=64 * BM.WAVES, lds_max=160 * 1024 - BM.lds_bytes(tables)))
    parts.append(_with_pool('kx_thermo_fp64_k', _emit_thermo_fp64, tables, info))
    from . import emit_mixed
    parts.append(_with_pool('kx_bk1_mixed_k', emit_mixed.emit, tables, info, o['mixed'],
                            lds=o['mixed_lds_pool']))
    if o['bk1_live']:
        parts.append(_with_pool('kx_bk1_mixed_valu_k', _emit_bk1_live, tables, info, 'mixed_valu',
                                o['mixed_arrh32'], order, o['mixed_waves'], lds=o['lds_pool']))
    else:
        parts.append(_with_pool('kx_bk1_mixed_valu_k', _emit_bk1_fp64, tables, info, 0, True, 'mixed_valu',
                                o['mixed_arrh32'], lds=o['lds_pool']))
    if 'cond' in tables:
        from . import emit_surrogate
        parts.append(_with_pool('kx_sur_k', emit_surrogate.emit, tables, info))
    return '\n\n'.join(parts) + '\n'


def write(out_dir, mech, tables, opts=None):
    """kernels.hip (every kernel, then the reference-structured kernels over the routine
    files) and the routine files themselves (out_dir/routines, emit_routines)."""
    from . import emit_routines
    from .tables import meta
    info = meta(mech)
    src = generate(tables, info, opts)
    emit_routines.write(out_dir, tables, info)
    with open(os.path.join(out_dir, 'kernels.hip'), 'w') as fh:
        fh.write(src + '\n' + emit_routines.kernels(tables, info))
