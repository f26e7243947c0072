"""The reference's generated device-routine files, written from the IR tables in this
generator's loop form (FP64).

A host code that calls the generated routines from its own per-state kernels (a CFD code's
chemistry source term; the reference's productionRates/transport kernels,
benchmark/okl/productionRates.okl:10-62 and transportProps.okl:11-47) includes these files
exactly as it includes the reference's (examples/KinetiX/Makefile:4-26): same file names,
same routine signatures (reaction_rates.py:1271-1273 and examples/KinetiX/kinetix_test.hpp),
same mech.h names, same qualifier macros (__KINETIX_DEVICE__, __KINETIX_INLINE__,
__KINETIX_CONST__, cfloat), so they compile for a HIP device or, with the reference
example's macros, for the CPU.

  mech.h               n_species, n_active_species, n_reactions, species_names,
                       kinetix_molar_mass[], kinetix_rcp_molar_mass[], __KINETIX_NSPECIES__
  rates.cpp            kinetix_species_rates(lnT, T, T2, T3, T4, rcpT, P, lnP, Ci, wdot)
                       (wdot += nu q: the caller zeroes it)
  enthalpy_RT.cpp      kinetix_enthalpy_RT(T, T2, T3, T4, rcpT, h_RT)
  heat_capacity_R.cpp  kinetix_molar_heat_capacity_R(T, T2, T3, T4, cp_R)
  conductivity.cpp     kinetix_conductivity(rcpMbar, lnT, lnT2, lnT3, lnT4, Xi) -> lambda/sqrt(T)
  viscosity.cpp        kinetix_viscosity(lnT, lnT2, lnT3, lnT4, Xi) -> mu/sqrt(T)
  diffusivity.cpp      kinetix_diffusivity(Mbar, p, TsqrT, lnT, lnT2, lnT3, lnT4, Xi, Dkm)

The arithmetic is this generator's loop form (emit_rolled.emit_bk1_core for the rates: the
w-form net rates, every reaction kind incl. PLOG); the transport routines follow
mix_transport.py:261-471 (conductivity mixing rule, Wilke viscosity, symmetric diffusion pair
sum). Tables are file-scope __KINETIX_CONST__ arrays with a kinetix_<file>_ prefix.
"""
from __future__ import annotations

import os

import numpy as np

from . import emit_hip as E
from . import emit_rolled as ER

HEADER = '''// Generated for mechanism {name}. The caller defines cfloat and the qualifier macros as
// for the reference's generated routines; the defaults below are those of a HIP build.
'''

DEFAULTS = '''#ifndef cfloat
#define cfloat double
#endif
#ifndef __KINETIX_DEVICE__
#define __KINETIX_DEVICE__ __device__
#endif
#ifndef __KINETIX_INLINE__
#define __KINETIX_INLINE__ static inline
#endif
#ifndef __KINETIX_CONST__
#define __KINETIX_CONST__ static __device__ const
#endif
'''


def _const(name, arr, ctype='cfloat'):
    a = np.asarray(arr).reshape(-1)
    if ctype == 'int':
        body = ', '.join(str(int(v)) for v in a)
    elif ctype == 'float':
        body = ', '.join(f'{float(np.float32(v))!r}f' for v in a)
    else:
        body = ', '.join(repr(float(v)) for v in a)
    return f'__KINETIX_CONST__ {ctype} {name}[{max(a.size, 1)}] = {{{body if a.size else "0"}}};'


def _requalify(lines):
    """The loop-form helpers and tables with the reference's qualifier macros."""
    out = []
    for ln in lines:
        ln = ln.replace('__device__ const double ', '__KINETIX_CONST__ cfloat ')
        ln = ln.replace('__device__ const int ', '__KINETIX_CONST__ int ')
        ln = ln.replace('__device__ const float ', '__KINETIX_CONST__ float ')
        ln = ln.replace('__device__ __forceinline__ double ', '__KINETIX_DEVICE__ __KINETIX_INLINE__ cfloat ')
        ln = ln.replace('__device__ double ', '__KINETIX_DEVICE__ __KINETIX_INLINE__ cfloat ')
        out.append(ln)
    return out


def mech_h(t, info):
    N = int(t['n_species'])
    names = ' '.join(info['species'])
    return '\n'.join([
        f'#define n_species {N}',
        f'#define n_active_species {int(t["n_active"])}',
        f'#define species_names_length {len(names) + 1}',
        f'#define n_reactions {int(t["n_reactions"])}',
        f'__KINETIX_CONST__ char species_names[species_names_length] = "{names}";',
        _const('kinetix_molar_mass', t['W']),
        _const('kinetix_rcp_molar_mass', t['rcpW']),
        '#define __KINETIX_NSPECIES__ n_species',
        '#define __KINETIX_NACTIVESPECIES__ n_active_species',
    ]) + '\n'


def rates_cpp(t, info):
    """kinetix_species_rates over the loop-form tables (FP64)."""
    N = int(t['n_species'])
    pre = 'kinetix_rates'
    tabs, nsets, MX, MS = ER._bk1_tables(t, pre)
    S = E._Src()
    for line in tabs:
        S(line, 0)
    S(f'__device__ __forceinline__ double {pre}_pow2(double y) {{ return exp2(y); }}', 0)
    helpers = ER.plog_helpers(pre)
    for line in helpers:
        S(line, 0)
    S('__KINETIX_DEVICE__ __KINETIX_INLINE__ void kinetix_species_rates(const cfloat lnT, const cfloat T, '
      'const cfloat T2, const cfloat T3, const cfloat T4, const cfloat rcpT, const cfloat P, '
      'const cfloat lnP, const cfloat* Ci, cfloat* wdot)', 0)
    S('{', 0)
    S('const double log2T = lnT * 1.4426950408889634;')
    S('const float Tf = (float)T, rcpTf = (float)rcpT, log2Tf = (float)log2T;')
    S('(void)Tf; (void)rcpTf; (void)log2Tf; (void)P; (void)lnP;')
    S(f'double w[{N}], wd[{N}];')
    S(f'for (int k = 0; k < {N}; ++k) w[k] = Ci[k];')
    ER.emit_bk1_core(S, t, pre, nsets, MX, MS, False)
    S(f'for (int k = 0; k < {N}; ++k) wdot[k] += wd[k];')
    S('}', 0)
    return '\n'.join(_requalify(S.text().split('\n'))) + '\n'


def thermo_cpp(t, which):
    N = int(t['n_species'])
    if which == 'h':
        lines = [_const('kinetix_enthalpy_RT_tmid', t['T_mid']),
                 _const('kinetix_enthalpy_RT_lo', t['h_lo']), _const('kinetix_enthalpy_RT_hi', t['h_hi']),
                 '__KINETIX_DEVICE__ __KINETIX_INLINE__ void kinetix_enthalpy_RT(const cfloat T, const cfloat T2, '
                 'const cfloat T3, const cfloat T4, const cfloat rcpT, cfloat* h_RT)',
                 '{',
                 f'  for (int k = 0; k < {N}; ++k) {{',
                 '    const cfloat* a = (T > kinetix_enthalpy_RT_tmid[k] ? kinetix_enthalpy_RT_hi : '
                 'kinetix_enthalpy_RT_lo) + 6 * k;',
                 '    h_RT[k] = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4 + a[5] * rcpT;',
                 '  }', '}']
    else:
        lines = [_const('kinetix_heat_capacity_R_tmid', t['T_mid']),
                 _const('kinetix_heat_capacity_R_lo', t['cp_lo']), _const('kinetix_heat_capacity_R_hi', t['cp_hi']),
                 '__KINETIX_DEVICE__ __KINETIX_INLINE__ void kinetix_molar_heat_capacity_R(const cfloat T, '
                 'const cfloat T2, const cfloat T3, const cfloat T4, cfloat* cp_R)',
                 '{',
                 f'  for (int k = 0; k < {N}; ++k) {{',
                 '    const cfloat* a = (T > kinetix_heat_capacity_R_tmid[k] ? kinetix_heat_capacity_R_hi : '
                 'kinetix_heat_capacity_R_lo) + 5 * k;',
                 '    cp_R[k] = a[0] + a[1] * T + a[2] * T2 + a[3] * T3 + a[4] * T4;',
                 '  }', '}']
    return '\n'.join(lines) + '\n'


def conductivity_cpp(t):
    N = int(t['n_species'])
    return '\n'.join([
        _const('kinetix_conductivity_b', t['cond']),
        '// mix_transport.py:261-294: lambda/sqrt(T) = (sum X_k l_k + 1 / sum X_k / l_k) / 2,',
        '// l_k a degree-4 fit in ln T; rcpMbar is not used',
        '__KINETIX_DEVICE__ __KINETIX_INLINE__ cfloat kinetix_conductivity(cfloat rcpMbar, cfloat lnT, '
        'cfloat lnT2, cfloat lnT3, cfloat lnT4, cfloat Xi[])',
        '{',
        '  (void)rcpMbar;',
        '  cfloat s1 = 0, s2 = 0;',
        f'  for (int k = 0; k < {N}; ++k) {{',
        '    const cfloat* b = kinetix_conductivity_b + 5 * k;',
        '    const cfloat l = b[0] + b[1] * lnT + b[2] * lnT2 + b[3] * lnT3 + b[4] * lnT4;',
        '    s1 += Xi[k] * l;',
        '    s2 += Xi[k] / l;',
        '  }',
        '  return 0.5 * (s1 + 1 / s2);',
        '}']) + '\n'


def viscosity_cpp(t):
    N = int(t['n_species'])
    return '\n'.join([
        _const('kinetix_viscosity_a', t['visc']),
        _const('kinetix_viscosity_C1', t['C1']),
        _const('kinetix_viscosity_C2', t['C2']),
        '// mix_transport.py:297-357, Wilke: mu/sqrt(T) = sum_k X_k e_k^2 / sum_j X_j (C1_kj + C2_kj e_k / e_j)^2,',
        '// e_k a degree-4 fit in ln T of sqrt(mu_k / sqrt(T))',
        '__KINETIX_DEVICE__ __KINETIX_INLINE__ cfloat kinetix_viscosity(cfloat lnT, cfloat lnT2, cfloat lnT3, '
        'cfloat lnT4, cfloat Xi[])',
        '{',
        f'  cfloat e[{N}], re[{N}];',
        f'  for (int k = 0; k < {N}; ++k) {{',
        '    const cfloat* a = kinetix_viscosity_a + 5 * k;',
        '    e[k] = a[0] + a[1] * lnT + a[2] * lnT2 + a[3] * lnT3 + a[4] * lnT4;',
        '    re[k] = 1 / e[k];',
        '  }',
        '  cfloat mu = 0;',
        f'  for (int k = 0; k < {N}; ++k) {{',
        '    cfloat s = 0;',
        f'    for (int j = 0; j < {N}; ++j) {{',
        f'      const cfloat phi = kinetix_viscosity_C1[{N} * k + j] + kinetix_viscosity_C2[{N} * k + j] * e[k] * re[j];',
        '      s += Xi[j] * phi * phi;',
        '    }',
        '    mu += Xi[k] * e[k] * e[k] / s;',
        '  }',
        '  return mu;',
        '}']) + '\n'


def diffusivity_cpp(t):
    N = int(t['n_species'])
    rcp = bool(int(t.get('rcp_diff', 0)))
    return '\n'.join([
        _const('kinetix_diffusivity_d', t['diff']),
        _const('kinetix_diffusivity_W', t['W']),
        '// mix_transport.py:414-471 (symmetric pair loop): S_k = sum_{j != k} X_j / D_kj,',
        '// D_km = T^1.5 (Mbar - W_k X_k) / (p Mbar S_k), D_kj from a degree-4 fit in ln T' +
        (' of its reciprocal' if rcp else ''),
        '__KINETIX_DEVICE__ __KINETIX_INLINE__ void kinetix_diffusivity(cfloat Mbar, cfloat p, cfloat TsqrT, '
        'cfloat lnT, cfloat lnT2, cfloat lnT3, cfloat lnT4, cfloat Xi[], cfloat* Dkm)',
        '{',
        f'  cfloat s1[{N}], s2[{N}];',
        f'  for (int k = 0; k < {N}; ++k) s1[k] = s2[k] = 0;',
        f'  for (int k = 1; k < {N}; ++k)',
        '    for (int j = 0; j < k; ++j) {',
        f'      const cfloat* d = kinetix_diffusivity_d + 5 * ({N} * k + j);',
        '      const cfloat f = d[0] + d[1] * lnT + d[2] * lnT2 + d[3] * lnT3 + d[4] * lnT4;',
        f'      const cfloat r = {"f" if rcp else "1 / f"};',
        '      s1[k] += Xi[j] * r;',
        '      s2[j] += Xi[k] * r;',
        '    }',
        f'  for (int k = 0; k < {N}; ++k)',
        '    Dkm[k] = TsqrT * (Mbar - kinetix_diffusivity_W[k] * Xi[k]) / (p * Mbar * (s1[k] + s2[k]));',
        '}']) + '\n'


FILES = ('mech.h', 'rates.cpp', 'enthalpy_RT.cpp', 'heat_capacity_R.cpp', 'conductivity.cpp',
         'viscosity.cpp', 'diffusivity.cpp')


def write(out_dir, t, info):
    """Writes the routine files into out_dir/routines/; returns that directory."""
    d = os.path.join(out_dir, 'routines')
    os.makedirs(d, exist_ok=True)
    head = HEADER.format(name=info['name'])
    texts = {'mech.h': mech_h(t, info), 'rates.cpp': rates_cpp(t, info),
             'enthalpy_RT.cpp': thermo_cpp(t, 'h'), 'heat_capacity_R.cpp': thermo_cpp(t, 'cp')}
    if 'cond' in t:
        texts.update({'conductivity.cpp': conductivity_cpp(t), 'viscosity.cpp': viscosity_cpp(t),
                      'diffusivity.cpp': diffusivity_cpp(t)})
    for f, body in texts.items():
        with open(os.path.join(d, f), 'w') as fh:
            fh.write(head + body)
    return d


def kernels(t, info):
    """Kernels structured as the reference's productionRates and transport kernels, calling the
    routine files (FP64): the drop-in form, for parity of the routines on the device."""
    has_tr = 'cond' in t
    S = E._Src()
    S('// ---- the reference-structured kernels over the generated routine files', 0)
    S(DEFAULTS.rstrip('\n'), 0)
    for f in FILES:
        if f in ('conductivity.cpp', 'viscosity.cpp', 'diffusivity.cpp') and not has_tr:
            continue
        S(f'#include "{f}"', 0)
    S(f'extern "C" __global__ void __launch_bounds__({E.BLOCK})', 0)
    S('kx_bk1_routines(const int n_states, const long long offsetT, const long long offset,', 0)
    S('                const double pressure_R, const double pressure_,', 0)
    S('                const double* __restrict__ state, double* __restrict__ rates, const double Tref)', 0)
    S('{', 0)
    S(f'const long long id = (long long)blockIdx.x * {E.BLOCK} + threadIdx.x;')
    S('if (id >= n_states) return;')
    S('const cfloat T = Tref * state[id], rcpT = 1 / T, lnT = log(T);')
    S('const cfloat T2 = T * T, T3 = T * T * T, T4 = T * T * T * T;')
    S('cfloat c[__KINETIX_NSPECIES__], wdot[__KINETIX_NSPECIES__];')
    S('cfloat rcpMbar = 0;')
    S('for (int k = 0; k < __KINETIX_NSPECIES__; ++k) {')
    S('  c[k] = fmax(0.0, state[id + offsetT + k * offset]) * kinetix_rcp_molar_mass[k];')
    S('  rcpMbar += c[k];')
    S('}')
    S('const cfloat rho = pressure_R * rcpT / rcpMbar;')
    S('for (int k = 0; k < __KINETIX_NSPECIES__; ++k) { c[k] *= rho; wdot[k] = 0; }')
    S('kinetix_species_rates(lnT, T, T2, T3, T4, rcpT, pressure_, log(pressure_), c, wdot);')
    S('kinetix_enthalpy_RT(T, T2, T3, T4, rcpT, c);')
    S('cfloat sum_h = 0;')
    S('for (int k = 0; k < __KINETIX_NSPECIES__; ++k) {')
    S('  rates[id + offsetT + k * offset] = kinetix_molar_mass[k] * wdot[k];')
    S('  sum_h += wdot[k] * c[k];')
    S('}')
    S(f'rates[id] = {float(-float(t["R"]))!r} * T * sum_h;')
    S('}', 0)
    if has_tr:
        S(f'extern "C" __global__ void __launch_bounds__({E.BLOCK})', 0)
        S('kx_bk2_routines(const int n_states, const long long offsetT, const long long offset,', 0)
        S('                const double pressure, const double* __restrict__ state,', 0)
        S('                double* __restrict__ conductivity, double* __restrict__ viscosity,', 0)
        S('                double* __restrict__ rhoD, const double Tref)', 0)
        S('{', 0)
        S(f'const long long id = (long long)blockIdx.x * {E.BLOCK} + threadIdx.x;')
        S('if (id >= n_states) return;')
        S('const cfloat T = Tref * state[id], lnT = log(T), rcpT = 1 / T, sqrT = sqrt(T);')
        S('const cfloat lnT2 = lnT * lnT, lnT3 = lnT * lnT * lnT, lnT4 = lnT * lnT * lnT * lnT;')
        S('cfloat X[__KINETIX_NSPECIES__], D[__KINETIX_NSPECIES__];')
        S('cfloat rcpMbar = 0;')
        S('for (int k = 0; k < __KINETIX_NSPECIES__; ++k) {')
        S('  X[k] = fmax(0.0, state[id + offsetT + k * offset]) * kinetix_rcp_molar_mass[k];')
        S('  rcpMbar += X[k];')
        S('}')
        S('const cfloat Mbar = 1 / rcpMbar;')
        S('for (int k = 0; k < __KINETIX_NSPECIES__; ++k) X[k] *= Mbar;')
        S('conductivity[id] = sqrT * kinetix_conductivity(rcpMbar, lnT, lnT2, lnT3, lnT4, X);')
        S('viscosity[id] = sqrT * kinetix_viscosity(lnT, lnT2, lnT3, lnT4, X);')
        S('kinetix_diffusivity(Mbar, pressure, T * sqrT, lnT, lnT2, lnT3, lnT4, X, D);')
        S(f'const cfloat rho = pressure * {1.0 / float(t["R"])!r} * rcpT * Mbar;')
        S('for (int k = 0; k < __KINETIX_NSPECIES__; ++k) rhoD[k * offset + id] = rho * D[k];')
        S('}', 0)
    return S.text() + '\n'
