"""Algorithmic work per state of each kernel, counted from a mechanism's IR.

Conventions: +, -, * = 1 flop, fused multiply-add = 2, divide = 1 flop (also tallied under
'divides'); exp/log/log10/exp10/pow/sqrt are tallied as 'transcendentals', not flops;
max() clipping is not counted. Bytes are the FP64 SoA bytes each kernel must move through
HBM (inputs read once, outputs written once). These are the figures bench.py divides by
measured launch time for roofline.achieved.
"""
from __future__ import annotations

import json
import os

import numpy as np

from . import mechanism as M


def _arrh_key(beta, Ta):
    beta, Ta = float(beta), float(Ta)
    if Ta == 0.0 and beta in (0.0, -2.0, -1.0, 1.0, 2.0):
        return None
    return (beta, Ta)


def bk1(t):
    N, R = int(t['n_species']), int(t['n_reactions'])
    net = t['nu_r'] - t['nu_f']
    rev = t['reversible'].astype(bool)
    kind = t['kind']
    f = 0
    tr = 0
    div = 0
    # prologue: rcpT, lnT, T2..T4, rcpT^2; Y->C (N mul, N-1 add, rho: 2 mul + 1 div, N mul)
    f += 1 + 3 + 1 + N + (N - 1) + 3 + N
    div += 2
    tr += 1
    # third-body sums
    f += N - 1
    effsets = {tuple(t['eff'][r]) for r in range(R) if t['has_eff'][r] and kind[r] != M.ELEMENTARY}
    for e in effsets:
        for v in e:
            if v != 1.0:
                f += 1 if v in (0.0, 2.0) else 2
    # Gibbs: poly (1 sub + 1 mul + 5 fma + 1 sub) + exp per needed species, reciprocal
    need_g = {k for r in range(R) if rev[r] for k in range(N) if net[r, k] != 0}
    need_rg = {k for r in range(R) if rev[r] for k in range(N) if net[r, k] < 0}
    f += 13 * len(need_g) + len(need_rg) + 2
    tr += len(need_g)
    div += len(need_rg)
    keys = set()
    for r in range(R):
        ks = [(t['beta'][r], t['Ta'][r])]
        if kind[r] in (M.LINDEMANN, M.TROE, M.SRI):
            ks.append((t['beta0'][r], t['Ta0'][r]))
        for b, Ta in ks:
            k = _arrh_key(b, Ta)
            if k is not None and k not in keys:
                keys.add(k)
                f += (1 if k[0] != 0 else 0) + (1 if k[1] != 0 else 0) + (1 if k[0] and k[1] else 0)
                tr += 1
            f += 1                                  # A * exp or A * T^n
        if kind[r] == M.THREE_BODY:
            f += 1
        elif kind[r] in (M.LINDEMANN, M.TROE, M.SRI):
            f += 3                                  # Pr = k0 M / (kinf + eps)
            div += 1
            if kind[r] == M.LINDEMANN:
                f += 3
                div += 1
            elif kind[r] == M.TROE:
                f += 8 + 2 + 9 + 3 + 4              # Fcent, log Pr, f1, F, k
                div += 2
                tr += 6 if t['troe'][r][3] >= 0 else 5
            else:
                f += 14
                div += 2
                tr += 5
        fwd = int(t['nu_f'][r].sum())
        f += max(fwd - 1, 0)
        if rev[r]:
            rv = int(t['nu_r'][r].sum())
            f += max(rv - 1, 0)
            kr_terms = int(np.abs(net[r]).sum()) + abs(int(net[r].sum()))
            f += max(kr_terms - 1, 0) + 3           # kr*rev, fwd - ., kf * (.)
        else:
            f += 1
        for v in net[r]:
            if v != 0:
                f += 1 if abs(v) == 1 else 2
    active = [k for k in range(N) if np.any(net[:, k] != 0)]
    f += N                                          # W_k * wdot_k
    f += len(active) * (10 + 1 + 2)                 # h/RT poly (5 fma), * wdot, accumulate
    f += 2
    return {'flops': int(f), 'transcendentals': int(tr), 'divides': int(div),
            'bytes': 16 * (N + 1)}


def bk2(t, pstate=False):
    N = int(t['n_species'])
    P = N * (N - 1) // 2
    rcp = bool(int(t.get('rcp_diff', 0)))
    f = 3 + 3                                        # lnT^2..4, rcpT, T*sqrt(T)
    div = 1
    tr = 2                                           # log, sqrt
    f += N + (N - 1) + 1 + N                         # X_k
    div += 1
    f += P * (8 + (0 if rcp else 1) + 4)             # poly, 1/poly, two FMAs per pair
    div += 0 if rcp else P
    f += N * 7                                       # D_km, rho*D
    div += N
    f += N * (8 + 1)                                 # eta_k, 1/eta_k
    div += N
    f += N * 8 + N * N * 6 + N * 4                   # e_k, pair terms, vis accumulation
    div += N
    f += N * (8 + 2 + 1 + 2) + 4                     # conductivity
    div += N + 1
    b = 8 * (N + 1) + 8 * (N + 2)
    if pstate:
        b += 8 + 8 * N
    return {'flops': int(f), 'transcendentals': int(tr), 'divides': int(div), 'bytes': int(b)}


def thermo(t):
    N = int(t['n_species'])
    return {'flops': int(5 + 3 * N + 1 + N * (10 + 2 + 4) + 4), 'transcendentals': 0,
            'divides': 2, 'bytes': 8 * (N + 1) + 8 * (N + 2)}


def all_counts(t):
    from . import emit_mixed
    mixed = dict(bk1(t))
    mixed.update(emit_mixed.mfma_per_state(t))
    out = {'kx_bk1_fp64': bk1(t), 'kx_bk1_mixed': mixed, 'kx_bk1_mixed_valu': bk1(t),
           'kx_thermo_fp64': thermo(t)}
    if 'cond' in t:
        out['kx_bk2_fp64'] = bk2(t)
        out['kx_bk2_fp64_pstate'] = bk2(t, pstate=True)
    return out


def write(out_dir, t):
    with open(os.path.join(out_dir, 'counts.json'), 'w') as fh:
        json.dump(all_counts(t), fh, indent=1)


def load(mech_dir):
    with open(os.path.join(mech_dir, 'counts.json')) as fh:
        return json.load(fh)
