"""Dense per-mechanism tables (the IR that the HIP emitter and the CPU oracle both read).

Every derived constant is computed the way the reference generator computes it before
printing it into the generated routines, so that both consumers see the same numbers:
  * Gibbs coefficients a_i * (1/(i+1) - 1/i) (kinetix/utils/write_utils.py:122-126),
    enthalpy a_i/(i+1) (:117-121), heat capacity a_i;
  * C0 = P_atm/R and 1/C0 = R/P_atm (kinetix/core/reaction_rates.py:1323-1324);
  * Troe 1/(T1+1e-300), 1/(T3+1e-300), T2 = -1 when absent (:963-973), SRI 1/(C+1e-300);
  * Wilke C1 = 1/sqrt(sqrt(8(1+W_k/W_j))), C2 = C1 sqrt(sqrt(W_j/W_k))
    (kinetix/core/mix_transport.py:309-316).
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

from . import mechanism as M
from . import physics as phys
from . import transport_fit as TF

FLOAT_MIN = 1e-300   # kinetix/utils/general_utils.py:235-236


def build(mech, fits=None):
    N, R = mech.n_species, mech.n_reactions
    t = {}
    t['n_species'] = np.array(N)
    t['n_active'] = np.array(mech.n_active)
    t['n_reactions'] = np.array(R)
    t['R'] = np.array(phys.R_GAS)
    t['one_atm_R'] = np.array(phys.ONE_ATM / phys.R_GAS)
    t['R_one_atm'] = np.array(phys.R_GAS / phys.ONE_ATM)
    t['W'] = mech.W.copy()
    t['rcpW'] = np.array([1.0 / w for w in mech.W])
    t['composition'] = mech.composition.copy()
    t['T_mid'] = mech.T_mid.astype(float)

    def gcoef(a):
        return [a[0], a[1] * (1 / 2 - 1), a[2] * (1 / 3 - 1 / 2), a[3] * (1 / 4 - 1 / 3),
                a[4] * (1 / 5 - 1 / 4), a[5], a[6]]

    def hcoef(a):
        return [a[0], a[1] / 2, a[2] / 3, a[3] / 4, a[4] / 5, a[5]]

    t['g_lo'] = np.array([gcoef(a) for a in mech.nasa_lo])
    t['g_hi'] = np.array([gcoef(a) for a in mech.nasa_hi])
    t['h_lo'] = np.array([hcoef(a) for a in mech.nasa_lo])
    t['h_hi'] = np.array([hcoef(a) for a in mech.nasa_hi])
    t['cp_lo'] = mech.nasa_lo[:, :5].copy()
    t['cp_hi'] = mech.nasa_hi[:, :5].copy()
    t['nasa_lo'] = mech.nasa_lo.copy()
    t['nasa_hi'] = mech.nasa_hi.copy()

    rx = mech.reactions
    t['kind'] = np.array([r.kind for r in rx], dtype=np.int32)
    t['reversible'] = np.array([int(r.reversible) for r in rx], dtype=np.int32)
    t['collider'] = np.array([r.collider for r in rx], dtype=np.int32)
    t['has_eff'] = np.array([int(r.efficiencies is not None) for r in rx], dtype=np.int32)
    t['A'] = np.array([r.A for r in rx])
    t['lnA'] = np.array([math.log(r.A) if r.A > 0 else 0.0 for r in rx])
    t['beta'] = np.array([float(r.beta) for r in rx])
    t['Ta'] = np.array([float(r.Ta) for r in rx])
    t['A0'] = np.array([r.A0 for r in rx])
    t['lnA0'] = np.array([math.log(r.A0) if r.A0 > 0 else 0.0 for r in rx])
    t['beta0'] = np.array([float(r.beta0) for r in rx])
    t['Ta0'] = np.array([float(r.Ta0) for r in rx])
    t['nu_f'] = np.array([r.nu_f for r in rx], dtype=np.int32).reshape(R, N)
    t['nu_r'] = np.array([r.nu_r for r in rx], dtype=np.int32).reshape(R, N)
    t['eff'] = np.array([r.efficiencies if r.efficiencies is not None else np.ones(N)
                         for r in rx]).reshape(R, N)
    troe = np.zeros((R, 4))
    sri = np.zeros((R, 5))
    for i, r in enumerate(rx):
        if r.kind == M.TROE:
            A, T3, T1, T2 = r.troe
            troe[i] = (A, 1 / (T3 + FLOAT_MIN), 1 / (T1 + FLOAT_MIN), T2 if T2 < math.inf else -1.0)
        if r.kind == M.SRI:
            a, b, c, d, e = r.sri
            sri[i] = (a, b, 1 / (c + FLOAT_MIN), d, e)
    t['troe'] = troe
    t['sri'] = sri
    plog_off, plog_P, plog_nk, plog_koff, plog_k, plog_A = [0], [], [], [], [], []
    for r in rx:
        if r.kind == M.PLOG:
            for P, rates in r.plog:
                plog_P.append(P)
                plog_nk.append(len(rates))
                plog_koff.append(len(plog_A))
                for (A, b, Ta) in rates:
                    plog_A.append(A)
                    plog_k.append((math.log(A) if A > 0 else 0.0, b, Ta))
        plog_off.append(len(plog_P))
    t['plog_off'] = np.array(plog_off, dtype=np.int32)
    t['plog_P'] = np.array(plog_P, dtype=float)
    t['plog_lnP'] = np.array([math.log(p) for p in plog_P], dtype=float)
    t['plog_nk'] = np.array(plog_nk, dtype=np.int32)
    t['plog_koff'] = np.array(plog_koff, dtype=np.int32)
    t['plog_k'] = np.array(plog_k, dtype=float).reshape(-1, 3)
    t['plog_A'] = np.array(plog_A, dtype=float)

    W = mech.W
    C1 = np.empty((N, N))
    C2 = np.empty((N, N))
    for k in range(N):
        for j in range(N):
            c1 = 1 / math.sqrt(math.sqrt(8 * (1. + W[k] / W[j])))
            C1[k, j] = c1
            C2[k, j] = c1 * math.sqrt(math.sqrt(W[j] / W[k]))
    t['C1'] = C1
    t['C2'] = C2
    if fits is not None:
        t['cond'] = fits.conductivity
        t['visc'] = fits.viscosity
        t['diff'] = fits.diffusivity
        t['rcp_diff'] = np.array(int(fits.reciprocal_diffusivity))
        t['T_fit'] = fits.T_fit
    return t


def meta(mech):
    return {
        'name': mech.name,
        'species': mech.species,
        'n_species': mech.n_species,
        'n_active_species': mech.n_active,
        'n_reactions': mech.n_reactions,
        'elements': mech.elements,
        'equations': [r.equation for r in mech.reactions],
    }


def save(out_dir, mech, tables):
    os.makedirs(out_dir, exist_ok=True)
    np.savez_compressed(os.path.join(out_dir, 'tables.npz'), **tables)
    with open(os.path.join(out_dir, 'mech.json'), 'w') as fh:
        json.dump(meta(mech), fh, indent=1)
    # plain-text metadata for the C++ host library (the role of mech.h / mech.okl)
    with open(os.path.join(out_dir, 'mech.txt'), 'w') as fh:
        fh.write(f'n_species {mech.n_species}\n')
        fh.write(f'n_active_species {mech.n_active}\n')
        fh.write(f'n_reactions {mech.n_reactions}\n')
        fh.write('species ' + ' '.join(mech.species) + '\n')
        fh.write('molar_mass ' + ' '.join(repr(float(w)) for w in mech.W) + '\n')


def load(mech_dir):
    with np.load(os.path.join(mech_dir, 'tables.npz'), allow_pickle=False) as z:
        tables = {k: z[k] for k in z.files}
    with open(os.path.join(mech_dir, 'mech.json')) as fh:
        info = json.load(fh)
    return info, tables
