"""Loader for the reference's Cantera known-answer records (kinetix/ci_data/*.cantera),
following benchmark/src/bk.cpp:271-374, and its pass/fail rules (bk.cpp:89-269)."""
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
STATES = ('initial', 'ignition', 'final', 'ignition.highP')
# max relative error per state (bk.cpp:148, 198-199, 258)
RTOL_THERMO = 5e-7
RTOL_RATES = {'initial': 2e-8, 'ignition': 2e-8, 'final': 5e-5, 'ignition.highP': 5e-5}
RTOL_TRANSPORT = 1e-3


def load(mech, state):
    path = os.path.join(HERE, 'golden', 'ci_data', f'{mech}.{state}.cantera')
    lines = open(path).read().split('\n')
    assert len(lines) >= 13, path
    num = lambda i: np.array([float(v) for v in lines[i].split()])
    d = {'species': lines[0].split(), 'W': num(1), 'T': float(lines[2]), 'P': float(lines[3])}
    X = num(4)
    d['X'] = X / X.sum()
    Mbar = float((d['X'] * d['W']).sum())
    d['Y'] = d['X'] * d['W'] / Mbar
    d['rho'] = float(lines[5])
    d['cp_mean'] = float(lines[6]) / Mbar         # J/kg/K
    d['cp_i'] = num(7) / d['W']                    # J/kg/K
    d['wdot'] = num(8)                             # mol/m^3/s
    d['hrr'] = float(lines[9])
    d['conductivity'] = float(lines[10])
    d['viscosity'] = float(lines[11])
    d['rhoD'] = num(12)
    return d


def rel(a, b):
    return abs((a - b) / b)


def rates_error(mass_rates, hrr, d, n_active):
    """bk.cpp:160-209: HRR and molar rates of active species; |ref| <= 1e-50 -> absolute."""
    errs = [rel(hrr, d['hrr'])]
    for k in range(n_active):
        molar = mass_rates[k] / d['W'][k]
        ref = d['wdot'][k]
        errs.append(rel(molar, ref) if abs(ref) > 1e-50 else abs(molar))
    return max(errs)


def transport_error(cond, visc, rhoD, d):
    errs = [rel(cond, d['conductivity']), rel(visc, d['viscosity'])]
    errs += [rel(rhoD[k], d['rhoD'][k]) for k in range(len(d['rhoD']))]
    return max(errs)


def thermo_error(rho, cp, rhocp, d):
    errs = [rel(rho, d['rho']), rel(rhocp, d['rho'] * d['cp_mean'])]
    errs += [rel(cp[k], d['cp_i'][k]) for k in range(len(d['cp_i']))]
    return max(errs)


def state_column(d):
    return np.concatenate([[d['T']], d['Y']])[:, None]
