"""The CPU oracle pinned against (1) the reference's own compiled GRI-3.0 routines (golden
vectors), (2) the reference's Cantera known-answer states with the reference's tolerances,
(3) conservation invariants. CPU only."""
import numpy as np
import pytest

import kat
from conftest import cond_err


def test_species_order_matches_reference(gri, golden):
    assert list(golden['species']) == gri[0]['species']


def test_rates_vs_reference_routines(oracle, golden):
    s, p = golden['state_bk1'], float(golden['p_bk1'])
    out, scale = oracle.production_rates(s, p, scale=True)
    # strict IEEE build of the reference routines: rounding-level agreement
    assert cond_err(out, golden['rates_strict'], scale) < 1e-13
    # the reference SERIAL flags (-ffast-math): its own fast-vs-strict spread is ~1e-12
    spread = cond_err(golden['rates_fast'], golden['rates_strict'], scale)
    assert spread < 2e-12
    assert cond_err(out, golden['rates_fast'], scale) < 2e-12


def test_transport_vs_reference_routines(oracle, golden):
    s = golden['state_bk2']
    cond, visc, rhoD, _ = oracle.transport(s, 1.0)
    # the transport fits are regenerated from the YAML (numpy polyfit); their values agree
    # with the reference's literal coefficients to ~3e-13 (LAPACK-level noise)
    assert cond_err(cond, golden['cond_strict']) < 1e-12
    assert cond_err(visc, golden['visc_strict']) < 1e-12
    assert cond_err(rhoD, golden['rhoD_strict']) < 1e-12


def test_thermo_vs_reference_routines(oracle, golden):
    s, p = golden['state_bk1'], float(golden['p_bk1'])
    rho, cp, rhocp = oracle.thermo(s, p)
    assert cond_err(rho, golden['rho_strict']) < 1e-14
    assert cond_err(cp, golden['cp_strict']) < 1e-14
    assert cond_err(rhocp, golden['rhocp_strict']) < 1e-14


@pytest.mark.parametrize('state', kat.STATES)
def test_known_answers(oracle, gri, state):
    info, t = gri
    d = kat.load('gri30', state)
    assert d['species'] == info['species']
    assert np.allclose(d['W'], t['W'], rtol=1e-7)          # bk.cpp:291-304
    s = kat.state_column(d)
    r = oracle.production_rates(s, d['P'])
    assert kat.rates_error(r[1:, 0], r[0, 0], d, info['n_active_species']) < kat.RTOL_RATES[state]
    cond, visc, rhoD, _ = oracle.transport(s, 1.0)
    assert kat.transport_error(cond[0], visc[0], rhoD[:, 0], d) < kat.RTOL_TRANSPORT
    rho, cp, rhocp = oracle.thermo(s, d['P'])
    assert kat.thermo_error(rho[0], cp[:, 0], rhocp[0], d) < kat.RTOL_THERMO


def test_mass_and_element_conservation(oracle, gri, golden):
    info, t = gri
    s, p = golden['state_bk1'], float(golden['p_bk1'])
    out, scale = oracle.production_rates(s, p, scale=True)
    mass = out[1:]                          # W_k wdot_k
    tot = np.abs(mass).sum(axis=0)
    assert np.all(np.abs(mass.sum(axis=0)) <= 1e-12 * np.maximum(tot, 1e-300))
    molar = mass / t['W'][:, None]
    elem = t['composition'] @ molar         # atoms of each element produced: zero
    atoms = t['composition'] @ np.abs(molar)
    assert np.all(np.abs(elem) <= 1e-12 * np.maximum(atoms, 1e-300))
    assert np.all(out[1 + info['species'].index('AR')] == 0.0)      # inert


def test_detailed_balance_from_gibbs(oracle, gri):
    """At a state where every reaction's concentration quotient equals its K_c the net rates
    vanish. Build one analytically: choose C_k = exp(-g_k(T) + ln(P_atm/RT)) (an 'ideal'
    equilibrium for element-unconstrained species), then prod C^nu = K_c for EVERY reaction
    with K_c = exp(-sum nu g) (P_atm/RT)^dnu. All q_r must be ~0 relative to q_fwd."""
    info, t = gri
    T = 1700.0
    R = float(t['R'])
    lnT = np.log(T)
    g = np.empty(len(info['species']))
    for k in range(len(g)):
        a = t['g_hi'][k] if T > t['T_mid'][k] else t['g_lo'][k]
        g[k] = a[0] * (1 - lnT) + a[1] * T + a[2] * T**2 + a[3] * T**3 + a[4] * T**4 + a[5] / T - a[6]
    C = np.exp(-g) * (101325.0 / (R * T))
    P = C.sum() * R * T
    Y = C * t['W'] / (C * t['W']).sum()
    s = np.concatenate([[T], Y])[:, None]
    q = oracle.reaction_rates(s, P)[:, 0]
    # forward rates for scale: same state, reverse term removed via irreversible copy
    t2 = dict(t)
    t2['reversible'] = np.zeros_like(t['reversible'])
    from oracle import pyoracle as po
    qf = po.Oracle(t2).reaction_rates(s, P)[:, 0]
    rev = t['reversible'].astype(bool)
    assert np.all(np.abs(q[rev]) <= 1e-9 * np.abs(qf[rev]))
