"""Seeded synthetic thermochemical states in the reference's structure-of-arrays layout.

Layout (benchmark/src/bk.cpp:606-648): state[0, id] = T/T_ref, state[1 + k, id] = Y_k,
i.e. a (N+1) x n row-major FP64 array with offsetT = offset = n.

Composition: Y = normalise(0.5 * Dirichlet(alpha=0.5 over all species) + 0.5 * Y_air), where
Y_air is the stoichiometric CH4/air mixture X(CH4:O2:N2) = 1:2:7.52 (Y_N2 ~ 0.72, the
composition of kinetix/ci_data/gri30.initial.cantera).
"""
from __future__ import annotations

import numpy as np

BASE_SEED = 20261202
ONE_ATM = 101325.0


def air_mass_fractions(species, W):
    X = np.zeros(len(species))
    for name, x in (('CH4', 1.0), ('O2', 2.0), ('N2', 7.52)):
        if name in species:
            X[species.index(name)] = x
    if X.sum() == 0:
        raise ValueError('mechanism has none of CH4/O2/N2')
    Y = X * W
    return Y / Y.sum()


def sample(n, species, W, T_range=(1000.0, 2500.0), seed=BASE_SEED, p_range_atm=None,
           dirichlet_alpha=0.5, mix=0.5, dtype=np.float64):
    """Return (state (N+1, n), p_state [Pa] (n,) or None)."""
    rng = np.random.default_rng(seed)
    N = len(species)
    T = rng.uniform(T_range[0], T_range[1], size=n)
    # Dirichlet(alpha) row by row via normalised Gamma(alpha) draws, generated species-major
    # so the SoA layout comes out without a transpose.
    G = rng.standard_gamma(dirichlet_alpha, size=(N, n))
    G /= G.sum(axis=0, keepdims=True)
    Yair = air_mass_fractions(species, W)
    state = np.empty((N + 1, n), dtype=dtype)
    state[0] = T
    Y = (1.0 - mix) * G + mix * Yair[:, None]
    Y /= Y.sum(axis=0, keepdims=True)
    state[1:] = Y
    p = None
    if p_range_atm is not None:
        lo, hi = np.log(p_range_atm[0]), np.log(p_range_atm[1])
        p = ONE_ATM * np.exp(rng.uniform(lo, hi, size=n))
    return state, p


def reference_mass_fractions(species, W):
    """The reference composition of kx_build: stoichiometric CH4/air where the mechanism
    has those species, else equal mass fractions."""
    try:
        return air_mass_fractions(species, W)
    except ValueError:
        return np.full(len(species), 1.0 / len(species))


def sample_any(n, species, W, T_range=(1000.0, 2500.0), seed=BASE_SEED, p_range_atm=None):
    """sample() for any mechanism: the CH4/air blend where it exists, else Dirichlet(0.5)
    alone (synthetic mechanisms have no CH4/O2/N2)."""
    try:
        air_mass_fractions(species, W)
        mix = 0.5
    except ValueError:
        mix = 0.0
    if mix:
        return sample(n, species, W, T_range=T_range, seed=seed, p_range_atm=p_range_atm)
    rng = np.random.default_rng(seed)
    N = len(species)
    T = rng.uniform(T_range[0], T_range[1], size=n)
    G = rng.standard_gamma(0.5, size=(N, n))
    G /= G.sum(axis=0, keepdims=True)
    state = np.empty((N + 1, n))
    state[0] = T
    state[1:] = G
    p = None
    if p_range_atm is not None:
        lo, hi = np.log(p_range_atm[0]), np.log(p_range_atm[1])
        p = ONE_ATM * np.exp(rng.uniform(lo, hi, size=n))
    return state, p


def sample_near_equilibrium_mix(n, species, W, t, T_range=(1000.0, 2500.0), seed=BASE_SEED,
                                frac=0.1, rel=0.01, pool=16384):
    """The mixed-precision workload's states: sample() for all states, then a seeded `frac` of
    them (at random positions) replaced by states within `rel` of chemical equilibrium:
    equilibria (kinetix_amd/equilibrium.py) of up to `pool` base states at their own T and
    1 atm, each reused with an independent per-species perturbation exp(U[-ln(1+rel),
    ln(1+rel)]). Returns (state, near_eq_mask)."""
    from . import equilibrium as eq
    state, _ = sample(n, species, W, T_range=T_range, seed=seed)
    rng = np.random.default_rng(seed + 1)
    m = int(round(frac * n))
    pos = np.sort(rng.choice(n, m, replace=False)) if m else np.zeros(0, dtype=np.int64)
    mask = np.zeros(n, dtype=bool)
    mask[pos] = True
    if m:
        k = min(pool, m)
        base = pos[:k]
        Wc = np.asarray(W, dtype=np.float64)[:, None]
        X = state[1:, base] / Wc
        X /= X.sum(0, keepdims=True)
        Xe = eq.equilibrate(t, state[0, base], np.full(k, ONE_ATM), X)
        pick = rng.integers(0, k, size=m)
        Xp = Xe[:, pick] * np.exp(rng.uniform(-np.log1p(rel), np.log1p(rel), size=(len(species), m)))
        Yp = Xp * Wc
        state[1:, pos] = Yp / Yp.sum(0, keepdims=True)
        state[0, pos] = state[0, base[pick]]
    return state, mask


def product_mass_fractions(species, W):
    """Complete-combustion products of stoichiometric CH4/air: CO2 + 2 H2O + 7.52 N2."""
    X = np.zeros(len(species))
    for name, x in (('CO2', 1.0), ('H2O', 2.0), ('N2', 7.52)):
        if name in species:
            X[species.index(name)] = x
    Y = X * W
    return Y / Y.sum()


def sample_blends(n, species, W, T_range=(300.0, 2500.0), p_range_atm=(0.5, 40.0), seed=BASE_SEED,
                  radical_max=1e-2):
    """Linear blends between CH4/air reactants and CO2/H2O/N2 products, plus the radicals
    OH, H, O at U[0, radical_max] mass fraction each, renormalised."""
    rng = np.random.default_rng(seed)
    N = len(species)
    T = rng.uniform(T_range[0], T_range[1], size=n)
    s = rng.uniform(0.0, 1.0, size=n)
    Ya, Yp = air_mass_fractions(species, W), product_mass_fractions(species, W)
    Y = (1.0 - s)[None, :] * Ya[:, None] + s[None, :] * Yp[:, None]
    for name in ('OH', 'H', 'O'):
        if name in species:
            Y[species.index(name)] += rng.uniform(0.0, radical_max, size=n)
    Y /= Y.sum(axis=0, keepdims=True)
    state = np.empty((N + 1, n))
    state[0] = T
    state[1:] = Y
    lo, hi = np.log(p_range_atm[0]), np.log(p_range_atm[1])
    p = ONE_ATM * np.exp(rng.uniform(lo, hi, size=n))
    return state, p


def sample_transport_set(n, species, W, seed=BASE_SEED, tables=None, near_eq_frac=0.0,
                         near_eq_pool=4096):
    """Surrogate training/inference distribution: half Dirichlet/air mixtures with
    T ~ U[300, 2500] K and p ~ logU[0.5, 40] atm, half reactant/product blends; with
    `tables` and near_eq_frac > 0, that fraction of the states (seeded positions) is
    replaced by states within 1 % of chemical equilibrium at their own T and p (the
    equilibria of up to near_eq_pool of them, each reused with independent perturbations)."""
    h = n // 2
    s1, p1 = sample(h, species, W, T_range=(300.0, 2500.0), seed=seed, p_range_atm=(0.5, 40.0))
    s2, p2 = sample_blends(n - h, species, W, seed=seed + 7919)
    state, p = np.concatenate([s1, s2], axis=1), np.concatenate([p1, p2])
    m = int(round(near_eq_frac * n))
    if tables is not None and m:
        from . import equilibrium as eq
        rng = np.random.default_rng(seed + 104729)
        pos = rng.choice(n, m, replace=False)
        k = min(near_eq_pool, m)
        base = pos[:k]
        Wc = np.asarray(W, dtype=np.float64)[:, None]
        X = state[1:, base] / Wc
        X /= X.sum(0, keepdims=True)
        Xe = eq.equilibrate(tables, state[0, base], p[base], X)
        pick = rng.integers(0, k, size=m)
        Xp = Xe[:, pick] * np.exp(rng.uniform(-np.log1p(0.01), np.log1p(0.01), size=(len(species), m)))
        Yp = Xp * Wc
        state[1:, pos] = Yp / Yp.sum(0, keepdims=True)
        state[0, pos] = state[0, base[pick]]
        p[pos] = p[base[pick]]
    return state, p
