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
