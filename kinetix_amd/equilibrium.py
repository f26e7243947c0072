"""Ideal-gas chemical equilibrium at fixed (T, p, element totals), vectorised over states:
Gibbs-energy minimisation.

Minimising G = sum_k n_k (g_k + ln(n_k / n) + ln(p / p0)) subject to A n = b (A: atoms of
element e in species k, b: element totals), with g_k = G_k(T, p0) / (R T) from the NASA-7
polynomials (the same Gibbs terms as the kernels' equilibrium constants,
K_c = exp(-sum_k nu_k g_k) (p0 / R T)^(sum nu)), by the iteration of Gordon and McBride
(NASA Reference Publication 1311, 1994), started from equal moles of every species.

Used to place states near equilibrium, where the net rates cancel (the mixed-precision
kernels' stress case): at an exact equilibrium every reversible reaction's net rate vanishes.
"""
from __future__ import annotations

import numpy as np

P0 = 101325.0


def gibbs_RT(t, T):
    """g_k(T) = G_k / (R T) at p0 for every species, (N, n)."""
    T = np.asarray(T, dtype=np.float64)
    lnT = np.log(T)
    out = np.empty((int(t['n_species']), T.size))
    glo, ghi = np.asarray(t['g_lo']).reshape(-1, 7), np.asarray(t['g_hi']).reshape(-1, 7)
    tmid = np.asarray(t['T_mid'])
    basis = np.stack([1.0 - lnT, T, T ** 2, T ** 3, T ** 4, 1.0 / T, -np.ones_like(T)])
    lo, hi = glo @ basis, ghi @ basis
    out[:] = np.where(T[None, :] > tmid[:, None], hi, lo)
    return out


def equilibrate(t, T, p, X, iters=300, tol=1e-12):
    """Equilibrium mole fractions (N, n) at T [K], p [Pa] (arrays of n) with the element
    totals of the mole fractions X (N, n).

    The Gibbs-minimisation iteration of Gordon and McBride (NASA RP-1311, 1994, ch. 2) for an
    ideal-gas mixture: with mu_k = g_k + ln(n_k / n) + ln(p / p0), each step solves for the
    Lagrange corrections pi_e and d ln n
        sum_j (sum_k a_ik a_jk n_k) pi_j + (sum_k a_ik n_k) dln n = b_i - sum_k a_ik n_k + sum_k a_ik n_k mu_k
        sum_j (sum_k a_jk n_k) pi_j + (sum_k n_k - n) dln n = n - sum_k n_k + sum_k n_k mu_k
    then d ln n_k = -mu_k + sum_j a_jk pi_j + d ln n, applied with the report's damping
    (2 / the largest correction of a major species or of ln n; trace species limited so they
    cannot jump above a 1e-4 mole fraction in one step)."""
    A = np.asarray(t['composition'], dtype=np.float64)            # (E, N)
    E, N = A.shape
    X = np.asarray(X, dtype=np.float64)
    n = X.shape[1]
    b = A @ X                                                      # (E, n) element totals per mole
    if np.any(b.max(0) <= 0):
        raise ValueError('empty composition')
    # absent or trace elements (no argon, or a 1e-15 argon share) would make the system
    # singular: their totals are floored at 1e-12 of the largest, a change far below every
    # other tolerance here (the species that carry them appear at that trace level)
    b = np.maximum(b, 1e-12 * b.max(0, keepdims=True))
    g = gibbs_RT(t, T) + np.log(np.asarray(p) / P0)[None, :]       # (N, n)
    AA = (A[:, None, :] * A[None, :, :]).reshape(E * E, N)
    lnk = np.full((N, n), np.log(0.1 / N))                         # start: equal moles
    lnn = np.log(np.full(n, 0.1))
    for _ in range(iters):
        nk = np.exp(lnk)
        ntot = np.exp(lnn)
        mu = g + lnk - lnn[None, :]
        M = np.zeros((n, E + 1, E + 1))
        M[:, :E, :E] = (AA @ nk).T.reshape(n, E, E)
        Ank = (A @ nk)                                             # (E, n)
        M[:, :E, E] = Ank.T
        M[:, E, :E] = Ank.T
        M[:, E, E] = nk.sum(0) - ntot
        rhs = np.zeros((n, E + 1))
        rhs[:, :E] = (b - Ank + A @ (nk * mu)).T
        rhs[:, E] = ntot - nk.sum(0) + (nk * mu).sum(0)
        sol = np.linalg.solve(M, rhs[..., None])[..., 0]           # (n, E+1)
        pi, dlnn = sol[:, :E].T, sol[:, E]
        dlnk = -mu + A.T @ pi + dlnn[None, :]
        frac = lnk - lnn[None, :]                                  # ln mole fraction
        major = frac > np.log(1e-8)
        big = np.maximum(np.abs(dlnn), np.where(major, np.abs(dlnk), 0.0).max(0))
        lam1 = 2.0 / np.maximum(big, 1e-300)
        # trace species: ln(n_k/n) + lam (dln n_k - dln n) <= ln 1e-4
        grow = (~major) & (dlnk - dlnn[None, :] > 0)
        with np.errstate(divide='ignore', invalid='ignore'):
            lim = np.where(grow, (np.log(1e-4) - frac) / (dlnk - dlnn[None, :]), np.inf)
        lam = np.minimum(1.0, np.minimum(lam1, np.abs(lim).min(0)))
        lnk += lam[None, :] * dlnk
        lnn += lam * dlnn
        lnk = np.maximum(lnk, -700.0)
        conv = np.maximum(np.abs(dlnn), (nk * np.abs(dlnk)).sum(0) / nk.sum(0))
        if conv.max() < tol:
            break
    x = np.exp(lnk - lnn[None, :])
    return x / x.sum(0, keepdims=True)


def near_equilibrium(t, T, p, Y, rel=0.01, seed=0, chunk=65536):
    """Mass fractions within `rel` of the equilibrium at (T, p) and Y's element totals: each
    equilibrium mole fraction is scaled by exp(U[-ln(1+rel), ln(1+rel)]), then renormalised."""
    W = np.asarray(t['W'])[:, None]
    X = (np.asarray(Y) / W)
    X /= X.sum(0, keepdims=True)
    T, p = np.asarray(T, dtype=np.float64), np.asarray(p, dtype=np.float64) * np.ones(X.shape[1])
    Xe = np.empty_like(X)
    for c in range(0, X.shape[1], chunk):
        Xe[:, c:c + chunk] = equilibrate(t, T[c:c + chunk], p[c:c + chunk], X[:, c:c + chunk])
    rng = np.random.default_rng(seed)
    Xe = Xe * np.exp(rng.uniform(-np.log1p(rel), np.log1p(rel), size=Xe.shape))
    Ye = Xe * W
    return Ye / Ye.sum(0, keepdims=True)
