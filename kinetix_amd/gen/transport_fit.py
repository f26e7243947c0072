"""Kinetic-theory pure-species / binary transport properties and their lnT polynomial fits.

Restates the reference generator's mixture-averaged transport model
(kinetix/core/mix_transport.py:73-207, kinetix/core/constants.py:155-344):

  * pair well depth / reduced dipole / polar-nonpolar correction xi (mix_transport.py:73-99);
  * reduced collision integrals Omega22*, A* (hence Omega11* = Omega22*/A*) by quadratic
    Lagrange interpolation in ln T* over three table rows, each row being the degree-6
    delta* fit evaluated with its first six coefficients only (mix_transport.py:101-153);
  * species viscosity, binary p*D_jk, and species conductivity with the Warnatz/Parker
    rotational-relaxation model (mix_transport.py:155-180);
  * 50-point fits over [max T_low, min T_high], degree 4 in ln T, weights 1/|y|
    (mix_transport.py:182-207, kinetix/utils/general_utils.py:224-231).

Evaluations are vectorised over the 50 fit temperatures; each element follows the same
arithmetic as the scalar reference formulas.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import physics as phys

N_FIT_POINTS = 50


def _cube(x):
    return x * x * x


class KineticTheory:
    def __init__(self, mech):
        self.m = mech
        self.W = mech.W
        self.eps = mech.well_depth
        self.sigma = mech.diameter
        self.mu = mech.dipole
        self.alpha = mech.polarizability
        self.rot = mech.rot_relax
        self.dof = mech.geometry_dof
        self.lnTs = np.log(np.array(phys.T_STAR))[1:-1]   # tabulated ln T*, T* in [0.1, 100]

    # --- pair parameters -------------------------------------------------------------
    def xi(self, j, k):
        mu, alpha, sigma, eps = self.mu, self.alpha, self.sigma, self.eps
        if (mu[j] > 0.0) == (mu[k] > 0.0):
            return 1.0
        p, q = (j, k) if mu[j] != 0.0 else (k, j)
        return (1.0 + 1.0 / 4.0 * alpha[q] / _cube(sigma[q]) *
                (mu[p] / np.sqrt(4.0 * np.pi * phys.EPSILON0 * eps[p] * _cube(sigma[p]))) ** 2
                * np.sqrt(eps[p] / eps[q]))

    def epsilon(self, j, k):
        return np.sqrt(self.eps[j] * self.eps[k]) * self.xi(j, k) ** 2

    def T_star(self, j, k, T):
        return T * phys.K_BOLTZMANN / self.epsilon(j, k)

    def reduced_mass(self, j, k):
        W, NA = self.W, phys.AVOGADRO
        return W[j] / NA * W[k] / NA / (W[j] / NA + W[k] / NA)

    def sigma_star(self, j, k):
        return (self.sigma[j] + self.sigma[k]) / 2.0 * pow(self.xi(j, k), -1.0 / 6.0)

    def delta_star(self, j, k):
        eps, mu, sigma = self.eps, self.mu, self.sigma
        return (0.5 * mu[j] * mu[k] / (4.0 * np.pi * phys.EPSILON0 * np.sqrt(eps[j] * eps[k]) *
                                       _cube((sigma[j] + sigma[k]) / 2.0)))

    # --- collision integrals ---------------------------------------------------------
    def _collision(self, row_offset, table, fit, j, k, T):
        """Quadratic Lagrange interpolation in ln T* over three tabulated rows."""
        lnT = np.log(self.T_star(j, k, T))
        lnTs = self.lnTs
        n = len(lnTs)
        out = np.empty_like(lnT)
        dstar = self.delta_star(j, k)
        for i, x0 in enumerate(lnT):
            first = 0
            for idx, v in enumerate(lnTs):
                if x0 < v:
                    first = idx
                    break
            s = max(first - 1, 0)
            if s + 3 > n - 1:
                s = n - 4
            x = lnTs[s:s + 3]
            if dstar == 0.0:
                y = [table[s + row_offset + r][0] for r in range(3)]
            else:
                y = [np.dot(fit[s + row_offset + r][:6], [pow(dstar, e) for e in range(6)])
                     for r in range(3)]
            L0 = ((x0 - x[1]) * (x0 - x[2])) / ((x[0] - x[1]) * (x[0] - x[2]))
            L1 = ((x0 - x[0]) * (x0 - x[2])) / ((x[1] - x[0]) * (x[1] - x[2]))
            L2 = ((x0 - x[0]) * (x0 - x[1])) / ((x[2] - x[0]) * (x[2] - x[1]))
            out[i] = L0 * y[0] + L1 * y[1] + L2 * y[2]
        return out

    def omega22(self, j, k, T):
        return self._collision(0, phys.OMEGA22_TABLE, phys.OMEGA22_FIT, j, k, T)

    def a_star(self, j, k, T):
        return self._collision(1, phys.A_STAR_TABLE, phys.A_STAR_FIT, j, k, T)

    def omega11(self, j, k, T):
        return self.omega22(j, k, T) / self.a_star(j, k, T)

    # --- pure species and binary properties ------------------------------------------
    def viscosity(self, k, T):
        W, sigma = self.W, self.sigma
        return (5.0 / 16.0 * np.sqrt(np.pi * W[k] / phys.AVOGADRO * phys.K_BOLTZMANN * T) /
                (np.pi * sigma[k] ** 2 * self.omega22(k, k, T)))

    def p_diffusivity(self, j, k, T):
        """p * D_jk [Pa m^2/s]."""
        return (3.0 / 16.0 * np.sqrt(2.0 * np.pi / self.reduced_mass(j, k)) *
                np.power(phys.K_BOLTZMANN * T, 3.0 / 2.0) /
                (np.pi * self.sigma_star(j, k) ** 2 * self.omega11(j, k, T)))

    def cp_R(self, k, T):
        a = np.where((T < self.m.T_mid[k])[:, None], self.m.nasa_lo[k][None, :],
                     self.m.nasa_hi[k][None, :])
        return a[:, 0] + a[:, 1] * T + a[:, 2] * T * T + a[:, 3] * T * T * T + a[:, 4] * T * T * T * T

    def conductivity(self, k, T):
        W, eps, rot, dof = self.W, self.eps, self.rot, self.dof
        kB, NA = phys.K_BOLTZMANN, phys.AVOGADRO
        visc = self.viscosity(k, T)
        f_vib = W[k] / NA / (kB * T) * self.p_diffusivity(k, k, T) / visc
        Ts = self.T_star(k, k, T)

        def parker(ts):
            return (1.0 + pow(np.pi, 3.0 / 2.0) / np.sqrt(ts) * (1.0 / 2.0 + 1.0 / ts)
                    + (1.0 / 4.0 * np.pi ** 2 + 2.0) / ts)

        A = 5.0 / 2.0 - f_vib
        B = rot[k] * parker(298.0 * kB / eps[k]) / parker(Ts) + 2.0 / np.pi * (5.0 / 3.0 * dof[k] + f_vib)
        f_rot = f_vib * (1.0 + 2.0 / np.pi * A / B)
        f_trans = 5.0 / 2.0 * (1.0 - 2.0 / np.pi * A / B * dof[k] / (3.0 / 2.0))
        Cv = self.cp_R(k, T) - 5.0 / 2.0 - dof[k]
        return (visc / (W[k] / NA)) * kB * (f_trans * 3.0 / 2.0 + f_rot * dof[k] + f_vib * Cv)


def fit_temperatures(mech):
    T_min = max(r[0] for r in mech.T_range)
    T_max = min(r[1] for r in mech.T_range)
    return np.linspace(T_min, T_max, N_FIT_POINTS)


def _polyfit(x, y, deg=4):
    return np.polynomial.polynomial.polyfit(x, y, deg=deg, w=[1.0 / abs(v) for v in y])


@dataclass
class TransportFits:
    conductivity: np.ndarray     # (N, 5): fit of lambda_k / sqrt(T)
    viscosity: np.ndarray        # (N, 5): fit of sqrt(mu_k / sqrt(T))
    diffusivity: np.ndarray      # (N, N, 5) symmetric: fit of p D_jk / T^1.5 (or its reciprocal)
    reciprocal_diffusivity: bool
    T_fit: np.ndarray


def fit(mech, reciprocal_diffusivity=False, progress=None):
    kt = KineticTheory(mech)
    T = fit_temperatures(mech)
    lnT = np.log(T)
    N = mech.n_species
    cond = np.array([_polyfit(lnT, kt.conductivity(k, T) / np.sqrt(T)) for k in range(N)])
    visc = np.array([_polyfit(lnT, np.sqrt(kt.viscosity(k, T) / np.sqrt(T))) for k in range(N)])
    diff = np.zeros((N, N, 5))
    T15 = T * np.sqrt(T)
    for k in range(N):
        for j in range(k):
            pD = kt.p_diffusivity(k, j, T)      # lower-triangle pair (k > j)
            y = T15 / pD if reciprocal_diffusivity else pD / T15
            c = _polyfit(lnT, y)
            diff[k, j] = c
            diff[j, k] = c
        if progress:
            progress(k, N)
    return TransportFits(cond, visc, diff, reciprocal_diffusivity, T)
