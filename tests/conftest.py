import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, 'tests', 'golden')


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs an MI355X (runs the HIP kernels)')


@pytest.fixture(scope='session')
def gri():
    from kinetix_amd.gen import tables
    info, t = tables.load(os.path.join(ROOT, 'kinetix_amd', 'mechs', 'gri30'))
    return info, t


@pytest.fixture(scope='session')
def oracle(gri):
    from oracle import pyoracle as po
    if not os.path.exists(os.path.join(ROOT, 'oracle', 'lib', 'libkxoracle.so')):
        po.build_libs(with_ref=False)
    return po.Oracle(gri[1])


@pytest.fixture(scope='session')
def golden():
    with np.load(os.path.join(GOLDEN, 'gri30_ref.npz'), allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def cond_err(x, y, scale=None):
    """max |x-y| / max(|y|, scale) with 0/0 := 0 (condition-aware relative error).
    NaN must sit at the same positions in both (the reference propagates NaN, e.g. the
    0/0 self-diffusion term of a pure species); a NaN in only one of them is an inf error."""
    x, y = np.asarray(x, float), np.asarray(y, float)
    nx, ny = np.isnan(x), np.isnan(y)
    if np.any(nx != ny):
        return float('inf')
    x, y = np.where(nx, 0.0, x), np.where(ny, 0.0, y)
    den = np.abs(y) if scale is None else np.maximum(np.abs(y), np.asarray(scale, float))
    d = np.abs(x - y)
    with np.errstate(divide='ignore', invalid='ignore'):
        r = np.where(d == 0, 0.0, d / np.where(den == 0, np.finfo(float).tiny, den))
    return float(r.max()) if r.size else 0.0


def flux_scale(orc, t, s, p):
    """Condition scale of BK1 outputs (N + 1, n): the reaction flux each output is a sum of,
    forward and reverse separately (net rates cancel near equilibrium and at high pressure):
    W_k sum_r |nu_kr| (q_fwd,r + q_rev,r) for the species rows, R T sum_k flux_k |h_k / RT|
    for the heat release rate."""
    q, qf = orc.reaction_rates(s, p, forward=True)
    net = np.asarray(t['nu_r']) - np.asarray(t['nu_f'])
    flux = np.abs(net).T @ (np.abs(qf) + np.abs(qf - q))                     # (N, n)
    T = s[0]
    powers = np.stack([np.ones_like(T), T, T * T, T ** 3, T ** 4, 1.0 / T])  # (6, n)
    hi = (T[None, :] > np.asarray(t['T_mid'])[:, None])
    hRT = np.where(hi, np.asarray(t['h_hi']) @ powers, np.asarray(t['h_lo']) @ powers)
    hrr = abs(float(t['R'])) * T * (flux * np.abs(hRT)).sum(axis=0)
    return np.vstack([hrr[None, :], np.asarray(t['W'])[:, None] * flux])


def fit_kappa(c, T):
    """Condition number sum_i |c_i| |ln T|^i / |sum_i c_i ln T^i| of degree-4 fits c (..., 5)
    at temperatures T (n,): (..., n)."""
    lT = np.log(T)
    P = np.stack([lT ** i for i in range(5)])
    c = np.asarray(c).reshape(-1, 5)
    c = c[np.any(c != 0.0, axis=1)]                    # pairs without a fit are all zero
    return (np.abs(c) @ np.abs(P)) / np.abs(c @ P)
