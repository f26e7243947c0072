"""The generated routine files (kinetix_amd/mechs/<m>/routines: mech.h, rates.cpp,
enthalpy_RT.cpp, heat_capacity_R.cpp, conductivity.cpp, viscosity.cpp, diffusivity.cpp) on the
CPU: compiled with g++ under the reference example's macros (examples/KinetiX/Makefile:4-26),
as a host code that includes the reference's generated files would compile them, and driven
like the reference's kernels (tests/routines_driver.cpp). Against the oracle:
production rates within 1e-12 of the reaction flux (forward and reverse) at pressures around
the PLOG points, heat capacities within 1e-14, transport within 1e-12 + 2e-15 kappa (kappa
the state's largest condition number of the ln T fits); for GRI-3.0 also against the
reference's own compiled routines (the golden vectors)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, cond_err, fit_kappa, flux_scale
from kinetix_amd import states
from kinetix_amd.gen import tables
from oracle import pyoracle as po

MECHS = ['gri30', 'LiDryer', 'NH3Konnov_edit', 'chempolimi_edit', 'EtOHKonnov', 'synth40cov']
P_ATM = 101325.0
FLAGS = ['-D__KINETIX_DEVICE__=', '-D__KINETIX_CONST__=const', '-D__KINETIX_INLINE__=static inline',
         '-Dcfloat=double']
_dp = C.POINTER(C.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


@pytest.fixture(scope='module', params=MECHS)
def routines(request, tmp_path_factory):
    name = request.param
    mdir = os.path.join(ROOT, 'kinetix_amd', 'mechs', name)
    rdir = os.path.join(mdir, 'routines')
    if not os.path.isdir(rdir):
        pytest.skip(f'{name}: routine files not generated')
    so = str(tmp_path_factory.mktemp(name) / 'librt.so')
    subprocess.run(['g++', '-std=c++17', '-O2', '-shared', '-fPIC', *FLAGS, '-I', rdir,
                    os.path.join(ROOT, 'tests', 'routines_driver.cpp'), '-o', so],
                   check=True, timeout=600)
    L = C.CDLL(so)
    L.rt_production_rates.argtypes = [C.c_int, C.c_double, C.c_double, _dp, _dp, C.c_double]
    L.rt_heat_capacity_R.argtypes = [C.c_int, _dp, _dp]
    L.rt_transport.argtypes = [C.c_int, C.c_double, _dp, _dp, _dp, _dp, C.c_double]
    info, t = tables.load(mdir)
    assert L.rt_n_species() == int(t['n_species'])
    return name, info, t, L, po.Oracle(t)


def _rates(L, t, s, p):
    s = np.ascontiguousarray(s)
    out = np.full_like(s, np.nan)
    R = float(t['R'])
    L.rt_production_rates(s.shape[1], p / R, p, _ptr(s), _ptr(out), R)
    return out


def test_rates_vs_oracle(routines):
    name, info, t, L, orc = routines
    s, _ = states.sample_any(256, info['species'], t['W'], seed=states.BASE_SEED + 12)
    for p_atm in (0.05, 1.0, 3.0, 20.0, 2000.0):
        p = p_atm * P_ATM
        out = _rates(L, t, s, p)
        ref = orc.production_rates(s, p)
        assert cond_err(out, ref, flux_scale(orc, t, s, p)) < 1e-12, p_atm


def test_thermo_and_transport_vs_oracle(routines):
    name, info, t, L, orc = routines
    N, n = int(t['n_species']), 256
    s, _ = states.sample_any(n, info['species'], t['W'], T_range=(300.0, 2500.0),
                             seed=states.BASE_SEED + 13)
    s = np.ascontiguousarray(s)
    cp = np.full((N, n), np.nan)
    L.rt_heat_capacity_R(n, _ptr(s), _ptr(cp))
    _, cp0, _ = orc.thermo(s, P_ATM)
    R = float(t['R'])
    # the oracle's cp_k is J/(kg K): cp_k / R * W_k is cp_R
    assert cond_err(cp, cp0 * np.asarray(t['W'])[:, None] / R) < 1e-14
    cond, visc, rhoD = np.full(n, np.nan), np.full(n, np.nan), np.full((N, n), np.nan)
    L.rt_transport(n, 1.0, _ptr(s), _ptr(cond), _ptr(visc), _ptr(rhoD), R)
    c0, v0, r0, _ = orc.transport(s, 1.0)
    kappa = np.maximum.reduce([fit_kappa(t[key], s[0]).max(axis=0) for key in ('cond', 'visc', 'diff')])
    tol = 1e-12 + 2e-15 * kappa
    for a, b in ((cond, c0), (visc, v0), (rhoD, r0)):
        a, b = np.atleast_2d(a), np.atleast_2d(b)
        assert np.all((np.abs(a - b) / np.abs(b)).max(axis=0) <= tol)


def test_gri_against_reference_routines(routines, golden):
    """The drop-in files against the reference's own compiled GRI-3.0 routines."""
    name, info, t, L, orc = routines
    if name != 'gri30':
        pytest.skip('golden vectors are GRI-3.0')
    s, p = golden['state_bk1'], float(golden['p_bk1'])
    out = _rates(L, t, s, p)
    _, scale = orc.production_rates(s, p, scale=True)
    assert cond_err(out, golden['rates_strict'], np.maximum(scale, flux_scale(orc, t, s, p))) < 2e-12
    s2 = np.ascontiguousarray(golden['state_bk2'])
    n, N = s2.shape[1], int(t['n_species'])
    cond, visc, rhoD = np.full(n, np.nan), np.full(n, np.nan), np.full((N, n), np.nan)
    L.rt_transport(n, 1.0, _ptr(s2), _ptr(cond), _ptr(visc), _ptr(rhoD), float(t['R']))
    assert cond_err(cond, golden['cond_strict']) < 1e-12
    assert cond_err(visc, golden['visc_strict']) < 1e-12
    assert cond_err(rhoD, golden['rhoD_strict']) < 1e-12
