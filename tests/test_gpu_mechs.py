"""The reference's other mechanisms on the GPU, through the C ABI, against the oracle:
LiDryer (9 species), NH3Konnov_edit (36, 3 PLOG reactions), chempolimi_edit (31, 21 PLOG),
both straight-line, NH3Konnov_edit in the loop form too, EtOHKonnov (129 species, 1,231
reactions, loop form), and the seeded 40-species coverage mechanism (kinetix_amd/synth.py
--coverage: every reaction form, incl. SRI, Lindemann, explicit colliders, irreversible
reactions and PLOG points with summed rates) in both forms. BK1 FP64 <= 1e-12 of the reaction flux each output sums (forward
and reverse separately) at pressures below, on, between and above the PLOG points; BK1 mixed <= 1e-5 relative L2 per species; BK2 and the
thermodynamic kernel <= 1e-12; and the reference's known-answer states end to end."""
import os

import numpy as np
import pytest
import torch

import kat
from conftest import ROOT, cond_err, fit_kappa, flux_scale
from kinetix_amd import states
from kinetix_amd.api import Kinetix
from kinetix_amd.gen import tables
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu
P_ATM = 101325.0
MECHS = ['LiDryer', 'NH3Konnov_edit', 'NH3Konnov_edit_rolled', 'chempolimi_edit', 'EtOHKonnov',
         'synth40cov', 'synth40cov_rolled']
# below, on and between NH3Konnov's points (0.1, 1, 10 atm); above both mechanisms' last point
PRESSURES_ATM = (0.05, 1.0, 3.0, 20.0, 2000.0)


@pytest.fixture(scope='module', params=MECHS)
def setup(request):
    name = request.param
    info, t = tables.load(os.path.join(ROOT, 'kinetix_amd', 'mechs', name))
    kx = Kinetix(name)
    sp, W = kx.speciesNames(), kx.molar_masses
    kx.build(P_ATM, 1.0, states.reference_mass_fractions(sp, W))
    yield name, info, t, kx, po.Oracle(t)
    kx.close()


def _rates(kx, s, p_norm, mixed=False):
    n = s.shape[1]
    ds = torch.from_numpy(np.ascontiguousarray(s)).cuda().reshape(-1)
    dr = torch.full_like(ds, np.nan)
    kx.productionRates(n, n, n, p_norm, ds, dr, mixed=mixed)
    torch.cuda.synchronize()
    return dr.cpu().numpy().reshape(s.shape)


def _l2(x, ref):
    out = [np.linalg.norm(x[k] - ref[k]) / np.linalg.norm(ref[k])
           for k in range(ref.shape[0]) if np.linalg.norm(ref[k]) > 0]
    return max(out)


@pytest.mark.parametrize('p_atm', PRESSURES_ATM)
def test_fp64_rates_vs_oracle(setup, p_atm):
    name, info, t, kx, orc = setup
    s, _ = states.sample_any(1024, info['species'], t['W'], seed=states.BASE_SEED + 5)
    out = _rates(kx, s, p_atm * P_ATM / kx.refPressure())
    ref = orc.production_rates(s, p_atm * P_ATM, threads=8)
    assert cond_err(out, ref, flux_scale(orc, t, s, p_atm * P_ATM)) < 1e-12


@pytest.mark.parametrize('p_atm', (0.05, 1.0, 20.0))
def test_mixed_rates_vs_oracle(setup, p_atm):
    name, info, t, kx, orc = setup
    s, _ = states.sample_any(2048, info['species'], t['W'], seed=states.BASE_SEED + 5)
    out = _rates(kx, s, p_atm * P_ATM / kx.refPressure(), mixed='valu')
    ref = orc.production_rates(s, p_atm * P_ATM, threads=8)
    assert np.all(np.isfinite(out))
    assert _l2(out[1:], ref[1:]) < 1e-5
    assert np.linalg.norm(out[0] - ref[0]) / np.linalg.norm(ref[0]) < 1e-5


def _rel_per_state(a, b):
    a, b = np.atleast_2d(a), np.atleast_2d(b)
    return (np.abs(a - b) / np.abs(b)).max(axis=0)


def test_transport_and_thermo_vs_oracle(setup):
    """Transport within 1e-12 + 2e-15 kappa of the oracle, kappa the state's largest
    condition number of the ln T fits (two FP64 evaluations of a degree-4 fit may differ by
    ~2 n u kappa): helium's conductivity fit reaches kappa ~ 1e5 in the NH3 and chempolimi
    mechanisms (GRI-3.0's largest is 3e3), where the two evaluation orders part by 2e-12."""
    name, info, t, kx, orc = setup
    N, n = kx.nSpecies(), 1024
    s, p = states.sample_any(n, info['species'], t['W'], T_range=(300.0, 2500.0),
                             seed=states.BASE_SEED + 6, p_range_atm=(0.5, 40.0))
    ds = torch.from_numpy(np.ascontiguousarray(s)).cuda().reshape(-1)
    z = lambda m: torch.full((m,), np.nan, dtype=torch.float64, device='cuda')
    v, c, r = z(n), z(n), z(N * n)
    kx.mixtureAvgTransportProps(n, n, n, 1.0, ds, v, c, r)
    torch.cuda.synchronize()
    c0, v0, r0, _ = orc.transport(s, 1.0, threads=8)
    T = s[0]
    kappa = np.maximum.reduce([fit_kappa(t[key], T).max(axis=0) for key in ('cond', 'visc', 'diff')])
    tol = 1e-12 + 2e-15 * kappa
    assert np.all(_rel_per_state(v.cpu().numpy(), v0) <= tol)
    assert np.all(_rel_per_state(c.cpu().numpy(), c0) <= tol)
    assert np.all(_rel_per_state(r.cpu().numpy().reshape(N, n), r0) <= tol)
    rho, cp, rhocp = z(n), z(N * n), z(n)
    kx.thermodynamicProps(n, n, n, 1.0, ds, rho, cp, rhocp)
    torch.cuda.synchronize()
    rho0, cp0, rhocp0 = orc.thermo(s, kx.refPressure())
    assert cond_err(rho.cpu().numpy(), rho0) < 1e-12
    assert cond_err(cp.cpu().numpy().reshape(N, n), cp0) < 1e-12
    assert cond_err(rhocp.cpu().numpy(), rhocp0) < 1e-12


def test_known_answer_states(setup):
    """The reference's Cantera records through the GPU kernels. The records sit near
    equilibrium (see test_oracle_mechs), where net rates cancel to ~1e-13 of the reaction
    flux sum_r |nu_kr| (q_fwd,r + q_rev,r); both comparisons are on that scale: against
    Cantera's rates and against the oracle (species rows W_k flux_k; the heat release rate
    R T sum_k flux_k |h_k / RT|)."""
    name, info, t, kx, orc = setup
    base = name.replace('_rolled', '')
    if base.startswith('synth'):
        pytest.skip('synthetic mechanism: no known-answer records')
    na = info['n_active_species']
    for state in ('initial', 'ignition', 'final'):
        d = kat.load(base, state)
        s = kat.state_column(d)
        out = _rates(kx, s, d['P'] / kx.refPressure())
        ref = orc.production_rates(s, d['P'])
        scale = flux_scale(orc, t, s, d['P'])
        assert cond_err(out, ref, scale) < 1e-12
        flux = scale[1:, 0] / t['W']
        molar = out[1:, 0] / t['W']
        assert np.all(np.abs(molar[:na] - d['wdot'][:na]) <= 1e-12 * flux[:na] + 1e-300)
