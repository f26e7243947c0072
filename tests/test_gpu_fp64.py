"""GPU parity of the plain FP64 HIP kernels (BK1 production rates, BK2 transport, thermo)
against the CPU oracle and the reference's own golden vectors, through the C ABI."""
import numpy as np
import pytest
import torch

import kat
from conftest import cond_err
from kinetix_amd import states
from kinetix_amd.api import Kinetix, KinetixError

pytestmark = pytest.mark.gpu
P_ATM = 101325.0


@pytest.fixture(scope='module')
def kx(gri):
    info, t = gri
    k = Kinetix('gri30', device=0)
    assert k.speciesNames() == info['species']
    # absolute units: T_ref = 1 K, p_ref = 1 atm, Y_ref = CH4/air
    k.build(P_ATM, 1.0, states.air_mass_fractions(info['species'], t['W']))
    yield k
    k.close()


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def run_bk1(kx, s, p, mixed=False):
    n = s.shape[1]
    ds = dev(s)
    dr = torch.full_like(ds, np.nan)
    kx.productionRates(n, n, n, p / kx.refPressure(), ds, dr, mixed=mixed)
    torch.cuda.synchronize()
    return dr.cpu().numpy()


def run_bk2(kx, s, p=1.0, mfma=False):
    n = s.shape[1]
    ds = dev(s)
    N = kx.nSpecies()
    visc, cond = torch.zeros(n, dtype=torch.float64, device='cuda'), torch.zeros(n, dtype=torch.float64, device='cuda')
    rhoD = torch.zeros(N * n, dtype=torch.float64, device='cuda')
    kx.mixtureAvgTransportProps(n, n, n, p, ds, visc, cond, rhoD, mfma=mfma)
    torch.cuda.synchronize()
    return cond.cpu().numpy(), visc.cpu().numpy(), rhoD.cpu().numpy().reshape(N, n)


def test_bk1_fp64_golden(kx, oracle, golden):
    s, p = golden['state_bk1'], float(golden['p_bk1'])
    out = run_bk1(kx, s, p)
    ref, scale = oracle.production_rates(s, p, scale=True)
    assert cond_err(out, ref, scale) < 1e-12
    assert cond_err(out, golden['rates_strict'], scale) < 1e-12
    assert cond_err(out, golden['rates_fast'], scale) < 2e-12


def test_bk2_fp64_golden(kx, oracle, golden):
    s = golden['state_bk2']
    cond, visc, rhoD = run_bk2(kx, s)
    c0, v0, r0, _ = oracle.transport(s, 1.0)
    assert cond_err(cond, c0) < 1e-12
    assert cond_err(visc, v0) < 1e-12
    assert cond_err(rhoD, r0) < 1e-12          # NaN (pure-species 0/0) at the same places
    assert cond_err(rhoD, golden['rhoD_strict']) < 2e-12


def test_bk2_pstate_dkm(kx, oracle, golden):
    s, pst = golden['state_bk2'], golden['p_bk2']
    n, N = s.shape[1], kx.nSpecies()
    ds, dp = dev(s), dev(pst)
    z = lambda m: torch.zeros(m, dtype=torch.float64, device='cuda')
    visc, cond, rhoD, dkm = z(n), z(n), z(N * n), z(N * n)
    kx.mixtureAvgTransportPropsPstate(n, n, n, dp, ds, visc, cond, rhoD, dkm)
    torch.cuda.synchronize()
    c0, v0, r0, d0 = oracle.transport(s, 1.0, p_state=pst)
    assert cond_err(cond.cpu().numpy(), c0) < 1e-12
    assert cond_err(visc.cpu().numpy(), v0) < 1e-12
    assert cond_err(rhoD.cpu().numpy().reshape(N, n), r0) < 1e-12
    assert cond_err(dkm.cpu().numpy().reshape(N, n), d0) < 1e-12


def test_thermo_fp64_golden(kx, golden):
    s, p = golden['state_bk1'], float(golden['p_bk1'])
    n, N = s.shape[1], kx.nSpecies()
    ds = dev(s)
    z = lambda m: torch.zeros(m, dtype=torch.float64, device='cuda')
    rho, cp, rhocp = z(n), z(N * n), z(n)
    kx.thermodynamicProps(n, n, n, p / kx.refPressure(), ds, rho, cp, rhocp)
    torch.cuda.synchronize()
    assert cond_err(rho.cpu().numpy(), golden['rho_strict']) < 1e-13
    assert cond_err(cp.cpu().numpy().reshape(N, n), golden['cp_strict']) < 1e-13
    assert cond_err(rhocp.cpu().numpy(), golden['rhocp_strict']) < 1e-13


@pytest.mark.parametrize('state', kat.STATES)
def test_known_answers_gpu(gri, state):
    """The reference's CI mode (bk.cpp --cimode): T_ref and p_ref from the first state;
    the highP state keeps p_ref = 101325 so that pressure != 1 (bk.cpp:639)."""
    info, t = gri
    d = kat.load('gri30', state)
    k = Kinetix('gri30')
    try:
        p_ref = 101325.0 if state == 'ignition.highP' else d['P']
        k.build(p_ref, d['T'], d['Y'])
        s = kat.state_column(d)
        s[0] /= d['T']
        r = run_bk1(k, s, d['P'])
        assert kat.rates_error(r[1:, 0], r[0, 0], d, info['n_active_species']) < kat.RTOL_RATES[state]
        cond, visc, rhoD = run_bk2(k, s, d['P'] / p_ref)
        assert kat.transport_error(cond[0], visc[0], rhoD[:, 0], d) < kat.RTOL_TRANSPORT
    finally:
        k.close()


def test_bk1_cfg0_seeded(kx, oracle, gri):
    """65,536 seeded states (T ~ U[1000,2500] K, 1 atm): a 2,048-state sample against the
    oracle, plus mass conservation on every state."""
    info, t = gri
    s, _ = states.sample(65536, info['species'], t['W'], T_range=(1000.0, 2500.0),
                         seed=states.BASE_SEED)
    out = run_bk1(kx, s, P_ATM)
    idx = np.random.default_rng(0).choice(s.shape[1], 2048, replace=False)
    ref, scale = oracle.production_rates(s[:, idx], P_ATM, scale=True, threads=8)
    assert cond_err(out[:, idx], ref, scale) < 1e-12
    mass = out[1:]
    assert np.all(np.abs(mass.sum(0)) <= 1e-11 * np.abs(mass).sum(0))


def test_bk2_cfg1_seeded(kx, oracle, gri):
    """The plain-transport workload at its full size: 1,048,576 states, T ~ U[300, 2500] K and
    a per-state p ~ logU[0.5, 40] atm through the per-state-pressure entry point (with D_km),
    against the oracle at the same per-state pressures on a sample; the scalar-p kernel on
    the same states gives bit-identical viscosity and conductivity and the same rho*D (none
    of them depends on p)."""
    info, t = gri
    n, N = 1 << 20, kx.nSpecies()
    s, p = states.sample(n, info['species'], t['W'], T_range=(300.0, 2500.0),
                         seed=states.BASE_SEED + 1, p_range_atm=(0.5, 40.0))
    assert p.shape == (n,) and p.min() >= 0.5 * P_ATM * (1 - 1e-12) and p.max() <= 40 * P_ATM * (1 + 1e-12)
    ds, dp = dev(s), dev(p)
    z = lambda m: torch.full((m,), np.nan, dtype=torch.float64, device='cuda')
    visc, cond, rhoD, dkm = z(n), z(n), z(N * n), z(N * n)
    kx.mixtureAvgTransportPropsPstate(n, n, n, dp, ds, visc, cond, rhoD, dkm)
    torch.cuda.synchronize()
    cond, visc = cond.cpu().numpy(), visc.cpu().numpy()
    rhoD, dkm = rhoD.cpu().numpy().reshape(N, n), dkm.cpu().numpy().reshape(N, n)
    assert np.all(np.isfinite(cond)) and np.all(np.isfinite(visc)) and np.all(np.isfinite(dkm))
    idx = np.random.default_rng(1).choice(n, 2048, replace=False)
    c0, v0, r0, d0 = oracle.transport(s[:, idx], 1.0, p_state=p[idx], threads=8)
    assert cond_err(cond[idx], c0) < 1e-12
    assert cond_err(visc[idx], v0) < 1e-12
    assert cond_err(rhoD[:, idx], r0) < 1e-12
    assert cond_err(dkm[:, idx], d0) < 1e-12
    # the scalar-p kernel on the same states: mu, lambda and rho*D do not depend on p
    c2, v2, r2 = run_bk2(kx, s)
    assert np.array_equal(c2, cond) and np.array_equal(v2, visc)
    # rho and D each carry p (the scalar kernel gets p / p_ref = 1): the product agrees to
    # rounding
    assert cond_err(r2, rhoD) < 1e-14


def test_subview_offsets_and_ragged(kx, golden):
    """Sub-view of a larger SoA buffer (pointer offset + global strides, the reference's
    offsetT/offset arguments) and a ragged state count give the same per-state results."""
    s, p = golden['state_bk1'], float(golden['p_bk1'])
    n = s.shape[1]
    full = run_bk1(kx, s, p)
    ds = dev(s).reshape(-1)
    dr = torch.full_like(ds, np.nan)
    lo, m = 37, 301                         # not multiples of the workgroup size
    kx.productionRates(m, n, n, p / kx.refPressure(), ds[lo:], dr[lo:])
    torch.cuda.synchronize()
    part = dr.cpu().numpy().reshape(s.shape)
    assert np.array_equal(part[:, lo:lo + m], full[:, lo:lo + m])
    assert np.all(np.isnan(part[:, :lo])) and np.all(np.isnan(part[:, lo + m:]))


def test_bk2_mfma_golden(kx, oracle, golden):
    """The Wilke sum on FP64 MFMA (kx_mixture_avg_transport_mfma) against the oracle and the
    reference's own routines, as the VALU kernel."""
    s = golden['state_bk2']
    cond, visc, rhoD = run_bk2(kx, s, mfma=True)
    c0, v0, r0, _ = oracle.transport(s, 1.0)
    assert cond_err(cond, c0) < 1e-12
    assert cond_err(visc, v0) < 1e-12
    assert cond_err(rhoD, r0) < 1e-12
    assert cond_err(visc, golden['visc_strict']) < 2e-12
    assert cond_err(rhoD, golden['rhoD_strict']) < 2e-12


@pytest.mark.parametrize('n', [131072, 1000, 257, 65, 1])
def test_bk2_mfma_seeded_and_ragged(kx, oracle, gri, n):
    """Seeded T ~ U[300, 2500] K states, state counts that end inside a wavefront (the lanes
    past the end take part in the MFMAs and store nothing): against the oracle on a sample
    and against the VALU kernel on every state."""
    info, t = gri
    s, _ = states.sample(n, info['species'], t['W'], T_range=(300.0, 2500.0),
                         seed=states.BASE_SEED + 11 + n, p_range_atm=(0.5, 40.0))
    cond, visc, rhoD = run_bk2(kx, s, mfma=True)
    c1, v1, r1 = run_bk2(kx, s)
    assert cond_err(cond, c1) < 1e-14 and cond_err(rhoD, r1) < 1e-14
    assert cond_err(visc, v1) < 1e-13
    idx = np.random.default_rng(2).choice(n, min(n, 1024), replace=False)
    c0, v0, r0, _ = oracle.transport(s[:, idx], 1.0, threads=8)
    assert cond_err(visc[idx], v0) < 1e-12
    assert cond_err(cond[idx], c0) < 1e-12


def test_bk2_mfma_subview(kx, golden):
    """A sub-view of a larger SoA buffer (pointer offset + global strides): the same per-state
    results, nothing written outside the view."""
    s = golden['state_bk2']
    n, N = s.shape[1], kx.nSpecies()
    full = run_bk2(kx, s, mfma=True)
    ds = dev(s).reshape(-1)
    lo, m = 37, 301
    z = lambda k: torch.full((k,), np.nan, dtype=torch.float64, device='cuda')
    visc, cond, rhoD = z(n), z(n), z(N * n)
    kx.mixtureAvgTransportProps(m, n, n, 1.0, ds[lo:], visc[lo:], cond[lo:], rhoD[lo:], mfma=True)
    torch.cuda.synchronize()
    visc, cond, rhoD = visc.cpu().numpy(), cond.cpu().numpy(), rhoD.cpu().numpy().reshape(N, n)
    assert np.array_equal(cond[lo:lo + m], full[0][lo:lo + m])
    assert np.array_equal(visc[lo:lo + m], full[1][lo:lo + m])
    assert np.array_equal(rhoD[:, lo:lo + m], full[2][:, lo:lo + m])
    for a in (visc, cond, rhoD.T):
        assert np.all(np.isnan(a[:lo])) and np.all(np.isnan(a[lo + m:]))


def test_reference_queries(kx, gri):
    """refPressure / refTemperature / refMassFractions / refMeanMolecularWeight / molecularWeights
    (kinetix.hpp:91-103) through the C ABI."""
    info, t = gri
    y = states.air_mass_fractions(info['species'], t['W'])
    assert kx.refPressure() == P_ATM and kx.refTemperature() == 1.0
    assert np.array_equal(kx.refMassFractions(), y)
    mbar = 1.0 / np.sum(y / t['W'])
    assert abs(kx.refMeanMolecularWeight() - mbar) <= 1e-15 * mbar
    np.testing.assert_allclose(kx.molecularWeights(), t['W'] / mbar, rtol=1e-15)


def test_zero_states_and_errors(kx):
    ds = torch.zeros(54, dtype=torch.float64, device='cuda')
    kx.productionRates(0, 0, 0, 1.0, ds, ds)           # no-op
    with pytest.raises(KinetixError):
        kx.productionRates(4, 2, 2, 1.0, ds, ds)        # overlapping layout rejected
    k2 = Kinetix('gri30')
    try:
        with pytest.raises(KinetixError):                # build() not called
            k2.productionRates(1, 1, 1, 1.0, ds, ds)
    finally:
        k2.close()
