"""Mixed-precision BK1 (FP32 Arrhenius/falloff, FP64 K_c/net/sums, FP32-MFMA q.nu) against the
FP64 oracle: relative L2 error per species over states <= 1e-5 (the stated accuracy bound)."""
import numpy as np
import pytest
import torch

from conftest import cond_err
from kinetix_amd import states
from kinetix_amd.api import Kinetix

pytestmark = pytest.mark.gpu
P_ATM = 101325.0
L2_BOUND = 1e-5


@pytest.fixture(scope='module')
def kx(gri):
    info, t = gri
    k = Kinetix('gri30')
    k.build(P_ATM, 1.0, states.air_mass_fractions(info['species'], t['W']))
    yield k
    k.close()


def run(kx, s, p, mixed):
    n = s.shape[1]
    ds = torch.from_numpy(np.ascontiguousarray(s)).cuda().reshape(-1)
    dr = torch.full_like(ds, np.nan)
    kx.productionRates(n, n, n, p / kx.refPressure(), ds, dr, mixed=mixed)
    torch.cuda.synchronize()
    return dr.cpu().numpy().reshape(s.shape)


def l2_per_species(x, ref):
    out = []
    for k in range(ref.shape[0]):
        nr = np.linalg.norm(ref[k])
        if nr > 0:
            out.append(np.linalg.norm(x[k] - ref[k]) / nr)
    return np.array(out)


ACCUM = ['mfma', 'valu']


def test_mixed_entry_point_runs_the_fastest_kernel(kx, golden):
    """kx_production_rates_mixed (mixed=True) is kx_bk1_mixed_valu: bit-identical outputs."""
    s, p = golden['state_bk1'], float(golden['p_bk1'])
    assert np.array_equal(run(kx, s, p, mixed=True), run(kx, s, p, mixed='valu'))


@pytest.mark.parametrize('accum', ACCUM)
def test_mixed_golden_states(kx, oracle, golden, accum):
    s, p = golden['state_bk1'], float(golden['p_bk1'])
    out = run(kx, s, p, mixed=accum)
    ref = oracle.production_rates(s, p)
    e = l2_per_species(out[1:], ref[1:])
    assert e.max() < L2_BOUND, e.max()
    hrr = np.linalg.norm(out[0] - ref[0]) / np.linalg.norm(ref[0])
    assert hrr < L2_BOUND


@pytest.mark.parametrize('accum', ACCUM)
@pytest.mark.parametrize('n', [65536, 1000, 257])
def test_mixed_seeded(kx, oracle, gri, n, accum):
    info, t = gri
    s, _ = states.sample(n, info['species'], t['W'], T_range=(1000.0, 2500.0),
                         seed=states.BASE_SEED + n)
    out = run(kx, s, P_ATM, mixed=accum)
    m = min(n, 4096)
    ref, scale = oracle.production_rates(s[:, :m], P_ATM, scale=True, threads=8)
    e = l2_per_species(out[1:, :m], ref[1:])
    assert e.max() < L2_BOUND, e.max()
    # every state, every species: error bounded by the FP32 accumulation scale
    assert cond_err(out[:, :m], ref, scale) < 2e-5
    mass = out[1:]
    assert np.all(np.abs(mass.sum(0)) <= 1e-5 * np.abs(mass).sum(0))
    assert np.all(out[1 + info['species'].index('AR')] == 0.0)


@pytest.mark.parametrize('accum', ACCUM)
def test_mixed_subview(kx, golden, accum):
    s, p = golden['state_bk1'], float(golden['p_bk1'])
    n = s.shape[1]
    full = run(kx, s, p, mixed=accum)
    ds = torch.from_numpy(np.ascontiguousarray(s)).cuda().reshape(-1)
    dr = torch.full_like(ds, np.nan)
    lo, m = 100, 150
    kx.productionRates(m, n, n, p / kx.refPressure(), ds[lo:], dr[lo:], mixed=accum)
    torch.cuda.synchronize()
    part = dr.cpu().numpy().reshape(s.shape)
    # a state's result depends only on that state (tile-mates only share the MFMA)
    np.testing.assert_allclose(part[:, lo:lo + m], full[:, lo:lo + m], rtol=1e-6, atol=0)
    assert np.all(np.isnan(part[:, :lo])) and np.all(np.isnan(part[:, lo + m:]))


@pytest.mark.parametrize('accum', ACCUM)
def test_mixed_near_equilibrium(kx, oracle, gri, accum):
    """States within 1 % of chemical equilibrium (net rates cancel to ~1 % of the forward
    rates): the FP64 K_c and net subtraction keep the error within the same bound."""
    info, t = gri
    s, mask = states.sample_near_equilibrium_mix(40960, info['species'], t['W'], t,
                                                  seed=states.BASE_SEED + 2, pool=1024)
    s = np.ascontiguousarray(s[:, mask])
    out = run(kx, s, P_ATM, mixed=accum)
    ref, scale = oracle.production_rates(s, P_ATM, scale=True, threads=8)
    e = l2_per_species(out[1:], ref[1:])
    assert e.max() < L2_BOUND, e.max()
    assert cond_err(out, ref, scale) < 2e-5
