"""The shipped GRI-3.0 transport surrogate (kinetix_amd/mechs/gri30/surrogate.npz, trained on
16,777,216 plain-kernel states) at its inference size: 4,194,304 states of the surrogate
distribution -- T ~ U[300, 2500] K and p ~ logU[0.5, 40] atm over Dirichlet/air mixtures,
plus reactant/product blends with O/H/OH radicals -- through the audited one-call path
kx_sur_transport, judged state by state against the plain FP64 kernel on the same inputs.

Bounds asserted (the plain kernel is the sole judge):
  * mu and lambda within 1e-3 relative on every state (fallback states carry the plain result);
  * rho*D_k within RHOD_MAX relative on every state and species, p99.9 within RHOD_P999;
  * the fallback fraction (out-of-hull + failed audits) below 1e-3, no whole-call fallback;
  * the learned part's gain: the median rho*D error at least 5x below the physics baseline's.
"""
import os

import numpy as np
import pytest
import torch

from kinetix_amd import states
from kinetix_amd.api import Kinetix
from kinetix_amd.surrogate import Surrogate, model_path

pytestmark = pytest.mark.gpu
dev = torch.device('cuda')
N_STATES = 1 << 22
RHOD_MAX = 3e-2
RHOD_P999 = 5e-3


@pytest.fixture(scope='module')
def shipped(gri):
    info, t = gri
    path = model_path('gri30')
    assert os.path.exists(path), path
    kx = Kinetix('gri30')
    kx.build(states.ONE_ATM, 1.0, states.air_mass_fractions(info['species'], t['W']))
    sur = Surrogate(kx, model=path)
    N, n = kx.nSpecies(), N_STATES
    s, p = states.sample_transport_set(n, info['species'], t['W'], seed=states.BASE_SEED + 404)
    ds = torch.from_numpy(np.ascontiguousarray(s)).to(dev).reshape(-1)
    dp = torch.from_numpy(np.ascontiguousarray(p)).to(dev)
    del s
    z = lambda m: torch.full((m,), np.nan, dtype=torch.float64, device=dev)
    out = [z(n), z(n), z(N * n), z(N * n)]
    res = sur.infer(n, n, n, dp, ds, *out)
    ref = sur.plain(n, n, n, dp, ds)
    torch.cuda.synchronize()
    yield kx, sur, res, out, ref, ds, dp
    kx.close()


def _rel(a, b):
    return (a / b - 1.0).abs()


def test_shipped_model_mu_lambda_every_state(shipped):
    _, _, res, (v, c, _, _), (pv, pc, _, _), _, _ = shipped
    ev, ec = _rel(v, pv), _rel(c, pc)
    assert bool(torch.isfinite(ev).all()) and bool(torch.isfinite(ec).all())
    print(f'\nshipped surrogate, {N_STATES} states: mu max {float(ev.max()):.3e}, '
          f'lambda max {float(ec.max()):.3e}, stats {res}')
    assert float(ev.max()) <= 1e-3 and float(ec.max()) <= 1e-3


def test_shipped_model_rhoD_bound_and_fallback(shipped):
    kx, _, res, (_, _, r, d), (_, _, pr, pd), _, _ = shipped
    N, n = kx.nSpecies(), N_STATES
    er = _rel(r.view(N, n), pr.view(N, n))
    # pure-species states have a 0/0 self term in the plain kernel (NaN at the same place)
    both_nan = torch.isnan(r.view(N, n)) & torch.isnan(pr.view(N, n))
    er = torch.where(both_nan, torch.zeros_like(er), er)
    assert bool(torch.isfinite(er).all())
    worst = er.max(dim=0).values                    # per state, over species
    q = torch.quantile(worst[torch.randperm(n, device=dev)[:1 << 20]].float(),
                       torch.tensor([0.5, 0.99, 0.999], device=dev)).tolist()
    print(f'\nrho*D max over species per state: median {q[0]:.2e} p99 {q[1]:.2e} '
          f'p99.9 {q[2]:.2e} max {float(worst.max()):.2e}; fallback {res["fallback_fraction"]:.2e}')
    assert float(worst.max()) <= RHOD_MAX
    assert q[2] <= RHOD_P999
    assert not res['whole_call_fallback']
    assert res['fallback_fraction'] < 1e-3
    assert res['audited'] == (n + 99) // 100


def test_shipped_model_gain_over_physics_baseline(shipped):
    """What the trained network adds: rho*D from the baseline alone (residual zero) against
    the full model, both judged by the plain kernel on a sample of the same states."""
    kx, sur, _, (_, _, r, _), (_, _, pr, pd), ds, dp = shipped
    N, n, m = kx.nSpecies(), N_STATES, 1 << 16
    st = ds.view(N + 1, n)[:, :m].contiguous().reshape(-1)
    xin, base, flags = sur.features(m, m, m, dp[:m].contiguous(), st)
    lnD_base = base[:, 2:2 + N].double().T + sur.ostat[2:2 + N].double()[:, None]
    ref_D = pd.view(N, n)[:, :m]
    ok = (flags == 0)
    e_base = _rel(torch.exp(lnD_base), ref_D)[:, ok]
    e_net = _rel(r.view(N, n)[:, :m], pr.view(N, n)[:, :m])[:, ok]
    med_base, med_net = float(e_base.nanmedian()), float(e_net.nanmedian())
    print(f'\nrho*D median error: physics baseline {med_base:.2e}, with the network {med_net:.2e}')
    assert med_net * 5 <= med_base
