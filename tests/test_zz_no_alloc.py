"""kx_sur_transport_noalloc (the surrogate transport call that allocates nothing: its workspace
is the library's per-GPU device scratch array) against kx_sur_transport with a caller workspace,
through the C ABI: bit-identical outputs and statistics on two calls in a row, the second with
fewer states (the scratch keeps the first call's contents), in both the statistics and the
enqueue-only form; a request beyond the scratch size is refused before any launch."""
import ctypes as C

import numpy as np
import pytest
import torch

from kinetix_amd import states
from kinetix_amd.api import Kinetix, lib
from kinetix_amd.surrogate import AUDIT_TOL, AUDIT_TOL_RHOD, SurStats, Surrogate, _p, _s, model_path

pytestmark = pytest.mark.gpu
dev = torch.device('cuda')
KX_ERR_TOO_LARGE = -7
SCRATCH_BYTES = 64 << 20


@pytest.fixture(scope='module')
def sur(gri):
    info, t = gri
    kx = Kinetix('gri30')
    kx.build(states.ONE_ATM, 1.0, states.air_mass_fractions(info['species'], t['W']))
    s = Surrogate(kx, model=model_path('gri30'))
    L = lib()
    vp, ci, cd = C.c_void_p, C.c_int, C.c_double
    L.kx_sur_transport_noalloc.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, ci, ci, cd, cd,
                                           cd, vp, vp, vp, vp, C.POINTER(SurStats), vp]
    yield s
    kx.close()


def _noalloc(sur, n, d_p, d_state, outs, stats=True, stride=100, offset=0):
    st = SurStats()
    rc = lib().kx_sur_transport_noalloc(
        sur.kx.h, n, n, n, _p(d_p), _p(d_state), _p(sur.norm), _p(sur.ostat), _p(sur.wbf16),
        _p(sur.params), stride, offset, AUDIT_TOL, AUDIT_TOL_RHOD, 0.01, *(_p(o) for o in outs),
        C.byref(st) if stats else None, _s())
    return rc, st


def _inputs(gri, n, seed):
    info, t = gri
    s, p = states.sample_transport_set(n, info['species'], t['W'], seed=seed)
    s[0, :3] = 3500.0                                          # out-of-hull states
    return (torch.from_numpy(np.ascontiguousarray(s)).to(dev).reshape(-1),
            torch.from_numpy(np.ascontiguousarray(p)).to(dev))


@pytest.mark.parametrize('stats', [True, False])
def test_noalloc_matches_workspace_call_twice(sur, gri, stats):
    N = sur.N
    for n, seed in ((300000, 71), (1000, 72)):             # second call smaller than the first
        d_state, d_p = _inputs(gri, n, seed)
        z = lambda m: torch.full((m,), np.nan, dtype=torch.float64, device=dev)
        a = [z(n), z(n), z(N * n), z(N * n)]
        b = [z(n), z(n), z(N * n), z(N * n)]
        ref = sur.infer(n, n, n, d_p, d_state, *a, stats=True)
        rc, st = _noalloc(sur, n, d_p, d_state, b, stats=stats)
        assert rc == 0, lib().kx_last_error()
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            assert torch.equal(x.nan_to_num(123.0), y.nan_to_num(123.0))
        assert bool(torch.isfinite(b[0]).all()) and bool(torch.isfinite(b[1]).all())
        if stats:
            assert (st.n_states, st.out_of_hull, st.audited, st.audit_failed, st.fallback) == \
                (ref['n'], ref['out_of_hull'], ref['audited'], ref['audit_failed'], ref['fallback'])
            assert st.out_of_hull >= 3


def test_noalloc_rejects_requests_beyond_the_scratch(sur):
    # 12 bytes per state at audit stride 1 (flags + audit id + queue slot): 6M states need
    # more than the 64 MiB scratch; refused before any launch or pointer use
    rc, _ = _noalloc(sur, 6_000_000, None, None, [None] * 4, stride=1)
    assert rc == KX_ERR_TOO_LARGE
    assert b'KX_SUR_SCRATCH_BYTES' in lib().kx_last_error()
