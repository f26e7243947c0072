"""The surrogate's hand-written training kernels against the FP64 CPU oracle (forward,
gradients, Adam), BF16-MFMA inference against the FP32 forward, and the surrogate's
fallback routes (hull check, audit, corrupted weights) against the plain FP64 kernel."""
import ctypes as C

import numpy as np
import pytest
import torch

from kinetix_amd import states
from kinetix_amd.api import Kinetix
from kinetix_amd.surrogate import IN, Surrogate, _bind, _p, hidden_width, init_params
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu
dev = torch.device('cuda')
# the rho*D audit tolerance for the smoke-sized model trained below (131,072 states, 4
# epochs, for speed): its rho*D errors are ~10x the shipped model's, whose own tolerance
# (kinetix_amd.surrogate.AUDIT_TOL_RHOD) the shipped-model tests use
SMOKE_TOL_RHOD = 0.2


# States the surrogate queues (audit samples, out-of-hull states) are recomputed by the plain
# FP64 pass with several lanes per state: its group sums over species (mu, lambda) add in
# another order than the one-lane-per-state plain kernel, so the two agree to rounding, not
# bit for bit.
QUEUE_RTOL = 1e-13


def same_as_plain(a, b):
    """a (queued states' outputs) equals b (the plain kernel's) to QUEUE_RTOL."""
    assert bool(torch.isfinite(a).all()) and bool(torch.isfinite(b).all())
    if a.numel() == 0:
        return True
    return float(((a - b).abs() / b.abs().clamp_min(1e-300)).max()) <= QUEUE_RTOL


def t32(a):
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32), device=dev)


@pytest.fixture(scope='module')
def L():
    return _bind()


@pytest.fixture(scope='module')
def case():
    rng = np.random.default_rng(3)
    p = init_params(1)
    p = p + rng.normal(0, 0.01, size=p.shape).astype(np.float32) * (p != 0)
    X = rng.normal(size=(128, IN)).astype(np.float32)
    X[:, 55:] = 0.0
    T = rng.normal(size=(128, IN)).astype(np.float32)
    return p, X, T


def test_forward_and_gradients_vs_oracle(L, case):
    p, X, T = case
    B = X.shape[0]
    dp, dX, dT = t32(p), t32(X), t32(T)
    H = hidden_width()
    acts = torch.empty(6 * B * H + B * IN, dtype=torch.float32, device=dev)
    grads = torch.empty_like(dp)
    work = torch.empty(2 * B * H, dtype=torch.float32, device=dev)
    part = torch.empty((B * IN + 255) // 256, dtype=torch.float32, device=dev)
    assert L.kx_mlp_forward(_p(dp), B, _p(dX), _p(acts), None) == 0
    assert L.kx_mlp_backward(_p(dp), B, _p(dX), _p(acts), _p(dT), _p(grads), _p(work), _p(part), None) == 0
    torch.cuda.synchronize()
    loss, Y, g = mo.loss_and_grad(p, X, T)
    Yg = acts[6 * B * H:].view(B, IN).cpu().numpy()
    assert np.max(np.abs(Yg - Y)) <= 1e-5 * np.max(np.abs(Y))
    gg = grads.cpu().numpy()
    assert np.max(np.abs(gg - g)) <= 1e-4 * np.max(np.abs(g))
    assert abs(float(part.sum()) / (B * 55) - loss) <= 1e-5 * loss


def test_adam_vs_oracle(L, case):
    p, _, _ = case
    rng = np.random.default_rng(4)
    g = rng.normal(size=p.shape).astype(np.float32)
    m = rng.normal(size=p.shape).astype(np.float32) * 0.1
    v = np.abs(rng.normal(size=p.shape)).astype(np.float32) * 0.01
    dp, dg, dm, dv = t32(p), t32(g), t32(m), t32(v)
    assert L.kx_mlp_adam(_p(dp), _p(dg), _p(dm), _p(dv), 1e-3, 0.9, 0.999, 1e-8, 7, None) == 0
    torch.cuda.synchronize()
    # the C ABI takes FP32 hyper-parameters: the oracle gets the same rounded values
    f32 = lambda x: float(np.float32(x))
    pr, mr, vr = mo.adam_step(*(a.astype(np.float64) for a in (p, g, m, v)), f32(1e-3), f32(0.9),
                              f32(0.999), f32(1e-8), 7)
    assert np.max(np.abs(dp.cpu().numpy() - pr)) <= 1e-6
    assert np.max(np.abs(dm.cpu().numpy() - mr)) <= 1e-6
    assert np.max(np.abs(dv.cpu().numpy() - vr)) <= 1e-7


@pytest.mark.parametrize('n', [128, 1000, 1])
def test_bf16_inference_vs_fp32(L, case, n):
    p, _, _ = case
    rng = np.random.default_rng(n)
    X = rng.normal(size=(n, IN)).astype(np.float32)
    X[:, 55:] = 0.0
    dp = t32(p)
    wb = torch.empty(L.kx_mlp_n_weights_bf16(), dtype=torch.int16, device=dev)
    assert L.kx_mlp_pack_bf16(_p(dp), _p(wb), None) == 0
    Y = torch.full((n, IN), np.nan, dtype=torch.float32, device=dev)
    assert L.kx_mlp_infer_bf16(_p(wb), _p(dp), n, _p(t32(X)), _p(Y), None) == 0
    torch.cuda.synchronize()
    ref, _ = mo.forward(p, X)
    err = np.abs(Y.cpu().numpy() - ref)
    # BF16 weights and activations (8-bit mantissa), FP32 accumulation
    assert np.all(np.isfinite(Y.cpu().numpy()))
    assert err.max() <= 0.05 * np.abs(ref).max()
    assert np.sqrt((err ** 2).mean()) <= 0.01 * np.sqrt((ref ** 2).mean())


@pytest.fixture(scope='module')
def trained(gri):
    """A small surrogate trained on the GPU by the library's own kernels (smoke-sized)."""
    info, t = gri
    kx = Kinetix('gri30')
    kx.build(states.ONE_ATM, 1.0, states.air_mass_fractions(info['species'], t['W']))
    sur = Surrogate(kx)
    meta = sur.train(n_train=1 << 17, epochs=4, batch=4096, holdout=8192, log=lambda s: None)
    yield kx, sur, meta
    kx.close()


def _run(sur, kx, s, p, **kw):
    N = kx.nSpecies()
    n = s.shape[1]
    ds = torch.from_numpy(np.ascontiguousarray(s)).to(dev).reshape(-1)
    dp = torch.from_numpy(np.ascontiguousarray(p)).to(dev)
    z = lambda m: torch.full((m,), np.nan, dtype=torch.float64, device=dev)
    v, c, r, d = z(n), z(n), z(N * n), z(N * n)
    kw.setdefault('audit_tol_rhoD', SMOKE_TOL_RHOD)
    info = sur.infer(n, n, n, dp, ds, v, c, r, d, **kw)
    pv, pc, pr, pd = sur.plain(n, n, n, dp, ds)
    torch.cuda.synchronize()
    return info, (v, c, r, d), (pv, pc, pr, pd)


def test_training_reduces_loss(trained):
    _, _, meta = trained
    h = meta['loss_history']
    assert h[-1] < 0.5 * h[0]


def test_surrogate_accuracy_and_gain(trained, gri):
    """mu and lambda within 1e-3 on every non-fallback state; the network must beat its own
    physics baseline for D_km (the gain of the learned part)."""
    info, t = gri
    kx, sur, _ = trained
    s, p = states.sample_transport_set(20000, info['species'], t['W'], seed=99)
    res, (v, c, r, d), (pv, pc, pr, pd) = _run(sur, kx, s, p)
    ok = torch.ones(s.shape[1], dtype=torch.bool, device=dev)
    ev = (v / pv - 1).abs()
    ec = (c / pc - 1).abs()
    assert float(ev.max()) <= 1e-3 and float(ec.max()) <= 1e-3
    N = kx.nSpecies()
    n = s.shape[1]
    ed_net = (d.view(N, n) / pd.view(N, n) - 1).abs()
    # baseline-only prediction: residual forced to zero
    xin, base, flags = sur.features(n, n, n, torch.from_numpy(p).to(dev),
                                    torch.from_numpy(s).to(dev).reshape(-1))
    lnD_base = base[:, 2:2 + N].double().T + sur.ostat[2:2 + N].double()[:, None]
    ed_base = (torch.exp(lnD_base) / pd.view(N, n) - 1).abs()
    assert float(ed_net.median()) < 0.5 * float(ed_base.median())
    assert res['fallback_fraction'] < 0.05


def test_out_of_hull_states_take_the_plain_result(trained, gri):
    info, t = gri
    kx, sur, _ = trained
    s, p = states.sample_transport_set(1000, info['species'], t['W'], seed=5)
    s[0, :10] = 3500.0                          # beyond the training temperatures
    res, got, ref = _run(sur, kx, s, p, audit_stride=0)
    assert res['out_of_hull'] >= 10
    for a, b in zip(got, ref):
        a_, b_ = a.view(-1, 1000)[:, :10], b.view(-1, 1000)[:, :10]
        assert same_as_plain(a_, b_)


def test_audit_failure_falls_back_whole_call(trained, gri):
    info, t = gri
    kx, sur, _ = trained
    s, p = states.sample_transport_set(3000, info['species'], t['W'], seed=6)
    res, got, ref = _run(sur, kx, s, p, audit_tol=0.0)       # every audited state "fails"
    assert res['whole_call_fallback'] and res['fallback_fraction'] == 1.0
    for a, b in zip(got, ref):
        assert torch.equal(a, b)


def test_rhoD_audit_failure_falls_back_whole_call(trained, gri):
    """A rho*D mismatch alone (mu and lambda tolerances wide open) fails the audit samples and
    routes the whole call to the plain kernel."""
    info, t = gri
    kx, sur, _ = trained
    s, p = states.sample_transport_set(3000, info['species'], t['W'], seed=6)
    res, got, ref = _run(sur, kx, s, p, audit_tol=1.0, audit_tol_rhoD=0.0)
    assert res['audit_failed'] == res['audited'] > 0
    assert res['whole_call_fallback'] and res['fallback_fraction'] == 1.0
    for a, b in zip(got, ref):
        assert torch.equal(a, b)
    res, _, _ = _run(sur, kx, s, p, audit_tol=1.0, audit_tol_rhoD=1.0)
    assert res['audit_failed'] == 0 and not res['whole_call_fallback']


def test_corrupted_weights_are_caught(trained, gri):
    info, t = gri
    kx, sur, _ = trained
    good = sur.params.clone()
    try:
        bad = good.cpu().numpy().copy()
        bad[:1000] = np.nan
        sur.set_model(bad, sur.norm.cpu().numpy(), sur.ostat.cpu().numpy(), sur.meta)
        s, p = states.sample_transport_set(3000, info['species'], t['W'], seed=7)
        res, got, ref = _run(sur, kx, s, p)
        assert res['whole_call_fallback']
        for a, b in zip(got, ref):
            assert torch.equal(a, b)
    finally:
        sur.set_model(good.cpu().numpy(), sur.norm.cpu().numpy(), sur.ostat.cpu().numpy(), sur.meta)


@pytest.mark.parametrize('n', [5000, 64, 1])
def test_fused_kernel_matches_composed_path(trained, gri, n):
    """kx_sur_infer (one kernel) against features -> kx_mlp_infer_bf16 -> outputs: the same
    model and BF16 arithmetic; they differ only in the final exp (FP32 vs FP64)."""
    info, t = gri
    kx, sur, _ = trained
    N = kx.nSpecies()
    s, p = states.sample_transport_set(n, info['species'], t['W'], seed=11 + n)
    if n > 1:
        s[0, 1] = 3500.0                                        # one out-of-hull state
    ds = torch.from_numpy(np.ascontiguousarray(s)).to(dev).reshape(-1)
    dp = torch.from_numpy(np.ascontiguousarray(p)).to(dev)
    z = lambda m: torch.full((m,), np.nan, dtype=torch.float64, device=dev)
    a = [z(n), z(n), z(N * n), z(N * n)]
    b = [z(n), z(n), z(N * n), z(N * n)]
    res = sur.infer(n, n, n, dp, ds, *a, audit_stride=0)
    fl = sur.infer_composed(n, n, n, dp, ds, *b)
    torch.cuda.synchronize()
    ok = (fl == 0)
    assert res['out_of_hull'] == int((~ok).sum())
    for x, y in zip(a, b):
        x_, y_ = x.view(-1, n)[:, ok], y.view(-1, n)[:, ok]
        assert torch.all(torch.isfinite(x_))
        assert float(((x_ - y_).abs() / y_.abs()).max()) <= 1e-5


def test_hull_flags_at_the_box_edges(trained, gri):
    """The out-of-hull flag against the training box evaluated on the host, on mass-fraction
    inputs (input k + 2 is Y_k rounded to FP32): Y_k exactly on either bound is inside, one
    FP32 ulp beyond either bound is outside, and NaN or +-inf flags the state. The feature kernel is checked state by state, the fused
    kernel by its out-of-hull count."""
    info, t = gri
    kx, sur, _ = trained
    n = 64
    s, p = states.sample_transport_set(n, info['species'], t['W'], seed=31)
    norm = sur.norm.cpu().numpy()
    lo, hi = norm[2 * IN:3 * IN], norm[3 * IN:4 * IN]

    def flags_of(st):
        ds = torch.from_numpy(np.ascontiguousarray(st)).to(dev).reshape(-1)
        dp = torch.from_numpy(np.ascontiguousarray(p)).to(dev)
        _, _, fl = sur.features(n, n, n, dp, ds)
        return fl.cpu().numpy(), ds, dp

    f0, _, _ = flags_of(s)
    inside = np.flatnonzero(f0 == 0)
    assert inside.size >= 8
    k = int(np.argmax(np.where(hi[2:2 + kx.nSpecies()] < 0.5, hi[2:2 + kx.nSpecies()], -1.0)))
    assert 0.0 < hi[k + 2] < 0.5 and lo[k + 2] < hi[k + 2]
    up = float(np.nextafter(np.float32(hi[k + 2]), np.float32(np.inf)))
    down = float(np.nextafter(np.float32(lo[k + 2]), np.float32(-np.inf)))
    edits = [(float(hi[k + 2]), 0), (up, 1), (float(lo[k + 2]), 0), (down, 1),
             (np.nan, 1), (np.inf, 1), (-np.inf, 1)]
    s2 = s.copy()
    want = f0.copy()
    for (v, w), j in zip(edits, inside):
        s2[1 + k, j] = v
        want[j] = w
    f2, ds, dp = flags_of(s2)
    assert np.array_equal(f2, want)
    z = lambda m: torch.full((m,), np.nan, dtype=torch.float64, device=dev)
    N = kx.nSpecies()
    res = sur.infer(n, n, n, dp, ds, z(n), z(n), z(N * n), z(N * n), audit_stride=0)
    torch.cuda.synchronize()
    assert res['out_of_hull'] == int(want.sum())


def test_out_of_hull_beyond_one_compact_pass(trained, gri):
    """Most states out of hull (4700 of 5000; the queue then holds far more than the audit):
    every one must equal the plain kernel's result."""
    info, t = gri
    kx, sur, _ = trained
    n = 5000
    s, p = states.sample_transport_set(n, info['species'], t['W'], seed=12)
    s[0, :4700] = 3500.0
    res, got, ref = _run(sur, kx, s, p)
    assert res['out_of_hull'] == 4700 and not res['whole_call_fallback']
    flagged = np.zeros(n, dtype=bool)
    flagged[:4700] = True
    for a, b in zip(got, ref):
        a_, b_ = a.view(-1, n)[:, flagged], b.view(-1, n)[:, flagged]
        assert same_as_plain(a_, b_)
    # audited states that were in hull also carry the plain result
    aud = np.zeros(n, dtype=bool)
    aud[::100] = True
    aud &= ~flagged
    for a, b in zip(got, ref):
        assert same_as_plain(a.view(-1, n)[:, aud], b.view(-1, n)[:, aud])


def test_enqueue_only_call_matches(trained, gri):
    """stats=False (kx_sur_transport with no statistics: no host wait, the capturable form)
    does the same device work: outputs bit-identical to the call that reads its statistics,
    with out-of-hull states and on a whole-call fallback."""
    info, t = gri
    kx, sur, _ = trained
    N, n = kx.nSpecies(), 3000
    s, p = states.sample_transport_set(n, info['species'], t['W'], seed=21)
    s[0, :5] = 3500.0
    ds = torch.from_numpy(np.ascontiguousarray(s)).to(dev).reshape(-1)
    dp = torch.from_numpy(np.ascontiguousarray(p)).to(dev)
    z = lambda m: torch.full((m,), np.nan, dtype=torch.float64, device=dev)
    for kw in ({}, {'audit_tol': 0.0}):
        kw = dict(kw)
        a = [z(n), z(n), z(N * n), z(N * n)]
        b = [z(n), z(n), z(N * n), z(N * n)]
        kw['audit_tol_rhoD'] = SMOKE_TOL_RHOD
        res = sur.infer(n, n, n, dp, ds, *a, **kw)
        assert sur.infer(n, n, n, dp, ds, *b, stats=False, **kw) is None
        torch.cuda.synchronize()
        assert res['out_of_hull'] >= 5
        assert res['whole_call_fallback'] == ('audit_tol' in kw)
        for x, y in zip(a, b):
            assert torch.equal(x, y)


def test_enqueue_only_call_replays_from_a_graph(trained, gri):
    """The enqueue-only call (no host wait, no allocation, no memset) captured once in a HIP
    graph and replayed on new inputs: every replay bit-identical to an eager call on the same
    inputs, with eager work and a sync between replays."""
    info, t = gri
    kx, sur, _ = trained
    N, n = kx.nSpecies(), 4096
    sets = [states.sample_transport_set(n, info['species'], t['W'], seed=30 + i) for i in range(3)]
    sets[1][0][0, :3] = 3500.0                                   # out-of-hull states in one set
    ds = torch.empty((N + 1) * n, dtype=torch.float64, device=dev)
    dp = torch.empty(n, dtype=torch.float64, device=dev)
    z = lambda m: torch.full((m,), np.nan, dtype=torch.float64, device=dev)
    got = [z(n), z(n), z(N * n), z(N * n)]
    ref = [z(n), z(n), z(N * n), z(N * n)]
    ds.copy_(torch.from_numpy(np.ascontiguousarray(sets[0][0])).reshape(-1))
    dp.copy_(torch.from_numpy(np.ascontiguousarray(sets[0][1])))
    sur.infer(n, n, n, dp, ds, *got, stats=False)                 # workspace allocated eagerly
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        sur.infer(n, n, n, dp, ds, *got, stats=False)
    torch.cuda.synchronize()
    for s, p in sets:
        ds.copy_(torch.from_numpy(np.ascontiguousarray(s)).reshape(-1))
        dp.copy_(torch.from_numpy(np.ascontiguousarray(p)))
        for x in got:
            x.fill_(np.nan)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        sur.infer(n, n, n, dp, ds, *ref, stats=False)
        torch.cuda.synchronize()
        for x, y in zip(got, ref):
            assert torch.equal(x, y)


@pytest.mark.parametrize('n,stride,offset', [(1001, 1, 0), (1001, 10, 7), (257, 300, 299), (1, 100, 0)])
def test_audit_sampling_edges(trained, gri, n, stride, offset):
    """The plain pass over the audit queue at its edges: every state audited (stride 1), an odd
    count with an offset, an offset past the last state, a single state. Audited states carry
    the plain kernel's result (to rounding); the count matches the sampling rule."""
    info, t = gri
    kx, sur, _ = trained
    s, p = states.sample_transport_set(n, info['species'], t['W'], seed=40 + n)
    res, got, ref = _run(sur, kx, s, p, audit_stride=stride, audit_offset=offset)
    aud = np.zeros(n, dtype=bool)
    aud[offset % stride::stride] = True
    assert res['audited'] == int(aud.sum())
    for a, b in zip(got, ref):
        a_, b_ = a.view(-1, n)[:, aud], b.view(-1, n)[:, aud]
        assert same_as_plain(a_, b_)
        assert torch.all(torch.isfinite(a.view(-1, n)))
