"""Generator and IR checks (CPU): GRI-3.0 structure as the reference describes it, the
emitted kernel set, the work counts, the surrogate baselines against the oracle, and a
numpy emulation of the MFMA fragment chain of kinetix_amd/csrc/kx_mlp_wave.h."""
import json
import os

import numpy as np
import pytest

from kinetix_amd.gen import emit_surrogate
from kinetix_amd.gen import mechanism as M

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MDIR = os.path.join(ROOT, 'kinetix_amd', 'mechs', 'gri30')


def test_gri30_structure(gri):
    info, t = gri
    assert int(t['n_species']) == 53 and int(t['n_active']) == 52
    assert int(t['n_reactions']) == 325
    assert info['species'][-1] == 'AR'                       # inert species last
    kind = np.asarray(t['kind'])
    assert int((kind == M.THREE_BODY).sum()) == 12
    assert int((kind == M.TROE).sum()) == 26
    assert int((kind == M.LINDEMANN).sum()) == 3
    assert int((kind == M.SRI).sum()) == 0 and int((kind == M.PLOG).sum()) == 0
    assert int(np.asarray(t['reversible']).sum()) == 309


def test_emitted_kernels_and_counts():
    src = open(os.path.join(MDIR, 'kernels.hip')).read()
    for k in ('kx_bk1_fp64', 'kx_bk1_mixed', 'kx_bk1_mixed_valu', 'kx_bk2_fp64',
              'kx_bk2_fp64_pstate', 'kx_thermo_fp64', 'kx_sur_features', 'kx_sur_targets',
              'kx_sur_outputs', 'kx_sur_infer', 'kx_sur_queue_init', 'kx_bk2_fp64_queue',
              'kx_bk2_fp64_mfma'):
        assert f'\n{k}(' in src, k
    counts = json.load(open(os.path.join(MDIR, 'counts.json')))
    assert counts['kx_bk1_fp64']['bytes'] == 16 * 54
    assert counts['kx_bk2_fp64']['bytes'] == 8 * 54 + 8 * 55
    assert counts['kx_bk1_fp64']['transcendentals'] > 300


def test_surrogate_baselines_against_oracle(gri, oracle):
    """mu baseline (rank-5 Wilke: 2.4e-6 truncation over 20,000 states) and lambda baseline
    (the reference's rule) against the oracle's FP64 transport."""
    from kinetix_amd import states
    info, t = gri
    s, p = states.sample_transport_set(512, info['species'], t['W'], seed=3)
    lmu, llam, _ = emit_surrogate.baseline_numpy(t, info['species'], s[0], p, s[1:])
    cond, visc, _, _ = oracle.transport(s, 1.0)
    assert np.max(np.abs(np.exp(lmu) / visc - 1)) < 1e-5
    assert np.max(np.abs(np.exp(llam) / cond - 1)) < 1e-12


# ---- emulation of v_mfma_f32_16x16x32_bf16 on the lane layouts the device test checks
def _mfma(a, b, c):
    """a, b: [64 lanes][8]; c: [64][4]. A[i][k] = a[i + 16 (k // 8)][k % 8],
    B[k][j] = b[j + 16 (k // 8)][k % 8], D[i][j] -> d[j + 16 (i // 4)][i % 4]."""
    A = np.zeros((16, 32)); B = np.zeros((32, 16))
    for k in range(32):
        for i in range(16):
            A[i, k] = a[i + 16 * (k // 8)][k % 8]
            B[k, i] = b[i + 16 * (k // 8)][k % 8]
    D = A @ B
    d = c.copy()
    for i in range(16):
        for j in range(16):
            d[j + 16 * (i // 4)][i % 4] += D[i, j]
    return d


def _kphi(kb, g, e):
    return 32 * kb + (4 * g + e if e < 4 else 16 + 4 * g + (e - 4))


def test_fragment_chain_keeps_activations_in_registers():
    """One hidden layer as kx_mlp_wave.h computes it: weights packed with kphi, the input
    read from LDS rows with kphi, the D fragments of output blocks 2kb and 2kb+1 reused
    as the next layer's B fragment kb. The result must be H W^T (before the activation)
    and its re-interpretation must index features by kphi."""
    rng = np.random.default_rng(0)
    K, OUT, S = 64, 64, 16                      # inputs, outputs, states (one block)
    W = rng.normal(size=(OUT, K)); H = rng.normal(size=(S, K))
    # pack weights: fragment (ob, kb), lane i + 16 g, element e = W[16 ob + i][kphi(kb, g, e)]
    frag = np.zeros((OUT // 16, K // 32, 64, 8))
    for ob in range(OUT // 16):
        for kb in range(K // 32):
            for lane in range(64):
                for e in range(8):
                    frag[ob, kb, lane, e] = W[16 * ob + lane % 16, _kphi(kb, lane // 16, e)]
    # B fragments of the input: lane j + 16 g, element e = H[j][kphi(kb, g, e)]
    bin_ = np.zeros((K // 32, 64, 8))
    for kb in range(K // 32):
        for lane in range(64):
            for e in range(8):
                bin_[kb, lane, e] = H[lane % 16, _kphi(kb, lane // 16, e)]
    Z = np.zeros((OUT // 16, 64, 4))
    for ob in range(OUT // 16):
        for kb in range(K // 32):
            Z[ob] = _mfma(frag[ob, kb], bin_[kb], Z[ob])
    ref = H @ W.T                                                  # [state][feature]
    for ob in range(OUT // 16):
        for lane in range(64):
            for e in range(4):
                assert Z[ob, lane, e] == pytest.approx(ref[lane % 16, 16 * ob + 4 * (lane // 16) + e])
    # next layer's B fragment kb: elements 0-3 from block 2kb, 4-7 from block 2kb + 1
    for kb in range(OUT // 32):
        for lane in range(64):
            nxt = np.concatenate([Z[2 * kb, lane], Z[2 * kb + 1, lane]])
            for e in range(8):
                assert nxt[e] == pytest.approx(ref[lane % 16, _kphi(kb, lane // 16, e)])


def _mfma_f32_16x16x4(a, b, c):
    """v_mfma_f32_16x16x4_f32 on the lane layout kx_selftest_mfma_layout checks on the device:
    a, b: [64 lanes] (one value each), c: [64][4]. A[i][k] = a[i + 16k], B[k][j] = b[j + 16k],
    D[i][j] -> c[j + 16 (i // 4)][i % 4]."""
    A = np.array([[a[i + 16 * k] for k in range(4)] for i in range(16)])
    B = np.array([[b[j + 16 * k] for j in range(16)] for k in range(4)])
    D = A @ B
    d = c.copy()
    for i in range(16):
        for j in range(16):
            d[j + 16 * (i // 4)][i % 4] += D[i, j]
    return d


def test_mixed_mfma_nu_fragments_accumulate_q_dot_nu(gri):
    """The q.nu accumulation of kx_bk1_mixed, emulated on the CPU exactly as the kernel
    addresses its operands: the A fragment of nu tile ti is read at flat float index
    (ti * 64 + lane) * 4 + s of the nu table (the same index into the LDS copy and into the
    global table), the B fragment from the q staging rows (state nb*16 + lane%16, reaction
    slot 4*(lane//16) + s), four K-steps per 16-reaction chunk, the block accumulators
    transposed to one row per state and summed. The result must be nu^T q for every species
    and state of a 64-state wave."""
    from kinetix_amd.gen import emit_mixed
    _, t = gri
    p = emit_mixed.plan(t)
    net = (t['nu_r'] - t['nu_f']).astype(np.float64)
    R, N = net.shape
    flat = p['table'].reshape(-1).astype(np.float64)
    tiles_by_chunk = {}
    for ti, (c, m) in enumerate(p['tiles']):
        tiles_by_chunk.setdefault(c, []).append((ti, m))
    rng = np.random.default_rng(0)
    q = rng.uniform(-1, 1, size=(R, 64)) * 2.0 ** rng.integers(-20, 20, size=(R, 64))
    wd = np.zeros((64, 16 * p['nmb']))
    for blk in p['blocks']:
        acc = np.zeros((p['nmb'], 4, 64, 4))
        for c in blk:
            ch = p['chunks'][c]
            qs = np.zeros((64, emit_mixed.CHUNK))               # one row per state
            qs[:, :len(ch)] = q[ch].T
            for nb in range(4):
                for s in range(4):
                    b = np.array([qs[nb * 16 + (lane & 15), 4 * (lane >> 4) + s] for lane in range(64)])
                    for ti, m in tiles_by_chunk.get(c, []):
                        a = np.array([flat[(ti * 64 + lane) * 4 + s] for lane in range(64)])
                        acc[m, nb] = _mfma_f32_16x16x4(a, b, acc[m, nb])
        for m in range(p['nmb']):
            for nb in range(4):
                for lane in range(64):
                    wd[nb * 16 + (lane & 15), m * 16 + 4 * (lane >> 4):m * 16 + 4 * (lane >> 4) + 4] += acc[m, nb, lane]
    ref = net.T @ q                                              # [species][state]
    for k in range(N):
        if k in p['pos']:
            np.testing.assert_allclose(wd[:, p['pos'][k]], ref[k], rtol=1e-12, atol=1e-12 * np.abs(q).max())
        else:
            assert np.all(ref[k] == 0)                           # inert: never a row


def _mfma_f64_16x16x4(a, b, c):
    """v_mfma_f64_16x16x4_f64 on the lane layout kx_selftest_mfma_f64_layout checks on the
    device: A[i][k] = a[i + 16k], B[k][j] = b[j + 16k], D[i][j] -> c[j + 16 (i % 4)][i // 4]."""
    A = np.array([[a[i + 16 * k] for k in range(4)] for i in range(16)])
    B = np.array([[b[j + 16 * k] for j in range(16)] for k in range(4)])
    D = A @ B
    d = c.copy()
    for i in range(16):
        for j in range(16):
            d[j + 16 * (i % 4)][i // 4] += D[i, j]
    return d


def test_wilke_mfma_fragments_reproduce_the_viscosity(gri, oracle):
    """kx_bk2_fp64_mfma's Wilke sum, emulated on the CPU exactly as the kernel lays out its
    operands: A fragments from emit_bk2_mfma.a_fragments (lane l of (ks, m, g) at flat index
    ((ks * M + m) * 3 + g) * 64 + l), B operands (x_j, x_j/eta_j, x_j/eta_j^2 of species
    4 ks + l // 16, state 16 sb + l % 16), the epilogue over the result layout (element r of
    lane l: species 16 m + 4 r + l // 16) and the four-lane-group sum. The viscosity of 64 seeded states must match the oracle's."""
    from kinetix_amd import states
    from kinetix_amd.gen import emit_bk2_mfma as BM
    info, t = gri
    N = int(t['n_species'])
    M, KS = BM.tiles(N)
    flat = BM.a_fragments(t).reshape(-1)
    sp = BM.species_table(t)
    s, _ = states.sample(64, info['species'], t['W'], T_range=(300.0, 2500.0), seed=5,
                         p_range_atm=(0.5, 40.0))
    T = s[0]
    lt = np.log(T)
    P = np.stack([lt ** i for i in range(5)])                   # (5, 64)
    x = np.maximum(s[1:], 0.0) * sp[:, 5:6]
    x = x / x.sum(0)                                             # mole fractions (N, 64)
    eta = sp[:, :5] @ P                                          # (N, 64)
    vis = np.zeros(64)
    for sb in range(4):
        acc = np.zeros((M, 3, 64, 4))
        for ks in range(KS):
            u = np.zeros(64); v = np.zeros(64); w = np.zeros(64)
            for lane in range(64):
                j, st = 4 * ks + lane // 16, 16 * sb + lane % 16
                if j < N:
                    u[lane] = x[j, st]
                    v[lane] = u[lane] / eta[j, st]
                    w[lane] = v[lane] / eta[j, st]
            for m in range(M):
                for g, b in enumerate((u, v, w)):
                    off = ((ks * M + m) * 3 + g) * 64
                    acc[m, g] = _mfma_f64_16x16x4(flat[off:off + 64], b, acc[m, g])
        part = np.zeros(64)
        for lane in range(64):
            st = 16 * sb + lane % 16
            for m in range(M):
                for r in range(4):
                    k = 16 * m + 4 * r + lane // 16
                    if k < N:
                        e = eta[k, st]
                        Sk = acc[m, 0, lane, r] + e * (acc[m, 1, lane, r] + e * acc[m, 2, lane, r])
                        part[lane] += x[k, st] * e * e / Sk
        tot = part[:16] + part[16:32] + part[32:48] + part[48:]
        vis[16 * sb:16 * sb + 16] = tot
    mu = np.sqrt(T) * vis
    _, v0, _, _ = oracle.transport(s, 1.0)
    assert np.max(np.abs(mu / v0 - 1)) < 1e-13


@pytest.mark.parametrize('mech', ['synth160', 'synth40cov_rolled', 'gri30_rolled'])
def test_cooperative_descriptors_round_trip(mech):
    """The packed 64-byte reaction descriptors of the cooperative loop-form kernels
    (emit_coop.packed) decode, field by field, to the loop-form tables they are built from:
    reactant/product molecules (padding = species N, the slot with w = 1 and g = 0), net
    species and signed coefficients, kind / reversibility / counts / third-body kind and
    index, falloff rows, and every reaction exactly once."""
    from kinetix_amd.gen import emit_coop, tables
    _, t = tables.load(os.path.join(ROOT, 'kinetix_amd', 'mechs', mech))
    N, R = int(t['n_species']), int(t['n_reactions'])
    P = emit_coop.packed(t)
    d = P['data']
    assert sorted(P['order']) == list(range(R))
    byte = lambda w, i: (int(w) >> (8 * i)) & 255
    nfo = 0
    for i, r in enumerate(P['order']):
        w = P['rxi'][i]
        kind, rev, nf, nr, nnet, dnu, mk, midx = (int(x) for x in d['meta'][r])
        m = int(w[2])
        assert (m & 15, (m >> 4) & 1, (m >> 5) & 7, (m >> 8) & 7, (m >> 11) & 15,
                ((m >> 15) & 15) - 8, (m >> 19) & 3) == (kind, rev, nf, nr, nnet, dnu, mk)
        fl = [byte(w[0], j) for j in range(4)]
        rl = [byte(w[1], j) for j in range(4)]
        assert fl[:nf] == [int(k) for k in d['fsp'][r][:nf]] and all(k == N for k in fl[nf:P['mx']])
        assert rl[:nr] == [int(k) for k in d['rsp'][r][:nr]] and all(k == N for k in rl[nr:P['mx']])
        sp = [byte(w[4 + j // 4], j % 4) for j in range(nnet)]
        nu = [byte(w[6 + j // 4], j % 4) - 256 * (byte(w[6 + j // 4], j % 4) > 127) for j in range(nnet)]
        assert sp == [int(k) for k in d['nsp'][r][:nnet]]
        assert nu == [int(round(float(v))) for v in d['nnu'][r][:nnet]]
        assert int(w[3]) == (r if kind == M.PLOG else midx)
        np.testing.assert_array_equal(P['rxd'][i], d['dbl'][r, :3])
        if kind in (M.LINDEMANN, M.TROE, M.SRI):
            fs = m >> 21
            assert fs == nfo
            np.testing.assert_array_equal(P['rxfo'][10 * fs + 3:10 * fs + 13], d['dbl'][r, 3:13])
            np.testing.assert_array_equal(P['rxff'][3 * fs + 3:3 * fs + 6], d['flt'][r, 3:6])
            nfo += 1
    # sorted by kind: the lanes of one iteration mostly share a branch
    kinds = [int(d['meta'][r, 0]) for r in P['order']]
    assert kinds == sorted(kinds)


def test_liveness_order(gri):
    """The BK1 reaction order is a permutation, deterministic, and lowers the modelled register
    peak below both the generator's grouped order and the all-live layout (w, g and wd of
    every species: 6N registers); the model itself is checked by hand on a two-reaction
    mechanism."""
    from kinetix_amd.gen import emit_hip, liveness
    info, t = gri
    R, N = int(t['n_reactions']), int(t['n_species'])
    order = liveness.order_reactions(t)
    assert sorted(order) == list(range(R))
    assert order == liveness.order_reactions(t)
    masks = liveness.use_masks(t)
    peak, _ = liveness.cost(masks, order)
    grouped, _ = liveness.cost(masks, emit_hip._RateEmitter(t, info).reaction_order())
    assert peak < grouped and peak < 6 * N
    # hand check: A + B <=> C (reversible), then C => D, over 4 species
    tiny = {'nu_f': np.array([[1, 1, 0, 0], [0, 0, 1, 0]]),
            'nu_r': np.array([[0, 0, 1, 0], [0, 0, 0, 1]]),
            'reversible': np.array([1, 0])}
    m = liveness.use_masks(tiny)
    # before reaction 0: no c (A, B, C have their first w use here; D never needs w, so
    # its C is dead after the third-body sums) + w A,B,C (6) + g A,B,C (6: g_C is computed
    # with w_C and read by reaction 1) + wd A,B,C (6) = 18; before reaction 1: w C (2)
    # + g C (2) + wd C,D (4) = 8
    assert list(liveness.profile(m, [0, 1])) == [18.0, 8.0]


def test_bk1_live_kernel_structure(gri):
    """Each reaction appears once in each w-form kernel, every species row is stored exactly
    once, and every w_k / g_k is defined before its first use."""
    import re
    info, t = gri
    src = open(os.path.join(MDIR, 'kernels.hip')).read()
    N, R = int(t['n_species']), int(t['n_reactions'])
    for kname in ('kx_bk1_fp64', 'kx_bk1_mixed_valu'):
        body = src.split(f'\n{kname}(')[1].split('\n}\n')[0]
        assert sorted(int(m) for m in re.findall(r'\bdouble q(\d+);', body)) == list(range(R))
        stores = re.findall(r'rates\[id \+ offsetT \+ (\d+) \* offset\]', body)
        assert sorted(int(m) for m in stores) == list(range(N))
        for var in ('w', 'g'):
            for k in range(N):
                uses = [m.start() for m in re.finditer(rf'\b{var}{k}\b', body)]
                if uses:
                    d = body.find(f'const double {var}{k} =')
                    assert d >= 0 and d + len('const double ') == min(uses), (kname, var, k)


def _queue_kernel_numpy(t, s, p):
    """The cooperative queue kernel's arithmetic (kinetix_amd/gen/emit_bk2_queue.py) in numpy,
    from the tables it stages in LDS: the triangular pair index, the C1/C2 row orientation and
    the species-table columns are those the kernel reads."""
    from kinetix_amd.gen import emit_bk2_queue as BQ
    tri, C1, C2, sp = BQ.tables(t)
    N = int(t['n_species'])
    T = s[0]
    lnT = np.log(T)
    P = np.stack([np.ones_like(lnT), lnT, lnT ** 2, lnT ** 3, lnT ** 4])   # [5][n]
    x = np.maximum(0.0, s[1:]) * sp[:, 10:11]
    Mbar = 1.0 / x.sum(0)
    x = x * Mbar
    e = sp[:, 0:5] @ P                                      # viscosity fit per species
    lam = sp[:, 5:10] @ P
    rho = p / float(t['R']) / T * Mbar
    rhoD, Dkm = np.zeros((N, s.shape[1])), np.zeros((N, s.shape[1]))
    mu_part = np.zeros(s.shape[1])
    for k in range(N):
        sk = np.zeros(s.shape[1])
        for j in range(N):
            if j == k:
                continue
            pi = k * (k - 1) // 2 + j if k > j else j * (j - 1) // 2 + k
            fit = tri[pi] @ P
            sk += x[j] * (fit if int(t.get('rcp_diff', 0)) else 1.0 / fit)
        Dkm[k] = T * np.sqrt(T) * (Mbar - sp[k, 11] * x[k]) / (p * Mbar * sk)
        rhoD[k] = rho * Dkm[k]
        f = C1[k][:, None] + C2[k][:, None] * e[k] / e
        mu_part += x[k] * e[k] ** 2 / (x * f * f).sum(0)
    mu = np.sqrt(T) * mu_part
    cond = np.sqrt(T) * 0.5 * ((x * lam).sum(0) + 1.0 / (x / lam).sum(0))
    return cond, mu, rhoD, Dkm


def test_cooperative_queue_kernel_arithmetic_vs_oracle(gri, oracle):
    """The queue kernel's table layout and formulas against the oracle's FP64 transport with
    per-state pressure; a transposed Wilke table or a wrong pair index fails by far more."""
    from kinetix_amd import states
    from kinetix_amd.gen import emit_bk2_queue as BQ
    info, t = gri
    assert BQ.lds_bytes(t) <= 160 * 1024
    src = open(os.path.join(MDIR, 'kernels.hip')).read()
    assert f'kx_bk2q_lanes = {BQ.LPS};' in src
    s, p = states.sample_transport_set(64, info['species'], t['W'], seed=21)
    got = _queue_kernel_numpy(t, s, p)
    ref = oracle.transport(s, 1.0, 1.0, p_state=p)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g / r - 1)) < 1e-12
    # the check has teeth: C1 transposed moves mu by far more than the tolerance
    t2 = dict(t)
    t2['C1'] = np.asarray(t['C1']).T.copy()
    assert np.max(np.abs(_queue_kernel_numpy(t2, s, p)[1] / ref[1] - 1)) > 1e-6


def _tile_kernel_numpy(t, s, p, wil=None):
    """The symmetric tile transport kernel (kinetix_amd/gen/emit_bk2_tile.py) in numpy, from the
    tables it reads: the upper-triangle pair fits at base(k) + j, and the [k][j] (C1, C2) records."""
    from kinetix_amd.gen import emit_bk2_tile as BT
    N = int(t['n_species'])
    fit = BT.upper_fits(t)
    wil = np.stack([np.asarray(t['C1']), np.asarray(t['C2'])], axis=-1) if wil is None else wil
    T = s[0]
    lnT = np.log(T)
    P = np.stack([np.ones_like(lnT), lnT, lnT ** 2, lnT ** 3, lnT ** 4])
    x = np.maximum(0.0, s[1:]) * np.asarray(t['rcpW'])[:, None]
    Mbar = 1.0 / x.sum(0)
    x = x * Mbar
    sums = np.zeros_like(x)
    for k in range(N):
        base = k * (N - 1) - k * (k - 1) // 2 - k - 1
        for j in range(k + 1, N):
            rD = 1.0 / (fit[base + j] @ P)
            sums[k] += x[j] * rD
            sums[j] += x[k] * rD
    Dkm = T * np.sqrt(T) * (Mbar - np.asarray(t['W'])[:, None] * x) / (p * Mbar * sums)
    rhoD = p / float(t['R']) / T * Mbar * Dkm
    e = np.asarray(t['visc']) @ P
    Sw = np.zeros_like(x)
    for k in range(N):
        f = wil[k, :, 0][:, None] + wil[k, :, 1][:, None] * e[k] / e
        Sw[k] = (x * f * f).sum(0)
    mu = np.sqrt(T) * (x * e * e / Sw).sum(0)
    lam_k = np.asarray(t['cond']) @ P
    cond = np.sqrt(T) * 0.5 * ((x * lam_k).sum(0) + 1.0 / (x / lam_k).sum(0))
    return cond, mu, rhoD, Dkm


def test_tile_transport_tables_vs_oracle(gri, oracle):
    """The tile kernel's pair-index arithmetic and Wilke record orientation against the oracle's
    FP64 transport (per-state pressure); a transposed Wilke record fails by far more."""
    from kinetix_amd import states
    from kinetix_amd.gen import emit_bk2_tile as BT
    info, t = gri
    s, p = states.sample_transport_set(48, info['species'], t['W'], seed=22)
    got = _tile_kernel_numpy(t, s, p)
    ref = oracle.transport(s, 1.0, 1.0, p_state=p)
    for g, r in zip(got, ref):
        assert np.max(np.abs(g / r - 1)) < 1e-12
    wil_t = np.stack([np.asarray(t['C1']).T, np.asarray(t['C2']).T], axis=-1)
    assert np.max(np.abs(_tile_kernel_numpy(t, s, p, wil_t)[1] / ref[1] - 1)) > 1e-6
    # the index formula enumerates the upper triangle row by row, at 160 species too
    for N in (53, 160):
        seen = [k * (N - 1) - k * (k - 1) // 2 - k - 1 + j for k in range(N) for j in range(k + 1, N)]
        assert seen == list(range(N * (N - 1) // 2))
    assert BT.lds_bytes({'n_species': 160}) == 160 * 1024


def test_thermo_table_vs_oracle(gri, oracle):
    """The thermodynamic kernel's coefficient table (both NASA-7 c_p pieces of every species in
    rows of 8 doubles, high piece first) read back from the generated source and evaluated with
    the kernel's index arithmetic -- base offset 0 when T > T_mid, else 8 N; species k at
    8 k -- against the oracle's c_p, on both sides of and exactly at every T_mid (the low piece
    at equality, as in thermoCoeffs). The same table with its pieces swapped fails."""
    import re
    from kinetix_amd import states
    from kinetix_amd.gen import emit_hip as E
    info, t = gri
    N = int(t['n_species'])
    src = E._emit_thermo_fp64(t, info)
    m = re.search(r'kx_thermo_fp64_cp\[(\d+)\] = \{([^}]*)\}', src)
    tab = np.array([float(v) for v in m.group(2).split(',')])
    assert tab.size == int(m.group(1)) == 2 * N * 8
    assert f'? 0 : {N * 8})' in src
    tmid = np.asarray(t['T_mid'], dtype=np.float64)
    Ts = [300.0, 750.0, 2500.0, 3500.0]
    for tm in sorted(set(tmid.tolist())):
        Ts += [np.nextafter(tm, 0.0), tm, np.nextafter(tm, np.inf), tm - 50.0, tm + 50.0]
    T = np.array(Ts)
    s, _ = states.sample_any(T.size, info['species'], t['W'], seed=23)
    s[0] = T
    _, cp0, _ = oracle.thermo(s, 101325.0)
    R, rcpW = float(t['R']), np.asarray(t['rcpW'])

    def kernel_cp(table):
        out = np.empty((N, T.size))
        for i, Ti in enumerate(T):
            for k in range(N):
                base = 0 if Ti > tmid[k] else 8 * N
                a = table[base + 8 * k: base + 8 * k + 5]
                c = a[0] + a[1] * Ti + a[2] * (Ti * Ti) + a[3] * (Ti * Ti * Ti) + a[4] * (Ti * Ti * Ti * Ti)
                out[k, i] = c * (R * rcpW[k])
        return out

    assert np.max(np.abs(kernel_cp(tab) / cp0 - 1)) < 1e-12
    swapped = np.concatenate([tab[8 * N:], tab[:8 * N]])
    assert np.max(np.abs(kernel_cp(swapped) / cp0 - 1)) > 1e-4
