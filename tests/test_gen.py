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
              'kx_sur_outputs', 'kx_sur_infer', 'kx_sur_queue_init', 'kx_sur_gather',
              'kx_sur_merge'):
        assert f'\n{k}(' in src, k
    counts = json.load(open(os.path.join(MDIR, 'counts.json')))
    assert counts['kx_bk1_fp64']['bytes'] == 16 * 54
    assert counts['kx_bk2_fp64']['bytes'] == 8 * 54 + 8 * 55
    assert counts['kx_bk1_fp64']['transcendentals'] > 300


def test_surrogate_baselines_against_oracle(gri, oracle):
    """mu baseline (rank-8 Wilke) and lambda baseline (the reference's rule) against the
    oracle's FP64 transport."""
    from kinetix_amd import states
    info, t = gri
    s, p = states.sample_transport_set(512, info['species'], t['W'], seed=3)
    lmu, llam, _ = emit_surrogate.baseline_numpy(t, info['species'], s[0], p, s[1:])
    cond, visc, _, _ = oracle.transport(s, 1.0)
    assert np.max(np.abs(np.exp(lmu) / visc - 1)) < 1e-6
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
