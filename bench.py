#!/usr/bin/env python3
"""Benchmark of the hot path: species production rates (BK1) + mixture-averaged transport
(BK2) over GRI-3.0 states on one MI355X.

One step = one BK1 launch + one BK2 launch over all N_STATES states, inputs resident in HBM
(the reference's timing loop, benchmark/src/bk.cpp:694-776, without H2D/D2H). Prints ONE
JSON line with the metric, the roofline of the dominant kernel, the CPU baseline
(the reference's own FP64 routines, oracle/_ref, SERIAL flags, timed on this host's cores on
a bounded sample), and the error of every output against the FP64 oracle.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--states N]
                    [--bk1 fp64|mixed] [--bk2 fp64|surrogate] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'Mstates/s for rates & transport (GRI-3.0, 4M states) at stated err vs FP64'
PEAK_FP64 = 78.6      # TFLOP/s, MI355X FP64 (vector = matrix)
PEAK_FP32 = 157.3     # TFLOP/s
PEAK_HBM = 8000.0     # GB/s
PEAK_BF16 = 2500.0    # TFLOP/s, dense BF16 MFMA
# FP32 transcendental rate (v_exp_f32, v_log_f32, ...: quarter rate, 16 of 64 lanes per cycle
# per SIMD): 256 CUs x 4 SIMDs x 16 lanes / 4 x 2.4 GHz. FP64 exp/log/pow have no hardware
# instruction (instruction sequences on the FP64 VALU); their rate is reported against the
# same figure for scale.
PEAK_TRANSC = 256 * 4 * 16 / 4 * 2.4e9 / 1e12   # T/s
CPU_SHARE = 16        # host threads one GPU's lease provides
def mlp_flops(H):
    """Useful MACs x 2 per state of the surrogate MLP 55 -> H -> H -> H -> 55."""
    return 2 * (55 * H + 2 * H * H + H * 55)


def SUR_BYTES(N):
    """HBM bytes per state of the fused surrogate kernel: T, Y_1..Y_N and p in; mu, lambda,
    rho*D_k and the hull flag out (FP64 except the flag)."""
    return 8 * (N + 1) + 8 + 8 * (N + 2) + 4
P_ATM = 101325.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--states', type=int, default=4194304)
    ap.add_argument('--mech', default='gri30')
    ap.add_argument('--bk1', choices=('fp64', 'mixed', 'mixed_mfma'), default='mixed',
                    help='headline BK1 kernel: fp64, mixed (FP32 rates, FP64 K_c/net/sum, '
                         'VALU accumulation) or mixed_mfma (q.nu on FP32 MFMA)')
    ap.add_argument('--bk2', choices=('fp64', 'surrogate'), default='surrogate',
                    help='headline BK2: fp64 (plain kernel) or surrogate (BF16-MFMA MLP with hull '
                         'check, 1%% audit and fallback to the plain kernel)')
    ap.add_argument('--surrogate-model', default=None,
                    help='trained model (.npz); default kinetix_amd/mechs/<mech>/surrogate.npz')
    ap.add_argument('--no-cpu', action='store_true', help='skip the CPU baseline')
    ap.add_argument('--cpu-seconds', type=float, default=10.0,
                    help='target wall time of the bounded CPU-baseline sample')
    ap.add_argument('--err-sample', type=int, default=4096)
    ap.add_argument('--profile-only', action='store_true',
                    help='for rocprofv3 --pmc: one device copy of a known size (HBM byte '
                         'calibration), then one launch of each kernel; no timing, no JSON')
    return ap.parse_args()


def affinity_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def host_cores():
    """Threads for the CPU baseline: the host's affinity set, at most the CPU share that one
    GPU's lease provides (the rest of the host belongs to other GPUs' jobs)."""
    return max(1, min(affinity_cores(), CPU_SHARE))


def cpu_baseline(state, p, cores, target_s, mech='gri30'):
    """Time the reference's own FP64 GRI routines (BK1 + BK2 per state) on this host; for
    other mechanisms (no reference build exists) the oracle port with the same flags."""
    from oracle import pyoracle as po
    kind = 'reference' if (mech == 'gri30' and po.Reference.available('fast')) else 'port'
    n_all = state.shape[1]
    if kind == 'reference':
        ref = po.Reference('fast')

        def run(s):
            ref.production_rates(s, p, threads=cores)
            ref.transport(s, 1.0, threads=cores)
    else:
        from kinetix_amd.gen import tables
        _, t = tables.load(os.path.join(ROOT, 'kinetix_amd', 'mechs', mech))
        orc = po.Oracle(t, fast=True)

        def run(s):
            orc.production_rates(s, p, threads=cores)
            orc.transport(s, 1.0, threads=cores)
    n = min(n_all, 4096 * cores)
    s = np.ascontiguousarray(state[:, :n])
    t0 = time.perf_counter(); run(s); dt = time.perf_counter() - t0
    # scale the sample to ~target_s of wall time on `cores` threads
    n = int(min(n_all, max(n, n * target_s / max(dt, 1e-6))))
    s = np.ascontiguousarray(state[:, :n])
    run(s[:, :min(n, 1024)])                     # warm
    t0 = time.perf_counter(); run(s); dt = time.perf_counter() - t0
    return {'value': round(n / dt / 1e6, 4), 'unit': 'Mstates/s', 'cores': cores, 'kind': kind,
            'per_core': round(n / dt / 1e6 / cores, 4), 'affinity_cores': affinity_cores(),
            'thread_cap': f'min(affinity, {CPU_SHARE}): the CPU share of one GPU',
            'sample': f'{n} of the {n_all} states, BK1 + BK2 FP64, '
                      f'{"reference GRI routines (SERIAL flags)" if kind == "reference" else "oracle port (-O3 -ffast-math)"}, '
                      f'{cores} threads, {dt:.1f} s'}


def main():
    a = parse()
    if a.gpus != 1:
        print(f'bench.py: this build measures one GPU only (asked for --gpus {a.gpus}); states '
              f'shard across GPUs with zero exchange, so run one process per GPU instead.',
              file=sys.stderr)
        return 2
    import torch
    from kinetix_amd import states
    from kinetix_amd.api import Kinetix
    from kinetix_amd.gen import counts as cnt
    from kinetix_amd.gen import tables

    mdir = os.path.join(ROOT, 'kinetix_amd', 'mechs', a.mech)
    info, t = tables.load(mdir)
    counts = cnt.load(mdir)
    kx = Kinetix(a.mech)
    sp, W = kx.speciesNames(), kx.molar_masses
    kx.build(P_ATM, 1.0, states.reference_mass_fractions(sp, W))
    N, n = kx.nSpecies(), a.states

    t0 = time.time()
    near_eq = None
    if 'CH4' in sp and 'O2' in sp and 'N2' in sp:
        # the mixed-precision workload: 10 % of the states within 1 % of chemical equilibrium
        st, near_eq = states.sample_near_equilibrium_mix(n, sp, W, t, T_range=(1000.0, 2500.0),
                                                         seed=states.BASE_SEED + 2)
    else:
        st, _ = states.sample_any(n, sp, W, T_range=(1000.0, 2500.0), seed=states.BASE_SEED + 2)
    gen_s = time.time() - t0
    dev = torch.device('cuda')
    d_state = torch.from_numpy(st).to(dev).reshape(-1)
    d_rates = torch.empty_like(d_state)
    d_visc = torch.empty(n, dtype=torch.float64, device=dev)
    d_cond = torch.empty_like(d_visc)
    d_rhoD = torch.empty(N * n, dtype=torch.float64, device=dev)

    BK1 = {'fp64': ('kx_bk1_fp64', False), 'mixed': ('kx_bk1_mixed_valu', 'valu'),
           'mixed_mfma': ('kx_bk1_mixed', 'mfma')}
    bk1_name, bk1_mode = BK1[a.bk1]
    bk2_name = 'kx_bk2_fp64' if a.bk2 == 'fp64' else 'kx_mlp_infer_bf16'
    d_p = torch.full((n,), P_ATM, dtype=torch.float64, device=dev)
    sur = None
    if a.bk2 == 'surrogate' and N + 2 > 64:
        print(f'bench.py: {a.mech} has {N} species; the surrogate covers at most 62, '
              f'timing the plain BK2', file=sys.stderr)
        a.bk2 = 'fp64'
        bk2_name = 'kx_bk2_fp64'
    if a.bk2 == 'surrogate':
        from kinetix_amd.surrogate import Surrogate, model_path
        path = a.surrogate_model or model_path(a.mech)
        if not os.path.exists(path):
            print(f'bench.py: no trained surrogate at {path}; train one with '
                  f'python -m kinetix_amd.surrogate or run --bk2 fp64', file=sys.stderr)
            return 2
        sur = Surrogate(kx, path)
    sur_width = sur_epochs = sur_ntrain = None
    if sur is not None:
        from kinetix_amd.surrogate import hidden_width
        sur_width = hidden_width()
        sur_epochs, sur_ntrain = sur.meta.get('epochs'), sur.meta.get('n_train')
    IN_FEATURES = N + 2
    sur_stats = []
    sur_calls = [0]
    from kinetix_amd.api import KinetixError

    def bk1(mode=None):
        kx.productionRates(n, n, n, 1.0, d_state, d_rates, mixed=bk1_mode if mode is None else mode)

    def bk2_plain():
        kx.mixtureAvgTransportProps(n, n, n, 1.0, d_state, d_visc, d_cond, d_rhoD)

    def bk2(stats=False):
        """One BK2 call. The surrogate call is enqueued without a host wait (kx_sur_transport
        with no statistics; the same device work); stats=True reads its counters."""
        if sur is None:
            bk2_plain()
        else:
            r = sur.infer(n, n, n, d_p, d_state, d_visc, d_cond, d_rhoD,
                          audit_offset=sur_calls[0], stats=stats)
            sur_calls[0] += 1
            if stats:
                sur_stats.append(r)

    if a.profile_only:
        # calibration: a plain copy moves exactly nbytes in and nbytes out
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev)      # 1 GiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        launched = []
        for vname, (kname, mode) in BK1.items():
            try:
                bk1(mode); torch.cuda.synchronize(); launched.append(kname)
            except KinetixError:
                pass
        bk2_plain(); torch.cuda.synchronize(); launched.append('kx_bk2_fp64')
        try:
            kx.mixtureAvgTransportProps(n, n, n, 1.0, d_state, d_visc, d_cond, d_rhoD, mfma=True)
            torch.cuda.synchronize(); launched.append('kx_bk2_fp64_mfma')
        except KinetixError:
            pass
        if sur is not None:
            bk2(); torch.cuda.synchronize(); launched.append('the surrogate')
        d_cp = torch.empty(N * n, dtype=torch.float64, device=dev)
        kx.thermodynamicProps(n, n, n, 1.0, d_state, d_visc, d_cp, d_cond)
        torch.cuda.synchronize(); launched.append('kx_thermo_fp64')
        print(f'profile-only: copy {src.numel() * 8} B, {", ".join(launched)} over {n} states')
        return 0

    for _ in range(a.warmup):
        bk1(); bk2()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(a.steps)]
    torch.cuda.synchronize()
    w0 = time.perf_counter()
    for e0, e1, e2 in ev:
        e0.record(); bk1(); e1.record(); bk2(); e2.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    t_bk1 = np.array([e0.elapsed_time(e1) for e0, e1, _ in ev])     # ms
    t_bk2 = np.array([e1.elapsed_time(e2) for _, e1, e2 in ev])
    ms_step = wall * 1e3 / a.steps

    # every BK1 variant, timed on the same inputs (outside the headline loop), and its error
    from oracle import pyoracle as po
    orc = po.Oracle(t)
    cores = host_cores()
    idx = np.random.default_rng(7).choice(n, min(a.err_sample, n), replace=False)
    ref, scale = orc.production_rates(st[:, idx], P_ATM, scale=True, threads=cores)
    variants = {}
    for vname, (kname, mode) in BK1.items():
        try:
            bk1(mode); torch.cuda.synchronize()
        except KinetixError as e:          # form not generated for this mechanism
            variants[vname] = {'kernel': kname, 'absent': str(e)}
            continue
        for _ in range(3):                  # back to full clocks after the CPU-only oracle phase
            bk1(mode)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            bk1(mode)
        e1.record(); torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        r_ = d_rates.reshape(N + 1, n)[:, idx].cpu().numpy()
        l2v = [float(np.linalg.norm(r_[1 + k] - ref[1 + k]) / np.linalg.norm(ref[1 + k]))
               for k in range(N) if np.linalg.norm(ref[1 + k]) > 0]
        variants[vname] = {'kernel': kname, 'ms': round(ms, 4), 'Mstates_s': round(n / ms / 1e3, 2),
                           'max_rel_l2_per_species_vs_fp64_oracle': max(l2v)}
    for v in variants.values():
        if 'ms' in v:
            v['speedup_vs_fp64'] = round(variants['fp64']['ms'] / v['ms'], 3)
    bk1(); torch.cuda.synchronize()               # leave the headline variant's output

    # BK2 forms on the same inputs; the surrogate's error is measured against the plain FP64
    # kernel over every state
    def timed(fn, reps=10):
        for _ in range(3):
            fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record(); torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    bk2_variants = {'fp64': {'kernel': 'kx_bk2_fp64', 'ms': round(timed(bk2_plain), 4)}}
    p_visc, p_cond, p_rhoD = d_visc.clone(), d_cond.clone(), d_rhoD.clone()

    def bk2_mfma():
        kx.mixtureAvgTransportProps(n, n, n, 1.0, d_state, d_visc, d_cond, d_rhoD, mfma=True)
    try:
        bk2_mfma(); torch.cuda.synchronize()
        ms_m = timed(bk2_mfma)
        bk2_variants['fp64_mfma'] = {
            'kernel': 'kx_bk2_fp64_mfma', 'ms': round(ms_m, 4),
            'viscosity_max_rel_vs_plain': float((d_visc / p_visc - 1).abs().max()),
            'conductivity_max_rel_vs_plain': float((d_cond / p_cond - 1).abs().max()),
            'rhoD_max_rel_vs_plain': float((d_rhoD / p_rhoD - 1).abs().nan_to_num(0.0).max())}
    except KinetixError as e:              # loop-form mechanisms have no MFMA form
        bk2_variants['fp64_mfma'] = {'kernel': 'kx_bk2_fp64_mfma', 'absent': str(e)}

    def bk2_pstate():
        kx.mixtureAvgTransportPropsPstate(n, n, n, d_p, d_state, d_visc, d_cond, d_rhoD)
    bk2_pstate(); torch.cuda.synchronize()
    ms_ps = timed(bk2_pstate)
    bk2_variants['fp64_pstate'] = {
        'kernel': 'kx_bk2_fp64_pstate', 'ms': round(ms_ps, 4),
        'note': 'per-state pressure (the transport workload form), here every state at 1 atm',
        'viscosity_max_rel_vs_plain': float((d_visc / p_visc - 1).abs().max()),
        'conductivity_max_rel_vs_plain': float((d_cond / p_cond - 1).abs().max()),
        'rhoD_max_rel_vs_plain': float((d_rhoD / p_rhoD - 1).abs().nan_to_num(0.0).max())}
    bk2_plain(); torch.cuda.synchronize()
    sur_err = None
    if sur is not None:
        ms_s = timed(bk2)
        bk2(stats=True)
        torch.cuda.synchronize()
        st_ = sur_stats[-1]
        ev_ = (d_visc / p_visc - 1).abs()
        ec_ = (d_cond / p_cond - 1).abs()
        er_ = (d_rhoD / p_rhoD - 1).abs()
        sur_err = {'viscosity_max_rel_vs_plain': float(ev_.max()),
                   'conductivity_max_rel_vs_plain': float(ec_.max()),
                   'rhoD_max_rel_vs_plain': float(er_.max()),
                   'rhoD_median_rel_vs_plain': float(er_.median()),
                   'rhoD_p99_rel_vs_plain': float(torch.quantile(er_[::97].float(), 0.99)),
                   'fallback_fraction': st_['fallback_fraction'],
                   'out_of_hull': st_['out_of_hull'], 'audited': st_['audited'],
                   'audit_failed': st_['audit_failed'],
                   'whole_call_fallback': st_['whole_call_fallback'],
                   'states_compared': n}
        bk2_variants['surrogate'] = dict({'kernel': 'kx_sur_transport (kx_sur_infer + 1% audit + fallback)',
                                          'ms': round(ms_s, 4)}, **sur_err)
    for v in bk2_variants.values():
        if 'ms' in v:
            v['Mstates_s'] = round(n / v['ms'] / 1e3, 2)
            v['speedup_vs_fp64'] = round(bk2_variants['fp64']['ms'] / v['ms'], 3)

    # errors of the headline outputs against the FP64 oracle on a sample of states
    rates = d_rates.reshape(N + 1, n)[:, idx].cpu().numpy()
    c0, v0, r0, _ = orc.transport(st[:, idx], 1.0, threads=cores)
    cond = d_cond[idx].cpu().numpy(); visc = d_visc[idx].cpu().numpy()
    rhoD = d_rhoD.reshape(N, n)[:, idx].cpu().numpy()

    def rel(x, y, s=None):
        den = np.abs(y) if s is None else np.maximum(np.abs(y), s)
        d = np.abs(x - y)
        return float(np.max(np.where(d == 0, 0, d / np.where(den == 0, 1e-300, den))))

    wref = ref[1:]
    l2 = [float(np.linalg.norm(rates[1 + k] - wref[k]) / np.linalg.norm(wref[k]))
          for k in range(N) if np.linalg.norm(wref[k]) > 0]
    if near_eq is not None:
        # the cancellation stress: errors on the near-equilibrium states alone, and how far
        # from equilibrium they are (|q_net| / q_fwd over reversible reactions)
        ne = np.nonzero(near_eq)[0][:a.err_sample]
        s_ne = np.ascontiguousarray(st[:, ne])
        ref_ne, sc_ne = orc.production_rates(s_ne, P_ATM, scale=True, threads=cores)
        q_ne, qf_ne = orc.reaction_rates(s_ne, P_ATM, forward=True)
        revm = np.asarray(t['reversible']).astype(bool)
        rq = np.abs(q_ne[revm]) / np.maximum(qf_ne[revm], 1e-300)
        rq = rq[qf_ne[revm] > 1e-200]
        out_ne = d_rates.reshape(N + 1, n)[:, torch.from_numpy(ne).to(dev)].cpu().numpy()
        l2_ne = [float(np.linalg.norm(out_ne[1 + k] - ref_ne[1 + k]) / np.linalg.norm(ref_ne[1 + k]))
                 for k in range(N) if np.linalg.norm(ref_ne[1 + k]) > 0]
        near_eq_report = {'fraction': float(near_eq.mean()), 'sample_states': int(len(ne)),
                          'qnet_over_qfwd_quantiles_10_50_90_99': [float(v) for v in np.quantile(rq, [0.1, 0.5, 0.9, 0.99])],
                          'bk1_max_rel_l2_per_species': max(l2_ne),
                          'bk1_max_rel_cond_aware': rel(out_ne, ref_ne, sc_ne)}
    else:
        near_eq_report = None
    err = {'bk1_max_rel_cond_aware': rel(rates, ref, scale),
           'bk1_max_rel_l2_per_species': max(l2),
           'bk2_conductivity_max_rel': rel(cond, c0), 'bk2_viscosity_max_rel': rel(visc, v0),
           'bk2_rhoD_max_rel': rel(rhoD, r0), 'sample_states': int(len(idx))}

    traffic_path = os.path.join(ROOT, 'profiles', 'traffic.json')
    traffic = {}
    if os.path.exists(traffic_path):
        with open(traffic_path) as fh:
            traffic = json.load(fh)

    def roof(name, ms):
        """Roofline record of one kernel from its event-timed launch and the generator's
        per-state counts; `traffic` is the HBM bytes per launch measured with rocprofv3
        (profiles/traffic.json) for this kernel at this state count, when one exists."""
        tr = traffic.get(f'{name}@{n}')
        if name == 'kx_sur_infer':
            from kinetix_amd.surrogate import hidden_width
            H = hidden_width()
            fl = mlp_flops(H) * n / (ms * 1e-3) / 1e12
            gbs = SUR_BYTES(N) * n / (ms * 1e-3) / 1e9
            return {'bound': 'mfma', 'pipe': 'BF16 MFMA (dense peak 2.5 PFLOP/s)',
                    'binding': f'VALU issue: the FP32 physics baselines and GELU beside an MLP of width {H}',
                    'achieved': round(fl, 3), 'peak': PEAK_BF16, 'unit': 'TFLOP/s',
                    'frac': round(fl / PEAK_BF16, 4), 'traffic': tr,
                    'flops_per_state': mlp_flops(H), 'bytes_per_state': SUR_BYTES(N),
                    'hbm_GBps': round(gbs, 1), 'hbm_frac': round(gbs / PEAK_HBM, 4),
                    'avg_launch_ms': round(ms, 4), 'states_per_launch': n}
        if name == 'kx_thermo_fp64':
            c = counts[name]
            gbs = c['bytes'] * n / (ms * 1e-3) / 1e9
            r = {'bound': 'hbm', 'pipe': 'HBM3E (8 TB/s)', 'achieved': round(gbs, 1),
                 'peak': PEAK_HBM, 'unit': 'GB/s', 'frac': round(gbs / PEAK_HBM, 4), 'traffic': tr,
                 'flops_per_state': c['flops'], 'bytes_per_state': c['bytes'],
                 'avg_launch_ms': round(ms, 4), 'states_per_launch': n}
            if tr:
                r['traffic_over_algorithmic'] = round(tr / (c['bytes'] * n), 3)
            return r
        c = counts[name if name in counts else ('kx_bk2_fp64' if 'bk2' in name else 'kx_bk1_fp64')]
        fl = c['flops'] * n / (ms * 1e-3) / 1e12
        gbs = c['bytes'] * n / (ms * 1e-3) / 1e9
        r = {'bound': 'valu', 'pipe': 'FP64 VALU (MI355X FP64 vector peak 78.6 TFLOP/s)'
             + (', Wilke sum on FP64 MFMA (same peak)' if name == 'kx_bk2_fp64_mfma' else ''),
             'achieved': round(fl, 3), 'peak': PEAK_FP64, 'unit': 'TFLOP/s',
             'frac': round(fl / PEAK_FP64, 4), 'traffic': tr,
             'flops_per_state': c['flops'], 'transcendentals_per_state': c['transcendentals'],
             'bytes_per_state': c['bytes'],
             'hbm_GBps': round(gbs, 1), 'hbm_frac': round(gbs / PEAK_HBM, 4),
             'avg_launch_ms': round(ms, 4), 'states_per_launch': n}
        if tr:
            r['traffic_over_algorithmic'] = round(tr / (c['bytes'] * n), 3)
            r['hbm_GBps_measured_traffic'] = round(tr / (ms * 1e-3) / 1e9, 1)
        if 'bk1' in name:
            ts = c['transcendentals'] * n / (ms * 1e-3) / 1e12
            r.update({'transcendentals_per_s_T': round(ts, 3), 'transcendental_peak_T': round(PEAK_TRANSC, 3),
                      'transcendental_frac': round(ts / PEAK_TRANSC, 4),
                      'transcendental_note': 'FP32 hardware rate as the scale; FP64 exp/log/pow are '
                                             'VALU instruction sequences' + (
                                                 ' (the mixed kernel runs its Arrhenius exponentials '
                                                 'on v_exp_f32)' if 'mixed' in name else '')})
        return r

    m1, m2 = float(t_bk1.mean()), float(t_bk2.mean())
    kern = {bk1_name: roof(bk1_name, m1)}
    for v in list(variants.values()) + list(bk2_variants.values()):
        if 'ms' in v and v['kernel'].startswith('kx_b') and v['kernel'] not in kern:
            kern[v['kernel']] = roof(v['kernel'], v['ms'])
    if sur is not None:
        # the fused surrogate kernel on its own, event-timed on the launch stream
        from kinetix_amd.surrogate import _p, _s
        fl_buf = torch.empty(n, dtype=torch.int32, device=dev)

        def sur_kernel():
            sur.L.kx_sur_infer(kx.h, n, n, n, _p(d_p), _p(d_state), _p(sur.norm), _p(sur.ostat),
                               _p(sur.wbf16), _p(sur.params), _p(fl_buf), _p(d_visc), _p(d_cond),
                               _p(d_rhoD), None, None, None, _s())
        m2k = timed(sur_kernel)
        kern['kx_sur_infer'] = roof('kx_sur_infer', m2k)
        bk2_kernel_ms = m2k
        bk2() ; torch.cuda.synchronize()              # restore the audited outputs
    else:
        kern[bk2_name] = roof(bk2_name, m2)
        bk2_kernel_ms = m2
    # the thermodynamic-properties kernel (thermoCoeffs.okl): rho, cp_k, rho*cp, HBM-bound
    d_rho = torch.empty(n, dtype=torch.float64, device=dev)
    d_rhocp = torch.empty_like(d_rho)
    d_cp = torch.empty(N * n, dtype=torch.float64, device=dev)

    def thermo():
        kx.thermodynamicProps(n, n, n, 1.0, d_state, d_rho, d_cp, d_rhocp)
    kern['kx_thermo_fp64'] = roof('kx_thermo_fp64', timed(thermo))
    del d_rho, d_rhocp, d_cp
    bk2_kernel = 'kx_sur_infer' if sur is not None else bk2_name
    dom = bk1_name if m1 >= bk2_kernel_ms else bk2_kernel
    roofline = dict(kern[dom]); roofline['kernel'] = dom

    out = {
        'metric': METRIC,
        'value': round(n / (ms_step * 1e-3) / 1e6, 3),
        'unit': 'Mstates/s',
        'n_gpus': 1, 'steps': a.steps, 'warmup': a.warmup,
        'ms_per_step': round(ms_step, 4),
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': ('f64' if a.bk1 == 'fp64' else 'f32+f64') + ('+bf16' if sur is not None else ''),
        'value_plain_fp64_path': round(n / (variants['fp64']['ms'] + bk2_variants['fp64']['ms']) / 1e3, 3),
        'bk1_variants': variants,
        'bk2_variants': bk2_variants,
        'data': 'synthetic',
        'config': {'workload': f'{a.mech} BK1 production rates + BK2 mixture-averaged transport, '
                               f'{n} states, T~U[1000,2500] K, 1 atm, Y=0.5 Dirichlet(0.5)+0.5 CH4/air'
                               + (', 10% of states within 1% of chemical equilibrium' if near_eq is not None
                                  else ''),
                   'mechanism': a.mech, 'n_states': n, 'bk1': a.bk1, 'bk2': a.bk2,
                   'n_species': N, 'n_reactions': kx.nReactions(),
                   **({'surrogate_model': f'MLP {IN_FEATURES} -> 3x{sur_width} GELU -> {IN_FEATURES}, BF16 weights, '
                                          f'FP32 accumulation, {sur_epochs} epochs on {sur_ntrain} plain-kernel states'}
                      if sur is not None else {})},
        'roofline': roofline,
        'kernels': kern,
        'bk1_Mstates_s': round(n / m1 / 1e3, 2), 'bk2_Mstates_s': round(n / m2 / 1e3, 2),
        'bk1_median_ms': round(float(np.median(t_bk1)), 4), 'bk2_median_ms': round(float(np.median(t_bk2)), 4),
        'bk1_GRXN_s': round(n * kx.nReactions() / m1 / 1e6, 2),
        'bk2_GDOF_s': round(n * (N + 2) / m2 / 1e6, 2),
        'err_vs_fp64_oracle': err,
        'near_equilibrium': near_eq_report,
        'state_gen_s': round(gen_s, 2),
    }
    if sur is not None:
        out['surrogate'] = sur_err
    if not a.no_cpu:
        out['cpu_baseline'] = cpu_baseline(st, P_ATM, cores, a.cpu_seconds, a.mech)
    print(json.dumps(out))
    return 0


if __name__ == '__main__':
    sys.exit(main())
