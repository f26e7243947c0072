"""Learned transport surrogate: training and audited inference (host side).

Model (per state): y = b(T, p, Y) + out_mean + out_std * MLP((x - in_mean) / in_std), with
y = (ln mu, ln lambda, ln D_1m .. ln D_Nm) and b the FP32 physics baseline of
kinetix_amd/gen/emit_surrogate.py. All arithmetic runs in this library's HIP kernels
(include/kinetix_mi_surrogate.h): features, targets, FP32 training GEMMs/Adam, BF16-MFMA
inference, outputs, and the plain FP64 transport kernel for labels, audits and fallback.
torch provides device memory, streams, random permutations and index gathers only.

Inference safeguards (the fallback route is the plain kernel, kx_mixture_avg_transport_pstate),
all inside one C-ABI call, kx_sur_transport:
  * hull check: any input outside the box spanned by the training inputs (small margin)
    -> plain kernel;
  * audit: every `audit_stride`-th state (1 % by default) is also computed by the plain
    kernel and keeps the plain result; one whose mu or lambda differs by more than
    `audit_tol`, or whose rho*D_k differs by more than `audit_tol_rhoD` for any k, counts as
    a failure; if more than `max_audit_fail` of the audited states fail, the whole call
    falls back to the plain kernel.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time

import numpy as np

from . import states as st
from .api import Kinetix, KinetixError, _check, _check_tensor, lib

IN = 64
VALID = 55
AUDIT_TOL = 1e-3          # mu, lambda: the accuracy bound the surrogate is built to
AUDIT_TOL_RHOD = 1e-2     # rho*D_k, max over k


class SurStats(C.Structure):
    _fields_ = [('n_states', C.c_int), ('out_of_hull', C.c_int), ('audited', C.c_int),
                ('audit_failed', C.c_int), ('whole_call_fallback', C.c_int), ('fallback', C.c_int)]


def _bind():
    L = lib()
    vp, ci, cf = C.c_void_p, C.c_int, C.c_float
    L.kx_sur_features.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp]
    L.kx_sur_targets.argtypes = [ci, ci, ci, vp, vp, vp, vp, vp, vp, vp]
    L.kx_sur_outputs.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.kx_sur_infer.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.kx_sur_workspace_bytes.argtypes = [ci, ci, ci, ci]
    L.kx_sur_workspace_bytes.restype = C.c_size_t
    L.kx_sur_transport.argtypes = [ci, ci, ci, ci, vp, vp, vp, vp, vp, vp, ci, ci, C.c_double,
                                   C.c_double, C.c_double, vp, vp, vp, vp, vp, C.c_size_t,
                                   C.POINTER(SurStats), vp]
    L.kx_mlp_width.restype = ci
    L.kx_mlp_forward.argtypes = [vp, ci, vp, vp, vp]
    L.kx_mlp_backward.argtypes = [vp, ci, vp, vp, vp, vp, vp, vp, vp]
    L.kx_mlp_adam.argtypes = [vp, vp, vp, vp, cf, cf, cf, cf, ci, vp]
    L.kx_mlp_pack_bf16.argtypes = [vp, vp, vp]
    L.kx_mlp_infer_bf16.argtypes = [vp, vp, ci, vp, vp, vp]
    return L


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _s():
    import torch
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def hidden_width():
    """The library's hidden width H (the MLP is 64 -> H -> H -> H -> 64)."""
    return int(lib().kx_mlp_width())


def init_params(seed=0):
    """Glorot-uniform weights, zero biases; padded input columns / output rows are zero."""
    rng = np.random.default_rng(seed)
    n = lib().kx_mlp_n_params()
    H = hidden_width()
    p = np.zeros(n, dtype=np.float32)
    o = 0
    for (rows, cols, vr, vc) in ((H, 64, H, VALID), (H, H, H, H), (H, H, H, H), (64, H, VALID, H)):
        lim = math.sqrt(6.0 / (vr + vc))
        W = np.zeros((rows, cols), dtype=np.float32)
        W[:vr, :vc] = rng.uniform(-lim, lim, size=(vr, vc))
        p[o:o + rows * cols] = W.ravel()
        o += rows * cols + rows              # weights then bias
    assert o == n
    return p


class Surrogate:
    """A trained surrogate bound to one loaded mechanism."""

    def __init__(self, kx: Kinetix, model=None):
        import torch
        self.kx = kx
        self.L = _bind()
        self.N = kx.nSpecies()
        if self.N + 2 > IN:
            raise KinetixError('surrogate supports at most 62 species')
        self.dev = torch.device('cuda')
        if model is not None:
            self.load(model)

    # ---------------------------------------------------------------- model state
    def set_model(self, params, norm, ostat, meta=None):
        import torch
        n = int(self.L.kx_mlp_n_params())
        if np.asarray(params).size != n or np.asarray(norm).size != 4 * IN or np.asarray(ostat).size != 2 * IN:
            raise ValueError(f'surrogate model shape: {np.asarray(params).size} parameters (this library: '
                             f'{n}, hidden width {hidden_width()}), {np.asarray(norm).size} input and '
                             f'{np.asarray(ostat).size} output statistics (expected {4 * IN}, {2 * IN})')
        self.params = torch.as_tensor(np.asarray(params, np.float32), device=self.dev).contiguous()
        self.norm = torch.as_tensor(np.asarray(norm, np.float32), device=self.dev).contiguous()
        self.ostat = torch.as_tensor(np.asarray(ostat, np.float32), device=self.dev).contiguous()
        self.wbf16 = torch.empty(self.L.kx_mlp_n_weights_bf16(), dtype=torch.int16, device=self.dev)
        _check(self.L.kx_mlp_pack_bf16(_p(self.params), _p(self.wbf16), _s()))
        self.meta = meta or {}

    def save(self, path):
        np.savez_compressed(path, params=self.params.cpu().numpy(), norm=self.norm.cpu().numpy(),
                            ostat=self.ostat.cpu().numpy(), meta=json.dumps(self.meta))

    def load(self, path):
        with np.load(path, allow_pickle=False) as z:
            self.set_model(z['params'], z['norm'], z['ostat'], json.loads(str(z['meta'])))

    # ---------------------------------------------------------------- kernels
    def features(self, n, offsetT, offset, d_p, d_state, norm=None):
        import torch
        xin = torch.empty((n, IN), dtype=torch.float32, device=self.dev)
        base = torch.empty((n, IN), dtype=torch.float32, device=self.dev)
        flags = torch.empty(n, dtype=torch.int32, device=self.dev)
        _check(self.L.kx_sur_features(self.kx.h, n, offsetT, offset, _p(d_p), _p(d_state),
                                      _p(self.norm if norm is None else norm), _p(xin), _p(base),
                                      _p(flags), _s()))
        return xin, base, flags

    def plain(self, n, offsetT, offset, d_p, d_state, Dkm=True):
        import torch
        z = lambda m: torch.empty(m, dtype=torch.float64, device=self.dev)
        visc, cond, rhoD = z(n), z(n), z(self.N * offset)
        dkm = z(self.N * offset) if Dkm else None
        self.kx.mixtureAvgTransportPropsPstate(n, offsetT, offset, d_p, d_state, visc, cond, rhoD, dkm)
        return visc, cond, rhoD, dkm

    def _tables(self):
        from .gen import tables
        return tables.load(self.kx.dir)[1]

    # ---------------------------------------------------------------- training
    def train(self, n_train=1 << 24, epochs=60, batch=8192, lr=2e-3, seed=st.BASE_SEED + 4,
              log=print, holdout=65536, hull_margin=0.02, near_eq_frac=0.1):
        import torch
        N, dev = self.N, self.dev
        sp, W = self.kx.speciesNames(), self.kx.molar_masses
        t0 = time.time()
        # training states (host-side seeded sampler, uploaded in chunks), labels from the
        # plain FP64 kernel on the GPU
        n = n_train + holdout
        d_state = torch.empty((N + 1) * n, dtype=torch.float64, device=dev)
        d_p = torch.empty(n, dtype=torch.float64, device=dev)
        view = d_state.view(N + 1, n)
        chunk = 1 << 21
        for c0 in range(0, n, chunk):
            m = min(chunk, n - c0)
            s_, p_ = st.sample_transport_set(m, sp, W, seed=seed + c0, tables=self._tables(),
                                             near_eq_frac=near_eq_frac)
            view[:, c0:c0 + m].copy_(torch.from_numpy(s_))
            d_p[c0:c0 + m].copy_(torch.from_numpy(p_))
        visc, cond, _, dkm = self.plain(n, n, n, d_p, d_state)
        torch.cuda.synchronize()
        log(f'surrogate: {n} states + plain-kernel labels in {time.time() - t0:.1f} s')
        # input statistics and hull from the raw features (identity normalisation)
        ident = torch.zeros(4 * IN, dtype=torch.float32, device=dev)
        ident[IN:2 * IN] = 1.0
        ident[2 * IN:3 * IN] = -float('inf')
        ident[3 * IN:] = float('inf')
        raw, base, _ = self.features(n, n, n, d_p, d_state, norm=ident)
        tr = raw[:n_train]
        mean, std = tr.mean(0), tr.std(0)
        lo, hi = tr.min(0).values, tr.max(0).values
        span = (hi - lo)
        norm = torch.cat([torch.where(std > 0, mean, torch.zeros_like(mean)),
                          torch.where(std > 1e-12, std, torch.ones_like(std)),
                          lo - hull_margin * span, hi + hull_margin * span]).contiguous()
        del raw
        xin, base, _ = self.features(n, n, n, d_p, d_state, norm=norm)
        # output statistics of the residuals, then standardised targets
        ost0 = torch.zeros(2 * IN, dtype=torch.float32, device=dev)
        ost0[IN:] = 1.0
        tgt = torch.empty((n, IN), dtype=torch.float32, device=dev)
        _check(self.L.kx_sur_targets(self.kx.h, n, n, _p(visc), _p(cond), _p(dkm), _p(base),
                                     _p(ost0), _p(tgt), _s()))
        om, osd = tgt[:n_train].mean(0), tgt[:n_train].std(0)
        # floor: a residual whose spread is FP32 rounding (mu, lambda: the baselines are
        # nearly exact) must not be blown up to unit variance -- that is noise to learn
        osd = torch.clamp(osd, min=1e-3)
        om[VALID:] = 0.0
        osd[VALID:] = 1.0
        ostat = torch.cat([om, osd]).contiguous()
        _check(self.L.kx_sur_targets(self.kx.h, n, n, _p(visc), _p(cond), _p(dkm), _p(base),
                                     _p(ostat), _p(tgt), _s()))
        torch.cuda.synchronize()
        log(f'surrogate: features/targets ready at {time.time() - t0:.1f} s')

        params = torch.as_tensor(init_params(seed), device=dev).contiguous()
        P = params.numel()
        grads = torch.zeros(P, dtype=torch.float32, device=dev)
        m1, m2 = torch.zeros_like(grads), torch.zeros_like(grads)
        H = hidden_width()
        acts = torch.empty(6 * batch * H + batch * IN, dtype=torch.float32, device=dev)
        work = torch.empty(2 * batch * H, dtype=torch.float32, device=dev)
        part = torch.empty((batch * IN + 255) // 256, dtype=torch.float32, device=dev)
        steps_per_epoch = n_train // batch
        total = epochs * steps_per_epoch
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        step = 0
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n_train, generator=gen, device=dev)
            ep_loss = torch.zeros((), dtype=torch.float64, device=dev)
            for b in range(steps_per_epoch):
                idx = perm[b * batch:(b + 1) * batch]
                Xb = xin.index_select(0, idx)
                Tb = tgt.index_select(0, idx)
                step += 1
                # one-cycle schedule: linear warm-up over 5 %, cosine decay to 1 % of lr
                f = step / total
                lr_t = lr * (f / 0.05 if f < 0.05 else 0.01 + 0.99 * 0.5 * (1 + math.cos(math.pi * (f - 0.05) / 0.95)))
                _check(self.L.kx_mlp_forward(_p(params), batch, _p(Xb), _p(acts), _s()))
                _check(self.L.kx_mlp_backward(_p(params), batch, _p(Xb), _p(acts), _p(Tb), _p(grads),
                                              _p(work), _p(part), _s()))
                _check(self.L.kx_mlp_adam(_p(params), _p(grads), _p(m1), _p(m2), lr_t, 0.9, 0.999,
                                          1e-8, step, _s()))
                ep_loss += part.sum()
            l_ = float(ep_loss) / (steps_per_epoch * batch * VALID)
            hist.append(l_)
            log(f'surrogate: epoch {ep + 1}/{epochs} mse {l_:.3e} ({time.time() - t0:.1f} s)')
        meta = {'hidden': hidden_width(), 'n_train': n_train, 'epochs': epochs, 'batch': batch, 'lr': lr, 'seed': seed,
                'near_eq_frac': near_eq_frac,
                'loss_history': hist, 'train_seconds': time.time() - t0,
                'mechanism': os.path.basename(self.kx.dir)}
        self.set_model(params.cpu().numpy(), norm.cpu().numpy(), ostat.cpu().numpy(), meta)
        return meta

    # ---------------------------------------------------------------- inference
    def infer(self, n, offsetT, offset, d_p, d_state, d_visc, d_cond, d_rhoD, d_Dkm=None,
              audit_stride=100, audit_tol=AUDIT_TOL, audit_tol_rhoD=AUDIT_TOL_RHOD,
              max_audit_fail=0.01, audit_offset=0, stats=True):
        """Audited surrogate transport for n states (FP64 outputs in the plain kernel's layout),
        one call of kx_sur_transport. Returns the fallback statistics; stats=False enqueues
        the call on the current stream without waiting and returns None (same device work).

        The workspace is kept per size and never freed while this object lives, so a graph
        that captured a call keeps a valid workspace for its replays."""
        import torch
        N = self.N
        if n:
            _check_tensor('state', d_state, offsetT + (N - 1) * offset + n)
            for nm, t_ in (('p', d_p), ('viscosity', d_visc), ('conductivity', d_cond)):
                _check_tensor(nm, t_, n)
            _check_tensor('rhoD', d_rhoD, (N - 1) * offset + n)
            if d_Dkm is not None:
                _check_tensor('Dkm', d_Dkm, (N - 1) * offset + n)
        need = int(self.L.kx_sur_workspace_bytes(self.kx.h, n, audit_stride, d_Dkm is not None))
        works = self.__dict__.setdefault('_works', [])
        work = next((w for w in works if w.numel() >= need), None)
        if work is None:
            work = torch.empty(max(need, 256), dtype=torch.uint8, device=self.dev)
            works.append(work)
        st = SurStats()
        _check(self.L.kx_sur_transport(self.kx.h, n, offsetT, offset, _p(d_p), _p(d_state),
                                       _p(self.norm), _p(self.ostat), _p(self.wbf16), _p(self.params),
                                       audit_stride, audit_offset, audit_tol, audit_tol_rhoD,
                                       max_audit_fail, _p(d_visc), _p(d_cond), _p(d_rhoD), _p(d_Dkm),
                                       _p(work), work.numel(), C.byref(st) if stats else None, _s()))
        if not stats:
            return None
        return {'n': st.n_states, 'out_of_hull': st.out_of_hull, 'audited': st.audited,
                'audit_failed': st.audit_failed, 'whole_call_fallback': bool(st.whole_call_fallback),
                'fallback': st.fallback, 'fallback_fraction': st.fallback / n if n else 0.0}

    def infer_composed(self, n, offsetT, offset, d_p, d_state, d_visc, d_cond, d_rhoD, d_Dkm=None):
        """The same model as three kernels (features -> kx_mlp_infer_bf16 -> outputs) with no
        audit or fallback: a cross-check of the fused kernel. Returns the flags."""
        import torch
        xin, base, flags = self.features(n, offsetT, offset, d_p, d_state)
        resid = torch.empty((n, IN), dtype=torch.float32, device=self.dev)
        _check(self.L.kx_mlp_infer_bf16(_p(self.wbf16), _p(self.params), n, _p(xin), _p(resid), _s()))
        _check(self.L.kx_sur_outputs(self.kx.h, n, offsetT, offset, _p(d_p), _p(d_state), _p(base),
                                     _p(resid), _p(self.ostat), _p(flags), _p(d_visc), _p(d_cond),
                                     _p(d_rhoD), _p(d_Dkm), _s()))
        return flags

def model_path(mech='gri30'):
    from .api import MECH_ROOT
    return os.path.join(MECH_ROOT, mech, 'surrogate.npz')


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description='train the transport surrogate on the GPU')
    ap.add_argument('--mech', default='gri30')
    ap.add_argument('--n-train', type=int, default=1 << 24)
    ap.add_argument('--epochs', type=int, default=60)
    ap.add_argument('--batch', type=int, default=8192)
    ap.add_argument('--lr', type=float, default=2e-3)
    ap.add_argument('--out', default=None)
    a = ap.parse_args(argv)
    kx = Kinetix(a.mech)
    kx.build(st.ONE_ATM, 1.0, st.air_mass_fractions(kx.speciesNames(), kx.molar_masses))
    sur = Surrogate(kx)
    meta = sur.train(n_train=a.n_train, epochs=a.epochs, batch=a.batch, lr=a.lr,
                     log=lambda s: print(s, flush=True))
    out = a.out or model_path(a.mech)
    sur.save(out)
    print(json.dumps({'saved': out, 'train_seconds': meta['train_seconds'],
                      'final_mse': meta['loss_history'][-1]}), flush=True)


if __name__ == '__main__':
    main()
