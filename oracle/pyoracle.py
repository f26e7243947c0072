"""TEST INFRASTRUCTURE ONLY: ctypes front-end of the CPU oracle (oracle/kx_oracle.c) and of
the reference's compiled GRI-3.0 routines (oracle/_ref).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only to check GPU results against or as the timed CPU baseline. The product path
(kinetix_amd) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class _Mech(C.Structure):
    _fields_ = [
        ('n_species', C.c_int), ('n_active', C.c_int), ('n_reactions', C.c_int),
        ('R', C.c_double), ('one_atm_R', C.c_double), ('R_one_atm', C.c_double),
        ('W', _dp), ('rcpW', _dp), ('g_lo', _dp), ('g_hi', _dp), ('h_lo', _dp), ('h_hi', _dp),
        ('cp_lo', _dp), ('cp_hi', _dp), ('T_mid', _dp),
        ('kind', _ip), ('reversible', _ip), ('collider', _ip), ('has_eff', _ip),
        ('lnA', _dp), ('A', _dp), ('beta', _dp), ('Ta', _dp),
        ('lnA0', _dp), ('A0', _dp), ('beta0', _dp), ('Ta0', _dp),
        ('nu_f', _ip), ('nu_r', _ip), ('eff', _dp), ('troe', _dp), ('sri', _dp),
        ('plog_off', _ip), ('plog_P', _dp), ('plog_lnP', _dp), ('plog_nk', _ip),
        ('plog_koff', _ip), ('plog_k', _dp), ('plog_A', _dp),
        ('cond', _dp), ('visc', _dp), ('diff', _dp), ('C1', _dp), ('C2', _dp),
        ('rcp_diff', C.c_int),
    ]


def build_libs(with_ref=True):
    """Compile the oracle (and oracle/_ref when the reference tree is present)."""
    targets = ['all']
    if with_ref and os.path.isdir(os.environ.get('REF', '/root/reference')):
        targets.append('ref')
    subprocess.run(['make', '-s', '-C', HERE] + targets, check=True)


def _load(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise FileNotFoundError(f'{path} is not built (run make -C oracle)')
    return C.CDLL(path)


def _ptr(a, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


class Oracle:
    """FP64 CPU restatement bound to one mechanism's tables (kinetix_amd/gen/tables.py)."""

    def __init__(self, tables, fast=False):
        self.lib = _load('lib/libkxoracle_fast.so' if fast else 'lib/libkxoracle.so')
        self.t = {}
        for k, v in tables.items():
            self.t[k] = np.asarray(v)
        t = self.t
        self.N = int(t['n_species'])
        self.R = int(t['n_reactions'])
        m = _Mech()
        m.n_species, m.n_active, m.n_reactions = self.N, int(t['n_active']), self.R
        m.R, m.one_atm_R, m.R_one_atm = float(t['R']), float(t['one_atm_R']), float(t['R_one_atm'])
        for name, ftype in _Mech._fields_[6:]:
            if name == 'rcp_diff':
                m.rcp_diff = int(t.get('rcp_diff', 0))
                continue
            is_int = ftype is _ip
            arr = t.get(name)
            if arr is None or arr.size == 0:
                arr = np.zeros(1, dtype=np.int32 if is_int else np.float64)
            else:
                arr = np.ascontiguousarray(arr, dtype=np.int32 if is_int else np.float64)
            self.t['_c_' + name] = arr          # keep the buffer alive
            setattr(m, name, _ptr(arr, C.c_int if is_int else C.c_double))
        self.m = m
        L = self.lib
        L.kxo_production_rates.argtypes = [C.POINTER(_Mech), C.c_int, C.c_int, C.c_int, C.c_double,
                                           _dp, _dp, _dp, _dp, C.c_double]
        L.kxo_transport.argtypes = [C.POINTER(_Mech), C.c_int, C.c_int, C.c_int, C.c_double,
                                    _dp, _dp, _dp, _dp, _dp, _dp, C.c_double]
        L.kxo_thermo.argtypes = [C.POINTER(_Mech), C.c_int, C.c_int, C.c_int, C.c_double,
                                 _dp, _dp, _dp, _dp, C.c_double]
        L.kxo_reaction_rates.argtypes = [C.POINTER(_Mech), C.c_int, C.c_int, C.c_int, C.c_double,
                                         _dp, _dp, _dp, _dp, C.c_double]

    # Threads: each call works on a contiguous chunk of states (ctypes releases the GIL).
    @staticmethod
    def _chunks(n, threads):
        threads = max(1, min(threads, n))
        step = (n + threads - 1) // threads
        return [(lo, min(n, lo + step)) for lo in range(0, n, step)]

    def production_rates(self, state, p, Tref=1.0, p_state=None, scale=False, threads=1):
        """state: SoA (N+1, n) [T/Tref, Y_0..Y_N-1]. p in Pa. Returns rates (N+1, n)
        [HRR, W_k wdot_k] and, if scale, the condition scale of the same shape."""
        state = np.ascontiguousarray(state, dtype=np.float64)
        n = state.shape[1]
        out = np.zeros_like(state)
        sc = np.zeros_like(state) if scale else None
        ps = np.ascontiguousarray(p_state, dtype=np.float64) if p_state is not None else None

        def run(lo, hi):
            # sub-view: pointer offset lo, global strides n (the reference's offset args)
            off = lo * 8
            self.lib.kxo_production_rates(
                C.byref(self.m), hi - lo, n, n, float(p),
                C.cast(ps.ctypes.data + off, _dp) if ps is not None else None,
                C.cast(state.ctypes.data + off, _dp), C.cast(out.ctypes.data + off, _dp),
                C.cast(sc.ctypes.data + off, _dp) if sc is not None else None, float(Tref))
        self._run(n, threads, run)
        return (out, sc) if scale else out

    def transport(self, state, p=1.0, Tref=1.0, p_state=None, threads=1):
        """Returns (conductivity[n], viscosity[n], rhoD[N, n], Dkm[N, n] or None)."""
        state = np.ascontiguousarray(state, dtype=np.float64)
        n = state.shape[1]
        cond = np.zeros(n)
        visc = np.zeros(n)
        rhoD = np.zeros((self.N, n))
        dkm = np.zeros((self.N, n)) if p_state is not None else None
        ps = np.ascontiguousarray(p_state, dtype=np.float64) if p_state is not None else None

        def run(lo, hi):
            off = lo * 8
            self.lib.kxo_transport(
                C.byref(self.m), hi - lo, n, n, float(p),
                C.cast(ps.ctypes.data + off, _dp) if ps is not None else None,
                C.cast(state.ctypes.data + off, _dp), C.cast(cond.ctypes.data + off, _dp),
                C.cast(visc.ctypes.data + off, _dp), C.cast(rhoD.ctypes.data + off, _dp),
                C.cast(dkm.ctypes.data + off, _dp) if dkm is not None else None, float(Tref))
        self._run(n, threads, run)
        return cond, visc, rhoD, dkm

    def thermo(self, state, p, Tref=1.0):
        state = np.ascontiguousarray(state, dtype=np.float64)
        n = state.shape[1]
        rho, cp, rhocp = np.zeros(n), np.zeros((self.N, n)), np.zeros(n)
        self.lib.kxo_thermo(C.byref(self.m), n, n, n, float(p) / float(self.t['R']),
                            _ptr(state), _ptr(rho), _ptr(cp), _ptr(rhocp), float(Tref))
        return rho, cp, rhocp

    def reaction_rates(self, state, p, Tref=1.0, p_state=None, forward=False):
        """Net rates q_r (R, n); with forward=True also the forward rates k_f prod C^nu'."""
        state = np.ascontiguousarray(state, dtype=np.float64)
        n = state.shape[1]
        q = np.zeros((self.R, n))
        qf = np.zeros((self.R, n)) if forward else None
        ps = np.ascontiguousarray(p_state, dtype=np.float64) if p_state is not None else None
        self.lib.kxo_reaction_rates(C.byref(self.m), n, n, n, float(p), _ptr(ps), _ptr(state),
                                    _ptr(q), _ptr(qf), float(Tref))
        return (q, qf) if forward else q

    def _run(self, n, threads, fn):
        if threads <= 1:
            fn(0, n)
            return
        with ThreadPoolExecutor(threads) as ex:
            list(ex.map(lambda c: fn(*c), self._chunks(n, threads)))


class Reference:
    """The reference's own compiled GRI-3.0 routines (oracle/_ref), FP64."""

    def __init__(self, flavour='fast'):
        self.lib = _load(f'_ref/libkxref_{flavour}.so')
        L = self.lib
        L.kxref_n_species.restype = C.c_int
        L.kxref_n_reactions.restype = C.c_int
        L.kxref_species_names.restype = C.c_char_p
        L.kxref_production_rates.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp,
                                             C.c_double, C.c_int]
        L.kxref_transport.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp,
                                      C.c_double, C.c_int]
        L.kxref_thermo.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _dp, _dp, _dp, _dp,
                                   C.c_double, C.c_int]
        self.N = L.kxref_n_species()
        self.species = L.kxref_species_names().decode().split()

    @staticmethod
    def available(flavour='fast'):
        return os.path.exists(os.path.join(HERE, f'_ref/libkxref_{flavour}.so'))

    def production_rates(self, state, p, Tref=1.0, threads=1):
        state = np.ascontiguousarray(state, dtype=np.float64)
        n = state.shape[1]
        out = np.zeros_like(state)
        R = 1.380649e-23 * 6.02214076e23
        self.lib.kxref_production_rates(n, n, n, float(p) / R, _ptr(state), _ptr(out),
                                        float(Tref), int(threads))
        return out

    def transport(self, state, p=1.0, Tref=1.0, threads=1):
        state = np.ascontiguousarray(state, dtype=np.float64)
        n = state.shape[1]
        cond, visc, rhoD = np.zeros(n), np.zeros(n), np.zeros((self.N, n))
        self.lib.kxref_transport(n, n, n, float(p), _ptr(state), _ptr(cond), _ptr(visc),
                                 _ptr(rhoD), float(Tref), int(threads))
        return cond, visc, rhoD

    def thermo(self, state, p, Tref=1.0, threads=1):
        state = np.ascontiguousarray(state, dtype=np.float64)
        n = state.shape[1]
        rho, cp, rhocp = np.zeros(n), np.zeros((self.N, n)), np.zeros(n)
        R = 1.380649e-23 * 6.02214076e23
        self.lib.kxref_thermo(n, n, n, float(p) / R, _ptr(state), _ptr(rho), _ptr(cp), _ptr(rhocp),
                              float(Tref), int(threads))
        return rho, cp, rhocp
