"""Python mirror of the reference host API (benchmark/src/kinetix.hpp:17-107) over the C ABI
of include/kinetix_mi.h (libkinetix_mi.so, loaded with ctypes).

Same names, argument meaning and order as the reference's kinetix:: functions; device
buffers are torch tensors (FP64, on the GPU) instead of occa::memory, and every call is
enqueued on torch's current HIP stream. Errors raise KinetixError (the reference asserts
or aborts). No CPU fallback exists: if the HIP library or a code object is missing, the
calls fail loudly.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, 'lib', 'libkinetix_mi.so')
MECH_ROOT = os.path.join(PKG, 'mechs')

_dp = C.POINTER(C.c_double)


class KinetixError(RuntimeError):
    pass


_lib = None


def lib():
    """The loaded libkinetix_mi.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise KinetixError(f'{LIB_PATH} not built; run `python -m kinetix_amd.build`')
        # torch bundles its own HIP runtime under the same soname (libamdhip64.so.7): load it
        # first so that the library binds to that one runtime and torch's streams and device
        # pointers are valid in both (loaded the other way round, torch finds no GPU).
        import torch  # noqa: F401
        L = C.CDLL(LIB_PATH)
        L.kx_last_error.restype = C.c_char_p
        L.kx_init.argtypes = [C.c_char_p, C.c_int, C.POINTER(C.c_int)]
        L.kx_init_image.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_size_t, C.c_int,
                                    C.POINTER(C.c_int)]
        L.kx_build.argtypes = [C.c_int, C.c_double, C.c_double, _dp]
        L.kx_finalize.argtypes = [C.c_int]
        for fn in ('kx_production_rates', 'kx_production_rates_mixed',
                   'kx_production_rates_mixed_valu', 'kx_production_rates_mixed_mfma',
                   'kx_production_rates_routines'):
            getattr(L, fn).argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                       C.c_void_p, C.c_void_p]
        for fn in ('kx_mixture_avg_transport', 'kx_mixture_avg_transport_routines',
                   'kx_mixture_avg_transport_mfma'):
            getattr(L, fn).argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.kx_mixture_avg_transport_pstate.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int,
                                                      C.c_void_p, C.c_void_p, C.c_void_p,
                                                      C.c_void_p, C.c_void_p, C.c_void_p,
                                                      C.c_void_p]
        L.kx_thermodynamic_props.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double,
                                             C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                             C.c_void_p]
        for fn in ('kx_n_species', 'kx_n_active_species', 'kx_n_reactions'):
            getattr(L, fn).argtypes = [C.c_int]
        L.kx_species_name.argtypes = [C.c_int, C.c_int, C.c_char_p, C.c_int]
        L.kx_species_index.argtypes = [C.c_int, C.c_char_p]
        L.kx_molecular_weights.argtypes = [C.c_int, _dp]
        L.kx_molar_masses.argtypes = [C.c_int, _dp]
        L.kx_ref_mass_fractions.argtypes = [C.c_int, _dp]
        for fn in ('kx_ref_pressure', 'kx_ref_temperature', 'kx_ref_mean_molecular_weight'):
            getattr(L, fn).argtypes = [C.c_int]
            getattr(L, fn).restype = C.c_double
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise KinetixError(f'kinetix_mi error {rc}: {lib().kx_last_error().decode()}')
    return rc


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _stream(t):
    import torch
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _check_tensor(name, t, numel=None):
    import torch
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise KinetixError(f'{name} must be a GPU tensor')
    if t.dtype != torch.float64:
        raise KinetixError(f'{name} must be float64')
    if not t.is_contiguous():
        raise KinetixError(f'{name} must be contiguous')
    if numel is not None and t.numel() < numel:
        raise KinetixError(f'{name} has {t.numel()} elements, needs {numel}')


def mech_dir(name):
    """Directory of a generated mechanism: one shipped with the package (e.g. 'gri30'), a
    directory, or a Cantera YAML file, generated and compiled into the build cache
    (kinetix_amd.cache: .cache/KINETIX/<stem>, regenerated only when its .hash is stale, as
    the reference's init does)."""
    if name.endswith(('.yaml', '.yml')) and os.path.isfile(name):
        from . import cache
        return cache.ensure(name)
    d = name if os.path.isdir(name) else os.path.join(MECH_ROOT, name)
    if not os.path.exists(os.path.join(d, 'kernels.hsaco')):
        raise KinetixError(f'no built code object in {d} (python -m kinetix_amd.build)')
    return d


class Kinetix:
    """One mechanism loaded on one device (the reference's process-global kinetix:: state)."""

    def __init__(self, mech='gri30', device=0, image=None):
        """image: optional (mech_txt bytes, code object bytes) loaded from memory through
        kx_init_image (the host-application hook) instead of the mechanism directory."""
        h = C.c_int(0)
        if image is None:
            self.dir = mech_dir(mech)
            _check(lib().kx_init(self.dir.encode(), int(device), C.byref(h)))
        else:
            self.dir = None
            meta, code = image
            _check(lib().kx_init_image(meta, len(meta), code, len(code), int(device), C.byref(h)))
        self.h = h.value
        self.device = device
        self.n_species = _check(lib().kx_n_species(self.h))
        self.n_active_species = _check(lib().kx_n_active_species(self.h))
        self.n_reactions = _check(lib().kx_n_reactions(self.h))
        buf = C.create_string_buffer(256)
        self.species_names = []
        for k in range(self.n_species):
            _check(lib().kx_species_name(self.h, k, buf, 256))
            self.species_names.append(buf.value.decode())
        w = np.zeros(self.n_species)
        _check(lib().kx_molar_masses(self.h, w.ctypes.data_as(_dp)))
        self.molar_masses = w
        self._built = False

    # kinetix::build
    def build(self, ref_pressure, ref_temperature, ref_mass_fractions):
        y = np.ascontiguousarray(ref_mass_fractions, dtype=np.float64)
        if y.size != self.n_species:
            raise KinetixError('ref_mass_fractions must have n_species entries')
        _check(lib().kx_build(self.h, float(ref_pressure), float(ref_temperature), y.ctypes.data_as(_dp)))
        self._ref_y = y.copy()
        self._built = True

    def _layout(self, n_states, offsetT, offset):
        return self.n_species * offset + offsetT  # minimum buffer length in elements

    # kinetix::productionRates
    def productionRates(self, n_states, offsetT, offset, pressure, o_state, o_rates, mixed=False):
        """mixed: False (FP64); True or 'valu' (mixed precision, q.nu as FP64 FMAs: the fastest
        mixed kernel, kx_production_rates_mixed); 'mfma' (mixed precision, q.nu on FP32 MFMA,
        kx_production_rates_mixed_mfma); 'routines' (FP64 through the generated routine
        files, as the reference's kernel calls them)."""
        need = offsetT + (self.n_species - 1) * offset + n_states if n_states else 0
        _check_tensor('state', o_state, need)
        _check_tensor('rates', o_rates, need)
        fn = {False: lib().kx_production_rates, True: lib().kx_production_rates_mixed,
              'mfma': lib().kx_production_rates_mixed_mfma,
              'valu': lib().kx_production_rates_mixed_valu,
              'routines': lib().kx_production_rates_routines}[mixed]
        _check(fn(self.h, int(n_states), int(offsetT), int(offset), float(pressure), _ptr(o_state),
                  _ptr(o_rates), _stream(o_state)))

    def productionRatesMixed(self, n_states, offsetT, offset, pressure, o_state, o_rates):
        self.productionRates(n_states, offsetT, offset, pressure, o_state, o_rates, mixed=True)

    # kinetix::mixtureAvgTransportProps
    def mixtureAvgTransportProps(self, n_states, offsetT, offset, pressure, o_state, o_viscosity,
                                 o_thermalConductivity, o_densityDiffCoeffs, routines=False,
                                 mfma=False):
        """routines=True: through the generated routine files, as the reference's transport
        kernel calls them (kx_mixture_avg_transport_routines). mfma=True: the Wilke sum on
        FP64 MFMA (kx_mixture_avg_transport_mfma), same outputs."""
        need = offsetT + (self.n_species - 1) * offset + n_states if n_states else 0
        _check_tensor('state', o_state, need)
        _check_tensor('viscosity', o_viscosity, n_states)
        _check_tensor('thermalConductivity', o_thermalConductivity, n_states)
        _check_tensor('densityDiffCoeffs', o_densityDiffCoeffs,
                      (self.n_species - 1) * offset + n_states if n_states else 0)
        fn = (lib().kx_mixture_avg_transport_routines if routines else
              lib().kx_mixture_avg_transport_mfma if mfma else lib().kx_mixture_avg_transport)
        _check(fn(self.h, int(n_states), int(offsetT), int(offset), float(pressure), _ptr(o_state),
                  _ptr(o_viscosity), _ptr(o_thermalConductivity), _ptr(o_densityDiffCoeffs),
                  _stream(o_state)))

    def mixtureAvgTransportPropsPstate(self, n_states, offsetT, offset, o_p, o_state, o_viscosity,
                                       o_thermalConductivity, o_densityDiffCoeffs, o_Dkm=None):
        need = offsetT + (self.n_species - 1) * offset + n_states if n_states else 0
        _check_tensor('state', o_state, need)
        _check_tensor('p', o_p, n_states)
        for nm, t in (('viscosity', o_viscosity), ('thermalConductivity', o_thermalConductivity)):
            _check_tensor(nm, t, n_states)
        rows = (self.n_species - 1) * offset + n_states if n_states else 0
        _check_tensor('densityDiffCoeffs', o_densityDiffCoeffs, rows)
        if o_Dkm is not None:
            _check_tensor('Dkm', o_Dkm, rows)
        _check(lib().kx_mixture_avg_transport_pstate(
            self.h, int(n_states), int(offsetT), int(offset), _ptr(o_p), _ptr(o_state),
            _ptr(o_viscosity), _ptr(o_thermalConductivity), _ptr(o_densityDiffCoeffs), _ptr(o_Dkm),
            _stream(o_state)))

    # kinetix::thermodynamicProps
    def thermodynamicProps(self, n_states, offsetT, offset, pressure, o_state, o_rho, o_cpi, o_rhoCp):
        need = offsetT + (self.n_species - 1) * offset + n_states if n_states else 0
        _check_tensor('state', o_state, need)
        _check_tensor('rho', o_rho, n_states)
        _check_tensor('rhoCp', o_rhoCp, n_states)
        _check_tensor('cp', o_cpi, (self.n_species - 1) * offset + n_states if n_states else 0)
        _check(lib().kx_thermodynamic_props(self.h, int(n_states), int(offsetT), int(offset),
                                            float(pressure), _ptr(o_state), _ptr(o_rho), _ptr(o_cpi),
                                            _ptr(o_rhoCp), _stream(o_state)))

    # queries (kinetix.hpp:85-103)
    def nSpecies(self):
        return self.n_species

    def nActiveSpecies(self):
        return self.n_active_species

    def nReactions(self):
        return self.n_reactions

    def speciesNames(self):
        return list(self.species_names)

    def speciesIndex(self, name):
        return lib().kx_species_index(self.h, name.encode())

    def molecularWeights(self):
        out = np.zeros(self.n_species)
        _check(lib().kx_molecular_weights(self.h, out.ctypes.data_as(_dp)))
        return out

    def refPressure(self):
        return lib().kx_ref_pressure(self.h)

    def refTemperature(self):
        return lib().kx_ref_temperature(self.h)

    def refMassFractions(self):
        out = np.zeros(self.n_species)
        _check(lib().kx_ref_mass_fractions(self.h, out.ctypes.data_as(_dp)))
        return out

    def refMeanMolecularWeight(self):
        return lib().kx_ref_mean_molecular_weight(self.h)

    def close(self):
        if getattr(self, 'h', None):
            lib().kx_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
