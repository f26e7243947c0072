"""The C-ABI library loads and exports every entry point include/*.h declares; host-only
calls and argument checks behave without a GPU (no kernel is launched here)."""
import ctypes as C
import glob
import os
import re

import pytest

from kinetix_amd import api

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECL = re.compile(r'^\s*(?:const\s+)?(?:int|size_t|double|void|char)\s*\*?\s*(kx_\w+)\s*\(', re.M)


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, 'include', '*.h')):
        names |= set(DECL.findall(open(h).read()))
    return sorted(names)


@pytest.fixture(scope='module')
def L():
    if not os.path.exists(api.LIB_PATH):
        pytest.skip('libkinetix_mi.so not built (python -m kinetix_amd.build)')
    return api.lib()


def test_headers_declare_the_boundary():
    names = declared()
    for must in ('kx_init', 'kx_build', 'kx_finalize', 'kx_production_rates',
                 'kx_production_rates_mixed', 'kx_production_rates_mixed_mfma',
                 'kx_ref_mass_fractions', 'kx_mixture_avg_transport',
                 'kx_mixture_avg_transport_pstate', 'kx_thermodynamic_props', 'kx_n_species',
                 'kx_species_name', 'kx_sur_transport', 'kx_sur_workspace_bytes',
                 'kx_mlp_forward', 'kx_mlp_backward', 'kx_mlp_adam', 'kx_mlp_infer_bf16'):
        assert must in names, must


def test_every_declared_symbol_is_exported(L):
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_host_only_queries(L):
    H = L.kx_mlp_width()
    assert H == 256
    assert L.kx_mlp_n_params() == H * 64 + H + 2 * (H * H + H) + 64 * H + 64 == 164672
    assert L.kx_mlp_n_weights_bf16() == (H // 16) * 2 * 512 + 2 * (H // 16) * (H // 32) * 512 \
        + 4 * (H // 32) * 512
    L.kx_sur_workspace_bytes.restype = C.c_size_t
    L.kx_sur_workspace_bytes.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int]
    assert L.kx_sur_workspace_bytes(987654, 1000, 100, 0) == 0         # unknown handle


def test_unknown_handle_is_an_error_not_a_launch(L):
    rc = L.kx_production_rates(987654, 16, 16, 16, 1.0, None, None, None)
    assert rc == -2                                                    # KX_ERR_NOT_INITIALIZED
    assert b'handle' in L.kx_last_error()
    assert L.kx_n_species(987654) == -2
    out = (C.c_double * 4)()
    assert L.kx_ref_mass_fractions(987654, out) == -2
    rc = L.kx_mlp_forward(None, 64, None, None, None)
    assert rc == -1                                                    # KX_ERR_INVALID_ARG
    rc = L.kx_mlp_forward(C.c_void_p(16), 65, C.c_void_p(16), C.c_void_p(16), None)
    assert rc == -1                                                    # batch not a multiple of 64


def test_init_image_rejects_malformed_text_before_any_device_call(L):
    """kx_init_image parses the mechanism text before it touches the device."""
    h = C.c_int(0)
    meta = b'n_species 2\nspecies A\nmolar_mass 0.002 0.032\n'          # one name for two species
    rc = L.kx_init_image(meta, len(meta), b'\x7fELF', 4, 0, C.byref(h))
    assert rc == -5                                                    # KX_ERR_IO
    assert b'malformed mechanism image text' in L.kx_last_error()
    assert L.kx_init_image(None, 0, b'\x7fELF', 4, 0, C.byref(h)) == -1
