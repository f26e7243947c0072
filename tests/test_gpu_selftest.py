"""The MFMA lane layout assumed by the generated mixed-precision kernels, checked on the
device against the host product of the same closed-form operands."""
import ctypes as C

import numpy as np
import pytest
import torch

from kinetix_amd.api import lib

pytestmark = pytest.mark.gpu


def test_mfma_f32_16x16x4_layout():
    L = lib()
    L.kx_selftest_mfma_layout.argtypes = [C.c_void_p, C.c_void_p]
    out = torch.full((256,), np.nan, dtype=torch.float32, device='cuda')
    assert L.kx_selftest_mfma_layout(C.c_void_p(out.data_ptr()), None) == 0
    torch.cuda.synchronize()
    A = np.array([[i + 16 * k + 1 for k in range(4)] for i in range(16)], dtype=np.float64)
    B = np.array([[(k + 1) * 100 + j for j in range(16)] for k in range(4)], dtype=np.float64)
    assert np.array_equal(out.cpu().numpy().reshape(16, 16).astype(np.float64), A @ B)


def test_mfma_bf16_16x16x16_layout():
    L = lib()
    L.kx_selftest_mfma_bf16_layout.argtypes = [C.c_void_p, C.c_void_p]
    out = torch.full((256,), np.nan, dtype=torch.float32, device='cuda')
    assert L.kx_selftest_mfma_bf16_layout(C.c_void_p(out.data_ptr()), None) == 0
    torch.cuda.synchronize()
    A = np.array([[(i + 16 * k + 1) / 16.0 for k in range(16)] for i in range(16)])
    B = np.array([[(k + 1) * 3 + j for j in range(16)] for k in range(16)], dtype=np.float64)
    assert np.array_equal(out.cpu().numpy().reshape(16, 16).astype(np.float64), A @ B)


def test_mfma_bf16_16x16x32_layout():
    """The layout the surrogate's inference kernels (kx_mlp_wave.h) are built on."""
    L = lib()
    L.kx_selftest_mfma_bf16_k32_layout.argtypes = [C.c_void_p, C.c_void_p]
    out = torch.full((256,), np.nan, dtype=torch.float32, device='cuda')
    assert L.kx_selftest_mfma_bf16_k32_layout(C.c_void_p(out.data_ptr()), None) == 0
    torch.cuda.synchronize()
    A = np.array([[(i * 7 + k * 3) % 17 - 8 for k in range(32)] for i in range(16)], dtype=np.float64)
    B = np.array([[(k * 5 + j * 11) % 13 - 6 for j in range(16)] for k in range(32)], dtype=np.float64)
    assert np.array_equal(out.cpu().numpy().reshape(16, 16).astype(np.float64), A @ B)
