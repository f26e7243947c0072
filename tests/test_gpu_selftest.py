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


def _f64_probe(mode):
    L = lib()
    L.kx_selftest_mfma_f64_layout.argtypes = [C.c_int, C.c_void_p, C.c_void_p]
    out = torch.full((256,), np.nan, dtype=torch.float64, device='cuda')
    assert L.kx_selftest_mfma_f64_layout(mode, C.c_void_p(out.data_ptr()), None) == 0
    torch.cuda.synchronize()
    return out.cpu().numpy().reshape(64, 4)


def _raw(D):
    """D (16x16) in the FP64 result layout: lane l, element r = D[4 r + l // 16][l % 16]."""
    return np.array([[D[4 * r + l // 16, l % 16] for r in range(4)] for l in range(64)])


def test_mfma_f64_16x16x4_result_layout():
    """Every result entry names its own (row, column): D[i][j] = 16 i + j."""
    got = _f64_probe(1)
    D = np.array([[16.0 * i + j for j in range(16)] for i in range(16)])
    assert np.array_equal(got, _raw(D)), got


def test_mfma_f64_16x16x4_operand_layout():
    got = _f64_probe(0)
    A = np.array([[i + 16 * k + 1 for k in range(4)] for i in range(16)], dtype=np.float64)
    B = np.array([[(k + 1) * 100 + j for j in range(16)] for k in range(4)], dtype=np.float64)
    assert np.array_equal(got, _raw(A @ B))


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


def test_newton_reciprocal_within_one_ulp():
    """kx_rcp, the transport kernels' reciprocal of the positive fit polynomials: within 1 ulp
    of the exact 1/y over 60 decades of both signs, exact on powers of two."""
    L = lib()
    L.kx_selftest_rcp.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    rng = np.random.default_rng(20261202)
    y = 10.0 ** rng.uniform(-30, 30, 1 << 20)
    y[: 1 << 18] = rng.uniform(1e-6, 1e3, 1 << 18)            # the fits' own range
    y[1 << 18: 1 << 19] *= -1.0
    pw = 2.0 ** np.arange(-60, 61)
    y = np.concatenate([y, pw, np.nextafter(pw, 0), np.nextafter(pw, np.inf), [1.0, 3.0, 0.1, 7.0]])
    d_in = torch.from_numpy(y).cuda()
    d_out = torch.full_like(d_in, np.nan)
    assert L.kx_selftest_rcp(y.size, C.c_void_p(d_in.data_ptr()), C.c_void_p(d_out.data_ptr()), None) == 0
    torch.cuda.synchronize()
    r = d_out.cpu().numpy()
    exact = 1.0 / y                                             # correctly rounded (IEEE)
    ulp = np.abs(r - exact) / np.spacing(np.abs(exact))
    assert np.all(np.isfinite(r)) and ulp.max() <= 1.0, ulp.max()
    assert np.array_equal(r[(1 << 20):(1 << 20) + pw.size], 1.0 / pw)
