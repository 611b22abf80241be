"""MFMA complex GEMM (C-ABI) against numpy: every operand form, ragged sizes, both
tile shapes (32x32 below 256 64x64 tiles, 64x64 from there: 1030 x 520 x 2)."""
import ctypes as C

import numpy as np
import pytest
import torch

from slakonet_amd import _lib

pytestmark = pytest.mark.gpu


def _op(x, op):
    if op & 1:
        x = np.swapaxes(x, -1, -2)
    if op & 2:
        x = np.conj(x)
    return x


@pytest.mark.parametrize("dtype", ["c128", "c64"])
@pytest.mark.parametrize("opa,opb", [(0, 0), (1, 0), (3, 2), (0, 3), (2, 1)])
@pytest.mark.parametrize("mnk", [(1, 1, 1), (17, 70, 33), (64, 64, 16), (129, 65, 100), (1030, 520, 40)])
def test_gemm_matches_numpy(dtype, opa, opb, mnk):
    M, N, K = mnk
    rng = np.random.default_rng(M * 7 + N + K)
    batch = 2
    npdt = np.complex128 if dtype == "c128" else np.complex64
    a_shape = (batch, K, M) if opa & 1 else (batch, M, K)
    b_shape = (batch, N, K) if opb & 1 else (batch, K, N)
    A = (rng.normal(size=a_shape) + 1j * rng.normal(size=a_shape)).astype(npdt)
    B = (rng.normal(size=b_shape) + 1j * rng.normal(size=b_shape)).astype(npdt)
    Cm = (rng.normal(size=(batch, M, N)) + 1j * rng.normal(size=(batch, M, N))).astype(npdt)
    alpha, beta = -0.75, 0.5
    ref = alpha * (_op(A.astype(np.complex128), opa) @ _op(B.astype(np.complex128), opb)) + beta * Cm
    dA, dB, dC = (torch.from_numpy(x).cuda() for x in (A, B, Cm))
    lib = _lib.lib()
    fn = lib.slk_zgemm if dtype == "c128" else lib.slk_cgemm
    rc = fn(batch, M, N, K, alpha, C.c_void_p(dA.data_ptr()), a_shape[2], a_shape[1] * a_shape[2], opa,
            C.c_void_p(dB.data_ptr()), b_shape[2], b_shape[1] * b_shape[2], opb, beta,
            C.c_void_p(dC.data_ptr()), N, M * N, _lib.stream_handle())
    assert rc == 0
    got = dC.cpu().numpy()
    tol = 1e-12 if dtype == "c128" else 2e-5
    assert np.abs(got - ref).max() <= tol * max(1.0, np.sqrt(K)) * max(1.0, np.abs(ref).max())
