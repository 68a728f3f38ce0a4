"""The matrix-core lane layouts the tile kernels rely on (kernels_common.h: mfmaRow): FP64 component r of lane l is
row 4 r + l / 16, FP32 component r is row 4 (l / 16) + r; column l % 16 for both. The bf16 16x16x32 product (lane l
supplies A[l % 16][8 (l / 16) + j], B[8 (l / 16) + j][l % 16]) leaves row 4 (l / 16) + r, column l % 16."""
import ctypes as C
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "probe", "libmfmaprobe.so")


def test_mfma_16x16x4_layouts():
    lib = C.CDLL(LIB)
    rng = np.random.default_rng(0)
    A = rng.integers(1, 9, size=(16, 4)).astype(np.float64)
    B = rng.integers(1, 9, size=(4, 16)).astype(np.float64)
    raw64 = np.zeros(256)
    raw32 = np.zeros(256, dtype=np.float32)
    f64p = np.ctypeslib.ndpointer(np.float64, flags="C")
    lib.probe_mfma.argtypes = [f64p, f64p, f64p, np.ctypeslib.ndpointer(np.float32, flags="C")]
    assert lib.probe_mfma(np.ascontiguousarray(A), np.ascontiguousarray(B), raw64, raw32) == 0
    D = A @ B
    for l in range(64):
        for r in range(4):
            col = l % 16
            assert raw64[4 * l + r] == D[4 * r + l // 16, col], ("f64", l, r)
            assert raw32[4 * l + r] == D[4 * (l // 16) + r, col], ("f32", l, r)


def test_mfma_16x16x32_bf16_layout():
    lib = C.CDLL(LIB)
    rng = np.random.default_rng(1)
    A = rng.integers(-4, 5, size=(16, 32)).astype(np.float32)
    B = rng.integers(-4, 5, size=(32, 16)).astype(np.float32)
    raw = np.zeros(256, dtype=np.float32)
    f32p = np.ctypeslib.ndpointer(np.float32, flags="C")
    lib.probe_mfma_bf16.argtypes = [f32p, f32p, f32p]
    assert lib.probe_mfma_bf16(np.ascontiguousarray(A), np.ascontiguousarray(B), raw) == 0
    D = A.astype(np.float64) @ B.astype(np.float64)
    for l in range(64):
        for r in range(4):
            assert raw[4 * l + r] == D[4 * (l // 16) + r, l % 16], ("bf16", l, r)
