"""The 16x16x4 MFMA lane layouts the H*X kernels are written for:
A(i,k) from lane i + 16 k, B(k,j) from lane j + 16 k; accumulator q of lane l
holds D(row, l % 16) with row = l / 16 + 4 q for FP64 (interleaved) and
row = 4 (l / 16) + q for FP32 (blocked)."""
import numpy as np
import pytest

from scalit_amd import _lib as L

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [L.FP64, L.FP32])
def test_mfma_16x16x4_layout(prec):
    out = np.zeros(256)
    L.check(L.lib().scalit_debug_mfma_probe(prec, out.ctypes.data_as(L._dp)))
    for lane in range(64):
        for q in range(4):
            v = out[4 * lane + q]
            i = int((v - 1) // 100) - 1
            j = int(round(v - 100 * (i + 1) - 1))
            row = lane // 16 + 4 * q if prec == L.FP64 else 4 * (lane // 16) + q
            assert (i, j) == (row, lane % 16), (lane, q, v)
