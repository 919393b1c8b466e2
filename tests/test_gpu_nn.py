"""GPU tests of the warm-start network kernels (BF16 matrix cores).

The operand layout of the 16x16x16 BF16 MFMA that every convolution kernel
relies on is checked first against a host matmul of the same BF16 values.
"""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bf16_bits(x):
    """Round-to-nearest-even FP32 -> BF16 bits (numpy)."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_value(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


def test_mfma_layout():
    import ctypes as C

    from pot3d_amd import _lib
    say("[gpu] mfma 16x16x16 bf16 layout")
    L = _lib.load()
    rng = np.random.default_rng(3)
    A = bf16_bits(rng.standard_normal((16, 16)))
    B = bf16_bits(rng.standard_normal((16, 16)))
    D = np.zeros((16, 16), np.float32)
    _lib.check(L.p3d_nn_mfma_check(A.ctypes.data_as(C.POINTER(C.c_ushort)),
                                   B.ctypes.data_as(C.POINTER(C.c_ushort)),
                                   D.ctypes.data_as(C.POINTER(C.c_float))))
    ref = bf16_value(A).astype(np.float64) @ bf16_value(B).astype(np.float64)
    assert np.max(np.abs(D - ref)) <= 1e-5 * np.max(np.abs(ref))
