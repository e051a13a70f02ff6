"""The pair-math helpers against the IEEE operations, on the device.

sqrt_rn_pair: every non-negative finite float; bitwise equal to the correctly
rounded square root from 2^-96 up, and within [0, 2^-40] below (where a length
feeds only q < 2^-25 and a zero gradient).
recip_pair: every mantissa at every exponent of 2^-23 .. 2^1; bitwise equal to
the IEEE 1 / b.  This is the precondition under which div_pair's single FMA
correction gives the correctly rounded quotient (Markstein's theorem).
div_pair: every divisor mantissa in [1, 2) at every exponent of 2^-23 .. 2^1
(divisors in [2^-23, 4)), against 16 numerators each: signed zeros, +-b, values
near 2^-59 and random magnitudes in [2^-59, b]; bitwise equal to IEEE a / b.
div_pair_zs (the gradient's division): the same quotients by value, a zero
quotient of either sign (the sums it feeds start at +0 and are never -0)."""
import ctypes as C
import os

import pytest

from fluidsolver_amd import build as fbuild

pytestmark = pytest.mark.gpu


def _lib():
    if not os.path.exists(fbuild.MATHCHECK_LIB):
        pytest.skip("math check library not built")
    L = C.CDLL(fbuild.MATHCHECK_LIB)
    L.fsa_mc_sqrt.restype = C.c_int
    L.fsa_mc_sqrt.argtypes = [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.fsa_mc_recip.restype = C.c_int
    L.fsa_mc_recip.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_uint64)]
    L.fsa_mc_div.restype = C.c_int
    L.fsa_mc_div.argtypes = [C.c_int, C.c_int, C.c_uint32, C.c_uint64, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.fsa_mc_int_abs.restype = C.c_int
    L.fsa_mc_int_abs.argtypes = [C.c_void_p, C.c_void_p, C.c_uint32]
    return L


def test_sqrt_pair_exhaustive():
    L = _lib()
    m, b = C.c_uint64(0), C.c_uint64(0)
    print("case sqrt exhaustive", flush=True)
    assert L.fsa_mc_sqrt(C.byref(m), C.byref(b)) == 0
    assert m.value == 0 and b.value == 0


def test_div_pair_all_divisor_mantissas():
    L = _lib()
    m, t = C.c_uint64(0), C.c_uint64(0)
    print("case div mantissas", flush=True)
    assert L.fsa_mc_div(-23, 1, 16, 12345, C.byref(m), C.byref(t)) == 0
    assert t.value == (1 << 23) * 25 * 16
    assert m.value == 0


def test_recip_pair_correctly_rounded():
    L = _lib()
    m = C.c_uint64(0)
    print("case recip mantissas", flush=True)
    assert L.fsa_mc_recip(-23, 1, C.byref(m)) == 0
    assert m.value == 0


def test_x86_int_abs_on_the_device(oracle_lib):
    """The 2D loop condition's integer abs() (IISPHFluidSolver.hpp:400-402) on the device
    equals the oracle's x86 form, NaN and out-of-range values (INT_MIN) included."""
    import numpy as np

    L = _lib()
    vals = np.array([0.0, 0.7, -0.7, 1.0, -2.5, 1e-3, 3.99, 2147483520.0, -2147483648.0, 2147483648.0, 3.0e9,
                     -3.0e9, float("nan"), float("inf"), float("-inf"), 1e30, -1e30, 65536.5, -123456.75],
                    dtype=np.float32)
    out = np.zeros_like(vals)
    print("case x86 int abs", flush=True)
    assert L.fsa_mc_int_abs(vals.ctypes.data, out.ctypes.data, len(vals)) == 0
    ref = np.array([oracle_lib.x86_int_abs(float(v)) for v in vals], dtype=np.float32)
    assert np.array_equal(out, ref), (vals, out, ref)
