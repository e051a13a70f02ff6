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
    L.fsa_mc_neighbor_order.restype = C.c_int
    L.fsa_mc_neighbor_order.argtypes = [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p]
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


def test_neighbor_order_is_the_reference_key_order(oracle_lib):
    """The list kernel reads a cell's 27 neighbor runs in the order of a 6-bit code
    instead of sorting their keys (CompressedNeighbors.cpp:536-560 sorts the 27
    cell keys).  On the device, for every cell of key levels 1-3 and for the
    edges (0, -1, +-2^L, 2^L - 1, powers of two) and random cells up to level 9:
    the 27 codes decode to the 27 offsets one to one, the cell's own code to
    offset (0, 0, 0), and the neighbors inside [-2^L, 2^L) ordered by code are
    ordered by the oracle's libmorton key of c + 2^31."""
    import itertools

    import numpy as np

    L = _lib()
    rng = np.random.default_rng(11)
    offsets = np.array(list(itertools.product((-1, 0, 1), repeat=3)), dtype=np.int64)  # index (dx+1)*9+(dy+1)*3+dz+1
    checked = 0
    for lv in range(1, 10):
        lo, hi = -(1 << lv), 1 << lv
        if lv <= 3:
            cells = np.array(list(itertools.product(range(lo, hi), repeat=3)), dtype=np.int32)
        else:
            edge = sorted({lo, lo + 1, -2, -1, 0, 1, 2, hi - 2, hi - 1, 1 << (lv - 1), -(1 << (lv - 1)),
                           (1 << (lv - 1)) - 1, 3 << (lv - 2)})
            cells = np.array(list(itertools.product(edge, repeat=3)), dtype=np.int32)
            cells = np.concatenate([cells, rng.integers(lo, hi, size=(600, 3), dtype=np.int32)])
        n = len(cells)
        out = np.zeros((n, 64), dtype=np.uint8)
        self_code = np.zeros(n, dtype=np.uint8)
        print(f"case neighbor order, level {lv}, {n} cells", flush=True)
        assert L.fsa_mc_neighbor_order(np.ascontiguousarray(cells).ctypes.data, lv, n, out.ctypes.data,
                                       self_code.ctypes.data) == 0
        for i in range(n):
            codes = out[i]
            valid = codes != 255
            assert valid.sum() == 27 and sorted(codes[valid].tolist()) == list(range(27)), cells[i]
            assert codes[self_code[i]] == 13, cells[i]
            order = codes[valid]  # offset indices in code order
            nb = cells[i].astype(np.int64) + offsets[order]
            inside = np.all((nb >= lo) & (nb < hi), axis=1)
            keys = [oracle_lib.cell_key(int(x), int(y), int(z)) for x, y, z in nb[inside]]
            assert all(a < b for a, b in zip(keys, keys[1:])), (lv, cells[i])
            checked += 1
    assert checked > 15000
