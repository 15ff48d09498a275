"""The walk's unscaled correctly rounded square root and quotients
(dmipy-sim_amd/csrc/rn_arith.h) against the compiler's IEEE sqrtf and '/' on the
GPU, bit for bit: every float in the square root's range, and 2^31 + 2^31 random
operand pairs for each quotient form (tests/gpu_selftest/arith_check.hip)."""
import ctypes as C
import os

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "gpu_selftest", "libarith_check.so")


@pytest.fixture(scope="module")
def arith(gpu):
    if not os.path.exists(LIB):
        pytest.fail("tests/gpu_selftest/libarith_check.so is not built (run build())")
    lib = C.CDLL(LIB)
    lib.arith_check_sqrt.argtypes = [C.c_uint32, C.c_uint32, C.POINTER(C.c_ulonglong),
                                     C.POINTER(C.c_uint32)]
    lib.arith_check_div.argtypes = [C.c_uint64, C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_ulonglong),
                                    C.POINTER(C.c_ulonglong), C.POINTER(C.c_uint32)]
    return lib


def test_sqrt_rn_exhaustive(arith):
    """+0 and every finite float >= 2^-96."""
    for lo, hi in ((0, 1), (0x0F800000, 0x7F800000)):
        bad, first = C.c_ulonglong(0), (C.c_uint32 * 2)()
        print(f"sqrt check [{lo:#x}, {hi:#x})", flush=True)
        assert arith.arith_check_sqrt(lo, hi, C.byref(bad), first) == 0
        assert bad.value == 0, f"{bad.value} mismatches, first at bits {first[0]:#x}"


@pytest.mark.parametrize("known", [0, 1], ids=["div_rn", "div_rn_r"])
@pytest.mark.parametrize("emax,seed", [(60, 1), (20, 2)])
def test_div_rn_random(arith, emax, seed, known):
    """Random signs, mantissas (one divisor in 16 with an all-ones mantissa, one
    numerator in 64 zero) and exponents in [-emax, emax]; div_rn_r with the divisor's
    correctly rounded reciprocal."""
    bad, skipped, first = C.c_ulonglong(0), C.c_ulonglong(0), (C.c_uint32 * 2)()
    print(f"div check emax={emax} seed={seed} known_divisor={known}", flush=True)
    assert arith.arith_check_div(1 << 31, seed, emax, known, C.byref(bad), C.byref(skipped), first) == 0
    assert bad.value == 0, (f"{bad.value} mismatches, first x={first[0]:#x} y={first[1]:#x}")
    assert skipped.value < (1 << 31) // 20     # quotients beyond 2^+-100 are outside the contract
