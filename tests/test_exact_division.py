"""The smoothing step of the differentiable tape divides each triangle's
corner sum by 3 with div3() in binarymeshfitting_amd/csrc/bmf_diff.hip:
q0 = x RN(1/3), one fma for the exact remainder x - 3 q0 and one fma for
q0 + r RN(1/3) (Markstein's correction step), zeros and infinities taking q0.
Bit-exact parity with the reference's FP32 smoothing needs that to equal the
IEEE quotient x / 3 for every x.  The host restatement in
tests/native/exact_division.c runs the same FP32 operations (fmaf is the
IEEE fused multiply-add, as v_fma_f32 is) over every one of the 2^32 floats.

The short form is not exact for every divisor: for 6 it rounds some
subnormal quotients wrongly, which the same scan finds, so the division by
the ring size stays an IEEE division.
"""
import ctypes
import shutil
import subprocess
from pathlib import Path

import pytest

SRC = Path(__file__).resolve().parent / "native" / "exact_division.c"


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    so = tmp_path_factory.mktemp("exdiv") / "libexdiv.so"
    subprocess.run([cc, "-O3", "-march=x86-64-v3", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-fno-math-errno", "-shared",
                    "-fPIC", "-o", str(so), str(SRC), "-lm"], check=True)
    h = ctypes.CDLL(str(so))
    h.exact_division_mismatches.restype = ctypes.c_long
    h.exact_division_mismatches.argtypes = [ctypes.c_int, ctypes.c_uint32, ctypes.c_uint32]
    return h


def test_every_float_divided_by_three(lib):
    assert lib.exact_division_mismatches(3, 0, 0x80000000) == 0  # +0 ... +inf, NaNs; negatives by symmetry


def test_the_scan_finds_a_divisor_where_the_short_form_fails(lib):
    bad = lib.exact_division_mismatches(6, 0, (127 - 100) << 23)  # |x| < 2^-100
    assert bad > 0
    assert lib.exact_division_mismatches(6, 127 << 23, 128 << 23) == 0  # [1, 2): normal quotients agree
