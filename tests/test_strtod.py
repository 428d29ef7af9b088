"""The text conversions of the GPU CSR reader (gooseberry_amd/csrc/hip/
gb_strtod.h) against the C library on the host.

The reference reads a column with strtoull(.., 10) and a value with strtod
(src/io.c:415-446), so the C library is the reference's own algorithm
here.  The header is compiled for the host (tests/native/strtod_check.cpp)
and every string below goes through both: value bits and end positions
must agree.  The cases cover every branch: random doubles in several
printed forms, random 64-bit mantissas over the whole exponent range,
exact halfway points between neighbouring doubles (and one digit either
side) that only the big-integer fallback settles, long digit strings,
hexadecimal forms, infinities, NaNs with payloads, and malformed text.
"""
import ctypes as C
import os
import random
import struct
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIP_DIR = os.path.join(ROOT, "gooseberry_amd", "csrc", "hip")


@pytest.fixture(scope="module")
def chk(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("strtod") / "strtod_check.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", HIP_DIR,
                    os.path.join(ROOT, "tests", "native", "strtod_check.cpp"), "-o", out],
                   check=True)
    lib = C.CDLL(out)
    lib.gbt_check.restype = C.c_long
    lib.gbt_check.argtypes = [C.c_char_p, C.c_long, C.POINTER(C.c_long), C.c_long,
                              C.POINTER(C.c_long)]
    lib.gbt_one.argtypes = [C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                            C.POINTER(C.c_int), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    return lib


def run(chk, strings):
    data = b"".join(s.encode("latin-1") + b"\0" for s in strings)
    bad = (C.c_long * 20)()
    nhard = C.c_long()
    nbad = chk.gbt_check(data, len(strings), bad, 20, C.byref(nhard))
    if nbad:
        libc = C.CDLL(None)
        libc.strtod.restype = C.c_double
        libc.strtod.argtypes = [C.c_char_p, C.POINTER(C.c_char_p)]
        lines = []
        for i in list(bad)[:min(nbad, 20)]:
            s = strings[i].encode("latin-1")
            v, e, h, u, ue = C.c_double(), C.c_uint64(), C.c_int(), C.c_uint64(), C.c_uint64()
            chk.gbt_one(s, C.byref(v), C.byref(e), C.byref(h), C.byref(u), C.byref(ue))
            buf = C.create_string_buffer(s)
            end = C.c_char_p()
            ref = libc.strtod(buf, C.byref(end))
            lines.append("%r: got %r end %d hard %d; libc %r end %d" % (
                strings[i][:80], v.value, e.value, h.value, ref,
                C.cast(end, C.c_void_p).value - C.addressof(buf)))
        pytest.fail("%d mismatches:\n%s" % (nbad, "\n".join(lines)))
    return nhard.value


def dbl(bits):
    return struct.unpack("<d", struct.pack("<Q", bits))[0]


def test_syntax_and_special_forms(chk):
    cases = [
        "", " ", "+", "-", ".", "e5", "1e", "1e+", "1e-5x", "  -0", "-0.0e-0", "0e", "0e+",
        "0e5", "00", "00x1", "0x", "0X", "0x.", "0x.p1", "0x0.p1", "0x0", "0X1P-1074",
        "0x1p-1075", "0x1p-1076", "0x1.8p-1074", "0x3p-1076", "0x1.fffffffffffffp1023",
        "0x1.fffffffffffff8p1023", "0x1p1024", "0x1p", "0x1p+", "0x1P-", "0x1.8p1", "0x.8p1",
        "0x8.p-3", "-0x1.000000000000080000001p0", "0x1.00000000000008p0",
        "0x1.00000000000018p0", "0x123456789abcdef0123p-20", "0x.000001p0",
        "inf", "-INF", "infinity", "infinit", "InFiNiTy", "infx", "in", "i",
        "nan", "-nan", "NaN", "nan()", "nan(123)", "nan(0x10)", "nan(010)", "nan(abc)",
        "nan(1.5)", "nan(", "nan(_)", "nan(0x)", "nan(0)", "nan(99999999999999999999999)",
        "-nan(0xfffffffffffff)", "na", "n",
        "1.5e9999999999999", "1e-9999999999999", "0e-9999999999999", "\t\n\v\f\r 42",
        "1_000", "1,5", "0.0000000000000000000000000000001e31", "+-1", "--1", "1..2",
        "1.2.3", "E5", ".5", "5.", "-.5e-3", "+.e1", "1e308", "1.8e308", "-1e309",
        "2.2250738585072011e-308", "2.2250738585072012e-308", "2.2250738585072013e-308",
        "4.9406564584124654e-324", "2.4703282292062327e-324", "2.4703282292062328e-324",
        "9007199254740993", "9007199254740992.5", "9007199254740993.0000000001",
        "1.7976931348623157e308", "1.7976931348623158e308", "1.7976931348623159e308",
        "18446744073709551615", "18446744073709551616", "-1", "-18446744073709551616",
        "-18446744073709551615", "+5", " 7x", "0x10", "0xg", "  ", "4294967296",
        "99999999999999999999999999999999", "0.1", "0.30000000000000004",
        "123456789012345678901234567890e-10", "1e23", "8.98846567431158e307",
        "7.2057594037927933e16", "1.00000000000000011102230246251565404236316680908203125",
        "1.00000000000000011102230246251565404236316680908203124",
        "1.00000000000000011102230246251565404236316680908203126",
    ]
    run(chk, cases)


def test_random_doubles_printed(chk):
    rng = random.Random(1)
    fmts = ["%r", "%.17g", "%.16e", "%.20e", "%.25g", "%g", "%.3e", "%.40e", "%.9g"]
    out = []
    for _ in range(60000):
        bits = rng.getrandbits(64)
        if rng.random() < 0.1:
            bits &= (1 << 52) - 1 | (1 << 63)        # subnormals
        x = dbl(bits)
        f = rng.choice(fmts)
        out.append(repr(x) if f == "%r" else f % x)
        if rng.random() < 0.2:
            out.append(x.hex())
    run(chk, out)


def test_random_mantissas_all_exponents(chk):
    rng = random.Random(2)
    out = []
    for _ in range(150000):
        w = rng.getrandbits(rng.randint(1, 64)) or 1
        q = rng.randint(-360, 330)
        out.append("%de%d" % (w, q))
    run(chk, out)


def test_halfway_points_need_the_exact_fallback(chk):
    getcontext().prec = 2000
    rng = random.Random(3)
    out = []
    for _ in range(4000):
        bits = rng.getrandbits(63)
        if rng.random() < 0.3:
            bits &= (1 << 52) - 1                    # subnormal neighbours
        if bits >= 0x7fefffffffffffff:
            continue
        a, b = Decimal(dbl(bits)), Decimal(dbl(bits + 1))
        h = (a + b) / 2                              # exact: a finite binary fraction
        s = format(h, "f") if rng.random() < 0.5 else format(h, "e")
        out.append(s)
        mant, _, exp = s.partition("e")
        out.append(mant + "1" + ("e" + exp if exp else "") if "." in mant else s + ".1")
        # one unit below in the last place
        digits = mant.rstrip("0")
        if digits.endswith("."):
            digits = digits[:-1]
        if digits and digits[-1] in "123456789":
            lower = digits[:-1] + str(int(digits[-1]) - 1) + "9" * 30
            out.append(lower + ("e" + exp if exp else ""))
    nhard = run(chk, out)
    assert nhard > 1000          # the fallback was exercised


def test_long_digit_strings(chk):
    rng = random.Random(4)
    out = []
    for _ in range(3000):
        nd = rng.randint(20, 1200)
        d = "".join(rng.choice("0123456789") for _ in range(nd))
        pt = rng.randint(0, nd)
        s = d[:pt] + "." + d[pt:]
        if rng.random() < 0.7:
            s += "e%d" % rng.randint(-400, 320)
        out.append(("-" if rng.random() < 0.3 else "") + s)
    # many leading zeros, and exact values of doubles written in full
    for _ in range(500):
        x = dbl(rng.getrandbits(63) & 0x7fefffffffffffff)
        out.append(format(Decimal(x), "f")[:5000])
        out.append("0." + "0" * rng.randint(300, 330) + "24703282292062327208828439643411068")
    run(chk, out)
