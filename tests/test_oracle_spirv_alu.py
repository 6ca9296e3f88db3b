"""Pins the shader-operation oracle (oracle/spirv_alu.py) against
independent references: Python integers and bit strings for the integer
and bit operations, struct's IEEE half codec, math.fmod and floor-modulo on
exact values, exact small-integer matrices against numpy's matmul (which
catches transposed operands), geometric identities, and the reference's
stated quirks (scalar Normalize, BitFieldInsert not shifting Insert)."""
from __future__ import annotations

import math
import struct
import sys
from pathlib import Path

import numpy as np
import pytest

sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "oracle"))
import spirv_alu as O  # noqa: E402

RNG = np.random.Generator(np.random.PCG64(77))


def words(n=400):
    w = RNG.integers(0, 2**32, size=n, dtype=np.uint64)
    w[:8] = [0, 1, 2, 0x80000000, 0xFFFFFFFF, 0x7FFFFFFF, 0x00010000, 0x80000001]
    return w.astype(np.uint32)


def as_vec(x):
    """(n,) uint32 words -> (n, 4) float32 carrying them in every lane"""
    return np.repeat(x.view(np.float32)[:, None], 4, axis=1)


def col(res):
    return res.view(np.uint32)[:, 0]


def test_bit_scans_match_python_integers():
    w = words()
    s = w.view(np.int32)
    lsb = col(O.evaluate("FindILsb", as_vec(w), as_vec(w), as_vec(w), as_vec(w))).view(np.int32)
    umsb = col(O.evaluate("FindUMsb", as_vec(w), as_vec(w), as_vec(w), as_vec(w))).view(np.int32)
    smsb = col(O.evaluate("FindSMsb", as_vec(w), as_vec(w), as_vec(w), as_vec(w))).view(np.int32)
    for k in range(len(w)):
        x = int(w[k])
        assert lsb[k] == ((x & -x).bit_length() - 1 if x else -1)
        assert umsb[k] == x.bit_length() - 1
        v = int(s[k])
        v = ~v if v < 0 else v
        assert smsb[k] == v.bit_length() - 1


def test_bit_reverse_and_count_match_bit_strings():
    w = words()
    rev = col(O.evaluate("BitReverse", as_vec(w), as_vec(w), as_vec(w), as_vec(w)))
    cnt = col(O.evaluate("BitCount", as_vec(w), as_vec(w), as_vec(w), as_vec(w)))
    for k in range(len(w)):
        bits = format(int(w[k]), "032b")
        assert int(rev[k]) == int(bits[::-1], 2)
        assert int(cnt[k]) == bits.count("1")


def test_integer_arithmetic_matches_python():
    a, b = words(), words()
    b[b == 0] = 7
    sa, sb = a.view(np.int32).astype(int), b.view(np.int32).astype(int)
    keep = ~((sa == -2**31) & (sb == -1))
    a, b, sa, sb = a[keep], b[keep], sa[keep], sb[keep]
    va, vb = as_vec(a), as_vec(b)
    got = {n: col(O.evaluate(n, va, vb, va, va)) for n in ("IAdd", "ISub", "IMul", "SDiv", "SRem", "SMod", "UDiv", "UMod")}
    for k in range(len(a)):
        x, y, ux, uy = int(sa[k]), int(sb[k]), int(a[k]), int(b[k])
        q = abs(x) // abs(y) * (1 if (x < 0) == (y < 0) else -1)     # C truncation
        assert got["IAdd"][k] == (ux + uy) % 2**32
        assert got["ISub"][k] == (ux - uy) % 2**32
        assert got["IMul"][k] == (ux * uy) % 2**32
        assert got["SDiv"][k] == q % 2**32
        assert got["SRem"][k] == (x - q * y) % 2**32
        assert got["SRem"][k] == int(math.fmod(x, y)) % 2**32
        # the reference lowers SMod as a - b * (a / b): the truncating remainder
        assert got["SMod"][k] == got["SRem"][k]
        assert got["UDiv"][k] == ux // uy
        assert got["UMod"][k] == ux % uy


def test_bitfield_insert_follows_the_reference_mask():
    """Bit i of the result is Insert's bit i inside [offset, offset + count)
    and Base's bit i outside (SPIRVCompiler.cpp:3099-3136 does not shift
    Insert; the SPIR-V specification would take Insert's bit i - offset)."""
    base, ins, oc = words(200), words(200), words(200)
    r = col(O.evaluate("BitFieldInsert", as_vec(base), as_vec(ins), as_vec(oc), as_vec(oc)))
    off = oc.view(np.int32) & 15
    # the case reads offset from c.x and count from c.y: both lanes hold the same word
    cnt = oc.view(np.int32) & 15
    for k in range(200):
        want = 0
        for i in range(32):
            src = ins if off[k] <= i < off[k] + cnt[k] else base
            want |= ((int(src[k]) >> i) & 1) << i
        assert int(r[k]) == want


def test_bitfield_extract_follows_the_specification():
    base, oc = words(200), words(200)
    ru = col(O.evaluate("BitFieldUExtract", as_vec(base), as_vec(base), as_vec(oc), as_vec(oc)))
    rs = col(O.evaluate("BitFieldSExtract", as_vec(base), as_vec(base), as_vec(oc), as_vec(oc))).view(np.int32)
    for k in range(200):
        off = int(oc.view(np.int32)[k]) & 15
        cnt = off            # c.x and c.y carry the same word here
        field = (int(base[k]) >> off) & ((1 << cnt) - 1)
        assert int(ru[k]) == field
        signed = field - (1 << cnt) if cnt and field >> (cnt - 1) else field
        assert int(rs[k]) == signed


def test_shifts_and_min_max_match_python():
    a, b = words(), words() % 32
    va, vb = as_vec(a), as_vec(b)
    sl, sr, sa = (col(O.evaluate(n, va, vb, va, va)) for n in ("ShiftLeftLogical", "ShiftRightLogical", "ShiftRightArithmetic"))
    s = a.view(np.int32)
    for k in range(len(a)):
        assert int(sl[k]) == (int(a[k]) << int(b[k])) % 2**32
        assert int(sr[k]) == int(a[k]) >> int(b[k])
        assert int(sa[k]) == (int(s[k]) >> int(b[k])) % 2**32
    x, y = words(), words()
    vx, vy = as_vec(x), as_vec(y)
    assert np.array_equal(col(O.evaluate("UMin", vx, vy, vx, vx)), np.minimum(x, y))
    assert np.array_equal(col(O.evaluate("SMax", vx, vy, vx, vx)).view(np.int32), np.maximum(x.view(np.int32), y.view(np.int32)))


def test_std_min_max_clamp_and_nan_rules():
    nan = np.float32("nan")
    one, two = np.float32(1), np.float32(2)
    # std::min(a, b) = b < a ? b : a: a NaN second operand loses, a NaN first operand wins
    assert O.std_min(one, nan) == one and np.isnan(O.std_min(nan, one))
    assert O.std_max(one, nan) == one and np.isnan(O.std_max(nan, one))
    # NMin / NMax return the other operand for a NaN first operand
    v = lambda *x: np.array([x], np.float32)
    assert O.evaluate("NMin", v(nan, 1, 2, nan), v(3, nan, 1, nan), v(0, 0, 0, 0), v(0, 0, 0, 0))[0].tolist()[:3] == [3.0, 1.0, 1.0]
    assert O.evaluate("FClamp", v(5, -5, 0.5, 2), v(0, 0, 0, 3), v(1, 1, 1, 1), v(0, 0, 0, 0))[0].tolist() == [1.0, 0.0, 0.5, 3.0]
    # clamp(x, lo, hi) with lo > hi: x < lo gives lo, else hi < x gives hi
    assert O.std_clamp(two, np.float32(3), one) == np.float32(3)


def test_fmod_and_frem():
    rng = np.random.Generator(np.random.PCG64(3))
    # exact integer-valued operands: the divisor-signed modulo of Python floats
    a = rng.integers(-1000, 1000, size=(300, 4)).astype(np.float32)
    b = rng.integers(1, 50, size=(300, 4)).astype(np.float32) * rng.choice([-1, 1], size=(300, 4)).astype(np.float32)
    got = O.evaluate("FMod", a, b, a, a)
    rem = O.evaluate("FRem", a, b, a, a)
    for (x, y, g, r) in zip(a.ravel(), b.ravel(), got.ravel(), rem.ravel()):
        assert g == float(x) % float(y)
        assert r == math.fmod(float(x), float(y))
    # non-integers: FRem is math.fmod exactly; FMod takes the divisor's sign
    a = rng.uniform(-100, 100, size=(300, 4)).astype(np.float32)
    b = rng.uniform(0.5, 10, size=(300, 4)).astype(np.float32) * rng.choice([-1, 1], size=(300, 4)).astype(np.float32)
    rem = O.evaluate("FRem", a, b, a, a)
    got = O.evaluate("FMod", a, b, a, a)
    for (x, y, g, r) in zip(a.ravel(), b.ravel(), got.ravel(), rem.ravel()):
        assert r == np.float32(math.fmod(float(x), float(y)))
        assert g == 0 or math.copysign(1.0, g) == math.copysign(1.0, y)
        assert abs(g) <= abs(y)


def test_rounding_functions():
    x = np.array([[0.5, 1.5, 2.5, -0.5], [-1.5, -2.5, 0.49999997, 8388609.0], [1e30, -0.0, 3.7, -3.7]], np.float32)
    assert O.evaluate("Round", x, x, x, x).tolist() == [[1, 2, 3, -1], [-2, -3, 0, 8388609], [float(np.float32(1e30)), -0.0, 4, -4]]
    assert O.evaluate("RoundEven", x, x, x, x).tolist()[0] == [0, 2, 2, -0.0]
    assert O.evaluate("Fract", x, x, x, x)[2, 2] == np.float32(3.7) - np.float32(3.0)


def test_half_and_unorm_packing_match_struct():
    rng = np.random.Generator(np.random.PCG64(9))
    a = rng.uniform(-2, 2, size=(200, 4)).astype(np.float32)
    p = O.evaluate("Pack", a, a, a, a).view(np.uint32)
    for k in range(200):
        h0 = struct.unpack("<H", struct.pack("<e", float(a[k, 0])))[0]
        h1 = struct.unpack("<H", struct.pack("<e", float(a[k, 1])))[0]
        assert int(p[k, 2]) == h0 | (h1 << 16)
        un = [min(max(float(v), 0.0), 1.0) * 255.0 for v in a[k]]
        want = sum(int(math.floor(np.float32(v) + 0.5)) << (8 * i) for i, v in enumerate(un))
        assert int(p[k, 0]) == want
    w = RNG.integers(0, 2**32, size=(200, 4), dtype=np.uint64).astype(np.uint32) & np.uint32(0xBBFFBBFF)
    u = O.evaluate("Unpack2x16", w.view(np.float32), w.view(np.float32), w.view(np.float32), w.view(np.float32))
    for k in range(200):
        assert u[k, 0] == struct.unpack("<e", struct.pack("<H", int(w[k, 0]) & 0xFFFF))[0]
        assert u[k, 1] == struct.unpack("<e", struct.pack("<H", int(w[k, 0]) >> 16))[0]


def test_quantize_to_f16_matches_struct():
    x = np.random.Generator(np.random.PCG64(4)).uniform(-70000, 70000, size=(100, 4)).astype(np.float32)
    q = O.evaluate("QuantizeToF16", x, x, x, x)
    for v, g in zip(x.ravel(), q.ravel()):
        try:
            want = struct.unpack("<e", struct.pack("<e", float(v)))[0]
        except OverflowError:
            want = math.copysign(math.inf, float(v))
        assert g == want


def test_matrix_products_against_numpy_matmul():
    """Small integer entries make every product and sum exact, so the order
    of summation cannot matter and only the operand layout is checked."""
    rng = np.random.Generator(np.random.PCG64(5))
    a, b, c, d = (rng.integers(-9, 10, size=(50, 4)).astype(np.float32) for _ in range(4))
    for k in range(50):
        M = np.stack([a[k], b[k], c[k], d[k]], axis=1).astype(np.float64)          # columns a, b, c, d
        N = np.stack([d[k], c[k], b[k], a[k]], axis=1).astype(np.float64)
        one = lambda n: O.evaluate(n, a[k:k + 1], b[k:k + 1], c[k:k + 1], d[k:k + 1])[0]
        assert np.array_equal(one("MatrixTimesVector4"), M @ a[k])
        assert np.array_equal(one("VectorTimesMatrix4"), a[k] @ M)
        assert np.array_equal(one("MatrixTimesMatrix4"), (M @ N)[:, 1])
        assert np.array_equal(one("Transpose4"), M.T[:, 0])
        assert np.array_equal(one("OuterProduct4"), np.outer(a[k], b[k])[:, 3])
        assert np.array_equal(one("MatrixTimesVector3")[:3], M[:3, :3] @ d[k][:3])
        assert np.array_equal(one("Dot")[:3], [a[k][:2] @ b[k][:2], a[k][:3] @ b[k][:3], a[k] @ b[k]])


def test_geometric_identities():
    v = lambda *x: np.array([x], np.float32)
    assert O.evaluate("Length", v(3, 4, 0, 0), v(0, 0, 0, 0), v(0, 0, 0, 0), v(0, 0, 0, 0))[0].tolist() == [5, 5, 5, 3]
    cr = O.evaluate("Cross", v(1, 0, 0, 0), v(0, 1, 0, 0), v(0, 0, 0, 0), v(0, 0, 0, 0))[0]
    assert cr.tolist() == [0, 0, 1, 0]
    rng = np.random.Generator(np.random.PCG64(6))
    a, b = rng.uniform(-3, 3, size=(100, 4)).astype(np.float32), rng.uniform(-3, 3, size=(100, 4)).astype(np.float32)
    cr = O.evaluate("Cross", a, b, a, a)[:, :3].astype(np.float64)
    assert np.allclose((cr * a[:, :3]).sum(1), 0, atol=1e-4) and np.allclose((cr * b[:, :3]).sum(1), 0, atol=1e-4)
    n = O.evaluate("Normalize", a, b, a, a).astype(np.float64)
    assert np.allclose((n * n).sum(1), 1.0, atol=1e-6)
    # reflecting twice about a unit normal returns the vector
    nrm = n.astype(np.float32)
    r2 = O.evaluate("Reflect", O.evaluate("Reflect", a, nrm, a, a), nrm, a, a)
    assert np.allclose(r2, a, atol=1e-4)
    # refract with eta = 1 passes straight through when the normal faces the ray
    i3 = np.array([[0.3, -0.8, 0.1, 0]], np.float32)
    eta = np.array([[1, 0, 0, 0]], np.float32)
    r = O.evaluate("Refract", i3, np.array([[0, 1, 0, 0]], np.float32), eta, eta)[0]
    assert np.allclose(r[:3], i3[0, :3], atol=1e-6)
    # the reference's scalar Normalise: +1 for negative, -1 otherwise (GlslFunctions.cpp:162-166)
    assert O.evaluate("NormalizeScalar", v(-2, 3, -0.0, 0.0), v(0, 0, 0, 0), v(0, 0, 0, 0), v(0, 0, 0, 0))[0].tolist() == [1, -1, 1, -1]


def test_ulp_distance():
    x = np.array([1.0, -1.0, 0.0, np.nan], np.float32)
    y = np.nextafter(x, np.float32(np.inf)).astype(np.float32)
    y[3] = np.nan
    assert O.ulp_distance(x, y).tolist() == [1, 1, 1, 0]
    assert O.ulp_distance(np.array([-0.0], np.float32), np.array([0.0], np.float32)).tolist() == [0]
