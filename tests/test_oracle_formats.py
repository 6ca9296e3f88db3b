"""Pins the oracle's pixel-format conversions (ImageCompiler.cpp semantics)
with known answers, exhaustive code sweeps and an independent exact-rational
model of fp32 rounding."""
import ctypes as C
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib
from cpvulkan_amd import abi

F = abi.FORMAT


def pack_f32(fmt, rgba):
    out = (C.c_uint8 * 16)()
    src = (C.c_float * 4)(*rgba)
    assert oracle_lib.lib().cpvref_pack_f32(fmt, src, out) == 0
    return bytes(out)[: oracle_lib.lib().cpvref_format_size(fmt)]


def unpack_f32(fmt, data):
    buf = (C.c_uint8 * 16)(*data)
    out = (C.c_float * 4)()
    assert oracle_lib.lib().cpvref_unpack_f32(fmt, buf, out) == 0
    return list(out)


def pack_ds(fmt, depth, stencil):
    out = (C.c_uint8 * 8)()
    assert oracle_lib.lib().cpvref_pack_depth_stencil(fmt, depth, stencil, out) == 0
    return bytes(out)[: oracle_lib.lib().cpvref_format_size(fmt)]


def unpack_depth(fmt, data):
    buf = (C.c_uint8 * 8)(*data)
    return oracle_lib.lib().cpvref_unpack_depth(fmt, buf)


# --- independent model: fp32 round-to-nearest-even and round-half-away, exact

def fp32_round(q: Fraction) -> Fraction:
    """Round an exact rational to the nearest fp32 value (ties to even)."""
    if q == 0:
        return Fraction(0)
    sign = -1 if q < 0 else 1
    a = abs(q)
    e = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** e > a:
        e -= 1
    e = max(e, -126)
    ulp = Fraction(2) ** (e - 23)
    n, r = divmod(a, ulp)
    n = int(n)
    if r * 2 > ulp or (r * 2 == ulp and n % 2 == 1):
        n += 1
    return sign * n * ulp


def round_half_away(q: Fraction) -> int:
    n = int(abs(q) + Fraction(1, 2))
    return n if q >= 0 else -n


def unorm8_model(v: float) -> int:
    if v != v:
        return 0
    v = min(max(v, 0.0), 1.0)
    return round_half_away(fp32_round(Fraction(float(np.float32(v))) * 255))


def test_unorm8_known_answers():
    fmt = F["R8G8B8A8_UNORM"]
    assert pack_f32(fmt, [0.2, 0.0, 1.0, 0.5]) == bytes([51, 0, 255, 128])
    assert pack_f32(fmt, [-0.5, 1.5, float("nan"), float("inf")]) == bytes([0, 255, 0, 255])
    assert pack_f32(fmt, [float("-inf"), 0.5 / 255.0, 1.5 / 255.0, 254.5 / 255.0]) == bytes(
        [0, unorm8_model(0.5 / 255.0), unorm8_model(1.5 / 255.0), unorm8_model(254.5 / 255.0)])


def test_unorm8_channel_order():
    # B8G8R8A8: R at byte 2, G at 1, B at 0, A at 3 (Formats.cpp:255)
    assert pack_f32(F["B8G8R8A8_UNORM"], [1.0, 0.0, 0.5, 0.2]) == bytes([128, 0, 255, 51])
    assert unpack_f32(F["B8G8R8A8_UNORM"], [10, 20, 30, 40])[:3] == [np.float32(30) / np.float32(255),
                                                                      np.float32(20) / np.float32(255),
                                                                      np.float32(10) / np.float32(255)]


def test_unorm8_roundtrip_all_codes():
    fmt = F["R8_UNORM"]
    for k in range(256):
        v = unpack_f32(fmt, [k])[0]
        assert pack_f32(fmt, [v, 0, 0, 0]) == bytes([k])


def test_unorm8_pack_matches_exact_model():
    rng = np.random.default_rng(0)
    vals = list(rng.uniform(-0.1, 1.1, 3000).astype(np.float32))
    # neighbours of every half-way point (k + 1/2) / 255
    for k in range(255):
        mid = np.float32((k + 0.5) / 255.0)
        vals += [np.nextafter(mid, np.float32(-1)), mid, np.nextafter(mid, np.float32(2))]
    fmt = F["R8_UNORM"]
    for v in vals:
        assert pack_f32(fmt, [float(v), 0, 0, 0])[0] == unorm8_model(float(v)), v


def test_unorm16_roundtrip_all_codes():
    fmt = F["R16G16B16A16_UNORM"]
    codes = np.arange(65536, dtype=np.uint16)
    for k in codes[::7].tolist() + [0, 1, 32767, 32768, 65534, 65535]:
        data = list(np.array([k, 0, 0, 0], np.uint16).view(np.uint8))
        v = unpack_f32(fmt, data)[0]
        assert v == np.float32(k) / np.float32(65535)
        out = pack_f32(fmt, [v, 0.0, 0.0, 0.0])
        assert np.frombuffer(out, np.uint16)[0] == k


def test_snorm8():
    fmt = F["R8G8B8A8_SNORM"]
    assert pack_f32(fmt, [-1.0, 1.0, -2.0, 0.0]) == bytes([0x81, 0x7F, 0x81, 0x00])
    # reading -128 is not clamped: -128 / 127 (ImageCompiler.cpp:43-47)
    assert unpack_f32(fmt, [0x80, 0, 0, 0])[0] == np.float32(-128) / np.float32(127)
    for k in range(-127, 128):
        v = unpack_f32(fmt, [k & 0xFF, 0, 0, 0])[0]
        assert np.frombuffer(pack_f32(fmt, [v, 0, 0, 0]), np.int8)[0] == k


def test_integer_saturation():
    L = oracle_lib.lib()
    out = (C.c_uint8 * 16)()
    assert L.cpvref_pack_u32(F["R8G8B8A8_UINT"], (C.c_uint32 * 4)(300, 255, 254, 0xFFFFFFFF), out) == 0
    assert bytes(out)[:4] == bytes([255, 255, 254, 255])
    assert L.cpvref_pack_i32(F["R8G8B8A8_SINT"], (C.c_int32 * 4)(-200, 200, -5, 127), out) == 0
    assert np.frombuffer(bytes(out)[:4], np.int8).tolist() == [-128, 127, -5, 127]


def test_d16_all_codes_roundtrip():
    fmt = F["D16_UNORM"]
    for k in range(0, 65536):
        d = unpack_depth(fmt, list(np.array([k], np.uint16).view(np.uint8)))
        assert np.frombuffer(pack_ds(fmt, d, 0), np.uint16)[0] == k


def test_d16_known_answers():
    fmt = F["D16_UNORM"]
    enc = lambda d: int(np.frombuffer(pack_ds(fmt, d, 0), np.uint16)[0])
    assert enc(1.0) == 65535 and enc(0.0) == 0 and enc(-1.0) == 0 and enc(2.0) == 65535
    assert enc(0.5) == 32768          # 32767.5 rounds half away from zero
    assert enc(float("nan")) == 0     # maxnum(NaN, 0) = 0


def test_d24_and_stencil_placement():
    d24s8 = pack_ds(F["D24_UNORM_S8_UINT"], 1.0, 0xAB)
    assert np.frombuffer(d24s8, np.uint32)[0] == 0xAB_FFFFFF
    x8 = pack_ds(F["X8_D24_UNORM_PACK32"], 1.0, 0xAB)
    assert np.frombuffer(x8, np.uint32)[0] == 0x00_FFFFFF
    for k in [0, 1, 12345, 8388607, 8388608, 16777214, 16777215]:
        raw = list(np.array([k | (0x5A << 24)], np.uint32).view(np.uint8))
        d = unpack_depth(F["D24_UNORM_S8_UINT"], raw)
        assert d == np.float32(k) / np.float32(16777215)
        assert np.frombuffer(pack_ds(F["D24_UNORM_S8_UINT"], d, 0x5A), np.uint32)[0] == k | (0x5A << 24)
    assert pack_ds(F["D32_SFLOAT_S8_UINT"], 0.25, 7)[4] == 7
    assert pack_ds(F["D16_UNORM_S8_UINT"], 0.0, 9)[2] == 9
    assert pack_ds(F["S8_UINT"], 0.5, 200) == bytes([200])


def test_d32_is_raw():
    fmt = F["D32_SFLOAT"]
    for v in [-0.0, 1.5, -3.0, float("inf"), 1e-40]:
        assert pack_ds(fmt, v, 0) == np.array([v], np.float32).tobytes()
    nan = np.array([0x7FC01234], np.uint32).view(np.float32)[0]
    assert np.frombuffer(pack_ds(fmt, float(nan), 0), np.uint32)[0] & 0x7FC00000 == 0x7FC00000


def test_image_layout_closed_form():
    L = oracle_lib.lib()
    offs = (C.c_uint64 * 16)()
    assert L.cpvref_image_layout(F["R8G8B8A8_UNORM"], 1024, 1024, 11, offs) == 5_592_404
    assert [offs[i] for i in range(3)] == [0, 4 * 1024 * 1024, 4 * (1024 * 1024 + 512 * 512)]
    # non-square: 8x2, 4x1, 2x1, 1x1 pixels
    assert L.cpvref_image_layout(F["R8G8B8A8_UNORM"], 8, 2, 4, offs) == 4 * (16 + 4 + 2 + 1)


@pytest.mark.parametrize("maxv", [255, 65535, 16777215])
def test_unorm_double_multiply_identity(maxv):
    """The device reads UNORM codes as (float)((double)u * (1.0 / max)); that
    equals the reference's fp32 division u / (float)max for every code."""
    u = np.arange(maxv + 1, dtype=np.uint64)
    a = u.astype(np.float32) / np.float32(maxv)
    b = (u.astype(np.float64) * (1.0 / maxv)).astype(np.float32)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
