"""CPU oracle of the shader operations the SPIR-V translator emits (test
infrastructure only: the GPU tests compare the translated conformance
shader, cpvulkan_amd.spirv.fs_alu_ops, against it; nothing in the product
imports it).

Each case evaluates one operation on per-fragment inputs a, b, c, d
(float32 arrays of shape (n, 4)) and returns the fragment's vec4 result as
float32 (integer results as their bits viewed as float32), restating the
code the reference runs for that operation:

  * GLSL.std.450 functions the reference registers (CPVulkan/GlslFunctions.cpp:
    Abs :19-35, SSign :37-52, Sin/Cos :54-84, Pow :86-100, Min/Max/Clamp
    :102-142, Mix :144-160, Normalise :162-172, Reflect :174-178,
    FindILsb/FindSMsb/FindUMsb :180-250, NMin/NMax/NClamp :252-300;
    registrations :930-1300),
  * core instructions lowered in LLVMRuntime/SPIRVCompiler.cpp (integer
    arithmetic :2620-2710, FRem/FMod :2706-2722, Any/All/IsNan :2770-2810,
    compares :2858-2990, shifts :2991-3091, BitFieldInsert :3099-3136,
    BitReverse/BitCount :3145-3155),
  * vector and matrix products through glm 0.9.5.3
    (LLVMRuntime/SpirvFunctions.cpp:5-81; Samples/utils/glm/detail/
    func_geometric.inl:56-83 dot, :88-172 length/distance, :217-229 cross,
    :233-280 normalize, :282-338 faceforward/reflect/refract;
    type_mat4x4.inl:651-711, 753-775; type_mat3x3.inl:587-611),
  * and, for functions the reference does not register (Round, Trunc, FSign,
    Floor, Ceil, Fract, the other transcendental functions, Step,
    SmoothStep, Fma, Ldexp, packing, BitFieldExtract, conversions), the
    GLSL.std.450 / SPIR-V specification: the reference cannot run shaders
    that use them, so those cases are "parity unpinned" against it and
    pinned only by the specification and the checks in
    tests/test_oracle_spirv_alu.py.

Every float32 operation is done one at a time in numpy float32 (each
rounds once, as the translated code compiled without contraction does),
in the order the restated code evaluates it.  Transcendental functions are
evaluated in float64 and rounded once; their results are compared within
TOLERANCE_ULPS, the accuracy the device library and the host libm each
promise, not bit for bit.
"""
from __future__ import annotations

import numpy as np

F32 = np.float32
ONE = F32(1.0)
TWO = F32(2.0)
ZERO = F32(0.0)

# ulps allowed against the correctly rounded value (0 = bit exact)
TOLERANCE_ULPS = dict(Sin=4, Cos=4, Tan=6, Asin=4, Acos=4, Atan=4, Sinh=4, Cosh=4, Tanh=4, Asinh=4, Acosh=4,
                      Atanh=4, Atan2=6, Pow=8, Exp=4, Log=4, Exp2=4, Log2=4, Fma=1)


def _f(x):
    return np.asarray(x, dtype=np.float32)


def _u(x):
    return np.ascontiguousarray(x, dtype=np.float32).view(np.uint32)


def _i(x):
    return np.ascontiguousarray(x, dtype=np.float32).view(np.int32)


def _as_f(bits):
    return np.ascontiguousarray(bits).astype(np.uint32).view(np.float32)


def _math64(fn):
    def g(*args):
        with np.errstate(all="ignore"):
            return fn(*[np.asarray(x, np.float64) for x in args]).astype(np.float32)
    return g


# ------------------------------------------------------------ scalar helpers

def std_min(a, b):
    """std::min(a, b) = (b < a) ? b : a"""
    return np.where(b < a, b, a)


def std_max(a, b):
    """std::max(a, b) = (a < b) ? b : a"""
    return np.where(a < b, b, a)


def std_clamp(v, lo, hi):
    """std::clamp(v, lo, hi) = v < lo ? lo : (hi < v ? hi : v)"""
    return np.where(v < lo, lo, np.where(hi < v, hi, v))


def round_half_away(x):
    """roundf: nearest integer, halves away from zero"""
    x = _f(x)
    t = np.trunc(x)
    with np.errstate(invalid="ignore"):
        frac = np.abs(x - t)
        return np.where(frac >= F32(0.5), t + np.sign(x), t).astype(np.float32)


def dot_glm(a, b):
    """glm::dot: 2 and 3 components left to right, 4 components (x + y) + (z + w)"""
    p = a * b
    n = a.shape[1]
    if n == 4:
        return (p[:, 0] + p[:, 1]) + (p[:, 2] + p[:, 3])
    s = p[:, 0]
    for k in range(1, n):
        s = s + p[:, k]
    return s


def sum_squares(v):
    """x.x * x.x + x.y * x.y + ..., left to right (glm length / normalize)"""
    s = v[:, 0] * v[:, 0]
    for k in range(1, v.shape[1]):
        s = s + v[:, k] * v[:, k]
    return s


def fmod_spirv(a, b):
    """OpFMod as the reference lowers it: r = frem(a, b); s = copysign(r, b);
    (r one s) ? r + b : s (SPIRVCompiler.cpp:2712-2722)"""
    with np.errstate(all="ignore"):
        r = np.fmod(a, b)
        s = np.copysign(r, b)
        return np.where((r < s) | (r > s), r + b, s).astype(np.float32)


def find_lsb(u):
    u = u.astype(np.uint64)
    out = np.full(u.shape, -1, np.int64)
    nz = u != 0
    low = u & (~u + np.uint64(1)) & np.uint64(0xFFFFFFFF)
    out[nz] = np.log2(low[nz].astype(np.float64)).astype(np.int64)
    return out.astype(np.int32)


def find_umsb(u):
    u = u.astype(np.uint64)
    out = np.full(u.shape, -1, np.int64)
    nz = u != 0
    out[nz] = np.floor(np.log2(u[nz].astype(np.float64))).astype(np.int64)
    # log2 of values just below a power of two can round up in float64
    fix = nz & (np.left_shift(np.uint64(1), np.maximum(out, 0).astype(np.uint64)) > u)
    out[fix] -= 1
    return out.astype(np.int32)


def find_smsb(i):
    i = i.astype(np.int64)
    v = np.where(i < 0, ~i, i)
    return find_umsb(v.astype(np.uint64))


def bit_reverse(u):
    u = u.astype(np.uint64)
    r = np.zeros_like(u)
    for k in range(32):
        r |= ((u >> np.uint64(k)) & np.uint64(1)) << np.uint64(31 - k)
    return r.astype(np.uint32)


def bit_count(u):
    u = u.astype(np.uint64)
    c = np.zeros_like(u)
    for k in range(32):
        c += (u >> np.uint64(k)) & np.uint64(1)
    return c.astype(np.uint32)


def bfi_reference(base, ins, offset, count):
    """mask = ((~0 << count) ^ ~0) << offset; (~mask & base) | (mask & insert)
    (SPIRVCompiler.cpp:3099-3136; Insert is not shifted)"""
    ones = np.uint64(0xFFFFFFFF)
    mask = ((ones << (count.astype(np.uint64) & np.uint64(31))) & ones) ^ ones
    mask = (mask << (offset.astype(np.uint64) & np.uint64(31))) & ones
    r = ((~mask & ones) & base.astype(np.uint64)) | (mask & ins.astype(np.uint64))
    return r.astype(np.uint32)


def bfe_u(base, offset, count):
    b = base.astype(np.uint64) >> (offset.astype(np.uint64) & np.uint64(31))
    m = (np.left_shift(np.uint64(1), count.astype(np.uint64)) - np.uint64(1)) & np.uint64(0xFFFFFFFF)
    return np.where(count == 0, 0, b & m).astype(np.uint32)


def bfe_s(base, offset, count):
    v = bfe_u(base, offset, count).astype(np.int64)
    sign = np.left_shift(np.int64(1), np.maximum(count.astype(np.int64) - 1, 0))
    v = np.where((count > 0) & (v & sign != 0), v - 2 * sign, v)
    return np.where(count == 0, 0, v).astype(np.int32)


def pack_unorm(v, scale):
    return round_half_away(std_clamp(v, ZERO, ONE) * F32(scale)).astype(np.uint32)


def pack_snorm(v, scale, bits):
    r = round_half_away(std_clamp(v, F32(-1.0), ONE) * F32(scale)).astype(np.int64)
    return (r & ((1 << bits) - 1)).astype(np.uint32)


def half_bits(x):
    return _f(x).astype(np.float16).view(np.uint16).astype(np.uint32)


def half_value(bits16):
    return bits16.astype(np.uint16).view(np.float16).astype(np.float32)


# ------------------------------------------------------------------- cases

def _vec(*cols):
    return np.stack([np.broadcast_to(_f(c), cols[0].shape) for c in cols], axis=1).astype(np.float32)


def _mat_vec4(cols, v):
    """glm mat4 * vec4: (m0 v0 + m1 v1) + (m2 v2 + m3 v3)"""
    return (cols[0] * v[:, 0:1] + cols[1] * v[:, 1:2]) + (cols[2] * v[:, 2:3] + cols[3] * v[:, 3:4])


def _mat_vec_ltr(cols, v):
    s = cols[0] * v[:, 0:1]
    for k in range(1, len(cols)):
        s = s + cols[k] * v[:, k:k + 1]
    return s


def _pack01(*bools):
    """dot(select(v, 1, 0), (1, 2, 4, 8)) per bvec4: (v0 + 2 v1) + (4 v2 + 8 v3)"""
    out = []
    for v in bools:
        f = v.astype(np.float32)
        out.append((f[:, 0] * ONE + f[:, 1] * TWO) + (f[:, 2] * F32(4.0) + f[:, 3] * F32(8.0)))
    return out


def evaluate(name: str, a, b, c, d) -> np.ndarray:
    """The vec4 result of conformance case `name` for each fragment."""
    a, b, c, d = _f(a), _f(b), _f(c), _f(d)
    n = a.shape[0]
    z = np.zeros(n, np.float32)
    with np.errstate(all="ignore"):
        simple1 = {
            "FAbs": np.abs, "Floor": np.floor, "Ceil": np.ceil, "Trunc": np.trunc, "Round": round_half_away,
            "RoundEven": np.rint, "Fract": lambda x: x - np.floor(x),
            "FSign": lambda x: np.where(x > 0, ONE, np.where(x < 0, F32(-1.0), x)),
            "Sin": _math64(np.sin), "Cos": _math64(np.cos), "Tan": _math64(np.tan), "Asin": _math64(np.arcsin),
            "Acos": _math64(np.arccos), "Atan": _math64(np.arctan), "Sinh": _math64(np.sinh), "Cosh": _math64(np.cosh),
            "Tanh": _math64(np.tanh), "Asinh": _math64(np.arcsinh), "Acosh": _math64(np.arccosh),
            "Atanh": _math64(np.arctanh), "Exp": _math64(np.exp), "Log": _math64(np.log), "Exp2": _math64(np.exp2),
            "Log2": _math64(np.log2), "Sqrt": np.sqrt, "InverseSqrt": lambda x: ONE / np.sqrt(x),
            "Radians": lambda x: x * F32(0.01745329251994329577), "Degrees": lambda x: x * F32(57.2957795130823208768),
            "FNegate": np.negative, "QuantizeToF16": lambda x: x.astype(np.float16).astype(np.float32),
        }
        if name in simple1:
            return _f(simple1[name](a))
        if name == "Normalize":
            inv = ONE / np.sqrt(sum_squares(a))
            return a * inv[:, None]
        if name == "NormalizeScalar":
            # the reference's scalar Normalise: signbit(v) ? 1 : -1 (GlslFunctions.cpp:162-166)
            return np.where(np.signbit(a), ONE, F32(-1.0)).astype(np.float32)
        two = {
            "FMin": std_min, "FMax": std_max,
            "NMin": lambda x, y: np.where(np.isnan(x), y, std_min(x, y)),
            "NMax": lambda x, y: np.where(np.isnan(x), y, std_max(x, y)),
            "Step": lambda e, x: np.where(x < e, ZERO, ONE), "Atan2": _math64(np.arctan2), "Pow": _math64(np.power),
            "FAdd": np.add, "FSub": np.subtract, "FMul": np.multiply, "FDiv": np.divide, "FRem": np.fmod,
            "FMod": fmod_spirv,
        }
        if name in two:
            return _f(two[name](a, b))
        if name == "Reflect":
            dd = dot_glm(b, a)[:, None]
            return a - b * dd * TWO
        if name == "FClamp":
            return std_clamp(a, b, c)
        if name == "NClamp":
            nmax = np.where(np.isnan(a), b, std_max(a, b))
            return np.where(np.isnan(nmax), c, std_min(nmax, c))
        if name == "FMix":
            return a * (ONE - c) + b * c
        if name == "SmoothStep":
            t = std_clamp((c - a) / (b - a), ZERO, ONE)
            return t * t * (F32(3.0) - TWO * t)
        if name == "Fma":
            return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)
        if name == "FaceForward":
            neg = (dot_glm(c, b) < 0)[:, None]
            return np.where(neg, a, -a)
        if name == "Ldexp":
            return np.ldexp(a, b.astype(np.int32)).astype(np.float32)
        if name == "Length":
            return _vec(np.sqrt(sum_squares(a)), np.sqrt(sum_squares(a[:, :3])), np.sqrt(sum_squares(a[:, :2])),
                        np.sqrt(a[:, 0] * a[:, 0]))
        if name == "Distance":
            dv = a - b
            return _vec(np.sqrt(sum_squares(dv)), np.sqrt(sum_squares(dv[:, :3])), np.sqrt(dv[:, 0] * dv[:, 0]), z)
        if name == "Cross":
            x, y = a, b
            return _vec(x[:, 1] * y[:, 2] - y[:, 1] * x[:, 2], x[:, 2] * y[:, 0] - y[:, 2] * x[:, 0],
                        x[:, 0] * y[:, 1] - y[:, 0] * x[:, 1], z)
        if name == "Refract":
            i3 = a[:, :3]
            b3 = b[:, :3]
            nrm = b3 * (ONE / np.sqrt(sum_squares(b3)))[:, None]
            eta = c[:, 0]
            dd = dot_glm(nrm, i3)
            k = ONE - eta * eta * (ONE - dd * dd)
            r = eta[:, None] * i3 - (eta * dd + np.sqrt(k))[:, None] * nrm
            r = np.where((k < 0)[:, None], ZERO, r)
            return _vec(r[:, 0], r[:, 1], r[:, 2], z)
        if name == "Dot":
            return _vec(dot_glm(a[:, :2], b[:, :2]), dot_glm(a[:, :3], b[:, :3]), dot_glm(a, b), z)
        if name == "FOrdCompare":
            return _vec(*_pack01(a < b, a == b, (a < b) | (a > b), a >= b))
        if name == "FUnordCompare":
            return _vec(*_pack01(~(a >= b), ~((a < b) | (a > b)), a != b, ~(a < b)))
        if name == "IsNanAnyAll":
            lt = a < b
            return _vec(_pack01(np.isnan(a))[0], lt.any(axis=1).astype(np.float32), lt.all(axis=1).astype(np.float32),
                        np.isinf(a[:, 0]).astype(np.float32))
        if name == "SelectVector":
            return np.where(a < b, c, d)
        cols = [a, b, c, d]
        if name == "MatrixTimesVector4":
            return _mat_vec4(cols, a)
        if name == "MatrixTimesVector3":
            r = _mat_vec_ltr([a[:, :3], b[:, :3], c[:, :3]], d[:, :3])
            return _vec(r[:, 0], r[:, 1], r[:, 2], z)
        if name == "VectorTimesMatrix4":
            # row j = m[j][0] v0 + m[j][1] v1 + m[j][2] v2 + m[j][3] v3, left to right
            return _vec(*[((cols[j][:, 0] * a[:, 0] + cols[j][:, 1] * a[:, 1]) + cols[j][:, 2] * a[:, 2]) + cols[j][:, 3] * a[:, 3]
                          for j in range(4)])
        if name == "MatrixTimesMatrix4":
            # column 1 of mat4(a, b, c, d) * mat4(d, c, b, a): A0 c.x + A1 c.y + A2 c.z + A3 c.w, left to right
            return _mat_vec_ltr(cols, c)
        if name == "MatrixTimesScalar4":
            return c * a[:, 0:1]
        if name == "Transpose4":
            return _vec(a[:, 0], b[:, 0], c[:, 0], d[:, 0])
        if name == "OuterProduct4":
            return a * b[:, 3:4]
        if name == "VectorTimesScalar":
            return a * b[:, 0:1]
        if name == "FunctionCall":
            y = b * TWO
            return (a + y) + y
        if name == "DynamicInsertExtract":
            ib = b[:, 0].astype(np.int32) & 3
            ic = c[:, 1].astype(np.int32) & 3
            v = a.copy()
            v[np.arange(n), ib] = c[:, 0]
            x = v[np.arange(n), ic]
            v[:, 3] = x
            return v
        if name == "Pack":
            un = pack_unorm(a, 255.0)
            sn = pack_snorm(a, 127.0, 8)
            x = un[:, 0] | (un[:, 1] << 8) | (un[:, 2] << 16) | (un[:, 3] << 24)
            y = sn[:, 0] | (sn[:, 1] << 8) | (sn[:, 2] << 16) | (sn[:, 3] << 24)
            zz = half_bits(a[:, 0]) | (half_bits(a[:, 1]) << 16)
            u2 = pack_unorm(a[:, 2:4], 65535.0)
            w = u2[:, 0] | (u2[:, 1] << 16)
            return _as_f(np.stack([x, y, zz, w], axis=1))
        if name == "Unpack4x8":
            x = _u(a)[:, 0]
            by = [(x >> (8 * k)) & 0xFF for k in range(4)]
            un = [_f(v.astype(np.float32) / F32(255.0)) for v in by]
            sn = [std_clamp(_f(v.astype(np.int8).astype(np.float32) / F32(127.0)), F32(-1.0), ONE) for v in by]
            return _vec(un[0], un[2], sn[0], sn[3])
        if name == "Unpack2x16":
            x, y = _u(a)[:, 0], _u(a)[:, 1]
            h0, h1 = half_value(x & 0xFFFF), half_value(x >> 16)
            u0 = _f((y & 0xFFFF).astype(np.float32) / F32(65535.0))
            u1 = _f((y >> 16).astype(np.float32) / F32(65535.0))
            s1 = std_clamp(_f((x >> 16).astype(np.uint16).view(np.int16).astype(np.float32) / F32(32767.0)), F32(-1.0), ONE)
            return _vec(h0, h1, u0, s1)
    return _as_f(_evaluate_int(name, a, b, c, d))


def _evaluate_int(name, a, b, c, d) -> np.ndarray:
    """Integer cases: uint32 bits of the result."""
    ua, ub, uc = _u(a).astype(np.uint64), _u(b).astype(np.uint64), _u(c).astype(np.uint64)
    ia, ib, ic = _i(a).astype(np.int64), _i(b).astype(np.int64), _i(c).astype(np.int64)
    M = np.uint64(0xFFFFFFFF)

    def wrap(x):
        return (np.asarray(x).astype(np.int64) & 0xFFFFFFFF).astype(np.uint32)

    def tdiv(x, y):
        q = np.abs(x) // np.abs(y)
        return np.where((x < 0) != (y < 0), -q, q)

    if name == "IAdd":
        return wrap(ia + ib)
    if name == "ISub":
        return wrap(ia - ib)
    if name == "IMul":
        return ((ua * ub) & M).astype(np.uint32)
    if name == "SDiv":
        return wrap(tdiv(ia, ib))
    if name == "SRem":
        return wrap(ia - tdiv(ia, ib) * ib)
    if name == "SMod":
        # (uint32)a - (uint32)b * (uint32)(a / b) (SPIRVCompiler.cpp:2696-2704)
        return wrap(ia - ib * tdiv(ia, ib))
    if name == "UDiv":
        return (ua // ub).astype(np.uint32)
    if name == "UMod":
        return (ua % ub).astype(np.uint32)
    if name == "BitwiseAnd":
        return (ua & ub).astype(np.uint32)
    if name == "BitwiseOr":
        return (ua | ub).astype(np.uint32)
    if name == "BitwiseXor":
        return (ua ^ ub).astype(np.uint32)
    if name == "ShiftLeftLogical":
        return ((ua << (ub & np.uint64(31))) & M).astype(np.uint32)
    if name == "ShiftRightLogical":
        return (ua >> (ub & np.uint64(31))).astype(np.uint32)
    if name == "ShiftRightArithmetic":
        return wrap(ia >> (ub & np.uint64(31)).astype(np.int64))
    if name == "SNegate":
        return wrap(-ia)
    if name == "Not":
        return (~ua & M).astype(np.uint32)
    if name == "BitReverse":
        return bit_reverse(ua)
    if name == "BitCount":
        return bit_count(ua)
    if name == "SAbs":
        return wrap(np.abs(ia))
    if name == "SSign":
        return wrap((ia > 0).astype(np.int64) - (ia < 0).astype(np.int64))
    if name == "FindILsb":
        return wrap(find_lsb(ua))
    if name == "FindSMsb":
        return wrap(find_smsb(ia))
    if name == "FindUMsb":
        return wrap(find_umsb(ua))
    if name == "SMin":
        return wrap(std_min(ia, ib))
    if name == "SMax":
        return wrap(std_max(ia, ib))
    if name == "UMin":
        return std_min(ua, ub).astype(np.uint32)
    if name == "UMax":
        return std_max(ua, ub).astype(np.uint32)
    if name == "SClamp":
        return wrap(std_clamp(ia, ib, ic))
    if name == "UClamp":
        return std_clamp(ua, ub, uc).astype(np.uint32)
    if name in ("BitFieldInsert", "BitFieldUExtract", "BitFieldSExtract"):
        off = np.broadcast_to((ic[:, 0:1] & 15), ua.shape)
        cnt = np.broadcast_to((ic[:, 1:2] & 15), ua.shape)
        if name == "BitFieldInsert":
            return bfi_reference(ua, ub, off, cnt)
        if name == "BitFieldUExtract":
            return bfe_u(ua, off, cnt)
        return wrap(bfe_s(ua, off, cnt))
    if name == "IntCompare":
        p = _pack01(ia < ib, ua < ub, ia == ib, ua >= ub)
        return _u(np.stack(p, axis=1))
    if name == "Convert":
        fs = wrap(np.trunc(a[:, 0].astype(np.float64)).astype(np.int64))
        fu = wrap(np.trunc(np.abs(a[:, 1]).astype(np.float64)).astype(np.int64))
        sf = _u(_i(b)[:, 2].astype(np.float32))
        uf = _u(_u(b)[:, 3].astype(np.float32))
        return np.stack([fs, fu, sf, uf], axis=1)
    raise KeyError(f"unknown conformance case {name}")


def ulp_distance(x, y) -> np.ndarray:
    """Distance in float32 ulps between x and y (0 for identical bits and for
    two NaNs)."""
    xi = _i(x).astype(np.int64)
    yi = _i(y).astype(np.int64)
    # map the sign-magnitude bit patterns onto a monotonic integer line
    xi = np.where(xi < 0, -(xi & 0x7FFFFFFF), xi)
    yi = np.where(yi < 0, -(yi & 0x7FFFFFFF), yi)
    d = np.abs(xi - yi)
    both_nan = np.isnan(_f(x)) & np.isnan(_f(y))
    return np.where(both_nan, 0, d)
