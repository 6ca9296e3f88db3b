"""Pins of the oracle's fixed-function and shader-math restatements, each
against a table written here from the reference's rules, on inputs chosen
so that a plausible misreading (a swapped factor, an unordered compare, the
back face's stencil ops, a transposed matrix, another summation order)
fails at least one case.

  blend     ApplyBlendFactor / ApplyBlend (CPVulkan/CommandBuffer.Draw.cpp:956-1262):
            every factor for colour and alpha, separate alpha factors and
            ops, SRC_ALPHA_SATURATE's alpha of 1, MIN / MAX ignoring factors
  depth     CompileFCompareTest's ordered predicates
            (LLVMRuntime/PipelineCompiler.cpp:1492-1508) with NaN, +-0, +-inf
  bounds    CompileDepthBoundsTest's unordered FCmpULT / FCmpUGT (:1128-1166)
  stencil   CompileICompareTest's unsigned compares of (ref & mask) and
            (cur & mask) (:1168-1223, 1510-1526), every stencil op with the
            i8 saturating clamps (:1382-1413), the write through the front
            state's ops and write mask for both faces with the facing
            state's reference (:1246-1380)
  glm       mat4 * vec4 (Samples/utils/glm/detail/type_mat4x4.inl:651-697),
            dot (func_geometric.inl:66-73), normalize (:256-266 with
            func_exponential.inl:55-60) in glm 0.9.5.3's order

Inputs for the blend table are dyadic rationals with few bits, so every
product and sum is exact in fp32 and the expected values are exact
rationals (fractions.Fraction)."""
import ctypes as C
import itertools
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib
from cpvulkan_amd import abi
from cpvulkan_amd import workloads as W

B, OP, CMP, ST, F = abi.BLEND, abi.BLEND_OP, abi.COMPARE, abi.STENCIL, abi.FORMAT
f32 = np.float32

# ------------------------------------------------------------------ blend

S_ = (Fraction(1, 2), Fraction(1, 4), Fraction(3, 4), Fraction(5, 8))      # source colour (FS output)
D_ = (Fraction(1, 8), Fraction(3, 8), Fraction(7, 8), Fraction(1, 4))      # destination
K_ = (Fraction(1, 16), Fraction(3, 16), Fraction(5, 16), Fraction(9, 16))  # blend constants


def factor(f, s, d, k):
    """The four channels of blend factor f (VkBlendFactor) as a colour factor."""
    one = Fraction(1)
    table = {
        B["ZERO"]: (0, 0, 0, 0), B["ONE"]: (1, 1, 1, 1),
        B["SRC_COLOR"]: s, B["ONE_MINUS_SRC_COLOR"]: tuple(one - v for v in s),
        B["DST_COLOR"]: d, B["ONE_MINUS_DST_COLOR"]: tuple(one - v for v in d),
        B["SRC_ALPHA"]: (s[3],) * 4, B["ONE_MINUS_SRC_ALPHA"]: (one - s[3],) * 4,
        B["DST_ALPHA"]: (d[3],) * 4, B["ONE_MINUS_DST_ALPHA"]: (one - d[3],) * 4,
        B["CONSTANT_COLOR"]: k, B["ONE_MINUS_CONSTANT_COLOR"]: tuple(one - v for v in k),
        B["CONSTANT_ALPHA"]: (k[3],) * 4, B["ONE_MINUS_CONSTANT_ALPHA"]: (one - k[3],) * 4,
    }
    if f == B["SRC_ALPHA_SATURATE"]:
        m = min(s[3], one - d[3])
        return (m, m, m, one)
    return tuple(Fraction(v) for v in table[f])


def alpha_factor(f, s, d, k):
    """The alpha channel a separate alpha factor gives: the colour forms
    take their alpha component, SRC_ALPHA_SATURATE gives 1."""
    if f == B["SRC_ALPHA_SATURATE"]:
        return Fraction(1)
    return factor(f, s, d, k)[3]


def expected_blend(sc, dc, op_c, sa, da, op_a, s=S_, d=D_, k=K_):
    fs = list(factor(sc, s, d, k))
    fd = list(factor(dc, s, d, k))
    if sa != sc:
        fs[3] = alpha_factor(sa, s, d, k)
    if da != dc:
        fd[3] = alpha_factor(da, s, d, k)

    def apply(op, i):
        if op == OP["ADD"]:
            return s[i] * fs[i] + d[i] * fd[i]
        if op == OP["SUBTRACT"]:
            return s[i] * fs[i] - d[i] * fd[i]
        if op == OP["REVERSE_SUBTRACT"]:
            return d[i] * fd[i] - s[i] * fs[i]
        if op == OP["MIN"]:
            return min(s[i], d[i])
        return max(s[i], d[i])

    out = [apply(op_c, i) for i in range(4)]
    if op_a != op_c:
        out[3] = apply(op_a, 3)
    return out


def oracle_blend(sc, dc, op_c, sa, da, op_a, s=S_, d=D_, k=K_):
    st = abi.BlendAttachment(1, sc, dc, op_c, sa, da, op_a, 15)
    arr = lambda v: (C.c_float * 4)(*[float(x) for x in v])
    out = (C.c_float * 4)()
    oracle_lib.lib().cpvref_apply_blend(arr(s), arr(d), arr(k), C.byref(st), out)
    return [Fraction(float(v)) for v in out]


FACTORS = list(range(15))   # ZERO .. SRC_ALPHA_SATURATE (the SRC1 factors stop the reference)


def test_blend_every_factor_pair():
    for sc, dc in itertools.product(FACTORS, FACTORS):
        want = expected_blend(sc, dc, OP["ADD"], sc, dc, OP["ADD"])
        got = oracle_blend(sc, dc, OP["ADD"], sc, dc, OP["ADD"])
        assert got == want, (sc, dc, got, want)


def test_blend_separate_alpha_factors():
    for sc, sa in itertools.product(FACTORS, FACTORS):
        dc, da = B["ONE_MINUS_SRC_ALPHA"], B["DST_ALPHA"]
        want = expected_blend(sc, dc, OP["ADD"], sa, da, OP["ADD"])
        got = oracle_blend(sc, dc, OP["ADD"], sa, da, OP["ADD"])
        assert got == want, (sc, sa, got, want)


def test_blend_saturate_alpha_is_one():
    # as a colour factor (colour == alpha factor) and as a separate alpha factor
    got = oracle_blend(B["SRC_ALPHA_SATURATE"], B["ZERO"], OP["ADD"], B["SRC_ALPHA_SATURATE"], B["ZERO"], OP["ADD"])
    assert got[3] == S_[3] * 1
    got = oracle_blend(B["ZERO"], B["ZERO"], OP["ADD"], B["SRC_ALPHA_SATURATE"], B["ZERO"], OP["ADD"])
    assert got[:3] == [0, 0, 0] and got[3] == S_[3]


@pytest.mark.parametrize("op_c", list(OP.values()))
@pytest.mark.parametrize("op_a", list(OP.values()))
def test_blend_ops(op_c, op_a):
    for sc, dc in [(B["SRC_ALPHA"], B["ONE_MINUS_SRC_ALPHA"]), (B["CONSTANT_COLOR"], B["DST_COLOR"]),
                   (B["ONE"], B["ONE"])]:
        want = expected_blend(sc, dc, op_c, sc, dc, op_a)
        got = oracle_blend(sc, dc, op_c, sc, dc, op_a)
        assert got == want, (op_c, op_a, sc, dc, got, want)


def test_blend_min_max_ignore_factors():
    for op in (OP["MIN"], OP["MAX"]):
        a = oracle_blend(B["ZERO"], B["ZERO"], op, B["ZERO"], B["ZERO"], op)
        b = oracle_blend(B["ONE"], B["CONSTANT_ALPHA"], op, B["DST_ALPHA"], B["ONE"], op)
        assert a == b == [min(x, y) if op == OP["MIN"] else max(x, y) for x, y in zip(S_, D_)]


# ------------------------------------------------------------ fragment tests


def _one_pixel_pass(ds_format, ds_bytes, colour_bytes):
    """A 1x1 pass over host buffers: R8G8B8A8 colour and the given
    depth/stencil format, both already holding the stored pixel."""
    cbuf = np.frombuffer(bytearray(colour_bytes), np.uint8).copy()
    dbuf = np.frombuffer(bytearray(ds_bytes), np.uint8).copy()
    p = abi.PassDesc()
    p.width = p.height = 1
    p.color_count = 1
    p.color[0].address = cbuf.ctypes.data
    p.color[0].format = F["R8G8B8A8_UNORM"]
    p.has_depth_stencil = 1
    p.depth_stencil.address = dbuf.ctypes.data
    p.depth_stencil.format = ds_format
    return p, cbuf, dbuf


def _pipeline(ds_format=F["D32_SFLOAT"], **kw):
    d = dict(W.soup_pipeline())
    d.pop("blend", None)
    d.update(kw, color_format=F["R8G8B8A8_UNORM"], ds_format=ds_format)
    return W.pipeline_desc(d)


def _fragment(pipe, pass_, depth, front=1, colour=(0.25, 0.5, 0.75, 1.0)):
    rc = oracle_lib.lib().cpvref_fragment(C.byref(pipe), C.byref(pass_), 0, 0, C.c_float(depth), front,
                                          (C.c_float * 4)(*colour))
    assert rc == 0


SPECIAL = [float("nan"), float("-inf"), -1.0, -0.0, 0.0, 0.25, 0.5, 1.0, float("inf")]


def ordered(op, a, b):
    """LLVM's ordered fcmp predicates (LangRef): false when either is NaN."""
    if math.isnan(a) or math.isnan(b):
        return op == CMP["ALWAYS"]
    return {CMP["NEVER"]: False, CMP["LESS"]: a < b, CMP["EQUAL"]: a == b, CMP["LESS_OR_EQUAL"]: a <= b,
            CMP["GREATER"]: a > b, CMP["NOT_EQUAL"]: a != b, CMP["GREATER_OR_EQUAL"]: a >= b,
            CMP["ALWAYS"]: True}[op]


@pytest.mark.parametrize("op", list(CMP.values()))
def test_depth_compare_ordered_predicates(op):
    """D32_SFLOAT, depth write on: the fragment's depth is stored, and its
    colour written, exactly when the ordered predicate holds; NaN on
    either side fails every op but ALWAYS, and -0 equals +0."""
    for frag, stored in itertools.product(SPECIAL, SPECIAL):
        p, cbuf, dbuf = _one_pixel_pass(F["D32_SFLOAT"], 4, 4)
        dbuf.view(np.float32)[0] = stored
        _fragment(_pipeline(depth_compare=op), p, frag)
        passed = ordered(op, frag, stored)
        now = dbuf.view(np.uint32)[0]
        want = np.float32(frag).view(np.uint32) if passed else np.float32(stored).view(np.uint32)
        assert now == want, (op, frag, stored)
        assert (cbuf.tolist() == [64, 128, 191, 255]) == passed, (op, frag, stored, cbuf)


def test_depth_bounds_are_unordered():
    """CompileDepthBoundsTest rejects when the stored depth is ULT min or UGT
    max: a NaN stored depth is outside every bound; the bounds are
    inclusive."""
    lo, hi = 0.25, 0.75
    for stored in [float("nan"), 0.0, 0.24, 0.25, 0.5, 0.75, 0.76, 1.0, float("-inf"), float("inf")]:
        p, cbuf, dbuf = _one_pixel_pass(F["D32_SFLOAT"], 4, 4)
        dbuf.view(np.float32)[0] = stored
        pipe = _pipeline(depth_compare=CMP["ALWAYS"], depth_bounds_test=1)
        pipe.min_depth_bounds, pipe.max_depth_bounds = lo, hi
        _fragment(pipe, p, 0.1)
        inside = (not math.isnan(stored)) and lo <= stored <= hi
        assert (cbuf.tolist() != [0, 0, 0, 0]) == inside, stored
        assert (dbuf.view(np.float32)[0] == np.float32(0.1)) == inside, stored


def icompare(op, a, b):
    return {CMP["NEVER"]: False, CMP["LESS"]: a < b, CMP["EQUAL"]: a == b, CMP["LESS_OR_EQUAL"]: a <= b,
            CMP["GREATER"]: a > b, CMP["NOT_EQUAL"]: a != b, CMP["GREATER_OR_EQUAL"]: a >= b,
            CMP["ALWAYS"]: True}[op]


def stencil_op(op, cur, ref):
    """VkStencilOp on an 8-bit value; the clamped forms saturate as signed
    bytes (llvm.sadd.sat / llvm.ssub.sat on i8)."""
    s = cur - 256 if cur >= 128 else cur
    if op == ST["KEEP"]:
        return cur
    if op == ST["ZERO"]:
        return 0
    if op == ST["REPLACE"]:
        return ref
    if op == ST["INC_CLAMP"]:
        return min(s + 1, 127) & 0xFF
    if op == ST["DEC_CLAMP"]:
        return max(s - 1, -128) & 0xFF
    if op == ST["INVERT"]:
        return (~cur) & 0xFF
    if op == ST["INC_WRAP"]:
        return (cur + 1) & 0xFF
    return (cur - 1) & 0xFF


def test_stencil_ops_saturate_as_signed_bytes():
    table = {ST["INC_CLAMP"]: {0: 1, 126: 127, 127: 127, 128: 129, 255: 0},
             ST["DEC_CLAMP"]: {0: 255, 1: 0, 128: 128, 129: 128, 255: 254},
             ST["INC_WRAP"]: {255: 0}, ST["DEC_WRAP"]: {0: 255}, ST["INVERT"]: {0x5A: 0xA5}}
    for op, cases in table.items():
        for cur, want in cases.items():
            assert stencil_op(op, cur, 0) == want
            p, cbuf, dbuf = _one_pixel_pass(F["S8_UINT"], 1, 4)
            dbuf[0] = cur
            state = (ST["KEEP"], op, ST["KEEP"], CMP["ALWAYS"], 0xFF, 0xFF, 0)
            _fragment(_pipeline(F["S8_UINT"], stencil_test=1, depth_test=0, front=state, back=state), p, 0.5)
            assert dbuf[0] == want, (op, cur, dbuf[0])


@pytest.mark.parametrize("front", [1, 0])
def test_stencil_compare_and_front_state_write(front):
    """The compare uses the facing state's op, reference and compare mask;
    the write uses the FRONT state's fail / pass / depth-fail ops and write
    mask for both faces, with the facing state's reference for REPLACE."""
    rng = np.random.Generator(np.random.PCG64(3 + front))
    ops = list(ST.values())
    cmps = list(CMP.values())
    for _ in range(400):
        f_state = (int(rng.choice(ops)), int(rng.choice(ops)), int(rng.choice(ops)), int(rng.choice(cmps)),
                   int(rng.integers(0, 256)), int(rng.integers(0, 256)), int(rng.integers(0, 256)))
        b_state = (int(rng.choice(ops)), int(rng.choice(ops)), int(rng.choice(ops)), int(rng.choice(cmps)),
                   int(rng.integers(0, 256)), int(rng.integers(0, 256)), int(rng.integers(0, 256)))
        cur = int(rng.integers(0, 256))
        stored_depth, frag_depth = 0.5, float(rng.choice([0.25, 0.75]))
        p, cbuf, dbuf = _one_pixel_pass(F["D24_UNORM_S8_UINT"], 4, 4)
        oracle_lib.lib().cpvref_pack_depth_stencil(F["D24_UNORM_S8_UINT"], C.c_float(stored_depth), cur, dbuf.ctypes.data)
        pipe = _pipeline(F["D24_UNORM_S8_UINT"], stencil_test=1, depth_compare=CMP["LESS"], front=f_state, back=b_state)
        _fragment(pipe, p, frag_depth, front=front)
        facing = f_state if front else b_state
        ref, cmask = facing[6], facing[4]
        s_ok = icompare(facing[3], ref & cmask, cur & cmask)
        d_ok = frag_depth < stored_depth
        w = f_state   # ops and write mask: always the front state's
        value = stencil_op(w[1], cur, ref) if (s_ok and d_ok) else stencil_op(w[2], cur, ref) if s_ok else stencil_op(w[0], cur, ref)
        value = (value & w[5]) | (cur & ~w[5] & 0xFF)
        got_s = oracle_lib.lib().cpvref_unpack_stencil(F["D24_UNORM_S8_UINT"], dbuf.ctypes.data)
        assert got_s == value, (front, f_state, b_state, cur, frag_depth, got_s, value)
        # the depth is written whenever the depth test passes, whatever the
        # stencil test gave: CompileStencilWrite's depth write is conditioned
        # on depthResult alone (:1316-1318), a quirk of the reference
        got_d = oracle_lib.lib().cpvref_unpack_depth(F["D24_UNORM_S8_UINT"], dbuf.ctypes.data)
        want_d = frag_depth if d_ok else stored_depth
        assert abs(got_d - want_d) < 1e-6, (got_d, want_d)
        assert (cbuf.tolist() != [0, 0, 0, 0]) == (s_ok and d_ok)


# ----------------------------------------------------------------- glm math


def glm_mat4_vec4(m, v):
    """glm 0.9.5.3: Add0 = m[0]*v0 + m[1]*v1, Add1 = m[2]*v2 + m[3]*v3, Add0 + Add1,
    per row, with m column-major (m[c] is column c), every operation in fp32."""
    m = np.asarray(m, f32).reshape(4, 4)   # m[c][r]
    out = []
    for r in range(4):
        add0 = f32(m[0][r] * v[0]) + f32(m[1][r] * v[1])
        add1 = f32(m[2][r] * v[2]) + f32(m[3][r] * v[3])
        out.append(f32(f32(add0) + f32(add1)))
    return np.asarray(out, f32)


def oracle_glm(op, a, b=None):
    a = np.ascontiguousarray(a, f32)
    b = np.ascontiguousarray(b if b is not None else np.zeros(4), f32)
    out = np.zeros(4, f32)
    oracle_lib.lib().cpvref_glm(op, a.ctypes.data_as(C.POINTER(C.c_float)), b.ctypes.data_as(C.POINTER(C.c_float)),
                                out.ctypes.data_as(C.POINTER(C.c_float)))
    return out


def test_mat4_vec4_glm_order_and_layout():
    rng = np.random.Generator(np.random.PCG64(11))
    distinguished_order = distinguished_layout = 0
    for _ in range(3000):
        # magnitudes spread over many binades so that summation order matters
        m = (rng.uniform(-1, 1, 16) * 2.0 ** rng.integers(-12, 12, 16)).astype(f32)
        v = (rng.uniform(-1, 1, 4) * 2.0 ** rng.integers(-12, 12, 4)).astype(f32)
        got = oracle_glm(0, m, v)
        want = glm_mat4_vec4(m, v)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), (m, v, got, want)
        # the alternatives a plausible misreading would give
        mm = m.reshape(4, 4)
        left = np.asarray([f32(f32(f32(mm[0][r] * v[0]) + f32(mm[1][r] * v[1])) + f32(mm[2][r] * v[2])) + f32(mm[3][r] * v[3])
                           for r in range(4)], f32)
        distinguished_order += not np.array_equal(left.view(np.uint32), got.view(np.uint32))
        distinguished_layout += not np.array_equal(glm_mat4_vec4(mm.T.reshape(-1), v).view(np.uint32), got.view(np.uint32))
    # the inputs did tell the orders and the layouts apart
    assert distinguished_order > 100 and distinguished_layout > 2500


def test_dot_and_normalize_glm_order():
    rng = np.random.Generator(np.random.PCG64(12))
    distinguished = 0
    for _ in range(3000):
        a = (rng.uniform(-1, 1, 3) * 2.0 ** rng.integers(-10, 10, 3)).astype(f32)
        b = (rng.uniform(-1, 1, 3) * 2.0 ** rng.integers(-10, 10, 3)).astype(f32)
        d = oracle_glm(1, a, b)[0]
        x, y, z = f32(a[0] * b[0]), f32(a[1] * b[1]), f32(a[2] * b[2])
        assert d.view(np.uint32) == f32(f32(x + y) + z).view(np.uint32)
        distinguished += f32(x + f32(y + z)).view(np.uint32) != d.view(np.uint32)
        n = oracle_glm(2, a)[:3]
        sqr = f32(f32(f32(a[0] * a[0]) + f32(a[1] * a[1])) + f32(a[2] * a[2]))
        inv = f32(f32(1.0) / f32(np.sqrt(sqr)))
        want = np.asarray([f32(a[i] * inv) for i in range(3)], f32)
        assert np.array_equal(n.view(np.uint32), want.view(np.uint32)), (a, n, want)
    assert distinguished > 100
