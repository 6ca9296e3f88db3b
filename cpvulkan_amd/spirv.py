"""A small SPIR-V module builder, and the shaders of the benchmark workloads
as SPIR-V.

The image has no GLSL compiler, so the SPIR-V that `cpv_program_create_spirv`
translates is written here, instruction by instruction, in the shape a GLSL
front end emits for the GLSL quoted with each shader (logical addressing,
GLSL450 memory model, one entry point, Function-storage locals, structured
control flow).  The numbers are those of the SPIR-V 1.0 specification and
the GLSL.std.450 extended instruction set.
"""
from __future__ import annotations

import struct

MAGIC = 0x07230203
VERSION = 0x00010000

# opcodes
OP = dict(
    Undef=1, Name=5, ExtInstImport=11, ExtInst=12, MemoryModel=14, EntryPoint=15, ExecutionMode=16, Capability=17,
    TypeVoid=19, TypeBool=20, TypeInt=21, TypeFloat=22, TypeVector=23, TypeMatrix=24, TypeImage=25,
    TypeSampler=26, TypeSampledImage=27, TypeArray=28, TypeStruct=30, TypePointer=32, TypeFunction=33,
    ConstantTrue=41, ConstantFalse=42, Constant=43, ConstantComposite=44, ConstantNull=46, SpecConstant=50,
    Function=54, FunctionParameter=55, FunctionEnd=56, FunctionCall=57, Variable=59, Load=61, Store=62,
    AccessChain=65, Decorate=71, MemberDecorate=72, VectorExtractDynamic=77, VectorInsertDynamic=78,
    VectorShuffle=79, CompositeConstruct=80, CompositeExtract=81, CompositeInsert=82, CopyObject=83,
    Transpose=84, ImageSampleImplicitLod=87, ImageSampleExplicitLod=88, ImageFetch=95, Image=100,
    ImageQuerySizeLod=103, ImageQuerySize=104, ImageQueryLevels=106, ConvertFToU=109, ConvertFToS=110,
    ConvertSToF=111, ConvertUToF=112, UConvert=113, SConvert=114, FConvert=115, QuantizeToF16=116, Bitcast=124,
    SNegate=126, FNegate=127, IAdd=128, FAdd=129, ISub=130, FSub=131, IMul=132, FMul=133, UDiv=134, SDiv=135,
    FDiv=136, UMod=137, SRem=138, SMod=139, FRem=140, FMod=141, VectorTimesScalar=142, MatrixTimesScalar=143,
    VectorTimesMatrix=144, MatrixTimesVector=145, MatrixTimesMatrix=146, OuterProduct=147, Dot=148, Any=154,
    All=155, IsNan=156, IsInf=157, LogicalEqual=164, LogicalNotEqual=165, LogicalOr=166, LogicalAnd=167,
    LogicalNot=168, Select=169, IEqual=170, INotEqual=171, UGreaterThan=172, SGreaterThan=173,
    UGreaterThanEqual=174, SGreaterThanEqual=175, ULessThan=176, SLessThan=177, ULessThanEqual=178,
    SLessThanEqual=179, FOrdEqual=180, FUnordEqual=181, FOrdNotEqual=182, FUnordNotEqual=183, FOrdLessThan=184,
    FUnordLessThan=185, FOrdGreaterThan=186, FUnordGreaterThan=187, FOrdLessThanEqual=188,
    FUnordLessThanEqual=189, FOrdGreaterThanEqual=190, FUnordGreaterThanEqual=191, ShiftRightLogical=194,
    ShiftRightArithmetic=195, ShiftLeftLogical=196, BitwiseOr=197, BitwiseXor=198, BitwiseAnd=199, Not=200,
    BitFieldInsert=201, BitFieldSExtract=202, BitFieldUExtract=203, BitReverse=204, BitCount=205, Phi=245,
    LoopMerge=246, SelectionMerge=247, Label=248, Branch=249, BranchConditional=250, Switch=251, Kill=252,
    Return=253, ReturnValue=254, Unreachable=255,
)
GLSL = dict(Round=1, RoundEven=2, Trunc=3, FAbs=4, SAbs=5, FSign=6, SSign=7, Floor=8, Ceil=9, Fract=10, Radians=11,
            Degrees=12, Sin=13, Cos=14, Tan=15, Asin=16, Acos=17, Atan=18, Sinh=19, Cosh=20, Tanh=21, Asinh=22,
            Acosh=23, Atanh=24, Atan2=25, Pow=26, Exp=27, Log=28, Exp2=29, Log2=30, Sqrt=31, InverseSqrt=32,
            FMin=37, UMin=38, SMin=39, FMax=40, UMax=41, SMax=42, FClamp=43, UClamp=44, SClamp=45, FMix=46,
            Step=48, SmoothStep=49, Fma=50, Ldexp=53, PackSnorm4x8=54, PackUnorm4x8=55, PackSnorm2x16=56,
            PackUnorm2x16=57, PackHalf2x16=58, UnpackSnorm2x16=60, UnpackUnorm2x16=61, UnpackHalf2x16=62,
            UnpackSnorm4x8=63, UnpackUnorm4x8=64, Length=66, Distance=67, Cross=68, Normalize=69, FaceForward=70,
            Reflect=71, Refract=72, FindILsb=73, FindSMsb=74, FindUMsb=75, NMin=79, NMax=80, NClamp=81)
DEC = dict(Block=2, ColMajor=5, ArrayStride=6, MatrixStride=7, BuiltIn=11, NoPerspective=13, Flat=14, Location=30,
           Binding=33, DescriptorSet=34, Offset=35)
BUILTIN = dict(Position=0, PointSize=1, ClipDistance=3, CullDistance=4, FragCoord=15, FrontFacing=17,
               VertexIndex=42, InstanceIndex=43)
SC = dict(UniformConstant=0, Input=1, Uniform=2, Output=3, Private=6, Function=7)
MODEL = dict(Vertex=0, Fragment=4)


def _string(s: str) -> list[int]:
    b = s.encode() + b"\0"
    b += b"\0" * (-len(b) % 4)
    return list(struct.unpack(f"<{len(b) // 4}I", b))


class Module:
    """Accumulates the logical sections of a module and assembles them."""

    def __init__(self):
        self.bound = 1
        self.caps, self.ext, self.mm, self.entry, self.modes = [], [], [], [], []
        self.debug, self.annot, self.types, self.funcs = [], [], [], []
        self._types: dict = {}
        self._consts: dict = {}
        self.glsl = None

    def id(self) -> int:
        i = self.bound
        self.bound += 1
        return i

    @staticmethod
    def _ins(sec: list, op: str, *ops) -> None:
        words = []
        for o in ops:
            words.extend(_string(o) if isinstance(o, str) else [int(o) & 0xFFFFFFFF])
        sec.append([((len(words) + 1) << 16) | OP[op]] + words)

    # types (deduplicated)
    def type(self, op: str, *ops) -> int:
        key = (op,) + tuple(ops)
        if key not in self._types:
            t = self.id()
            self._ins(self.types, op, t, *ops)
            self._types[key] = t
        return self._types[key]

    def void(self): return self.type("TypeVoid")
    def bool(self): return self.type("TypeBool")
    def f32(self): return self.type("TypeFloat", 32)
    def i32(self): return self.type("TypeInt", 32, 1)
    def u32(self): return self.type("TypeInt", 32, 0)
    def vec(self, n): return self.type("TypeVector", self.f32(), n)
    def ivec(self, n): return self.type("TypeVector", self.i32(), n)
    def uvec(self, n): return self.type("TypeVector", self.u32(), n)
    def bvec(self, n): return self.type("TypeVector", self.bool(), n)
    def mat(self, cols, rows): return self.type("TypeMatrix", self.vec(rows), cols)
    def ptr(self, sc, t): return self.type("TypePointer", SC[sc], t)
    def fn(self, ret, *params): return self.type("TypeFunction", ret, *params)

    def sampled_image_2d(self):
        img = self.type("TypeImage", self.f32(), 1, 0, 0, 0, 1, 0)   # 2D, not depth/arrayed/MS, sampled, Unknown
        return self.type("TypeSampledImage", img)

    def struct(self, *members) -> int:
        t = self.id()   # structs are not deduplicated (decorations differ)
        self._ins(self.types, "TypeStruct", t, *members)
        return t

    def const_f(self, v: float) -> int:
        key = ("f", struct.pack("<f", v))
        if key not in self._consts:
            c = self.id()
            self._ins(self.types, "Constant", self.f32(), c, struct.unpack("<I", struct.pack("<f", v))[0])
            self._consts[key] = c
        return self._consts[key]

    def const_i(self, v: int) -> int:
        key = ("i", v)
        if key not in self._consts:
            c = self.id()
            self._ins(self.types, "Constant", self.i32(), c, v)
            self._consts[key] = c
        return self._consts[key]

    def const_u(self, v: int) -> int:
        key = ("u", v)
        if key not in self._consts:
            c = self.id()
            self._ins(self.types, "Constant", self.u32(), c, v)
            self._consts[key] = c
        return self._consts[key]

    def const_b(self, v: bool) -> int:
        key = ("b", bool(v))
        if key not in self._consts:
            c = self.id()
            self._ins(self.types, "ConstantTrue" if v else "ConstantFalse", self.bool(), c)
            self._consts[key] = c
        return self._consts[key]

    def const_composite(self, t: int, *parts) -> int:
        key = ("c", t) + tuple(parts)
        if key not in self._consts:
            c = self.id()
            self._ins(self.types, "ConstantComposite", t, c, *parts)
            self._consts[key] = c
        return self._consts[key]

    def variable(self, sc: str, pointee: int, name: str | None = None) -> int:
        v = self.id()
        self._ins(self.types, "Variable", self.ptr(sc, pointee), v, SC[sc])
        if name:
            self._ins(self.debug, "Name", v, name)
        return v

    def decorate(self, target, dec: str, *args) -> None:
        self._ins(self.annot, "Decorate", target, DEC[dec], *args)

    def member_decorate(self, target, member, dec: str, *args) -> None:
        self._ins(self.annot, "MemberDecorate", target, member, DEC[dec], *args)

    def import_glsl(self) -> int:
        if self.glsl is None:
            self.glsl = self.id()
            self._ins(self.ext, "ExtInstImport", self.glsl, "GLSL.std.450")
        return self.glsl

    def assemble(self, model: str, entry_fn: int, interface: list[int], origin_upper_left: bool) -> list[int]:
        self._ins(self.caps, "Capability", 1)                 # Shader
        self._ins(self.mm, "MemoryModel", 0, 1)               # Logical, GLSL450
        self._ins(self.entry, "EntryPoint", MODEL[model], entry_fn, "main", *interface)
        if model == "Fragment" and origin_upper_left:
            self._ins(self.modes, "ExecutionMode", entry_fn, 7)
        words = [MAGIC, VERSION, 0, self.bound, 0]
        for sec in (self.caps, self.ext, self.mm, self.entry, self.modes, self.debug, self.annot, self.types,
                    self.funcs):
            for ins in sec:
                words.extend(ins)
        return words


class Function:
    """Body builder of one function: the entry point (no arguments), or a
    callee with a return type and parameter types (pointer types for
    inout parameters); its parameter ids are in `params`."""

    def __init__(self, m: Module, ret: int | None = None, params: tuple = ()):
        self.m = m
        self.fid = m.id()
        self.ret = m.void() if ret is None else ret
        m._ins(m.funcs, "Function", self.ret, self.fid, 0, m.fn(self.ret, *params))
        self.params = []
        for p in params:
            pid = m.id()
            m._ins(m.funcs, "FunctionParameter", p, pid)
            self.params.append(pid)
        self.label()

    def label(self, lid: int | None = None) -> int:
        lid = lid or self.m.id()
        self.m._ins(self.m.funcs, "Label", lid)
        return lid

    def op(self, op: str, rtype: int, *ops) -> int:
        r = self.m.id()
        self.m._ins(self.m.funcs, op, rtype, r, *ops)
        return r

    def raw(self, op: str, *ops) -> None:
        self.m._ins(self.m.funcs, op, *ops)

    def load(self, t, ptr): return self.op("Load", t, ptr)
    def store(self, ptr, val): self.raw("Store", ptr, val)

    def chain(self, sc: str, t, base, *idx):
        return self.op("AccessChain", self.m.ptr(sc, t), base, *[self.m.const_i(i) for i in idx])

    def ext(self, t, inst: str, *args):
        return self.op("ExtInst", t, self.m.import_glsl(), GLSL[inst], *args)

    def local(self, t: int) -> int:
        """A Function-storage variable (declared where it is created: the
        translator accepts variables anywhere in the function)."""
        v = self.m.id()
        self.m._ins(self.m.funcs, "Variable", self.m.ptr("Function", t), v, SC["Function"])
        return v

    def end(self, value: int | None = None) -> int:
        if value is None:
            self.raw("Return")
        else:
            self.raw("ReturnValue", value)
        self.raw("FunctionEnd")
        return self.fid


def _per_vertex(m: Module):
    """out gl_PerVertex { vec4 gl_Position; float gl_PointSize; float gl_ClipDistance[1]; }"""
    arr = m.type("TypeArray", m.f32(), m.type("Constant", m.u32(), 1) if False else _u32_const(m, 1))
    s = m.struct(m.vec(4), m.f32(), arr)
    m.decorate(s, "Block")
    m.member_decorate(s, 0, "BuiltIn", BUILTIN["Position"])
    m.member_decorate(s, 1, "BuiltIn", BUILTIN["PointSize"])
    m.member_decorate(s, 2, "BuiltIn", BUILTIN["ClipDistance"])
    return s, m.variable("Output", s, "gl_PerVertex")


def _u32_const(m: Module, v: int) -> int:
    key = ("u", v)
    if key not in m._consts:
        c = m.id()
        m._ins(m.types, "Constant", m.u32(), c, v)
        m._consts[key] = c
    return m._consts[key]


def _io(m: Module, sc: str, t: int, location: int, name: str) -> int:
    v = m.variable(sc, t, name)
    m.decorate(v, "Location", location)
    return v


def _ubo(m: Module, binding: int, member_t: int, name: str, matrix: bool) -> int:
    s = m.struct(member_t)
    m.decorate(s, "Block")
    m.member_decorate(s, 0, "Offset", 0)
    if matrix:
        m.member_decorate(s, 0, "ColMajor")
        m.member_decorate(s, 0, "MatrixStride", 16)
    v = m.variable("Uniform", s, name)
    m.decorate(v, "DescriptorSet", 0)
    m.decorate(v, "Binding", binding)
    return v


def vs_passthrough() -> list[int]:
    """layout(location=0) in vec4 pos; layout(location=1) in vec4 inColor;
    layout(location=0) out vec4 outColor;
    void main() { gl_Position = pos; outColor = inColor; }"""
    m = Module()
    v4 = m.vec(4)
    pv_t, pv = _per_vertex(m)
    pos = _io(m, "Input", v4, 0, "pos")
    col = _io(m, "Input", v4, 1, "inColor")
    out = _io(m, "Output", v4, 0, "outColor")
    f = Function(m)
    f.store(f.chain("Output", v4, pv, 0), f.load(v4, pos))
    f.store(out, f.load(v4, col))
    return m.assemble("Vertex", f.end(), [pv, pos, col, out], False)


def vs_mvp_color() -> list[int]:
    """The 15-draw_cube vertex shader (Samples/15-draw_cube/15-draw_cube.cpp:38-51):
    layout(std140, binding=0) uniform bufferVals { mat4 mvp; } myBufferVals;
    layout(location=0) in vec4 pos; layout(location=1) in vec4 inColor;
    layout(location=0) out vec4 outColor;
    void main() { outColor = inColor; gl_Position = myBufferVals.mvp * pos; }"""
    m = Module()
    v4, m4 = m.vec(4), m.mat(4, 4)
    pv_t, pv = _per_vertex(m)
    ubo = _ubo(m, 0, m4, "myBufferVals", True)
    pos = _io(m, "Input", v4, 0, "pos")
    col = _io(m, "Input", v4, 1, "inColor")
    out = _io(m, "Output", v4, 0, "outColor")
    f = Function(m)
    f.store(out, f.load(v4, col))
    mvp = f.load(m4, f.chain("Uniform", m4, ubo, 0))
    f.store(f.chain("Output", v4, pv, 0), f.op("MatrixTimesVector", v4, mvp, f.load(v4, pos)))
    return m.assemble("Vertex", f.end(), [pv, pos, col, out], False)


def fs_color() -> list[int]:
    """The 15-draw_cube fragment shader (15-draw_cube.cpp:53-61):
    layout(location=0) in vec4 color; layout(location=0) out vec4 outColor;
    void main() { outColor = color; }"""
    m = Module()
    v4 = m.vec(4)
    col = _io(m, "Input", v4, 0, "color")
    out = _io(m, "Output", v4, 0, "outColor")
    f = Function(m)
    f.store(out, f.load(v4, col))
    return m.assemble("Fragment", f.end(), [col, out], True)


def vs_mvp_lambert() -> list[int]:
    """layout(std140, binding=0) uniform Transform { mat4 mvp; };
    layout(location=0) in vec3 pos; layout(location=1) in vec3 normal; layout(location=2) in vec2 uv;
    layout(location=0) out vec3 outNormal; layout(location=1) out vec2 outUv;
    void main() { gl_Position = mvp * vec4(pos, 1.0); outNormal = normal; outUv = uv; }"""
    m = Module()
    v2, v3, v4, m4 = m.vec(2), m.vec(3), m.vec(4), m.mat(4, 4)
    pv_t, pv = _per_vertex(m)
    ubo = _ubo(m, 0, m4, "Transform", True)
    pos = _io(m, "Input", v3, 0, "pos")
    nrm = _io(m, "Input", v3, 1, "normal")
    uv = _io(m, "Input", v2, 2, "uv")
    onrm = _io(m, "Output", v3, 0, "outNormal")
    ouv = _io(m, "Output", v2, 1, "outUv")
    f = Function(m)
    mvp = f.load(m4, f.chain("Uniform", m4, ubo, 0))
    p = f.load(v3, pos)
    x, y, z = (f.op("CompositeExtract", m.f32(), p, i) for i in range(3))
    p4 = f.op("CompositeConstruct", v4, x, y, z, m.const_f(1.0))
    f.store(f.chain("Output", v4, pv, 0), f.op("MatrixTimesVector", v4, mvp, p4))
    f.store(onrm, f.load(v3, nrm))
    f.store(ouv, f.load(v2, uv))
    return m.assemble("Vertex", f.end(), [pv, pos, nrm, uv, onrm, ouv], False)


def fs_tex_lambert() -> list[int]:
    """layout(binding=1) uniform sampler2D tex; layout(std140, binding=2) uniform Light { vec4 dir; };
    layout(location=0) in vec3 normal; layout(location=1) in vec2 uv; layout(location=0) out vec4 outColor;
    void main() { outColor = texture(tex, uv) * max(dot(normalize(normal), dir.xyz), 0.0); }"""
    m = Module()
    f32, v2, v3, v4 = m.f32(), m.vec(2), m.vec(3), m.vec(4)
    si = m.sampled_image_2d()
    tex = m.variable("UniformConstant", si, "tex")
    m.decorate(tex, "DescriptorSet", 0)
    m.decorate(tex, "Binding", 1)
    light = _ubo(m, 2, v4, "Light", False)
    nrm = _io(m, "Input", v3, 0, "normal")
    uv = _io(m, "Input", v2, 1, "uv")
    out = _io(m, "Output", v4, 0, "outColor")
    f = Function(m)
    t = f.op("ImageSampleImplicitLod", v4, f.load(si, tex), f.load(v2, uv))
    n = f.ext(v3, "Normalize", f.load(v3, nrm))
    d4 = f.load(v4, f.chain("Uniform", v4, light, 0))
    d3 = f.op("VectorShuffle", v3, d4, d4, 0, 1, 2)
    l = f.op("Dot", f32, n, d3)
    mx = f.ext(f32, "FMax", l, m.const_f(0.0))
    f.store(out, f.op("VectorTimesScalar", v4, t, mx))
    return m.assemble("Fragment", f.end(), [nrm, uv, out], True)


def fs_discard_dark() -> list[int]:
    """layout(location=0) in vec4 color; layout(location=0) out vec4 outColor;
    void main() { if (color.r < 0.25) discard; outColor = color; }
    (exercises OpKill and structured control flow)"""
    m = Module()
    f32, v4 = m.f32(), m.vec(4)
    col = _io(m, "Input", v4, 0, "color")
    out = _io(m, "Output", v4, 0, "outColor")
    f = Function(m)
    c = f.load(v4, col)
    r = f.op("CompositeExtract", f32, c, 0)
    cond = f.op("FOrdLessThan", m.bool(), r, m.const_f(0.25))
    then_l, merge_l = m.id(), m.id()
    f.raw("SelectionMerge", merge_l, 0)
    f.raw("BranchConditional", cond, then_l, merge_l)
    f.label(then_l)
    f.raw("Kill")
    f.label(merge_l)
    f.store(out, c)
    return m.assemble("Fragment", f.end(), [col, out], True)


def vs_point_size() -> list[int]:
    """layout(location=0) in vec4 pos; layout(location=1) in vec4 inColor;
    layout(location=2) in float size; layout(location=0) out vec4 outColor;
    void main() { gl_Position = pos; gl_PointSize = size; outColor = inColor; }"""
    m = Module()
    v4, f32 = m.vec(4), m.f32()
    pv_t, pv = _per_vertex(m)
    pos = _io(m, "Input", v4, 0, "pos")
    col = _io(m, "Input", v4, 1, "inColor")
    size = _io(m, "Input", f32, 2, "size")
    out = _io(m, "Output", v4, 0, "outColor")
    f = Function(m)
    f.store(f.chain("Output", v4, pv, 0), f.load(v4, pos))
    f.store(f.chain("Output", f32, pv, 1), f.load(f32, size))
    f.store(out, f.load(v4, col))
    return m.assemble("Vertex", f.end(), [pv, pos, col, size, out], False)


def _fs_color_decorated(dec: str) -> list[int]:
    m = Module()
    v4 = m.vec(4)
    col = _io(m, "Input", v4, 0, "color")
    m.decorate(col, dec)
    out = _io(m, "Output", v4, 0, "outColor")
    f = Function(m)
    f.store(out, f.load(v4, col))
    return m.assemble("Fragment", f.end(), [col, out], True)


def fs_color_flat() -> list[int]:
    """layout(location=0) flat in vec4 color; layout(location=0) out vec4 outColor;
    void main() { outColor = color; }"""
    return _fs_color_decorated("Flat")


def fs_color_noperspective() -> list[int]:
    """layout(location=0) noperspective in vec4 color; layout(location=0) out vec4 outColor;
    void main() { outColor = color; }"""
    return _fs_color_decorated("NoPerspective")


SHADERS = {
    "vs_passthrough": vs_passthrough,
    "vs_point_size": vs_point_size,
    "fs_color_flat": fs_color_flat,
    "fs_color_noperspective": fs_color_noperspective,
    "vs_mvp_color": vs_mvp_color,
    "vs_mvp_lambert": vs_mvp_lambert,
    "fs_color": fs_color,
    "fs_tex_lambert": fs_tex_lambert,
    "fs_discard_dark": fs_discard_dark,
}


# ---------------------------------------------------------------- conformance shaders
#
# A vertex shader that forwards four vec4 attributes and an integer case
# number as flat varyings, and a fragment shader that evaluates one
# operation per fragment, chosen by the case number, over those inputs:
# integer results leave the shader bit-cast to float, so an
# R32G32B32A32_SFLOAT attachment holds every result's exact bits.

def vs_flat_inputs() -> list[int]:
    """layout(location=0) in vec4 pos; layout(location=1..4) in vec4 a, b, c, d;
    layout(location=5) in vec4 e;
    layout(location=0..3) flat out vec4 oa, ob, oc, od; layout(location=4) flat out int sel;
    void main() { gl_Position = pos; oa = a; ob = b; oc = c; od = d; sel = int(e.x); }"""
    m = Module()
    v4, i32 = m.vec(4), m.i32()
    pv_t, pv = _per_vertex(m)
    pos = _io(m, "Input", v4, 0, "pos")
    ins = [_io(m, "Input", v4, 1 + k, "abcd"[k]) for k in range(4)]
    e = _io(m, "Input", v4, 5, "e")
    outs = [_io(m, "Output", v4, k, "o" + "abcd"[k]) for k in range(4)]
    sel = _io(m, "Output", i32, 4, "sel")
    f = Function(m)
    f.store(f.chain("Output", v4, pv, 0), f.load(v4, pos))
    for i, o in zip(ins, outs):
        f.store(o, f.load(v4, i))
    ex = f.op("CompositeExtract", m.f32(), f.load(v4, e), 0)
    f.store(sel, f.op("ConvertFToS", i32, ex))
    return m.assemble("Vertex", f.end(), [pv, pos] + ins + [e] + outs + [sel], False)


def _alu_cases():
    """(name, builder) per case; a builder gets (f, m, a, b, c, d, helper)
    and returns the vec4 result id."""
    cases = []

    def add(name):
        def deco(fn):
            cases.append((name, fn))
            return fn
        return deco

    def ext(inst, n=1, t=None):
        def b(f, m, a, bb, c, d, h):
            args = (a, bb, c)[:n]
            return f.ext(t or m.vec(4), inst, *args)
        return b

    def splat(f, m, x):
        return f.op("CompositeConstruct", m.vec(4), x, x, x, x)

    def bits(f, m, v, t):     # vec4 float -> t (ivec4 / uvec4) by bit cast
        return f.op("Bitcast", t, v)

    def back(f, m, v):        # integer vec4 -> vec4 float bits
        return f.op("Bitcast", m.vec(4), v)

    def sel01(f, m, cond):    # bvec4 -> vec4 of 1.0 / 0.0
        one = m.const_composite(m.vec(4), *[m.const_f(1.0)] * 4)
        zero = m.const_composite(m.vec(4), *[m.const_f(0.0)] * 4)
        return f.op("Select", m.vec(4), cond, one, zero)

    def comp(f, m, v, i):
        return f.op("CompositeExtract", m.f32(), v, i)

    def vec4_of(f, m, *xs):
        return f.op("CompositeConstruct", m.vec(4), *xs)

    for name in ("FAbs", "Floor", "Ceil", "Trunc", "Round", "RoundEven", "Fract", "FSign", "Sin", "Cos", "Tan", "Asin",
                 "Acos", "Atan", "Sinh", "Cosh", "Tanh", "Asinh", "Acosh", "Atanh", "Exp", "Log", "Exp2", "Log2",
                 "Sqrt", "InverseSqrt", "Radians", "Degrees", "Normalize"):
        add(name)(ext(name, 1))
    for name in ("FMin", "FMax", "NMin", "NMax", "Step", "Atan2", "Pow", "Reflect"):
        add(name)(ext(name, 2))
    for name in ("FClamp", "NClamp", "FMix", "SmoothStep", "Fma", "FaceForward"):
        add(name)(ext(name, 3))
    for name in ("FAdd", "FSub", "FMul", "FDiv", "FRem", "FMod"):
        add(name)((lambda op: lambda f, m, a, b, c, d, h: f.op(op, m.vec(4), a, b))(name))
    add("FNegate")(lambda f, m, a, b, c, d, h: f.op("FNegate", m.vec(4), a))
    add("QuantizeToF16")(lambda f, m, a, b, c, d, h: f.op("QuantizeToF16", m.vec(4), a))

    @add("Ldexp")
    def _(f, m, a, b, c, d, h):
        return f.ext(m.vec(4), "Ldexp", a, f.op("ConvertFToS", m.ivec(4), b))

    @add("NormalizeScalar")
    def _(f, m, a, b, c, d, h):
        return vec4_of(f, m, *[f.ext(m.f32(), "Normalize", comp(f, m, a, i)) for i in range(4)])

    @add("Length")
    def _(f, m, a, b, c, d, h):
        a3 = f.op("VectorShuffle", m.vec(3), a, a, 0, 1, 2)
        a2 = f.op("VectorShuffle", m.vec(2), a, a, 0, 1)
        return vec4_of(f, m, f.ext(m.f32(), "Length", a), f.ext(m.f32(), "Length", a3), f.ext(m.f32(), "Length", a2),
                       f.ext(m.f32(), "Length", comp(f, m, a, 0)))

    @add("Distance")
    def _(f, m, a, b, c, d, h):
        a3 = f.op("VectorShuffle", m.vec(3), a, a, 0, 1, 2)
        b3 = f.op("VectorShuffle", m.vec(3), b, b, 0, 1, 2)
        return vec4_of(f, m, f.ext(m.f32(), "Distance", a, b), f.ext(m.f32(), "Distance", a3, b3),
                       f.ext(m.f32(), "Distance", comp(f, m, a, 0), comp(f, m, b, 0)), m.const_f(0.0))

    @add("Cross")
    def _(f, m, a, b, c, d, h):
        a3 = f.op("VectorShuffle", m.vec(3), a, a, 0, 1, 2)
        b3 = f.op("VectorShuffle", m.vec(3), b, b, 0, 1, 2)
        r = f.ext(m.vec(3), "Cross", a3, b3)
        return vec4_of(f, m, *[f.op("CompositeExtract", m.f32(), r, i) for i in range(3)], m.const_f(0.0))

    @add("Refract")
    def _(f, m, a, b, c, d, h):
        a3 = f.op("VectorShuffle", m.vec(3), a, a, 0, 1, 2)
        b3 = f.ext(m.vec(3), "Normalize", f.op("VectorShuffle", m.vec(3), b, b, 0, 1, 2))
        r = f.ext(m.vec(3), "Refract", a3, b3, comp(f, m, c, 0))
        return vec4_of(f, m, *[f.op("CompositeExtract", m.f32(), r, i) for i in range(3)], m.const_f(0.0))

    @add("Dot")
    def _(f, m, a, b, c, d, h):
        a3 = f.op("VectorShuffle", m.vec(3), a, a, 0, 1, 2)
        b3 = f.op("VectorShuffle", m.vec(3), b, b, 0, 1, 2)
        a2 = f.op("VectorShuffle", m.vec(2), a, a, 0, 1)
        b2 = f.op("VectorShuffle", m.vec(2), b, b, 0, 1)
        return vec4_of(f, m, f.op("Dot", m.f32(), a2, b2), f.op("Dot", m.f32(), a3, b3), f.op("Dot", m.f32(), a, b),
                       m.const_f(0.0))

    @add("FOrdCompare")
    def _(f, m, a, b, c, d, h):
        bv = m.bvec(4)
        lt, eq, ne, ge = (f.op(op, bv, a, b) for op in ("FOrdLessThan", "FOrdEqual", "FOrdNotEqual", "FOrdGreaterThanEqual"))
        # pack four bvec4 results into one vec4: x = sum of lanes of lt, ...
        return vec4_of(f, m, *[f.op("Dot", m.f32(), sel01(f, m, v), m.const_composite(m.vec(4), *[m.const_f(s) for s in (1.0, 2.0, 4.0, 8.0)]))
                               for v in (lt, eq, ne, ge)])

    @add("FUnordCompare")
    def _(f, m, a, b, c, d, h):
        bv = m.bvec(4)
        vs = [f.op(op, bv, a, b) for op in ("FUnordLessThan", "FUnordEqual", "FUnordNotEqual", "FUnordGreaterThanEqual")]
        return vec4_of(f, m, *[f.op("Dot", m.f32(), sel01(f, m, v), m.const_composite(m.vec(4), *[m.const_f(s) for s in (1.0, 2.0, 4.0, 8.0)]))
                               for v in vs])

    @add("IsNanAnyAll")
    def _(f, m, a, b, c, d, h):
        bv = m.bvec(4)
        nan = sel01(f, m, f.op("IsNan", bv, a))
        lt = f.op("FOrdLessThan", bv, a, b)
        anyv = f.op("Select", m.f32(), f.op("Any", m.bool(), lt), m.const_f(1.0), m.const_f(0.0))
        allv = f.op("Select", m.f32(), f.op("All", m.bool(), lt), m.const_f(1.0), m.const_f(0.0))
        inf = f.op("Select", m.f32(), f.op("IsInf", m.bool(), comp(f, m, a, 0)), m.const_f(1.0), m.const_f(0.0))
        w = f.op("Dot", m.f32(), nan, m.const_composite(m.vec(4), *[m.const_f(s) for s in (1.0, 2.0, 4.0, 8.0)]))
        return vec4_of(f, m, w, anyv, allv, inf)

    @add("SelectVector")
    def _(f, m, a, b, c, d, h):
        return f.op("Select", m.vec(4), f.op("FOrdLessThan", m.bvec(4), a, b), c, d)

    def mat4_of(f, m, *cols):
        return f.op("CompositeConstruct", m.mat(4, 4), *cols)

    @add("MatrixTimesVector4")
    def _(f, m, a, b, c, d, h):
        return f.op("MatrixTimesVector", m.vec(4), mat4_of(f, m, a, b, c, d), a)

    @add("MatrixTimesVector3")
    def _(f, m, a, b, c, d, h):
        v3 = m.vec(3)
        sh = lambda v: f.op("VectorShuffle", v3, v, v, 0, 1, 2)
        m3 = f.op("CompositeConstruct", m.mat(3, 3), sh(a), sh(b), sh(c))
        r = f.op("MatrixTimesVector", v3, m3, sh(d))
        return vec4_of(f, m, *[f.op("CompositeExtract", m.f32(), r, i) for i in range(3)], m.const_f(0.0))

    @add("VectorTimesMatrix4")
    def _(f, m, a, b, c, d, h):
        return f.op("VectorTimesMatrix", m.vec(4), a, mat4_of(f, m, a, b, c, d))

    @add("MatrixTimesMatrix4")
    def _(f, m, a, b, c, d, h):
        p = f.op("MatrixTimesMatrix", m.mat(4, 4), mat4_of(f, m, a, b, c, d), mat4_of(f, m, d, c, b, a))
        return f.op("CompositeExtract", m.vec(4), p, 1)

    @add("MatrixTimesScalar4")
    def _(f, m, a, b, c, d, h):
        p = f.op("MatrixTimesScalar", m.mat(4, 4), mat4_of(f, m, a, b, c, d), comp(f, m, a, 0))
        return f.op("CompositeExtract", m.vec(4), p, 2)

    @add("Transpose4")
    def _(f, m, a, b, c, d, h):
        return f.op("CompositeExtract", m.vec(4), f.op("Transpose", m.mat(4, 4), mat4_of(f, m, a, b, c, d)), 0)

    @add("OuterProduct4")
    def _(f, m, a, b, c, d, h):
        return f.op("CompositeExtract", m.vec(4), f.op("OuterProduct", m.mat(4, 4), a, b), 3)

    @add("VectorTimesScalar")
    def _(f, m, a, b, c, d, h):
        return f.op("VectorTimesScalar", m.vec(4), a, comp(f, m, b, 0))

    @add("FunctionCall")
    def _(f, m, a, b, c, d, h):
        # vec4 helper(vec4 x, inout vec4 y) { y = y * 2.0; return x + y; }; result helper(a, t) + t with t = b
        t = f.local(m.vec(4))
        f.store(t, b)
        r = f.op("FunctionCall", m.vec(4), h, a, t)
        return f.op("FAdd", m.vec(4), r, f.load(m.vec(4), t))

    @add("DynamicInsertExtract")
    def _(f, m, a, b, c, d, h):
        i32 = m.i32()
        ib = f.op("BitwiseAnd", i32, f.op("ConvertFToS", i32, comp(f, m, b, 0)), m.const_i(3))
        ic = f.op("BitwiseAnd", i32, f.op("ConvertFToS", i32, comp(f, m, c, 1)), m.const_i(3))
        v = f.op("VectorInsertDynamic", m.vec(4), a, comp(f, m, c, 0), ib)
        x = f.op("VectorExtractDynamic", m.f32(), v, ic)
        w = f.op("CompositeInsert", m.vec(4), x, v, 3)
        return f.op("CopyObject", m.vec(4), w)

    @add("Pack")
    def _(f, m, a, b, c, d, h):
        u32 = m.u32()
        a2 = f.op("VectorShuffle", m.vec(2), a, a, 0, 1)
        a2b = f.op("VectorShuffle", m.vec(2), a, a, 2, 3)
        us = [f.ext(u32, "PackUnorm4x8", a), f.ext(u32, "PackSnorm4x8", a), f.ext(u32, "PackHalf2x16", a2),
              f.ext(u32, "PackUnorm2x16", a2b)]
        return back(f, m, f.op("CompositeConstruct", m.uvec(4), *us))

    @add("Unpack4x8")
    def _(f, m, a, b, c, d, h):
        ua = bits(f, m, a, m.uvec(4))
        x = f.op("CompositeExtract", m.u32(), ua, 0)
        r1 = f.ext(m.vec(4), "UnpackUnorm4x8", x)
        r2 = f.ext(m.vec(4), "UnpackSnorm4x8", x)
        return f.op("VectorShuffle", m.vec(4), r1, r2, 0, 2, 4, 7)

    @add("Unpack2x16")
    def _(f, m, a, b, c, d, h):
        ua = bits(f, m, a, m.uvec(4))
        x = f.op("CompositeExtract", m.u32(), ua, 0)
        y = f.op("CompositeExtract", m.u32(), ua, 1)
        r1 = f.ext(m.vec(2), "UnpackHalf2x16", x)
        r2 = f.ext(m.vec(2), "UnpackUnorm2x16", y)
        r3 = f.ext(m.vec(2), "UnpackSnorm2x16", x)
        t = f.op("VectorShuffle", m.vec(4), r1, r2, 0, 1, 2, 3)
        return f.op("CompositeInsert", m.vec(4), f.op("CompositeExtract", m.f32(), r3, 1), t, 3)

    # integer operations on the bit patterns of a, b, c
    def ibin(op, signed=True):
        def bld(f, m, a, b, c, d, h):
            t = m.ivec(4) if signed else m.uvec(4)
            return back(f, m, f.op(op, t, bits(f, m, a, t), bits(f, m, b, t)))
        return bld

    for op in ("IAdd", "ISub", "IMul", "SDiv", "SRem", "SMod", "BitwiseAnd", "BitwiseOr", "BitwiseXor"):
        add(op)(ibin(op, True))
    for op in ("UDiv", "UMod"):
        add(op)(ibin(op, False))
    for op in ("ShiftLeftLogical", "ShiftRightLogical", "ShiftRightArithmetic"):
        add(op)(ibin(op, op == "ShiftRightArithmetic"))

    def iun(op, signed=True, glsl=False):
        def bld(f, m, a, b, c, d, h):
            t = m.ivec(4) if signed else m.uvec(4)
            x = bits(f, m, a, t)
            return back(f, m, f.ext(t, op, x) if glsl else f.op(op, t, x))
        return bld

    for op in ("SNegate", "Not"):
        add(op)(iun(op, True))
    for op in ("BitReverse", "BitCount"):
        add(op)(iun(op, False))
    for op in ("SAbs", "SSign", "FindILsb", "FindSMsb"):
        add(op)(iun(op, True, True))
    add("FindUMsb")(iun("FindUMsb", False, True))

    def iext2(op, signed):
        def bld(f, m, a, b, c, d, h):
            t = m.ivec(4) if signed else m.uvec(4)
            return back(f, m, f.ext(t, op, bits(f, m, a, t), bits(f, m, b, t)))
        return bld

    def iext3(op, signed):
        def bld(f, m, a, b, c, d, h):
            t = m.ivec(4) if signed else m.uvec(4)
            return back(f, m, f.ext(t, op, bits(f, m, a, t), bits(f, m, b, t), bits(f, m, c, t)))
        return bld

    add("SMin")(iext2("SMin", True))
    add("SMax")(iext2("SMax", True))
    add("UMin")(iext2("UMin", False))
    add("UMax")(iext2("UMax", False))
    add("SClamp")(iext3("SClamp", True))
    add("UClamp")(iext3("UClamp", False))

    def bitfield(op):
        def bld(f, m, a, b, c, d, h):
            t, i32 = m.uvec(4), m.i32()
            cc = bits(f, m, c, m.ivec(4))
            off = f.op("BitwiseAnd", i32, f.op("CompositeExtract", i32, cc, 0), m.const_i(15))
            cnt = f.op("BitwiseAnd", i32, f.op("CompositeExtract", i32, cc, 1), m.const_i(15))
            if op == "BitFieldInsert":
                r = f.op(op, t, bits(f, m, a, t), bits(f, m, b, t), off, cnt)
            elif op == "BitFieldSExtract":
                r = f.op(op, m.ivec(4), bits(f, m, a, m.ivec(4)), off, cnt)
            else:
                r = f.op(op, t, bits(f, m, a, t), off, cnt)
            return back(f, m, r)
        return bld

    for op in ("BitFieldInsert", "BitFieldUExtract", "BitFieldSExtract"):
        add(op)(bitfield(op))

    @add("IntCompare")
    def _(f, m, a, b, c, d, h):
        ia, ib = bits(f, m, a, m.ivec(4)), bits(f, m, b, m.ivec(4))
        ua, ub = bits(f, m, a, m.uvec(4)), bits(f, m, b, m.uvec(4))
        bv = m.bvec(4)
        vs = [f.op("SLessThan", bv, ia, ib), f.op("ULessThan", bv, ua, ub), f.op("IEqual", bv, ia, ib),
              f.op("UGreaterThanEqual", bv, ua, ub)]
        w = m.const_composite(m.vec(4), *[m.const_f(s) for s in (1.0, 2.0, 4.0, 8.0)])
        return vec4_of(f, m, *[f.op("Dot", m.f32(), sel01(f, m, v), w) for v in vs])

    @add("Convert")
    def _(f, m, a, b, c, d, h):
        i32, u32 = m.i32(), m.u32()
        ib, ub = bits(f, m, b, m.ivec(4)), bits(f, m, b, m.uvec(4))
        fs = f.op("ConvertFToS", i32, comp(f, m, a, 0))
        fu = f.op("ConvertFToU", u32, f.ext(m.f32(), "FAbs", comp(f, m, a, 1)))
        sf = f.op("ConvertSToF", m.f32(), f.op("CompositeExtract", i32, ib, 2))
        uf = f.op("ConvertUToF", m.f32(), f.op("CompositeExtract", u32, ub, 3))
        return vec4_of(f, m, f.op("Bitcast", m.f32(), fs), f.op("Bitcast", m.f32(), fu), sf, uf)

    return cases


ALU_CASES = [name for name, _ in _alu_cases()]


def fs_alu_ops() -> list[int]:
    """layout(location=0..3) flat in vec4 a, b, c, d; layout(location=4) flat in int sel;
    layout(location=0) out vec4 outColor;
    vec4 helper(vec4 x, inout vec4 y) { y = y * 2.0; return x + y; }
    void main() { vec4 r = vec4(0); switch (sel) { case k: r = <case k>(a, b, c, d); break; ... } outColor = r; }"""
    m = Module()
    v4, i32 = m.vec(4), m.i32()
    ins = [_io(m, "Input", v4, k, "abcd"[k]) for k in range(4)]
    sel = _io(m, "Input", i32, 4, "sel")
    for v in ins + [sel]:
        m.decorate(v, "Flat")
    out = _io(m, "Output", v4, 0, "outColor")
    # the helper, defined before the entry point
    hf = Function(m, v4, (v4, m.ptr("Function", v4)))
    x, yp = hf.params
    y2 = hf.op("VectorTimesScalar", v4, hf.load(v4, yp), m.const_f(2.0))
    hf.store(yp, y2)
    helper = hf.end(hf.op("FAdd", v4, x, y2))
    f = Function(m)
    r = f.local(v4)
    f.store(r, m.const_composite(v4, *[m.const_f(0.0)] * 4))
    a, b, c, d = (f.load(v4, v) for v in ins)
    s = f.load(i32, sel)
    cases = _alu_cases()
    labels = [m.id() for _ in cases]
    default_l, merge_l = m.id(), m.id()
    f.raw("SelectionMerge", merge_l, 0)
    sw = [s, default_l]
    for k, lab in enumerate(labels):
        sw += [k, lab]
    f.raw("Switch", *sw)
    for (name, bld), lab in zip(cases, labels):
        f.label(lab)
        f.store(r, bld(f, m, a, b, c, d, helper))
        f.raw("Branch", merge_l)
    f.label(default_l)
    f.raw("Branch", merge_l)
    f.label(merge_l)
    f.store(out, f.load(v4, r))
    return m.assemble("Fragment", f.end(), ins + [sel, out], True)


SHADERS["vs_flat_inputs"] = vs_flat_inputs
SHADERS["fs_alu_ops"] = fs_alu_ops
