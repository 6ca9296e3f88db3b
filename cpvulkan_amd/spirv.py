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
    Name=5, ExtInstImport=11, ExtInst=12, MemoryModel=14, EntryPoint=15, ExecutionMode=16, Capability=17,
    TypeVoid=19, TypeBool=20, TypeInt=21, TypeFloat=22, TypeVector=23, TypeMatrix=24, TypeImage=25,
    TypeSampler=26, TypeSampledImage=27, TypeArray=28, TypeStruct=30, TypePointer=32, TypeFunction=33,
    ConstantTrue=41, ConstantFalse=42, Constant=43, ConstantComposite=44,
    Function=54, FunctionParameter=55, FunctionEnd=56, FunctionCall=57, Variable=59, Load=61, Store=62,
    AccessChain=65, Decorate=71, MemberDecorate=72, VectorShuffle=79, CompositeConstruct=80,
    CompositeExtract=81, ImageSampleImplicitLod=87, ConvertFToU=109, ConvertFToS=110, ConvertSToF=111,
    ConvertUToF=112, Bitcast=124, SNegate=126, FNegate=127, IAdd=128, FAdd=129, ISub=130, FSub=131, IMul=132,
    FMul=133, UDiv=134, SDiv=135, FDiv=136, VectorTimesScalar=142, MatrixTimesVector=145, Dot=148,
    LogicalNot=168, Select=169, IEqual=170, INotEqual=171, UGreaterThan=172, SGreaterThan=173, ULessThan=176,
    SLessThan=177, FOrdEqual=180, FOrdNotEqual=182, FOrdLessThan=184, FOrdGreaterThan=186,
    FOrdLessThanEqual=188, FOrdGreaterThanEqual=190, Phi=245, LoopMerge=246, SelectionMerge=247, Label=248,
    Branch=249, BranchConditional=250, Kill=252, Return=253, ReturnValue=254,
)
GLSL = dict(FAbs=4, Floor=8, Fract=10, Sqrt=31, InverseSqrt=32, FMin=37, FMax=40, FClamp=43, Length=66,
            Normalize=69)
DEC = dict(Block=2, ColMajor=5, MatrixStride=7, BuiltIn=11, NoPerspective=13, Flat=14, Location=30, Binding=33,
           DescriptorSet=34, Offset=35)
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
    """Body builder of one function (the entry point)."""

    def __init__(self, m: Module):
        self.m = m
        self.fid = m.id()
        m._ins(m.funcs, "Function", m.void(), self.fid, 0, m.fn(m.void()))
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

    def end(self) -> int:
        self.raw("Return")
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


SHADERS = {
    "vs_passthrough": vs_passthrough,
    "vs_mvp_color": vs_mvp_color,
    "vs_mvp_lambert": vs_mvp_lambert,
    "fs_color": fs_color,
    "fs_tex_lambert": fs_tex_lambert,
    "fs_discard_dark": fs_discard_dark,
}
