// SPIR-V -> HIP translation of a (vertex, fragment) shader pair, compiled
// for gfx950 with hiprtc.
//
// This replaces the reference's shader JIT (LLVMRuntime/SPIRVCompiler.cpp,
// PipelineCompiler.cpp:499-1799, which turn SPIR-V into x86 through LLVM):
// each module becomes a C++ struct with the interface of the built-in
// programs (VS::run / FS::run, see programs_builtin.h), and the source
// instantiates the same vertex and raster kernel templates the library is
// built from, so a translated program runs the identical raster pipeline.
//
// Arithmetic follows the reference's mapping of SPIR-V to host code:
//   OpMatrixTimesVector -> glm mat4*vec4 (SpirvFunctions.cpp:37-47, 78)
//   OpDot               -> glm::dot      (SpirvFunctions.cpp:5-9, 67-69)
//   GLSL.std.450 FMax / FMin / FClamp -> std::max / std::min / std::clamp
//                          (GlslFunctions.cpp:118-142, 1096-1099, 1143)
//   Normalize           -> glm::normalize, v * (1 / sqrt(dot(v, v)))
//   OpImageSampleImplicitLod -> ImageSampleImplicitLod, lambda = 0
//                          (GlslFunctions.cpp:656-671)
// with every float operation rounding once (the kernels are compiled with
// contraction off and correctly rounded division and square root).
//
// Supported: logical addressing; 32-bit float, int and uint scalars and
// vectors, booleans, column-major float matrices, arrays and structs
// (function variables, constants, OpConstantNull, specialization
// constants at their defaults); uniform blocks of set 0, bindings 0-3, with
// nested structs, arrays and matrices at their declared Offset /
// ArrayStride / MatrixStride; combined image samplers of set 0, bindings
// 0-3 (implicit and explicit-LOD sampling, OpImageFetch, size and level
// queries); vertex inputs, float and Flat integer varyings by location,
// Flat / NoPerspective interpolation; gl_PerVertex Position and
// PointSize, VertexIndex, InstanceIndex, FragCoord, FrontFacing; integer
// and float fragment outputs at location 0; function calls (each function
// a lambda, callees first, no recursion); structured or unstructured
// control flow (labels become goto targets, OpPhi assignments on the
// incoming edges, OpSwitch); OpKill; the core arithmetic, bit, logical,
// relational, conversion, composite and matrix instructions with LLVM's
// integer semantics; GLSL.std.450 instructions 1-81 except Determinant,
// MatrixInverse, Modf / ModfStruct, IMix, Frexp / FrexpStruct, the double
// packing pair and InterpolateAt*.  Anything else is reported as
// unsupported with what it was.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <functional>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cpv.h"
#include "cpv_program.h"
#include "cpv_types.h"

extern const char* const cpv_rtc_header_names[];
extern const char* const cpv_rtc_header_sources[];
extern const int cpv_rtc_header_count;

namespace {

enum : uint32_t {
    OpName = 5, OpExtInstImport = 11, OpExtInst = 12, OpMemoryModel = 14, OpEntryPoint = 15, OpExecutionMode = 16,
    OpCapability = 17, OpTypeVoid = 19, OpTypeBool = 20, OpTypeInt = 21, OpTypeFloat = 22, OpTypeVector = 23,
    OpTypeMatrix = 24, OpTypeImage = 25, OpTypeSampler = 26, OpTypeSampledImage = 27, OpTypeArray = 28,
    OpTypeStruct = 30, OpTypePointer = 32, OpTypeFunction = 33, OpConstantTrue = 41, OpConstantFalse = 42,
    OpConstant = 43, OpConstantComposite = 44, OpFunction = 54, OpFunctionParameter = 55, OpFunctionEnd = 56,
    OpFunctionCall = 57, OpVariable = 59, OpLoad = 61, OpStore = 62, OpAccessChain = 65, OpDecorate = 71,
    OpMemberDecorate = 72, OpVectorShuffle = 79, OpCompositeConstruct = 80, OpCompositeExtract = 81,
    OpImageSampleImplicitLod = 87, OpConvertFToU = 109, OpConvertFToS = 110, OpConvertSToF = 111,
    OpConvertUToF = 112, OpBitcast = 124, OpSNegate = 126, OpFNegate = 127, OpIAdd = 128, OpFAdd = 129,
    OpISub = 130, OpFSub = 131, OpIMul = 132, OpFMul = 133, OpUDiv = 134, OpSDiv = 135, OpFDiv = 136,
    OpVectorTimesScalar = 142, OpMatrixTimesVector = 145, OpDot = 148, OpLogicalOr = 166, OpLogicalAnd = 167,
    OpLogicalNot = 168, OpSelect = 169, OpIEqual = 170, OpINotEqual = 171, OpUGreaterThan = 172,
    OpSGreaterThan = 173, OpUGreaterThanEqual = 174, OpSGreaterThanEqual = 175, OpULessThan = 176,
    OpSLessThan = 177, OpULessThanEqual = 178, OpSLessThanEqual = 179, OpFOrdEqual = 180, OpFOrdNotEqual = 182,
    OpFOrdLessThan = 184, OpFOrdGreaterThan = 186, OpFOrdLessThanEqual = 188, OpFOrdGreaterThanEqual = 190,
    OpPhi = 245, OpLoopMerge = 246, OpSelectionMerge = 247, OpLabel = 248, OpBranch = 249,
    OpBranchConditional = 250, OpKill = 252, OpReturn = 253, OpReturnValue = 254, OpSource = 3,
    OpSourceExtension = 4, OpMemberName = 6, OpString = 7, OpLine = 8, OpNoLine = 317, OpModuleProcessed = 330,
    OpUndef = 1, OpExtension = 10, OpConstantNull = 46, OpSpecConstantTrue = 48, OpSpecConstantFalse = 49,
    OpSpecConstant = 50, OpSpecConstantComposite = 51, OpVectorExtractDynamic = 77, OpVectorInsertDynamic = 78,
    OpCompositeInsert = 82, OpCopyObject = 83, OpTranspose = 84, OpSampledImage = 86, OpImageSampleExplicitLod = 88,
    OpImageFetch = 95, OpImage = 100, OpImageQuerySizeLod = 103, OpImageQuerySize = 104, OpImageQueryLevels = 106,
    OpUConvert = 113, OpSConvert = 114, OpFConvert = 115, OpQuantizeToF16 = 116, OpUMod = 137, OpSRem = 138,
    OpSMod = 139, OpFRem = 140, OpFMod = 141, OpMatrixTimesScalar = 143, OpVectorTimesMatrix = 144,
    OpMatrixTimesMatrix = 146, OpOuterProduct = 147, OpAny = 154, OpAll = 155, OpIsNan = 156, OpIsInf = 157,
    OpLogicalEqual = 164, OpLogicalNotEqual = 165, OpFUnordEqual = 181, OpFUnordNotEqual = 183,
    OpFUnordLessThan = 185, OpFUnordGreaterThan = 187, OpFUnordLessThanEqual = 189,
    OpFUnordGreaterThanEqual = 191, OpShiftRightLogical = 194, OpShiftRightArithmetic = 195,
    OpShiftLeftLogical = 196, OpBitwiseOr = 197, OpBitwiseXor = 198, OpBitwiseAnd = 199, OpNot = 200,
    OpBitFieldInsert = 201, OpBitFieldSExtract = 202, OpBitFieldUExtract = 203, OpBitReverse = 204, OpBitCount = 205,
    OpSwitch = 251, OpUnreachable = 255, OpDecorationGroup = 73, OpGroupDecorate = 74,
};

enum : uint32_t {
    DecBlock = 2, DecColMajor = 5, DecMatrixStride = 7, DecBuiltIn = 11, DecNoPerspective = 13, DecFlat = 14,
    DecLocation = 30, DecBinding = 33, DecDescriptorSet = 34, DecOffset = 35, DecArrayStride = 6,
};
enum : uint32_t { BiPosition = 0, BiPointSize = 1, BiClipDistance = 3, BiCullDistance = 4, BiFragCoord = 15,
                  BiFrontFacing = 17, BiVertexIndex = 42, BiInstanceIndex = 43 };
enum : uint32_t { ScUniformConstant = 0, ScInput = 1, ScUniform = 2, ScOutput = 3, ScPrivate = 6, ScFunction = 7 };

struct Fail {
    std::string msg;
};

[[noreturn]] void bad(const std::string& m) { throw Fail{m}; }

struct Type {
    uint32_t op = 0;
    uint32_t a = 0, b = 0;              // int: width, signedness; vector/matrix: element/column type, count;
                                        // array: element type, length constant; pointer: storage class, pointee
    std::vector<uint32_t> members;      // struct
};

struct MemberDec {
    int offset = -1, builtin = -1, matrix_stride = -1;
};

struct Decor {
    int location = -1, binding = -1, set = -1, builtin = -1, array_stride = -1;
    bool flat = false, noperspective = false, block = false;
    std::map<uint32_t, MemberDec> members;
};

struct Inst {
    uint32_t op;
    std::vector<uint32_t> w;   // operands
};

struct Function {
    uint32_t ret = 0;                    // result type
    std::vector<Inst> body;              // parameters, then blocks
};

struct Module {
    uint32_t model = 0xFFFFFFFFu;
    uint32_t entry = 0;
    bool origin_upper_left = false;
    uint32_t glsl = 0;
    std::map<uint32_t, Type> types;
    std::map<uint32_t, Decor> dec;
    std::map<uint32_t, std::pair<uint32_t, std::vector<uint32_t>>> consts;   // id -> (type, words or constituents)
    std::map<uint32_t, bool> const_composite;
    std::set<uint32_t> const_null;                                           // OpConstantNull / module-level OpUndef
    std::map<uint32_t, std::pair<uint32_t, uint32_t>> globals;               // var id -> (pointer type, storage class)
    std::map<uint32_t, uint32_t> global_init;                                // Private variable -> initializer
    std::map<uint32_t, Function> functions;                                  // every function, by id
};

std::string lower(const char* s) { return s ? std::string(s) : std::string("main"); }

Module parse(const uint32_t* w, size_t n, const char* entry_name)
{
    if (n < 5 || w[0] != 0x07230203u) bad("not a SPIR-V module (magic)");
    Module m;
    size_t i = 5;
    bool in_fn = false, found = false;
    const std::string want = lower(entry_name);
    std::vector<Inst> fn_insts;
    uint32_t fn_id = 0, fn_ret = 0;
    while (i < n) {
        const uint32_t len = w[i] >> 16, op = w[i] & 0xFFFFu;
        if (len == 0 || i + len > n) bad("truncated instruction");
        Inst in{op, std::vector<uint32_t>(w + i + 1, w + i + len)};
        i += len;
        if (in_fn) {
            if (op == OpFunctionEnd) {
                m.functions[fn_id] = Function{fn_ret, fn_insts};
                fn_insts.clear();
                in_fn = false;
            } else {
                fn_insts.push_back(in);
            }
            continue;
        }
        const std::vector<uint32_t>& o = in.w;
        switch (op) {
        case OpCapability: case OpMemoryModel: case OpName: case OpMemberName: case OpSource: case OpSourceExtension:
        case OpString: case OpLine: case OpNoLine: case OpModuleProcessed: case OpExtension:
            break;
        case OpExtInstImport: {
            const char* s = (const char*)&o[1];
            if (std::strncmp(s, "GLSL.std.450", 12) != 0) bad(std::string("extended instruction set ") + s);
            m.glsl = o[0];
            break;
        }
        case OpEntryPoint: {
            const char* s = (const char*)&o[2];
            if (want == s) {
                if (found) bad("entry point '" + want + "' declared twice");
                m.model = o[0];
                m.entry = o[1];
                found = true;
            }
            break;
        }
        case OpExecutionMode:
            if (o[0] == m.entry && o[1] == 7) m.origin_upper_left = true;
            break;
        case OpDecorate: {
            Decor& d = m.dec[o[0]];
            switch (o[1]) {
            case DecLocation: d.location = (int)o[2]; break;
            case DecBinding: d.binding = (int)o[2]; break;
            case DecDescriptorSet: d.set = (int)o[2]; break;
            case DecBuiltIn: d.builtin = (int)o[2]; break;
            case DecFlat: d.flat = true; break;
            case DecNoPerspective: d.noperspective = true; break;
            case DecBlock: d.block = true; break;
            case DecArrayStride: d.array_stride = (int)o[2]; break;
            default: break;
            }
            break;
        }
        case OpMemberDecorate: {
            MemberDec& md = m.dec[o[0]].members[o[1]];
            if (o[2] == DecOffset) md.offset = (int)o[3];
            else if (o[2] == DecBuiltIn) md.builtin = (int)o[3];
            else if (o[2] == DecMatrixStride) md.matrix_stride = (int)o[3];
            break;
        }
        case OpTypeVoid: case OpTypeBool: case OpTypeSampler:
            m.types[o[0]] = Type{op, 0, 0, {}};
            break;
        case OpTypeInt: case OpTypeVector: case OpTypeMatrix: case OpTypeArray: case OpTypePointer:
            m.types[o[0]] = Type{op, o[1], o[2], {}};
            break;
        case OpTypeFloat:
            if (o[1] != 32) bad("only 32-bit floats are supported");
            m.types[o[0]] = Type{op, o[1], 0, {}};
            break;
        case OpTypeImage: case OpTypeSampledImage:
            m.types[o[0]] = Type{op, o[1], 0, {}};
            break;
        case OpTypeStruct:
            m.types[o[0]] = Type{op, 0, 0, std::vector<uint32_t>(o.begin() + 1, o.end())};
            break;
        case OpTypeFunction:
            m.types[o[0]] = Type{op, o[1], 0, {}};
            break;
        // Specialization constants take their default values: programs are
        // created without specialization data.
        case OpConstant: case OpSpecConstant:
            m.consts[o[1]] = {o[0], {o[2]}};
            break;
        case OpConstantTrue: case OpConstantFalse: case OpSpecConstantTrue: case OpSpecConstantFalse:
            m.consts[o[1]] = {o[0], {(op == OpConstantTrue || op == OpSpecConstantTrue) ? 1u : 0u}};
            break;
        case OpConstantComposite: case OpSpecConstantComposite:
            m.consts[o[1]] = {o[0], std::vector<uint32_t>(o.begin() + 2, o.end())};
            m.const_composite[o[1]] = true;
            break;
        case OpConstantNull: case OpUndef:
            m.consts[o[1]] = {o[0], {}};
            m.const_null.insert(o[1]);
            break;
        case OpVariable:
            m.globals[o[1]] = {o[0], o[2]};
            if (o.size() > 3) m.global_init[o[1]] = o[3];
            break;
        case OpFunction:
            in_fn = true;
            fn_ret = o[0];
            fn_id = o[1];
            break;
        default: {
            std::ostringstream e;
            e << "unsupported module-level instruction, opcode " << op;
            bad(e.str());
        }
        }
    }
    if (!found) bad("entry point '" + want + "' not found");
    if (m.model != 0 && m.model != 4) bad("only vertex and fragment entry points are translated");
    if (!m.functions.count(m.entry)) bad("entry function body missing");
    for (const Inst& in : m.functions[m.entry].body)
        if (in.op == OpFunctionParameter) bad("entry function takes parameters");
    return m;
}

// ------------------------------------------------------------ code generation

struct UPtr {
    int binding = 0;
    std::string off;           // byte offset expression from the start of the buffer
    int mstride = 16;          // MatrixStride of the member the pointer is inside
};

struct Gen {
    const Module& m;
    bool vertex;
    std::ostringstream types, decl, fns, pro, epi, code;
    std::ostringstream* od = &decl;          // declarations of the function being generated
    std::ostringstream* oc = &code;          // statements of the function being generated
    bool in_entry = true;
    std::map<uint32_t, std::string> lval;    // pointer id -> lvalue expression
    std::map<uint32_t, uint32_t> ptr_pointee;
    std::map<uint32_t, UPtr> uptr;           // pointers into uniform buffers
    std::set<uint32_t> structs_done;
    int out_total = 0;                       // vertex: varying components written
    int psize = -1;                          // vertex: output component holding gl_PointSize
    std::string psize_expr;                  // vertex: the variable written as gl_PointSize
    std::map<int, int> loc_offset;           // location -> first varying component
    std::map<int, int> loc_size;
    bool discard = false;
    std::vector<int> interp;                 // fragment: mode per input component
    std::map<uint32_t, uint32_t> rtype;      // value id -> type id, every function

    Gen(const Module& mm, bool vs) : m(mm), vertex(vs)
    {
        // result types of every instruction, so a value may be used in a
        // block that precedes its definition in the module's block order
        for (const auto& f : m.functions)
            for (const Inst& in : f.second.body)
                if (has_result(in.op) && in.w.size() >= 2) rtype[in.w[1]] = in.w[0];
    }

    static bool has_result(uint32_t op)
    {
        switch (op) {
        case OpLabel: case OpStore: case OpBranch: case OpBranchConditional: case OpSwitch: case OpReturn:
        case OpReturnValue: case OpKill: case OpUnreachable: case OpSelectionMerge: case OpLoopMerge: case OpLine:
        case OpNoLine:
            return false;
        default:
            return true;
        }
    }

    const Type& T(uint32_t id) const
    {
        auto it = m.types.find(id);
        if (it == m.types.end()) bad("unknown type id");
        return it->second;
    }

    int comps(uint32_t t) const
    {
        const Type& ty = T(t);
        if (ty.op == OpTypeVector) return (int)ty.b;
        if (ty.op == OpTypeFloat || ty.op == OpTypeInt || ty.op == OpTypeBool) return 1;
        bad("not a scalar or vector type");
    }

    bool is_vec(uint32_t t) const { return T(t).op == OpTypeVector; }
    uint32_t scalar_of(uint32_t t) const { return is_vec(t) ? T(t).a : t; }
    bool is_float(uint32_t t) const { return T(scalar_of(t)).op == OpTypeFloat; }
    bool is_int(uint32_t t) const { return T(scalar_of(t)).op == OpTypeInt; }

    std::string ctype(uint32_t t)
    {
        const Type& ty = T(t);
        switch (ty.op) {
        case OpTypeVoid: return "void";
        case OpTypeBool: return "bool";
        case OpTypeFloat: return "float";
        case OpTypeInt:
            if (ty.a != 32) bad("only 32-bit integers are supported");
            return ty.b ? "int32_t" : "uint32_t";
        case OpTypeVector: return "V<" + ctype(ty.a) + ", " + std::to_string(ty.b) + ">";
        case OpTypeMatrix: {
            const Type& col = T(ty.a);
            if (T(col.a).op != OpTypeFloat) bad("non-float matrix");
            return "M<" + std::to_string(col.b) + ", " + std::to_string(ty.b) + ">";
        }
        case OpTypeArray: return "A<" + ctype(ty.a) + ", " + std::to_string(const_u32(ty.b)) + ">";
        case OpTypeStruct: {
            const std::string name = "S" + std::to_string(t);
            if (!structs_done.count(t)) {
                structs_done.insert(t);
                std::string body;
                for (size_t k = 0; k < ty.members.size(); k++) body += "        " + ctype(ty.members[k]) + " m" + std::to_string(k) + ";\n";
                types << "    struct " << name << " {\n" << body << "    };\n";
            }
            return name;
        }
        case OpTypeImage: case OpTypeSampledImage: return "const TextureDesc*";
        default: bad("unsupported value type");
        }
    }

    uint32_t const_u32(uint32_t id) const
    {
        auto it = m.consts.find(id);
        if (it == m.consts.end() || m.const_composite.count(id)) bad("expected a scalar constant");
        if (m.const_null.count(id)) return 0;
        return it->second.second[0];
    }

    std::string val(uint32_t id)
    {
        auto it = m.consts.find(id);
        if (it != m.consts.end()) return const_expr(id);
        return "r" + std::to_string(id);
    }

    std::string const_expr(uint32_t id)
    {
        const auto& c = m.consts.at(id);
        const Type& ty = T(c.first);
        if (m.const_null.count(id)) return ctype(c.first) + "{}";
        if (m.const_composite.count(id)) {
            std::string s;
            for (size_t i = 0; i < c.second.size(); i++) s += (i ? ", " : "") + const_expr(c.second[i]);
            return ty.op == OpTypeStruct ? ctype(c.first) + "{" + s + "}" : ctype(c.first) + "{{" + s + "}}";
        }
        const uint32_t v = c.second[0];
        char buf[64];
        if (ty.op == OpTypeFloat) {
            std::snprintf(buf, sizeof(buf), "__uint_as_float(0x%08xu)", v);
            return buf;
        }
        if (ty.op == OpTypeBool) return v ? "true" : "false";
        if (ty.op == OpTypeInt) return ty.b ? "(int32_t)" + std::to_string((int32_t)v) : std::to_string(v) + "u";
        bad("unsupported constant type");
    }

    // element access into a value of type t
    std::string elem(uint32_t t, const std::string& base, const std::string& idx, uint32_t* et) const
    {
        const Type& ty = T(t);
        if (ty.op == OpTypeVector) { *et = ty.a; return base + ".c[" + idx + "]"; }
        if (ty.op == OpTypeMatrix) { *et = ty.a; return base + ".col[" + idx + "]"; }
        if (ty.op == OpTypeArray) { *et = ty.a; return base + ".e[" + idx + "]"; }
        if (ty.op == OpTypeStruct) {
            const uint32_t k = (uint32_t)std::stoul(idx);
            if (k >= ty.members.size()) bad("struct member index");
            *et = ty.members[k];
            return base + ".m" + idx;
        }
        bad("element access into a non-composite");
    }

    // element access with an index value id (constant for structs)
    std::string elem_id(uint32_t t, const std::string& base, uint32_t idx, uint32_t* et)
    {
        if (T(t).op == OpTypeStruct) return elem(t, base, std::to_string(const_u32(idx)), et);
        return elem(t, base, val(idx), et);
    }

    // a value of type t laid out in uniform buffer `b` at byte offset `off`
    // (Offset, ArrayStride and MatrixStride decorations)
    std::string uniform_load(uint32_t t, int b, const std::string& off, int mstride)
    {
        const Type& ty = T(t);
        const std::string base = "env.uniform[" + std::to_string(b) + "]";
        switch (ty.op) {
        case OpTypeFloat: case OpTypeInt: case OpTypeVector:
            if (T(scalar_of(t)).op == OpTypeBool) bad("bool in a uniform block");
            return "(*(const " + ctype(t) + "*)(" + base + " + (" + off + ")))";
        case OpTypeBool:
            return "((*(const uint32_t*)(" + base + " + (" + off + "))) != 0u)";
        case OpTypeMatrix: {
            std::string s = ctype(t) + "{{";
            for (uint32_t c = 0; c < ty.b; c++)
                s += (c ? ", " : "") + uniform_load(ty.a, b, "(" + off + ") + " + std::to_string(c * mstride), mstride);
            return s + "}}";
        }
        case OpTypeArray: {
            auto d = m.dec.find(t);
            if (d == m.dec.end() || d->second.array_stride <= 0) bad("uniform array without an ArrayStride");
            std::string s = ctype(t) + "{{";
            for (uint32_t i = 0; i < const_u32(ty.b); i++)
                s += (i ? ", " : "") + uniform_load(ty.a, b, "(" + off + ") + " + std::to_string(i * (uint32_t)d->second.array_stride), mstride);
            return s + "}}";
        }
        case OpTypeStruct: {
            std::string s = ctype(t) + "{";
            for (size_t k = 0; k < ty.members.size(); k++) {
                const MemberDec md = member_dec(t, (uint32_t)k);
                s += (k ? ", " : "") + uniform_load(ty.members[k], b, "(" + off + ") + " + std::to_string(md.offset),
                                                    md.matrix_stride > 0 ? md.matrix_stride : 16);
            }
            return s + "}";
        }
        default: bad("unsupported type in a uniform block");
        }
    }

    // True when some function stores to `var` (member < 0) or takes an
    // access chain to its member `member`: the shader writes that output.
    bool writes(uint32_t var, int member) const
    {
        for (const auto& f : m.functions)
            for (const Inst& in : f.second.body) {
                if (member < 0 && in.op == OpStore && !in.w.empty() && in.w[0] == var) return true;
                if (member >= 0 && in.op == OpAccessChain && in.w.size() >= 4 && in.w[2] == var) {
                    auto c = m.consts.find(in.w[3]);
                    if (c != m.consts.end() && !c->second.second.empty() && c->second.second[0] == (uint32_t)member) return true;
                }
            }
        return false;
    }

    MemberDec member_dec(uint32_t st, uint32_t k) const
    {
        auto d = m.dec.find(st);
        if (d == m.dec.end()) bad("uniform struct without member offsets");
        auto md = d->second.members.find(k);
        if (md == d->second.members.end() || md->second.offset < 0) bad("uniform member without an Offset");
        return md->second;
    }

    void globals()
    {
        // varying layout: locations in increasing order, components packed
        std::vector<std::pair<int, uint32_t>> io;   // (location * 2 + is_output, var id)
        for (const auto& g : m.globals) {
            const uint32_t var = g.first, ptype = g.second.first, sc = g.second.second;
            const uint32_t pointee = T(ptype).b;
            ptr_pointee[var] = pointee;
            auto dit = m.dec.find(var);
            const Decor d = dit == m.dec.end() ? Decor{} : dit->second;
            const std::string name = "g" + std::to_string(var);
            if (sc == ScInput || sc == ScOutput) {
                const Type& pt = T(pointee);
                if (pt.op == OpTypeStruct) {
                    // gl_PerVertex
                    auto sd = m.dec.find(pointee);
                    if (sc != ScOutput || !vertex || sd == m.dec.end()) bad("input / output blocks are not supported");
                    for (size_t k = 0; k < pt.members.size(); k++) {
                        auto md = sd->second.members.find((uint32_t)k);
                        const int bi = md == sd->second.members.end() ? -1 : md->second.builtin;
                        const std::string mn = name + "_m" + std::to_string(k);
                        decl << "    " << ctype(pt.members[k]) << " " << mn << "{};\n";
                        if (bi == BiPosition) epi << "    pos[0] = " << mn << ".c[0]; pos[1] = " << mn << ".c[1]; pos[2] = " << mn
                                                  << ".c[2]; pos[3] = " << mn << ".c[3];\n";
                        else if (bi == BiPointSize) { if (writes(var, (int)k)) psize_expr = mn; }
                        else if (bi != BiClipDistance && bi != BiCullDistance)
                            bad("unsupported gl_PerVertex member");
                    }
                    lval[var] = name;
                    continue;
                }
                if (d.builtin >= 0) {
                    decl << "    " << ctype(pointee) << " " << name << "{};\n";
                    lval[var] = name;
                    if (vertex && sc == ScInput && d.builtin == BiVertexIndex) pro << "    " << name << " = (int32_t)vid;\n";
                    else if (vertex && sc == ScInput && d.builtin == BiInstanceIndex) pro << "    " << name << " = (int32_t)iid;\n";
                    else if (vertex && sc == ScOutput && d.builtin == BiPosition)
                        epi << "    pos[0] = " << name << ".c[0]; pos[1] = " << name << ".c[1]; pos[2] = " << name << ".c[2]; pos[3] = "
                            << name << ".c[3];\n";
                    else if (vertex && sc == ScOutput && d.builtin == BiPointSize) { if (writes(var, -1)) psize_expr = name; }
                    else if (!vertex && sc == ScInput && d.builtin == BiFragCoord)
                        pro << "    " << name << " = V<float, 4>{{fc[0], fc[1], fc[2], fc[3]}};\n";
                    else if (!vertex && sc == ScInput && d.builtin == BiFrontFacing) pro << "    " << name << " = front;\n";
                    else bad("unsupported built-in variable");
                    continue;
                }
                if (d.location < 0) bad("interface variable without a location");
                const uint32_t et = scalar_of(pointee);
                if (T(et).op != OpTypeFloat && T(et).op != OpTypeInt) bad("only float and 32-bit integer interface variables are supported");
                if (T(et).op == OpTypeInt && vertex && sc == ScInput) bad("integer vertex inputs are not supported");
                if (T(et).op == OpTypeInt && sc == ScInput && !d.flat) bad("integer fragment inputs must be Flat");
                decl << "    " << ctype(pointee) << " " << name << "{};\n";
                lval[var] = name;
                io.push_back({d.location * 2 + (sc == ScOutput ? 1 : 0), var});
                continue;
            }
            if (sc == ScUniform) {
                if (d.set > 0 || d.binding < 0 || d.binding > 3) bad("uniform blocks must be set 0, binding 0-3");
                if (T(pointee).op != OpTypeStruct) bad("uniform variable is not a block");
                uptr[var] = UPtr{d.binding, "0", 16};
                continue;
            }
            if (sc == ScUniformConstant) {
                if (T(pointee).op != OpTypeSampledImage) bad("only combined image samplers are supported as UniformConstant");
                if (d.set > 0 || d.binding < 0 || d.binding > 3) bad("samplers must be set 0, binding 0-3");
                decl << "    const TextureDesc* " << name << " = &env.tex[" << d.binding << "];\n";
                lval[var] = name;
                continue;
            }
            if (sc == ScPrivate) {
                decl << "    " << ctype(pointee) << " " << name << "{};\n";
                auto init = m.global_init.find(var);
                if (init != m.global_init.end()) pro << "    " << name << " = " << val(init->second) << ";\n";
                lval[var] = name;
                continue;
            }
            bad("unsupported storage class");
        }
        std::sort(io.begin(), io.end());
        int off = 0;
        for (const auto& e : io) {
            const uint32_t var = e.second;
            const uint32_t pointee = ptr_pointee[var];
            const int n = comps(pointee);
            const int loc = e.first / 2;
            const bool out = e.first & 1;
            const bool integer = is_int(pointee);
            const std::string name = "g" + std::to_string(var);
            auto sub = [&](int k) { return name + (is_vec(pointee) ? ".c[" + std::to_string(k) + "]" : ""); };
            // integer varyings travel bit for bit in the float varying array
            // (flat interpolation copies the provoking vertex's words)
            auto store = [&](int k) { return integer ? "__uint_as_float((uint32_t)" + sub(k) + ")" : sub(k); };
            if (vertex && !out) {
                // vertex attribute of `n` float components (EmitCopyInput, PipelineCompiler.cpp:821-896)
                if (is_vec(pointee)) pro << "    fetch_attribute(env, " << loc << ", " << n << ", vid, iid, " << name << ".c);\n";
                else pro << "    { float t_[1]; fetch_attribute(env, " << loc << ", 1, vid, iid, t_); " << name << " = t_[0]; }\n";
            } else if (vertex && out) {
                loc_offset[loc] = off;
                loc_size[loc] = n;
                for (int k = 0; k < n; k++) epi << "    out[" << off + k << "] = " << store(k) << ";\n";
                off += n;
            } else if (!vertex && !out) {
                loc_offset[loc] = -1;
                loc_size[loc] = n;
            } else {
                // @setPixelF32 / U32 / I32 by the output's base type (PipelineCompiler.cpp:1528-1653):
                // integer outputs reach an integer attachment as raw words
                if (loc != 0) bad("fragment outputs other than location 0 are not supported");
                if (n != 4) bad("the fragment colour output must have four components");
                for (int k = 0; k < 4; k++) epi << "    out[" << k << "] = " << store(k) << ";\n";
            }
        }
        // gl_PointSize travels after the varyings (the size ProcessPoints
        // reads from the vertex's built-in outputs, CommandBuffer.Draw.cpp:1350)
        if (vertex && !psize_expr.empty()) {
            epi << "    out[" << off << "] = " << psize_expr << ";\n";
            psize = off++;
        }
        out_total = off;
    }

    // fragment inputs laid out at the vertex shader's varying offsets
    void bind_fragment_inputs(const std::map<int, int>& vs_offset, const std::map<int, int>& vs_size)
    {
        int total = 0;
        for (const auto& g : m.globals) {
            const uint32_t var = g.first, sc = g.second.second;
            auto dit = m.dec.find(var);
            if (sc != ScInput || dit == m.dec.end() || dit->second.location < 0 || dit->second.builtin >= 0) continue;
            const Decor& d = dit->second;
            const uint32_t pointee = ptr_pointee[var];
            const int n = comps(pointee);
            auto it = vs_offset.find(d.location);
            if (it == vs_offset.end()) bad("fragment input location " + std::to_string(d.location) + " has no vertex output");
            if (vs_size.at(d.location) < n) bad("fragment input wider than the vertex output at its location");
            const int off = it->second;
            const std::string name = "g" + std::to_string(var);
            const bool integer = is_int(pointee);
            const std::string et = ctype(scalar_of(pointee));
            for (int k = 0; k < n; k++) {
                const std::string src = "in[" + std::to_string(off + k) + "]";
                pro << "    " << name << (is_vec(pointee) ? ".c[" + std::to_string(k) + "]" : "") << " = "
                    << (integer ? "(" + et + ")__float_as_uint(" + src + ")" : src) << ";\n";
            }
            total = std::max(total, off + n);
            if ((int)interp.size() < off + n) interp.resize(off + n, 0);
            for (int k = 0; k < n; k++) interp[off + k] = d.flat ? 2 : d.noperspective ? 1 : 0;
        }
        if ((int)interp.size() < total) interp.resize(total, 0);
    }

    // access chain: lvalue expression and pointee type
    std::string chain(uint32_t base, const std::vector<uint32_t>& idx, uint32_t* out_t)
    {
        auto it = lval.find(base);
        if (it == lval.end()) bad("access chain on an unknown pointer");
        std::string s = it->second;
        uint32_t t = ptr_pointee.at(base);
        size_t i = 0;
        const Type& bt = T(t);
        if (bt.op == OpTypeStruct && m.dec.count(t) && !m.dec.at(t).members.empty() &&
            m.dec.at(t).members.begin()->second.builtin >= 0) {
            // gl_PerVertex members are separate variables
            if (idx.empty()) bad("access to a whole block");
            const uint32_t mi = const_u32(idx[0]);
            s = s + "_m" + std::to_string(mi);
            t = bt.members.at(mi);
            i = 1;
        }
        for (; i < idx.size(); i++) s = elem_id(t, s, idx[i], &t);
        *out_t = t;
        return s;
    }

    // access chain into a uniform buffer: byte offset arithmetic
    UPtr uchain(const UPtr& base, uint32_t t, const std::vector<uint32_t>& idx, uint32_t* out_t)
    {
        UPtr p = base;
        for (uint32_t ix : idx) {
            const Type& ty = T(t);
            if (ty.op == OpTypeStruct) {
                const uint32_t k = const_u32(ix);
                if (k >= ty.members.size()) bad("uniform member index");
                const MemberDec md = member_dec(t, k);
                p.off = "(" + p.off + " + " + std::to_string(md.offset) + ")";
                if (md.matrix_stride > 0) p.mstride = md.matrix_stride;
                t = ty.members[k];
            } else if (ty.op == OpTypeArray) {
                auto d = m.dec.find(t);
                if (d == m.dec.end() || d->second.array_stride <= 0) bad("uniform array without an ArrayStride");
                p.off = "(" + p.off + " + " + std::to_string(d->second.array_stride) + "u * (uint32_t)(" + val(ix) + "))";
                t = ty.a;
            } else if (ty.op == OpTypeMatrix) {
                p.off = "(" + p.off + " + " + std::to_string(p.mstride) + "u * (uint32_t)(" + val(ix) + "))";
                t = ty.a;
            } else if (ty.op == OpTypeVector) {
                p.off = "(" + p.off + " + 4u * (uint32_t)(" + val(ix) + "))";
                t = ty.a;
            } else {
                bad("access chain into a uniform scalar");
            }
        }
        *out_t = t;
        return p;
    }

    // component k of value id (scalar values broadcast)
    std::string comp(uint32_t id, uint32_t t, int k)
    {
        return is_vec(t) ? val(id) + ".c[" + std::to_string(k) + "]" : val(id);
    }

    std::string res(uint32_t id, uint32_t t, int k) const
    {
        return "r" + std::to_string(id) + (is_vec(t) ? ".c[" + std::to_string(k) + "]" : "");
    }

    uint32_t type_of(uint32_t id) const
    {
        auto c = m.consts.find(id);
        if (c != m.consts.end()) return c->second.first;
        auto r = rtype.find(id);
        if (r != rtype.end()) return r->second;
        bad("value " + std::to_string(id) + " has no type");
    }

    // Binary operations with the reference's LLVM semantics
    // (SPIRVCompiler.cpp:2620-3091): integer arithmetic wraps (done in
    // uint32_t), signed / unsigned division, remainder and comparison
    // reinterpret the operand words, shift amounts are taken mod 32.
    std::string binexpr(uint32_t op, const std::string& et, const std::string& a, const std::string& b)
    {
        auto U = [](const std::string& x) { return "(uint32_t)(" + x + ")"; };
        auto S = [](const std::string& x) { return "(int32_t)(" + x + ")"; };
        auto as = [&](const std::string& x) { return "(" + et + ")(" + x + ")"; };
        switch (op) {
        case OpFAdd: return a + " + " + b;
        case OpFSub: return a + " - " + b;
        case OpFMul: return a + " * " + b;
        case OpFDiv: return a + " / " + b;
        case OpFRem: return "fmodf(" + a + ", " + b + ")";
        case OpFMod: return "spv_fmod(" + a + ", " + b + ")";
        case OpIAdd: return as(U(a) + " + " + U(b));
        case OpISub: return as(U(a) + " - " + U(b));
        case OpIMul: return as(U(a) + " * " + U(b));
        case OpUDiv: return as(U(a) + " / " + U(b));
        case OpSDiv: return as(S(a) + " / " + S(b));
        case OpUMod: return as(U(a) + " % " + U(b));
        case OpSRem: return as(S(a) + " % " + S(b));
        case OpSMod: return as(U(a) + " - " + U(b) + " * (uint32_t)(" + S(a) + " / " + S(b) + ")");   // :2696-2704
        case OpIEqual: return U(a) + " == " + U(b);
        case OpINotEqual: return U(a) + " != " + U(b);
        case OpUGreaterThan: return U(a) + " > " + U(b);
        case OpSGreaterThan: return S(a) + " > " + S(b);
        case OpUGreaterThanEqual: return U(a) + " >= " + U(b);
        case OpSGreaterThanEqual: return S(a) + " >= " + S(b);
        case OpULessThan: return U(a) + " < " + U(b);
        case OpSLessThan: return S(a) + " < " + S(b);
        case OpULessThanEqual: return U(a) + " <= " + U(b);
        case OpSLessThanEqual: return S(a) + " <= " + S(b);
        case OpFOrdEqual: return a + " == " + b;
        case OpFOrdNotEqual: return "(" + a + " < " + b + " || " + a + " > " + b + ")";
        case OpFOrdLessThan: return a + " < " + b;
        case OpFOrdGreaterThan: return a + " > " + b;
        case OpFOrdLessThanEqual: return a + " <= " + b;
        case OpFOrdGreaterThanEqual: return a + " >= " + b;
        case OpFUnordEqual: return "!(" + a + " < " + b + " || " + a + " > " + b + ")";
        case OpFUnordNotEqual: return a + " != " + b;
        case OpFUnordLessThan: return "!(" + a + " >= " + b + ")";
        case OpFUnordGreaterThan: return "!(" + a + " <= " + b + ")";
        case OpFUnordLessThanEqual: return "!(" + a + " > " + b + ")";
        case OpFUnordGreaterThanEqual: return "!(" + a + " < " + b + ")";
        case OpLogicalAnd: return a + " && " + b;
        case OpLogicalOr: return a + " || " + b;
        case OpLogicalEqual: return a + " == " + b;
        case OpLogicalNotEqual: return a + " != " + b;
        case OpBitwiseAnd: return as(U(a) + " & " + U(b));
        case OpBitwiseOr: return as(U(a) + " | " + U(b));
        case OpBitwiseXor: return as(U(a) + " ^ " + U(b));
        case OpShiftLeftLogical: return as(U(a) + " << (" + U(b) + " & 31u)");
        case OpShiftRightLogical: return as(U(a) + " >> (" + U(b) + " & 31u)");
        case OpShiftRightArithmetic: return as(S(a) + " >> (" + U(b) + " & 31u)");
        default: return "";
        }
    }

    void declare(uint32_t type, uint32_t id) { *od << "    " << ctype(type) << " r" << id << "{};\n"; }

    // one function: the entry point's statements, or a lambda for a callee
    void function(uint32_t fid, bool entry)
    {
        const Function& fn = m.functions.at(fid);
        std::ostringstream fdecl, fcode;
        od = entry ? &decl : &fdecl;
        oc = entry ? &code : &fcode;
        in_entry = entry;
        // OpPhi: per target block, per incoming block, the (result, value)
        // assignments made on that edge
        std::map<uint32_t, std::map<uint32_t, std::vector<std::pair<uint32_t, uint32_t>>>> phis;
        std::vector<std::string> params;
        {
            uint32_t blk = 0;
            for (const Inst& in : fn.body) {
                if (in.op == OpLabel) blk = in.w[0];
                if (in.op == OpPhi)
                    for (size_t k = 2; k + 1 < in.w.size(); k += 2) phis[blk][in.w[k + 1]].push_back({in.w[1], in.w[k]});
                if (in.op == OpFunctionParameter) {
                    const Type& pt = T(in.w[0]);
                    if (pt.op == OpTypePointer) {
                        const std::string p = "p" + std::to_string(in.w[1]);
                        params.push_back(ctype(pt.b) + "& " + p);
                        lval[in.w[1]] = p;
                        ptr_pointee[in.w[1]] = pt.b;
                    } else {
                        params.push_back(ctype(in.w[0]) + " r" + std::to_string(in.w[1]));
                    }
                }
            }
        }
        uint32_t cur = 0;
        auto edge = [&](uint32_t target) {
            std::string s;
            auto it = phis.find(target);
            if (it != phis.end()) {
                auto jt = it->second.find(cur);
                if (jt != it->second.end())
                    for (const auto& pv : jt->second) s += "r" + std::to_string(pv.first) + " = " + val(pv.second) + "; ";
            }
            return s + "goto L" + std::to_string(target) + ";";
        };
        for (const Inst& in : fn.body) {
            if (in.op == OpFunctionParameter) continue;
            if (in.op == OpLabel) cur = in.w[0];
            instruction(in, edge);
        }
        if (!entry) {
            fns << "    auto f" << fid << " = [&](";
            for (size_t k = 0; k < params.size(); k++) fns << (k ? ", " : "") << params[k];
            fns << ") -> " << ctype(fn.ret) << " {\n" << fdecl.str() << fcode.str() << "    };\n";
        }
        od = &decl;
        oc = &code;
        in_entry = true;
    }

    // callees before callers, each once
    void functions()
    {
        std::vector<uint32_t> order;
        std::set<uint32_t> seen, active;
        std::function<void(uint32_t)> visit = [&](uint32_t f) {
            if (seen.count(f)) return;
            if (active.count(f)) bad("recursive function calls");
            active.insert(f);
            auto it = m.functions.find(f);
            if (it == m.functions.end()) bad("call of an undefined function");
            for (const Inst& in : it->second.body)
                if (in.op == OpFunctionCall) visit(in.w[2]);
            active.erase(f);
            seen.insert(f);
            order.push_back(f);
        };
        visit(m.entry);
        for (uint32_t f : order) function(f, f == m.entry);
    }

    template <class Edge>
    void instruction(const Inst& in, Edge& edge)
    {
        const std::vector<uint32_t>& o = in.w;
        const uint32_t op = in.op;
        std::ostream& code = *oc;
        switch (op) {
        case OpLabel:
            code << "L" << o[0] << ":;\n";
            break;
        case OpVariable: {
            if (o[2] != ScFunction) bad("non-Function variable inside a function");
            const uint32_t pointee = T(o[0]).b;
            *od << "    " << ctype(pointee) << " v" << o[1] << "{};\n";
            lval[o[1]] = "v" + std::to_string(o[1]);
            ptr_pointee[o[1]] = pointee;
            if (o.size() > 3) code << "    v" << o[1] << " = " << val(o[3]) << ";\n";
            break;
        }
        case OpLoad: {
            declare(o[0], o[1]);
            auto u = uptr.find(o[2]);
            if (u != uptr.end()) {
                code << "    r" << o[1] << " = " << uniform_load(o[0], u->second.binding, u->second.off, u->second.mstride) << ";\n";
                break;
            }
            auto it = lval.find(o[2]);
            if (it == lval.end()) bad("load through an unknown pointer");
            code << "    r" << o[1] << " = " << it->second << ";\n";
            break;
        }
        case OpStore: {
            auto it = lval.find(o[0]);
            if (it == lval.end() || uptr.count(o[0])) bad("store through a read-only or unknown pointer");
            code << "    " << it->second << " = " << val(o[1]) << ";\n";
            break;
        }
        case OpAccessChain: {
            uint32_t t = 0;
            const std::vector<uint32_t> idx(o.begin() + 3, o.end());
            auto u = uptr.find(o[2]);
            if (u != uptr.end()) {
                uptr[o[1]] = uchain(u->second, ptr_pointee.at(o[2]), idx, &t);
            } else {
                lval[o[1]] = chain(o[2], idx, &t);
            }
            ptr_pointee[o[1]] = t;
            break;
        }
        case OpCopyObject:
            declare(o[0], o[1]);
            code << "    r" << o[1] << " = " << val(o[2]) << ";\n";
            break;
        case OpUndef:
            declare(o[0], o[1]);
            break;
        case OpCompositeExtract: {
            declare(o[0], o[1]);
            uint32_t t = type_of(o[2]);
            std::string s = val(o[2]);
            for (size_t k = 3; k < o.size(); k++) s = elem(t, s, std::to_string(o[k]), &t);
            code << "    r" << o[1] << " = " << s << ";\n";
            break;
        }
        case OpCompositeInsert: {
            declare(o[0], o[1]);
            uint32_t t = o[0];
            std::string s = "r" + std::to_string(o[1]);
            for (size_t k = 4; k < o.size(); k++) s = elem(t, s, std::to_string(o[k]), &t);
            code << "    r" << o[1] << " = " << val(o[3]) << ";\n    " << s << " = " << val(o[2]) << ";\n";
            break;
        }
        case OpVectorExtractDynamic: {
            // an index past the end is undefined; it reads the last component
            declare(o[0], o[1]);
            const int n = comps(type_of(o[2]));
            code << "    r" << o[1] << " = " << val(o[2]) << ".c[min((uint32_t)(" << val(o[3]) << "), " << n - 1 << "u)];\n";
            break;
        }
        case OpVectorInsertDynamic: {
            declare(o[0], o[1]);
            const int n = comps(o[0]);
            code << "    r" << o[1] << " = " << val(o[2]) << ";\n    { const uint32_t i_ = (uint32_t)(" << val(o[4]) << "); if (i_ < "
                 << n << "u) r" << o[1] << ".c[i_] = " << val(o[3]) << "; }\n";
            break;
        }
        case OpCompositeConstruct: {
            declare(o[0], o[1]);
            const Type& rt = T(o[0]);
            if (rt.op == OpTypeVector) {
                int k = 0;
                for (size_t a = 2; a < o.size(); a++) {
                    const uint32_t at = type_of(o[a]);
                    for (int c = 0; c < comps(at); c++) code << "    r" << o[1] << ".c[" << k++ << "] = " << comp(o[a], at, c) << ";\n";
                }
            } else if (rt.op == OpTypeMatrix) {
                for (size_t a = 2; a < o.size(); a++) code << "    r" << o[1] << ".col[" << a - 2 << "] = " << val(o[a]) << ";\n";
            } else if (rt.op == OpTypeArray) {
                for (size_t a = 2; a < o.size(); a++) code << "    r" << o[1] << ".e[" << a - 2 << "] = " << val(o[a]) << ";\n";
            } else if (rt.op == OpTypeStruct) {
                for (size_t a = 2; a < o.size(); a++) code << "    r" << o[1] << ".m" << a - 2 << " = " << val(o[a]) << ";\n";
            } else {
                bad("composite construct of an unsupported type");
            }
            break;
        }
        case OpVectorShuffle: {
            declare(o[0], o[1]);
            const uint32_t t1 = type_of(o[2]), t2 = type_of(o[3]);
            const int n1 = comps(t1);
            for (size_t k = 4; k < o.size(); k++) {
                const uint32_t s = o[k];
                if (s == 0xFFFFFFFFu) continue;
                code << "    r" << o[1] << ".c[" << k - 4 << "] = " << ((int)s < n1 ? comp(o[2], t1, (int)s) : comp(o[3], t2, (int)s - n1))
                     << ";\n";
            }
            break;
        }
        case OpFNegate: case OpSNegate: case OpNot: case OpLogicalNot: case OpBitReverse: case OpBitCount: case OpIsNan:
        case OpIsInf: {
            declare(o[0], o[1]);
            const uint32_t st = type_of(o[2]);
            const std::string et = ctype(scalar_of(o[0]));
            for (int k = 0; k < comps(o[0]); k++) {
                const std::string a = comp(o[2], st, k);
                std::string e;
                switch (op) {
                case OpFNegate: e = "-" + a; break;
                case OpSNegate: e = "(" + et + ")(0u - (uint32_t)(" + a + "))"; break;
                case OpNot: e = "(" + et + ")(~(uint32_t)(" + a + "))"; break;
                case OpLogicalNot: e = "!" + a; break;
                case OpBitReverse: e = "(" + et + ")__builtin_bitreverse32((uint32_t)(" + a + "))"; break;
                case OpBitCount: e = "(" + et + ")__builtin_popcount((uint32_t)(" + a + "))"; break;
                case OpIsNan: e = "(" + a + " != " + a + ")"; break;
                default: e = "__builtin_isinf(" + a + ")"; break;
                }
                code << "    " << res(o[1], o[0], k) << " = " << e << ";\n";
            }
            break;
        }
        case OpAny: case OpAll: {
            declare(o[0], o[1]);
            const uint32_t st = type_of(o[2]);
            code << "    r" << o[1] << " = ";
            for (int k = 0; k < comps(st); k++) code << (k ? (op == OpAny ? " || " : " && ") : "") << comp(o[2], st, k);
            code << ";\n";
            break;
        }
        case OpConvertFToS: case OpConvertFToU: case OpConvertSToF: case OpConvertUToF: case OpBitcast: case OpUConvert:
        case OpSConvert: case OpFConvert: case OpQuantizeToF16: {
            declare(o[0], o[1]);
            const int n = comps(o[0]);
            const uint32_t st = type_of(o[2]);
            if (comps(st) != n) bad("conversion between vectors of different sizes");
            const std::string dt = ctype(scalar_of(o[0])), sty = ctype(scalar_of(st));
            for (int k = 0; k < n; k++) {
                const std::string src = comp(o[2], st, k);
                std::string e;
                if (op == OpConvertSToF) e = "(float)(int32_t)(" + src + ")";
                else if (op == OpConvertUToF) e = "(float)(uint32_t)(" + src + ")";
                else if (op == OpQuantizeToF16) e = "(float)(_Float16)(" + src + ")";
                else if (op == OpBitcast && dt == "float" && sty != "float") e = "__uint_as_float((uint32_t)(" + src + "))";
                else if (op == OpBitcast && dt != "float" && sty == "float") e = "(" + dt + ")__float_as_uint(" + src + ")";
                else e = "(" + dt + ")(" + src + ")";
                code << "    " << res(o[1], o[0], k) << " = " << e << ";\n";
            }
            break;
        }
        case OpVectorTimesScalar: {
            declare(o[0], o[1]);
            for (int k = 0; k < comps(o[0]); k++)
                code << "    r" << o[1] << ".c[" << k << "] = " << val(o[2]) << ".c[" << k << "] * " << val(o[3]) << ";\n";
            break;
        }
        case OpMatrixTimesScalar: {
            // glm m * s, column by column (SpirvFunctions.cpp:10-21)
            declare(o[0], o[1]);
            const Type& mt = T(o[0]);
            for (uint32_t c = 0; c < mt.b; c++)
                for (uint32_t r = 0; r < T(mt.a).b; r++)
                    code << "    r" << o[1] << ".col[" << c << "].c[" << r << "] = " << val(o[2]) << ".col[" << c << "].c[" << r << "] * "
                         << val(o[3]) << ";\n";
            break;
        }
        case OpMatrixTimesVector: {
            declare(o[0], o[1]);
            const Type& mt = T(type_of(o[2]));
            const uint32_t cols = mt.b, rows = T(mt.a).b;
            if (cols == 4 && rows == 4) {
                // glm mat4 * vec4: (m0 v0 + m1 v1) + (m2 v2 + m3 v3) (SpirvFunctions.cpp:37-47)
                code << "    mat4_mul_vec4(" << val(o[2]) << ".col[0].c, " << val(o[3]) << ".c, r" << o[1] << ".c);\n";
            } else {
                // glm mat3 * vec3 and the other shapes: left to right
                for (uint32_t r = 0; r < rows; r++) {
                    code << "    r" << o[1] << ".c[" << r << "] = ";
                    for (uint32_t c = 0; c < cols; c++)
                        code << (c ? " + " : "") << val(o[2]) << ".col[" << c << "].c[" << r << "] * " << val(o[3]) << ".c[" << c << "]";
                    code << ";\n";
                }
            }
            break;
        }
        case OpVectorTimesMatrix: {
            // glm v * m: row j = m[j][0] v0 + m[j][1] v1 + ..., left to right (SpirvFunctions.cpp:23-34)
            declare(o[0], o[1]);
            const Type& mt = T(type_of(o[3]));
            const uint32_t cols = mt.b, rows = T(mt.a).b;
            for (uint32_t c = 0; c < cols; c++) {
                code << "    r" << o[1] << ".c[" << c << "] = ";
                for (uint32_t r = 0; r < rows; r++)
                    code << (r ? " + " : "") << val(o[3]) << ".col[" << c << "].c[" << r << "] * " << val(o[2]) << ".c[" << r << "]";
                code << ";\n";
            }
            break;
        }
        case OpMatrixTimesMatrix: {
            // glm m1 * m2: column j = m1[0] m2[j][0] + m1[1] m2[j][1] + ..., left to right (SpirvFunctions.cpp:49-60)
            declare(o[0], o[1]);
            const Type& at = T(type_of(o[2]));
            const Type& bt = T(type_of(o[3]));
            const uint32_t rows = T(at.a).b, inner = at.b, cols = bt.b;
            for (uint32_t c = 0; c < cols; c++)
                for (uint32_t r = 0; r < rows; r++) {
                    code << "    r" << o[1] << ".col[" << c << "].c[" << r << "] = ";
                    for (uint32_t k = 0; k < inner; k++)
                        code << (k ? " + " : "") << val(o[2]) << ".col[" << k << "].c[" << r << "] * " << val(o[3]) << ".col[" << c << "].c["
                             << k << "]";
                    code << ";\n";
                }
            break;
        }
        case OpOuterProduct: {
            declare(o[0], o[1]);
            const Type& mt = T(o[0]);
            for (uint32_t c = 0; c < mt.b; c++)
                for (uint32_t r = 0; r < T(mt.a).b; r++)
                    code << "    r" << o[1] << ".col[" << c << "].c[" << r << "] = " << val(o[2]) << ".c[" << r << "] * " << val(o[3]) << ".c["
                         << c << "];\n";
            break;
        }
        case OpTranspose: {
            declare(o[0], o[1]);
            const Type& mt = T(o[0]);
            for (uint32_t c = 0; c < mt.b; c++)
                for (uint32_t r = 0; r < T(mt.a).b; r++)
                    code << "    r" << o[1] << ".col[" << c << "].c[" << r << "] = " << val(o[2]) << ".col[" << r << "].c[" << c << "];\n";
            break;
        }
        case OpDot: {
            declare(o[0], o[1]);
            const int n = comps(type_of(o[2]));
            code << "    r" << o[1] << " = dot_n(" << val(o[2]) << ".c, " << val(o[3]) << ".c, " << n << ");\n";
            break;
        }
        case OpSelect: {
            declare(o[0], o[1]);
            const uint32_t ct = type_of(o[2]);
            const Type& rt = T(o[0]);
            if (rt.op != OpTypeVector || !is_vec(ct)) {
                code << "    r" << o[1] << " = " << val(o[2]) << " ? " << val(o[3]) << " : " << val(o[4]) << ";\n";
            } else {
                for (int k = 0; k < comps(o[0]); k++)
                    code << "    " << res(o[1], o[0], k) << " = " << comp(o[2], ct, k) << " ? " << comp(o[3], o[0], k) << " : "
                         << comp(o[4], o[0], k) << ";\n";
            }
            break;
        }
        case OpBitFieldInsert: case OpBitFieldSExtract: case OpBitFieldUExtract: {
            declare(o[0], o[1]);
            const uint32_t bt = type_of(o[2]);
            const std::string et = ctype(scalar_of(o[0]));
            for (int k = 0; k < comps(o[0]); k++) {
                if (op == OpBitFieldInsert)
                    code << "    " << res(o[1], o[0], k) << " = (" << et << ")spv_bfi((uint32_t)(" << comp(o[2], bt, k) << "), (uint32_t)("
                         << comp(o[3], type_of(o[3]), k) << "), (uint32_t)(" << val(o[4]) << "), (uint32_t)(" << val(o[5]) << "));\n";
                else
                    code << "    " << res(o[1], o[0], k) << " = (" << et << ")" << (op == OpBitFieldSExtract ? "spv_bfe_s" : "spv_bfe_u")
                         << "((uint32_t)(" << comp(o[2], bt, k) << "), (uint32_t)(" << val(o[3]) << "), (uint32_t)(" << val(o[4]) << "));\n";
            }
            break;
        }
        case OpSampledImage:
            bad("separate images and samplers are not supported (use combined image samplers)");
        case OpImage:
            declare(o[0], o[1]);
            code << "    r" << o[1] << " = " << val(o[2]) << ";\n";
            break;
        case OpImageSampleImplicitLod: {
            declare(o[0], o[1]);
            if (o.size() > 4) bad("image operands on OpImageSampleImplicitLod are not supported");
            code << "    sample_implicit(*" << val(o[2]) << ", " << val(o[3]) << ".c[0], " << val(o[3]) << ".c[1], r" << o[1] << ".c);\n";
            break;
        }
        case OpImageSampleExplicitLod: {
            // ImageSampleExplicitLod (GlslFunctions.cpp:598-654): Lod operand only
            declare(o[0], o[1]);
            if (o.size() != 6 || o[4] != 0x2u) bad("OpImageSampleExplicitLod needs exactly the Lod image operand");
            code << "    sample_explicit(*" << val(o[2]) << ", " << val(o[3]) << ".c[0], " << val(o[3]) << ".c[1], " << val(o[5]) << ", r"
                 << o[1] << ".c);\n";
            break;
        }
        case OpImageFetch: {
            // ImageFetch (GlslFunctions.cpp:673-734): the view's base level, a Lod operand is not used
            declare(o[0], o[1]);
            if (!is_float(o[0])) bad("OpImageFetch of a non-float image");
            code << "    fetch_texel(*" << val(o[2]) << ", (int32_t)" << val(o[3]) << ".c[0], (int32_t)" << val(o[3]) << ".c[1], r" << o[1]
                 << ".c);\n";
            break;
        }
        case OpImageQuerySize: case OpImageQuerySizeLod: {
            declare(o[0], o[1]);
            const std::string lod = op == OpImageQuerySizeLod ? "(uint32_t)(" + val(o[3]) + ")" : "0u";
            code << "    { const uint32_t l_ = min(" << lod << ", 15u); r" << o[1] << ".c[0] = (int32_t)" << val(o[2]) << "->level_w[l_]; r"
                 << o[1] << ".c[1] = (int32_t)" << val(o[2]) << "->level_h[l_]; }\n";
            break;
        }
        case OpImageQueryLevels:
            declare(o[0], o[1]);
            code << "    r" << o[1] << " = (int32_t)" << val(o[2]) << "->level_count;\n";
            break;
        case OpExtInst: {
            if (o[2] != m.glsl) bad("extended instruction from an unknown set");
            declare(o[0], o[1]);
            ext(o);
            break;
        }
        case OpFunctionCall: {
            if (o[0] != 0 && T(o[0]).op != OpTypeVoid) declare(o[0], o[1]);
            std::string args;
            for (size_t k = 3; k < o.size(); k++) {
                std::string a;
                if (uptr.count(o[k])) bad("a pointer into a uniform buffer passed to a function");
                const bool pointer = m.globals.count(o[k]) || (rtype.count(o[k]) && T(rtype.at(o[k])).op == OpTypePointer);
                if (pointer) {
                    auto it = lval.find(o[k]);
                    if (it == lval.end()) bad("function argument through an unknown pointer");
                    a = it->second;
                } else {
                    a = val(o[k]);
                }
                args += (k > 3 ? ", " : "") + a;
            }
            code << "    " << (T(o[0]).op == OpTypeVoid ? "" : "r" + std::to_string(o[1]) + " = ") << "f" << o[2] << "(" << args << ");\n";
            break;
        }
        case OpPhi:
            declare(o[0], o[1]);
            break;
        case OpSelectionMerge: case OpLoopMerge: case OpLine: case OpNoLine:
            break;
        case OpBranch:
            code << "    " << edge(o[0]) << "\n";
            break;
        case OpBranchConditional:
            code << "    if (" << val(o[0]) << ") { " << edge(o[1]) << " } else { " << edge(o[2]) << " }\n";
            break;
        case OpSwitch: {
            code << "    switch ((uint32_t)(" << val(o[0]) << ")) {\n";
            for (size_t k = 2; k + 1 < o.size(); k += 2) code << "    case " << o[k] << "u: " << edge(o[k + 1]) << "\n";
            code << "    default: " << edge(o[1]) << "\n    }\n";
            break;
        }
        case OpKill:
            if (vertex) bad("OpKill in a vertex shader");
            if (!in_entry) bad("OpKill inside a called function is not supported");
            discard = true;
            code << "    return true;\n";
            break;
        case OpReturn:
            code << (in_entry ? "    goto L_end;\n" : "    return;\n");
            break;
        case OpReturnValue:
            if (in_entry) bad("OpReturnValue in the entry point");
            code << "    return " << val(o[0]) << ";\n";
            break;
        case OpUnreachable:
            code << (in_entry ? "    goto L_end;\n" : "    __builtin_unreachable();\n");
            break;
        default: {
            if (o.size() < 4 || binexpr(op, "T", "a", "b").empty()) {
                std::ostringstream e;
                e << "unsupported instruction, opcode " << op;
                bad(e.str());
            }
            const uint32_t at = type_of(o[2]), bt2 = type_of(o[3]);
            const std::string et = ctype(scalar_of(o[0]));
            declare(o[0], o[1]);
            const int n = std::max(comps(at), comps(bt2));
            for (int k = 0; k < n; k++)
                code << "    " << res(o[1], o[0], k) << " = " << binexpr(op, et, comp(o[2], at, k), comp(o[3], bt2, k)) << ";\n";
            break;
        }
        }
    }

    // GLSL.std.450.  The reference registers host functions for FAbs, SAbs,
    // SSign, Sin, Cos, Pow, [FUS]Min, [FUS]Max, [FUS]Clamp, FMix, Normalize,
    // Reflect, FindILsb, FindSMsb, FindUMsb and NMin / NMax / NClamp
    // (GlslFunctions.cpp:930-1300); those follow its code.  The others
    // follow the GLSL.std.450 specification (glm 0.9.5.3 where glm defines
    // them: length, distance, cross, faceforward, refract).
    void ext(const std::vector<uint32_t>& o)
    {
        std::ostream& code = *oc;
        const uint32_t inst = o[3];
        const uint32_t rt = o[0];
        const std::string r = "r" + std::to_string(o[1]);
        const int n = comps_or_one(rt);
        auto c = [&](size_t a, int k) { return comp(o[a], type_of(o[a]), k); };
        auto out = [&](int k) { return r + (is_vec(rt) ? ".c[" + std::to_string(k) + "]" : ""); };
        auto each = [&](const std::string& f, int args) {
            // f with #0, #1, #2 standing for the arguments' components
            for (int k = 0; k < n; k++) {
                std::string e = f;
                for (int a = 0; a < args; a++) {
                    const std::string tag = "#" + std::to_string(a);
                    for (size_t p = e.find(tag); p != std::string::npos; p = e.find(tag, p + 1)) e.replace(p, tag.size(), "(" + c(4 + a, k) + ")");
                }
                code << "    " << out(k) << " = " << e << ";\n";
            }
        };
        const std::string et = is_vec(rt) || T(rt).op != OpTypeStruct ? ctype(scalar_of(rt)) : "";
        auto sumsq = [&](uint32_t id) {
            // x.x * x.x + x.y * x.y + ..., left to right (glm length / normalize)
            const uint32_t t = type_of(id);
            std::string s;
            for (int k = 0; k < comps(t); k++) s += (k ? " + " : "") + comp(id, t, k) + " * " + comp(id, t, k);
            return s;
        };
        switch (inst) {
        case 1: each("roundf(#0)", 1); break;                                               // Round
        case 2: each("rintf(#0)", 1); break;                                                // RoundEven
        case 3: each("truncf(#0)", 1); break;                                               // Trunc
        case 4: each("fabsf(#0)", 1); break;                                                // FAbs
        case 5: each("(int32_t)((int32_t)#0 < 0 ? 0u - (uint32_t)#0 : (uint32_t)#0)", 1); break;   // SAbs (std::abs)
        case 6: each("(#0 > 0.0f ? 1.0f : (#0 < 0.0f ? -1.0f : #0))", 1); break;            // FSign
        case 7: each("(int32_t)(0 < (int32_t)#0) - (int32_t)((int32_t)#0 < 0)", 1); break;   // SSign
        case 8: each("floorf(#0)", 1); break;                                               // Floor
        case 9: each("ceilf(#0)", 1); break;                                                // Ceil
        case 10: each("spv_fract(#0)", 1); break;                                           // Fract
        case 11: each("#0 * 0.01745329251994329577f", 1); break;                            // Radians
        case 12: each("#0 * 57.2957795130823208768f", 1); break;                            // Degrees
        case 13: each("sinf(#0)", 1); break;                                                // Sin
        case 14: each("cosf(#0)", 1); break;                                                // Cos
        case 15: each("tanf(#0)", 1); break;                                                // Tan
        case 16: each("asinf(#0)", 1); break;                                               // Asin
        case 17: each("acosf(#0)", 1); break;                                               // Acos
        case 18: each("atanf(#0)", 1); break;                                               // Atan
        case 19: each("sinhf(#0)", 1); break;                                               // Sinh
        case 20: each("coshf(#0)", 1); break;                                               // Cosh
        case 21: each("tanhf(#0)", 1); break;                                               // Tanh
        case 22: each("asinhf(#0)", 1); break;                                              // Asinh
        case 23: each("acoshf(#0)", 1); break;                                              // Acosh
        case 24: each("atanhf(#0)", 1); break;                                              // Atanh
        case 25: each("atan2f(#0, #1)", 2); break;                                          // Atan2
        case 26: each("powf(#0, #1)", 2); break;                                            // Pow
        case 27: each("expf(#0)", 1); break;                                                // Exp
        case 28: each("logf(#0)", 1); break;                                                // Log
        case 29: each("exp2f(#0)", 1); break;                                               // Exp2
        case 30: each("log2f(#0)", 1); break;                                               // Log2
        case 31: each("sqrtf(#0)", 1); break;                                               // Sqrt
        case 32: each("1.0f / sqrtf(#0)", 1); break;                                        // InverseSqrt (glm inversesqrt)
        case 37: each("spv_min(#0, #1)", 2); break;                                         // FMin
        case 38: each("spv_min((uint32_t)#0, (uint32_t)#1)", 2); break;                     // UMin
        case 39: each("spv_min((int32_t)#0, (int32_t)#1)", 2); break;                       // SMin
        case 40: each("spv_max(#0, #1)", 2); break;                                         // FMax
        case 41: each("spv_max((uint32_t)#0, (uint32_t)#1)", 2); break;                     // UMax
        case 42: each("spv_max((int32_t)#0, (int32_t)#1)", 2); break;                       // SMax
        case 43: each("spv_clamp(#0, #1, #2)", 3); break;                                   // FClamp
        case 44: each("spv_clamp((uint32_t)#0, (uint32_t)#1, (uint32_t)#2)", 3); break;     // UClamp
        case 45: each("spv_clamp((int32_t)#0, (int32_t)#1, (int32_t)#2)", 3); break;        // SClamp
        case 46: each("#0 * (1.0f - #2) + #1 * #2", 3); break;                              // FMix (Mix :145-149)
        case 48: each("(#1 < #0 ? 0.0f : 1.0f)", 2); break;                                 // Step
        case 49: each("spv_smoothstep(#0, #1, #2)", 3); break;                              // SmoothStep
        case 50: each("fmaf(#0, #1, #2)", 3); break;                                        // Fma
        case 53: each("ldexpf(#0, (int32_t)#1)", 2); break;                                 // Ldexp
        case 54: case 55: case 56: case 57: case 58: {                                      // Pack*
            static const char* f[] = {"spv_pack_snorm4x8", "spv_pack_unorm4x8", "spv_pack_snorm2x16", "spv_pack_unorm2x16", "spv_pack_half2x16"};
            code << "    " << r << " = " << f[inst - 54] << "(" << val(o[4]) << ");\n";
            break;
        }
        case 60: case 61: case 62: case 63: case 64: {                                      // Unpack*
            static const char* f[] = {"spv_unpack_snorm2x16", "spv_unpack_unorm2x16", "spv_unpack_half2x16", "spv_unpack_snorm4x8",
                                      "spv_unpack_unorm4x8"};
            code << "    " << r << " = " << f[inst - 60] << "(" << val(o[4]) << ");\n";
            break;
        }
        case 66:                                                                            // Length
            code << "    " << r << " = sqrtf(" << sumsq(o[4]) << ");\n";
            break;
        case 67: {                                                                          // Distance: length(p0 - p1)
            const uint32_t t = type_of(o[4]);
            const int m2 = comps(t);
            code << "    { " << ctype(t) << " d_{};\n";
            for (int k = 0; k < m2; k++) code << "      d_" << (is_vec(t) ? ".c[" + std::to_string(k) + "]" : "") << " = " << c(4, k) << " - " << c(5, k) << ";\n";
            std::string s;
            for (int k = 0; k < m2; k++) {
                const std::string dk = "d_" + (is_vec(t) ? ".c[" + std::to_string(k) + "]" : std::string());
                s += (k ? " + " : "") + dk + " * " + dk;
            }
            code << "      " << r << " = sqrtf(" << s << "); }\n";
            break;
        }
        case 68:                                                                            // Cross (glm)
            code << "    " << r << ".c[0] = " << c(4, 1) << " * " << c(5, 2) << " - " << c(5, 1) << " * " << c(4, 2) << ";\n"
                 << "    " << r << ".c[1] = " << c(4, 2) << " * " << c(5, 0) << " - " << c(5, 2) << " * " << c(4, 0) << ";\n"
                 << "    " << r << ".c[2] = " << c(4, 0) << " * " << c(5, 1) << " - " << c(5, 0) << " * " << c(4, 1) << ";\n";
            break;
        case 69:                                                                            // Normalize
            if (!is_vec(rt)) {
                // the reference's scalar Normalise (GlslFunctions.cpp:151-155)
                each("(__builtin_signbit(#0) ? 1.0f : -1.0f)", 1);
            } else {
                code << "    { const float inv_ = 1.0f / sqrtf(" << sumsq(o[4]) << ");\n";
                for (int k = 0; k < n; k++) code << "      " << r << ".c[" << k << "] = " << c(4, k) << " * inv_;\n";
                code << "    }\n";
            }
            break;
        case 70: {                                                                          // FaceForward (glm)
            const std::string d = dotexpr(o[6], o[5]);
            code << "    { const bool neg_ = " << d << " < 0.0f;\n";
            for (int k = 0; k < n; k++) code << "      " << out(k) << " = neg_ ? " << c(4, k) << " : -" << c(4, k) << ";\n";
            code << "    }\n";
            break;
        }
        case 71: {                                                                          // Reflect: I - N * dot(N, I) * 2 (glm)
            const std::string d = dotexpr(o[5], o[4]);
            code << "    { const float d_ = " << d << ";\n";
            for (int k = 0; k < n; k++) code << "      " << out(k) << " = " << c(4, k) << " - " << c(5, k) << " * d_ * 2.0f;\n";
            code << "    }\n";
            break;
        }
        case 72: {                                                                          // Refract (glm)
            const std::string d = dotexpr(o[5], o[4]);
            const std::string eta = val(o[6]);
            code << "    { const float d_ = " << d << "; const float k_ = 1.0f - " << eta << " * " << eta << " * (1.0f - d_ * d_);\n";
            for (int k = 0; k < n; k++)
                code << "      " << out(k) << " = k_ < 0.0f ? 0.0f : " << eta << " * " << c(4, k) << " - (" << eta << " * d_ + sqrtf(k_)) * "
                     << c(5, k) << ";\n";
            code << "    }\n";
            break;
        }
        case 73: each("(" + et + ")spv_find_lsb((uint32_t)#0)", 1); break;                  // FindILsb
        case 74: each("(" + et + ")spv_find_smsb((int32_t)#0)", 1); break;                  // FindSMsb
        case 75: each("(" + et + ")spv_find_umsb((uint32_t)#0)", 1); break;                 // FindUMsb
        case 79: each("spv_nmin(#0, #1)", 2); break;                                        // NMin
        case 80: each("spv_nmax(#0, #1)", 2); break;                                        // NMax
        case 81: each("spv_nmin(spv_nmax(#0, #1), #2)", 3); break;                          // NClamp
        default: {
            std::ostringstream e;
            e << "unsupported GLSL.std.450 instruction " << inst;
            bad(e.str());
        }
        }
    }

    int comps_or_one(uint32_t t) const
    {
        const Type& ty = T(t);
        return ty.op == OpTypeVector ? (int)ty.b : 1;
    }

    // glm::dot of two values (a scalar "dot" is the product)
    std::string dotexpr(uint32_t a, uint32_t b)
    {
        const uint32_t t = type_of(a);
        if (!is_vec(t)) return val(a) + " * " + val(b);
        return "dot_n(" + val(a) + ".c, " + val(b) + ".c, " + std::to_string(comps(t)) + ")";
    }
};

const char* kPrelude = R"(#include "raster_kernels.h"
namespace cpv {
template <class T, int N> struct V { T c[N]; };
template <int R, int C> struct M { V<float, R> col[C]; };
template <class T, int N> struct A { T e[N]; };
// std::min / std::max / std::clamp (GlslFunctions.cpp:118-142)
template <class T> __device__ __forceinline__ T spv_min(T a, T b) { return b < a ? b : a; }
template <class T> __device__ __forceinline__ T spv_max(T a, T b) { return a < b ? b : a; }
template <class T> __device__ __forceinline__ T spv_clamp(T v, T lo, T hi) { return v < lo ? lo : (hi < v ? hi : v); }
// NMin / NMax (GlslFunctions.cpp:266-300): a NaN first operand yields the second
__device__ __forceinline__ float spv_nmin(float a, float b) { return a != a ? b : spv_min(a, b); }
__device__ __forceinline__ float spv_nmax(float a, float b) { return a != a ? b : spv_max(a, b); }
__device__ __forceinline__ float spv_fract(float v) { return v - floorf(v); }
// OpFMod (SPIRVCompiler.cpp:2712-2722): frem, then the divisor's sign
__device__ __forceinline__ float spv_fmod(float a, float b)
{
    const float r = fmodf(a, b);
    const float s = copysignf(r, b);
    return (r < s || r > s) ? r + b : s;
}
// OpBitFieldInsert as the reference builds it (SPIRVCompiler.cpp:3099-3136):
// the mask selects bits [offset, offset + count) of Insert itself
__device__ __forceinline__ uint32_t spv_bfi(uint32_t base, uint32_t ins, uint32_t offset, uint32_t count)
{
    uint32_t mask = (0xFFFFFFFFu << (count & 31u)) ^ 0xFFFFFFFFu;
    mask <<= (offset & 31u);
    return (~mask & base) | (mask & ins);
}
__device__ __forceinline__ uint32_t spv_bfe_u(uint32_t base, uint32_t offset, uint32_t count)
{
    if (count == 0u) return 0u;
    const uint32_t v = base >> (offset & 31u);
    return count >= 32u ? v : v & ((1u << count) - 1u);
}
__device__ __forceinline__ int32_t spv_bfe_s(uint32_t base, uint32_t offset, uint32_t count)
{
    if (count == 0u) return 0;
    const uint32_t sh = 32u - (count < 32u ? count : 32u);
    return (int32_t)(spv_bfe_u(base, offset, count) << sh) >> sh;
}
// FindILsb / FindSMsb / FindUMsb (GlslFunctions.cpp:185-250)
__device__ __forceinline__ int32_t spv_find_lsb(uint32_t v) { return v == 0u ? -1 : __builtin_ctz(v); }
__device__ __forceinline__ int32_t spv_find_umsb(uint32_t v) { return v == 0u ? -1 : 31 - __builtin_clz(v); }
__device__ __forceinline__ int32_t spv_find_smsb(int32_t v) { return spv_find_umsb((uint32_t)(v < 0 ? ~v : v)); }
__device__ __forceinline__ float spv_smoothstep(float e0, float e1, float x)
{
    const float t = spv_clamp((x - e0) / (e1 - e0), 0.0f, 1.0f);
    return t * t * (3.0f - 2.0f * t);
}
// GLSL.std.450 packing: round to nearest, halves away from zero
__device__ __forceinline__ uint32_t spv_pack_unorm4x8(V<float, 4> v)
{
    uint32_t r = 0u;
    for (int i = 0; i < 4; i++) r |= (uint32_t)roundf(spv_clamp(v.c[i], 0.0f, 1.0f) * 255.0f) << (8 * i);
    return r;
}
__device__ __forceinline__ uint32_t spv_pack_snorm4x8(V<float, 4> v)
{
    uint32_t r = 0u;
    for (int i = 0; i < 4; i++) r |= ((uint32_t)(int32_t)roundf(spv_clamp(v.c[i], -1.0f, 1.0f) * 127.0f) & 0xFFu) << (8 * i);
    return r;
}
__device__ __forceinline__ uint32_t spv_pack_unorm2x16(V<float, 2> v)
{
    return (uint32_t)roundf(spv_clamp(v.c[0], 0.0f, 1.0f) * 65535.0f) | ((uint32_t)roundf(spv_clamp(v.c[1], 0.0f, 1.0f) * 65535.0f) << 16);
}
__device__ __forceinline__ uint32_t spv_pack_snorm2x16(V<float, 2> v)
{
    const uint32_t a = (uint32_t)(int32_t)roundf(spv_clamp(v.c[0], -1.0f, 1.0f) * 32767.0f) & 0xFFFFu;
    const uint32_t b = (uint32_t)(int32_t)roundf(spv_clamp(v.c[1], -1.0f, 1.0f) * 32767.0f) & 0xFFFFu;
    return a | (b << 16);
}
__device__ __forceinline__ uint32_t spv_pack_half2x16(V<float, 2> v)
{
    const _Float16 a = (_Float16)v.c[0], b = (_Float16)v.c[1];
    return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
}
__device__ __forceinline__ V<float, 4> spv_unpack_unorm4x8(uint32_t p)
{
    V<float, 4> r;
    for (int i = 0; i < 4; i++) r.c[i] = (float)((p >> (8 * i)) & 0xFFu) / 255.0f;
    return r;
}
__device__ __forceinline__ V<float, 4> spv_unpack_snorm4x8(uint32_t p)
{
    V<float, 4> r;
    for (int i = 0; i < 4; i++) r.c[i] = spv_clamp((float)(int8_t)((p >> (8 * i)) & 0xFFu) / 127.0f, -1.0f, 1.0f);
    return r;
}
__device__ __forceinline__ V<float, 2> spv_unpack_unorm2x16(uint32_t p)
{
    return V<float, 2>{{(float)(p & 0xFFFFu) / 65535.0f, (float)(p >> 16) / 65535.0f}};
}
__device__ __forceinline__ V<float, 2> spv_unpack_snorm2x16(uint32_t p)
{
    return V<float, 2>{{spv_clamp((float)(int16_t)(p & 0xFFFFu) / 32767.0f, -1.0f, 1.0f),
                        spv_clamp((float)(int16_t)(p >> 16) / 32767.0f, -1.0f, 1.0f)}};
}
__device__ __forceinline__ V<float, 2> spv_unpack_half2x16(uint32_t p)
{
    return V<float, 2>{{(float)__builtin_bit_cast(_Float16, (uint16_t)(p & 0xFFFFu)), (float)__builtin_bit_cast(_Float16, (uint16_t)(p >> 16))}};
}
)";

// Template arguments of the raster kernel variants in Program::raster order:
// the seven tile configurations, then the 16x16 / four-wave kernel that
// also rasterises point and line records, then the seven configurations
// for the fixed formats.
std::vector<std::string> raster_instances()
{
    return {"8, 1",  "16, 1", "16, 2", "16, 4", "32, 1", "32, 2", "32, 4", "16, 4, true",
            "8, 1, false, true",  "16, 1, false, true", "16, 2, false, true", "16, 4, false, true",
            "32, 1, false, true", "32, 2, false, true", "32, 4, false, true"};
}

struct Translation {
    std::string source;
    int vs_outputs = 0, fs_inputs = 0, psize = -1;
};

Translation translate(const uint32_t* vs, size_t vs_n, const char* vs_entry, const uint32_t* fs, size_t fs_n, const char* fs_entry)
{
    Module mv = parse(vs, vs_n, vs_entry);
    Module mf = parse(fs, fs_n, fs_entry);
    if (mv.model != 0) bad("the first module's entry point is not a vertex shader");
    if (mf.model != 4) bad("the second module's entry point is not a fragment shader");
    Gen gv(mv, true), gf(mf, false);
    gv.globals();
    gv.functions();
    gf.globals();
    gf.bind_fragment_inputs(gv.loc_offset, gv.loc_size);
    gf.functions();
    if (!mf.origin_upper_left) bad("fragment shaders must declare OriginUpperLeft (Vulkan requires it)");
    std::ostringstream s;
    s << kPrelude;
    s << "struct SpvVs {\n    static constexpr int kOutputs = " << gv.out_total << ";\n" << gv.types.str()
      << "    __device__ __forceinline__ static void run(const DrawEnv& env, uint32_t vid, uint32_t iid, float pos[4], float* out)\n    {\n"
      << gv.decl.str() << gv.fns.str() << gv.pro.str() << gv.code.str() << "L_end:;\n" << gv.epi.str()
      << "    (void)env; (void)vid; (void)iid; (void)out; (void)pos;\n    }\n};\n";
    const int fin = (int)gf.interp.size();
    s << "struct SpvFs {\n    static constexpr int kInputs = " << fin << ";\n" << gf.types.str()
      << "    static constexpr bool kMayDiscard = " << (gf.discard ? "true" : "false") << ";\n"
      << "    __device__ static int interp(int c)\n    {\n";
    for (int c = 0; c < fin; c++)
        if (gf.interp[c]) s << "        if (c == " << c << ") return " << gf.interp[c] << ";\n";
    s << "        (void)c;\n        return 0;\n    }\n"
      << "    __device__ __forceinline__ static bool run(const DrawEnv& env, const float* in, const float* fc, bool front, float out[4])\n    {\n"
      << gf.decl.str() << gf.fns.str() << gf.pro.str() << gf.code.str() << "L_end:;\n" << gf.epi.str()
      << "    (void)env; (void)in; (void)fc; (void)front;\n        return false;\n    }\n};\n";
    s << "}  // namespace cpv\n";
    s << "template __global__ void cpv::vertex_kernel<cpv::SpvVs>(cpv::VertexArgs);\n";
    for (const std::string& k : raster_instances())
        s << "template __global__ void cpv::raster_kernel<cpv::SpvFs, " << k << ">(cpv::RasterArgs);\n";
    Translation t;
    t.source = s.str();
    t.vs_outputs = gv.out_total;
    t.fs_inputs = fin;
    t.psize = gv.psize;
    return t;
}

std::vector<std::string> kernel_names()
{
    std::vector<std::string> n{"cpv::vertex_kernel<cpv::SpvVs>"};
    for (const std::string& k : raster_instances()) n.push_back("cpv::raster_kernel<cpv::SpvFs, " + k + ">");
    return n;
}

const char* const kRtcOptions[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fno-fast-math",
                                   "-fhip-fp32-correctly-rounded-divide-sqrt", "-fno-gpu-flush-denormals-to-zero",
                                   "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
constexpr int kRtcOptionCount = (int)(sizeof(kRtcOptions) / sizeof(kRtcOptions[0]));

// Code-object cache (the role of the reference's pipeline cache of compiled
// modules, PipelineCache.cpp:18-147): compiled programs are stored under a
// 64-bit FNV-1a hash of the generated source, the compiler options and the
// kernel headers, in $CPV_PROGRAM_CACHE or rtc_cache/ next to the library.
uint64_t fnv1a(uint64_t h, const void* p, size_t n)
{
    const unsigned char* b = (const unsigned char*)p;
    for (size_t i = 0; i < n; i++) {
        h ^= b[i];
        h *= 0x100000001B3ull;
    }
    return h;
}

std::string cache_key(const std::string& src)
{
    uint64_t h = 0xCBF29CE484222325ull;
    h = fnv1a(h, src.data(), src.size());
    for (int i = 0; i < kRtcOptionCount; i++) h = fnv1a(h, kRtcOptions[i], std::strlen(kRtcOptions[i]) + 1);
    for (int i = 0; i < cpv_rtc_header_count; i++) h = fnv1a(h, cpv_rtc_header_sources[i], std::strlen(cpv_rtc_header_sources[i]));
    char buf[32];
    std::snprintf(buf, sizeof(buf), "%016llx", (unsigned long long)h);
    return buf;
}

std::string cache_dir()
{
    const char* env = std::getenv("CPV_PROGRAM_CACHE");
    if (env && *env) return env;
    Dl_info info;
    if (dladdr((const void*)&cache_key, &info) && info.dli_fname) {
        std::string p = info.dli_fname;
        const size_t slash = p.rfind('/');
        return (slash == std::string::npos ? std::string(".") : p.substr(0, slash)) + "/rtc_cache";
    }
    return "";
}

bool cache_load(const std::string& key, std::vector<char>* code, std::vector<std::string>* lowered)
{
    const std::string dir = cache_dir();
    if (dir.empty()) return false;
    std::ifstream co(dir + "/" + key + ".co", std::ios::binary), nm(dir + "/" + key + ".names");
    if (!co || !nm) return false;
    code->assign(std::istreambuf_iterator<char>(co), std::istreambuf_iterator<char>());
    std::string line;
    while (std::getline(nm, line))
        if (!line.empty()) lowered->push_back(line);
    return !code->empty() && lowered->size() == 1 + (size_t)cpv::kRasterVariants;
}

void cache_store(const std::string& key, const std::vector<char>& code, const std::vector<std::string>& lowered)
{
    const std::string dir = cache_dir();
    if (dir.empty()) return;
    mkdir(dir.c_str(), 0755);
    // write under temporary names, then rename: readers never see partial files
    const std::string tmp = dir + "/" + key + ".tmp" + std::to_string((long long)getpid());
    {
        std::ofstream co(tmp + ".co", std::ios::binary);
        co.write(code.data(), (std::streamsize)code.size());
        std::ofstream nm(tmp + ".names");
        for (const std::string& n : lowered) nm << n << "\n";
        if (!co || !nm) return;
    }
    std::rename((tmp + ".names").c_str(), (dir + "/" + key + ".names").c_str());
    std::rename((tmp + ".co").c_str(), (dir + "/" + key + ".co").c_str());
}

// hiprtc: the kernels with the library's floating-point flags; returns the
// code object and the lowered kernel names (from the cache when present).
int rtc_compile(const std::string& src, std::vector<char>* code, std::vector<std::string>* lowered, std::string* log, std::string* err,
                bool* cached)
{
    const std::string key = cache_key(src);
    *cached = cache_load(key, code, lowered);
    if (*cached) return CPV_OK;
    code->clear();
    lowered->clear();
    hiprtcProgram prog;
    if (hiprtcCreateProgram(&prog, src.c_str(), "cpv_spirv_program.hip", cpv_rtc_header_count, cpv_rtc_header_sources,
                            cpv_rtc_header_names) != HIPRTC_SUCCESS) {
        *err = "hiprtcCreateProgram failed";
        return CPV_ERROR_COMPILE;
    }
    const std::vector<std::string> names = kernel_names();
    for (const std::string& n : names) hiprtcAddNameExpression(prog, n.c_str());
    const hiprtcResult rc = hiprtcCompileProgram(prog, kRtcOptionCount, kRtcOptions);
    size_t log_n = 0;
    hiprtcGetProgramLogSize(prog, &log_n);
    if (log_n > 1) {
        log->resize(log_n);
        hiprtcGetProgramLog(prog, &(*log)[0]);
    }
    if (rc != HIPRTC_SUCCESS) {
        *err = "hiprtc compilation failed: " + *log;
        hiprtcDestroyProgram(&prog);
        return CPV_ERROR_COMPILE;
    }
    for (const std::string& n : names) {
        const char* low = nullptr;
        if (hiprtcGetLoweredName(prog, n.c_str(), &low) != HIPRTC_SUCCESS || !low) {
            *err = "no lowered name for " + n;
            hiprtcDestroyProgram(&prog);
            return CPV_ERROR_COMPILE;
        }
        lowered->push_back(low);
    }
    size_t size = 0;
    hiprtcGetCodeSize(prog, &size);
    code->resize(size);
    hiprtcGetCode(prog, code->data());
    hiprtcDestroyProgram(&prog);
    cache_store(key, *code, *lowered);
    return CPV_OK;
}

thread_local std::string g_translate_error;

}  // namespace

int cpv_spirv_compile(cpv_context* c, const uint32_t* vs, size_t vs_n, const char* vs_entry, const uint32_t* fs, size_t fs_n,
                      const char* fs_entry, cpv_program** out, std::string* err)
{
    (void)c;
    Translation t;
    try {
        t = translate(vs, vs_n, vs_entry, fs, fs_n, fs_entry);
    } catch (const Fail& f) {
        *err = "SPIR-V translation: " + f.msg;
        return CPV_ERROR_UNSUPPORTED;
    }
    std::vector<char> code;
    std::vector<std::string> lowered;
    std::string log;
    bool cached = false;
    int rc = rtc_compile(t.source, &code, &lowered, &log, err, &cached);
    if (rc) return rc;
    cpv_program* p = new cpv_program();
    p->source = t.source;
    p->log = log;
    p->vs_outputs = t.vs_outputs;
    p->fs_inputs = t.fs_inputs;
    p->psize = t.psize;
    hipError_t e = hipModuleLoadData(&p->module, code.data());
    if (e != hipSuccess) {
        *err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
        delete p;
        return CPV_ERROR_DEVICE;
    }
    e = hipModuleGetFunction(&p->vertex.module_fn, p->module, lowered[0].c_str());
    for (int i = 0; i < cpv::kRasterVariants && e == hipSuccess; i++)
        e = hipModuleGetFunction(&p->raster[i].module_fn, p->module, lowered[1 + i].c_str());
    if (e != hipSuccess) {
        *err = std::string("hipModuleGetFunction: ") + hipGetErrorString(e);
        (void)hipModuleUnload(p->module);
        delete p;
        return CPV_ERROR_DEVICE;
    }
    *out = p;
    return CPV_OK;
}

extern "C" int cpv_spirv_translate(const uint32_t* vs_words, size_t vs_word_count, const char* vs_entry, const uint32_t* fs_words,
                                   size_t fs_word_count, const char* fs_entry, char* source, size_t capacity, size_t* length)
{
    if (!vs_words || !fs_words || !length) return CPV_ERROR_INVALID;
    try {
        const Translation t = translate(vs_words, vs_word_count, vs_entry, fs_words, fs_word_count, fs_entry);
        *length = t.source.size();
        if (source && capacity) {
            const size_t n = std::min(capacity - 1, t.source.size());
            std::memcpy(source, t.source.data(), n);
            source[n] = '\0';
        }
        g_translate_error.clear();
        return CPV_OK;
    } catch (const Fail& f) {
        g_translate_error = f.msg;
        *length = 0;
        return CPV_ERROR_UNSUPPORTED;
    }
}

extern "C" const char* cpv_spirv_translate_error(void) { return g_translate_error.c_str(); }

extern "C" int cpv_spirv_build(const uint32_t* vs_words, size_t vs_word_count, const char* vs_entry, const uint32_t* fs_words,
                               size_t fs_word_count, const char* fs_entry, size_t* code_size)
{
    if (!vs_words || !fs_words || !code_size) return CPV_ERROR_INVALID;
    try {
        const Translation t = translate(vs_words, vs_word_count, vs_entry, fs_words, fs_word_count, fs_entry);
        std::vector<char> code;
        std::vector<std::string> lowered;
        std::string log, err;
        bool cached = false;
        const int rc = rtc_compile(t.source, &code, &lowered, &log, &err, &cached);
        g_translate_error = err;
        *code_size = rc ? 0 : code.size();
        return rc;
    } catch (const Fail& f) {
        g_translate_error = f.msg;
        *code_size = 0;
        return CPV_ERROR_UNSUPPORTED;
    }
}
