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
// Supported: logical addressing; float scalars, vectors, column-major
// matrices (MatrixStride = rows * 4); uniform blocks and combined image
// samplers of descriptor set 0, bindings 0-3; vertex inputs and
// varyings by location; gl_PerVertex / Position, VertexIndex,
// InstanceIndex, FragCoord, FrontFacing; structured or unstructured
// control flow (labels become goto targets, OpPhi becomes assignments on
// the incoming edges); OpKill.  Anything else is reported with its
// opcode.
#include <dlfcn.h>
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <cstdint>
#include <cstring>
#include <map>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cpv.h"
#include "cpv_program.h"

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
};

enum : uint32_t {
    DecBlock = 2, DecColMajor = 5, DecMatrixStride = 7, DecBuiltIn = 11, DecNoPerspective = 13, DecFlat = 14,
    DecLocation = 30, DecBinding = 33, DecDescriptorSet = 34, DecOffset = 35,
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
    int location = -1, binding = -1, set = -1, builtin = -1;
    bool flat = false, noperspective = false, block = false;
    std::map<uint32_t, MemberDec> members;
};

struct Inst {
    uint32_t op;
    std::vector<uint32_t> w;   // operands
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
    std::map<uint32_t, std::pair<uint32_t, uint32_t>> globals;               // var id -> (pointer type, storage class)
    std::vector<Inst> body;                                                  // the entry function's instructions
    std::map<uint32_t, uint32_t> result_type;                                // id -> type id (body results)
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
    uint32_t fn_id = 0;
    std::map<uint32_t, std::vector<Inst>> functions;
    while (i < n) {
        const uint32_t len = w[i] >> 16, op = w[i] & 0xFFFFu;
        if (len == 0 || i + len > n) bad("truncated instruction");
        Inst in{op, std::vector<uint32_t>(w + i + 1, w + i + len)};
        i += len;
        if (in_fn) {
            if (op == OpFunctionEnd) {
                functions[fn_id] = fn_insts;
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
        case OpString: case OpLine: case OpNoLine: case OpModuleProcessed:
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
        case OpConstant:
            m.consts[o[1]] = {o[0], {o[2]}};
            break;
        case OpConstantTrue: case OpConstantFalse:
            m.consts[o[1]] = {o[0], {op == OpConstantTrue ? 1u : 0u}};
            break;
        case OpConstantComposite:
            m.consts[o[1]] = {o[0], std::vector<uint32_t>(o.begin() + 2, o.end())};
            m.const_composite[o[1]] = true;
            break;
        case OpVariable:
            m.globals[o[1]] = {o[0], o[2]};
            break;
        case OpFunction:
            in_fn = true;
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
    if (functions.size() != 1) bad("shaders with function calls are not supported (expected only the entry function)");
    if (!functions.count(m.entry)) bad("entry function body missing");
    m.body = functions[m.entry];
    for (const Inst& in : m.body) {
        if (in.op == OpFunctionParameter) bad("entry function takes parameters");
        if (in.op == OpFunctionCall) bad("OpFunctionCall is not supported");
    }
    return m;
}

// ------------------------------------------------------------ code generation

struct Gen {
    const Module& m;
    bool vertex;
    std::ostringstream decl, code, pro, epi;
    std::map<uint32_t, std::string> lval;    // pointer id -> lvalue expression
    std::map<uint32_t, uint32_t> ptr_pointee;
    int out_total = 0;                       // vertex: varying components written
    std::map<int, int> loc_offset;           // location -> first varying component
    std::map<int, int> loc_size;
    bool discard = false;
    std::vector<int> interp;                 // fragment: mode per input component

    Gen(const Module& mm, bool vs) : m(mm), vertex(vs) {}

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

    std::string ctype(uint32_t t) const
    {
        const Type& ty = T(t);
        switch (ty.op) {
        case OpTypeVoid: return "void";
        case OpTypeBool: return "bool";
        case OpTypeFloat: return "float";
        case OpTypeInt: return ty.b ? "int32_t" : "uint32_t";
        case OpTypeVector: return "V<" + ctype(ty.a) + ", " + std::to_string(ty.b) + ">";
        case OpTypeMatrix: {
            const Type& col = T(ty.a);
            if (T(col.a).op != OpTypeFloat) bad("non-float matrix");
            return "M<" + std::to_string(col.b) + ", " + std::to_string(ty.b) + ">";
        }
        case OpTypeArray: return "A<" + ctype(ty.a) + ", " + std::to_string(const_u32(ty.b)) + ">";
        case OpTypeSampledImage: return "const TextureDesc*";
        default: bad("unsupported value type");
        }
    }

    uint32_t const_u32(uint32_t id) const
    {
        auto it = m.consts.find(id);
        if (it == m.consts.end() || m.const_composite.count(id)) bad("expected a scalar constant");
        return it->second.second[0];
    }

    std::string val(uint32_t id) const
    {
        auto it = m.consts.find(id);
        if (it != m.consts.end()) return const_expr(id);
        return "r" + std::to_string(id);
    }

    std::string const_expr(uint32_t id) const
    {
        const auto& c = m.consts.at(id);
        const Type& ty = T(c.first);
        if (m.const_composite.count(id)) {
            std::string s = ctype(c.first) + "{{";
            for (size_t i = 0; i < c.second.size(); i++) s += (i ? ", " : "") + const_expr(c.second[i]);
            return s + "}}";
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
        bad("element access into a non-composite");
    }

    void globals()
    {
        // varying layout: locations in increasing order, components packed
        std::vector<std::pair<int, uint32_t>> io;   // (location, var id)
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
                        else if (bi != BiPointSize && bi != BiClipDistance && bi != BiCullDistance)
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
                    else if (!vertex && sc == ScInput && d.builtin == BiFragCoord)
                        pro << "    " << name << " = V<float, 4>{{fc[0], fc[1], fc[2], fc[3]}};\n";
                    else if (!vertex && sc == ScInput && d.builtin == BiFrontFacing) pro << "    " << name << " = front;\n";
                    else bad("unsupported built-in variable");
                    continue;
                }
                if (d.location < 0) bad("interface variable without a location");
                const uint32_t et = is_vec(pointee) ? T(pointee).a : pointee;
                if (T(et).op != OpTypeFloat) bad("only float interface variables are supported");
                decl << "    " << ctype(pointee) << " " << name << "{};\n";
                lval[var] = name;
                io.push_back({d.location * 2 + (sc == ScOutput ? 1 : 0), var});
                continue;
            }
            if (sc == ScUniform) {
                if (d.set > 0 || d.binding < 0 || d.binding > 3) bad("uniform blocks must be set 0, binding 0-3");
                if (T(pointee).op != OpTypeStruct) bad("uniform variable is not a block");
                lval[var] = "@U" + std::to_string(d.binding) + ":" + std::to_string(pointee);
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
            const std::string name = "g" + std::to_string(var);
            if (vertex && !out) {
                // vertex attribute of `n` float components (EmitCopyInput, PipelineCompiler.cpp:821-896)
                if (is_vec(pointee)) pro << "    fetch_attribute(env, " << loc << ", " << n << ", vid, iid, " << name << ".c);\n";
                else pro << "    { float t_[1]; fetch_attribute(env, " << loc << ", 1, vid, iid, t_); " << name << " = t_[0]; }\n";
            } else if (vertex && out) {
                loc_offset[loc] = off;
                loc_size[loc] = n;
                for (int k = 0; k < n; k++)
                    epi << "    out[" << off + k << "] = " << name << (is_vec(pointee) ? ".c[" + std::to_string(k) + "]" : "") << ";\n";
                off += n;
            } else if (!vertex && !out) {
                loc_offset[loc] = -1;
                loc_size[loc] = n;
            } else {
                if (loc != 0) bad("fragment outputs other than location 0 are not supported");
                if (n != 4) bad("the fragment colour output must be a vec4");
                for (int k = 0; k < 4; k++) epi << "    out[" << k << "] = " << name << ".c[" << k << "];\n";
            }
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
            for (int k = 0; k < n; k++)
                pro << "    " << name << (is_vec(pointee) ? ".c[" + std::to_string(k) + "]" : "") << " = in[" << off + k << "];\n";
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
        if (s.size() > 2 && s[0] == '@' && s[1] == 'U') {
            // uniform block: byte offset arithmetic on env.uniform[binding]
            const size_t colon = s.find(':');
            const int binding = std::stoi(s.substr(2, colon - 2));
            const uint32_t block = (uint32_t)std::stoul(s.substr(colon + 1));
            if (idx.empty()) bad("load of a whole uniform block");
            const uint32_t mi = const_u32(idx[0]);
            const Type& bt = T(block);
            if (mi >= bt.members.size()) bad("uniform member index");
            const auto& md = m.dec.at(block).members;
            auto mdi = md.find(mi);
            if (mdi == md.end() || mdi->second.offset < 0) bad("uniform member without an offset");
            std::string off = std::to_string(mdi->second.offset);
            t = bt.members[mi];
            if (T(t).op == OpTypeMatrix) {
                const Type& col = T(T(t).a);
                if (mdi->second.matrix_stride != (int)col.b * 4) bad("uniform matrix stride must equal 4 * rows");
            }
            for (i = 1; i < idx.size(); i++) {
                const Type& ty = T(t);
                const std::string ix = val(idx[i]);
                if (ty.op == OpTypeMatrix) {
                    off = "(" + off + " + 16 * (" + ix + "))";
                    t = ty.a;
                } else if (ty.op == OpTypeVector) {
                    off = "(" + off + " + 4 * (" + ix + "))";
                    t = ty.a;
                } else {
                    bad("unsupported uniform member type");
                }
            }
            *out_t = t;
            return "(*(const " + ctype(t) + "*)(env.uniform[" + std::to_string(binding) + "] + " + off + "))";
        }
        const Type& bt = T(t);
        if (bt.op == OpTypeStruct) {
            // gl_PerVertex members
            if (idx.empty()) bad("access to a whole block");
            const uint32_t mi = const_u32(idx[0]);
            s = s + "_m" + std::to_string(mi);
            t = bt.members.at(mi);
            i = 1;
        }
        for (; i < idx.size(); i++) s = elem(t, s, val(idx[i]), &t);
        *out_t = t;
        return s;
    }

    static bool is_float_op(uint32_t op)
    {
        return op == OpFAdd || op == OpFSub || op == OpFMul || op == OpFDiv || op == OpIAdd || op == OpISub || op == OpIMul ||
               op == OpUDiv || op == OpSDiv;
    }

    static const char* binop(uint32_t op)
    {
        switch (op) {
        case OpFAdd: case OpIAdd: return "+";
        case OpFSub: case OpISub: return "-";
        case OpFMul: case OpIMul: return "*";
        case OpFDiv: case OpUDiv: case OpSDiv: return "/";
        case OpFOrdEqual: case OpIEqual: return "==";
        case OpFOrdNotEqual: case OpINotEqual: return "!=";
        case OpFOrdLessThan: case OpULessThan: case OpSLessThan: return "<";
        case OpFOrdGreaterThan: case OpUGreaterThan: case OpSGreaterThan: return ">";
        case OpFOrdLessThanEqual: case OpULessThanEqual: case OpSLessThanEqual: return "<=";
        case OpFOrdGreaterThanEqual: case OpUGreaterThanEqual: case OpSGreaterThanEqual: return ">=";
        case OpLogicalAnd: return "&&";
        case OpLogicalOr: return "||";
        default: return nullptr;
        }
    }

    // component k of value id (scalar values broadcast)
    std::string comp(uint32_t id, uint32_t t, int k) const
    {
        return is_vec(t) ? val(id) + ".c[" + std::to_string(k) + "]" : val(id);
    }

    void body()
    {
        // OpPhi: per target block, per incoming block, the (result, value)
        // assignments made on that edge
        std::map<uint32_t, std::map<uint32_t, std::vector<std::pair<uint32_t, uint32_t>>>> phis;
        {
            uint32_t blk = 0;
            for (const Inst& in : m.body) {
                if (in.op == OpLabel) blk = in.w[0];
                if (in.op == OpPhi)
                    for (size_t k = 2; k + 1 < in.w.size(); k += 2) phis[blk][in.w[k + 1]].push_back({in.w[1], in.w[k]});
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
        for (const Inst& in : m.body) {
            const std::vector<uint32_t>& o = in.w;
            const uint32_t op = in.op;
            // declare results
            auto result = [&](uint32_t type, uint32_t id) {
                decl << "    " << ctype(type) << " r" << id << "{};\n";
            };
            switch (op) {
            case OpLabel:
                cur = o[0];
                code << "L" << o[0] << ":;\n";
                break;
            case OpVariable: {
                if (o[2] != ScFunction) bad("non-Function variable inside a function");
                const uint32_t pointee = T(o[0]).b;
                decl << "    " << ctype(pointee) << " v" << o[1] << "{};\n";
                lval[o[1]] = "v" + std::to_string(o[1]);
                ptr_pointee[o[1]] = pointee;
                if (o.size() > 3) code << "    v" << o[1] << " = " << val(o[3]) << ";\n";
                break;
            }
            case OpLoad: {
                auto it = lval.find(o[2]);
                if (it == lval.end()) bad("load through an unknown pointer");
                result(o[0], o[1]);
                code << "    r" << o[1] << " = " << it->second << ";\n";
                break;
            }
            case OpStore: {
                auto it = lval.find(o[0]);
                if (it == lval.end() || it->second[0] == '(') bad("store through a read-only or unknown pointer");
                code << "    " << it->second << " = " << val(o[1]) << ";\n";
                break;
            }
            case OpAccessChain: {
                uint32_t t = 0;
                const std::string s = chain(o[2], std::vector<uint32_t>(o.begin() + 3, o.end()), &t);
                lval[o[1]] = s;
                ptr_pointee[o[1]] = t;
                break;
            }
            case OpCompositeExtract: {
                result(o[0], o[1]);
                uint32_t t = m.result_type.count(o[2]) ? m.result_type.at(o[2]) : type_of(o[2]);
                std::string s = val(o[2]);
                for (size_t k = 3; k < o.size(); k++) s = elem(t, s, std::to_string(o[k]), &t);
                code << "    r" << o[1] << " = " << s << ";\n";
                break;
            }
            case OpCompositeConstruct: {
                result(o[0], o[1]);
                const Type& rt = T(o[0]);
                if (rt.op == OpTypeVector) {
                    int k = 0;
                    for (size_t a = 2; a < o.size(); a++) {
                        const uint32_t at = type_of(o[a]);
                        for (int c = 0; c < comps(at); c++)
                            code << "    r" << o[1] << ".c[" << k++ << "] = " << comp(o[a], at, c) << ";\n";
                    }
                } else if (rt.op == OpTypeMatrix) {
                    for (size_t a = 2; a < o.size(); a++) code << "    r" << o[1] << ".col[" << a - 2 << "] = " << val(o[a]) << ";\n";
                } else {
                    bad("composite construct of an unsupported type");
                }
                break;
            }
            case OpVectorShuffle: {
                result(o[0], o[1]);
                const uint32_t t1 = type_of(o[2]), t2 = type_of(o[3]);
                const int n1 = comps(t1);
                for (size_t k = 4; k < o.size(); k++) {
                    const uint32_t s = o[k];
                    if (s == 0xFFFFFFFFu) continue;
                    code << "    r" << o[1] << ".c[" << k - 4 << "] = "
                         << ((int)s < n1 ? comp(o[2], t1, (int)s) : comp(o[3], t2, (int)s - n1)) << ";\n";
                }
                break;
            }
            case OpFNegate: case OpSNegate: {
                result(o[0], o[1]);
                const int n = comps(o[0]);
                for (int k = 0; k < n; k++)
                    code << "    r" << o[1] << (is_vec(o[0]) ? ".c[" + std::to_string(k) + "]" : "") << " = -" << comp(o[2], o[0], k) << ";\n";
                break;
            }
            case OpLogicalNot: {
                result(o[0], o[1]);
                const int n = comps(o[0]);
                for (int k = 0; k < n; k++)
                    code << "    r" << o[1] << (is_vec(o[0]) ? ".c[" + std::to_string(k) + "]" : "") << " = !" << comp(o[2], o[0], k) << ";\n";
                break;
            }
            case OpConvertFToS: case OpConvertFToU: case OpConvertSToF: case OpConvertUToF: case OpBitcast: {
                result(o[0], o[1]);
                const int n = comps(o[0]);
                const uint32_t et = is_vec(o[0]) ? T(o[0]).a : o[0];
                const uint32_t st = type_of(o[2]);
                for (int k = 0; k < n; k++) {
                    const std::string src = comp(o[2], st, k);
                    const std::string dt = ctype(et), sty = ctype(is_vec(st) ? T(st).a : st);
                    std::string e = "(" + dt + ")(" + src + ")";
                    if (op == OpBitcast && dt == "float" && sty != "float") e = "__uint_as_float((uint32_t)(" + src + "))";
                    if (op == OpBitcast && dt != "float" && sty == "float") e = "(" + dt + ")__float_as_uint(" + src + ")";
                    code << "    r" << o[1] << (is_vec(o[0]) ? ".c[" + std::to_string(k) + "]" : "") << " = " << e << ";\n";
                }
                break;
            }
            case OpVectorTimesScalar: {
                result(o[0], o[1]);
                for (int k = 0; k < comps(o[0]); k++)
                    code << "    r" << o[1] << ".c[" << k << "] = " << val(o[2]) << ".c[" << k << "] * " << val(o[3]) << ";\n";
                break;
            }
            case OpMatrixTimesVector: {
                result(o[0], o[1]);
                const Type& mt = T(type_of(o[2]));
                if (mt.b != 4 || T(mt.a).b != 4) bad("OpMatrixTimesVector is supported for mat4 * vec4");
                code << "    mat4_mul_vec4(" << val(o[2]) << ".col[0].c, " << val(o[3]) << ".c, r" << o[1] << ".c);\n";
                break;
            }
            case OpDot: {
                result(o[0], o[1]);
                const int n = comps(type_of(o[2]));
                code << "    r" << o[1] << " = dot_n(" << val(o[2]) << ".c, " << val(o[3]) << ".c, " << n << ");\n";
                break;
            }
            case OpSelect: {
                result(o[0], o[1]);
                const uint32_t ct = type_of(o[2]);
                for (int k = 0; k < comps(o[0]); k++)
                    code << "    r" << o[1] << (is_vec(o[0]) ? ".c[" + std::to_string(k) + "]" : "") << " = " << comp(o[2], ct, k)
                         << " ? " << comp(o[3], o[0], k) << " : " << comp(o[4], o[0], k) << ";\n";
                break;
            }
            case OpImageSampleImplicitLod: {
                result(o[0], o[1]);
                if (o.size() > 4) bad("image operands on OpImageSampleImplicitLod are not supported");
                code << "    sample_implicit(*" << val(o[2]) << ", " << val(o[3]) << ".c[0], " << val(o[3]) << ".c[1], r" << o[1] << ".c);\n";
                break;
            }
            case OpExtInst: {
                if (o[2] != m.glsl) bad("extended instruction from an unknown set");
                result(o[0], o[1]);
                ext(o);
                break;
            }
            case OpPhi:
                result(o[0], o[1]);
                break;
            case OpSelectionMerge: case OpLoopMerge:
                break;
            case OpBranch:
                code << "    " << edge(o[0]) << "\n";
                break;
            case OpBranchConditional:
                code << "    if (" << val(o[0]) << ") { " << edge(o[1]) << " } else { " << edge(o[2]) << " }\n";
                break;
            case OpKill:
                if (vertex) bad("OpKill in a vertex shader");
                discard = true;
                code << "    return true;\n";
                break;
            case OpReturn:
                code << "    goto L_end;\n";
                break;
            default: {
                const char* b = binop(op);
                if (!b) {
                    std::ostringstream e;
                    e << "unsupported instruction, opcode " << op;
                    bad(e.str());
                }
                result(o[0], o[1]);
                const uint32_t at = type_of(o[2]), bt2 = type_of(o[3]);
                const int n = std::max(comps(at), comps(bt2));
                for (int k = 0; k < n; k++)
                    code << "    r" << o[1] << (is_vec(o[0]) ? ".c[" + std::to_string(k) + "]" : "") << " = " << comp(o[2], at, k) << " "
                         << b << " " << comp(o[3], bt2, k) << ";\n";
                break;
            }
            }
            if (op != OpLabel && op != OpVariable && op != OpStore && op != OpAccessChain && op != OpSelectionMerge &&
                op != OpLoopMerge && op != OpBranch && op != OpBranchConditional && op != OpKill && op != OpReturn && o.size() >= 2) {
                const_cast<Module&>(m).result_type[o[1]] = o[0];
            }
        }
    }

    uint32_t type_of(uint32_t id) const
    {
        auto c = m.consts.find(id);
        if (c != m.consts.end()) return c->second.first;
        auto r = m.result_type.find(id);
        if (r != m.result_type.end()) return r->second;
        bad("value " + std::to_string(id) + " used before its definition");
    }

    void ext(const std::vector<uint32_t>& o)
    {
        const uint32_t inst = o[3];
        const uint32_t rt = o[0];
        const std::string r = "r" + std::to_string(o[1]);
        const int n = comps(rt);
        auto c = [&](size_t a, int k) { return comp(o[a], type_of(o[a]), k); };
        auto each = [&](const std::string& f) {
            for (int k = 0; k < n; k++) code << "    " << r << (is_vec(rt) ? ".c[" + std::to_string(k) + "]" : "") << " = " << f.substr(0, f.find('#'))
                                             << c(4, k) << f.substr(f.find('#') + 1) << ";\n";
        };
        switch (inst) {
        case 4: each("fabsf(#)"); break;                                                    // FAbs
        case 8: each("floorf(#)"); break;                                                   // Floor
        case 10: each("spv_fract(#)"); break;                                               // Fract
        case 31: each("sqrtf(#)"); break;                                                   // Sqrt
        case 32: each("1.0f / sqrtf(#)"); break;                                            // InverseSqrt
        case 37: case 40: case 43:                                                          // FMin FMax FClamp
            for (int k = 0; k < n; k++) {
                code << "    " << r << (is_vec(rt) ? ".c[" + std::to_string(k) + "]" : "") << " = "
                     << (inst == 37 ? "spv_fmin(" : inst == 40 ? "spv_fmax(" : "spv_fclamp(") << c(4, k) << ", " << c(5, k);
                if (inst == 43) code << ", " << c(6, k);
                code << ");\n";
            }
            break;
        case 66: {                                                                          // Length
            const uint32_t at = type_of(o[4]);
            if (is_vec(at)) code << "    " << r << " = sqrtf(dot_n(" << val(o[4]) << ".c, " << val(o[4]) << ".c, " << comps(at) << "));\n";
            else code << "    " << r << " = fabsf(" << val(o[4]) << ");\n";
            break;
        }
        case 69: {                                                                          // Normalize
            if (!is_vec(rt)) bad("Normalize of a scalar");
            code << "    { const float inv_ = 1.0f / sqrtf(dot_n(" << val(o[4]) << ".c, " << val(o[4]) << ".c, " << n << "));\n";
            for (int k = 0; k < n; k++) code << "      " << r << ".c[" << k << "] = " << val(o[4]) << ".c[" << k << "] * inv_;\n";
            code << "    }\n";
            break;
        }
        default: {
            std::ostringstream e;
            e << "unsupported GLSL.std.450 instruction " << inst;
            bad(e.str());
        }
        }
    }
};

const char* kPrelude = R"(#include "raster_kernels.h"
namespace cpv {
template <class T, int N> struct V { T c[N]; };
template <int R, int C> struct M { V<float, R> col[C]; };
template <class T, int N> struct A { T e[N]; };
// std::max / std::min / std::clamp (GlslFunctions.cpp:118-142)
__device__ __forceinline__ float spv_fmax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float spv_fmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float spv_fclamp(float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); }
__device__ __forceinline__ float spv_fract(float v) { return v - floorf(v); }
)";

struct Translation {
    std::string source;
    int vs_outputs = 0, fs_inputs = 0;
};

Translation translate(const uint32_t* vs, size_t vs_n, const char* vs_entry, const uint32_t* fs, size_t fs_n, const char* fs_entry)
{
    Module mv = parse(vs, vs_n, vs_entry);
    Module mf = parse(fs, fs_n, fs_entry);
    if (mv.model != 0) bad("the first module's entry point is not a vertex shader");
    if (mf.model != 4) bad("the second module's entry point is not a fragment shader");
    Gen gv(mv, true), gf(mf, false);
    gv.globals();
    gv.body();
    gf.globals();
    gf.bind_fragment_inputs(gv.loc_offset, gv.loc_size);
    gf.body();
    if (!mf.origin_upper_left) bad("fragment shaders must declare OriginUpperLeft (Vulkan requires it)");
    std::ostringstream s;
    s << kPrelude;
    s << "struct SpvVs {\n    static constexpr int kOutputs = " << gv.out_total << ";\n"
      << "    __device__ static void run(const DrawEnv& env, uint32_t vid, uint32_t iid, float pos[4], float* out)\n    {\n"
      << gv.decl.str() << gv.pro.str() << gv.code.str() << "L_end:;\n" << gv.epi.str()
      << "    (void)env; (void)vid; (void)iid; (void)out; (void)pos;\n    }\n};\n";
    const int fin = (int)gf.interp.size();
    s << "struct SpvFs {\n    static constexpr int kInputs = " << fin << ";\n"
      << "    static constexpr bool kMayDiscard = " << (gf.discard ? "true" : "false") << ";\n"
      << "    __device__ static int interp(int c)\n    {\n";
    for (int c = 0; c < fin; c++)
        if (gf.interp[c]) s << "        if (c == " << c << ") return " << gf.interp[c] << ";\n";
    s << "        (void)c;\n        return 0;\n    }\n"
      << "    __device__ static bool run(const DrawEnv& env, const float* in, const float* fc, bool front, float out[4])\n    {\n"
      << gf.decl.str() << gf.pro.str() << gf.code.str() << "L_end:;\n" << gf.epi.str()
      << "    (void)env; (void)in; (void)fc; (void)front;\n        return false;\n    }\n};\n";
    s << "}  // namespace cpv\n";
    s << "template __global__ void cpv::vertex_kernel<cpv::SpvVs>(cpv::VertexArgs);\n";
    static const int cfg[7][2] = {{8, 1}, {16, 1}, {16, 2}, {16, 4}, {32, 1}, {32, 2}, {32, 4}};
    for (const auto& c : cfg)
        s << "template __global__ void cpv::raster_kernel<cpv::SpvFs, " << c[0] << ", " << c[1] << ">(cpv::RasterArgs);\n";
    Translation t;
    t.source = s.str();
    t.vs_outputs = gv.out_total;
    t.fs_inputs = fin;
    return t;
}

std::vector<std::string> kernel_names()
{
    std::vector<std::string> n{"cpv::vertex_kernel<cpv::SpvVs>"};
    static const int cfg[7][2] = {{8, 1}, {16, 1}, {16, 2}, {16, 4}, {32, 1}, {32, 2}, {32, 4}};
    for (const auto& c : cfg)
        n.push_back("cpv::raster_kernel<cpv::SpvFs, " + std::to_string(c[0]) + ", " + std::to_string(c[1]) + ">");
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
    return !code->empty() && lowered->size() == 8;
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
    hipError_t e = hipModuleLoadData(&p->module, code.data());
    if (e != hipSuccess) {
        *err = std::string("hipModuleLoadData: ") + hipGetErrorString(e);
        delete p;
        return CPV_ERROR_DEVICE;
    }
    e = hipModuleGetFunction(&p->vertex.module_fn, p->module, lowered[0].c_str());
    for (int i = 0; i < 7 && e == hipSuccess; i++) e = hipModuleGetFunction(&p->raster[i].module_fn, p->module, lowered[1 + i].c_str());
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
