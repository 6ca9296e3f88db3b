// Built-in shader programs: device versions of the GLSL shaders the
// benchmark workloads use.  Each follows the operation order the
// reference's JIT produces for the corresponding SPIR-V
// (OpMatrixTimesVector -> glm mat4*vec4, OpDot -> glm::dot, GLSL.std.450
// Normalize -> glm::normalize, FMax -> std::max; LLVMRuntime/SpirvFunctions.cpp:5-81,
// CPVulkan/GlslFunctions.cpp:122-126, 179-183).
#pragma once
#pragma clang fp contract(off)

#include "raster_device.h"

namespace cpv {

// layout(location=0) in vec4 pos; layout(location=1) in vec4 color;
// gl_Position = pos; out vec4 = color.
struct VsPassthrough {
    static constexpr int kOutputs = 4;
    __device__ __forceinline__ static void run(const DrawEnv& env, uint32_t vid, uint32_t iid, float pos[4], float* out)
    {
        fetch_attribute(env, 0, 4, vid, iid, pos);
        fetch_attribute(env, 1, 4, vid, iid, out);
    }
};

// in vec4 pos @0, vec4 colour @1, float size @2: gl_Position = pos,
// gl_PointSize = size, the colour passed through (output 4 is the size).
struct VsPassthroughPoint {
    static constexpr int kOutputs = 5;
    static constexpr int kPointSize = 4;
    __device__ __forceinline__ static void run(const DrawEnv& env, uint32_t vid, uint32_t iid, float pos[4], float* out)
    {
        fetch_attribute(env, 0, 4, vid, iid, pos);
        fetch_attribute(env, 1, 4, vid, iid, out);
        fetch_attribute(env, 2, 1, vid, iid, out + 4);
    }
};

// Samples/15-draw_cube/15-draw_cube.cpp:38-51: gl_Position = mvp * pos.
struct VsMvpColor {
    static constexpr int kOutputs = 4;
    __device__ __forceinline__ static void run(const DrawEnv& env, uint32_t vid, uint32_t iid, float pos[4], float* out)
    {
        float p[4];
        fetch_attribute(env, 0, 4, vid, iid, p);
        mat4_mul_vec4((const float*)env.uniform[0], p, pos);
        fetch_attribute(env, 1, 4, vid, iid, out);
    }
};

// in vec3 pos @0, vec3 normal @1, vec2 uv @2: gl_Position = mvp * vec4(pos, 1).
struct VsMvpLambert {
    static constexpr int kOutputs = 5;
    __device__ __forceinline__ static void run(const DrawEnv& env, uint32_t vid, uint32_t iid, float pos[4], float* out)
    {
        float p[4];
        fetch_attribute(env, 0, 3, vid, iid, p);
        p[3] = 1.0f;
        mat4_mul_vec4((const float*)env.uniform[0], p, pos);
        fetch_attribute(env, 1, 3, vid, iid, out);
        fetch_attribute(env, 2, 2, vid, iid, out + 3);
    }
};

// Samples/15-draw_cube/15-draw_cube.cpp:53-61: outColor = color.
struct FsColor {
    static constexpr int kInputs = 4;
    static constexpr bool kMayDiscard = false;
    static constexpr int kWavesPerEU = 6;
    __device__ static int interp(int) { return 0; }
    __device__ __forceinline__ static bool run(const DrawEnv&, const float* in, const float*, bool, float out[4])
    {
        out[0] = in[0]; out[1] = in[1]; out[2] = in[2]; out[3] = in[3];
        return false;
    }
};

// out = color with the input Flat (mode 2) or NoPerspective (mode 1).
template <int MODE>
struct FsColorInterp {
    static constexpr int kInputs = 4;
    static constexpr bool kMayDiscard = false;
    static constexpr int kWavesPerEU = 6;
    __device__ static int interp(int) { return MODE; }
    __device__ __forceinline__ static bool run(const DrawEnv&, const float* in, const float*, bool, float out[4])
    {
        out[0] = in[0]; out[1] = in[1]; out[2] = in[2]; out[3] = in[3];
        return false;
    }
};
using FsColorFlat = FsColorInterp<2>;
using FsColorNoPerspective = FsColorInterp<1>;

// if (color.r < 0.25) discard; out = color.
struct FsColorDiscard {
    static constexpr int kInputs = 4;
    static constexpr bool kMayDiscard = true;
    static constexpr int kWavesPerEU = 6;
    __device__ static int interp(int) { return 0; }
    __device__ __forceinline__ static bool run(const DrawEnv&, const float* in, const float*, bool, float out[4])
    {
        if (in[0] < 0.25f) return true;
        out[0] = in[0]; out[1] = in[1]; out[2] = in[2]; out[3] = in[3];
        return false;
    }
};

// out = texture(tex@1, uv) * max(dot(normalize(n), L@2.xyz), 0).
struct FsTexLambert {
    static constexpr int kInputs = 5;
    static constexpr bool kMayDiscard = false;
    __device__ static int interp(int) { return 0; }
    __device__ __forceinline__ static bool run(const DrawEnv& env, const float* in, const float*, bool, float out[4])
    {
        float t[4];
        sample_implicit(env.tex[1], in[3], in[4], t);
        const float dd = dot_n(in, in, 3);
        const float inv = 1.0f / sqrtf(dd);
        const float n[3] = {in[0] * inv, in[1] * inv, in[2] * inv};
        const float l = dot_n(n, (const float*)env.uniform[2], 3);
        const float m = l < 0.0f ? 0.0f : l;
        out[0] = t[0] * m; out[1] = t[1] * m; out[2] = t[2] * m; out[3] = t[3] * m;
        return false;
    }
};

}  // namespace cpv
