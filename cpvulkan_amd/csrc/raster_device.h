// Device functions of the draw path: pixel format conversion, texture
// sampling, blend and the per-fragment depth/stencil/colour operations.
//
// Every floating-point expression is written in the operation order the
// reference executes (no FMA contraction: the file turns contraction off and
// the build passes -ffp-contract=off), so the GPU produces the same bits as
// the reference's CPU code.  The reference locations are cited per function
// (paths relative to the reference repository).
#pragma once
#pragma clang fp contract(off)

#include "cpv_types.h"

namespace cpv {

#define CPV_DEV __device__ __forceinline__

// ---------------------------------------------------------------- helpers

// static_cast<int32_t>(float) as the reference's x86-64 build executes it
// (cvttss2si): truncation toward zero, INT32_MIN for NaN or out of range.
CPV_DEV int32_t trunc_to_i32(float v)
{
    return (v >= -2147483648.0f && v < 2147483648.0f) ? (int32_t)v : (int32_t)0x80000000;
}

// llvm.maxnum / llvm.minnum: a NaN operand yields the other operand.
CPV_DEV float maxnum(float a, float b) { return fmaxf(a, b); }
CPV_DEV float minnum(float a, float b) { return fminf(a, b); }

// Byte extraction/insertion in a pixel held as 4 dwords (up to 16 bytes).
CPV_DEV uint32_t pick_word(const uint32_t px[4], int w)
{
    return w == 0 ? px[0] : w == 1 ? px[1] : w == 2 ? px[2] : px[3];
}
CPV_DEV uint32_t get_field(const uint32_t px[4], int byte_off, int nbytes)
{
    const uint32_t word = pick_word(px, byte_off >> 2);
    const int shift = (byte_off & 3) * 8;
    const uint32_t v = word >> shift;
    return nbytes >= 4 ? v : (v & ((1u << (nbytes * 8)) - 1u));
}
CPV_DEV void set_field(uint32_t px[4], int byte_off, int nbytes, uint32_t v)
{
    const int w = byte_off >> 2;
    const int shift = (byte_off & 3) * 8;
    const uint32_t mask = nbytes >= 4 ? 0xFFFFFFFFu : (((1u << (nbytes * 8)) - 1u) << shift);
    const uint32_t bits = (v << shift) & mask;
    if (w == 0) px[0] = (px[0] & ~mask) | bits;
    else if (w == 1) px[1] = (px[1] & ~mask) | bits;
    else if (w == 2) px[2] = (px[2] & ~mask) | bits;
    else px[3] = (px[3] & ~mask) | bits;
}

// ------------------------------------------------------ format conversion

// GetPixelF32 of a normal format (LLVMRuntime/ImageCompiler.cpp:161-230,
// 303-347): channels absent from the format read as (0, 0, 0, 1); integer
// channels convert with an fdiv by the type maximum.
CPV_DEV void unpack_color(const FormatDesc& f, const uint32_t px[4], float out[4])
{
    out[0] = 0.0f; out[1] = 0.0f; out[2] = 0.0f; out[3] = 1.0f;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (f.off[i] < 0) continue;
        const uint32_t raw = get_field(px, f.off[i], f.element);
        float v;
        if (f.base == kBaseSFloat) {
            v = __uint_as_float(raw);
        } else if (f.base == kBaseUNorm || f.base == kBaseUScaled || f.base == kBaseUInt) {
            // (float)u / 255.0f and (float)u / 65535.0f, computed as one double
            // multiply rounded to float: identical for every 8- and 16-bit code
            // (checked exhaustively), a quarter of the instructions of a
            // correctly rounded fp32 division.
            if (f.element == 1) v = (float)((double)raw * (1.0 / 255.0));
            else if (f.element == 2) v = (float)((double)raw * (1.0 / 65535.0));
            else v = (float)raw / 4294967296.0f;
        } else {
            const int32_t s = f.element == 1 ? (int32_t)(int8_t)raw : f.element == 2 ? (int32_t)(int16_t)raw : (int32_t)raw;
            const float maxv = f.element == 1 ? 127.0f : f.element == 2 ? 32767.0f : 2147483648.0f;
            v = (float)s / maxv;
        }
        out[i] = v;
    }
}

// SetPixelF32 of a normal format (ImageCompiler.cpp:1077-1114, 1179-1197):
// UNORM clamps with maxnum/minnum, scales by the type maximum, rounds half
// away from zero (llvm.round) and converts.
CPV_DEV void pack_color_f32(const FormatDesc& f, const float in[4], uint32_t px[4])
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (f.off[i] < 0) continue;
        float v = in[i];
        uint32_t raw;
        if (f.base == kBaseUNorm) {
            v = minnum(maxnum(v, 0.0f), 1.0f);
            raw = (uint32_t)roundf(v * (f.element == 1 ? 255.0f : 65535.0f));
        } else if (f.base == kBaseSNorm) {
            v = minnum(maxnum(v, -1.0f), 1.0f);
            raw = (uint32_t)(int32_t)roundf(v * (f.element == 1 ? 127.0f : 32767.0f));
        } else {
            raw = __float_as_uint(v);
        }
        set_field(px, f.off[i], f.element, raw);
    }
}

// SetPixelU32 / SetPixelI32 (ImageCompiler.cpp:1121-1175).
CPV_DEV void pack_color_u32(const FormatDesc& f, const uint32_t in[4], uint32_t px[4])
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (f.off[i] < 0) continue;
        uint32_t v = in[i];
        if (f.element == 1) v = v < 255u ? v : 255u;
        else if (f.element == 2) v = v < 65535u ? v : 65535u;
        set_field(px, f.off[i], f.element, v);
    }
}
CPV_DEV void pack_color_i32(const FormatDesc& f, const int32_t in[4], uint32_t px[4])
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        if (f.off[i] < 0) continue;
        int32_t v = in[i];
        if (f.element == 1) { v = v > -128 ? v : -128; v = v < 127 ? v : 127; }
        else if (f.element == 2) { v = v > -32768 ? v : -32768; v = v < 32767 ? v : 32767; }
        set_field(px, f.off[i], f.element, (uint32_t)v);
    }
}

// GetPixelDepth (ImageCompiler.cpp:541-606).  A depth/stencil pixel is held
// as two dwords: dword 0 = bytes 0-3, dword 1 = bytes 4-7.
CPV_DEV float unpack_depth(const FormatDesc& f, uint32_t d0)
{
    // same double-multiply form as unpack_color, exhaustively equal to the
    // fp32 divisions u / 65535.0f and u / 16777215.0f
    if (f.depth_kind == 16) return (float)((double)(d0 & 0xFFFFu) * (1.0 / 65535.0));
    if (f.depth_kind == 24) return (float)((double)(d0 & 0x00FFFFFFu) * (1.0 / 16777215.0));
    if (f.depth_kind == 32) return __uint_as_float(d0);
    return 0.0f;
}
CPV_DEV uint32_t unpack_stencil(const FormatDesc& f, uint32_t d0, uint32_t d1)
{
    if (f.stencil_off < 0) return 0u;
    if (f.stencil_off >= 4) return (d1 >> ((f.stencil_off - 4) * 8)) & 0xFFu;
    return (d0 >> (f.stencil_off * 8)) & 0xFFu;
}

// SetDSPixel (ImageCompiler.cpp:869-947).
CPV_DEV void pack_depth_stencil(const FormatDesc& f, float depth, uint32_t stencil, uint32_t& d0, uint32_t& d1)
{
    if (f.depth_kind == 16) {
        const float v = minnum(maxnum(depth, 0.0f), 1.0f);
        d0 = (d0 & 0xFFFF0000u) | ((uint32_t)roundf(v * 65535.0f) & 0xFFFFu);
    } else if (f.depth_kind == 24) {
        const float v = minnum(maxnum(depth, 0.0f), 1.0f) * 16777215.0f;
        uint32_t u = (uint32_t)roundf(v);
        if (f.stencil_off == 3) u |= (stencil & 0xFFu) << 24;   // D24_UNORM_S8_UINT
        d0 = u;
    } else if (f.depth_kind == 32) {
        d0 = __float_as_uint(depth);                              // raw store, no clamp
    }
    if (f.stencil_off >= 0 && !(f.depth_kind == 24 && f.stencil_off == 3)) {
        if (f.stencil_off >= 4) {
            const int sh = (f.stencil_off - 4) * 8;
            d1 = (d1 & ~(0xFFu << sh)) | ((stencil & 0xFFu) << sh);
        } else {
            const int sh = f.stencil_off * 8;
            d0 = (d0 & ~(0xFFu << sh)) | ((stencil & 0xFFu) << sh);
        }
    }
}

// ---------------------------------------------------------------- sampling

struct SamplerDesc {
    int32_t mag_filter, min_filter, mipmap_mode;
    int32_t address_u, address_v;
    float mip_lod_bias, min_lod, max_lod;
    int32_t border_color;
    int32_t pad;
};

struct TextureDesc {
    uint64_t address;
    FormatDesc fmt;
    uint32_t level_count;
    uint32_t pad;
    uint64_t level_offset[16];
    uint32_t level_w[16];
    uint32_t level_h[16];
    SamplerDesc sampler;
};

// wrap (CPVulkan/ImageSampler.cpp:12-38)
CPV_DEV int32_t wrap_coord(int32_t v, int32_t size, int32_t mode)
{
    if (mode == 0) return (v % size + size) % size;                       // REPEAT
    if (mode == 1) {                                                       // MIRRORED_REPEAT
        const int32_t two = 2 * size;
        const int32_t n = (v % two + two) % two - size;
        return size - 1 - (n >= 0 ? n : -(1 + n));
    }
    if (mode == 2) return v < 0 ? 0 : (v > size - 1 ? size - 1 : v);       // CLAMP_TO_EDGE
    if (mode == 3) return v < -1 ? -1 : (v > size ? size : v);             // CLAMP_TO_BORDER
    const int32_t m = v >= 0 ? v : -(1 + v);                               // MIRROR_CLAMP_TO_EDGE
    return m < 0 ? 0 : (m > size - 1 ? size - 1 : m);
}

CPV_DEV void border_colour(int32_t b, float out[4])
{
    const float rgb = (b >= 4) ? 1.0f : 0.0f;
    const float a = (b >= 2) ? 1.0f : 0.0f;
    out[0] = rgb; out[1] = rgb; out[2] = rgb; out[3] = a;
}

// GetPixel<fvec4, 2D> (ImageSampler.cpp:105-145).
CPV_DEV void texel(const TextureDesc& t, uint32_t level, int32_t x, int32_t y, float out[4])
{
    const uint32_t w = t.level_w[level], h = t.level_h[level];
    if (x < 0 || (uint32_t)x >= w || y < 0 || (uint32_t)y >= h) {
        border_colour(t.sampler.border_color, out);
        return;
    }
    const uint8_t* p = (const uint8_t*)t.address + t.level_offset[level] +
                       ((uint64_t)y * w + (uint64_t)x) * (uint64_t)t.fmt.total;
    uint32_t px[4] = {0u, 0u, 0u, 0u};
    if (t.fmt.total == 4) {
        px[0] = *(const uint32_t*)p;
    } else {
        // unrolled over the largest pixel so px[] keeps constant indices (registers)
#pragma unroll
        for (int i = 0; i < 16; i++)
            if (i < t.fmt.total) px[i >> 2] |= (uint32_t)p[i] << ((i & 3) * 8);
    }
    unpack_color(t.fmt, px, out);
}

// lerp (ImageSampler.cpp:51-55): the difference in float, the rest in double.
CPV_DEV void lerp4(const float a[4], const float b[4], double t, float out[4])
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float d = b[i] - a[i];
        out[i] = (float)((double)a[i] + (double)d * t);
    }
}

// SampleImageOfLevel, 2D, weighted average (ImageSampler.cpp:461-579).
CPV_DEV void sample_level(const TextureDesc& t, uint32_t level, float u, float v, int32_t filter, float out[4])
{
    const float fw = (float)t.level_w[level], fh = (float)t.level_h[level];
    const int32_t iw = (int32_t)t.level_w[level], ih = (int32_t)t.level_h[level];
    if (filter == 0) {
        int32_t cx = (int32_t)floorf(u * fw + 0.0f);
        int32_t cy = (int32_t)floorf(v * fh + 0.0f);
        cx = wrap_coord(cx, iw, t.sampler.address_u);
        cy = wrap_coord(cy, ih, t.sampler.address_v);
        texel(t, level, cx, cy, out);
        return;
    }
    int32_t x0 = (int32_t)floorf(u * fw - 0.5f);
    int32_t y0 = (int32_t)floorf(v * fh - 0.5f);
    const int32_t x1 = wrap_coord(x0 + 1, iw, t.sampler.address_u);
    const int32_t y1 = wrap_coord(y0 + 1, ih, t.sampler.address_v);
    x0 = wrap_coord(x0, iw, t.sampler.address_u);
    y0 = wrap_coord(y0, ih, t.sampler.address_v);
    const float qx = u * fw - 0.5f, qy = v * fh - 0.5f;
    const float tx = qx - floorf(qx), ty = qy - floorf(qy);
    float p00[4], p01[4], p10[4], p11[4], a[4], b[4];
    texel(t, level, x0, y0, p00);
    texel(t, level, x0, y1, p01);
    texel(t, level, x1, y0, p10);
    texel(t, level, x1, y1, p11);
    lerp4(p00, p10, (double)tx, a);
    lerp4(p01, p11, (double)tx, b);
    lerp4(a, b, (double)ty, out);
}

// ImageSampleExplicitLod (CPVulkan/GlslFunctions.cpp:598-654) and SampleImage
// (ImageSampler.cpp:581-673): lambda = clamp(lod + clamp(bias, +-32), minLod,
// maxLod).
CPV_DEV void sample_explicit(const TextureDesc& t, float u, float v, float lod, float out[4])
{
    float bias = t.sampler.mip_lod_bias + 0.0f;
    bias = bias < -32.0f ? -32.0f : (bias > 32.0f ? 32.0f : bias);
    float lambda = lod + bias;
    lambda = lambda < t.sampler.min_lod ? t.sampler.min_lod : (lambda > t.sampler.max_lod ? t.sampler.max_lod : lambda);
    if (lambda <= 0) {
        sample_level(t, 0, u, v, t.sampler.mag_filter, out);
    } else {
        const float maxl = (float)(t.level_count - 1);
        const float ml = lambda < 0.0f ? 0.0f : (lambda > maxl ? maxl : lambda);
        if (t.sampler.mipmap_mode == 0) {
            const uint32_t real = (uint32_t)ceilf(ml + 0.5f) - 1u;
            sample_level(t, real, u, v, t.sampler.min_filter, out);
        } else {
            const uint32_t l1 = (uint32_t)floorf(ml);
            const float delta = ml - (float)l1;
            float p1[4];
            sample_level(t, l1, u, v, t.sampler.min_filter, p1);
            if (delta == 0) {
                out[0] = p1[0]; out[1] = p1[1]; out[2] = p1[2]; out[3] = p1[3];
            } else {
                float p2[4];
                sample_level(t, l1 + 1, u, v, t.sampler.min_filter, p2);
                lerp4(p1, p2, (double)delta, out);
            }
        }
    }
    if (t.fmt.off[0] < 0) out[0] = 0.0f;
    if (t.fmt.off[1] < 0) out[1] = 0.0f;
    if (t.fmt.off[2] < 0) out[2] = 0.0f;
    if (t.fmt.off[3] < 0) out[3] = 1.0f;
}

// ImageSampleImplicitLod (CPVulkan/GlslFunctions.cpp:656-671): lambda base 0.
CPV_DEV void sample_implicit(const TextureDesc& t, float u, float v, float out[4])
{
    sample_explicit(t, u, v, 0.0f, out);
}

// ImageFetch (CPVulkan/GlslFunctions.cpp:673-734): the texel of the view's
// base level; coordinates outside it read ReturnType{} = (0, 0, 0, 0).
CPV_DEV void fetch_texel(const TextureDesc& t, int32_t x, int32_t y, float out[4])
{
    if (x < 0 || (uint32_t)x >= t.level_w[0] || y < 0 || (uint32_t)y >= t.level_h[0]) {
        out[0] = out[1] = out[2] = out[3] = 0.0f;
        return;
    }
    texel(t, 0, x, y, out);
}

// ------------------------------------------------------------------ blend

// ApplyBlendFactor (CPVulkan/CommandBuffer.Draw.cpp:956-1103).
CPV_DEV void blend_factor(const float s[4], const float d[4], const float c[4], int32_t cf, int32_t af, float o[4])
{
    float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f, v3 = 0.0f;
    switch (cf) {
    case 1: v0 = v1 = v2 = v3 = 1.0f; break;
    case 2: v0 = s[0]; v1 = s[1]; v2 = s[2]; v3 = s[3]; break;
    case 3: v0 = 1.0f - s[0]; v1 = 1.0f - s[1]; v2 = 1.0f - s[2]; v3 = 1.0f - s[3]; break;
    case 4: v0 = d[0]; v1 = d[1]; v2 = d[2]; v3 = d[3]; break;
    case 5: v0 = 1.0f - d[0]; v1 = 1.0f - d[1]; v2 = 1.0f - d[2]; v3 = 1.0f - d[3]; break;
    case 6: v0 = v1 = v2 = v3 = s[3]; break;
    case 7: v0 = v1 = v2 = v3 = 1.0f - s[3]; break;
    case 8: v0 = v1 = v2 = v3 = d[3]; break;
    case 9: v0 = v1 = v2 = v3 = 1.0f - d[3]; break;
    case 10: v0 = c[0]; v1 = c[1]; v2 = c[2]; v3 = c[3]; break;
    case 11: v0 = 1.0f - c[0]; v1 = 1.0f - c[1]; v2 = 1.0f - c[2]; v3 = 1.0f - c[3]; break;
    case 12: v0 = v1 = v2 = v3 = c[3]; break;
    case 13: v0 = v1 = v2 = v3 = 1.0f - c[3]; break;
    case 14: {
        const float b = 1.0f - d[3];
        const float m = b < s[3] ? b : s[3];
        v0 = v1 = v2 = m; v3 = 1.0f;
        break;
    }
    default: break;
    }
    if (cf != af) {
        switch (af) {
        case 0: v3 = 0.0f; break;
        case 1: case 14: v3 = 1.0f; break;
        case 2: case 6: v3 = s[3]; break;
        case 3: case 7: v3 = 1.0f - s[3]; break;
        case 4: case 8: v3 = d[3]; break;
        case 5: case 9: v3 = 1.0f - d[3]; break;
        case 10: case 12: v3 = c[3]; break;
        case 11: case 13: v3 = 1.0f - c[3]; break;
        default: break;
        }
    }
    o[0] = v0; o[1] = v1; o[2] = v2; o[3] = v3;
}

// ApplyBlend (CommandBuffer.Draw.cpp:1105-1262).
CPV_DEV void apply_blend(const FragmentState& st, const float s[4], const float d[4], float o[4])
{
    float sf[4], df[4];
    blend_factor(s, d, st.blend_constants, st.src_color, st.src_alpha, sf);
    blend_factor(s, d, st.blend_constants, st.dst_color, st.dst_alpha, df);
#pragma unroll
    for (int i = 0; i < 4; i++) {
        float v;
        switch (st.color_op) {
        case 0: v = s[i] * sf[i] + d[i] * df[i]; break;
        case 1: v = s[i] * sf[i] - d[i] * df[i]; break;
        case 2: v = d[i] * df[i] - s[i] * sf[i]; break;
        case 3: v = d[i] < s[i] ? d[i] : s[i]; break;   // glm::min
        default: v = s[i] < d[i] ? d[i] : s[i]; break;  // glm::max
        }
        o[i] = v;
    }
    if (st.color_op != st.alpha_op) {
        switch (st.alpha_op) {
        case 0: o[3] = s[3] * sf[3] + d[3] * df[3]; break;
        case 1: o[3] = s[3] * sf[3] - d[3] * df[3]; break;
        case 2: o[3] = d[3] * df[3] - s[3] * sf[3]; break;
        case 3: o[3] = d[3] < s[3] ? d[3] : s[3]; break;
        default: o[3] = s[3] < d[3] ? d[3] : s[3]; break;
        }
    }
}

// ---------------------------------------------------------- fragment ops

// CompileFCompareTest / CompileICompareTest (LLVMRuntime/PipelineCompiler.cpp:1492-1526).
// A VkCompareOp's three low bits select less, equal and greater (NEVER 0,
// LESS 1, EQUAL 2, LESS_OR_EQUAL 3, GREATER 4, NOT_EQUAL 5, GREATER_OR_EQUAL
// 6), so the ordered predicates are unions of the three ordered relations;
// ALWAYS (7) also passes unordered (NaN) operands, through bit 3 of both
// sides.  Computed in vector registers from the op's bits, so no branch and
// no per-op scalar mask depends on it.
CPV_DEV bool fcompare(float a, float b, int32_t op)
{
    const uint32_t rel = (a < b ? 1u : 0u) | (a == b ? 2u : 0u) | (a > b ? 4u : 0u) | (op == 7 ? 8u : 0u);
    return (rel & ((uint32_t)op | 8u)) != 0u;
}
CPV_DEV bool icompare(uint32_t a, uint32_t b, int32_t op)
{
    const uint32_t rel = (a < b ? 1u : 0u) | (a == b ? 2u : 0u) | (a > b ? 4u : 0u) | (op == 7 ? 8u : 0u);
    return (rel & ((uint32_t)op | 8u)) != 0u;
}

// CompileGetStencilResult (PipelineCompiler.cpp:1382-1413); the clamped
// variants saturate as signed bytes (llvm.sadd.sat / llvm.ssub.sat on i8).
CPV_DEV uint32_t stencil_result(int32_t op, uint32_t cur, uint32_t ref)
{
    switch (op) {
    case 0: return cur;
    case 1: return 0u;
    case 2: return ref;
    case 3: { int v = (int)(int8_t)cur + 1; v = v > 127 ? 127 : v; return (uint32_t)(uint8_t)(int8_t)v; }
    case 4: { int v = (int)(int8_t)cur - 1; v = v < -128 ? -128 : v; return (uint32_t)(uint8_t)(int8_t)v; }
    case 5: return (~cur) & 0xFFu;
    case 6: return (cur + 1u) & 0xFFu;
    case 7: return (cur - 1u) & 0xFFu;
    default: return cur;
    }
}

// Pixel state held by a raster lane for the duration of a tile.
struct PixelState {
    uint32_t color[4];   // packed colour pixel (attachment bytes)
    uint32_t ds0, ds1;   // packed depth/stencil pixel
};

// Depth/stencil part of the JIT'd fragment wrapper
// (PipelineCompiler.cpp:1087-1380) for a fragment the shader did not
// discard.  Returns true when the colour attachment must be written.
CPV_DEV bool depth_stencil_ops(const FragmentState& st, PixelState& px, float depth, bool front)
{
    const FormatDesc& f = st.ds;
    const bool has_depth = st.has_ds && f.depth_off >= 0;
    const bool has_stencil = st.has_ds && f.stencil_off >= 0;
    const bool not_s8 = st.has_ds && f.vk_format != 127u;   // VK_FORMAT_S8_UINT

    const float current_depth = ((st.depth_bounds_test || st.depth_test) && has_depth) ? unpack_depth(f, px.ds0) : 0.0f;
    if (st.depth_bounds_test && st.has_ds) {
        const bool less = !(current_depth >= st.min_depth_bounds);     // FCmpULT
        const bool greater = !(current_depth <= st.max_depth_bounds);  // FCmpUGT
        if (less || greater) return false;
    }
    const bool depth_ok = (st.depth_test && not_s8) ? fcompare(depth, current_depth, st.depth_compare) : true;
    const bool attempt_depth = st.depth_test && st.depth_write && not_s8;
    if (st.stencil_test && has_stencil) {
        const uint32_t cur = unpack_stencil(f, px.ds0, px.ds1);
        const int32_t* sf = front ? st.st_front : st.st_back;
        const uint32_t ref = (uint32_t)sf[6] & 0xFFu, mask = (uint32_t)sf[4] & 0xFFu;
        const bool stencil_ok = icompare(ref & mask, cur & mask, sf[3]);
        // The write uses the front state's ops and mask (the reference
        // passes front=true to CompileDepthStencilWrite in both branches).
        const int32_t* wf = st.st_front;
        const uint32_t fail = stencil_result(wf[0], cur, ref);
        const uint32_t pass = stencil_result(wf[1], cur, ref);
        const uint32_t zfail = stencil_result(wf[2], cur, ref);
        uint32_t value = stencil_ok ? (depth_ok ? pass : zfail) : fail;
        const uint32_t wmask = (uint32_t)wf[5] & 0xFFu;
        value = ((value & wmask) | (cur & ~wmask)) & 0xFFu;
        // depth and stencil written together; a failed depth rewrites the
        // stored depth value decoded and re-encoded (SetStencilPixelXXX).
        const float d = (depth_ok && attempt_depth) ? depth : (f.depth_off >= 0 ? unpack_depth(f, px.ds0) : 0.0f);
        pack_depth_stencil(f, d, value, px.ds0, px.ds1);
        return stencil_ok && depth_ok;
    }
    if (attempt_depth && depth_ok) {
        const uint32_t s = unpack_stencil(f, px.ds0, px.ds1);
        pack_depth_stencil(f, depth, s, px.ds0, px.ds1);
    }
    return depth_ok;
}

// True when every depth in [lo, hi] fails the compare op against every
// stored depth in [bmin, bmax] (NaN operands make the relations false: only
// NEVER then rejects).  Bit `op` of a word of the per-op answers: NEVER
// always, LESS lo >= bmax, EQUAL lo > bmax or hi < bmin, LESS_OR_EQUAL
// lo > bmax, GREATER hi <= bmin, GREATER_OR_EQUAL hi < bmin; NOT_EQUAL and
// ALWAYS never.  No branch depends on the op.
CPV_DEV bool interval_fails(int32_t op, float lo, float hi, float bmin, float bmax)
{
    const uint32_t ge = lo >= bmax ? 1u : 0u, gt = lo > bmax ? 1u : 0u, le = hi <= bmin ? 1u : 0u, lt = hi < bmin ? 1u : 0u;
    const uint32_t bits = 1u | (ge << 1) | ((gt | lt) << 2) | (gt << 3) | (le << 4) | (lt << 6);
    return ((bits >> (uint32_t)(op & 7)) & 1u) != 0u;
}

// Conservative early depth rejection.  The reference computes
//   w_i = e_i / area (correctly rounded), depth = (z0*w0 + z1*w1) + z2*w2,
//   dt = (maxD - minD) * depth + minD
// and then tests dt against the stored depth.  Here the same depth is
// approximated with one reciprocal (v_rcp_f32, <= 2 ulp assumed) instead of
// three divisions.  Both values lie within 12.2u * S of the real-number
// depth sum(z_i e_i / area), u = 2^-24, S = sum |z_i w_i|, so the exact
// depth lies in [da - B, da + B] with B = 32u * S + 2^-120 (the constant
// covers subnormal results; the fp32 rounding of da +- B is absorbed by the
// factor-2.6 margin).  dt is a monotone function of depth, so the interval
// maps to [tmin, tmax].  Returns true only when the depth test fails for
// every depth in the interval; any NaN or infinity returns false and the
// caller takes the exact path.
CPV_DEV bool depth_surely_fails(int32_t op, float cur, float e0, float e1, float e2, float area,
                                float z0, float z1, float z2, float vp_min, float vp_max)
{
    const float r = __builtin_amdgcn_rcpf(area);
    const float t0 = z0 * (e0 * r), t1 = z1 * (e1 * r), t2 = z2 * (e2 * r);
    const float da = (t0 + t1) + t2;
    const float S = (fabsf(t0) + fabsf(t1)) + fabsf(t2);
    const float B = S * 0x1p-19f + 0x1p-120f;
    const float lo = da - B, hi = da + B;
    const float sc = vp_max - vp_min;
    const float tl = sc * lo + vp_min, th = sc * hi + vp_min;
    if (!(__builtin_isfinite(tl) && __builtin_isfinite(th))) return false;
    const float tmin = fminf(tl, th), tmax = fmaxf(tl, th);
    return interval_fails(op, tmin, tmax, cur, cur);
}

// depth_surely_fails for the viewport depth range [0, 1], where
// (1 - 0) * depth + 0 equals depth for every finite depth.
CPV_DEV bool depth_surely_fails_unit(int32_t op, float cur, float e0, float e1, float e2, float area, float z0, float z1, float z2)
{
    const float r = __builtin_amdgcn_rcpf(area);
    const float t0 = z0 * (e0 * r), t1 = z1 * (e1 * r), t2 = z2 * (e2 * r);
    const float da = (t0 + t1) + t2;
    const float S = (fabsf(t0) + fabsf(t1)) + fabsf(t2);
    const float B = S * 0x1p-19f + 0x1p-120f;
    const float lo = da - B, hi = da + B;
    if (!(__builtin_isfinite(lo) && __builtin_isfinite(hi))) return false;
    return interval_fails(op, lo, hi, cur, cur);
}

// Bounds [lo, hi] of the depth the raster loop computes for a triangle at
// every pixel centre in the rectangle [xlo, xhi] x [ylo, yhi] (edge values
// e_i as in the raster loop, w_i = e_i / area correctly rounded,
// depth = (z0 w0 + z1 w1) + z2 w2).  Returns false when it gives none (area
// below 2^-40, a NaN or infinity anywhere, K beyond 2^100).
//
// In real arithmetic let E_i(x, y) = (x - ax_i) dy_i - (y - ay_i) dx_i,
// M_i >= the largest |x - ax_i||dy_i| + |y - ay_i||dx_i| over the rectangle
// and K = sum |z_i| M_i / area.  The five roundings of an edge value give
// |e_i - E_i| <= 3.0001u M_i (u = 2^-24), plus 3 * 2^-149 where products
// are subnormal, so the affine function D = sum z_i E_i / area is within
// 3.0001u K + 2^-106 sum |z_i| of sum z_i e_i / area at every point of the
// rectangle (area >= 2^-40).  The raster loop's depth, and the estimate da
// of depth_surely_fails at a corner, lie within 12.2u S + 2^-120 of that
// sum, with S <= K (1 + 8u).  D takes its extremes over the rectangle at
// the corners, so every pixel's depth lies in [min da - B, max da + B] with
// B = 2 (3.0001u + 12.2u (1 + 8u)) K + 2^-105 sum |z_i| + 2^-119, which
// 2^-18 K + 2^-100 sum |z_i| + 2^-118 exceeds by more than the rounding of
// K, B, lo and hi.
CPV_DEV bool block_depth_range(const float ax[3], const float ay[3], const float dy[3], const float dx[3], float area,
                               const float z[3], float xlo, float xhi, float ylo, float yhi, float& lo, float& hi)
{
    if (!(area >= 0x1p-40f)) return false;
    const float r = __builtin_amdgcn_rcpf(area);
    // corners (xlo, ylo), (xhi, ylo), (xlo, yhi), (xhi, yhi)
    float d0 = 0.0f, d1 = 0.0f, d2 = 0.0f, d3 = 0.0f, K = 0.0f, Z = 0.0f;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const float pl = (xlo - ax[i]) * dy[i], ph = (xhi - ax[i]) * dy[i];
        const float ql = (ylo - ay[i]) * dx[i], qh = (yhi - ay[i]) * dx[i];
        d0 += z[i] * ((pl - ql) * r);
        d1 += z[i] * ((ph - ql) * r);
        d2 += z[i] * ((pl - qh) * r);
        d3 += z[i] * ((ph - qh) * r);
        K += fabsf(z[i]) * (fmaxf(fabsf(pl), fabsf(ph)) + fmaxf(fabsf(ql), fabsf(qh)));
        Z += fabsf(z[i]);
    }
    K = K * r;
    if (!(K < 0x1p100f) || !__builtin_isfinite((d0 + d1) + (d2 + d3))) return false;
    const float B = (K * 0x1p-18f + Z * 0x1p-100f) + 0x1p-118f;
    lo = fminf(fminf(d0, d1), fminf(d2, d3)) - B;
    hi = fmaxf(fmaxf(d0, d1), fmaxf(d2, d3)) + B;
    return __builtin_isfinite(lo) && __builtin_isfinite(hi);
}

// Colour write of CompileWriteFragmentAttachment/Blend (PipelineCompiler.cpp:1528-1727)
// with blending per ApplyBlend.
CPV_DEV void color_write(const FragmentState& st, PixelState& px, const float colour[4])
{
    const FormatDesc& f = st.color;
    if (f.base == kBaseUInt) {
        uint32_t u[4] = {__float_as_uint(colour[0]), __float_as_uint(colour[1]), __float_as_uint(colour[2]), __float_as_uint(colour[3])};
        pack_color_u32(f, u, px.color);
        return;
    }
    if (f.base == kBaseSInt) {
        int32_t s[4] = {(int32_t)__float_as_uint(colour[0]), (int32_t)__float_as_uint(colour[1]),
                        (int32_t)__float_as_uint(colour[2]), (int32_t)__float_as_uint(colour[3])};
        pack_color_i32(f, s, px.color);
        return;
    }
    if (st.blend_enable) {
        float dst[4], out[4];
        unpack_color(f, px.color, dst);
        if (st.blend_kind == 1) {
            // apply_blend with sf = (s.a, s.a, s.a, s.a), df = (1 - s.a, ...), ADD
            const float sa = colour[3], da = 1.0f - colour[3];
#pragma unroll
            for (int i = 0; i < 4; i++) out[i] = colour[i] * sa + dst[i] * da;
        } else {
            apply_blend(st, colour, dst, out);
        }
        pack_color_f32(f, out, px.color);
    } else {
        pack_color_f32(f, colour, px.color);
    }
}

// -------------------------------------------------------- shader environment

struct AttribDesc {
    FormatDesc fmt;
    uint32_t binding;
    uint32_t offset;
    int32_t valid;
    int32_t pad;
};

struct BindingDesc {
    uint64_t address;
    uint64_t bytes;
    uint32_t stride;
    uint32_t input_rate;
};

// Resources of one draw as its shaders see them.
struct DrawEnv {
    AttribDesc attr[16];       // by shader input location
    BindingDesc bind[16];      // by binding number
    uint64_t uniform[4];       // uniform buffers of set 0 by binding
    TextureDesc tex[4];        // combined image samplers of set 0 by binding
};

// EmitCopyInput (PipelineCompiler.cpp:821-896) for a float shader input of
// `comps` components: an attribute of the matching R32..._SFLOAT format is a
// straight load; normalised formats go through GetPixel.  Out-of-range
// reads return zeros instead of touching memory past the buffer.
CPV_DEV void fetch_attribute(const DrawEnv& env, int location, int comps, uint32_t vertex_id, uint32_t instance_id, float* out)
{
    const AttribDesc& a = env.attr[location];
    const BindingDesc& b = env.bind[a.binding & 15u];
    const uint64_t index = b.input_rate == 0 ? vertex_id : instance_id;
    const uint64_t off = (uint64_t)b.stride * index + a.offset;
    const uint8_t* p = (const uint8_t*)b.address + off;
    const bool straight = a.fmt.base == kBaseSFloat && a.fmt.element == 4 && a.fmt.total == 4 * comps;
    const uint64_t need = straight ? (uint64_t)(4 * comps) : (uint64_t)a.fmt.total;
    if (!a.valid || off + need > b.bytes) {
        for (int i = 0; i < comps; i++) out[i] = 0.0f;
        return;
    }
    if (straight) {
        for (int i = 0; i < comps; i++) out[i] = ((const float*)p)[i];
        return;
    }
    uint32_t px[4] = {0u, 0u, 0u, 0u};
    for (int i = 0; i < a.fmt.total; i++) px[i >> 2] |= (uint32_t)p[i] << ((i & 3) * 8);
    float v[4];
    unpack_color(a.fmt, px, v);
    for (int i = 0; i < comps; i++) out[i] = v[i];
}

// glm mat4 * vec4 (column major): (m0*v0 + m1*v1) + (m2*v2 + m3*v3)
// (Samples/utils/glm/detail/type_mat4x4.inl:651-697, reached through
// MultiplicationMV, LLVMRuntime/SpirvFunctions.cpp:37-47).
CPV_DEV void mat4_mul_vec4(const float* m, const float v[4], float out[4])
{
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const float mul0 = m[0 * 4 + r] * v[0];
        const float mul1 = m[1 * 4 + r] * v[1];
        const float add0 = mul0 + mul1;
        const float mul2 = m[2 * 4 + r] * v[2];
        const float mul3 = m[3 * 4 + r] * v[3];
        const float add1 = mul2 + mul3;
        out[r] = add0 + add1;
    }
}

// glm::dot (Samples/utils/glm/detail/func_geometric.inl:55-83): products
// first, then vec2/vec3 summed left to right, vec4 summed pairwise.
CPV_DEV float dot_n(const float* a, const float* b, int n)
{
    if (n == 4) {
        const float x = a[0] * b[0], y = a[1] * b[1], z = a[2] * b[2], w = a[3] * b[3];
        return (x + y) + (z + w);
    }
    float s = a[0] * b[0];
    for (int i = 1; i < n; i++) s = s + a[i] * b[i];
    return s;
}

}  // namespace cpv
