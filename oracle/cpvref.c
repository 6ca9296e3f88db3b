/*
 * cpvref.c -- CPU oracle for the CPVulkan draw path (test infrastructure).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this; the product never does.  It restates, single-threaded and in the
 * reference's loop order, what the reference executes for vkCmdDraw /
 * vkCmdDrawIndexed replayed by vkQueueSubmit:
 *
 *   clears            CPVulkan/CommandBuffer.cpp:591-640,
 *                     CPVulkan/CommandBuffer.Draw.cpp:117-149, 2398-2419
 *   input assembly    CommandBuffer.Draw.cpp:567-760
 *   vertex stage      CommandBuffer.Draw.cpp:776-814,
 *                     LLVMRuntime/PipelineCompiler.cpp:513-582, 821-981
 *   triangle raster   CommandBuffer.Draw.cpp:410-418, 816-954, 1300-1313, 1510-1697
 *   fragment ops      LLVMRuntime/PipelineCompiler.cpp:1020-1799
 *   blend             CommandBuffer.Draw.cpp:956-1262 (ApplyBlend; dead code in
 *                     the reference, whose JIT aborts on blendEnable at
 *                     PipelineCompiler.cpp:1674-1684 -- it is the only blend
 *                     specification the reference has)
 *   pixel formats     LLVMRuntime/ImageCompiler.cpp:14-100, 161-230, 303-382,
 *                     541-606, 869-947, 1000-1240; CPVulkanBase/Formats.cpp:209-587
 *   sampling          CPVulkan/ImageSampler.cpp:12-55, 105-145, 363-410, 461-673;
 *                     CPVulkan/GlslFunctions.cpp:598-671
 *   shader math       LLVMRuntime/SpirvFunctions.cpp:5-47 (glm mat4*vec4,
 *                     dot), CPVulkan/GlslFunctions.cpp:122-126, 179-183
 *
 * Built with -O2 -ffp-contract=off and no fast-math on x86-64 (SSE scalar
 * float), so every float operation below rounds exactly once, as the
 * reference's host code and its un-flagged LLVM JIT do.
 */
#include "cpvref.h"

#include <stdint.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------ */
/* Format information (CPVulkanBase/Formats.cpp:209-444)                */
/* ------------------------------------------------------------------ */

enum { BT_NONE, BT_UNORM, BT_SNORM, BT_USCALED, BT_SSCALED, BT_UINT, BT_SINT, BT_SFLOAT };
enum { FT_NORMAL, FT_DEPTH_STENCIL };

typedef struct {
    int type;
    int base;
    uint32_t total;     /* bytes per pixel */
    uint32_t element;   /* bytes per channel */
    int off[4];         /* byte offset of R,G,B,A or -1 */
    int depth_off, stencil_off;
} fmt_t;

static int fmt_get(uint32_t f, fmt_t* o)
{
    static const int RGBA[4] = {0, 1, 2, 3}, BGRA[4] = {2, 1, 0, 3};
    memset(o, 0, sizeof(*o));
    o->off[0] = o->off[1] = o->off[2] = o->off[3] = -1;
    o->depth_off = o->stencil_off = -1;
    o->type = FT_NORMAL;
    int ch = 0, el = 1, base = BT_NONE;
    const int* order = RGBA;
    switch (f) {
    case CPV_FORMAT_R8_UNORM: ch = 1; base = BT_UNORM; break;
    case CPV_FORMAT_R8_SNORM: ch = 1; base = BT_SNORM; break;
    case CPV_FORMAT_R8_USCALED: ch = 1; base = BT_USCALED; break;
    case CPV_FORMAT_R8_SSCALED: ch = 1; base = BT_SSCALED; break;
    case CPV_FORMAT_R8_UINT: ch = 1; base = BT_UINT; break;
    case CPV_FORMAT_R8_SINT: ch = 1; base = BT_SINT; break;
    case CPV_FORMAT_R8G8_UNORM: ch = 2; base = BT_UNORM; break;
    case CPV_FORMAT_R8G8_SNORM: ch = 2; base = BT_SNORM; break;
    case CPV_FORMAT_R8G8_UINT: ch = 2; base = BT_UINT; break;
    case CPV_FORMAT_R8G8_SINT: ch = 2; base = BT_SINT; break;
    case CPV_FORMAT_R8G8B8_UNORM: ch = 3; base = BT_UNORM; break;
    case CPV_FORMAT_R8G8B8_SNORM: ch = 3; base = BT_SNORM; break;
    case CPV_FORMAT_R8G8B8_UINT: ch = 3; base = BT_UINT; break;
    case CPV_FORMAT_R8G8B8_SINT: ch = 3; base = BT_SINT; break;
    case CPV_FORMAT_B8G8R8_UNORM: ch = 3; base = BT_UNORM; order = BGRA; break;
    case CPV_FORMAT_B8G8R8_SNORM: ch = 3; base = BT_SNORM; order = BGRA; break;
    case CPV_FORMAT_B8G8R8_UINT: ch = 3; base = BT_UINT; order = BGRA; break;
    case CPV_FORMAT_B8G8R8_SINT: ch = 3; base = BT_SINT; order = BGRA; break;
    case CPV_FORMAT_R8G8B8A8_UNORM: ch = 4; base = BT_UNORM; break;
    case CPV_FORMAT_R8G8B8A8_SNORM: ch = 4; base = BT_SNORM; break;
    case CPV_FORMAT_R8G8B8A8_USCALED: ch = 4; base = BT_USCALED; break;
    case CPV_FORMAT_R8G8B8A8_SSCALED: ch = 4; base = BT_SSCALED; break;
    case CPV_FORMAT_R8G8B8A8_UINT: ch = 4; base = BT_UINT; break;
    case CPV_FORMAT_R8G8B8A8_SINT: ch = 4; base = BT_SINT; break;
    case CPV_FORMAT_B8G8R8A8_UNORM: ch = 4; base = BT_UNORM; order = BGRA; break;
    case CPV_FORMAT_B8G8R8A8_SNORM: ch = 4; base = BT_SNORM; order = BGRA; break;
    case CPV_FORMAT_B8G8R8A8_UINT: ch = 4; base = BT_UINT; order = BGRA; break;
    case CPV_FORMAT_B8G8R8A8_SINT: ch = 4; base = BT_SINT; order = BGRA; break;
    case CPV_FORMAT_R16_UNORM: ch = 1; el = 2; base = BT_UNORM; break;
    case CPV_FORMAT_R16_SNORM: ch = 1; el = 2; base = BT_SNORM; break;
    case CPV_FORMAT_R16_UINT: ch = 1; el = 2; base = BT_UINT; break;
    case CPV_FORMAT_R16_SINT: ch = 1; el = 2; base = BT_SINT; break;
    case CPV_FORMAT_R16G16_UNORM: ch = 2; el = 2; base = BT_UNORM; break;
    case CPV_FORMAT_R16G16_UINT: ch = 2; el = 2; base = BT_UINT; break;
    case CPV_FORMAT_R16G16_SINT: ch = 2; el = 2; base = BT_SINT; break;
    case CPV_FORMAT_R16G16B16A16_UNORM: ch = 4; el = 2; base = BT_UNORM; break;
    case CPV_FORMAT_R16G16B16A16_SNORM: ch = 4; el = 2; base = BT_SNORM; break;
    case CPV_FORMAT_R16G16B16A16_UINT: ch = 4; el = 2; base = BT_UINT; break;
    case CPV_FORMAT_R16G16B16A16_SINT: ch = 4; el = 2; base = BT_SINT; break;
    case CPV_FORMAT_R32_UINT: ch = 1; el = 4; base = BT_UINT; break;
    case CPV_FORMAT_R32_SINT: ch = 1; el = 4; base = BT_SINT; break;
    case CPV_FORMAT_R32_SFLOAT: ch = 1; el = 4; base = BT_SFLOAT; break;
    case CPV_FORMAT_R32G32_UINT: ch = 2; el = 4; base = BT_UINT; break;
    case CPV_FORMAT_R32G32_SINT: ch = 2; el = 4; base = BT_SINT; break;
    case CPV_FORMAT_R32G32_SFLOAT: ch = 2; el = 4; base = BT_SFLOAT; break;
    case CPV_FORMAT_R32G32B32_UINT: ch = 3; el = 4; base = BT_UINT; break;
    case CPV_FORMAT_R32G32B32_SINT: ch = 3; el = 4; base = BT_SINT; break;
    case CPV_FORMAT_R32G32B32_SFLOAT: ch = 3; el = 4; base = BT_SFLOAT; break;
    case CPV_FORMAT_R32G32B32A32_UINT: ch = 4; el = 4; base = BT_UINT; break;
    case CPV_FORMAT_R32G32B32A32_SINT: ch = 4; el = 4; base = BT_SINT; break;
    case CPV_FORMAT_R32G32B32A32_SFLOAT: ch = 4; el = 4; base = BT_SFLOAT; break;
    /* MakeDepthFormatInformation(format, total, element, base, depthOffset, stencilOffset) */
    case CPV_FORMAT_D16_UNORM: o->type = FT_DEPTH_STENCIL; o->total = 2; o->element = 2; o->base = BT_UNORM; o->depth_off = 0; return 1;
    case CPV_FORMAT_X8_D24_UNORM_PACK32: o->type = FT_DEPTH_STENCIL; o->total = 4; o->element = 4; o->base = BT_UNORM; o->depth_off = 0; return 1;
    case CPV_FORMAT_D32_SFLOAT: o->type = FT_DEPTH_STENCIL; o->total = 4; o->element = 4; o->base = BT_SFLOAT; o->depth_off = 0; return 1;
    case CPV_FORMAT_S8_UINT: o->type = FT_DEPTH_STENCIL; o->total = 1; o->element = 1; o->base = BT_UINT; o->stencil_off = 0; return 1;
    case CPV_FORMAT_D16_UNORM_S8_UINT: o->type = FT_DEPTH_STENCIL; o->total = 3; o->element = 2; o->base = BT_UNORM; o->depth_off = 0; o->stencil_off = 2; return 1;
    case CPV_FORMAT_D24_UNORM_S8_UINT: o->type = FT_DEPTH_STENCIL; o->total = 4; o->element = 3; o->base = BT_UNORM; o->depth_off = 0; o->stencil_off = 3; return 1;
    case CPV_FORMAT_D32_SFLOAT_S8_UINT: o->type = FT_DEPTH_STENCIL; o->total = 8; o->element = 4; o->base = BT_SFLOAT; o->depth_off = 0; o->stencil_off = 4; return 1;
    default: return 0;
    }
    o->base = base;
    o->element = (uint32_t)el;
    o->total = (uint32_t)(el * ch);
    for (int i = 0; i < ch; i++) o->off[i] = order[i] * el;
    return 1;
}

uint32_t cpvref_format_size(uint32_t format)
{
    fmt_t f;
    return fmt_get(format, &f) ? f.total : 0;
}

/* ------------------------------------------------------------------ */
/* Conversions (LLVMRuntime/ImageCompiler.cpp:14-100)                   */
/* ------------------------------------------------------------------ */

/* llvm.maxnum / llvm.minnum: a NaN operand yields the other operand. */
static float maxnum(float a, float b) { return fmaxf(a, b); }
static float minnum(float a, float b) { return fminf(a, b); }

/* EmitConvertFloatUInt: fmul by (float)max, llvm.round (half away from
 * zero), fptoui. */
static uint32_t f2unorm(float v, uint32_t maxv) { return (uint32_t)roundf(v * (float)maxv); }
/* EmitConvertFloatInt: fmul, llvm.round, fptosi. */
static int32_t f2snorm(float v, int32_t maxv) { return (int32_t)roundf(v * (float)maxv); }

static uint32_t load_uint(const uint8_t* p, uint32_t el)
{
    if (el == 1) return p[0];
    if (el == 2) { uint16_t v; memcpy(&v, p, 2); return v; }
    uint32_t v; memcpy(&v, p, 4); return v;
}
static int32_t load_sint(const uint8_t* p, uint32_t el)
{
    if (el == 1) return (int8_t)p[0];
    if (el == 2) { int16_t v; memcpy(&v, p, 2); return v; }
    int32_t v; memcpy(&v, p, 4); return v;
}
static void store_uint(uint8_t* p, uint32_t el, uint32_t v)
{
    if (el == 1) p[0] = (uint8_t)v;
    else if (el == 2) { uint16_t t = (uint16_t)v; memcpy(p, &t, 2); }
    else memcpy(p, &v, 4);
}

/* GetPixelF32 of a normal format: EmitGetNormalPixel / EmitGetNormalChannel<float>
 * (ImageCompiler.cpp:161-230, 303-347).  Missing channels are (0,0,0,1).
 * Unsigned integer channels of any kind are converted with an fdiv by the
 * type maximum (EmitConvertUIntFloat), signed ones likewise (EmitConvertIntFloat). */
int cpvref_unpack_f32(uint32_t format, const uint8_t* src, float rgba[4])
{
    fmt_t f;
    if (!fmt_get(format, &f) || f.type != FT_NORMAL) return CPV_ERROR_UNSUPPORTED;
    rgba[0] = 0.0f; rgba[1] = 0.0f; rgba[2] = 0.0f; rgba[3] = 1.0f;
    for (int i = 0; i < 4; i++) {
        if (f.off[i] < 0) continue;
        const uint8_t* p = src + f.off[i];
        switch (f.base) {
        case BT_UNORM: case BT_USCALED: case BT_UINT: {
            uint32_t maxv = f.element == 1 ? 255u : f.element == 2 ? 65535u : 4294967295u;
            rgba[i] = (float)load_uint(p, f.element) / (float)maxv;
            break;
        }
        case BT_SNORM: case BT_SSCALED: case BT_SINT: {
            int32_t maxv = f.element == 1 ? 127 : f.element == 2 ? 32767 : 2147483647;
            rgba[i] = (float)load_sint(p, f.element) / (float)maxv;
            break;
        }
        case BT_SFLOAT: memcpy(&rgba[i], p, 4); break;
        default: return CPV_ERROR_UNSUPPORTED;
        }
    }
    return CPV_OK;
}

/* SetPixelF32 of a normal format: CompileSetNormal/CompileSetNormalChannel
 * (ImageCompiler.cpp:1000-1037, 1077-1114, 1179-1197). */
int cpvref_pack_f32(uint32_t format, const float rgba[4], uint8_t* dst)
{
    fmt_t f;
    if (!fmt_get(format, &f) || f.type != FT_NORMAL) return CPV_ERROR_UNSUPPORTED;
    for (int i = 0; i < 4; i++) {
        if (f.off[i] < 0) continue;
        uint8_t* p = dst + f.off[i];
        float v = rgba[i];
        switch (f.base) {
        case BT_UNORM: {
            v = minnum(maxnum(v, 0.0f), 1.0f);
            store_uint(p, f.element, f2unorm(v, f.element == 1 ? 255u : 65535u));
            break;
        }
        case BT_SNORM: {
            v = minnum(maxnum(v, -1.0f), 1.0f);
            store_uint(p, f.element, (uint32_t)f2snorm(v, f.element == 1 ? 127 : 32767));
            break;
        }
        case BT_SFLOAT: memcpy(p, &v, 4); break;
        default: return CPV_ERROR_UNSUPPORTED;   /* USCALED/SSCALED: TODO_ERROR in the reference */
        }
    }
    return CPV_OK;
}

/* SetPixelU32 (ImageCompiler.cpp:1121-1143): 8/16-bit channels saturate
 * with an unsigned compare-select, 32-bit store as is. */
int cpvref_pack_u32(uint32_t format, const uint32_t rgba[4], uint8_t* dst)
{
    fmt_t f;
    if (!fmt_get(format, &f) || f.type != FT_NORMAL || f.base != BT_UINT) return CPV_ERROR_UNSUPPORTED;
    for (int i = 0; i < 4; i++) {
        if (f.off[i] < 0) continue;
        uint32_t v = rgba[i];
        if (f.element == 1) v = v < 255u ? v : 255u;
        else if (f.element == 2) v = v < 65535u ? v : 65535u;
        store_uint(dst + f.off[i], f.element, v);
    }
    return CPV_OK;
}

/* SetPixelI32 (ImageCompiler.cpp:1146-1175): signed clamp then truncate. */
int cpvref_pack_i32(uint32_t format, const int32_t rgba[4], uint8_t* dst)
{
    fmt_t f;
    if (!fmt_get(format, &f) || f.type != FT_NORMAL || f.base != BT_SINT) return CPV_ERROR_UNSUPPORTED;
    for (int i = 0; i < 4; i++) {
        if (f.off[i] < 0) continue;
        int32_t v = rgba[i];
        if (f.element == 1) { v = v > -128 ? v : -128; v = v < 127 ? v : 127; }
        else if (f.element == 2) { v = v > -32768 ? v : -32768; v = v < 32767 ? v : 32767; }
        store_uint(dst + f.off[i], f.element, (uint32_t)v);
    }
    return CPV_OK;
}

/* SetDSPixel (ImageCompiler.cpp:869-947). */
int cpvref_pack_depth_stencil(uint32_t format, float depth, uint32_t stencil, uint8_t* dst)
{
    fmt_t f;
    if (!fmt_get(format, &f) || f.type != FT_DEPTH_STENCIL) return CPV_ERROR_UNSUPPORTED;
    if (f.depth_off >= 0) {
        float v = depth;
        switch (format) {
        case CPV_FORMAT_D16_UNORM: case CPV_FORMAT_D16_UNORM_S8_UINT: {
            v = minnum(maxnum(v, 0.0f), 1.0f);
            uint16_t u = (uint16_t)f2unorm(v, 65535u);
            memcpy(dst, &u, 2);
            break;
        }
        case CPV_FORMAT_D24_UNORM_S8_UINT: case CPV_FORMAT_X8_D24_UNORM_PACK32: {
            v = minnum(maxnum(v, 0.0f), 1.0f);
            v = v * (float)0x00FFFFFF;
            uint32_t u = (uint32_t)roundf(v);
            if (format == CPV_FORMAT_D24_UNORM_S8_UINT) u |= (uint32_t)(uint8_t)stencil << 24;
            memcpy(dst, &u, 4);
            break;
        }
        default: memcpy(dst, &v, 4); break;      /* D32: raw store, no clamp */
        }
    }
    if (f.stencil_off >= 0 && format != CPV_FORMAT_D24_UNORM_S8_UINT) dst[f.stencil_off] = (uint8_t)stencil;
    return CPV_OK;
}

/* GetPixelDepth (ImageCompiler.cpp:541-606). */
float cpvref_unpack_depth(uint32_t format, const uint8_t* src)
{
    switch (format) {
    case CPV_FORMAT_D16_UNORM: case CPV_FORMAT_D16_UNORM_S8_UINT: {
        uint16_t u; memcpy(&u, src, 2);
        return (float)u / 65535.0f;
    }
    case CPV_FORMAT_D24_UNORM_S8_UINT: case CPV_FORMAT_X8_D24_UNORM_PACK32: {
        uint32_t u; memcpy(&u, src, 4);
        return (float)(u & 0x00FFFFFFu) / (float)0x00FFFFFF;
    }
    case CPV_FORMAT_D32_SFLOAT: case CPV_FORMAT_D32_SFLOAT_S8_UINT: {
        float v; memcpy(&v, src, 4);
        return v;
    }
    default: return 0.0f;
    }
}

/* GetPixelStencil (ImageCompiler.cpp:608-645). */
uint32_t cpvref_unpack_stencil(uint32_t format, const uint8_t* src)
{
    fmt_t f;
    if (!fmt_get(format, &f) || f.stencil_off < 0) return 0;
    return src[f.stencil_off];
}

/* GetNormalImageSize (Formats.cpp:455-482): contiguous mip chain, stride
 * = pixel size * width, halving with a floor of 1. */
uint64_t cpvref_image_layout(uint32_t format, uint32_t width, uint32_t height, uint32_t levels,
                             uint64_t* level_offsets)
{
    uint64_t px = cpvref_format_size(format), layer = 0;
    for (uint32_t i = 0; i < levels; i++) {
        if (level_offsets) level_offsets[i] = layer;
        layer += px * width * height;
        width = width / 2 > 1 ? width / 2 : 1;
        height = height / 2 > 1 ? height / 2 : 1;
    }
    return layer;
}

/* ------------------------------------------------------------------ */
/* Sampling (CPVulkan/ImageSampler.cpp)                                */
/* ------------------------------------------------------------------ */

/* wrap (ImageSampler.cpp:12-38) */
static int32_t wrap_coord(int32_t v, int32_t size, uint32_t mode)
{
    switch (mode) {
    case CPV_ADDRESS_REPEAT: return (v % size + size) % size;
    case CPV_ADDRESS_MIRRORED_REPEAT: {
        int32_t two = 2 * size;
        int32_t n = (v % two + two) % two - size;
        return size - 1 - (n >= 0 ? n : -(1 + n));
    }
    case CPV_ADDRESS_CLAMP_TO_EDGE: return v < 0 ? 0 : v > size - 1 ? size - 1 : v;
    case CPV_ADDRESS_CLAMP_TO_BORDER: return v < -1 ? -1 : v > size ? size : v;
    case CPV_ADDRESS_MIRROR_CLAMP_TO_EDGE: {
        int32_t m = v >= 0 ? v : -(1 + v);
        return m < 0 ? 0 : m > size - 1 ? size - 1 : m;
    }
    default: return 0;
    }
}

/* lerp (ImageSampler.cpp:51-55): dvec4(a) + dvec4(b - a) * delta, the
 * difference taken in float, the rest in double, rounded back to float. */
static void lerp4(const float a[4], const float b[4], double t, float out[4])
{
    for (int i = 0; i < 4; i++) {
        float d = b[i] - a[i];
        out[i] = (float)((double)a[i] + (double)d * t);
    }
}

typedef struct { const uint8_t* data; uint32_t w, h; } level_t;

/* GetPixel<fvec4, 2D> (ImageSampler.cpp:105-145). */
static void texel(uint32_t format, const level_t* lv, int32_t x, int32_t y, const float border[4], float out[4])
{
    if (x < 0 || (uint32_t)x >= lv->w || y < 0 || (uint32_t)y >= lv->h) {
        memcpy(out, border, 16);
        return;
    }
    uint32_t px = cpvref_format_size(format);
    cpvref_unpack_f32(format, lv->data + (uint64_t)y * lv->w * px + (uint64_t)x * px, out);
}

static const float k_border[6][4] = {
    {0, 0, 0, 0}, {0, 0, 0, 0}, {0, 0, 0, 1}, {0, 0, 0, 1}, {1, 1, 1, 1}, {1, 1, 1, 1}};

/* SampleImageOfLevel, 2D, weighted average (ImageSampler.cpp:461-579). */
static void sample_level(uint32_t format, const level_t* lv, float u, float v, uint32_t filter,
                         const cpv_sampler_desc* s, float out[4])
{
    const float* border = k_border[s->border_color <= 5 ? s->border_color : 0];
    const uint32_t modes[2] = {s->address_u, s->address_v};
    const float coords[2] = {u, v};
    const uint32_t range[2] = {lv->w, lv->h};
    if (filter == CPV_FILTER_NEAREST) {
        int32_t c[2];
        for (int i = 0; i < 2; i++) {
            c[i] = (int32_t)floorf(coords[i] * (float)range[i] + 0.0f);
            c[i] = wrap_coord(c[i], (int32_t)range[i], modes[i]);
        }
        texel(format, lv, c[0], c[1], border, out);
        return;
    }
    int32_t c0[2], c1[2];
    float t[2];
    for (int i = 0; i < 2; i++) {
        c0[i] = (int32_t)floorf(coords[i] * (float)range[i] - 0.5f);
        c1[i] = wrap_coord(c0[i] + 1, (int32_t)range[i], modes[i]);
        c0[i] = wrap_coord(c0[i], (int32_t)range[i], modes[i]);
    }
    for (int i = 0; i < 2; i++) {
        float q = coords[i] * (float)range[i] - 0.5f;
        t[i] = q - floorf(q);
    }
    /* GetPixelLinear, length 2 (ImageSampler.cpp:363-410) */
    float p00[4], p01[4], p10[4], p11[4], a[4], b[4];
    texel(format, lv, c0[0], c0[1], border, p00);
    texel(format, lv, c0[0], c1[1], border, p01);
    texel(format, lv, c1[0], c0[1], border, p10);
    texel(format, lv, c1[0], c1[1], border, p11);
    lerp4(p00, p10, (double)t[0], a);
    lerp4(p01, p11, (double)t[0], b);
    lerp4(a, b, (double)t[1], out);
}

static void get_level(const cpv_texture_desc* t, uint32_t view_level, level_t* lv)
{
    uint64_t offs[CPV_MAX_MIP_LEVELS];
    uint32_t levels = t->levels > CPV_MAX_MIP_LEVELS ? CPV_MAX_MIP_LEVELS : t->levels;
    cpvref_image_layout(t->format, t->width, t->height, levels, offs);
    uint32_t level = t->base_level + view_level;
    uint32_t w = t->width, h = t->height;
    for (uint32_t i = 0; i < level; i++) { w = w / 2 > 1 ? w / 2 : 1; h = h / 2 > 1 ? h / 2 : 1; }
    lv->data = (const uint8_t*)(uintptr_t)t->address + offs[level];
    lv->w = w;
    lv->h = h;
}

/* ImageSampleExplicitLod (GlslFunctions.cpp:598-654) -> SampleImage
 * (ImageSampler.cpp:581-673): lambda = clamp(lod + clamp(bias, +-32), minLod, maxLod). */
int cpvref_sample_explicit(const cpv_texture_desc* t, float u, float v, float lod, float out[4])
{
    const cpv_sampler_desc* s = &t->sampler;
    fmt_t f;
    if (!fmt_get(t->format, &f) || f.type != FT_NORMAL) return CPV_ERROR_UNSUPPORTED;
    float bias = s->mip_lod_bias + 0.0f;
    bias = bias < -32.0f ? -32.0f : bias > 32.0f ? 32.0f : bias;   /* MAX_SAMPLER_LOD_BIAS */
    float lambda = lod + bias;
    lambda = lambda < s->min_lod ? s->min_lod : lambda > s->max_lod ? s->max_lod : lambda;
    uint32_t levels = t->level_count;
    level_t lv;
    if (lambda <= 0) {
        get_level(t, 0, &lv);
        sample_level(t->format, &lv, u, v, s->mag_filter, s, out);
    } else {
        float maxl = (float)(levels - 1);
        float ml = lambda < 0.0f ? 0.0f : lambda > maxl ? maxl : lambda;
        if (s->mipmap_mode == CPV_MIPMAP_NEAREST) {
            uint32_t real = (uint32_t)ceilf(ml + 0.5f) - 1;
            get_level(t, real, &lv);
            sample_level(t->format, &lv, u, v, s->min_filter, s, out);
        } else {
            uint32_t l1 = (uint32_t)floorf(ml);
            float delta = ml - (float)l1;
            float p1[4];
            get_level(t, l1, &lv);
            sample_level(t->format, &lv, u, v, s->min_filter, s, p1);
            if (delta == 0) {
                memcpy(out, p1, 16);
            } else {
                float p2[4];
                get_level(t, l1 + 1, &lv);
                sample_level(t->format, &lv, u, v, s->min_filter, s, p2);
                lerp4(p1, p2, (double)delta, out);
            }
        }
    }
    if (f.off[0] < 0) out[0] = 0;
    if (f.off[1] < 0) out[1] = 0;
    if (f.off[2] < 0) out[2] = 0;
    if (f.off[3] < 0) out[3] = 1;
    return CPV_OK;
}

/* ImageSampleImplicitLod: lambda base 0 (GlslFunctions.cpp:656-671). */
int cpvref_sample_implicit(const cpv_texture_desc* t, float u, float v, float out[4])
{
    return cpvref_sample_explicit(t, u, v, 0.0f, out);
}

/* ImageFetch (GlslFunctions.cpp:673-734): the texel of the view's base
 * level; outside it, ReturnType{} = (0, 0, 0, 0). */
int cpvref_fetch(const cpv_texture_desc* t, int32_t x, int32_t y, float out[4])
{
    fmt_t f;
    if (!fmt_get(t->format, &f) || f.type != FT_NORMAL) return CPV_ERROR_UNSUPPORTED;
    level_t lv;
    get_level(t, 0, &lv);
    static const float zero[4] = {0, 0, 0, 0};
    texel(t->format, &lv, x, y, zero, out);
    return CPV_OK;
}

/* ------------------------------------------------------------------ */
/* Blend (CommandBuffer.Draw.cpp:956-1262)                             */
/* ------------------------------------------------------------------ */

static void blend_factor(const float s[4], const float d[4], const float c[4], uint32_t colour_f,
                         uint32_t alpha_f, float out[4])
{
    int i;
    for (i = 0; i < 4; i++) out[i] = 0.0f;
    switch (colour_f) {
    case CPV_BLEND_ZERO: break;
    case CPV_BLEND_ONE: for (i = 0; i < 4; i++) out[i] = 1.0f; break;
    case CPV_BLEND_SRC_COLOR: for (i = 0; i < 4; i++) out[i] = s[i]; break;
    case CPV_BLEND_ONE_MINUS_SRC_COLOR: for (i = 0; i < 4; i++) out[i] = 1.0f - s[i]; break;
    case CPV_BLEND_DST_COLOR: for (i = 0; i < 4; i++) out[i] = d[i]; break;
    case CPV_BLEND_ONE_MINUS_DST_COLOR: for (i = 0; i < 4; i++) out[i] = 1.0f - d[i]; break;
    case CPV_BLEND_SRC_ALPHA: for (i = 0; i < 4; i++) out[i] = s[3]; break;
    case CPV_BLEND_ONE_MINUS_SRC_ALPHA: for (i = 0; i < 4; i++) out[i] = 1.0f - s[3]; break;
    case CPV_BLEND_DST_ALPHA: for (i = 0; i < 4; i++) out[i] = d[3]; break;
    case CPV_BLEND_ONE_MINUS_DST_ALPHA: for (i = 0; i < 4; i++) out[i] = 1.0f - d[3]; break;
    case CPV_BLEND_CONSTANT_COLOR: for (i = 0; i < 4; i++) out[i] = c[i]; break;
    case CPV_BLEND_ONE_MINUS_CONSTANT_COLOR: for (i = 0; i < 4; i++) out[i] = 1.0f - c[i]; break;
    case CPV_BLEND_CONSTANT_ALPHA: for (i = 0; i < 4; i++) out[i] = c[3]; break;
    case CPV_BLEND_ONE_MINUS_CONSTANT_ALPHA: for (i = 0; i < 4; i++) out[i] = 1.0f - c[3]; break;
    case CPV_BLEND_SRC_ALPHA_SATURATE: {
        /* std::min(source.a, 1 - destination.a) */
        float a = s[3], b = 1.0f - d[3];
        float m = b < a ? b : a;
        out[0] = m; out[1] = m; out[2] = m; out[3] = 1.0f;
        break;
    }
    default: break;
    }
    if (colour_f != alpha_f) {
        switch (alpha_f) {
        case CPV_BLEND_ZERO: out[3] = 0.0f; break;
        case CPV_BLEND_ONE: case CPV_BLEND_SRC_ALPHA_SATURATE: out[3] = 1.0f; break;
        case CPV_BLEND_SRC_COLOR: case CPV_BLEND_SRC_ALPHA: out[3] = s[3]; break;
        case CPV_BLEND_ONE_MINUS_SRC_COLOR: case CPV_BLEND_ONE_MINUS_SRC_ALPHA: out[3] = 1.0f - s[3]; break;
        case CPV_BLEND_DST_COLOR: case CPV_BLEND_DST_ALPHA: out[3] = d[3]; break;
        case CPV_BLEND_ONE_MINUS_DST_COLOR: case CPV_BLEND_ONE_MINUS_DST_ALPHA: out[3] = 1.0f - d[3]; break;
        case CPV_BLEND_CONSTANT_COLOR: case CPV_BLEND_CONSTANT_ALPHA: out[3] = c[3]; break;
        case CPV_BLEND_ONE_MINUS_CONSTANT_COLOR: case CPV_BLEND_ONE_MINUS_CONSTANT_ALPHA: out[3] = 1.0f - c[3]; break;
        default: break;
        }
    }
}

/* glm::min / glm::max per component: min(x,y) = y < x ? y : x,
 * max(x,y) = x < y ? y : x. */
static float gmin(float x, float y) { return y < x ? y : x; }
static float gmax(float x, float y) { return x < y ? y : x; }

void cpvref_apply_blend(const float s[4], const float d[4], const float c[4],
                        const cpv_blend_attachment* st, float out[4])
{
    float sf[4], df[4];
    blend_factor(s, d, c, st->src_color_factor, st->src_alpha_factor, sf);
    blend_factor(s, d, c, st->dst_color_factor, st->dst_alpha_factor, df);
    for (int i = 0; i < 4; i++) {
        switch (st->color_op) {
        case CPV_BLEND_OP_ADD: out[i] = s[i] * sf[i] + d[i] * df[i]; break;
        case CPV_BLEND_OP_SUBTRACT: out[i] = s[i] * sf[i] - d[i] * df[i]; break;
        case CPV_BLEND_OP_REVERSE_SUBTRACT: out[i] = d[i] * df[i] - s[i] * sf[i]; break;
        case CPV_BLEND_OP_MIN: out[i] = gmin(s[i], d[i]); break;
        case CPV_BLEND_OP_MAX: out[i] = gmax(s[i], d[i]); break;
        default: out[i] = 0.0f; break;
        }
    }
    if (st->color_op != st->alpha_op) {
        switch (st->alpha_op) {
        case CPV_BLEND_OP_ADD: out[3] = s[3] * sf[3] + d[3] * df[3]; break;
        case CPV_BLEND_OP_SUBTRACT: out[3] = s[3] * sf[3] - d[3] * df[3]; break;
        case CPV_BLEND_OP_REVERSE_SUBTRACT: out[3] = d[3] * df[3] - s[3] * sf[3]; break;
        /* std::min / std::max on the alpha channel */
        case CPV_BLEND_OP_MIN: out[3] = d[3] < s[3] ? d[3] : s[3]; break;
        case CPV_BLEND_OP_MAX: out[3] = s[3] < d[3] ? d[3] : s[3]; break;
        default: break;
        }
    }
}

/* ------------------------------------------------------------------ */
/* Shaders of the benchmark workloads                                  */
/* ------------------------------------------------------------------ */

/* glm mat4 * vec4 (Samples/utils/glm/detail/type_mat4x4.inl:651-697):
 * (m0*v0 + m1*v1) + (m2*v2 + m3*v3), column major. */
static void mat4_mul_vec4(const float m[16], const float v[4], float out[4])
{
    for (int r = 0; r < 4; r++) {
        float mul0 = m[0 * 4 + r] * v[0];
        float mul1 = m[1 * 4 + r] * v[1];
        float add0 = mul0 + mul1;
        float mul2 = m[2 * 4 + r] * v[2];
        float mul3 = m[3 * 4 + r] * v[3];
        float add1 = mul2 + mul3;
        out[r] = add0 + add1;
    }
}

/* glm::dot for vec3: (x*x' + y*y') + z*z'. */
static float dot3(const float a[3], const float b[3])
{
    float x = a[0] * b[0], y = a[1] * b[1], z = a[2] * b[2];
    return (x + y) + z;
}

/* glm::normalize for vec3 (Samples/utils/glm/detail/func_geometric.inl:256-266,
 * func_exponential.inl:55-60): sqr = (x*x + y*y) + z*z, then v * (1 / sqrt(sqr)). */
static void normalize3(const float v[3], float out[3])
{
    const float sqr = (v[0] * v[0] + v[1] * v[1]) + v[2] * v[2];
    const float inv = 1.0f / sqrtf(sqr);
    out[0] = v[0] * inv;
    out[1] = v[1] * inv;
    out[2] = v[2] * inv;
}

typedef struct {
    int nout;               /* varying floats written by the VS, in declaration order */
    int nin_attr;           /* number of vertex attributes read */
    int loc[3];             /* attribute locations */
    int comps[3];           /* float components per attribute */
} vs_info_t;

static int vs_info(uint32_t kind, vs_info_t* o)
{
    memset(o, 0, sizeof(*o));
    switch (kind) {
    case CPV_VS_PASSTHROUGH:
    case CPV_VS_MVP_COLOR:
        o->nout = 4; o->nin_attr = 2;
        o->loc[0] = 0; o->comps[0] = 4; o->loc[1] = 1; o->comps[1] = 4;
        return 1;
    case CPV_VS_MVP_LAMBERT:
        o->nout = 5; o->nin_attr = 3;
        o->loc[0] = 0; o->comps[0] = 3; o->loc[1] = 1; o->comps[1] = 3; o->loc[2] = 2; o->comps[2] = 2;
        return 1;
    case CPV_VS_PASSTHROUGH_POINT:
        o->nout = 5; o->nin_attr = 3;
        o->loc[0] = 0; o->comps[0] = 4; o->loc[1] = 1; o->comps[1] = 4; o->loc[2] = 2; o->comps[2] = 1;
        return 1;
    default: return 0;
    }
}

static int fs_inputs(uint32_t kind)
{
    switch (kind) {
    case CPV_FS_COLOR: return 4;
    case CPV_FS_COLOR_DISCARD: return 4;
    case CPV_FS_TEX_LAMBERT: return 5;
    case CPV_FS_COLOR_FLAT: return 4;
    case CPV_FS_COLOR_NOPERSPECTIVE: return 4;
    default: return -1;
    }
}

/* InterpolationType of the fragment shader's inputs (every input of a
 * built-in shader shares one): 0 Perspective, 1 Linear (NoPerspective),
 * 2 Flat. */
static int fs_interp(uint32_t kind)
{
    if (kind == CPV_FS_COLOR_FLAT) return 2;
    if (kind == CPV_FS_COLOR_NOPERSPECTIVE) return 1;
    return 0;
}

static uint32_t float_format_for(int comps)
{
    switch (comps) {
    case 1: return CPV_FORMAT_R32_SFLOAT;
    case 2: return CPV_FORMAT_R32G32_SFLOAT;
    case 3: return CPV_FORMAT_R32G32B32_SFLOAT;
    default: return CPV_FORMAT_R32G32B32A32_SFLOAT;
    }
}

/* EmitCopyInput (PipelineCompiler.cpp:821-896) for float shader inputs:
 * identical formats are a straight load, normalised/scaled formats go
 * through GetPixel and the first components are kept. */
static int fetch_attribute(const cpv_pipeline_desc* p, const cpv_draw_desc* d, int location, int comps,
                           uint32_t vertex_id, uint32_t instance_id, float* out)
{
    const cpv_vertex_attribute_desc* a = NULL;
    const cpv_vertex_binding_desc* b = NULL;
    for (uint32_t i = 0; i < p->attribute_count; i++)
        if ((int)p->attributes[i].location == location) { a = &p->attributes[i]; break; }
    if (!a) return CPV_ERROR_INVALID;
    for (uint32_t i = 0; i < p->binding_count; i++)
        if (p->bindings[i].binding == a->binding) { b = &p->bindings[i]; break; }
    if (!b || b->binding >= CPV_MAX_VERTEX_BINDINGS) return CPV_ERROR_INVALID;
    uint64_t index = b->input_rate == CPV_INPUT_RATE_VERTEX ? vertex_id : instance_id;
    uint64_t off = (uint64_t)b->stride * index + a->offset;
    const uint8_t* base = (const uint8_t*)(uintptr_t)d->vertex_address[b->binding];
    if (off + 4u * (uint32_t)comps > d->vertex_bytes[b->binding] && a->format == float_format_for(comps))
        return CPV_ERROR_INVALID;
    if (a->format == float_format_for(comps)) {
        memcpy(out, base + off, 4u * (uint32_t)comps);
        return CPV_OK;
    }
    fmt_t f;
    if (!fmt_get(a->format, &f) || f.type != FT_NORMAL) return CPV_ERROR_UNSUPPORTED;
    if (f.base != BT_UNORM && f.base != BT_SNORM && f.base != BT_USCALED && f.base != BT_SSCALED)
        return CPV_ERROR_UNSUPPORTED;
    if (off + f.total > d->vertex_bytes[b->binding]) return CPV_ERROR_INVALID;
    float v[4];
    cpvref_unpack_f32(a->format, base + off, v);
    memcpy(out, v, 4u * (uint32_t)comps);
    return CPV_OK;
}

static const float* uniform_ptr(const cpv_draw_desc* d, int binding)
{
    return (const float*)(uintptr_t)d->uniform_address[binding];
}

/* Vertex shader of one vertex: writes clip position + varyings. */
static int run_vs(const cpv_pipeline_desc* p, const cpv_draw_desc* d, uint32_t vertex_id,
                  uint32_t instance_id, float pos[4], float* vary)
{
    vs_info_t info;
    if (!vs_info(p->vs_kind, &info)) return CPV_ERROR_UNSUPPORTED;
    float in[3][4];
    for (int i = 0; i < info.nin_attr; i++) {
        int rc = fetch_attribute(p, d, info.loc[i], info.comps[i], vertex_id, instance_id, in[i]);
        if (rc) return rc;
    }
    switch (p->vs_kind) {
    case CPV_VS_PASSTHROUGH:
        memcpy(pos, in[0], 16);
        memcpy(vary, in[1], 16);
        break;
    case CPV_VS_PASSTHROUGH_POINT:
        /* gl_Position = pos, gl_PointSize = size; the size is output 4 */
        memcpy(pos, in[0], 16);
        memcpy(vary, in[1], 16);
        vary[4] = in[2][0];
        break;
    case CPV_VS_MVP_COLOR: {
        const float* mvp = uniform_ptr(d, 0);
        if (!mvp) return CPV_ERROR_INVALID;
        mat4_mul_vec4(mvp, in[0], pos);
        memcpy(vary, in[1], 16);
        break;
    }
    case CPV_VS_MVP_LAMBERT: {
        const float* mvp = uniform_ptr(d, 0);
        if (!mvp) return CPV_ERROR_INVALID;
        float p4[4] = {in[0][0], in[0][1], in[0][2], 1.0f};
        mat4_mul_vec4(mvp, p4, pos);
        memcpy(vary, in[1], 12);
        memcpy(vary + 3, in[2], 8);
        break;
    }
    default: return CPV_ERROR_UNSUPPORTED;
    }
    return CPV_OK;
}

/* Fragment shader: returns 1 when the fragment is discarded (OpKill). */
static int run_fs(const cpv_pipeline_desc* p, const cpv_draw_desc* d, const float* in, float out[4])
{
    switch (p->fs_kind) {
    case CPV_FS_COLOR:
    case CPV_FS_COLOR_FLAT:
    case CPV_FS_COLOR_NOPERSPECTIVE:
        memcpy(out, in, 16);
        return 0;
    case CPV_FS_COLOR_DISCARD:
        /* if (color.r < 0.25) discard: OpFOrdLessThan then OpKill */
        if (in[0] < 0.25f) return 1;
        memcpy(out, in, 16);
        return 0;
    case CPV_FS_TEX_LAMBERT: {
        float t[4];
        cpvref_sample_implicit(&d->textures[1], in[3], in[4], t);
        float n[3];
        normalize3(in, n);
        const float* L = uniform_ptr(d, 2);
        float l = dot3(n, L);
        /* FMax -> std::max(l, 0.0f) (GlslFunctions.cpp:122-126) */
        float m = l < 0.0f ? 0.0f : l;
        for (int i = 0; i < 4; i++) out[i] = t[i] * m;
        return 0;
    }
    default: return 0;
    }
}

/* ------------------------------------------------------------------ */
/* Raster (CommandBuffer.Draw.cpp)                                     */
/* ------------------------------------------------------------------ */

/* static_cast<int32_t>(float) as x86-64 cvttss2si executes it: truncation,
 * INT32_MIN for NaN and out-of-range values. */
static int32_t f2i_trunc(float v)
{
    if (v >= -2147483648.0f && v < 2147483648.0f) return (int32_t)v;
    return INT32_MIN;
}

/* EdgeFunction (CommandBuffer.Draw.cpp:410-418) */
static float edge(const float* a, const float* b, float cx, float cy)
{
    return (cx - a[0]) * (b[1] - a[1]) - (cy - a[1]) * (b[0] - a[0]);
}

typedef struct {
    const cpv_pass_desc* pass;
    const cpv_pipeline_desc* pipe;
    const cpv_draw_desc* draw;
    fmt_t cfmt, dsfmt;
    int has_color, has_ds;
    uint64_t fragments;
    /* Pixel window [wx0, wx1) x [wy0, wy1) the primitive loops are limited
     * to (the whole plane for the reference's single-threaded replay; one
     * tile per task in the tile-parallel mode, where every pixel still sees
     * the same primitives in the same order with the same arithmetic). */
    int32_t wx0, wy0, wx1, wy1;
    int threads;   /* > 1: tile-parallel replay (OpenMP) */
} ctx_t;

static uint8_t* color_px(ctx_t* c, int x, int y)
{
    return (uint8_t*)(uintptr_t)c->pass->color[0].address + ((uint64_t)y * c->pass->width + (uint64_t)x) * c->cfmt.total;
}
static uint8_t* ds_px(ctx_t* c, int x, int y)
{
    return (uint8_t*)(uintptr_t)c->pass->depth_stencil.address + ((uint64_t)y * c->pass->width + (uint64_t)x) * c->dsfmt.total;
}

static int fcompare(float a, float b, uint32_t op)
{
    switch (op) {
    case CPV_COMPARE_NEVER: return 0;
    case CPV_COMPARE_LESS: return a < b;
    case CPV_COMPARE_EQUAL: return a == b;
    case CPV_COMPARE_LESS_OR_EQUAL: return a <= b;
    case CPV_COMPARE_GREATER: return a > b;
    case CPV_COMPARE_NOT_EQUAL: return a < b || a > b;     /* FCmpONE: ordered and unequal */
    case CPV_COMPARE_GREATER_OR_EQUAL: return a >= b;
    default: return 1;
    }
}

static int icompare(uint32_t a, uint32_t b, uint32_t op)
{
    switch (op) {
    case CPV_COMPARE_NEVER: return 0;
    case CPV_COMPARE_LESS: return a < b;
    case CPV_COMPARE_EQUAL: return a == b;
    case CPV_COMPARE_LESS_OR_EQUAL: return a <= b;
    case CPV_COMPARE_GREATER: return a > b;
    case CPV_COMPARE_NOT_EQUAL: return a != b;
    case CPV_COMPARE_GREATER_OR_EQUAL: return a >= b;
    default: return 1;
    }
}

/* CompileGetStencilResult (PipelineCompiler.cpp:1382-1413); INC/DEC_CLAMP
 * are llvm.sadd.sat / llvm.ssub.sat on i8. */
static uint8_t stencil_result(uint32_t op, uint8_t cur, uint8_t ref)
{
    switch (op) {
    case CPV_STENCIL_KEEP: return cur;
    case CPV_STENCIL_ZERO: return 0;
    case CPV_STENCIL_REPLACE: return ref;
    case CPV_STENCIL_INC_CLAMP: { int v = (int8_t)cur + 1; if (v > 127) v = 127; return (uint8_t)(int8_t)v; }
    case CPV_STENCIL_DEC_CLAMP: { int v = (int8_t)cur - 1; if (v < -128) v = -128; return (uint8_t)(int8_t)v; }
    case CPV_STENCIL_INVERT: return (uint8_t)~cur;
    case CPV_STENCIL_INC_WRAP: return (uint8_t)(cur + 1);
    case CPV_STENCIL_DEC_WRAP: return (uint8_t)(cur - 1);
    default: return cur;
    }
}

/* The JIT'd fragment wrapper (PipelineCompiler.cpp:1020-1085) after the
 * shader returned (not discarded). */
static void fragment_ops(ctx_t* c, int x, int y, float depth, int front, const float colour[4])
{
    const cpv_pipeline_desc* p = c->pipe;
    const int ds_attached = c->has_ds;
    const int has_depth = ds_attached && c->dsfmt.depth_off >= 0;
    const int has_stencil = ds_attached && c->dsfmt.stencil_off >= 0;
    const int not_s8 = ds_attached && p->depth_stencil_format != CPV_FORMAT_S8_UINT;

    /* CompileGetCurrentData (:1087-1126) */
    float current_depth = 0.0f;
    uint8_t current_stencil = 0;
    if ((p->depth_bounds_test || p->depth_test) && has_depth)
        current_depth = cpvref_unpack_depth(p->depth_stencil_format, ds_px(c, x, y));
    if (p->stencil_test && has_stencil)
        current_stencil = (uint8_t)cpvref_unpack_stencil(p->depth_stencil_format, ds_px(c, x, y));

    /* CompileDepthBoundsTest (:1128-1166): FCmpULT / FCmpUGT */
    if (p->depth_bounds_test && ds_attached) {
        int less = !(current_depth >= p->min_depth_bounds);
        int greater = !(current_depth <= p->max_depth_bounds);
        if (less || greater) return;
    }

    int stencil_ok = 1, depth_ok = 1;
    /* CompileDepthTest (:1225-1244) */
#define DEPTH_TEST() do { \
        if (p->depth_test && not_s8) depth_ok = fcompare(depth, current_depth, p->depth_compare); \
        else depth_ok = 1; } while (0)

    if (p->stencil_test && has_stencil) {
        /* CompileStencilTest (:1168-1192): the compare uses the facing
         * state; the write below always uses the front state's ops and
         * write mask, with the facing state's reference for REPLACE. */
        const cpv_stencil_state* st = front ? &p->front : &p->back;
        uint8_t ref = (uint8_t)st->reference, mask = (uint8_t)st->compare_mask;
        stencil_ok = icompare((uint32_t)(ref & mask), (uint32_t)(current_stencil & mask), st->compare_op);
        DEPTH_TEST();
        /* CompileStencilWrite (:1294-1380) */
        const cpv_stencil_state* ws = &p->front;
        uint8_t fail = stencil_result(ws->fail_op, current_stencil, ref);
        uint8_t zfail = stencil_result(ws->depth_fail_op, current_stencil, ref);
        uint8_t pass = stencil_result(ws->pass_op, current_stencil, ref);
        uint8_t value = stencil_ok ? (depth_ok ? pass : zfail) : fail;
        uint8_t wmask = (uint8_t)ws->write_mask;
        value = (uint8_t)((value & wmask) | (current_stencil & (uint8_t)~wmask));
        int attempt_depth = p->depth_test && p->depth_write && not_s8;
        if (depth_ok && attempt_depth) {
            cpvref_pack_depth_stencil(p->depth_stencil_format, depth, value, ds_px(c, x, y));
        } else {
            /* SetStencilPixelXXX re-reads the depth (GlslFunctions.cpp:898-912) */
            float d = c->dsfmt.depth_off < 0 ? 0.0f : cpvref_unpack_depth(p->depth_stencil_format, ds_px(c, x, y));
            cpvref_pack_depth_stencil(p->depth_stencil_format, d, value, ds_px(c, x, y));
        }
    } else {
        stencil_ok = 1;
        DEPTH_TEST();
        /* CompileDepthStencilWrite, depth only (:1255-1291); SetDepthPixelXXX
         * keeps the stored stencil (GlslFunctions.cpp:880-896). */
        if (p->depth_test && p->depth_write && not_s8 && depth_ok) {
            uint32_t s = c->dsfmt.stencil_off < 0 ? 0u : cpvref_unpack_stencil(p->depth_stencil_format, ds_px(c, x, y));
            cpvref_pack_depth_stencil(p->depth_stencil_format, depth, s, ds_px(c, x, y));
        }
    }
#undef DEPTH_TEST

    /* CompileWriteFragment (:1415-1429) */
    if (!(stencil_ok && depth_ok) || !c->has_color) return;
    const cpv_blend_attachment* b = &p->blend[0];
    uint8_t* px = color_px(c, x, y);
    float value[4] = {colour[0], colour[1], colour[2], colour[3]};
    if (c->cfmt.base == BT_UINT) {
        uint32_t u[4];
        memcpy(u, colour, 16);
        cpvref_pack_u32(p->color_formats[0], u, px);
        return;
    }
    if (c->cfmt.base == BT_SINT) {
        int32_t s[4];
        memcpy(s, colour, 16);
        cpvref_pack_i32(p->color_formats[0], s, px);
        return;
    }
    if (b->blend_enable) {
        float dst[4], out[4];
        cpvref_unpack_f32(p->color_formats[0], px, dst);
        cpvref_apply_blend(value, dst, p->blend_constants, b, out);
        memcpy(value, out, 16);
    }
    cpvref_pack_f32(p->color_formats[0], value, px);
}

/* SetDatum (CommandBuffer.Draw.cpp:816-837) for one float component over
 * n vertices: perspective (mode 0) or linear (mode 1). */
static float set_datum(int mode, int n, const float* pw, const float* v, const float* w)
{
    if (mode == 1) {
        float r = 0.0f;
        for (int i = 0; i < n; i++) r += w[i] * v[i];
        return r;
    }
    float num = 0.0f, den = 0.0f;
    for (int i = 0; i < n; i++) {
        num += w[i] * v[i] / pw[i];
        den += w[i] / pw[i];
    }
    return num / den;
}

/* One triangle after the per-primitive part of ProcessTriangles
 * (CommandBuffer.Draw.cpp:1526-1569): the CW swap, the perspective divide and
 * the truncated pixel box [x0, x1) x [y0, y1). */
typedef struct {
    float P[3][4];
    const float* q[3];
    uint32_t provoking;
    int32_t x0, y0, x1, y1;
} tri_t;

typedef struct {
    float vpw, vph, half_x, half_y;
    int32_t limit_x, limit_y;
} tri_view_t;

static void tri_view(const ctx_t* c, tri_view_t* v)
{
    const float* vp = c->draw->viewport;
    v->vpw = vp[2];
    v->vph = vp[3];
    v->half_x = 1.0f / v->vpw * 0.5f;
    v->half_y = 1.0f / v->vph * 0.5f;
    v->limit_x = f2i_trunc(v->vpw);
    v->limit_y = f2i_trunc(v->vph);
    /* Guard the attachment extent (a viewport larger than the framebuffer
     * writes out of bounds in the reference). */
    if (v->limit_x > (int32_t)c->pass->width) v->limit_x = (int32_t)c->pass->width;
    if (v->limit_y > (int32_t)c->pass->height) v->limit_y = (int32_t)c->pass->height;
}

static void tri_setup(const ctx_t* c, const tri_view_t* v, const uint32_t prim[4], const float* vout, int vstride, tri_t* t)
{
    t->provoking = prim[0];
    uint32_t i0 = prim[1], i1 = prim[2], i2 = prim[3];
    if (c->pipe->front_face == CPV_FRONT_FACE_CW) { uint32_t s = i0; i0 = i2; i2 = s; }
    t->q[0] = vout + (uint64_t)i0 * vstride;
    t->q[1] = vout + (uint64_t)i1 * vstride;
    t->q[2] = vout + (uint64_t)i2 * vstride;
    for (int k = 0; k < 3; k++) {
        t->P[k][0] = t->q[k][0] / t->q[k][3];
        t->P[k][1] = t->q[k][1] / t->q[k][3];
        t->P[k][2] = t->q[k][2] / t->q[k][3];
        t->P[k][3] = t->q[k][3];
    }
    int32_t sx[3], sy[3];
    for (int k = 0; k < 3; k++) {
        sx[k] = f2i_trunc((t->P[k][0] + 1.0f) * 0.5f * v->vpw);
        sy[k] = f2i_trunc((t->P[k][1] + 1.0f) * 0.5f * v->vph);
    }
    int32_t mnx = sx[0], mny = sy[0], mxx = sx[0], mxy = sy[0];
    for (int k = 1; k < 3; k++) {
        if (sx[k] < mnx) mnx = sx[k];
        if (sy[k] < mny) mny = sy[k];
        if (sx[k] > mxx) mxx = sx[k];
        if (sy[k] > mxy) mxy = sy[k];
    }
    t->x0 = mnx > 0 ? mnx : 0;
    t->y0 = mny > 0 ? mny : 0;
    t->x1 = (int32_t)((int64_t)mxx + 1 < v->limit_x ? (int64_t)mxx + 1 : v->limit_x);
    t->y1 = (int32_t)((int64_t)mxy + 1 < v->limit_y ? (int64_t)mxy + 1 : v->limit_y);
}

/* The per-pixel part of ProcessTriangles over the triangle's box clipped to
 * the context's window. */
static void tri_raster(ctx_t* c, const tri_view_t* v, const tri_t* t, const float* vout, int vstride, int nvary)
{
    const cpv_pipeline_desc* p = c->pipe;
    const float* vp = c->draw->viewport;
    const int32_t ys = t->y0 > c->wy0 ? t->y0 : c->wy0, ye = t->y1 < c->wy1 ? t->y1 : c->wy1;
    const int32_t xs = t->x0 > c->wx0 ? t->x0 : c->wx0, xe = t->x1 < c->wx1 ? t->x1 : c->wx1;
    const float (*P)[4] = t->P;
    for (int32_t y = ys; y < ye; y++) {
        const float yf = ((float)y / v->vph + v->half_y) * 2.0f - 1.0f;
        for (int32_t x = xs; x < xe; x++) {
            const float xf = ((float)x / v->vpw + v->half_x) * 2.0f - 1.0f;
            /* GetFragmentInput (:874-954) */
            float area = edge(P[0], P[1], P[2][0], P[2][1]);
            float w0, w1, w2;
            int front;
            if (area < 0) {
                area = -area;
                front = 0;
                w0 = edge(P[2], P[1], xf, yf);
                w1 = edge(P[0], P[2], xf, yf);
                w2 = edge(P[1], P[0], xf, yf);
            } else {
                front = 1;
                w0 = edge(P[1], P[2], xf, yf);
                w1 = edge(P[2], P[0], xf, yf);
                w2 = edge(P[0], P[1], xf, yf);
            }
            if (w0 < 0 || w1 < 0 || w2 < 0 || ((p->cull_mode & CPV_CULL_BACK) && !front) ||
                ((p->cull_mode & CPV_CULL_FRONT) && front))
                continue;
            c->fragments++;
            w0 /= area;
            w1 /= area;
            w2 /= area;
            float depth = P[0][2] * w0 + P[1][2] * w1 + P[2][2] * w2;
            float in[16];
            const float pw[3] = {P[0][3], P[1][3], P[2][3]};
            const float w[3] = {w0, w1, w2};
            const int mode = fs_interp(p->fs_kind);
            for (int k = 0; k < nvary; k++) {
                const float vv[3] = {t->q[0][4 + k], t->q[1][4 + k], t->q[2][4 + k]};
                /* Flat: the provoking vertex's value (:943-946) */
                in[k] = mode == 2 ? vout[(uint64_t)t->provoking * vstride + 4 + k] : set_datum(mode, 3, pw, vv, w);
            }
            /* DrawPixel (:1300-1313) */
            float dt = (vp[5] - vp[4]) * depth + vp[4];
            float colour[4];
            if (run_fs(p, c->draw, in, colour)) continue;
            fragment_ops(c, x, y, dt, front, colour);
        }
    }
}

/* ProcessTriangles (CommandBuffer.Draw.cpp:1510-1594) over primitives
 * [pb, pe) of the assembled list, in order. */
static void process_triangles(ctx_t* c, const uint32_t (*prims)[4], uint64_t nprims, const float* vout,
                              int vstride, int nvary, uint64_t pb, uint64_t pe)
{
    tri_view_t v;
    tri_view(c, &v);
    if (pe > nprims) pe = nprims;
    for (uint64_t pi = pb; pi < pe; pi++) {
        tri_t t;
        tri_setup(c, &v, prims[pi], vout, vstride, &t);
        tri_raster(c, &v, &t, vout, vstride, nvary);
    }
}

/* Tile-parallel ProcessTriangles: the framebuffer is cut into kParTile x
 * kParTile pixel tiles; the triangles are set up once and binned to the
 * tiles their boxes meet, in primitive order; each tile then runs the
 * per-pixel loop of its triangles over its own pixels.  Pixels are
 * independent and each sees the same triangles in the same order with the
 * same arithmetic, so the result is bit-identical to process_triangles. */
enum { kParTile = 64 };

static void process_triangles_parallel(ctx_t* c, const uint32_t (*prims)[4], uint64_t nprims, const float* vout,
                                       int vstride, int nvary, uint64_t pb, uint64_t pe)
{
    tri_view_t v;
    tri_view(c, &v);
    if (pe > nprims) pe = nprims;
    if (pb >= pe) return;
    const uint64_t n = pe - pb;
    tri_t* tris = (tri_t*)malloc(sizeof(tri_t) * n);
    const int32_t tx = (int32_t)((c->pass->width + kParTile - 1) / kParTile), ty = (int32_t)((c->pass->height + kParTile - 1) / kParTile);
    const uint64_t ntiles = (uint64_t)tx * (uint64_t)ty;
    uint64_t* start = (uint64_t*)calloc(ntiles + 1, sizeof(uint64_t));
#pragma omp parallel for schedule(static) num_threads(c->threads)
    for (int64_t i = 0; i < (int64_t)n; i++) tri_setup(c, &v, prims[pb + (uint64_t)i], vout, vstride, &tris[i]);
    /* bin: count, scan, fill (primitive order within every tile) */
    for (uint64_t i = 0; i < n; i++) {
        const tri_t* t = &tris[i];
        if (t->x0 >= t->x1 || t->y0 >= t->y1) continue;
        for (int32_t b = t->y0 / kParTile; b <= (t->y1 - 1) / kParTile; b++)
            for (int32_t a = t->x0 / kParTile; a <= (t->x1 - 1) / kParTile; a++) start[(uint64_t)b * tx + a + 1]++;
    }
    for (uint64_t k = 0; k < ntiles; k++) start[k + 1] += start[k];
    uint64_t* fill = (uint64_t*)malloc(sizeof(uint64_t) * (ntiles ? ntiles : 1));
    memcpy(fill, start, sizeof(uint64_t) * ntiles);
    uint32_t* list = (uint32_t*)malloc(sizeof(uint32_t) * (start[ntiles] ? start[ntiles] : 1));
    for (uint64_t i = 0; i < n; i++) {
        const tri_t* t = &tris[i];
        if (t->x0 >= t->x1 || t->y0 >= t->y1) continue;
        for (int32_t b = t->y0 / kParTile; b <= (t->y1 - 1) / kParTile; b++)
            for (int32_t a = t->x0 / kParTile; a <= (t->x1 - 1) / kParTile; a++) list[fill[(uint64_t)b * tx + a]++] = (uint32_t)i;
    }
    uint64_t frags = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : frags) num_threads(c->threads)
    for (int64_t k = 0; k < (int64_t)ntiles; k++) {
        ctx_t tc = *c;
        tc.fragments = 0;
        tc.wx0 = (int32_t)(k % tx) * kParTile;
        tc.wy0 = (int32_t)(k / tx) * kParTile;
        tc.wx1 = tc.wx0 + kParTile;
        tc.wy1 = tc.wy0 + kParTile;
        for (uint64_t j = start[k]; j < start[k + 1]; j++) tri_raster(&tc, &v, &tris[list[j]], vout, vstride, nvary);
        frags += tc.fragments;
    }
    c->fragments += frags;
    free(list);
    free(fill);
    free(start);
    free(tris);
}

/* int32 sum with two's-complement wrap, as the reference's compiled int
 * arithmetic behaves. */
static int32_t wrap_add(int32_t a, int32_t b) { return (int32_t)((uint32_t)a + (uint32_t)b); }
static int32_t wrap_sub(int32_t a, int32_t b) { return (int32_t)((uint32_t)a - (uint32_t)b); }

/* ProcessPoints (CommandBuffer.Draw.cpp:1315-1383) over points [pb, pe):
 * the square around (int)((p.x + 1) * 0.5f * (width - 1)) of half extent
 * (int)ceil(pointSize / 2), pixels with s, t in [0, 1], the vertex's
 * outputs copied and its depth. */
static void process_points(ctx_t* c, uint64_t nverts, const float* vout, int vstride, int nvary, int psize,
                           uint64_t pb, uint64_t pe)
{
    const float* vp = c->draw->viewport;
    const float vpw = vp[2], vph = vp[3];
    int32_t limit_x = f2i_trunc(vpw), limit_y = f2i_trunc(vph);
    if (limit_x > (int32_t)c->pass->width) limit_x = (int32_t)c->pass->width;
    if (limit_y > (int32_t)c->pass->height) limit_y = (int32_t)c->pass->height;
    if (pe > nverts) pe = nverts;
    for (uint64_t pi = pb; pi < pe; pi++) {
        const float* q = vout + pi * (uint64_t)vstride;
        const float P[4] = {q[0] / q[3], q[1] / q[3], q[2] / q[3], q[3]};
        const int32_t cx = f2i_trunc((P[0] + 1.0f) * 0.5f * (vpw - 1.0f));
        const int32_t cy = f2i_trunc((P[1] + 1.0f) * 0.5f * (vph - 1.0f));
        const float size = psize >= 0 ? q[4 + psize] : 0.0f;
        const int32_t h = f2i_trunc(ceilf(size / 2.0f));
        int32_t sx = wrap_sub(cx, h), sy = wrap_sub(cy, h);
        int32_t ex = wrap_add(wrap_add(cx, h), 1), ey = wrap_add(wrap_add(cy, h), 1);
        if (sx < 0) sx = 0;
        if (sy < 0) sy = 0;
        if (ex > limit_x) ex = limit_x;
        if (ey > limit_y) ey = limit_y;
        if (sx < c->wx0) sx = c->wx0;
        if (sy < c->wy0) sy = c->wy0;
        if (ex > c->wx1) ex = c->wx1;
        if (ey > c->wy1) ey = c->wy1;
        for (int32_t y = sy; y < ey; y++) {
            for (int32_t x = sx; x < ex; x++) {
                const float s = 0.5f + (float)wrap_sub(x, cx) / size;
                const float t = 0.5f + (float)wrap_sub(y, cy) / size;
                if (!(s >= 0 && t >= 0 && s <= 1 && t <= 1)) continue;
                c->fragments++;
                float in[16];
                for (int k = 0; k < nvary; k++) in[k] = q[4 + k];
                const float dt = (vp[5] - vp[4]) * P[2] + vp[4];
                float colour[4];
                if (run_fs(c->pipe, c->draw, in, colour)) continue;
                fragment_ops(c, x, y, dt, 1, colour);
            }
        }
    }
}

/* ProcessLines (CommandBuffer.Draw.cpp:1385-1508) over lines [pb, pe):
 * every pixel of the viewport (x < width, y < height as floats) whose centre
 * passes the four edge tests of the quad around p0p1. */
static void process_lines(ctx_t* c, uint64_t nverts, const float* vout, int vstride, int nvary, uint64_t pb, uint64_t pe)
{
    const cpv_pipeline_desc* p = c->pipe;
    const float* vp = c->draw->viewport;
    const float vpw = vp[2], vph = vp[3];
    const float half_x = 1.0f / vpw * 0.5f, half_y = 1.0f / vph * 0.5f;
    uint64_t n = 0;
    if (p->topology == CPV_TOPOLOGY_LINE_LIST) n = nverts / 2;
    else if (nverts > 1) n = nverts - 1;
    if (pe > n) pe = n;
    /* the attachment extent guards the scan (the reference writes out of
     * bounds for a viewport larger than the framebuffer) */
    uint32_t lx = 0, ly = 0;
    while (lx < c->pass->width && (float)lx < vpw) lx++;
    while (ly < c->pass->height && (float)ly < vph) ly++;
    const float lw = p->line_width;
    for (uint64_t pi = pb; pi < pe; pi++) {
        const uint64_t i0 = p->topology == CPV_TOPOLOGY_LINE_LIST ? 2 * pi : pi;
        const float* q0 = vout + i0 * (uint64_t)vstride;
        const float* q1 = vout + (i0 + 1) * (uint64_t)vstride;
        const float P0[4] = {q0[0] / q0[3], q0[1] / q0[3], q0[2] / q0[3], q0[3]};
        const float P1[4] = {q1[0] / q1[3], q1[1] / q1[3], q1[2] / q1[3], q1[3]};
        const float lwx = lw / vpw, lwy = lw / vph;
        const float d[4] = {P1[0] - P0[0], P1[1] - P0[1], P1[2] - P0[2], P1[3] - P0[3]};
        /* glm::normalize: v * (1 / sqrt(x*x + y*y + z*z + w*w)), left to right */
        const float inv = 1.0f / sqrtf(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3]);
        const float dir[2] = {d[0] * inv, d[1] * inv};
        const float perp[2] = {dir[1], -dir[0]};
        const float p00[2] = {P0[0] + perp[0] * lwx, P0[1] + perp[1] * lwy};
        const float p01[2] = {P0[0] - perp[0] * lwx, P0[1] - perp[1] * lwy};
        const float p10[2] = {P1[0] + perp[0] * lwx, P1[1] + perp[1] * lwy};
        const float p11[2] = {P1[0] - perp[0] * lwx, P1[1] - perp[1] * lwy};
        const float len = sqrtf(d[0] * d[0] + d[1] * d[1]);
        const uint32_t ys = c->wy0 > 0 ? (uint32_t)c->wy0 : 0u, xs = c->wx0 > 0 ? (uint32_t)c->wx0 : 0u;
        const uint32_t ye = c->wy1 < (int32_t)ly ? (uint32_t)(c->wy1 > 0 ? c->wy1 : 0) : ly;
        const uint32_t xe = c->wx1 < (int32_t)lx ? (uint32_t)(c->wx1 > 0 ? c->wx1 : 0) : lx;
        for (uint32_t y = ys; y < ye; y++) {
            const float yf = ((float)y / vph + half_y) * 2.0f - 1.0f;
            for (uint32_t x = xs; x < xe; x++) {
                const float xf = ((float)x / vpw + half_x) * 2.0f - 1.0f;
                if (!(edge(p00, p01, xf, yf) >= 0 && edge(p11, p10, xf, yf) >= 0 && edge(p10, p00, xf, yf) >= 0 &&
                      edge(p01, p11, xf, yf) >= 0))
                    continue;
                c->fragments++;
                const float t = ((xf - P0[0]) * d[0] + (yf - P0[1]) * d[1]) / (len * len);
                const float w[2] = {1 - t, t};
                const float pw[2] = {P0[3], P1[3]};
                const int mode = fs_interp(p->fs_kind);
                float in[16];
                for (int k = 0; k < nvary; k++) {
                    /* SetDatum<..., 2>; Flat reads the provoking vertex, p0 (:1480-1482) */
                    const float v[2] = {q0[4 + k], q1[4 + k]};
                    in[k] = mode == 2 ? q0[4 + k] : set_datum(mode, 2, pw, v, w);
                }
                const float depth = P0[2] * t + P1[2] * (1 - t);
                const float dt = (vp[5] - vp[4]) * depth + vp[4];
                float colour[4];
                if (run_fs(p, c->draw, in, colour)) continue;
                fragment_ops(c, (int)x, (int)y, dt, 1, colour);
            }
        }
    }
}

/* CalculatePrimitives for triangle topologies (:567-673): rows of
 * {provoking, v0, v1, v2} indexing the assembled vertex list. */
static uint32_t (*assemble(uint32_t topology, uint64_t nverts, uint64_t* nprims))[4]
{
    uint64_t n = 0;
    if (topology == CPV_TOPOLOGY_TRIANGLE_LIST) n = nverts / 3;
    else if (topology == CPV_TOPOLOGY_TRIANGLE_STRIP || topology == CPV_TOPOLOGY_TRIANGLE_FAN) n = nverts > 2 ? nverts - 2 : 0;
    uint32_t (*pr)[4] = (uint32_t (*)[4])malloc(sizeof(uint32_t[4]) * (n ? n : 1));
    for (uint64_t i = 0; i < n; i++) {
        uint32_t ii = (uint32_t)i;
        if (topology == CPV_TOPOLOGY_TRIANGLE_LIST) {
            pr[i][0] = ii * 3; pr[i][1] = ii * 3; pr[i][2] = ii * 3 + 1; pr[i][3] = ii * 3 + 2;
        } else if (topology == CPV_TOPOLOGY_TRIANGLE_STRIP) {
            pr[i][0] = ii; pr[i][1] = ii; pr[i][2] = ii + 1; pr[i][3] = ii + 2;
        } else {
            pr[i][0] = ii + 1; pr[i][1] = 0; pr[i][2] = ii + 1; pr[i][3] = ii + 2;
        }
    }
    *nprims = n;
    return pr;
}

static int draw_one(ctx_t* c, uint64_t pb, uint64_t pe)
{
    const cpv_pipeline_desc* p = c->pipe;
    const cpv_draw_desc* d = c->draw;
    if (p->topology > CPV_TOPOLOGY_TRIANGLE_FAN) return CPV_ERROR_UNSUPPORTED;
    if (p->polygon_mode != CPV_POLYGON_FILL || p->rasterizer_discard || p->depth_clamp || p->logic_op_enable)
        return CPV_ERROR_UNSUPPORTED;
    if (p->color_count && p->blend[0].color_write_mask != CPV_COLOR_RGBA) return CPV_ERROR_UNSUPPORTED;
    vs_info_t vi;
    if (!vs_info(p->vs_kind, &vi) || fs_inputs(p->fs_kind) < 0) return CPV_ERROR_UNSUPPORTED;
    int nvary = fs_inputs(p->fs_kind);
    if (nvary > vi.nout) return CPV_ERROR_INVALID;
    const int vstride = 4 + vi.nout;
    for (uint32_t inst = 0; inst < d->instance_count; inst++) {
        uint64_t n = d->count;
        /* ProcessInputAssembler(Indexed): vertex ids per raw index */
        uint32_t* vid = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
        for (uint64_t i = 0; i < n; i++) {
            if (d->indexed) {
                const uint8_t* ib = (const uint8_t*)(uintptr_t)d->index_address;
                uint64_t at = d->first + i;
                uint32_t index;
                if (d->index_size == 1) index = ib[at];
                else if (d->index_size == 2) { uint16_t t; memcpy(&t, ib + at * 2, 2); index = t; }
                else { memcpy(&index, ib + at * 4, 4); }
                if (p->primitive_restart &&
                    index == (d->index_size == 1 ? 0xFFu : d->index_size == 2 ? 0xFFFFu : 0xFFFFFFFFu)) {
                    free(vid);
                    return CPV_ERROR_UNSUPPORTED;
                }
                vid[i] = (uint32_t)((int64_t)d->vertex_offset + index);
            } else {
                vid[i] = d->first + (uint32_t)i;
            }
        }
        /* ProcessVertexShader: one invocation per assembled vertex */
        float* vout = (float*)malloc(sizeof(float) * (size_t)vstride * (n ? n : 1));
        int vs_rc = 0;   /* the lowest (most negative) error code of any invocation */
#pragma omp parallel for schedule(static) reduction(min : vs_rc) num_threads(c->threads > 1 ? c->threads : 1)
        for (int64_t i = 0; i < (int64_t)n; i++) {
            int rc = run_vs(p, d, vid[i], d->first_instance + inst, vout + i * vstride, vout + i * vstride + 4);
            if (rc < vs_rc) vs_rc = rc;
        }
        if (vs_rc) { free(vid); free(vout); return vs_rc; }
        if (p->topology < CPV_TOPOLOGY_TRIANGLE_LIST && c->threads > 1) {
            /* points and lines: every tile runs the whole loop over its own pixels */
            const int32_t tx = (int32_t)((c->pass->width + kParTile - 1) / kParTile);
            const int64_t ntiles = (int64_t)tx * (int64_t)((c->pass->height + kParTile - 1) / kParTile);
            uint64_t frags = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : frags) num_threads(c->threads)
            for (int64_t k = 0; k < ntiles; k++) {
                ctx_t tc = *c;
                tc.fragments = 0;
                tc.wx0 = (int32_t)(k % tx) * kParTile;
                tc.wy0 = (int32_t)(k / tx) * kParTile;
                tc.wx1 = tc.wx0 + kParTile;
                tc.wy1 = tc.wy0 + kParTile;
                if (p->topology == CPV_TOPOLOGY_POINT_LIST)
                    process_points(&tc, n, vout, vstride, nvary, p->vs_kind == CPV_VS_PASSTHROUGH_POINT ? 4 : -1, pb, pe);
                else
                    process_lines(&tc, n, vout, vstride, nvary, pb, pe);
                frags += tc.fragments;
            }
            c->fragments += frags;
        } else if (p->topology == CPV_TOPOLOGY_POINT_LIST) {
            process_points(c, n, vout, vstride, nvary, p->vs_kind == CPV_VS_PASSTHROUGH_POINT ? 4 : -1, pb, pe);
        } else if (p->topology < CPV_TOPOLOGY_TRIANGLE_LIST) {
            process_lines(c, n, vout, vstride, nvary, pb, pe);
        } else {
            uint64_t nprims;
            uint32_t (*prims)[4] = assemble(p->topology, n, &nprims);
            if (c->threads > 1)
                process_triangles_parallel(c, (const uint32_t (*)[4])prims, nprims, vout, vstride, nvary, pb, pe);
            else
                process_triangles(c, (const uint32_t (*)[4])prims, nprims, vout, vstride, nvary, pb, pe);
            free(prims);
        }
        free(vout);
        free(vid);
    }
    return CPV_OK;
}

static int clear_pass(const cpv_pass_desc* pass)
{
    uint64_t npx = (uint64_t)pass->width * pass->height;
    for (uint32_t a = 0; a < pass->color_count; a++) {
        const cpv_attachment_desc* at = &pass->color[a];
        if (at->load_op != CPV_LOAD_OP_CLEAR) continue;
        fmt_t f;
        if (!fmt_get(at->format, &f) || f.type != FT_NORMAL) return CPV_ERROR_UNSUPPORTED;
        uint8_t px[16];
        int rc;
        /* SetPixel(VkClearColorValue) by base type (ImageSampler.cpp:749-774) */
        if (f.base == BT_UINT) rc = cpvref_pack_u32(at->format, at->clear.u32, px);
        else if (f.base == BT_SINT) rc = cpvref_pack_i32(at->format, at->clear.i32, px);
        else rc = cpvref_pack_f32(at->format, at->clear.f32, px);
        if (rc) return rc;
        uint8_t* dst = (uint8_t*)(uintptr_t)at->address;
        for (uint64_t i = 0; i < npx; i++) memcpy(dst + i * f.total, px, f.total);
    }
    if (pass->has_depth_stencil && pass->depth_stencil.load_op == CPV_LOAD_OP_CLEAR) {
        const cpv_attachment_desc* at = &pass->depth_stencil;
        fmt_t f;
        if (!fmt_get(at->format, &f) || f.type != FT_DEPTH_STENCIL) return CPV_ERROR_UNSUPPORTED;
        /* SetPixelDepthStencil per pixel, in place: the depth and stencil
         * bytes only (D32_SFLOAT_S8_UINT keeps its padding, ImageCompiler.cpp:895-945) */
        uint8_t* dst = (uint8_t*)(uintptr_t)at->address;
        for (uint64_t i = 0; i < npx; i++) cpvref_pack_depth_stencil(at->format, at->clear.ds.depth, at->clear.ds.stencil, dst + i * f.total);
    }
    return CPV_OK;
}

static int render_pass_impl(const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines,
                            const cpv_draw_desc* draws, const uint32_t* draw_pipeline,
                            uint32_t draw_count, uint64_t prim_begin, uint64_t prim_end, int threads,
                            uint64_t* fragments)
{
    if (!pass || (draw_count && (!pipelines || !draws))) return CPV_ERROR_INVALID;
    if (pass->color_count > CPV_MAX_COLOR_ATTACHMENTS) return CPV_ERROR_UNSUPPORTED;
    int rc = clear_pass(pass);
    if (rc) return rc;
    uint64_t frags = 0;
    for (uint32_t i = 0; i < draw_count; i++) {
        ctx_t c;
        memset(&c, 0, sizeof(c));
        c.pass = pass;
        c.wx0 = c.wy0 = INT32_MIN;
        c.wx1 = c.wy1 = INT32_MAX;
        c.threads = threads;
        c.pipe = &pipelines[draw_pipeline ? draw_pipeline[i] : 0];
        c.draw = &draws[i];
        c.has_color = pass->color_count > 0 && c.pipe->color_count > 0;
        if (c.has_color && !fmt_get(pass->color[0].format, &c.cfmt)) return CPV_ERROR_UNSUPPORTED;
        c.has_ds = pass->has_depth_stencil != 0;
        if (c.has_ds && !fmt_get(pass->depth_stencil.format, &c.dsfmt)) return CPV_ERROR_UNSUPPORTED;
        rc = draw_one(&c, prim_begin, prim_end);
        if (rc) return rc;
        frags += c.fragments;
    }
    if (fragments) *fragments = frags;
    return CPV_OK;
}

int cpvref_render_pass_range(const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines,
                             const cpv_draw_desc* draws, const uint32_t* draw_pipeline,
                             uint32_t draw_count, uint64_t prim_begin, uint64_t prim_end,
                             uint64_t* fragments)
{
    return render_pass_impl(pass, pipelines, draws, draw_pipeline, draw_count, prim_begin, prim_end, 1, fragments);
}

int cpvref_render_pass(const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines,
                       const cpv_draw_desc* draws, const uint32_t* draw_pipeline,
                       uint32_t draw_count, uint64_t* fragments)
{
    return render_pass_impl(pass, pipelines, draws, draw_pipeline, draw_count, 0, UINT64_MAX, 1, fragments);
}

int cpvref_render_pass_parallel(const cpv_pass_desc* pass, const cpv_pipeline_desc* pipelines,
                                const cpv_draw_desc* draws, const uint32_t* draw_pipeline,
                                uint32_t draw_count, int threads, uint64_t* fragments)
{
    return render_pass_impl(pass, pipelines, draws, draw_pipeline, draw_count, 0, UINT64_MAX, threads, fragments);
}

int cpvref_pixel_coverage(const float positions[12], uint32_t cull_mode, uint32_t front_face,
                          float vp_width, float vp_height, int32_t x, int32_t y,
                          float* depth, int* front, float weights[3])
{
    int o[3] = {0, 1, 2};
    if (front_face == CPV_FRONT_FACE_CW) { o[0] = 2; o[2] = 0; }
    float P[3][4];
    for (int k = 0; k < 3; k++) {
        const float* q = positions + 4 * o[k];
        P[k][0] = q[0] / q[3]; P[k][1] = q[1] / q[3]; P[k][2] = q[2] / q[3]; P[k][3] = q[3];
    }
    const float half_x = 1.0f / vp_width * 0.5f, half_y = 1.0f / vp_height * 0.5f;
    const float yf = ((float)y / vp_height + half_y) * 2.0f - 1.0f;
    const float xf = ((float)x / vp_width + half_x) * 2.0f - 1.0f;
    float area = edge(P[0], P[1], P[2][0], P[2][1]);
    float w0, w1, w2;
    int fr;
    if (area < 0) {
        area = -area; fr = 0;
        w0 = edge(P[2], P[1], xf, yf); w1 = edge(P[0], P[2], xf, yf); w2 = edge(P[1], P[0], xf, yf);
    } else {
        fr = 1;
        w0 = edge(P[1], P[2], xf, yf); w1 = edge(P[2], P[0], xf, yf); w2 = edge(P[0], P[1], xf, yf);
    }
    *front = fr;
    if (w0 < 0 || w1 < 0 || w2 < 0 || ((cull_mode & CPV_CULL_BACK) && !fr) || ((cull_mode & CPV_CULL_FRONT) && fr))
        return 0;
    w0 /= area; w1 /= area; w2 /= area;
    weights[0] = w0; weights[1] = w1; weights[2] = w2;
    *depth = P[0][2] * w0 + P[1][2] * w1 + P[2][2] * w2;
    return 1;
}


/* ------------------------------------------------------------------ */
/* Transfers (CPVulkan/CommandBuffer.cpp, CommandBuffer.Draw.cpp)      */
/* ------------------------------------------------------------------ */

/* FillBufferCommand::Process (CommandBuffer.cpp:307-326): whole words,
 * then the size % 4 tail bytes at byte size / 4 + i. */
void cpvref_fill_buffer(uint8_t* dst, uint64_t size, uint32_t data)
{
    uint64_t words = size / 4;
    for (uint64_t i = 0; i < words; i++) memcpy(dst + 4 * i, &data, 4);
    const uint8_t* bytes = (const uint8_t*)&data;
    for (uint64_t i = 0; i < size % 4; i++) dst[size / 4 + i] = bytes[i];
}

/* SetPixel of a VkClearColorValue / VkClearDepthStencilValue
 * (ImageSampler.cpp:723-800), chosen by the format's base type. */
static int pack_clear(uint32_t format, const uint32_t value[4], uint8_t* px)
{
    fmt_t f;
    if (!fmt_get(format, &f)) return CPV_ERROR_UNSUPPORTED;
    if (f.type == FT_DEPTH_STENCIL) {
        float depth;
        memcpy(&depth, &value[0], 4);
        return cpvref_pack_depth_stencil(format, depth, value[1], px);
    }
    if (f.base == BT_UINT) return cpvref_pack_u32(format, value, px);
    if (f.base == BT_SINT) return cpvref_pack_i32(format, (const int32_t*)value, px);
    float c[4];
    memcpy(c, value, 16);
    return cpvref_pack_f32(format, c, px);
}

/* ClearImage / ClearAttachmentsCommand (CommandBuffer.Draw.cpp:117-149,
 * 2176-2220): SetPixel of the clear value at every pixel of each rect. */
int cpvref_clear_rects(const cpv_image_level* image, const uint32_t value[4], const cpv_rect* rects, uint32_t count)
{
    uint32_t bytes = cpvref_format_size(image->format);
    uint8_t* base = (uint8_t*)(uintptr_t)image->address;
    /* packed in place: a depth/stencil store writes only its depth and
     * stencil bytes (D32_SFLOAT_S8_UINT keeps its padding) */
    for (uint32_t i = 0; i < count; i++)
        for (uint32_t y = 0; y < rects[i].height; y++)
            for (uint32_t x = 0; x < rects[i].width; x++) {
                int rc = pack_clear(image->format, value,
                                    base + ((uint64_t)(rects[i].y + (int32_t)y) * image->width + (uint64_t)(rects[i].x + (int32_t)x)) * bytes);
                if (rc) return rc;
            }
    return CPV_OK;
}

/* BlitImageCommand::Process (CommandBuffer.cpp:57-238) for one layer:
 * colour formats sample the source level at (u / W, v / H) with the blit
 * filter and CLAMP_TO_EDGE (SampleImage, ImageSampler.cpp:676-681), integer
 * formats take the nearest texel's channels, depth/stencil formats read the
 * texel at the truncated (u, v); each destination pixel is written with SetPixel. */
int cpvref_blit_image(const cpv_image_level* src, const cpv_image_level* dst, const cpv_blit_region* regions, uint32_t count,
                      uint32_t filter)
{
    fmt_t sf, df;
    if (!fmt_get(src->format, &sf) || !fmt_get(dst->format, &df)) return CPV_ERROR_UNSUPPORTED;
    const uint32_t sbytes = cpvref_format_size(src->format), dbytes = cpvref_format_size(dst->format);
    const uint8_t* sbase = (const uint8_t*)(uintptr_t)src->address;
    uint8_t* dbase = (uint8_t*)(uintptr_t)dst->address;
    level_t lv = {sbase, src->width, src->height};
    cpv_sampler_desc s;
    memset(&s, 0, sizeof(s));
    s.address_u = s.address_v = CPV_ADDRESS_CLAMP_TO_EDGE;
    for (uint32_t r = 0; r < count; r++) {
        const cpv_blit_region* g = &regions[r];
        int32_t w = g->dst[2] - g->dst[0], h = g->dst[3] - g->dst[1];
        const int negx = w < 0, negy = h < 0;
        if (negx) w = -w;
        if (negy) h = -h;
        const float rx = (float)(g->src[2] - g->src[0]) / (float)(g->dst[2] - g->dst[0]);
        const float ry = (float)(g->src[3] - g->src[1]) / (float)(g->dst[3] - g->dst[1]);
        for (int32_t y = 0; y < h; y++) {
            for (int32_t x = 0; x < w; x++) {
                const int32_t dx = negx ? x + g->dst[2] : x + g->dst[0];
                const int32_t dy = negy ? y + g->dst[3] : y + g->dst[1];
                const float u = ((float)dx + 0.5f - (float)g->dst[0]) * rx + (float)g->src[0];
                const float v = ((float)dy + 0.5f - (float)g->dst[1]) * ry + (float)g->src[1];
                uint8_t* out = dbase + ((uint64_t)dy * dst->width + (uint64_t)dx) * dbytes;
                if (sf.type == FT_DEPTH_STENCIL) {
                    uint32_t sx = (uint32_t)u, sy = (uint32_t)v;
                    const uint8_t* in = sbase + ((uint64_t)sy * src->width + sx) * sbytes;
                    float depth = df.depth_off >= 0 ? cpvref_unpack_depth(dst->format, out) : 0.0f;
                    uint32_t stencil = df.stencil_off >= 0 ? cpvref_unpack_stencil(dst->format, out) : 0u;
                    if (g->aspect & 2u) depth = cpvref_unpack_depth(src->format, in);
                    if (g->aspect & 4u) stencil = sf.stencil_off >= 0 ? cpvref_unpack_stencil(src->format, in) : 0u;
                    cpvref_pack_depth_stencil(dst->format, depth, stencil, out);
                } else if (sf.base == BT_UINT || sf.base == BT_SINT) {
                    int32_t cx = (int32_t)floorf(u / (float)src->width * (float)src->width);
                    int32_t cy = (int32_t)floorf(v / (float)src->height * (float)src->height);
                    cx = wrap_coord(cx, (int32_t)src->width, CPV_ADDRESS_CLAMP_TO_EDGE);
                    cy = wrap_coord(cy, (int32_t)src->height, CPV_ADDRESS_CLAMP_TO_EDGE);
                    const uint8_t* in = sbase + ((uint64_t)cy * src->width + (uint64_t)cx) * sbytes;
                    uint32_t ch[4] = {0, 0, 0, 1};
                    for (int k = 0; k < 4; k++) {
                        if (sf.off[k] < 0) continue;
                        ch[k] = sf.base == BT_SINT ? (uint32_t)load_sint(in + sf.off[k], sf.element) : load_uint(in + sf.off[k], sf.element);
                    }
                    if (sf.base == BT_UINT) cpvref_pack_u32(dst->format, ch, out);
                    else cpvref_pack_i32(dst->format, (const int32_t*)ch, out);
                } else {
                    float c[4];
                    sample_level(src->format, &lv, u / (float)src->width, v / (float)src->height, filter, &s, c);
                    if (sf.off[0] < 0) c[0] = 0;
                    if (sf.off[1] < 0) c[1] = 0;
                    if (sf.off[2] < 0) c[2] = 0;
                    if (sf.off[3] < 0) c[3] = 1;
                    cpvref_pack_f32(dst->format, c, out);
                }
            }
        }
    }
    return CPV_OK;
}

/* ------------------------------------------------------------------ */
/* Test hooks                                                          */
/* ------------------------------------------------------------------ */

int cpvref_fragment(const cpv_pipeline_desc* p, const cpv_pass_desc* pass, int32_t x, int32_t y, float depth, int front,
                    const float colour[4])
{
    if (!p || !pass || !colour) return CPV_ERROR_INVALID;
    ctx_t c;
    memset(&c, 0, sizeof(c));
    c.pass = pass;
    c.pipe = p;
    c.has_color = pass->color_count > 0 && p->color_count > 0;
    if (c.has_color && !fmt_get(pass->color[0].format, &c.cfmt)) return CPV_ERROR_UNSUPPORTED;
    c.has_ds = pass->has_depth_stencil != 0;
    if (c.has_ds && !fmt_get(pass->depth_stencil.format, &c.dsfmt)) return CPV_ERROR_UNSUPPORTED;
    if (x < 0 || y < 0 || (uint32_t)x >= pass->width || (uint32_t)y >= pass->height) return CPV_ERROR_INVALID;
    fragment_ops(&c, x, y, depth, front, colour);
    return CPV_OK;
}

void cpvref_glm(uint32_t op, const float* a, const float* b, float* out)
{
    switch (op) {
    case 0: mat4_mul_vec4(a, b, out); break;
    case 1: out[0] = dot3(a, b); break;
    case 2: normalize3(a, out); break;
    default: break;
    }
}
