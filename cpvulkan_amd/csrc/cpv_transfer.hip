// Transfer and clear commands on the context stream: region copies
// (vkCmdCopyBuffer / CopyImage / CopyBufferToImage / CopyImageToBuffer
// after the front end has turned their regions into byte rows), buffer
// fills and updates, rectangle clears (vkCmdClearColorImage,
// vkCmdClearDepthStencilImage, vkCmdClearAttachments) and blits.
//
// Each is enqueued on the same stream as the render passes, so a replayed
// command buffer keeps its order without host synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "cpv_internal.h"
#include "raster_kernels.h"

using namespace cpv;
using namespace cpvrt;

namespace {

// FillBufferCommand::Process (CPVulkan/CommandBuffer.cpp:307-326): the word
// fills every whole 4-byte unit; the bytes of a size % 4 tail go to byte
// size / 4 + i, as the reference indexes them.
__global__ void fill_words_kernel(uint32_t* dst, uint64_t words, uint32_t data)
{
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < words; i += (uint64_t)gridDim.x * blockDim.x) dst[i] = data;
}

__global__ void fill_tail_kernel(uint8_t* dst, uint64_t size, uint32_t data)
{
    const uint32_t i = threadIdx.x;
    if (i < (uint32_t)(size % 4)) dst[size / 4 + i] = (uint8_t)(data >> (8 * i));
}

struct ClearArgs {
    uint8_t* base;          // first byte of the image level
    FormatDesc fmt;
    uint32_t width;         // pixels per row
    int32_t x0, y0;
    uint32_t w, h;
    uint32_t value[4];      // VkClearValue words
};

// Bytes a pixel store writes: the reference's SetPixelDepthStencil stores
// the depth and the stencil byte of D32_SFLOAT_S8_UINT and leaves its three
// padding bytes (ImageCompiler.cpp:895-945); every other format is written whole.
CPV_DEV int stored_bytes(const FormatDesc& f)
{
    return (f.total == 8 && f.depth_off == 0 && f.stencil_off == 4) ? 5 : f.total;
}

// The packed pixel of a clear value (SetPixel overloads for
// VkClearColorValue and VkClearDepthStencilValue, ImageSampler.cpp:723-800).
CPV_DEV void pack_clear(const FormatDesc& f, const uint32_t value[4], uint32_t words[4])
{
    words[0] = words[1] = words[2] = words[3] = 0u;
    if (f.depth_off >= 0 || f.stencil_off >= 0) {
        pack_depth_stencil(f, __uint_as_float(value[0]), value[1], words[0], words[1]);
    } else if (f.base == kBaseUInt) {
        pack_color_u32(f, value, words);
    } else if (f.base == kBaseSInt) {
        const int32_t v[4] = {(int32_t)value[0], (int32_t)value[1], (int32_t)value[2], (int32_t)value[3]};
        pack_color_i32(f, v, words);
    } else {
        const float v[4] = {__uint_as_float(value[0]), __uint_as_float(value[1]), __uint_as_float(value[2]), __uint_as_float(value[3])};
        pack_color_f32(f, v, words);
    }
}

// SetPixel of the packed clear value over a rectangle (ClearImage,
// CommandBuffer.Draw.cpp:117-149; ClearAttachmentsCommand :2176-2220).
__global__ void clear_rect_kernel(ClearArgs a)
{
    uint32_t words[4];
    pack_clear(a.fmt, a.value, words);
    // rectangle pixels in row-major order, grid-strided; the stride's row
    // and column steps are divided out once
    const uint64_t n = (uint64_t)a.w * a.h;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t rx = (uint32_t)(i % a.w), ry = (uint32_t)(i / a.w);
    const uint32_t sx = (uint32_t)(step % a.w), sy = (uint32_t)(step / a.w);
    for (; i < n; i += step) {
        const uint32_t x = (uint32_t)a.x0 + rx, y = (uint32_t)a.y0 + ry;
        store_pixel(a.base + ((uint64_t)y * a.width + x) * (uint32_t)a.fmt.total, stored_bytes(a.fmt), words);
        rx += sx;
        ry += sy;
        if (rx >= a.w) { rx -= a.w; ry++; }
    }
}

struct BlitArgs {
    TextureDesc src;        // the source level as level 0, CLAMP_TO_EDGE
    uint8_t* dst;
    FormatDesc dfmt;
    uint32_t dst_width;
    int32_t s0x, s0y, s1x, s1y, d0x, d0y, d1x, d1y;
    uint32_t w, h;          // |d1 - d0|
    uint32_t filter;
    uint32_t aspect;        // VkImageAspectFlags of the region
    uint32_t src_kind;      // 0 float, 1 uint, 2 sint, 3 depth/stencil, 4 8-bit UNORM x4 to 8-bit UNORM x4
};

CPV_DEV uint32_t raw_channel(const FormatDesc& f, const uint32_t px[4], int c)
{
    return f.off[c] < 0 ? 0u : get_field(px, f.off[c], f.element);
}

CPV_DEV void read_pixel(const uint8_t* p, int bytes, uint32_t w[4])
{
    w[0] = w[1] = w[2] = w[3] = 0u;
    // unrolled over the largest pixel so w[] keeps constant indices (registers)
#pragma unroll
    for (int b = 0; b < 16; b++)
        if (b < bytes) w[b >> 2] |= (uint32_t)p[b] << ((b & 3) * 8);
}

// BlitImageCommand::Process (CPVulkan/CommandBuffer.cpp:57-238): per
// destination pixel, u = (dstX + 0.5 - dst0.x) * ((float)(src1.x - src0.x) /
// (dst1.x - dst0.x)) + src0.x (likewise v); colour formats are sampled at
// (u / W, v / H) with the blit filter and CLAMP_TO_EDGE (SampleImage,
// ImageSampler.cpp:676-681) and written with SetPixel; depth/stencil
// formats read the texel at the truncated (u, v).
// KIND is a.src_kind and FILTER a.filter, fixed per instantiation so each
// carries only its own path; KIND 4 is a colour blit whose source and
// destination formats both hold four 8-bit UNORM channels in 4 bytes, for
// which the generic format code folds to fixed shifts (the same arithmetic).
template <int KIND, int FILTER>
__global__ void blit_kernel(BlitArgs a)
{
    TextureDesc src = a.src;
    FormatDesc dfmt = a.dfmt;
    if (KIND == 4) {
        src.fmt.base = dfmt.base = kBaseUNorm;
        src.fmt.element = dfmt.element = 1;
        src.fmt.total = dfmt.total = 4;
        src.fmt.depth_off = src.fmt.stencil_off = dfmt.depth_off = dfmt.stencil_off = -1;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            __builtin_assume(src.fmt.off[k] >= 0 && src.fmt.off[k] < 4);
            __builtin_assume(dfmt.off[k] >= 0 && dfmt.off[k] < 4);
        }
    }
    const bool negx = a.d1x < a.d0x, negy = a.d1y < a.d0y;
    const float rx = (float)(a.s1x - a.s0x) / (float)(a.d1x - a.d0x);
    const float ry = (float)(a.s1y - a.s0y) / (float)(a.d1y - a.d0y);
    // destination pixels in row-major order, grid-strided as clear_rect_kernel
    const uint64_t n = (uint64_t)a.w * a.h;
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t ux = (uint32_t)(i % a.w), uy = (uint32_t)(i / a.w);
    const uint32_t sx = (uint32_t)(step % a.w), sy = (uint32_t)(step / a.w);
    for (; i < n; i += step, ux += sx, uy += sy) {
        if (ux >= a.w) { ux -= a.w; uy++; }
        const int32_t x = (int32_t)ux, y = (int32_t)uy;
        const int32_t dx = negx ? x + a.d1x : x + a.d0x;
        const int32_t dy = negy ? y + a.d1y : y + a.d0y;
        const float u = ((float)dx + 0.5f - (float)a.d0x) * rx + (float)a.s0x;
        const float v = ((float)dy + 0.5f - (float)a.d0y) * ry + (float)a.s0y;
        uint8_t* out = a.dst + ((uint64_t)dy * a.dst_width + (uint64_t)dx) * dfmt.total;
        uint32_t px[4] = {0u, 0u, 0u, 0u};
        if (KIND == 3) {
            const TextureDesc& t = a.src;
            // (uint32_t)u as GetDepthPixel receives it; held inside the level
            const uint32_t sx = min((uint32_t)fmaxf(u, 0.0f), t.level_w[0] - 1u), sy = min((uint32_t)fmaxf(v, 0.0f), t.level_h[0] - 1u);
            uint32_t sw[4];
            read_pixel((const uint8_t*)t.address + t.level_offset[0] + ((uint64_t)sy * t.level_w[0] + sx) * t.fmt.total, t.fmt.total, sw);
            uint32_t dw[4];
            read_pixel(out, a.dfmt.total, dw);
            float depth = a.dfmt.depth_off >= 0 ? unpack_depth(a.dfmt, dw[0]) : 0.0f;
            uint32_t stencil = a.dfmt.stencil_off >= 0 ? get_field(dw, a.dfmt.stencil_off, 1) : 0u;
            if (a.aspect & 2u) depth = unpack_depth(t.fmt, sw[0]);                        // DEPTH
            if (a.aspect & 4u) stencil = t.fmt.stencil_off >= 0 ? get_field(sw, t.fmt.stencil_off, 1) : 0u;   // STENCIL
            if ((a.aspect & 2u) && !(a.aspect & 4u) && a.dfmt.stencil_off < 0) stencil = 0u;
            if ((a.aspect & 4u) && !(a.aspect & 2u) && a.dfmt.depth_off < 0) depth = 0.0f;
            uint32_t d0 = 0u, d1 = 0u;
            pack_depth_stencil(a.dfmt, depth, stencil, d0, d1);
            px[0] = d0;
            px[1] = d1;
        } else if (KIND == 0 || KIND == 4) {
            float c[4];
            sample_level(src, 0, u / (float)src.level_w[0], v / (float)src.level_h[0], FILTER, c);
            if (src.fmt.off[0] < 0) c[0] = 0.0f;
            if (src.fmt.off[1] < 0) c[1] = 0.0f;
            if (src.fmt.off[2] < 0) c[2] = 0.0f;
            if (src.fmt.off[3] < 0) c[3] = 1.0f;
            pack_color_f32(dfmt, c, px);
        } else {
            // integer formats: the nearest texel's channel values
            const TextureDesc& t = a.src;
            int32_t cx = (int32_t)floorf(u / (float)t.level_w[0] * (float)t.level_w[0]);
            int32_t cy = (int32_t)floorf(v / (float)t.level_h[0] * (float)t.level_h[0]);
            cx = wrap_coord(cx, (int32_t)t.level_w[0], 2);
            cy = wrap_coord(cy, (int32_t)t.level_h[0], 2);
            uint32_t sw[4];
            read_pixel((const uint8_t*)t.address + t.level_offset[0] + ((uint64_t)cy * t.level_w[0] + (uint64_t)cx) * t.fmt.total, t.fmt.total, sw);
            uint32_t ch[4];
            for (int k = 0; k < 4; k++) ch[k] = raw_channel(t.fmt, sw, k);
            if (t.fmt.off[3] < 0) ch[3] = 1u;
            if (KIND == 1) {
                pack_color_u32(a.dfmt, ch, px);
            } else {
                int32_t si[4];
                for (int k = 0; k < 4; k++) {
                    const int sh = 32 - 8 * t.fmt.element;
                    si[k] = (t.fmt.off[k] < 0 || sh <= 0) ? (int32_t)ch[k] : ((int32_t)(ch[k] << sh) >> sh);
                }
                pack_color_i32(a.dfmt, si, px);
            }
        }
        store_pixel(out, stored_bytes(dfmt), px);
    }
}

uint32_t grid_for(uint64_t n)
{
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((n + 255) / 256, 16384));
}

}  // namespace

extern "C" int cpv_copy_regions(cpv_context* c, const cpv_copy_region* r, uint32_t count)
{
    if (!c || (!r && count)) return fail(CPV_ERROR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->lock);
    note_enqueue(c);
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    for (uint32_t i = 0; i < count; i++) {
        if (!r[i].rows || !r[i].row_bytes) continue;
        if (r[i].rows > 1 && (r[i].src_pitch < r[i].row_bytes || r[i].dst_pitch < r[i].row_bytes))
            return fail(CPV_ERROR_INVALID, "copy pitch below the row size");
        if (r[i].rows == 1 || (r[i].src_pitch == r[i].row_bytes && r[i].dst_pitch == r[i].row_bytes)) {
            HIP_TRY(hipMemcpyAsync((void*)(uintptr_t)r[i].dst, (const void*)(uintptr_t)r[i].src, r[i].row_bytes * r[i].rows,
                                   hipMemcpyDefault, c->stream));
        } else {
            HIP_TRY(hipMemcpy2DAsync((void*)(uintptr_t)r[i].dst, r[i].dst_pitch, (const void*)(uintptr_t)r[i].src, r[i].src_pitch,
                                     r[i].row_bytes, r[i].rows, hipMemcpyDefault, c->stream));
        }
    }
    return CPV_OK;
}

extern "C" int cpv_fill_buffer(cpv_context* c, uint64_t dst, uint64_t size, uint32_t data)
{
    if (!c || (!dst && size)) return fail(CPV_ERROR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->lock);
    note_enqueue(c);
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    if (size / 4 && (dst & 3u) == 0) {
        uint32_t* p = (uint32_t*)(uintptr_t)dst;
        uint64_t words = size / 4;
        void* args[] = {&p, &words, &data};
        HIP_TRY(hipLaunchKernel((const void*)fill_words_kernel, dim3(grid_for(words)), dim3(256), args, 0, c->stream));
    } else if (size / 4) {
        return fail(CPV_ERROR_INVALID, "fill destination must be 4-byte aligned");
    }
    if (size % 4) {
        uint8_t* p = (uint8_t*)(uintptr_t)dst;
        void* args[] = {&p, &size, &data};
        HIP_TRY(hipLaunchKernel((const void*)fill_tail_kernel, dim3(1), dim3(64), args, 0, c->stream));
    }
    return CPV_OK;
}

extern "C" int cpv_update_buffer(cpv_context* c, uint64_t dst, const void* data, uint64_t size)
{
    if (!c || ((!dst || !data) && size)) return fail(CPV_ERROR_INVALID, "null argument");
    std::lock_guard<std::mutex> g(c->lock);
    note_enqueue(c);
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    // through the pinned staging ring, so the caller's memory is free on return
    return stage_upload(c, (void*)(uintptr_t)dst, data, size);
}

extern "C" int cpv_clear_rects(cpv_context* c, const cpv_image_level* image, const uint32_t value[4], const cpv_rect* rects,
                               uint32_t count)
{
    if (!c || !image || !value || (!rects && count)) return fail(CPV_ERROR_INVALID, "null argument");
    FormatDesc f;
    if (!format_desc(image->format, &f)) return fail(CPV_ERROR_UNSUPPORTED, "clear of an unsupported format");
    std::lock_guard<std::mutex> g(c->lock);
    note_enqueue(c);
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    ClearArgs a{};
    a.base = (uint8_t*)(uintptr_t)image->address;
    a.fmt = f;
    a.width = image->width;
    std::memcpy(a.value, value, 16);
    for (uint32_t i = 0; i < count; i++) {
        const cpv_rect& r = rects[i];
        if (r.x < 0 || r.y < 0 || (uint64_t)r.x + r.width > image->width || (uint64_t)r.y + r.height > image->height)
            return fail(CPV_ERROR_INVALID, "clear rectangle outside the image");
        if (!r.width || !r.height) continue;
        a.x0 = r.x;
        a.y0 = r.y;
        a.w = r.width;
        a.h = r.height;
        void* args[] = {&a};
        HIP_TRY(hipLaunchKernel((const void*)clear_rect_kernel, dim3(grid_for((uint64_t)r.width * r.height)), dim3(256), args, 0, c->stream));
    }
    return CPV_OK;
}

extern "C" int cpv_blit_image(cpv_context* c, const cpv_image_level* src, const cpv_image_level* dst, const cpv_blit_region* regions,
                              uint32_t count, uint32_t filter)
{
    if (!c || !src || !dst || (!regions && count)) return fail(CPV_ERROR_INVALID, "null argument");
    FormatDesc sf, df;
    if (!format_desc(src->format, &sf) || !format_desc(dst->format, &df)) return fail(CPV_ERROR_UNSUPPORTED, "blit of an unsupported format");
    const bool sds = sf.depth_off >= 0 || sf.stencil_off >= 0, dds = df.depth_off >= 0 || df.stencil_off >= 0;
    if (sds != dds) return fail(CPV_ERROR_INVALID, "blit between colour and depth/stencil formats");
    if (sf.total > 16 || df.total > 16 || (!sds && sf.element > 4)) return fail(CPV_ERROR_UNSUPPORTED, "blit of a format wider than 4 bytes per channel");
    const bool integer = sf.base == kBaseUInt || sf.base == kBaseSInt;
    if ((sds || integer) && filter != 0) return fail(CPV_ERROR_INVALID, "depth/stencil and integer blits use VK_FILTER_NEAREST");
    std::lock_guard<std::mutex> g(c->lock);
    note_enqueue(c);
    if (int rc = resolve_pending(c)) return rc;
    HIP_TRY(hipSetDevice(c->device));
    BlitArgs a;
    std::memset(&a, 0, sizeof(a));
    a.src.address = src->address;
    a.src.fmt = sf;
    a.src.level_count = 1;
    a.src.level_offset[0] = 0;
    a.src.level_w[0] = src->width;
    a.src.level_h[0] = src->height;
    a.src.sampler.address_u = a.src.sampler.address_v = 2;   // CLAMP_TO_EDGE
    a.src.sampler.mag_filter = a.src.sampler.min_filter = (int32_t)filter;
    a.dst = (uint8_t*)(uintptr_t)dst->address;
    a.dfmt = df;
    a.dst_width = dst->width;
    a.filter = filter;
    a.src_kind = sds ? 3u : sf.base == kBaseUInt ? 1u : sf.base == kBaseSInt ? 2u : 0u;
    auto unorm8x4 = [](const FormatDesc& f) {
        return f.base == kBaseUNorm && f.element == 1 && f.total == 4 && f.depth_off < 0 && f.stencil_off < 0 &&
               f.off[0] >= 0 && f.off[1] >= 0 && f.off[2] >= 0 && f.off[3] >= 0;
    };
    if (a.src_kind == 0 && unorm8x4(sf) && unorm8x4(df)) a.src_kind = 4;
    for (uint32_t i = 0; i < count; i++) {
        const cpv_blit_region& r = regions[i];
        const int32_t w = std::abs(r.dst[2] - r.dst[0]), h = std::abs(r.dst[3] - r.dst[1]);
        if (!w || !h) continue;
        const int32_t dx0 = std::min(r.dst[0], r.dst[2]), dy0 = std::min(r.dst[1], r.dst[3]);
        if (dx0 < 0 || dy0 < 0 || dx0 + w > (int32_t)dst->width || dy0 + h > (int32_t)dst->height)
            return fail(CPV_ERROR_INVALID, "blit destination outside the image");
        const int32_t sx0 = std::min(r.src[0], r.src[2]), sy0 = std::min(r.src[1], r.src[3]);
        const int32_t sx1 = std::max(r.src[0], r.src[2]), sy1 = std::max(r.src[1], r.src[3]);
        if (sx0 < 0 || sy0 < 0 || sx1 > (int32_t)src->width || sy1 > (int32_t)src->height)
            return fail(CPV_ERROR_INVALID, "blit source outside the image");
        a.s0x = r.src[0]; a.s0y = r.src[1]; a.s1x = r.src[2]; a.s1y = r.src[3];
        a.d0x = r.dst[0]; a.d0y = r.dst[1]; a.d1x = r.dst[2]; a.d1y = r.dst[3];
        a.w = (uint32_t)w;
        a.h = (uint32_t)h;
        a.aspect = r.aspect;
        void* args[] = {&a};
        const void* fn = a.src_kind == 3 ? (const void*)blit_kernel<3, 0>
                         : a.src_kind == 2 ? (const void*)blit_kernel<2, 0>
                         : a.src_kind == 1 ? (const void*)blit_kernel<1, 0>
                         : a.src_kind == 4 ? (filter ? (const void*)blit_kernel<4, 1> : (const void*)blit_kernel<4, 0>)
                         : filter ? (const void*)blit_kernel<0, 1> : (const void*)blit_kernel<0, 0>;
        HIP_TRY(hipLaunchKernel(fn, dim3(grid_for((uint64_t)w * h)), dim3(256), args, 0, c->stream));
    }
    return CPV_OK;
}
