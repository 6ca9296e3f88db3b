// Host/device shared layouts of the raster pipeline.
//
// Everything here is plain data that the host fills and the kernels read:
// the per-format descriptor, the per-triangle setup record written by the
// setup kernel and read by the fine raster kernel, and the kernel parameter
// blocks.  Included by the AOT kernels, the hiprtc-compiled program modules
// and the host runtime.
#pragma once

#ifndef __HIPCC_RTC__
#include <stdint.h>
#else
// hiprtc declares the fixed-width integer types in its own namespace
using __hip_internal::int8_t;
using __hip_internal::int16_t;
using __hip_internal::int32_t;
using __hip_internal::int64_t;
using __hip_internal::uint8_t;
using __hip_internal::uint16_t;
using __hip_internal::uint32_t;
using __hip_internal::uint64_t;
#endif

namespace cpv {

// Work queues of the persistent raster kernel.  Tile i of the launch's
// queue order belongs to queue i % kRasterQueues, and a wave starts on the
// queue of the XCD its workgroup was dispatched to (workgroups go to the
// MI355X's 8 XCDs round-robin by index), so the wave slots of one tile run
// under one L2 and share its copies of the tile's triangle records.  A
// wave whose queue is empty takes work from the others in turn.  Counters
// are kRasterQueueStride words apart, one 128-byte line each.
constexpr uint32_t kRasterQueues = 8;
constexpr uint32_t kRasterQueueStride = 32;

// Channel base types of FormatInformation (CPVulkanBase/Formats.h:30-103).
enum BaseType : int32_t {
    kBaseNone = 0,
    kBaseUNorm = 1,
    kBaseSNorm = 2,
    kBaseUScaled = 3,
    kBaseSScaled = 4,
    kBaseUInt = 5,
    kBaseSInt = 6,
    kBaseSFloat = 7,
};

// Pixel format as the kernels see it: byte offsets of R,G,B,A (or -1),
// element size, total pixel size; depth/stencil byte offsets for
// depth/stencil formats.  Mirrors the normal and depth/stencil rows of
// the reference's format table (CPVulkanBase/Formats.cpp:209-444).
struct FormatDesc {
    uint32_t vk_format;
    int32_t base;
    int32_t element;
    int32_t total;
    int32_t off[4];
    int32_t depth_off;
    int32_t stencil_off;
    int32_t depth_kind;   // 0 none, 16 = D16, 24 = D24, 32 = D32
    int32_t pad;
};

// Triangle setup record: 32 dwords (128 B, one L2 line).  Edge k is
// E_k(p) = (p.x - ax[k]) * dy[k] - (p.y - ay[k]) * dx[k], the reference's
// EdgeFunction (CPVulkan/CommandBuffer.Draw.cpp:410-418) with the
// per-triangle differences (b.y - a.y), (b.x - a.x) hoisted; the
// weight order and orientation already follow GetFragmentInput's sign
// branch (:879-898), so the raster kernel never re-derives them.
struct TriSetup {
    float ax[3], ay[3];       // bytes 0-23
    float dy[3], dx[3];       // 24-47
    int32_t x0, y0, x1, y1;   // 48-63: pixel bounding box, [x0, x1) x [y0, y1)
    float dt_lo, dt_hi;       // 64-71: bounds of the viewport-transformed depth of every covered pixel
    float area;               // 72: |EdgeFunction(p0, p1, p2)|
    uint32_t draw;            // 76: draw index within the launch
    float z[3];               // 80-91: p_k.z / p_k.w
    uint32_t flags;           // 92: bit 0 front facing, the primitive kind
    // Bytes 0-95 are everything a primitive's visit of a pixel block reads
    // (kVisitParts 16-byte parts); the rest only the colour work reads.
    float w[3];               // 96-107: clip-space w of p_k
    uint32_t v[3];            // 108-119: vertex-output indices of p0, p1, p2 (after the CW swap)
    uint32_t provoking;       // 120: vertex-output index of the provoking vertex
    uint32_t pad;
};
static_assert(sizeof(TriSetup) == 128, "TriSetup must be 128 bytes");
constexpr int kVisitParts = 6;
static_assert(__builtin_offsetof(TriSetup, flags) == 92 && __builtin_offsetof(TriSetup, w) == 16 * kVisitParts,
              "the visit fields fill the first kVisitParts parts");

// TriSetup.flags: bit 0 front facing; the primitive kind.  A point record
// holds its screen centre (int bits of ax[0], ay[0]), its size (dx[0]), its
// clamped square [x0, x1) x [y0, y1) and its vertex in v[0]; a line record
// holds three edges of its quad (the fourth in LineExt), |p1.xy - p0.xy|^2
// in area, z = (p0.z, p1.z), w = (p0.w, p1.w) and v = (p0, p1).
enum : uint32_t { kPrimFront = 1u, kPrimPoint = 2u, kPrimLine = 4u };

// A (tile, primitive) pair's value: the primitive index in the low 28 bits
// (launches hold fewer than 2^28 primitives) and, for 16x16 and 32x32
// tiles, in bits 28-31 which of the tile's four quarters (8x8 blocks or
// 16x16 quadrants, bit by * 2 + bx) the primitive's pixel box overlaps,
// less, for a triangle, the quarters where an edge function is negative at
// every pixel.
constexpr uint32_t kPairPrim = 0x0FFFFFFFu;
constexpr int kPairBlockShift = 28;

// Raster kernel variants of a program: the tile configurations 8x1, 16x1,
// 16x2, 16x4, 32x1, 32x2, 32x4, then 16x4 with point and line records, then
// the seven configurations again for launches whose every pipeline and
// target has the fixed formats (fold_fixed_formats).
constexpr int kRasterVariants = 15;
constexpr int kRasterAnyPrim = 7;
constexpr int kRasterFixedFormats = 8;   // + tile configuration index

// Per-primitive extension of a line record (only lines write it).
struct LineExt {
    float ax3, ay3, dy3, dx3;   // fourth edge
    float p0x, p0y, ddx, ddy;   // p0.xy and p1.xy - p0.xy for t
};
static_assert(sizeof(LineExt) == 32, "LineExt must be 32 bytes");

enum { kMaxVaryings = 16 };

// Fixed-function state of one pipeline as the fragment stage needs it.
struct FragmentState {
    FormatDesc color;
    FormatDesc ds;
    int32_t has_color, has_ds;
    int32_t depth_test, depth_write, depth_compare, depth_bounds_test, stencil_test;
    int32_t blend_enable;
    int32_t src_color, dst_color, color_op, src_alpha, dst_alpha, alpha_op;
    float min_depth_bounds, max_depth_bounds;
    float blend_constants[4];
    // front / back stencil: fail, pass, depth_fail, compare, compare_mask, write_mask, reference
    int32_t st_front[7];
    int32_t st_back[7];
    int32_t origin_upper_left;
    // 1 when blending is SRC_ALPHA / ONE_MINUS_SRC_ALPHA / ADD for both colour
    // and alpha (set by the host; the kernel then skips the factor and op
    // selection), else 0
    int32_t blend_kind;
};

// Per-render-target parameters (one per job in a batch).
struct TargetParams {
    uint64_t color_addr;
    uint64_t ds_addr;
    uint32_t width, height;        // attachment extent
    uint32_t tiles_x, tiles_y;
    uint32_t tile_base;            // first global tile index of this target
    int32_t color_load, ds_load;   // CPV_LOAD_OP_*
    int32_t has_color, has_ds;
    int32_t color_bytes, ds_bytes; // bytes per pixel
    uint32_t clear_color[4];       // VkClearColorValue bits (float32 / uint32 / int32 by format)
    uint32_t clear_ds[2];          // VkClearDepthStencilValue: depth bits, stencil
    FormatDesc color_fmt, ds_fmt;
};

#if defined(__HIPCC__) || defined(__HIPCC_RTC__)
#define CPV_HOST_DEVICE __host__ __device__
#else
#define CPV_HOST_DEVICE
#endif

// The fixed formats: a colour attachment of four 8-bit UNORM channels in 4
// bytes (any channel order) and a D32_SFLOAT depth attachment, no stencil
// test.  Setting these fields to their values leaves such a state unchanged,
// which the host checks before it picks the fixed-format raster variant; in
// that variant the kernel sets them again, so the compiler sees constants
// and folds the generic format code (same arithmetic).
CPV_HOST_DEVICE inline void fold_fixed_formats(FormatDesc& c, FormatDesc& d)
{
    c.base = kBaseUNorm;
    c.element = 1;
    c.total = 4;
    c.depth_off = c.stencil_off = -1;
    c.depth_kind = 0;
    d.vk_format = 126u;   // VK_FORMAT_D32_SFLOAT
    d.base = kBaseSFloat;
    d.element = 4;
    d.total = 4;
    d.off[0] = d.off[1] = d.off[2] = d.off[3] = -1;
    d.depth_off = 0;
    d.stencil_off = -1;
    d.depth_kind = 32;
    d.pad = 0;
}
CPV_HOST_DEVICE inline void fold_fixed_formats(FragmentState& s)
{
    fold_fixed_formats(s.color, s.ds);
    s.has_color = s.has_ds = 1;
    s.stencil_test = 0;
}
CPV_HOST_DEVICE inline void fold_fixed_formats(TargetParams& t)
{
    fold_fixed_formats(t.color_fmt, t.ds_fmt);
    t.has_color = t.has_ds = 1;
    t.color_bytes = t.ds_bytes = 4;
}

// Per-draw parameters.  Vertex-output slots and primitive indices of all
// draws of one launch live in shared arrays; a draw owns the ranges
// [vout_base, vout_base + vout_count) and [prim_base, prim_base + prim_count).
struct DrawDesc {
    uint64_t env;              // device address of the draw's DrawEnv
    uint64_t index_addr;       // index buffer (indexed draws)
    uint32_t target;           // TargetParams index
    uint32_t fstate;           // FragmentState index
    uint32_t vout_base, vout_count;
    uint32_t prim_base, prim_count;
    uint32_t vblock_base, sblock_base;   // first block of this draw in the vertex / setup grids
    uint32_t topology, front_face, cull_mode;
    uint32_t vmode;            // vertex slot mode, see VertexSlotMode
    uint32_t first;            // firstVertex or firstIndex
    int32_t vertex_offset;
    uint32_t index_size;
    uint32_t instance_id;
    float vp_w, vp_h, vp_min_depth, vp_max_depth;
    float half_x, half_y;      // (1 / width) * 0.5, (1 / height) * 0.5 (ProcessTriangles :1524)
    int32_t limit_x, limit_y;  // min(viewport extent, attachment extent)
    float line_width;          // RasterizationState.LineWidth
    int32_t psize;             // varying component holding gl_PointSize, -1 when the shader writes none
};

// How the slot of an assembled vertex is found.
enum VertexSlotMode : uint32_t {
    kSlotSequential = 0,   // non-indexed: slot i shades vertex firstVertex + i
    kSlotVertexId = 1,     // indexed: slot = vertexOffset + index (every buffer vertex shaded once)
    kSlotPerIndex = 2,     // indexed: slot i shades index i (the reference's per-index shading)
};

}  // namespace cpv
