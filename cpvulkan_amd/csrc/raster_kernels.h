// Program-dependent kernels of the draw path: the vertex kernel (runs the
// vertex shader) and the fine raster / fragment kernel (coverage, depth,
// stencil, fragment shader, blend, pack), both templated on the shader
// program.  The ahead-of-time library instantiates them for the built-in
// programs; programs translated from SPIR-V instantiate them through hiprtc
// from this same header.
//
// Reference semantics (paths relative to the reference repository):
//   vertex kernel   CPVulkan/CommandBuffer.Draw.cpp:675-814,
//                   LLVMRuntime/PipelineCompiler.cpp:513-582, 898-981
//   raster kernel   CPVulkan/CommandBuffer.Draw.cpp:816-954, 1300-1313, 1510-1594,
//                   LLVMRuntime/PipelineCompiler.cpp:1020-1085
#pragma once
#pragma clang fp contract(off)

#include "raster_device.h"

namespace cpv {

// ------------------------------------------------------------ vertex kernel

struct VertexArgs {
    const DrawDesc* draws;
    const uint32_t* block_draw;    // vertex-grid block -> draw index
    float* vout;                   // slots x stride floats
    uint32_t stride;               // 4 + varyings
    uint32_t pad;
};

// Vertex id of slot `slot` of draw d (ProcessInputAssembler(Indexed),
// CommandBuffer.Draw.cpp:675-760).
__device__ __forceinline__ uint32_t read_index(const DrawDesc& d, uint32_t at)
{
    if (d.index_size == 1) return ((const uint8_t*)d.index_addr)[at];
    if (d.index_size == 2) return ((const uint16_t*)d.index_addr)[at];
    return ((const uint32_t*)d.index_addr)[at];
}

template <class VS>
__global__ __launch_bounds__(256) void vertex_kernel(VertexArgs a)
{
    const uint32_t di = a.block_draw[blockIdx.x];
    const DrawDesc& d = a.draws[di];
    const uint32_t slot = (blockIdx.x - d.vblock_base) * 256u + threadIdx.x;
    if (slot >= d.vout_count) return;
    uint32_t vid;
    if (d.vmode == kSlotSequential) vid = d.first + slot;
    else if (d.vmode == kSlotVertexId) vid = slot;
    else vid = (uint32_t)((int64_t)d.vertex_offset + (int64_t)read_index(d, d.first + slot));
    const DrawEnv& env = *(const DrawEnv*)d.env;
    float pos[4];
    float out[VS::kOutputs > 0 ? VS::kOutputs : 1];
    VS::run(env, vid, d.instance_id, pos, out);
    float* dst = a.vout + (uint64_t)(d.vout_base + slot) * a.stride;
    dst[0] = pos[0]; dst[1] = pos[1]; dst[2] = pos[2]; dst[3] = pos[3];
#pragma unroll
    for (int k = 0; k < VS::kOutputs; k++) dst[4 + k] = out[k];
}

// ------------------------------------------------------------ raster kernel

struct RasterArgs {
    const TriSetup* recs;
    const uint32_t* tile_prims;     // primitive ids grouped by tile, submission order within a tile
    const uint32_t* tile_start;     // per global tile, size tiles + 1
    const TargetParams* targets;
    const DrawDesc* draws;
    const FragmentState* fstates;
    const float* vout;
    const LineExt* line_ext;        // per-primitive line extension (launches with lines)
    const uint32_t* order;          // queue order of tiles (nullptr: index order)
    uint32_t* queue;                // kRasterQueues work counters, kRasterQueueStride words apart
    unsigned long long* fragments;  // covered (primitive, pixel) pairs
    float* trace_cycles;            // shader cycles per (tile, wave slot) (nullptr: not recorded)
    const uint32_t* overflow;       // set when the launch's pairs outgrew their arrays: then nothing is drawn
    uint32_t vout_stride;
    uint32_t tiles;                 // total tiles over all targets
    uint32_t target_count;
    uint32_t persistent;            // 0: one workgroup per tile; 1: work queue
    int32_t hiz_op;                 // depth compare op shared by every draw when depth culling is exact, else -1
    int32_t hiz_depth_kind;         // 16, 24 or 32
    int32_t uniform_fstate;         // the fragment state every draw of the launch has, -1 when they differ
};

// True when a depth in [lo, hi] fails `op` against every stored depth in
// [bmin, bmax] (NaN stored depths fail every ordered compare anyway).
__device__ __forceinline__ bool hiz_rejects(int32_t op, float lo, float hi, float bmin, float bmax)
{
    return interval_fails(op, lo, hi, bmin, bmax);
}

__device__ __forceinline__ float stored_depth(int32_t kind, uint32_t d0)
{
    if (kind == 32) return __uint_as_float(d0);
    if (kind == 16) return (float)((double)(d0 & 0xFFFFu) * (1.0 / 65535.0));
    return (float)((double)(d0 & 0x00FFFFFFu) * (1.0 / 16777215.0));
}

// SetDatum (CommandBuffer.Draw.cpp:816-837) for the three interpolation modes.
__device__ __forceinline__ float interpolate(int mode, const float w[3], const float pw[3], float v0, float v1, float v2, float flat)
{
    if (mode == 2) return flat;
    if (mode == 1) {
        float r = 0.0f;
        r += w[0] * v0;
        r += w[1] * v1;
        r += w[2] * v2;
        return r;
    }
    float num = 0.0f, den = 0.0f;
    num += w[0] * v0 / pw[0];
    den += w[0] / pw[0];
    num += w[1] * v1 / pw[1];
    den += w[1] / pw[1];
    num += w[2] * v2 / pw[2];
    den += w[2] / pw[2];
    return num / den;
}

// SetDatum<..., 2> of a line (CommandBuffer.Draw.cpp:1450-1485): two terms.
__device__ __forceinline__ float interpolate_line(int mode, const float w[3], const float pw[3], float v0, float v1, float flat)
{
    if (mode == 2) return flat;
    if (mode == 1) {
        float r = 0.0f;
        r += w[0] * v0;
        r += w[1] * v1;
        return r;
    }
    float num = 0.0f, den = 0.0f;
    num += w[0] * v0 / pw[0];
    den += w[0] / pw[0];
    num += w[1] * v1 / pw[1];
    den += w[1] / pw[1];
    return num / den;
}

__device__ __forceinline__ LineExt load_line_ext(const LineExt* ext, uint32_t gprim)
{
    const uint4* p = (const uint4*)(ext + gprim);
    const uint4 a = p[0], b = p[1];
    LineExt L;
    L.ax3 = __uint_as_float(a.x); L.ay3 = __uint_as_float(a.y); L.dy3 = __uint_as_float(a.z); L.dx3 = __uint_as_float(a.w);
    L.p0x = __uint_as_float(b.x); L.p0y = __uint_as_float(b.y); L.ddx = __uint_as_float(b.z); L.ddy = __uint_as_float(b.w);
    return L;
}

// Perspective SetDatum with all three clip w equal to 1: (w*v)/1 and w/1 are
// exact, so only the final division remains.
__device__ __forceinline__ float interpolate_unit_w(const float w[3], float v0, float v1, float v2)
{
    float num = 0.0f, den = 0.0f;
    num += w[0] * v0;
    den += w[0];
    num += w[1] * v1;
    den += w[1];
    num += w[2] * v2;
    den += w[2];
    return num / den;
}

// Colour work of one fragment that passed the depth/stencil tests:
// interpolation (SetDatum), the fragment shader and the colour write with
// blend (PipelineCompiler.cpp:1415-1727).  The raster kernel calls it from
// one place (the colour queue's drain), so it is inlined: an out-of-line call
// would pass the pixel state through scratch memory.
// color_write with the fixed formats folded in when FIXED.
template <bool FIXED>
__device__ __forceinline__ void color_write_state(const FragmentState& st, PixelState& px, const float colour[4])
{
    if constexpr (FIXED) {
        FragmentState s = st;
        fold_fixed_formats(s);
        color_write(s, px, colour);
    } else {
        color_write(st, px, colour);
    }
}

// Fragment-shader inputs of one covered pixel of primitive g with weights
// (w0, w1, w2): SetDatum over the three vertices of a triangle, over the two
// of a line, and a point's vertex outputs as they are (ProcessPoints copies
// them, :1364-1368).
template <class FS, bool ANYPRIM, class Rec>
__device__ __forceinline__ void fragment_inputs(const Rec& g, const float* vout, uint32_t stride, float w0, float w1, float w2,
                                                float* in)
{
    const float w[3] = {w0, w1, w2};
    const float pw[3] = {g.w[0], g.w[1], g.w[2]};
    const float* q0 = vout + (uint64_t)g.v[0] * stride + 4;
    const float* q1 = vout + (uint64_t)g.v[1] * stride + 4;
    const float* q2 = vout + (uint64_t)g.v[2] * stride + 4;
    const float* qp = vout + (uint64_t)g.provoking * stride + 4;
    if (ANYPRIM && (g.flags & kPrimPoint)) {
#pragma unroll
        for (int c = 0; c < FS::kInputs; c++) in[c] = q0[c];
        return;
    }
    if (ANYPRIM && (g.flags & kPrimLine)) {
#pragma unroll
        for (int c = 0; c < FS::kInputs; c++) in[c] = interpolate_line(FS::interp(c), w, pw, q0[c], q1[c], qp[c]);
        return;
    }
    // x / 1.0f == x exactly, so clip w == 1 drops the divisions.  One
    // branch around all inputs: with a select per input the compiler
    // computes the perspective denominator (three divisions) for every
    // fragment.
    const bool unit_w = pw[0] == 1.0f && pw[1] == 1.0f && pw[2] == 1.0f;
    if (unit_w) {
#pragma unroll
        for (int c = 0; c < FS::kInputs; c++) {
            const int mode = FS::interp(c);
            in[c] = mode == 0 ? interpolate_unit_w(w, q0[c], q1[c], q2[c]) : interpolate(mode, w, pw, q0[c], q1[c], q2[c], qp[c]);
        }
    } else {
#pragma unroll
        for (int c = 0; c < FS::kInputs; c++) in[c] = interpolate(FS::interp(c), w, pw, q0[c], q1[c], q2[c], qp[c]);
    }
}

// The shader and the colour write of one queued fragment (lanes hold
// fragments of different primitives).  When every draw of the launch has
// the same fragment state (uniform_fstate >= 0) it is read once for the
// wave, so the blend and format state are scalar and its branches uniform.
template <class FS, bool ANYPRIM = false, bool FIXED = false>
__device__ __forceinline__ PixelState shade_fragment(const TriSetup* recs, const DrawDesc* draws, const FragmentState* fstates,
                                                  int32_t uniform_fstate, const float* vout, uint32_t stride, uint32_t gprim,
                                                  float w0, float w1, float w2, int x, int y, PixelState ps)
{
    const TriSetup& g = recs[gprim];
    const DrawDesc& d = draws[g.draw];
    const DrawEnv& de = *(const DrawEnv*)d.env;
    float in[FS::kInputs > 0 ? FS::kInputs : 1];
    fragment_inputs<FS, ANYPRIM>(g, vout, stride, w0, w1, w2, in);
    // a point's depth is its vertex's; a line's is p0.z * t + p1.z * (1 - t)
    // with t = w1 (ProcessLines :1494)
    const float depth = (ANYPRIM && (g.flags & kPrimPoint)) ? g.z[0]
                        : (ANYPRIM && (g.flags & kPrimLine)) ? g.z[0] * w1 + g.z[1] * (1.0f - w1)
                                                               : g.z[0] * w0 + g.z[1] * w1 + g.z[2] * w2;
    float colour[4];
    // one copy of the shader body (it reads no fragment state); only the
    // colour write is taken twice
    const int32_t upper_left = uniform_fstate >= 0 ? fstates[uniform_fstate].origin_upper_left : fstates[d.fstate].origin_upper_left;
    const float fc[4] = {upper_left ? (float)x : d.vp_w - (float)x - 1.0f, (float)y, depth, 1.0f};
    FS::run(de, in, fc, (g.flags & 1u) != 0, colour);
    if (uniform_fstate >= 0) color_write_state<FIXED>(fstates[uniform_fstate], ps, colour);
    else color_write_state<FIXED>(fstates[d.fstate], ps, colour);
    return ps;
}

struct DepthStencilResult {
    PixelState px;
    uint32_t write;
};

// depth_stencil_ops with the pixel state passed and returned by value (one
// call site per raster kernel instantiation, inlined).  OP >= 0: the draw
// has the depth test with compare op OP and no stencil or depth-bounds
// test, folded into the state as constants.
template <bool FIXED = false, int OP = -1>
__device__ __forceinline__ DepthStencilResult depth_stencil_call(const FragmentState* st, PixelState px, float depth, uint32_t front)
{
    DepthStencilResult r;
    r.px = px;
    if constexpr (FIXED) {
        FragmentState s = *st;
        fold_fixed_formats(s);
        if constexpr (OP >= 0) {
            s.depth_test = 1;
            s.depth_compare = OP;
            s.stencil_test = 0;
            s.depth_bounds_test = 0;
        }
        r.write = depth_stencil_ops(s, r.px, depth, front != 0) ? 1u : 0u;
    } else {
        r.write = depth_stencil_ops(*st, r.px, depth, front != 0) ? 1u : 0u;
    }
    return r;
}


__device__ __forceinline__ uint32_t find_target(const TargetParams* t, uint32_t n, uint32_t tile)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (t[mid].tile_base <= tile) lo = mid; else hi = mid;
    }
    return lo;
}

}  // namespace cpv

#include "raster_tile.h"
