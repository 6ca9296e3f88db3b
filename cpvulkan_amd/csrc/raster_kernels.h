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
    const uint32_t* order;          // queue order of tiles (nullptr: index order)
    uint32_t* queue;                // work counter for the persistent schedule
    unsigned long long* fragments;  // covered (primitive, pixel) pairs
    float* trace_cycles;            // per tile shader cycles (nullptr: not recorded)
    uint32_t vout_stride;
    uint32_t tiles;                 // total tiles over all targets
    uint32_t target_count;
    uint32_t persistent;            // 0: one workgroup per tile; 1: work queue
};

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

__device__ __forceinline__ uint32_t find_target(const TargetParams* t, uint32_t n, uint32_t tile)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (t[mid].tile_base <= tile) lo = mid; else hi = mid;
    }
    return lo;
}

template <class FS, int TILE, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void raster_kernel(RasterArgs a)
{
    constexpr int NT = 64 * WAVES;
    constexpr int BW = TILE / 8;               // 8x8 blocks per tile row
    constexpr int BLOCKS = BW * BW;
    constexpr int BPW = BLOCKS / WAVES;        // blocks (pixels) per lane
    static_assert(BLOCKS % WAVES == 0, "tile blocks must divide among the waves");

    // The record batch and the framebuffer staging area are never live at
    // the same time, so they share LDS.
    constexpr int kRecWords = NT * 32;
    constexpr int kStageWords = TILE * TILE * 4;     // up to 16 bytes per pixel
    __shared__ uint32_t s_lds[kRecWords > kStageWords ? kRecWords : kStageWords];
    __shared__ uint32_t s_item;
    TriSetup* s_rec = (TriSetup*)s_lds;
    uint32_t* s_stage = s_lds;

    const int tid = threadIdx.x;
    const int wave = tid >> 6;
    const int lane = tid & 63;
    uint32_t item = a.persistent ? 0u : blockIdx.x;

    for (;;) {
        if (a.persistent) {
            __syncthreads();
            if (tid == 0) s_item = atomicAdd(a.queue, 1u);
            __syncthreads();
            item = s_item;
        }
        if (item >= a.tiles) break;
        const uint32_t tile = a.order ? a.order[item] : item;
        const uint64_t t_begin = clock64();

        const uint32_t ti = find_target(a.targets, a.target_count, tile);
        const TargetParams& tp = a.targets[ti];
        const uint32_t local = tile - tp.tile_base;
        const int tile_x0 = (int)(local % tp.tiles_x) * TILE;
        const int tile_y0 = (int)(local / tp.tiles_x) * TILE;
        const int cb = tp.color_bytes, db = tp.ds_bytes;

        // ---- pixel state: load (LOAD) or clear, staged through LDS
        // loadOp CLEAR: the clear value packed as ClearImage -> SetPixel does
        // (CPVulkan/CommandBuffer.Draw.cpp:117-149, ImageSampler.cpp:723-774)
        PixelState clear;
        clear.color[0] = clear.color[1] = clear.color[2] = clear.color[3] = 0u;
        clear.ds0 = clear.ds1 = 0u;
        if (tp.has_color && tp.color_load == 1) {
            if (tp.color_fmt.base == kBaseUInt) {
                pack_color_u32(tp.color_fmt, tp.clear_color, clear.color);
            } else if (tp.color_fmt.base == kBaseSInt) {
                const int32_t ci[4] = {(int32_t)tp.clear_color[0], (int32_t)tp.clear_color[1], (int32_t)tp.clear_color[2], (int32_t)tp.clear_color[3]};
                pack_color_i32(tp.color_fmt, ci, clear.color);
            } else {
                const float cf[4] = {__uint_as_float(tp.clear_color[0]), __uint_as_float(tp.clear_color[1]),
                                     __uint_as_float(tp.clear_color[2]), __uint_as_float(tp.clear_color[3])};
                pack_color_f32(tp.color_fmt, cf, clear.color);
            }
        }
        if (tp.has_ds && tp.ds_load == 1) pack_depth_stencil(tp.ds_fmt, __uint_as_float(tp.clear_ds[0]), tp.clear_ds[1], clear.ds0, clear.ds1);
        PixelState px[BPW];
        int pxx[BPW], pxy[BPW];
#pragma unroll
        for (int k = 0; k < BPW; k++) {
            const int blk = wave + k * WAVES;
            pxx[k] = tile_x0 + (blk % BW) * 8 + (lane & 7);
            pxy[k] = tile_y0 + (blk / BW) * 8 + (lane >> 3);
            px[k] = clear;
        }
        const int valid_w = min(TILE, (int)tp.width - tile_x0);
        const int valid_h = min(TILE, (int)tp.height - tile_y0);
        // One attachment at a time: rows are copied HBM -> LDS with dword
        // accesses along the row, then each lane picks its pixel's bytes.
        for (int att = 0; att < 2; att++) {
            const bool is_color = att == 0;
            const int bytes = is_color ? cb : db;
            const bool present = is_color ? tp.has_color : tp.has_ds;
            const int load = is_color ? tp.color_load : tp.ds_load;
            if (!present || load != 0 || bytes == 0) continue;   // only LOAD reads memory
            const uint8_t* base = (const uint8_t*)(is_color ? tp.color_addr : tp.ds_addr);
            const int row_bytes = valid_w * bytes;
            const int row_words = (row_bytes + 3) >> 2;
            __syncthreads();
            for (int i = tid; i < valid_h * row_words; i += NT) {
                const int r = i / row_words, wd = i % row_words;
                const uint8_t* src = base + ((uint64_t)(tile_y0 + r) * tp.width + tile_x0) * bytes + wd * 4;
                uint32_t v;
                if (wd * 4 + 4 <= row_bytes) {
                    v = *(const uint32_t*)src;
                } else {
                    v = 0;
                    for (int b = 0; wd * 4 + b < row_bytes; b++) v |= (uint32_t)src[b] << (b * 8);
                }
                s_stage[r * (TILE * 4) + wd] = v;
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < BPW; k++) {
                const int lx = pxx[k] - tile_x0, ly = pxy[k] - tile_y0;
                if (lx >= valid_w || ly >= valid_h) continue;
                uint32_t w4[4] = {0u, 0u, 0u, 0u};
                const uint8_t* row = (const uint8_t*)&s_stage[ly * (TILE * 4)];
                for (int b = 0; b < bytes; b++) w4[b >> 2] |= (uint32_t)row[lx * bytes + b] << ((b & 3) * 8);
                if (is_color) { px[k].color[0] = w4[0]; px[k].color[1] = w4[1]; px[k].color[2] = w4[2]; px[k].color[3] = w4[3]; }
                else { px[k].ds0 = w4[0]; px[k].ds1 = w4[1]; }
            }
        }

        // ---- walk the tile's primitives in submission order
        const uint32_t lbeg = a.tile_start[tile], lend = a.tile_start[tile + 1];
        uint32_t cur_draw = 0xFFFFFFFFu;
        float xf[BPW], yf[BPW];
        float vp_min = 0.0f, vp_max = 1.0f;
        bool vp_identity = true;
        const FragmentState* fst = nullptr;
        const DrawEnv* env = nullptr;
        unsigned long long nfrag = 0;
        bool early = false;   // conservative depth rejection allowed for this draw
        int32_t depth_op = 7;
        int32_t depth_kind = 0;

        // Deferred colour work: a fragment that passed the depth/stencil
        // tests has its depth/stencil written at once, and its colour
        // (interpolation, fragment shader, blend, pack) queued.  Queues drain
        // in order, so every pixel still sees its colour writes in
        // primitive order; draining for all lanes at once keeps the shading
        // path from running once per primitive whenever any lane passes.
        constexpr int Q = FS::kMayDiscard ? 1 : 4;
        uint32_t q_prim[Q];
        float q_w0[Q], q_w1[Q], q_w2[Q];
        int qn = 0;

        auto shade = [&](uint32_t gprim, int k, float w0, float w1, float w2) {
            const TriSetup& g = a.recs[gprim];
            const DrawDesc& d = a.draws[g.draw];
            const FragmentState& fs = a.fstates[d.fstate];
            const DrawEnv& de = *(const DrawEnv*)d.env;
            const float w[3] = {w0, w1, w2};
            const float pw[3] = {g.w[0], g.w[1], g.w[2]};
            const float* q0 = a.vout + (uint64_t)g.v[0] * a.vout_stride + 4;
            const float* q1 = a.vout + (uint64_t)g.v[1] * a.vout_stride + 4;
            const float* q2 = a.vout + (uint64_t)g.v[2] * a.vout_stride + 4;
            const float* qp = a.vout + (uint64_t)g.provoking * a.vout_stride + 4;
            float in[FS::kInputs > 0 ? FS::kInputs : 1];
            // x / 1.0f == x exactly, so clip w == 1 drops the divisions
            const bool unit_w = pw[0] == 1.0f && pw[1] == 1.0f && pw[2] == 1.0f;
#pragma unroll
            for (int c = 0; c < FS::kInputs; c++) {
                const int mode = FS::interp(c);
                in[c] = (mode == 0 && unit_w) ? interpolate_unit_w(w, q0[c], q1[c], q2[c])
                                              : interpolate(mode, w, pw, q0[c], q1[c], q2[c], qp[c]);
            }
            const float depth = g.z[0] * w0 + g.z[1] * w1 + g.z[2] * w2;
            const float fc[4] = {fs.origin_upper_left ? (float)pxx[k] : d.vp_w - (float)pxx[k] - 1.0f, (float)pxy[k], depth, 1.0f};
            float colour[4];
            FS::run(de, in, fc, (g.flags & 1u) != 0, colour);
#pragma unroll
            for (int kk = 0; kk < BPW; kk++)
                if (kk == k) color_write(fs, px[kk], colour);
        };
        auto drain = [&]() {
#pragma unroll
            for (int e = 0; e < Q; e++) {
                if (!__any(e < qn)) break;
                if (e < qn) shade(q_prim[e] & 0x0FFFFFFFu, (int)(q_prim[e] >> 28), q_w0[e], q_w1[e], q_w2[e]);
            }
            qn = 0;
        };

        // wave bounds (union of its blocks) for the per-wave primitive filter
        int wx0 = 1 << 30, wy0 = 1 << 30, wx1 = -(1 << 30), wy1 = -(1 << 30);
#pragma unroll
        for (int k = 0; k < BPW; k++) {
            const int blk = wave + k * WAVES;
            const int bx0 = tile_x0 + (blk % BW) * 8, by0 = tile_y0 + (blk / BW) * 8;
            wx0 = min(wx0, bx0); wy0 = min(wy0, by0); wx1 = max(wx1, bx0 + 8); wy1 = max(wy1, by0 + 8);
        }

        for (uint32_t base = lbeg; base < lend; base += NT) {
            const int n = (int)min((uint32_t)NT, lend - base);
            __syncthreads();
            if (tid < n) {
                const uint4* src = (const uint4*)&a.recs[a.tile_prims[base + tid]];
                uint4* dst = (uint4*)&s_rec[tid];
#pragma unroll
                for (int q = 0; q < 8; q++) dst[q] = src[q];
            }
            __syncthreads();
            for (int c0 = 0; c0 < n; c0 += 64) {
                // each lane tests one primitive of the chunk against this wave's area
                bool ov = false;
                if (c0 + lane < n) {
                    const TriSetup& t = s_rec[c0 + lane];
                    ov = !(t.x1 <= wx0 || t.x0 >= wx1 || t.y1 <= wy0 || t.y0 >= wy1);
                }
                uint64_t mask = __ballot(ov);
                while (mask) {
                    const int j = c0 + (int)__builtin_ctzll(mask);
                    mask &= mask - 1;
                    const TriSetup& r = s_rec[j];
                    const uint32_t rdraw = r.draw;
                    if (rdraw != cur_draw) {
                        cur_draw = rdraw;
                        const DrawDesc& d = a.draws[cur_draw];
                        // pixel centres (ProcessTriangles :1571-1580)
#pragma unroll
                        for (int k = 0; k < BPW; k++) {
                            xf[k] = ((float)pxx[k] / d.vp_w + d.half_x) * 2.0f - 1.0f;
                            yf[k] = ((float)pxy[k] / d.vp_h + d.half_y) * 2.0f - 1.0f;
                        }
                        vp_min = d.vp_min_depth; vp_max = d.vp_max_depth;
                        vp_identity = vp_min == 0.0f && vp_max == 1.0f;
                        fst = &a.fstates[d.fstate];
                        env = (const DrawEnv*)d.env;
                        // Rejecting on the depth test alone is exact when nothing
                        // else observes the fragment: no depth-bounds or stencil
                        // test (and so no stencil write).
                        early = fst->depth_test && fst->has_ds && fst->ds.depth_off >= 0 && fst->ds.vk_format != 127u &&
                                !fst->depth_bounds_test && !fst->stencil_test;
                        depth_op = fst->depth_compare;
                        depth_kind = fst->ds.depth_kind;
                    }
                    const int rx0 = r.x0, ry0 = r.y0, rx1 = r.x1, ry1 = r.y1;
                    const uint32_t gprim = a.tile_prims[base + j];
#pragma unroll
                    for (int k = 0; k < BPW; k++) {
                        // every lane is active here; a lane queues at most one
                        // entry per iteration, so a free slot always exists
                        if (!FS::kMayDiscard && __any(qn == Q)) drain();
                        const int blk = wave + k * WAVES;
                        const int bx0 = tile_x0 + (blk % BW) * 8, by0 = tile_y0 + (blk / BW) * 8;
                        if (BPW > 1 && (rx1 <= bx0 || rx0 >= bx0 + 8 || ry1 <= by0 || ry0 >= by0 + 8)) continue;   // wave-uniform
                        const int x = pxx[k], y = pxy[k];
                        if (x < rx0 || x >= rx1 || y < ry0 || y >= ry1) continue;
                        // GetFragmentInput (:874-903); NaN edge values are not < 0 and pass
                        const float e0 = (xf[k] - r.ax[0]) * r.dy[0] - (yf[k] - r.ay[0]) * r.dx[0];
                        const float e1 = (xf[k] - r.ax[1]) * r.dy[1] - (yf[k] - r.ay[1]) * r.dx[1];
                        const float e2 = (xf[k] - r.ax[2]) * r.dy[2] - (yf[k] - r.ay[2]) * r.dx[2];
                        if (e0 < 0 || e1 < 0 || e2 < 0) continue;
                        nfrag++;
                        if (early) {
                            const float cur = depth_kind == 32 ? __uint_as_float(px[k].ds0) : unpack_depth(fst->ds, px[k].ds0);
                            if (vp_identity ? depth_surely_fails_unit(depth_op, cur, e0, e1, e2, r.area, r.z[0], r.z[1], r.z[2])
                                            : depth_surely_fails(depth_op, cur, e0, e1, e2, r.area, r.z[0], r.z[1], r.z[2], vp_min, vp_max))
                                continue;
                        }
                        const float w0 = e0 / r.area, w1 = e1 / r.area, w2 = e2 / r.area;
                        const float depth = r.z[0] * w0 + r.z[1] * w1 + r.z[2] * w2;
                        // DrawPixel (:1300-1313)
                        const float dt = (vp_max - vp_min) * depth + vp_min;
                        const bool front = (r.flags & 1u) != 0;
                        if (FS::kMayDiscard) {
                            // the shader decides first whether anything is written
                            const float wv[3] = {w0, w1, w2};
                            const float pw[3] = {r.w[0], r.w[1], r.w[2]};
                            const float* q0 = a.vout + (uint64_t)r.v[0] * a.vout_stride + 4;
                            const float* q1 = a.vout + (uint64_t)r.v[1] * a.vout_stride + 4;
                            const float* q2 = a.vout + (uint64_t)r.v[2] * a.vout_stride + 4;
                            const float* qp = a.vout + (uint64_t)r.provoking * a.vout_stride + 4;
                            float in[FS::kInputs > 0 ? FS::kInputs : 1];
#pragma unroll
                            for (int c = 0; c < FS::kInputs; c++) in[c] = interpolate(FS::interp(c), wv, pw, q0[c], q1[c], q2[c], qp[c]);
                            const float fc[4] = {fst->origin_upper_left ? (float)x : a.draws[cur_draw].vp_w - (float)x - 1.0f, (float)y, depth, 1.0f};
                            float colour[4];
                            if (FS::run(*env, in, fc, front, colour)) continue;
                            if (depth_stencil_ops(*fst, px[k], dt, front) && fst->has_color) color_write(*fst, px[k], colour);
                        } else if (depth_stencil_ops(*fst, px[k], dt, front) && fst->has_color) {
                            const uint32_t tag = gprim | ((uint32_t)k << 28);
#pragma unroll
                            for (int e = 0; e < Q; e++) {
                                if (e == qn) { q_prim[e] = tag; q_w0[e] = w0; q_w1[e] = w1; q_w2[e] = w2; }
                            }
                            qn++;
                        }
                    }
                }
            }
        }
        if (!FS::kMayDiscard) drain();
        (void)env;

        // ---- write back through LDS, rows stored with dword accesses
        for (int att = 0; att < 2; att++) {
            const bool is_color = att == 0;
            const int bytes = is_color ? cb : db;
            const bool present = is_color ? tp.has_color : tp.has_ds;
            if (!present || bytes == 0) continue;
            uint8_t* base = (uint8_t*)(is_color ? tp.color_addr : tp.ds_addr);
            __syncthreads();
#pragma unroll
            for (int k = 0; k < BPW; k++) {
                const int lx = pxx[k] - tile_x0, ly = pxy[k] - tile_y0;
                if (lx >= valid_w || ly >= valid_h) continue;
                const uint32_t w4[4] = {is_color ? px[k].color[0] : px[k].ds0, is_color ? px[k].color[1] : px[k].ds1,
                                        is_color ? px[k].color[2] : 0u, is_color ? px[k].color[3] : 0u};
                uint8_t* row = (uint8_t*)&s_stage[ly * (TILE * 4)];
                for (int b = 0; b < bytes; b++) row[lx * bytes + b] = (uint8_t)(w4[b >> 2] >> ((b & 3) * 8));
            }
            __syncthreads();
            const int row_bytes = valid_w * bytes;
            const int row_words = (row_bytes + 3) >> 2;
            for (int i = tid; i < valid_h * row_words; i += NT) {
                const int r = i / row_words, wd = i % row_words;
                uint8_t* dst = base + ((uint64_t)(tile_y0 + r) * tp.width + tile_x0) * bytes + wd * 4;
                const uint32_t v = s_stage[r * (TILE * 4) + wd];
                if (wd * 4 + 4 <= row_bytes) {
                    *(uint32_t*)dst = v;
                } else {
                    for (int b = 0; wd * 4 + b < row_bytes; b++) dst[b] = (uint8_t)(v >> (b * 8));
                }
            }
        }

        // ---- statistics
        for (int off = 32; off > 0; off >>= 1) nfrag += __shfl_down(nfrag, off, 64);
        if (lane == 0 && nfrag) atomicAdd(a.fragments, nfrag);
        if (a.trace_cycles && tid == 0) a.trace_cycles[tile] = (float)(clock64() - t_begin);

        if (!a.persistent) break;
    }
}

}  // namespace cpv
