// Fine raster / fragment kernel.
//
// Work unit: one wavefront owns BPW 8x8 pixel blocks of one tile and walks
// that tile's primitive list in submission order.  Wavefronts never
// synchronise with each other: the framebuffer pixels live in registers for
// the whole tile (loaded or cleared at the start, stored at the end), the
// primitive records are read with scalar loads (one record serves all 64
// lanes), and the per-chunk filter reads bounding boxes with vector gathers.
//
// Per primitive and pixel the reference's order of operations is kept
// (CPVulkan/CommandBuffer.Draw.cpp:874-954, 1300-1313 and
// LLVMRuntime/PipelineCompiler.cpp:1020-1085); the shortcuts taken are
// exact: hierarchical depth culling and the per-fragment early depth
// rejection use proven error bounds, and the colour work of passing
// fragments is queued and drained in order per pixel.
#pragma once
#pragma clang fp contract(off)

namespace cpv {

// A wave-uniform record through the constant address space, which lets the
// compiler use scalar loads (the records are read-only during the kernel).
__device__ __forceinline__ void load_record_uniform(const TriSetup* recs, uint32_t gp, TriSetup& r)
{
    uint32_t* dst = (uint32_t*)&r;
#if defined(__HIP_DEVICE_COMPILE__)
    typedef const __attribute__((address_space(4))) uint32_t* ScalarU32Ptr;
    ScalarU32Ptr p = (ScalarU32Ptr)(recs + gp);
#else
    const uint32_t* p = (const uint32_t*)(recs + gp);
#endif
#pragma unroll
    for (int q = 0; q < 32; q++) dst[q] = p[q];
}

__device__ __forceinline__ void load_pixel(const uint8_t* p, int bytes, uint32_t w[4])
{
    w[0] = w[1] = w[2] = w[3] = 0u;
    if (bytes == 4) {
        w[0] = *(const uint32_t*)p;
    } else if (bytes == 2) {
        w[0] = *(const uint16_t*)p;
    } else if (bytes == 8) {
        const uint2 v = *(const uint2*)p;
        w[0] = v.x; w[1] = v.y;
    } else if (bytes == 16) {
        const uint4 v = *(const uint4*)p;
        w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
    } else {
        // unrolled over the largest pixel so w[] keeps constant indices (registers)
#pragma unroll
        for (int b = 0; b < 16; b++)
            if (b < bytes) w[b >> 2] |= (uint32_t)p[b] << ((b & 3) * 8);
    }
}

__device__ __forceinline__ void store_pixel(uint8_t* p, int bytes, const uint32_t w[4])
{
    if (bytes == 4) {
        *(uint32_t*)p = w[0];
    } else if (bytes == 2) {
        *(uint16_t*)p = (uint16_t)w[0];
    } else if (bytes == 8) {
        *(uint2*)p = make_uint2(w[0], w[1]);
    } else if (bytes == 16) {
        *(uint4*)p = make_uint4(w[0], w[1], w[2], w[3]);
    } else {
#pragma unroll
        for (int b = 0; b < 16; b++)
            if (b < bytes) p[b] = (uint8_t)(w[b >> 2] >> ((b & 3) * 8));
    }
}

// Resident wavefronts per SIMD the raster kernel is compiled for: the
// kernel is latency bound, and more resident waves hide more of it until
// the register budget (512 / waves VGPRs) forces spills.  A fragment shader
// may set kWavesPerEU; measured on MI355X: 6 for the colour shaders, 5 for
// the textured Lambert shader (the default).
template <class FS, class = void>
struct RasterWavesPerEU {
    static constexpr int value = 5;
};
template <class FS>
struct RasterWavesPerEU<FS, decltype((void)FS::kWavesPerEU)> {
    static constexpr int value = FS::kWavesPerEU;
};

// The target applies when a lane holds one 8x8 block's pixel (8x8 tiles,
// or 16x16 tiles with four waves); a lane holding several blocks' pixels
// needs the registers of the compiler's default budget.
template <class FS, int TILE, int WAVES>
constexpr int raster_waves_per_eu()
{
    return (TILE / 8) * (TILE / 8) / WAVES == 1 ? RasterWavesPerEU<FS>::value : 1;
}


// Orders one wave's LDS writes before its lanes read each other's words.
__device__ __forceinline__ void wave_lds_fence()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Staging area of a workgroup: the visit fields of a chunk's candidates,
// one region per wave, used with several blocks per lane (see GATHER).
template <class FS, int TILE, int WAVES>
struct RasterStage {
    static constexpr int kBlocksPerLane = (TILE / 8) * (TILE / 8) / WAVES;
    static constexpr bool kOn = !FS::kMayDiscard && kBlocksPerLane > 1;
    uint4 v[kOn ? WAVES : 1][kOn ? kVisitParts : 1][64];
};

// The kernel body (raster_kernel below).  OP >= 0 is the launch's depth
// compare op as a compile-time constant, for launches whose every draw has
// the depth test with that op and no stencil or depth-bounds test (the
// launches with hierarchical culling): the depth interval tests and the
// depth compare then fold to single comparisons.  OP = -1 reads the ops at
// run time.
template <class FS, int TILE, int WAVES, bool ANYPRIM, bool FIXED, int OP>
__device__ __forceinline__ void raster_body(const RasterArgs& a, RasterStage<FS, TILE, WAVES>& stage_area)
{
    constexpr int BW = TILE / 8;               // 8x8 blocks per tile row
    constexpr int BLOCKS = BW * BW;
    constexpr int BPW = BLOCKS / WAVES;        // blocks (pixels) per lane
    static_assert(BLOCKS % WAVES == 0, "tile blocks must divide among the waves");

    const int wave = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const uint32_t home = blockIdx.x % kRasterQueues;
    // Pairs that outgrew their arrays were dropped: the launch draws nothing
    // and leaves the attachments as they were, so that it can be run again
    // with larger arrays.
    if (*a.overflow) return;

    // With several blocks per lane, a chunk's candidate primitives have the
    // visit fields of their records (the first kVisitParts parts) loaded by
    // all their lanes at once straight into the wave's LDS region
    // (global_load_lds_dwordx4, which takes no registers), before the
    // chunk's depth test, and the survivor loop reads them from there: the
    // chunk pays one memory latency for the records of all its survivors.
    // With one block per lane each survivor's record is read with scalar
    // loads (measured on MI355X: the LDS staging made the 16x16 / four-wave
    // soup frame's raster 9% slower, 0.85 -> 0.93 ms, and the 16x16 / one-wave
    // and 32x32 / four-wave ones 2-4% faster).  A shader that may discard
    // reads the whole record per fragment and keeps the scalar loads too.
    // Every configuration loads the next chunk's pairs while it works on the
    // current one.
    constexpr bool GATHER = RasterStage<FS, TILE, WAVES>::kOn;
    constexpr bool STAGE = GATHER;
    // block-local depth bounds of triangles, where a lane holds one pixel
    // of an 8x8 block (lane = row * 8 + column)
    constexpr bool BLOCK_DEPTH = !ANYPRIM && !GATHER && BPW == 1;
    auto& stage = stage_area.v;
    // the hierarchical depth op (-1: no hierarchical culling)
    const int32_t hiz_op = OP >= 0 ? OP : a.hiz_op;
    uint32_t tried = 0;   // queues found empty, starting at home

    for (;;) {
        uint32_t item = blockIdx.x;
        int slot = wave;
        if (a.persistent) {
            // Unit u of queue q is wave slot u % WAVES of the queue's
            // (u / WAVES)-th tile, queue item (u / WAVES) * kRasterQueues + q.
            // Every wave ends after at most one failed take per queue.
            for (; tried < kRasterQueues; tried++) {
                const uint32_t q = (home + tried) % kRasterQueues;
                const uint32_t per_q = a.tiles > q ? (a.tiles - q + kRasterQueues - 1) / kRasterQueues : 0u;
                uint32_t u = 0;
                if (lane == 0) u = atomicAdd(a.queue + q * kRasterQueueStride, 1u);
                u = __builtin_amdgcn_readfirstlane(u);
                if (u < per_q * (uint32_t)WAVES) {
                    item = (u / WAVES) * kRasterQueues + q;
                    slot = (int)(u % WAVES);
                    break;
                }
            }
            if (tried == kRasterQueues) break;
        }
        if (item >= a.tiles) break;
        const uint32_t tile = a.order ? a.order[item] : item;
        const uint64_t t_begin = clock64();

        const uint32_t ti = find_target(a.targets, a.target_count, tile);
        const TargetParams& tp = a.targets[ti];
        const uint32_t local = tile - tp.tile_base;
        const int tile_x0 = (int)(local % tp.tiles_x) * TILE;
        const int tile_y0 = (int)(local / tp.tiles_x) * TILE;

        // ---- pixel state: loadOp CLEAR packs the clear value as ClearImage ->
        // SetPixel does (CommandBuffer.Draw.cpp:117-149, ImageSampler.cpp:723-774);
        // LOAD / DONT_CARE read the attachment.
        PixelState clear;
        clear.color[0] = clear.color[1] = clear.color[2] = clear.color[3] = 0u;
        clear.ds0 = clear.ds1 = 0u;
        if (tp.has_color && tp.color_load == 1) {
            if (tp.color_fmt.base == kBaseUInt) {
                pack_color_u32(tp.color_fmt, tp.clear_color, clear.color);
            } else if (tp.color_fmt.base == kBaseSInt) {
                const int32_t ci[4] = {(int32_t)tp.clear_color[0], (int32_t)tp.clear_color[1], (int32_t)tp.clear_color[2],
                                       (int32_t)tp.clear_color[3]};
                pack_color_i32(tp.color_fmt, ci, clear.color);
            } else {
                const float cf[4] = {__uint_as_float(tp.clear_color[0]), __uint_as_float(tp.clear_color[1]),
                                     __uint_as_float(tp.clear_color[2]), __uint_as_float(tp.clear_color[3])};
                pack_color_f32(tp.color_fmt, cf, clear.color);
            }
        }
        if (tp.has_ds && tp.ds_load == 1)
            pack_depth_stencil(tp.ds_fmt, __uint_as_float(tp.clear_ds[0]), tp.clear_ds[1], clear.ds0, clear.ds1);

        PixelState px[BPW];
        int pxx[BPW], pxy[BPW];
        bool inside[BPW];
#pragma unroll
        for (int k = 0; k < BPW; k++) {
            const int blk = slot + k * WAVES;
            const int bx0 = tile_x0 + (blk % BW) * 8, by0 = tile_y0 + (blk / BW) * 8;
            pxx[k] = bx0 + (lane & 7);
            pxy[k] = by0 + (lane >> 3);
            inside[k] = pxx[k] < (int)tp.width && pxy[k] < (int)tp.height;
            px[k] = clear;
            const uint64_t pix = (uint64_t)pxy[k] * tp.width + (uint64_t)pxx[k];
            if (inside[k] && tp.has_color && tp.color_load != 1)
                load_pixel((const uint8_t*)tp.color_addr + pix * tp.color_bytes, tp.color_bytes, px[k].color);
            if (inside[k] && tp.has_ds && tp.ds_load != 1) {
                uint32_t w4[4];
                load_pixel((const uint8_t*)tp.ds_addr + pix * tp.ds_bytes, tp.ds_bytes, w4);
                px[k].ds0 = w4[0];
                px[k].ds1 = w4[1];
            } else if (inside[k] && tp.has_ds && tp.ds_bytes == 8) {
                // D32_SFLOAT_S8_UINT cleared: SetPixelDepthStencil writes the
                // depth and the stencil byte and keeps the three padding bytes
                // (ImageCompiler.cpp:895-945)
                uint32_t w4[4];
                load_pixel((const uint8_t*)tp.ds_addr + pix * tp.ds_bytes, tp.ds_bytes, w4);
                px[k].ds1 = (w4[1] & 0xFFFFFF00u) | (clear.ds1 & 0xFFu);
            }
        }

        // ---- per-draw state, refreshed when the draw changes
        uint32_t cur_draw = 0xFFFFFFFFu;
        float xf[BPW], yf[BPW];
        float vp_min = 0.0f, vp_max = 1.0f;
        bool vp_identity = true;
        const FragmentState* fst = nullptr;
        const DrawEnv* env = nullptr;
        // Conservative depth rejection allowed for this draw, and its op.
        // With the op a constant (OP >= 0: every draw has the depth test
        // with op OP and no stencil or depth-bounds test) both are constants
        // from the start, so the rejection test folds to one compare.
        bool early = OP >= 0;
        int32_t depth_op = OP >= 0 ? OP : 7;
        int32_t depth_kind = 0;
        unsigned long long nfrag = 0;

        // Deferred colour work: a fragment that passed the depth/stencil
        // tests has its depth/stencil written at once, and its colour
        // (interpolation, fragment shader, blend, pack) queued.  Queues drain
        // in order, so every pixel still sees its colour writes in primitive
        // order; draining for all lanes at once keeps the shading path from
        // running once per primitive whenever any lane passes.
        constexpr int Q = FS::kMayDiscard ? 1 : 4;
        uint32_t q_prim[Q];
        float q_w0[Q], q_w1[Q], q_w2[Q];
        int qn = 0;

        auto shade = [&](uint32_t gprim, int k, float w0, float w1, float w2) {
            PixelState ps = px[0];
            int sx = pxx[0], sy = pxy[0];
#pragma unroll
            for (int kk = 1; kk < BPW; kk++)
                if (kk == k) { ps = px[kk]; sx = pxx[kk]; sy = pxy[kk]; }
            ps = shade_fragment<FS, ANYPRIM, FIXED>(a.recs, a.draws, a.fstates, a.uniform_fstate, a.vout, a.vout_stride, gprim, w0, w1,
                                                    w2, sx, sy, ps);
#pragma unroll
            for (int kk = 0; kk < BPW; kk++)
                if (kk == k) px[kk] = ps;
        };
        // Shades the oldest queued entry of every lane that has one and
        // shifts the queues: called when some lane's queue is full, this
        // keeps the other lanes' queues filled, so each shading pass runs
        // with as many lanes as possible.
        auto drain_one = [&]() {
            if (qn > 0) shade(q_prim[0] & 0x0FFFFFFFu, (int)(q_prim[0] >> 28), q_w0[0], q_w1[0], q_w2[0]);
#pragma unroll
            for (int e = 0; e + 1 < Q; e++) {
                q_prim[e] = q_prim[e + 1];
                q_w0[e] = q_w0[e + 1];
                q_w1[e] = q_w1[e + 1];
                q_w2[e] = q_w2[e + 1];
            }
            qn = qn > 0 ? qn - 1 : 0;
        };

        // ---- walk the tile's primitives in submission order
        const uint32_t lbeg = a.tile_start[tile], lend = a.tile_start[tile + 1];
        uint32_t e_next = lbeg + (uint32_t)lane < lend ? a.tile_prims[lbeg + lane] : 0u;
        for (uint32_t base = lbeg; base < lend; base += 64) {
            const int n = (int)min(64u, lend - base);
            // Depth range stored in this wave's pixels, for hierarchical
            // culling: a primitive whose depth interval fails the test
            // against all of them cannot change any of them.
            float bmin = __builtin_inff(), bmax = -__builtin_inff();
            if (hiz_op >= 0) {
#pragma unroll
                for (int k = 0; k < BPW; k++) {
                    if (inside[k]) {
                        const float cd = stored_depth(FIXED ? 32 : a.hiz_depth_kind, px[k].ds0);
                        bmin = fminf(bmin, cd);
                        bmax = fmaxf(bmax, cd);
                    }
                }
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    bmin = fminf(bmin, __shfl_xor(bmin, off, 64));
                    bmax = fmaxf(bmax, __shfl_xor(bmax, off, 64));
                }
            }
            // each lane tests one primitive of the chunk against the wave's area
            const uint32_t e = e_next;
            uint32_t gp_l = 0, blocks_l = 0;
            bool ov = false;
            if (lane < n) {
                gp_l = e & kPairPrim;
                blocks_l = e >> kPairBlockShift;
                if (TILE == 8) {
                    // one block per tile: a pair exists only where the box meets the tile
                    ov = true;
                } else if (TILE == 16) {
                    // the pair carries the blocks the primitive may cover
                    uint32_t mine = 0u;
#pragma unroll
                    for (int k = 0; k < BPW; k++) mine |= 1u << (slot + k * WAVES);
                    ov = (blocks_l & mine) != 0u;
                } else {
                    // the pair carries the 16x16 quadrants the primitive may
                    // cover; a wave's blocks lie in the quadrants of mine
                    uint32_t mine = 0u;
#pragma unroll
                    for (int k = 0; k < BPW; k++) {
                        const int blk = slot + k * WAVES;
                        mine |= 1u << ((blk / BW) / 2 * 2 + (blk % BW) / 2);
                    }
                    ov = (blocks_l & mine) != 0u;
                }
            }
            // the candidates' visit fields, all lanes at once, straight into
            // the wave's LDS region (global_load_lds: no registers)
            float4 dz = make_float4(0.0f, 0.0f, 0.0f, 0.0f);   // dt_lo, dt_hi, area, draw
            if constexpr (GATHER) {
                wave_lds_fence();   // the previous chunk's reads come first
                if (ov) {
                    const uint4* src = (const uint4*)&a.recs[gp_l];
#pragma unroll
                    for (int p = 0; p < kVisitParts; p++)
                        __builtin_amdgcn_global_load_lds((const void*)(src + p),
                                                         (__attribute__((address_space(3))) void*)&stage[wave][p][0], 16, 0, 0);
                    dz = *(const float4*)&a.recs[gp_l].dt_lo;
                }
            }
            // the next chunk's pairs, loaded while this chunk is worked on
            {
                const uint32_t nb = base + 64u + (uint32_t)lane;
                e_next = nb < lend ? a.tile_prims[nb] : 0u;
            }
            if (hiz_op >= 0) {
                if constexpr (BLOCK_DEPTH) {
                    // A triangle's depth bounded over the pixels it may cover
                    // in this wave's block (block_depth_range): its box
                    // clipped to the block, far tighter than its bounds over
                    // all its pixels.  Taken for the candidates of the draw
                    // last visited, whose pixel centres the lanes hold, when
                    // that draw's depth range is [0, 1] (the depth then is
                    // the tested value).  The record's visit fields are
                    // loaded with the depth interval, at once: one memory
                    // latency per chunk; before the tile's first visit
                    // (all of a short list) only the interval is loaded.
                    const bool near = cur_draw != 0xFFFFFFFFu && vp_identity;   // wave-uniform
                    uint4 p0 = make_uint4(0u, 0u, 0u, 0u), p1 = p0, p2 = p0, p3 = p0, p5 = p0;
                    if (ov) {
                        const uint4* src = (const uint4*)&a.recs[gp_l];
                        if (near) {
                            p0 = src[0]; p1 = src[1]; p2 = src[2]; p3 = src[3];
                            p5 = src[5];
                        }
                        dz = *(const float4*)(src + 4);
                        ov = !hiz_rejects(hiz_op, dz.x, dz.y, bmin, bmax);
                    }
                    if (near) {
                        // lane c holds column bx0 + c, lane 8 r row by0 + r
                        const int bx0 = __builtin_amdgcn_readfirstlane(pxx[0]), by0 = __builtin_amdgcn_readfirstlane(pxy[0]);
                        const int cl = min(max((int)p3.x - bx0, 0), 7), ch = min(max((int)p3.z - 1 - bx0, 0), 7);
                        const int rl = min(max((int)p3.y - by0, 0), 7), rh = min(max((int)p3.w - 1 - by0, 0), 7);
                        const float xa = __shfl(xf[0], cl, 64), xb = __shfl(xf[0], ch, 64);
                        const float ya = __shfl(yf[0], rl * 8, 64), yb = __shfl(yf[0], rh * 8, 64);
                        if (ov && __float_as_uint(dz.w) == cur_draw) {
                            const float ax[3] = {__uint_as_float(p0.x), __uint_as_float(p0.y), __uint_as_float(p0.z)};
                            const float ay[3] = {__uint_as_float(p0.w), __uint_as_float(p1.x), __uint_as_float(p1.y)};
                            const float dy[3] = {__uint_as_float(p1.z), __uint_as_float(p1.w), __uint_as_float(p2.x)};
                            const float dx[3] = {__uint_as_float(p2.y), __uint_as_float(p2.z), __uint_as_float(p2.w)};
                            const float z[3] = {__uint_as_float(p5.x), __uint_as_float(p5.y), __uint_as_float(p5.z)};
                            float lo, hi;
                            if (block_depth_range(ax, ay, dy, dx, dz.z, z, fminf(xa, xb), fmaxf(xa, xb), fminf(ya, yb), fmaxf(ya, yb),
                                                  lo, hi))
                                ov = !hiz_rejects(hiz_op, lo, hi, bmin, bmax);
                        }
                    }
                } else if (ov) {
                    if constexpr (!GATHER) dz = *(const float4*)&a.recs[gp_l].dt_lo;
                    ov = !hiz_rejects(hiz_op, dz.x, dz.y, bmin, bmax);
                }
            }
            uint64_t mask = __ballot(ov);
            if constexpr (STAGE) {
                __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): the LDS loads have landed
                wave_lds_fence();
            }
            while (mask) {
                const int j = (int)__builtin_ctzll(mask);
                mask &= mask - 1;
                const uint32_t gprim = __builtin_amdgcn_readlane(gp_l, j);
                const uint32_t blocks = (TILE >= 16 && BPW > 1) ? __builtin_amdgcn_readlane(blocks_l, j) : 0u;
                TriSetup r;
                if constexpr (STAGE) {
                    uint32_t* dst = (uint32_t*)&r;
#pragma unroll
                    for (int p = 0; p < kVisitParts; p++) {
                        // uniform words: kept in scalar registers
                        const uint4 v = stage[wave][p][j];
                        dst[4 * p + 0] = __builtin_amdgcn_readfirstlane(v.x);
                        dst[4 * p + 1] = __builtin_amdgcn_readfirstlane(v.y);
                        dst[4 * p + 2] = __builtin_amdgcn_readfirstlane(v.z);
                        dst[4 * p + 3] = __builtin_amdgcn_readfirstlane(v.w);
                    }
                } else {
                    load_record_uniform(a.recs, gprim, r);
                    // the loads of all visit fields issued before the first of
                    // them is used: one memory latency per survivor, not one
                    // for the draw index and another for the rest
                    __asm__ volatile("" ::"s"(r.ax[0]), "s"(r.y1), "s"(r.dt_lo), "s"(r.draw), "s"(r.z[0]), "s"(r.flags));
                }
                if (r.draw != cur_draw) {
                    cur_draw = r.draw;
                    const DrawDesc& d = a.draws[cur_draw];
                    // pixel centres (ProcessTriangles :1571-1580)
#pragma unroll
                    for (int k = 0; k < BPW; k++) {
                        xf[k] = ((float)pxx[k] / d.vp_w + d.half_x) * 2.0f - 1.0f;
                        yf[k] = ((float)pxy[k] / d.vp_h + d.half_y) * 2.0f - 1.0f;
                    }
                    vp_min = d.vp_min_depth;
                    vp_max = d.vp_max_depth;
                    vp_identity = vp_min == 0.0f && vp_max == 1.0f;
                    fst = &a.fstates[d.fstate];
                    env = (const DrawEnv*)d.env;
                    // Rejecting on the depth test alone is exact when nothing
                    // else observes the fragment: no depth-bounds or stencil
                    // test (and so no stencil write).
                    if constexpr (OP >= 0) {
                        depth_kind = 32;
                    } else if constexpr (FIXED) {
                        early = fst->depth_test && !fst->depth_bounds_test;
                        depth_kind = 32;
                    } else {
                        early = fst->depth_test && fst->has_ds && fst->ds.depth_off >= 0 && fst->ds.vk_format != 127u &&
                                !fst->depth_bounds_test && !fst->stencil_test;
                        depth_kind = fst->ds.depth_kind;
                    }
                    depth_op = OP >= 0 ? OP : fst->depth_compare;
                }
#pragma unroll
                for (int k = 0; k < BPW; k++) {
                    // every lane is active here; a lane queues at most one
                    // entry per iteration, so a free slot always exists
                    if (!FS::kMayDiscard && __any(qn == Q)) drain_one();
                    const int blk = slot + k * WAVES;
                    const int bx0 = tile_x0 + (blk % BW) * 8, by0 = tile_y0 + (blk / BW) * 8;
                    if (TILE == 16 && BPW > 1) {
                        if (!((blocks >> blk) & 1u)) continue;   // wave-uniform
                    } else if (BPW > 1) {
                        // wave-uniform: the block's quadrant may be covered
                        // and the block meets the primitive's box
                        if (TILE == 32 && !((blocks >> ((blk / BW) / 2 * 2 + (blk % BW) / 2)) & 1u)) continue;
                        if (r.x1 <= bx0 || r.x0 >= bx0 + 8 || r.y1 <= by0 || r.y0 >= by0 + 8) continue;
                    }
                    const int x = pxx[k], y = pxy[k];
                    if (x < r.x0 || x >= r.x1 || y < r.y0 || y >= r.y1) continue;
                    if (hiz_op >= 0) {
                        const float cd = stored_depth(FIXED ? 32 : a.hiz_depth_kind, px[k].ds0);
                        if (hiz_rejects(hiz_op, r.dt_lo, r.dt_hi, cd, cd)) continue;
                    }
                    float w0, w1, w2, depth;
                    if (ANYPRIM && (r.flags & kPrimPoint)) {
                        // ProcessPoints (:1358-1378): s, t of the pixel in the
                        // point's square, the vertex's outputs and depth as they are
                        const float size = r.dx[0];
                        const float s = 0.5f + (float)(int32_t)((uint32_t)x - (uint32_t)__float_as_int(r.ax[0])) / size;
                        const float t = 0.5f + (float)(int32_t)((uint32_t)y - (uint32_t)__float_as_int(r.ay[0])) / size;
                        if (!(s >= 0.0f && t >= 0.0f && s <= 1.0f && t <= 1.0f)) continue;
                        nfrag++;
                        w0 = 1.0f; w1 = 0.0f; w2 = 0.0f;
                        depth = r.z[0];
                    } else if (ANYPRIM && (r.flags & kPrimLine)) {
                        // ProcessLines (:1420-1497): all four quad edges >= 0
                        // (NaN fails), t = dot(p - p0, p1 - p0) / |p1 - p0|^2
                        const LineExt L = load_line_ext(a.line_ext, gprim);
                        const float e0 = (xf[k] - r.ax[0]) * r.dy[0] - (yf[k] - r.ay[0]) * r.dx[0];
                        const float e1 = (xf[k] - r.ax[1]) * r.dy[1] - (yf[k] - r.ay[1]) * r.dx[1];
                        const float e2 = (xf[k] - r.ax[2]) * r.dy[2] - (yf[k] - r.ay[2]) * r.dx[2];
                        const float e3 = (xf[k] - L.ax3) * L.dy3 - (yf[k] - L.ay3) * L.dx3;
                        if (!(e0 >= 0.0f && e1 >= 0.0f && e2 >= 0.0f && e3 >= 0.0f)) continue;
                        nfrag++;
                        const float t = ((xf[k] - L.p0x) * L.ddx + (yf[k] - L.p0y) * L.ddy) / r.area;
                        w0 = 1.0f - t; w1 = t; w2 = 0.0f;
                        depth = r.z[0] * t + r.z[1] * (1.0f - t);
                    } else {
                        // GetFragmentInput (:874-903); NaN edge values are not < 0 and pass
                        const float e0 = (xf[k] - r.ax[0]) * r.dy[0] - (yf[k] - r.ay[0]) * r.dx[0];
                        const float e1 = (xf[k] - r.ax[1]) * r.dy[1] - (yf[k] - r.ay[1]) * r.dx[1];
                        const float e2 = (xf[k] - r.ax[2]) * r.dy[2] - (yf[k] - r.ay[2]) * r.dx[2];
                        if (e0 < 0 || e1 < 0 || e2 < 0) continue;
                        nfrag++;
                        if (early) {
                            const float cur = (FIXED || depth_kind == 32) ? __uint_as_float(px[k].ds0) : unpack_depth(fst->ds, px[k].ds0);
                            if (vp_identity ? depth_surely_fails_unit(depth_op, cur, e0, e1, e2, r.area, r.z[0], r.z[1], r.z[2])
                                            : depth_surely_fails(depth_op, cur, e0, e1, e2, r.area, r.z[0], r.z[1], r.z[2], vp_min, vp_max))
                                continue;
                        }
                        w0 = e0 / r.area; w1 = e1 / r.area; w2 = e2 / r.area;
                        depth = r.z[0] * w0 + r.z[1] * w1 + r.z[2] * w2;
                    }
                    // DrawPixel (:1300-1313)
                    const float dt = (vp_max - vp_min) * depth + vp_min;
                    const bool front = (r.flags & 1u) != 0;
                    if (FS::kMayDiscard) {
                        // the shader decides first whether anything is written
                        float in[FS::kInputs > 0 ? FS::kInputs : 1];
                        fragment_inputs<FS, ANYPRIM>(r, a.vout, a.vout_stride, w0, w1, w2, in);
                        const float fc[4] = {fst->origin_upper_left ? (float)x : a.draws[cur_draw].vp_w - (float)x - 1.0f, (float)y,
                                             depth, 1.0f};
                        float colour[4];
                        if (FS::run(*env, in, fc, front, colour)) continue;
                        const DepthStencilResult dr = depth_stencil_call<FIXED, OP>(fst, px[k], dt, front ? 1u : 0u);
                        px[k] = dr.px;
                        if (dr.write && (FIXED || fst->has_color)) color_write_state<FIXED>(*fst, px[k], colour);
                    } else {
                        const DepthStencilResult dr = depth_stencil_call<FIXED, OP>(fst, px[k], dt, front ? 1u : 0u);
                        px[k] = dr.px;
                        if (!(dr.write && (FIXED || fst->has_color))) continue;
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
        if (!FS::kMayDiscard) {
            while (__any(qn > 0)) drain_one();
        }
        (void)env;

        // ---- write back
#pragma unroll
        for (int k = 0; k < BPW; k++) {
            if (!inside[k]) continue;
            const uint64_t pix = (uint64_t)pxy[k] * tp.width + (uint64_t)pxx[k];
            if (tp.has_color) store_pixel((uint8_t*)tp.color_addr + pix * tp.color_bytes, FIXED ? 4 : tp.color_bytes, px[k].color);
            if (tp.has_ds) {
                const uint32_t w4[4] = {px[k].ds0, px[k].ds1, 0u, 0u};
                store_pixel((uint8_t*)tp.ds_addr + pix * tp.ds_bytes, FIXED ? 4 : tp.ds_bytes, w4);
            }
        }

        // ---- statistics
        for (int off = 32; off > 0; off >>= 1) nfrag += __shfl_down(nfrag, off, 64);
        if (lane == 0 && nfrag) atomicAdd(a.fragments, nfrag);
        if (a.trace_cycles && lane == 0) a.trace_cycles[(uint64_t)tile * WAVES + slot] = (float)(clock64() - t_begin);

        if (!a.persistent) break;
    }
}

// ANYPRIM: the launch may hold point and line records besides triangles
// (one 16x16 / four-wave instantiation per fragment shader carries it).
// FIXED: every pipeline and target of the launch has the fixed formats
// (fold_fixed_formats).  The fixed-format kernels carry a second copy of
// the body for depth op LESS, the common one.
template <class FS, int TILE, int WAVES, bool ANYPRIM = false, bool FIXED = false>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(raster_waves_per_eu<FS, TILE, WAVES>()))) void
raster_kernel(RasterArgs a)
{
    __shared__ RasterStage<FS, TILE, WAVES> stage;
    if constexpr (FIXED) {
        if (a.hiz_op == 1) {
            raster_body<FS, TILE, WAVES, ANYPRIM, FIXED, 1>(a, stage);
            return;
        }
    }
    raster_body<FS, TILE, WAVES, ANYPRIM, FIXED, -1>(a, stage);
}

}  // namespace cpv
