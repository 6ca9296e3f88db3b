// Tile scheduling kernels: the per-tile binning statistics that the cost
// predictor reads (and that cost traces record next to the measured cycles),
// the threshold from which a tile counts as heavy, and the sort keys that
// turn predicted costs into the raster kernel's dispatch order.  None of
// them changes what is drawn: they only decide when tiles are taken.
#pragma once

#include "cpv_types.h"

namespace cpv {

struct FeatureArgs {
    const TriSetup* recs;
    const uint32_t* tile_prims;
    const uint32_t* tile_start;
    const TargetParams* targets;
    const DrawDesc* draws;
    const FragmentState* fstates;
    float* features;          // tiles x 16 floats; columns 12-15 are zero
    uint32_t tiles;
    uint32_t target_count;
    int32_t tile;             // tile edge in pixels
    int32_t waves;            // wavefronts per tile
    int32_t hiz;              // hierarchical depth culling on (1) or off (0)
    int32_t pad;
};

__device__ __forceinline__ uint32_t feature_target(const TargetParams* t, uint32_t n, uint32_t tile)
{
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (t[mid].tile_base <= tile) lo = mid; else hi = mid;
    }
    return lo;
}

// One wavefront per tile.  Lists of up to 64 primitives are read whole;
// longer lists are sampled at 64 evenly spaced entries and the sums scaled
// by n / 64, which bounds the kernel's cost per tile (the statistics only
// steer the queue order).
//   f0  log2(1 + n)                         n = primitives binned to the tile
//   f1  n / 1024
//   f2  log2(1 + sum of bounding-box pixels inside the tile)
//   f3  (sum of bounding-box pixels inside the tile) / (64 * tile pixels)
//   f4  log2(1 + primitives whose bounding box covers the whole tile)
//   f5  log2(1 + sum of min(triangle pixels, tile pixels))
//   f6  fraction of the tile inside the attachment
//   f7  tile edge / 32
//   f8  wavefronts per tile / 4
//   f9  hierarchical depth culling on
//   f10 fraction of primitives whose draw blends
//   f11 mean depth span (dt_hi - dt_lo) of the primitives
__global__ __launch_bounds__(256) void tile_features_kernel(FeatureArgs a)
{
    const uint32_t tile = blockIdx.x * 4u + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (tile >= a.tiles) return;
    const TargetParams& tp = a.targets[feature_target(a.targets, a.target_count, tile)];
    const uint32_t local = tile - tp.tile_base;
    const int x0 = (int)(local % tp.tiles_x) * a.tile, y0 = (int)(local / tp.tiles_x) * a.tile;
    const int x1 = min(x0 + a.tile, (int)tp.width), y1 = min(y0 + a.tile, (int)tp.height);
    const float tile_px = (float)(a.tile * a.tile);
    const uint32_t b = a.tile_start[tile], e = a.tile_start[tile + 1];
    float s_box = 0.0f, s_full = 0.0f, s_tri = 0.0f, s_blend = 0.0f, s_span = 0.0f;
    const uint32_t n_all = e - b;
    const uint32_t m = min(n_all, 64u);
    if ((uint32_t)lane < m) {
        const uint32_t i = b + (uint32_t)(((uint64_t)lane * n_all) / m);
        const TriSetup& r = a.recs[a.tile_prims[i] & kPairPrim];
        const int cx0 = max(r.x0, x0), cy0 = max(r.y0, y0), cx1 = min(r.x1, x1), cy1 = min(r.y1, y1);
        const float box = (float)(max(cx1 - cx0, 0) * max(cy1 - cy0, 0));
        s_box += box;
        s_full += (r.x0 <= x0 && r.y0 <= y0 && r.x1 >= x1 && r.y1 >= y1) ? 1.0f : 0.0f;
        const DrawDesc& d = a.draws[r.draw];
        const float tri = fabsf(r.area) * d.vp_w * d.vp_h * 0.125f;   // edge function = twice the NDC area
        s_tri += fminf(tri, tile_px);
        s_blend += a.fstates[d.fstate].blend_enable ? 1.0f : 0.0f;
        s_span += fminf(fmaxf(r.dt_hi - r.dt_lo, 0.0f), 1.0f);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        s_box += __shfl_xor(s_box, off, 64);
        s_full += __shfl_xor(s_full, off, 64);
        s_tri += __shfl_xor(s_tri, off, 64);
        s_blend += __shfl_xor(s_blend, off, 64);
        s_span += __shfl_xor(s_span, off, 64);
    }
    const float scale = m ? (float)n_all / (float)m : 0.0f;
    s_box *= scale;
    s_full *= scale;
    s_tri *= scale;
    if (lane < 16) {
        const float n = (float)n_all;
        const float inv_n = m ? 1.0f / (float)m : 0.0f;   // s_blend, s_span are sample sums
        float v = 0.0f;
        switch (lane) {
        case 0: v = log2f(1.0f + n); break;
        case 1: v = n * (1.0f / 1024.0f); break;
        case 2: v = log2f(1.0f + s_box); break;
        case 3: v = s_box / (64.0f * tile_px); break;
        case 4: v = log2f(1.0f + s_full); break;
        case 5: v = log2f(1.0f + s_tri); break;
        case 6: v = (float)(max(x1 - x0, 0) * max(y1 - y0, 0)) / tile_px; break;
        case 7: v = (float)a.tile * (1.0f / 32.0f); break;
        case 8: v = (float)a.waves * 0.25f; break;
        case 9: v = (float)a.hiz; break;
        case 10: v = s_blend * inv_n; break;
        case 11: v = s_span * inv_n; break;
        default: v = 0.0f; break;
        }
        a.features[(uint64_t)tile * 16u + lane] = v;
    }
}

// The predicted cost from which a tile counts as heavy: within a factor of
// 4 (2 in log2 cost) of the costliest prediction.  The heavy tiles are the
// frame's critical path: they are dispatched first.  max_bits is zeroed
// beforehand; the maximum of the finite predictions is taken with an atomic
// max on their ordered integer images.
__device__ __forceinline__ uint32_t ordered_bits(float f)
{
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float from_ordered_bits(uint32_t o)
{
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

__global__ __launch_bounds__(256) void pred_max_kernel(const float* pred, uint32_t tiles, uint32_t* max_bits)
{
    uint32_t m = 0u;
    for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < tiles; i += gridDim.x * 256u) {
        const float p = pred[i];
        if (__builtin_isfinite(p)) m = max(m, ordered_bits(p));
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, off, 64));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(max_bits, m);
}

__global__ void heavy_threshold_kernel(const uint32_t* max_bits, float* out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = *max_bits ? from_ordered_bits(*max_bits) - 2.0f : __builtin_inff();
}

// Dispatch order keys: the stable radix sort of (key, tile) by a 16-bit key
// puts the heavy tiles first, costliest first, and keeps every other tile
// in index order, so that neighbouring tiles, which share their primitives'
// records, still run together (ordering all tiles by predicted cost costs
// more in record re-reads than it saves in tail).
__global__ __launch_bounds__(256) void order_keys_kernel(const float* pred, const float* heavy_from, uint32_t tiles, uint32_t* keys,
                                                         uint32_t* vals)
{
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i >= tiles) return;
    const float p = pred[i];
    uint32_t k = 65535u;
    if (p >= *heavy_from) {
        const float q = fminf(fmaxf(p, 0.0f), 63.98f) * 1024.0f;   // NaN fails the test above
        k = 65534u - (uint32_t)q;
    }
    keys[i] = k;
    vals[i] = i;
}

}  // namespace cpv
